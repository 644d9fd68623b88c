"""Benchmark of the MI355X BD timestep: particle-steps/s at 2M particles with a
2048^2 TSC mesh (L = 1414 sigma, N1 = N2 = 10^6, lambda = 50, q2/q1 = 0.5,
kT = 1, dt = 1e-4, FP32 forces with FP64 positions), synthetic clipped
Gaussian-blob start (thinned hexagonal lattice of spacing 0.95 sigma, so every pair
is farther apart than the 0.9 sigma exclusion).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--no-ensemble] [--no-config0]

Prints one JSON line.  The timed region is K graph-replayed steps with the
particle state resident in HBM, bracketed by stream synchronisation.  The
line also carries the ensemble figures: a batched mix of 256 uneven
replicas (one launch per stage) against the plain one-at-a-time GPU path,
and a served Poisson job stream (continuous batching) with p50/p99 job
latency; see bd-simulations_amd/ensemble.py for the mixes; and the CPU
reference configuration (2000 particles, CIC 128^2, 10,000 FP64 steps) on
the GPU beside the FP64 CPU restatement on one host thread.
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bd-simulations_amd"))

PEAK_HBM_GBS = 8000.0

WORKLOAD = dict(name="2M particles, TSC 2048^2, L=1414, lambda=50, blob start",
                n1=1_000_000, n2=1_000_000, L=1414.0, Z=2048, lam=50.0, kap=1.0, f2_f1=0.5, T=1.0,
                dt=1e-4, scheme=2, blob_width=1414.0 / 6.0, background=0.1, dmin=0.95, seed=1)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(kernel, N, Z, rb=4):
    """Compulsory HBM bytes of one launch (FP32 forces/mesh, FP64 positions):
    P = 16 B position state, f = 8 B force pair, g = 4 B mesh real."""
    P, f, g = 16, 8, rb
    table = {
        # read positions, write the WCA force (neighbour reads hit the caches)
        "k_wca": P * N + f * N,
        # read positions, write the deposited mesh
        "k_deposit": P * N + g * Z * Z,
        # read real mesh, write the half spectrum (Z x Z/2 complex)
        "k_r2c_rows": g * Z * Z + g * Z * Z,
        # read half spectrum, write the Fx and Fy half spectra
        "k_col": g * Z * Z + 2 * g * Z * Z,
        # read both half spectra, write the real (Fx, Fy) mesh
        "k_c2r_rows": 2 * g * Z * Z + 2 * g * Z * Z,
        # read positions + WCA force + (Fx, Fy) mesh, write positions
        "k_update": (P + f) * N + 2 * g * Z * Z + P * N,
    }
    return table.get(kernel)


def canonical_step_bytes(N, Z):
    """Per-step compulsory traffic of the whole path, 64 N + 36 Z^2 bytes:
    positions read twice and written once, forces written and read, the mesh
    written by the deposit, read and written by a real-to-complex and two
    complex-to-real transforms, and read by the gather."""
    return 64 * N + 36 * Z * Z


# the kernels of one graph-replayed step (fixed dt, FP32 forces, TSC)
STEP_KERNELS = ("k_update", "k_rank_scan", "k_fill", "k_sort_cells", "k_wca_f32", "k_deposit", "k_r2c_rows",
                "k_col", "k_c2r_rows")
# bench names of the kernels timed one at a time -> names in the profiler's records
TRACE_NAME = {"k_wca": "k_wca_f32"}


def host_cpu():
    """CPU model, logical CPUs and the CPUs this process may run on."""
    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "nproc": os.cpu_count() or 1, "usable_cpus": usable}


def load_traffic(kernel):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch), each scaled by the same counter's reading
    for a copy kernel of known size; committed under profiles/ as JSON."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        k = d["kernels"].get(kernel)
        if not k:
            return None
        return 1024.0 * (k["fetch_kib"] * d["fetch_scale"] + k["write_kib"] * d["write_scale"])
    except (OSError, KeyError, ValueError, TypeError):
        return None


PROBE_COPY_BYTES = 256 << 20


def kernel_probe(sim, reps):
    """For rocprofv3 --pmc passes: each step kernel `reps` times at the
    benchmark state, then `reps` device-to-device copies of PROBE_COPY_BYTES
    (the runtime's copy kernel reads and writes each byte once), whose
    counter readings calibrate FETCH_SIZE and WRITE_SIZE."""
    import ctypes as C

    for k in ("k_wca", "k_deposit", "k_r2c_rows", "k_col", "k_c2r_rows", "k_gather", "k_update"):
        sim.profile_kernel(k, reps)
    sim.sync()
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    n = C.c_size_t(PROBE_COPY_BYTES)
    for buf in (a, b):
        if hip.hipMalloc(C.byref(buf), n) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, n) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(reps):
        if hip.hipMemcpy(b, a, n, 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log("kernel probe: %d launches per kernel, %d copies of %d bytes" % (reps, reps, PROBE_COPY_BYTES))


def cpu_baseline(p, git, rel, budget_s=20.0):
    """The FP64 CPU restatement (oracle/, one thread, like the reference's
    single-threaded build) timed on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    op = O.make_params(n1=p.n1, n2=p.n2, L=p.L, Z=p.Z, scheme=p.scheme, lam=p.lam, kap=p.kap, T=p.T,
                       dt_max=p.dt_max, mrd=float("inf"), f2_f1=p.f2_f1, noise_mode=1, seed=p.seed)
    o = O.Oracle(op)
    o.set_all(git, rel)
    log("cpu baseline: warm-up step")
    t0 = time.perf_counter()
    o.step(1.0)
    t1 = time.perf_counter()
    per = t1 - t0
    n = max(1, min(20, int(budget_s / max(per, 1e-3))))
    log("cpu baseline: %d timed steps (%.2f s per step)" % (n, per))
    t0 = time.perf_counter()
    for _ in range(n):
        o.step(1.0)
    dt = time.perf_counter() - t0
    o.close()
    N = p.n1 + p.n2
    hc = host_cpu()
    return {"value": N * n / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": "FP64 C restatement, 1 thread, %d steps after 1 warm-up of the same 2M-particle "
                      "TSC 2048^2 workload (%.1f s); host %s, %d CPUs (%d usable)"
                      % (n, dt, hc["cpu_model"], hc["nproc"], hc["usable_cpus"]), **hc}


def cpu_ensemble_baseline(mix, budget_s=10.0):
    """All host cores over independent replicas of the batched mix: the FP64
    oracle, one replica per worker process (at most 16), each stepped for
    budget_s after one warm-up step; aggregate particle-steps/s."""
    import multiprocessing as mp

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    import pool

    O.build()
    cores = min(16, int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1)
    jobs = []
    for p, git, rel in mix[:cores]:
        kw = dict(n1=p.n1, n2=p.n2, L=p.L, Z=p.Z, scheme=p.scheme, lam=p.lam, kap=p.kap, T=p.T,
                  dt_max=p.dt_max, mrd=float("inf"), f2_f1=p.f2_f1, noise_mode=1, seed=p.seed)
        jobs.append((kw, git, rel, budget_s))
    with mp.get_context("spawn").Pool(len(jobs)) as workers:
        res = workers.map(pool.run_replica, jobs)
    value = sum(ps / t for ps, t in res)
    hc = host_cpu()
    return {"value": value, "unit": "particle-steps/s", "cores": len(jobs), "kind": "port",
            "sample": "FP64 C restatement, %d processes side by side, each one replica of the batched mix "
                      "(the first %d, %d particles in all) stepped %.0f s after 1 warm-up step; host %s"
                      % (len(jobs), len(jobs), sum(j[0]["n1"] + j[0]["n2"] for j in jobs), budget_s,
                         hc["cpu_model"]), **hc}


CONFIG0 = dict(n1=1000, n2=1000, L=64.0, Z=128, scheme=1, lam=8.0, kap=1.0, f2_f1=0.5, T=1.0, dt=1e-4, seed=5,
               steps=10000)


def config0_figures(bdk, cpu=True):
    """The reference's CPU configuration (N1 = N2 = 1000, L = 64, CIC 128^2,
    lambda 8, q2/q1 = 0.5, kT = 1, dt = 1e-4 fixed, 10,000 steps, FP64, RSA
    start with 0.9 exclusion): the GPU in FP64 mode, 10,000 graph-replayed
    steps, and with cpu the FP64 oracle on one host thread over the same
    10,000 steps (the reference's single-threaded run)."""
    c = CONFIG0
    p = bdk.make_params(n1=c["n1"], n2=c["n2"], L=c["L"], Z=c["Z"], scheme=c["scheme"], lam=c["lam"], kap=c["kap"],
                        f2_f1=c["f2_f1"], T=c["T"], dt_max=c["dt"], precision=bdk.FP64, noise_mode=bdk.NOISE_DISC,
                        seed=c["seed"])
    git, rel = bdk.init_uniform(p, c["seed"], 0.9)
    N = c["n1"] + c["n2"]
    log("config0: GPU, FP64, %d steps" % c["steps"])
    sim = bdk.RunDynamik(p)
    sim.set_positions(git, rel)
    sim.run_steps(100)
    sim.sync()
    t0 = time.perf_counter()
    sim.run_steps(c["steps"])
    sim.sync()
    gpu_s = time.perf_counter() - t0
    sim.close()
    out = {"workload": "N1 = N2 = 1000, L = 64, CIC 128^2, lambda 8, kT 1, dt 1e-4 fixed, FP64, RSA start",
           "steps": c["steps"], "gpu_value": N * c["steps"] / gpu_s, "gpu_ms_per_step": gpu_s / c["steps"] * 1e3,
           "unit": "particle-steps/s"}
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        O.build()
        op = O.make_params(n1=c["n1"], n2=c["n2"], L=c["L"], Z=c["Z"], scheme=c["scheme"], lam=c["lam"],
                           kap=c["kap"], T=c["T"], dt_max=c["dt"], mrd=float("inf"), f2_f1=c["f2_f1"],
                           noise_mode=1, seed=c["seed"])
        o = O.Oracle(op)
        o.set_all(git, rel)
        log("config0: CPU oracle, FP64, 1 thread, %d steps" % c["steps"])
        t0 = time.perf_counter()
        for k in range(c["steps"]):
            o.step(1.0)
            if k % 2000 == 1999:
                log("config0: CPU step %d" % (k + 1))
        cpu_s = time.perf_counter() - t0
        o.close()
        hc = host_cpu()
        out["cpu_baseline"] = {"value": N * c["steps"] / cpu_s, "unit": "particle-steps/s", "cores": 1,
                               "kind": "port", "seconds": cpu_s,
                               "sample": "FP64 C restatement, 1 thread, all %d steps; host %s" % (c["steps"],
                                                                                                  hc["cpu_model"]),
                               **hc}
        out["gpu_over_cpu"] = out["gpu_value"] / out["cpu_baseline"]["value"]
    return out


def ensemble_figures(log_fn, cpu=False):
    """Batched mix (256 replicas) vs the plain path, and the served stream;
    with cpu, the all-core host baseline of the batched mix."""
    import numpy as np

    import ensemble

    log_fn("ensemble: building the 256-replica mix")
    mix = ensemble.batched_mix(256)
    sims = ensemble.make_sims(mix)
    n_mix = sum(s.N for s in sims)
    plain_steps, batch_steps = 100, 500
    log_fn("ensemble: plain path, %d steps per replica, one replica at a time" % plain_steps)
    ps_plain, t_plain = ensemble.time_plain(sims, plain_steps)
    log_fn("ensemble: batched, %d steps (autotuned)" % batch_steps)
    ps_b, t_b, ept, stages = ensemble.time_batched(sims, batch_steps)
    for s in sims:
        s.close()
    batched = {"replicas": len(mix), "particles": n_mix, "steps": batch_steps, "value": ps_b,
               "unit": "particle-steps/s", "plain_value": ps_plain, "plain_steps": plain_steps,
               "speedup_vs_plain": ps_b / ps_plain, "tuning": ept,
               "stage_ms": {k: v for k, v in stages.items()},
               "mix": "N log-uniform [2e3, 6.4e4], density 0.5, Z = pow2(L/0.5) in [64, 512], "
                      "lambda U[5,50], CIC/TSC, FP32 forces, uniform start"}
    if cpu:
        log_fn("ensemble: all-core CPU baseline of the mix")
        batched["cpu_baseline"] = cpu_ensemble_baseline(mix)
    jobs = ensemble.served_jobs(64)
    work = float(np.mean([(p.n1 + p.n2) * st for p, _, _, st in jobs]))
    rate = 1.5 * ps_b / work
    log_fn("ensemble: serving %d jobs, Poisson arrivals at %.2f jobs/s" % (len(jobs), rate))
    tl = []
    lat, ps_srv, wall, _ = ensemble.serve(jobs, rate, max_jobs=64, timeline=tl)
    served = {"jobs": len(jobs), "arrival_rate_per_s": rate, "p50_latency_s": float(np.percentile(lat, 50)),
              "p99_latency_s": float(np.percentile(lat, 99)), "aggregate_value": ps_srv,
              "unit": "particle-steps/s", "wall_s": wall, "phases": ensemble.serve_phases(tl),
              "longest_job_particle_steps": max((p.n1 + p.n2) * st for p, _, _, st in jobs),
              "mix": "N log-uniform [1e3, 2.56e5], split U[0.2,0.8], density U[0.3,0.7], lambda U[5,100], "
                     "NGP/CIC/TSC, steps log-uniform [1e3, 1e5], uniform or blob start, travel cap 0.24 "
                     "(the reference's adaptive dt, dt_max 1e-4); offered load 1.5x "
                     "the batched throughput; budget 64 GiB, at most 64 jobs at once, 256-step ticks"}
    return batched, served


def adaptive_figure(bdk, w, git, rel, steps, warmup, N):
    """The same workload with the reference's default travel cap
    (max_reisedistanz = 0.24, parameter.h:43): dt per step from the device
    minimum of optimaler_zeitschritt's candidates (dm:1420-1491), dt_max 1e-4."""
    log("bench: adaptive dt (max_travel 0.24), %d steps" % steps)
    p = bdk.make_params(n1=w["n1"], n2=w["n2"], L=w["L"], Z=w["Z"], scheme=w["scheme"], lam=w["lam"],
                        kap=w["kap"], f2_f1=w["f2_f1"], T=w["T"], dt_max=w["dt"], precision=bdk.FP32,
                        noise_mode=bdk.NOISE_DISC, seed=w["seed"], max_travel=0.24)
    sim = bdk.RunDynamik(p)
    sim.set_positions(git, rel)
    sim.run_steps(warmup)
    t0 = time.perf_counter()
    t_sim = sim.run_steps(steps)
    elapsed = time.perf_counter() - t0
    sim.close()
    return {"max_travel": 0.24, "steps": steps, "ms_per_step": elapsed / steps * 1e3,
            "particle_steps_per_s": N * steps / elapsed, "mean_dt": t_sim / steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched and served figures")
    ap.add_argument("--no-config0", action="store_true", help="skip the CPU reference configuration's figures")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--kernel-probe", type=int, default=0, metavar="REPS",
                    help="run each kernel REPS times plus a calibration copy and exit (for --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the library runs one simulation on one GPU; --gpus %d is not supported "
              "(multi-GPU sharding is design-only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import bdk

    w = WORKLOAD
    p = bdk.make_params(n1=w["n1"], n2=w["n2"], L=w["L"], Z=w["Z"], scheme=w["scheme"], lam=w["lam"],
                        kap=w["kap"], f2_f1=w["f2_f1"], T=w["T"], dt_max=w["dt"], precision=bdk.FP32,
                        noise_mode=bdk.NOISE_DISC, seed=w["seed"])
    log("bench: initial state (blob, %d particles)" % (w["n1"] + w["n2"]))
    git, rel, clipped = bdk.init_blob(p, w["seed"], w["blob_width"], w["background"], w["dmin"])
    sim = bdk.RunDynamik(p)
    sim.set_positions(git, rel)
    N = w["n1"] + w["n2"]
    if args.kernel_probe:
        sim.run_steps(args.warmup)
        kernel_probe(sim, args.kernel_probe)
        sim.close()
        return
    log("bench: warm-up %d steps" % args.warmup)
    sim.run_steps(args.warmup)
    sim.sync()
    log("bench: timing %d steps" % args.steps)
    t0 = time.perf_counter()
    sim.run_steps(args.steps)
    sim.sync()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = N * args.steps / elapsed

    log("bench: per-kernel timing (%d back-to-back launches each, HIP events)" % args.profile_steps)
    kernels = {}
    for k in ("k_wca", "k_deposit", "k_r2c_rows", "k_col", "k_c2r_rows", "k_update"):
        kernels[k] = sim.profile_kernel(k, args.profile_steps)
    step_ms_events = sim.profile_kernel("step", args.profile_steps)
    own_bytes = sum(algorithmic_bytes(k, N, w["Z"]) for k in kernels)
    canon = canonical_step_bytes(N, w["Z"])
    # The headline roofline is one fixed quantity: the whole graph-replayed
    # step (every kernel of it, on both streams) against the path's
    # canonical compulsory bytes, 64 N + 36 Z^2, timed with HIP events over
    # back-to-back step launches on the simulation's stream; its traffic is
    # the measured HBM bytes of all the step's kernels (profiles/).
    tr = [load_traffic(k) for k in STEP_KERNELS]
    achieved = canon / (step_ms_events * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "whole step (graph replay of " + ", ".join(STEP_KERNELS) + ")",
            "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
            "traffic": sum(tr) if all(t is not None for t in tr) else None,
            "algorithmic_bytes": canon, "avg_ms": step_ms_events}
    # the fused force + update path (gather, noise, Euler step, refresh) on its own
    fu_ab = algorithmic_bytes("k_update", N, w["Z"])
    fu = fu_ab / (kernels["k_update"] * 1e-3) / 1e9
    fused = {"bound": "hbm", "kernel": "k_update", "achieved": fu, "peak": PEAK_HBM_GBS, "unit": "GB/s",
             "frac": fu / PEAK_HBM_GBS, "traffic": load_traffic("k_update"), "algorithmic_bytes": fu_ab,
             "avg_ms": kernels["k_update"]}
    out = {
        "metric": "particle-steps/s at 2M particles, 2048\u00b2 TSC mesh; p50/p99 ensemble job latency",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 forces / f64 positions", "data": "synthetic (seeded clipped Gaussian blob, %.1f%% clipped)"
                                                      % (100 * clipped),
        "config": {"workload": w["name"], "n_particles": N, "mesh": "%dx%d" % (w["Z"], w["Z"]),
                   "scheme": "TSC", "L": w["L"], "lambda": w["lam"], "dt": w["dt"], "kT": w["T"],
                   "parallelism": "single GPU"},
        "roofline": roof,
        "fused_update_roofline": fused,
        "step_roofline": {"canonical_bytes": canon, "kernel_bytes": own_bytes,
                          "achieved_GBs": canon / (ms_per_step * 1e-3) / 1e9,
                          "frac": canon / (ms_per_step * 1e-3) / 1e9 / PEAK_HBM_GBS,
                          "step_ms_wall": ms_per_step, "step_ms_events": step_ms_events},
        "kernels_ms": kernels,
        # every kernel against the HBM peak by its algorithmic bytes; k_update is
        # the fused gather + noise + Euler update (40 N + 8 Z^2 bytes)
        "kernels_roofline": {k: {"algorithmic_bytes": algorithmic_bytes(k, N, w["Z"]),
                                 "achieved_GBs": algorithmic_bytes(k, N, w["Z"]) / (ms * 1e-3) / 1e9,
                                 "frac": algorithmic_bytes(k, N, w["Z"]) / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
                                 "traffic": load_traffic(TRACE_NAME.get(k, k))}
                             for k, ms in kernels.items()},
        # measured HBM bytes of the list-rebuild kernels (implementation
        # overhead outside the canonical bytes)
        "rebuild_traffic": {k: load_traffic(k) for k in ("k_rank_scan", "k_fill", "k_sort_cells")},
    }
    sim.close()
    out["adaptive_dt"] = adaptive_figure(bdk, w, git, rel, args.steps, args.warmup, N)
    if not args.no_config0:
        out["config0"] = config0_figures(bdk, cpu=not args.no_cpu_baseline)
    if not args.no_ensemble:
        out["batched"], out["served"] = ensemble_figures(log, cpu=not args.no_cpu_baseline)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, git, rel)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
