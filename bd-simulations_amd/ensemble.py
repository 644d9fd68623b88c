"""Ensemble workloads: many independent replicas of uneven size advanced as
one batch (one launch per stage), and a served job stream with continuous
batching, each measured beside the plain one-simulation-at-a-time GPU path.

The replica mixes follow the reference's use of `runs` and parameter sweeps
(main.cpp:126-300 runs independent RunZustand objects one after another):
  batched mix - per replica total N log-uniform in [2000, 64000] split 50/50,
                area density 0.5, mesh the power of two nearest L/0.5
                clamped to [64, 512], lambda uniform in [5, 50], CIC or TSC,
                kT = 1, dt = 1e-4, uniform random start (exclusion 0.9);
  served jobs - total N log-uniform in [1000, 256000], species split uniform
                in [0.2, 0.8], density uniform in [0.3, 0.7], lambda uniform
                in [5, 100], NGP/CIC/TSC, steps log-uniform in [1e3, 1e5],
                uniform or Gaussian-blob start, the reference's default
                travel cap max_reisedistanz = 0.24 (parameter.h:43: the
                adaptive dt of dm:1420-1491, at most dt_max = 1e-4);
                Poisson arrivals.
All in FP32 forces with FP64 positions.  Without the travel cap a blob
start at dt = 1e-4 turns unstable after a few thousand steps (a close pair
under the explicit Euler step), and the jobs of up to 1e5 steps then spend
most of their steps in a blown-up state that costs many times a normal step.
"""
import math
import time

import numpy as np

import bdk


def mesh_for(L):
    z = 2 ** int(round(math.log2(L / 0.5)))
    return int(min(512, max(64, z)))


def batched_mix(nrep=256, seed=3):
    """[(params, git, rel)] of the batched mix."""
    rng = np.random.default_rng(seed)
    out = []
    for r in range(nrep):
        N = int(round(math.exp(rng.uniform(math.log(2000), math.log(64000)))))
        N -= N % 2
        L = math.sqrt(N / 0.5)
        p = bdk.make_params(n1=N // 2, n2=N // 2, L=L, Z=mesh_for(L), scheme=int(rng.choice([bdk.CIC, bdk.TSC])),
                            lam=float(rng.uniform(5, 50)), kap=1.0, f2_f1=0.5, T=1.0, dt_max=1e-4,
                            precision=bdk.FP32, seed=int(seed * 100003 + r))
        git, rel = bdk.init_uniform(p, int(seed * 7919 + r), 0.9)
        out.append((p, git, rel))
    return out


def served_jobs(njobs=64, seed=4, max_steps=100000, step_scale=1.0, max_travel=0.24):
    """[(params, git, rel, steps)] of the served job stream; step_scale
    multiplies every job's step count (at least 1 step); max_travel is every
    job's travel cap (inf: fixed dt)."""
    rng = np.random.default_rng(seed)
    out = []
    for j in range(njobs):
        N = int(round(math.exp(rng.uniform(math.log(1000), math.log(256000)))))
        n1 = int(round(N * rng.uniform(0.2, 0.8)))
        n1 = min(max(n1, 1), N - 1)
        rho = rng.uniform(0.3, 0.7)
        L = math.sqrt(N / rho)
        steps = int(round(math.exp(rng.uniform(math.log(1000), math.log(max_steps)))))
        steps = max(1, int(round(steps * step_scale)))
        p = bdk.make_params(n1=n1, n2=N - n1, L=L, Z=mesh_for(L), scheme=int(rng.choice([bdk.NGP, bdk.CIC, bdk.TSC])),
                            lam=float(rng.uniform(5, 100)), kap=1.0, f2_f1=0.5, T=1.0, dt_max=1e-4,
                            max_travel=max_travel, precision=bdk.FP32, seed=int(seed * 100003 + j))
        if rng.random() < 0.5:
            git, rel = bdk.init_uniform(p, int(seed * 7919 + j), 0.9)
        else:
            git, rel, _ = bdk.init_blob(p, int(seed * 7919 + j), L / 6, 0.1, 0.95)
        out.append((p, git, rel, steps))
    return out


def make_sims(mix):
    sims = []
    for p, git, rel in mix:
        s = bdk.RunDynamik(p)
        s.set_positions(git, rel)
        sims.append(s)
    return sims


def time_plain(sims, steps):
    """One simulation at a time (graph-replayed steps each): particle-steps/s."""
    for s in sims:
        s.run_steps(2)
    t0 = time.perf_counter()
    for s in sims:
        s.run_steps(steps)
    dt = time.perf_counter() - t0
    return sum(s.N for s in sims) * steps / dt, dt


def time_batched(sims, steps, tune_reps=5):
    """All replicas as one batch: particle-steps/s, the autotuner's choice
    and per-stage times."""
    b = bdk.Batch(sims)
    ept = b.tune(tune_reps)
    b.run_steps(8)
    b.sync()
    t0 = time.perf_counter()
    b.run_steps(steps)
    b.sync()
    dt = time.perf_counter() - t0
    stages = {k: b.profile(k, 5) for k in ("k_wca", "k_deposit", "k_r2c_rows", "k_col", "k_c2r_rows")}
    b.close()
    return sum(s.N for s in sims) * steps / dt, dt, ept, stages


def serve(jobs, rate, hbm_budget=64 << 30, max_jobs=256, tick=256, seed=5, timeline=None):
    """Poisson arrivals at `rate` jobs/s into a Server; returns per-job
    latencies (arrival to completion, s), aggregate particle-steps/s over the
    run and the server handle's results by job index.  timeline (a list)
    receives (seconds, active, queued, particle-steps so far, jobs still to
    arrive) after every tick."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(jobs)))
    srv = bdk.Server(hbm_budget, max_jobs=max_jobs, tick_steps=tick)
    ids, submitted = {}, {}
    t0 = time.perf_counter()
    nxt, done = 0, 0
    while done < len(jobs):
        now = time.perf_counter() - t0
        while nxt < len(jobs) and arrivals[nxt] <= now:
            p, git, rel, steps = jobs[nxt]
            ids[nxt] = srv.submit(p, git, rel, steps)
            submitted[nxt] = time.perf_counter() - t0
            nxt += 1
        active, queued, done = srv.step()
        if timeline is not None:
            timeline.append((time.perf_counter() - t0, active, queued, srv.stats()[0], len(jobs) - nxt))
        if active == 0 and queued == 0 and nxt < len(jobs):
            time.sleep(max(0.0, min(0.01, arrivals[nxt] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    lat, results = [], {}
    for j in range(len(jobs)):
        git, rel, ta, td = srv.result(ids[j])
        lat.append(submitted[j] - arrivals[j] + td)
        results[j] = (git, rel)
    ps, busy, _ = srv.stats()
    if timeline is not None:
        timeline.append(srv.timing())
    srv.close()
    return np.array(lat), ps / wall, wall, results


def serve_phases(timeline):
    """Split a serve() timeline where the last job has been admitted (no job
    left to arrive and none queued): particle-steps/s before that point (the
    device always had work offered) and in the tail after it (only the
    admitted jobs left, the longest ones last), with the tail's length and
    its share of the particle-steps."""
    timing = timeline[-1] if timeline and isinstance(timeline[-1], dict) else {}
    timeline = [t for t in timeline if not isinstance(t, dict)]
    if not timeline:
        return {}
    cut = next((k for k, (_, _, q, _, left) in enumerate(timeline) if q == 0 and left == 0), len(timeline) - 1)
    t_cut, ps_cut = timeline[cut][0], timeline[cut][3]
    t_end, ps_end = timeline[-1][0], timeline[-1][3]
    return {"offered_phase_s": t_cut, "offered_phase_value": ps_cut / t_cut if t_cut > 0 else None,
            "tail_s": t_end - t_cut, "tail_value": (ps_end - ps_cut) / (t_end - t_cut) if t_end > t_cut else None,
            "tail_share_of_particle_steps": (ps_end - ps_cut) / ps_end if ps_end else None,
            "active_at_tail_start": timeline[cut][1], "server_host_seconds": timing}
