"""GPU tests at the workloads of the benchmark configurations.

- CPU reference configuration (N1 = N2 = 1000, L = 64, CIC 128^2, lambda 8,
  q2/q1 = 0.5, dt 1e-4): the noiseless (kT = 0) trajectory over 1000 steps
  (zeitschritt, dynamik_methoden.cpp:74-163) against the FP64 oracle: FP64
  mode with identical cells and positions within 1e-10, FP32 forces with the
  drift stated and bounded.
- Precision ladder (L = 2000, N1 = N2 = 2,000,000, Z = 4096, NGP/CIC/TSC):
  WCA and capillary forces of every rung (FP64, FP32 forces, FP32 in-cell
  positions, Kahan-compensated FP32 positions) on identical positions
  against the FP64 oracle per component (relative RMS <= 1e-12 for FP64,
  <= 1e-5 for the FP32 rungs), and the trajectory drift of the FP32 rungs
  against FP64 mode over 100 steps at kT = 0 and kT = 1 (TSC).
- Batched ensemble (256 replicas: N log-uniform in [2000, 64000], density
  0.5, mesh the power of two nearest L/0.5 in [64, 512], lambda in [5, 50],
  CIC/TSC): 100 batched steps; every member's invariants, and 16 sampled
  members equal to their runs alone bit for bit.
- Served stream (64 jobs of the served mix with step counts scaled by
  1/100 and the reference's travel cap 0.24, i.e. each with its own
  adaptive dt, plus three fixed-dt FP64 jobs, Poisson arrivals): every FP32
  job equals its run alone bit for bit; the FP64 jobs follow the oracle's
  trajectory (the same Philox noise) with identical cells and positions
  within 1e-10.
"""
import math
import multiprocessing as mp

import numpy as np
import pytest

import bdk
import ensemble
import oracle as O
import pool
from common import params, positions

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def absolute(g, r, B):
    return g * B + r


def min_image(d, L):
    return d - L * np.round(d / L)


# ---------------------------------------------------------------- CPU reference configuration, kT = 0

C0 = dict(n1=1000, n2=1000, L=64.0, Z=128, scheme=bdk.CIC, lam=8.0, kap=1.0, f2_f1=0.5, T=0.0, dt_max=1e-4,
          seed=101)
C0_STEPS = 1000


@pytest.fixture(scope="module")
def c0_reference():
    bp, op = params(precision=bdk.FP64, **C0)
    git, rel = positions(bp, 101, 0.9)
    o = O.Oracle(op)
    o.set_all(git, rel)
    for _ in range(C0_STEPS):
        o.step(1.0)
    og, orl = o.all_positions()
    B = o.B
    o.close()
    return git, rel, og, orl, B


def test_config0_T0_1000_steps_fp64(c0_reference):
    say("case: config-0 kT=0 trajectory, FP64, %d steps" % C0_STEPS)
    git, rel, og, orl, B = c0_reference
    bp, _ = params(precision=bdk.FP64, **C0)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(C0_STEPS)
    gg, gr = s.positions()
    err = float(np.abs(gr - orl).max())
    say("FP64: max position error after %d steps %.3e" % (C0_STEPS, err))
    assert np.array_equal(gg, og)
    assert err < 1e-10


def test_config0_T0_1000_steps_fp32(c0_reference):
    say("case: config-0 kT=0 trajectory, FP32 forces, %d steps" % C0_STEPS)
    git, rel, og, orl, B = c0_reference
    bp, _ = params(precision=bdk.FP32, **C0)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(C0_STEPS)
    gg, gr = s.positions()
    d = min_image(absolute(gg, gr, B) - absolute(og, orl, B), C0["L"])
    a = np.sqrt((d * d).sum(1))
    moved = np.sqrt((min_image(absolute(og, orl, B) - absolute(git, rel, B), C0["L"]) ** 2).sum(1))
    say("FP32: drift after %d steps: max %.3e rms %.3e sigma (mean displacement %.3e)"
        % (C0_STEPS, a.max(), np.sqrt((a * a).mean()), moved.mean()))
    # FP32 forces (relative RMS ~1e-6) over 1000 noiseless steps
    assert a.max() < 1e-5
    assert np.sqrt((a * a).mean()) < 1e-6


# ---------------------------------------------------------------- precision ladder at full size

C2 = dict(n1=2_000_000, n2=2_000_000, L=2000.0, Z=4096, lam=50.0, kap=1.0, f2_f1=0.5, dt_max=1e-4, seed=7)
SCHEMES = {"NGP": bdk.NGP, "CIC": bdk.CIC, "TSC": bdk.TSC}
RUNGS = {"fp64": bdk.FP64, "fp32": bdk.FP32, "fp32_pos32": bdk.FP32_POS32, "fp32_pos32k": bdk.FP32_POS32K}


def c2_params(prec, scheme, T=1.0):
    return params(precision=prec, scheme=scheme, T=T, **C2)


@pytest.fixture(scope="module")
def c2_state():
    """Identical seeded positions of every rung: density 1.0 from a thinned
    0.95 lattice, in-cell positions rounded to FP32 values (which every rung
    stores exactly), and the oracle's forces per scheme (one process each)."""
    bp, _ = c2_params(bdk.FP64, bdk.TSC)
    git, rel, _ = bdk.init_blob(bp, 7, C2["L"], 1.0, 0.95)
    B = bdk.derived(bp)["B"]
    r32 = rel.astype(np.float32)
    top = np.float32(B)
    if float(top) >= B:
        top = np.nextafter(top, np.float32(0.0))
    rel = np.minimum(r32, top).astype(np.float64)
    jobs = []
    for name, sch in SCHEMES.items():
        _, op = c2_params(bdk.FP64, sch)
        kw = {f: getattr(op, f) for f, _ in O.Params._fields_ if f != "pad0"}
        jobs.append((kw, git, rel))
    say("config-2: oracle forces of %d particles, 3 schemes in 3 processes" % len(git))
    with mp.get_context("spawn").Pool(len(jobs)) as workers:
        res = workers.map(pool.forces, jobs)
    return git, rel, dict(zip(SCHEMES, res))


def per_component(a, b):
    out = []
    for k in range(2):
        d = a[:, k] - b[:, k]
        out.append((float(np.sqrt((d * d).sum() / (b[:, k] ** 2).sum())),
                    float(np.abs(d).max() / np.abs(b[:, k]).max())))
    return out


@pytest.mark.timeout(900)  # the first case builds the oracle's forces (three processes, ~40 s each)
@pytest.mark.parametrize("scheme", list(SCHEMES))
def test_config2_forces_every_rung_vs_oracle(c2_state, scheme):
    say("case: config-2 forces, %s, 4M particles, Z=4096" % scheme)
    git, rel, ref = c2_state
    ow, ok = ref[scheme]
    forces = {}
    for rung, prec in RUNGS.items():
        bp, _ = c2_params(prec, SCHEMES[scheme])
        s = bdk.RunDynamik(bp)
        s.set_positions(git, rel)
        w, k = s.berechne_WCA(), s.berechne_kapkraefte()
        s.close()
        forces[rung] = (w, k)
        tol = 1e-12 if prec == bdk.FP64 else 1e-5
        for what, mine, theirs in (("wca", w, ow), ("capillary", k, ok)):
            (rx, mx), (ry, my) = per_component(mine, theirs)
            say("%s %s %s: x rel RMS %.2e (max %.2e), y rel RMS %.2e (max %.2e)" % (scheme, rung, what, rx, mx, ry, my))
            assert rx <= tol and ry <= tol, (scheme, rung, what)
    # the FP32 rungs differ only in how they store positions: on identical
    # (FP32-representable) positions their forces are the same
    for rung in ("fp32_pos32", "fp32_pos32k"):
        assert np.array_equal(forces[rung][0], forces["fp32"][0])
        assert np.array_equal(forces[rung][1], forces["fp32"][1])


@pytest.mark.timeout(900)
@pytest.mark.parametrize("T", [0.0, 1.0])
def test_config2_rung_drift_vs_fp64(c2_state, T):
    say("case: config-2 drift of the FP32 rungs, TSC, kT=%g, 100 steps" % T)
    git, rel, _ = c2_state
    steps = 100
    out = {}
    for rung, prec in RUNGS.items():
        bp, _ = c2_params(prec, bdk.TSC, T=T)
        s = bdk.RunDynamik(bp)
        s.set_positions(git, rel)
        s.run_steps(steps)
        g, r = s.positions()
        out[rung] = absolute(g, r, s.B)
        s.close()
    for rung in ("fp32", "fp32_pos32", "fp32_pos32k"):
        d = min_image(out[rung] - out["fp64"], C2["L"])
        a = np.sqrt((d * d).sum(1))
        say("kT=%g %s: drift after %d steps max %.3e rms %.3e" % (T, rung, steps, a.max(), np.sqrt((a * a).mean())))
        assert a.max() < 1e-3 and np.sqrt((a * a).mean()) < 1e-5


# ---------------------------------------------------------------- batched ensemble, 256 replicas

def test_config3_batched_256_replicas():
    say("case: config-3 batch of 256 replicas, 100 steps")
    mix = ensemble.batched_mix(256, seed=3)
    members = ensemble.make_sims(mix)
    b = bdk.Batch(members)
    steps = 100
    b.run_steps(steps)
    b.sync()
    assert list(b.status()) == [bdk.OK] * 256
    for (p, git0, rel0), s in zip(mix, members):
        assert s.step_counter() == steps
        g, r = s.positions()
        d = bdk.derived(p)
        assert np.all((g >= 0) & (g < d["nc"]))
        assert np.all(np.isfinite(r)) and r.min() >= 0.0 and r.max() <= d["B"]
        start, ids = s.cells()
        N = p.n1 + p.n2
        assert start[-1] == N and np.array_equal(np.sort(ids), np.arange(N))
        # membership: particle id k sits in the cell of its position
        cell_of = np.repeat(np.arange(len(start) - 1), np.diff(start))
        gid = np.empty(N, np.int64)
        gid[ids] = cell_of
        assert np.array_equal(gid, g[:, 0].astype(np.int64) * d["nc"] + g[:, 1])
    b.close()
    rng = np.random.default_rng(17)
    sample = sorted(rng.choice(256, 16, replace=False))
    for i in sample:
        p, git0, rel0 = mix[i]
        solo = bdk.RunDynamik(p)
        solo.set_positions(git0, rel0)
        solo.run_steps(steps)
        ga, ra = solo.positions()
        gm, rm = members[i].positions()
        assert np.array_equal(ga, gm) and np.array_equal(ra, rm), "member %d" % i
        solo.close()
    say("batch: 16 sampled members equal their runs alone")


# ---------------------------------------------------------------- served stream

FP64_JOBS = [
    # n1, n2, L, Z, scheme, lam, seed, steps
    (700, 500, 50.0, 64, bdk.TSC, 12.0, 31, 150),
    (400, 800, 48.0, 128, bdk.NGP, 30.0, 32, 90),
    (1000, 1000, 64.0, 128, bdk.CIC, 8.0, 33, 200),
]


def test_config4_served_stream():
    say("case: config-4 served stream, 64 FP32 jobs + 3 FP64 jobs")
    jobs = ensemble.served_jobs(64, seed=4, step_scale=0.01)
    fp64 = []
    for n1, n2, L, Z, scheme, lam, seed, steps in FP64_JOBS:
        bp, op = params(n1=n1, n2=n2, L=L, Z=Z, scheme=scheme, lam=lam, precision=bdk.FP64, seed=seed)
        git, rel = positions(bp, seed, 0.9)
        fp64.append((op, len(jobs)))
        jobs.append((bp, git, rel, steps))
    order = np.random.default_rng(9).permutation(len(jobs))  # FP64 jobs arrive among the others
    stream = [jobs[i] for i in order]
    lat, rate, wall, results = ensemble.serve(stream, rate=40.0, hbm_budget=48 << 30, max_jobs=48, tick=64)
    say("served %d jobs in %.2f s: p50 %.3f s, p99 %.3f s, %.3e particle-steps/s"
        % (len(stream), wall, np.percentile(lat, 50), np.percentile(lat, 99), rate))
    res = {int(order[k]): results[k] for k in range(len(stream))}
    for j in range(64):
        p, git0, rel0, steps = jobs[j]
        solo = bdk.RunDynamik(p)
        solo.set_positions(git0, rel0)
        solo.run_steps(steps)
        ga, ra = solo.positions()
        solo.close()
        gs, rs = res[j]
        assert np.array_equal(ga, gs) and np.array_equal(ra, rs), "served job %d" % j
    for op, j in fp64:
        p, git0, rel0, steps = jobs[j]
        o = O.Oracle(op)
        o.set_all(git0, rel0)
        for _ in range(steps):
            o.step(1.0)
        og, orl = o.all_positions()
        o.close()
        gs, rs = res[j]
        err = float(np.abs(rs - orl).max())
        say("served FP64 job %d (%d steps): max position error vs oracle %.3e" % (j, steps, err))
        assert np.array_equal(gs, og) and err < 1e-10
