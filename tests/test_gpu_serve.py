"""Job server (bdk_serve_*): jobs of uneven size submitted at different
times, admitted by HBM footprint under a budget that forces queueing, and
advanced together in batches; every job's result equals the same job run
alone (bit-identical positions)."""
import numpy as np
import pytest

import bdk
from common import params, positions

pytestmark = pytest.mark.gpu

JOBS = [
    # n1, n2, L, Z, scheme, lam, seed, steps
    (400, 400, 40.0, 64, bdk.CIC, 5.0, 21, 120),
    (1200, 600, 64.0, 128, bdk.TSC, 30.0, 22, 75),
    (300, 500, 40.0, 64, bdk.NGP, 9.0, 23, 200),
    (900, 900, 60.0, 128, bdk.TSC, 15.0, 24, 33),
    (150, 250, 30.0, 64, bdk.CIC, 4.0, 25, 160),
    (2000, 1000, 80.0, 128, bdk.CIC, 45.0, 26, 90),
]


def job(spec):
    n1, n2, L, Z, scheme, lam, seed, steps = spec
    bp, _ = params(n1=n1, n2=n2, L=L, Z=Z, scheme=scheme, lam=lam, precision=bdk.FP32, seed=seed)
    git, rel = positions(bp, seed, 0.9)
    return bp, git, rel, steps


def test_served_jobs_equal_alone():
    print("case: served jobs vs alone", flush=True)
    jobs = [job(s) for s in JOBS]
    fps = [bdk.footprint(bp) for bp, _, _, _ in jobs]
    # room for about two jobs at a time: later jobs wait in the queue
    srv = bdk.Server(hbm_budget=2 * max(fps) + 1, max_jobs=3, tick_steps=40)
    ids = []
    for k, (bp, git, rel, steps) in enumerate(jobs):
        ids.append(srv.submit(bp, git, rel, steps))
        if k % 2 == 1:
            srv.step()  # jobs arrive while others run
    for _ in range(1000):
        active, queued, done = srv.step()
        if done == len(jobs):
            break
    assert done == len(jobs)
    ps, busy, used = srv.stats()
    assert used == 0 and ps == sum((bp.n1 + bp.n2) * st for bp, _, _, st in jobs)
    for jid, (bp, git, rel, steps) in zip(ids, jobs):
        g_srv, r_srv, t_admit, t_done = srv.result(jid)
        assert 0.0 <= t_admit <= t_done
        s = bdk.RunDynamik(bp)
        s.set_positions(git, rel)
        s.run_steps(steps)
        g_alone, r_alone = s.positions()
        assert np.array_equal(g_srv, g_alone) and np.array_equal(r_srv, r_alone), "job %d" % jid
    srv.close()


def test_server_rejects_oversized_and_uploaded_noise_jobs():
    bp, git, rel, steps = job(JOBS[0])
    srv = bdk.Server(hbm_budget=bdk.footprint(bp) // 2, max_jobs=2, tick_steps=10)
    with pytest.raises(bdk.BdkError):
        srv.submit(bp, git, rel, steps)
    srv.close()
    bpu, _ = params(n1=400, n2=400, L=40.0, Z=64, noise_mode=bdk.NOISE_UPLOAD, precision=bdk.FP32)
    srv = bdk.Server(hbm_budget=1 << 34, max_jobs=2, tick_steps=10)
    with pytest.raises(bdk.BdkError):
        srv.submit(bpu, git, rel, 10)
    srv.close()


def test_served_adaptive_jobs():
    """Jobs with the reference's adaptive dt (max_reisedistanz, travel caps
    0.24 and 0.05; dt_max 2e-3 so that the cap binds) served beside
    fixed-dt jobs (dt 1e-4), FP32 and FP64 together (one batch per precision): every
    job equals its run alone, and the FP64 adaptive job follows the oracle's
    trajectory (the same Philox draws), cells identical and positions within
    1e-10."""
    import oracle as O
    print("case: served adaptive-dt jobs", flush=True)
    specs = [  # n1, n2, L, Z, scheme, lam, seed, steps, cap, precision
        (400, 400, 40.0, 64, bdk.CIC, 5.0, 41, 90, 0.24, bdk.FP32),
        (600, 300, 48.0, 64, bdk.TSC, 9.0, 42, 60, None, bdk.FP32),
        (300, 500, 40.0, 64, bdk.TSC, 12.0, 43, 120, 0.05, bdk.FP32),
        (500, 500, 45.0, 64, bdk.CIC, 8.0, 44, 80, 0.05, bdk.FP64),
        (700, 200, 50.0, 128, bdk.NGP, 20.0, 45, 40, None, bdk.FP64),
    ]
    jobs, oracles = [], {}
    for k, (n1, n2, L, Z, scheme, lam, seed, steps, cap, prec) in enumerate(specs):
        kw = dict(n1=n1, n2=n2, L=L, Z=Z, scheme=scheme, lam=lam, precision=prec, seed=seed, T=2.0,
                  dt_max=1e-4 if cap is None else 2e-3)
        if cap is not None:
            kw["max_travel"] = cap
        bp, op = params(**kw)
        git, rel = positions(bp, seed, 0.9)
        jobs.append((bp, git, rel, steps))
        if prec == bdk.FP64 and cap is not None:
            oracles[k] = op
    srv = bdk.Server(hbm_budget=1 << 34, max_jobs=3, tick_steps=25)
    ids = []
    for k, j in enumerate(jobs):
        ids.append(srv.submit(*j))
        if k % 2:
            srv.step()
    for _ in range(200):
        if srv.step()[2] == len(jobs):
            break
    for k, (jid, (bp, git, rel, steps)) in enumerate(zip(ids, jobs)):
        g_srv, r_srv, _, _ = srv.result(jid)
        s = bdk.RunDynamik(bp)
        s.set_positions(git, rel)
        s.run_steps(steps)
        g_alone, r_alone = s.positions()
        assert np.array_equal(g_srv, g_alone) and np.array_equal(r_srv, r_alone), "job %d" % k
        if k in oracles:
            o = O.Oracle(oracles[k])
            o.set_all(git, rel)
            for _ in range(steps):
                o.step(1.0)
            og, orl = o.all_positions()
            assert np.array_equal(g_srv, og) and np.abs(r_srv - orl).max() < 1e-10, "job %d vs oracle" % k
    srv.close()


def overlapping(bp, git, rel, i=0, j=1, gap=1e-3):
    """Particle j moved to gap from particle i along x (same y): with an
    unbounded WCA cap the pair force overflows, and the update holds such
    particles in place and raises the run-away flag."""
    d = bdk.derived(bp)
    B, nc = d["B"], d["nc"]
    git, rel = git.copy(), rel.copy()
    x = (git[i] * B + rel[i]) + np.array([gap, 0.0])
    x %= bp.L
    git[j] = np.floor(x / B).astype(np.int32) % nc
    rel[j] = x - git[j] * B
    return git, rel


def test_runaway_flag_cleared_by_new_positions():
    print("case: run-away flag and recovery", flush=True)
    bp, _ = params(n1=300, n2=200, L=30.0, Z=32, precision=bdk.FP32, max_pair_wca=1e300, seed=8)
    git, rel = positions(bp, 8, 0.9)
    s = bdk.RunDynamik(bp)
    s.set_positions(*overlapping(bp, git, rel))
    with pytest.raises(bdk.BdkError) as e:
        s.run_steps(3)
    assert e.value.code == bdk.ERR_NONFINITE
    with pytest.raises(bdk.BdkError) as e:
        s.positions()
    assert e.value.code == bdk.ERR_NONFINITE
    s.set_positions(git, rel)  # a new state starts without the flag
    s.run_steps(3)
    g1, r1 = s.positions()
    # the same new state on a handle at the same step counter (the Philox
    # draws are keyed by it): the same trajectory
    t = bdk.RunDynamik(bp)
    t.set_positions(git, rel)
    t.run_steps(3)
    t.set_positions(git, rel)
    t.run_steps(3)
    g2, r2 = t.positions()
    assert s.step_counter() == t.step_counter() == 6
    assert np.array_equal(g1, g2) and np.array_equal(r1, r2)


def test_runaway_job_fails_alone_others_finish():
    print("case: served run-away job", flush=True)
    jobs = [job(JOBS[0]), job(JOBS[4]), job(JOBS[2])]
    bp, git, rel, steps = jobs[2]
    bad = bdk.make_params(**{f: getattr(bp, f) for f, _ in bdk.Params._fields_})
    bad.max_pair_wca = 1e300
    jobs[2] = (bad,) + overlapping(bad, git, rel) + (steps,)
    srv = bdk.Server(hbm_budget=1 << 34, max_jobs=3, tick_steps=25)
    ids = [srv.submit(*j) for j in jobs]
    for _ in range(100):
        active, queued, done = srv.step()
        if done == 3:
            break
    assert done == 3
    with pytest.raises(bdk.BdkError) as e:
        srv.result(ids[2])
    assert e.value.code == bdk.ERR_NONFINITE
    for jid, (p, g0, r0, st) in zip(ids[:2], jobs[:2]):
        g_srv, r_srv, _, _ = srv.result(jid)
        s = bdk.RunDynamik(p)
        s.set_positions(g0, r0)
        s.run_steps(st)
        g_alone, r_alone = s.positions()
        assert np.array_equal(g_srv, g_alone) and np.array_equal(r_srv, r_alone), "job %d" % jid
    assert srv.stats()[2] == 0
    srv.close()
