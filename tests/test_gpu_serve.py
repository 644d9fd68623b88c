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


def test_server_rejects_oversized_and_adaptive_jobs():
    bp, git, rel, steps = job(JOBS[0])
    srv = bdk.Server(hbm_budget=bdk.footprint(bp) // 2, max_jobs=2, tick_steps=10)
    with pytest.raises(bdk.BdkError):
        srv.submit(bp, git, rel, steps)
    srv.close()
    bpa, _ = params(n1=400, n2=400, L=40.0, Z=64, max_travel=0.1, precision=bdk.FP32)
    srv = bdk.Server(hbm_budget=1 << 34, max_jobs=2, tick_steps=10)
    with pytest.raises(bdk.BdkError):
        srv.submit(bpa, git, rel, 10)
    srv.close()
