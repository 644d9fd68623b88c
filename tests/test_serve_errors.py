"""Error paths of the job server that need no GPU: jobs on a device ordinal
that does not exist cannot be admitted (bdk_create fails); each such job is
finished as failed with its error code, the queue drains instead of
blocking at its head, nothing stays charged against the HBM budget, and the
server's tables own every job exactly once (destroy after failures, and an
idle tick, run cleanly)."""
import numpy as np
import pytest

import bdk


def test_failed_admissions_finish_as_failed_jobs():
    p = bdk.make_params(n1=60, n2=40, L=20.0, Z=16, scheme=bdk.CIC, lam=4.0, precision=bdk.FP32, seed=3)
    git, rel = bdk.init_uniform(p, 5, 0.9)
    fp = bdk.footprint(p)
    srv = bdk.Server(hbm_budget=4 * fp, max_jobs=2, tick_steps=10, device=99)
    ids = [srv.submit(p, git, rel, 20 + k) for k in range(5)]
    active, queued, done = srv.step()
    assert (active, queued, done) == (0, 0, 5)
    for jid in ids:
        with pytest.raises(bdk.BdkError) as e:
            srv.result(jid)
        assert e.value.code in (bdk.ERR_HIP, bdk.ERR_ARG)
        assert "admit" in str(e.value)
    ps, busy, used = srv.stats()
    assert ps == 0 and used == 0
    assert srv.step() == (0, 0, 5)
    # unknown ids stay an error of their own
    with pytest.raises(bdk.BdkError) as e:
        srv.result(max(ids) + 1)
    assert e.value.code == bdk.ERR_STATE
    srv.close()


def test_submit_validation():
    p = bdk.make_params(n1=60, n2=40, L=20.0, Z=16, precision=bdk.FP32)
    git, rel = bdk.init_uniform(p, 5, 0.9)
    srv = bdk.Server(hbm_budget=bdk.footprint(p) // 2, max_jobs=2, tick_steps=10, device=99)
    with pytest.raises(bdk.BdkError):
        srv.submit(p, git, rel, 10)  # footprint above the budget
    srv.close()
    bad = bdk.make_params(n1=60, n2=40, L=20.0, Z=22, precision=bdk.FP32)  # Z has a factor 11
    srv = bdk.Server(hbm_budget=1 << 34, max_jobs=2, tick_steps=10, device=99)
    with pytest.raises(bdk.BdkError):
        srv.submit(bad, git, rel, 10)
    assert srv.step() == (0, 0, 0)
    srv.close()
