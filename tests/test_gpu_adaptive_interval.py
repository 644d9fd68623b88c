"""Adaptive dt from the step graphs: bdk_run_interval replays the steps that
are certain to be taken (the clock is more than k + 2 dt_max short of obs_dt)
and adds their logged dt's in long double; bdk_run_steps replays all of
them.  Both must take the same steps as the reference's loop of
zeitschritt calls (dm:444-455: while t < obs_dt, t += dt in long double):
same step count, same clock (bdk_run_interval_clock returns it exactly),
positions bit for bit.  A small travel cap
makes dt vary from step to step."""
import numpy as np
import pytest

import bdk
from common import params, positions

pytestmark = pytest.mark.gpu

SPEC = dict(n1=1500, n2=1000, L=70.0, Z=64, scheme=bdk.CIC, lam=20.0, T=1.0, dt_max=1e-4, seed=31)


@pytest.mark.parametrize("prec,travel,obs_dt", [(bdk.FP32, 0.02, 0.05), (bdk.FP64, 0.01, 0.03),
                                                (bdk.FP32, 0.24, 0.1)])
def test_interval_from_graphs_equals_step_loop(prec, travel, obs_dt):
    print("case: adaptive interval prec=%d travel=%g obs_dt=%g" % (prec, travel, obs_dt), flush=True)
    bp, _ = params(precision=prec, max_travel=travel, **SPEC)
    git, rel = positions(bp, 31, 0.9)
    a, b = bdk.RunDynamik(bp), bdk.RunDynamik(bp)
    for s in (a, b):
        s.set_positions(git, rel)
    dts = []
    for nr in range(2):
        if nr == 0:
            n_a, t_a = a.zeitschritte_bis_obs(obs_dt), None
        else:
            n_a, t_a = a.run_interval_clock(obs_dt)  # the clock, exactly
        t = np.longdouble(0.0)
        n_b = 0
        while t < np.longdouble(obs_dt):
            dt = b.zeitschritt(1.0)
            dts.append(dt)
            t += np.longdouble(dt)
            n_b += 1
        assert n_a == n_b, (nr, n_a, n_b)
        assert t_a is None or t_a == t
        ga, ra = a.positions()
        gb, rb = b.positions()
        assert np.array_equal(ga, gb) and np.array_equal(ra, rb)
    if travel < 0.1:
        assert min(dts) < 0.9 * SPEC["dt_max"], "dt did not vary"
    # bdk_run_steps: same state and the same long-double sum of the dt's
    t_a = a.run_steps(300)
    t = np.longdouble(0.0)
    for _ in range(300):
        t += np.longdouble(b.zeitschritt(1.0))
    assert t_a == float(t)
    ga, ra = a.positions()
    gb, rb = b.positions()
    assert np.array_equal(ga, gb) and np.array_equal(ra, rb)
    a.close()
    b.close()
