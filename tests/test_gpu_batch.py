"""Batched simulations (bdk_batch_*): each member of a batch of uneven
replicas (different N, L, mesh size, scheme, lambda, seed) follows exactly
the trajectory it follows alone: positions bit-identical after the same
number of steps, in FP32 and FP64 modes, before and after the FFT autotuner
changes the kernels' tiling."""
import numpy as np
import pytest

import bdk
from common import params, positions

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


MIX = [
    # n1, n2, L, Z, scheme, lam, seed
    (400, 400, 40.0, 64, bdk.CIC, 5.0, 11),
    (1500, 700, 70.0, 128, bdk.TSC, 20.0, 12),
    (300, 900, 48.0, 64, bdk.NGP, 8.0, 13),
    (2500, 2500, 100.0, 256, bdk.TSC, 50.0, 14),
    (900, 300, 50.0, 128, bdk.CIC, 12.0, 15),
    (100, 60, 16.0, 32, bdk.TSC, 3.0, 16),
]


def make(prec, spec):
    n1, n2, L, Z, scheme, lam, seed = spec
    bp, _ = params(n1=n1, n2=n2, L=L, Z=Z, scheme=scheme, lam=lam, precision=prec, seed=seed, f2_f1=0.5)
    git, rel = positions(bp, seed, 0.9)
    return bp, git, rel


def sims(prec):
    out = []
    for spec in MIX:
        bp, git, rel = make(prec, spec)
        s = bdk.RunDynamik(bp)
        s.set_positions(git, rel)
        out.append(s)
    return out


@pytest.mark.parametrize("prec", [bdk.FP32, bdk.FP64])
def test_batch_members_equal_alone(prec):
    say("case: batch vs alone prec=%d" % prec)
    alone, members = sims(prec), sims(prec)
    b = bdk.Batch(members)
    b.run_steps(37)
    b.sync()
    for s in alone:
        s.run_steps(37)
    for i, (a, m) in enumerate(zip(alone, members)):
        ga, ra = a.positions()
        gm, rm = m.positions()
        assert np.array_equal(ga, gm), "member %d cells differ" % i
        assert np.array_equal(ra, rm), "member %d positions differ (max %.3e)" % (i, np.abs(ra - rm).max())
        assert a.step_counter() == m.step_counter() == 37
    # the autotuner re-tiles the stages; no choice changes a result
    choice = b.tune(3)
    say("tuned", choice)
    assert all(choice[k] in (8, 16, 32) for k in ("ept_r2c", "ept_col", "ept_c2r"))
    assert choice["wca_block"] in (128, 256, 512) and choice["deposit_tile"] in (24, 32, 38, 48, 64, 80)
    b.run_steps(11)
    b.sync()
    for s in alone:
        s.run_steps(11)
    for i, (a, m) in enumerate(zip(alone, members)):
        assert np.array_equal(a.positions()[1], m.positions()[1]), "member %d after tune" % i
    # members stay usable through their own handles
    f_alone, f_member = alone[3].berechne_WCA(), members[3].berechne_WCA()
    assert np.array_equal(f_alone, f_member)
    b.close()


@pytest.mark.parametrize("choice", [
    dict(ept_r2c=32, ept_col=32, ept_c2r=32, wca_block=512, deposit_tile=24, persist=64),
    dict(ept_r2c=16, ept_col=8, ept_c2r=32, wca_block=128, deposit_tile=48, persist=1),
    dict(ept_r2c=8, ept_col=16, ept_c2r=16, wca_block=256, deposit_tile=32, persist=0),
    dict(ept_r2c=8, ept_col=8, ept_c2r=8, wca_block=256, deposit_tile=80, persist=0),
])
def test_batch_tuning_choices_keep_results(choice):
    """Every knob of the autotuner (FFT elements per thread, WCA block size,
    deposit tile, persistent grid down to a single block looping over all
    work items) leaves each member's trajectory bit-identical to its run
    alone, FP32, with one adaptive-dt member."""
    say("case: tuning choice %s" % choice)

    def with_adaptive():
        out = sims(bdk.FP32)
        bp, _ = params(n1=500, n2=300, L=40.0, Z=64, scheme=bdk.TSC, lam=6.0, precision=bdk.FP32, seed=19,
                       T=2.0, dt_max=2e-3, max_travel=0.05)
        git, rel = positions(bp, 19, 0.9)
        s = bdk.RunDynamik(bp)
        s.set_positions(git, rel)
        return out + [s]

    alone, members = with_adaptive(), with_adaptive()
    b = bdk.Batch(members)
    b.set_tuning(**choice)
    b.run_steps(13)
    b.sync()
    for s in alone:
        s.run_steps(13)
    for i, (a, m) in enumerate(zip(alone, members)):
        ga, ra = a.positions()
        gm, rm = m.positions()
        assert np.array_equal(ga, gm) and np.array_equal(ra, rm), "member %d" % i
    with pytest.raises(bdk.BdkError):
        b.set_tuning(wca_block=96)
    b.close()


def test_batch_profile_and_errors():
    say("case: batch profile and errors")
    members = sims(bdk.FP32)
    b = bdk.Batch(members)
    for st in ("k_wca", "k_deposit", "k_r2c_rows", "k_col", "k_c2r_rows", "step"):
        assert b.profile(st, 3) > 0.0
    b.close()
    # mixed precision, uploaded noise and a handle given twice are refused
    bp64, git, rel = make(bdk.FP64, MIX[0])
    s64 = bdk.RunDynamik(bp64)
    s64.set_positions(git, rel)
    with pytest.raises(bdk.BdkError):
        bdk.Batch([members[0], s64])
    bpu, _ = params(n1=400, n2=400, L=40.0, Z=64, noise_mode=bdk.NOISE_UPLOAD, precision=bdk.FP32)
    su = bdk.RunDynamik(bpu)
    su.set_positions(git, rel)
    with pytest.raises(bdk.BdkError):
        bdk.Batch([su])
    with pytest.raises(bdk.BdkError):
        bdk.Batch([members[0], members[0]])


@pytest.mark.parametrize("prec", [bdk.FP32, bdk.FP64])
def test_batch_adaptive_members_equal_alone(prec):
    """Members with the adaptive dt of optimaler_zeitschritt (travel caps
    0.05 and 0.02, each member its own dt per step) beside fixed-dt members:
    each member's trajectory and its own clock (the long-double sum of its
    logged dt's, read back through bdk_run_interval_clock afterwards) equal
    its run alone."""
    say("case: batch with adaptive-dt members prec=%d" % prec)
    caps = [0.05, None, 0.02, None, 0.05, 0.02]

    def make_all():
        out = []
        for spec, cap in zip(MIX, caps):
            n1, n2, L, Z, scheme, lam, seed = spec
            # adaptive members: dt_max 2e-3, so that the travel cap binds;
            # fixed-dt members at the usual 1e-4
            kw = dict(n1=n1, n2=n2, L=L, Z=Z, scheme=scheme, lam=lam, precision=prec, seed=seed, f2_f1=0.5,
                      T=2.0, dt_max=1e-4 if cap is None else 2e-3)
            if cap is not None:
                kw["max_travel"] = cap
            bp, _ = params(**kw)
            git, rel = positions(bp, seed, 0.9)
            s = bdk.RunDynamik(bp)
            s.set_positions(git, rel)
            out.append(s)
        return out

    alone, members = make_all(), make_all()
    b = bdk.Batch(members)
    b.run_steps(29)
    b.sync()
    for s in alone:
        s.run_steps(29)
    for i, (a, m) in enumerate(zip(alone, members)):
        ga, ra = a.positions()
        gm, rm = m.positions()
        assert np.array_equal(ga, gm) and np.array_equal(ra, rm), "member %d (cap %s)" % (i, caps[i])
    b.close()
    # the members go on alone: their next interval (clock included) equals
    for i, (a, m) in enumerate(zip(alone, members)):
        assert a.run_interval_clock(0.02) == m.run_interval_clock(0.02), "member %d" % i
        assert np.array_equal(a.positions()[1], m.positions()[1])
