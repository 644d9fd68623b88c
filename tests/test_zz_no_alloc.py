"""bdk_pair_correlation_noalloc (no device allocation; a fixed scratch array
per GPU) against bdk_pair_correlation on the same states: two calls in a
row, the second on a smaller system with fewer bins (the array keeps the
first call's counts, so the second must clear what it uses), and its
refusals (bins beyond the limit, state not yet on the device).  The counts
are integers, so the outputs are equal bit for bit."""
import numpy as np
import pytest

import bdk
from common import params, positions

pytestmark = pytest.mark.gpu


def sim(n1, n2, L, seed):
    bp, _ = params(n1=n1, n2=n2, L=L, Z=64, precision=bdk.FP32, seed=seed)
    git, rel = positions(bp, seed, 0.9)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(5)  # the state is on the device
    return s


def test_noalloc_pair_correlation_equals_allocating_entry():
    print("case: no-allocation pair correlation", flush=True)
    big, small = sim(1500, 1200, 60.0, 3), sim(300, 200, 30.0, 4)
    for s, nbins, dr in ((big, 150, 0.2), (small, 40, 0.25)):
        want = s.pair_correlation(nbins, dr)
        got = s.pair_correlation_noalloc(nbins, dr)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        assert want[0].sum() > 0 and want[1].sum() > 0
    # repeated calls on the same state give the same counts
    again = big.pair_correlation_noalloc(150, 0.2)
    for a, b in zip(again, big.pair_correlation(150, 0.2)):
        assert np.array_equal(a, b)


def test_noalloc_pair_correlation_refusals():
    print("case: no-allocation refusals", flush=True)
    s = sim(300, 200, 30.0, 5)
    with pytest.raises(bdk.BdkError) as e:
        s.pair_correlation_noalloc(4097, 0.01)
    assert e.value.code == bdk.ERR_ARG
    bp, _ = params(n1=300, n2=200, L=30.0, Z=64, precision=bdk.FP32, seed=6)
    git, rel = positions(bp, 6, 0.9)
    fresh = bdk.RunDynamik(bp)
    fresh.set_positions(git, rel)
    with pytest.raises(bdk.BdkError) as e:
        fresh.pair_correlation_noalloc(40, 0.25)
    assert e.value.code == bdk.ERR_STATE
    fresh.berechne_WCA()  # now on the device
    for a, b in zip(fresh.pair_correlation_noalloc(40, 0.25), fresh.pair_correlation(40, 0.25)):
        assert np.array_equal(a, b)
