"""Per-kernel timing (bdk_profile_kernel) leaves the simulation as it was:
after timing every step kernel, including the update that moves particles
between cells, the run continues exactly like a twin that was never timed."""
import numpy as np
import pytest

import bdk
from common import params, positions

pytestmark = pytest.mark.gpu

KERNELS = ("k_wca", "k_deposit", "k_r2c_rows", "k_col", "k_c2r_rows", "k_gather", "k_update")


@pytest.mark.parametrize("prec", [bdk.FP32, bdk.FP64])
def test_profiled_run_equals_untimed_twin(prec):
    print("case: profile then step, precision %d" % prec, flush=True)
    bp, _ = params(n1=3000, n2=2000, L=100.0, Z=128, scheme=bdk.TSC, lam=20.0, precision=prec, seed=5)
    git, rel = positions(bp, 5, 0.9)
    sims = []
    for _ in range(2):
        s = bdk.RunDynamik(bp)
        s.set_positions(git, rel)
        s.run_steps(3)
        sims.append(s)
    for k in KERNELS:
        assert sims[0].profile_kernel(k, 4) > 0.0
    for s in sims:
        s.run_steps(25)
    ga, ra = sims[0].positions()
    gb, rb = sims[1].positions()
    assert np.array_equal(ga, gb) and np.array_equal(ra, rb)
    for s in sims:
        s.close()
