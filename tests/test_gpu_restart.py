"""Device snapshot (bdk_save_state / bdk_load_state): a run saved after k
steps and resumed in a new handle continues bit for bit like the run that
never stopped (noise counter, list order and positions restored), in FP32
and FP64 modes; a state file of other parameters is refused."""
import numpy as np
import pytest

import bdk
from common import params, positions

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [bdk.FP64, bdk.FP32])
def test_resume_is_bit_exact(tmp_path, prec):
    print("case: resume prec=%d" % prec, flush=True)
    bp, _ = params(n1=1200, n2=800, L=60.0, Z=64, scheme=bdk.TSC, precision=prec, seed=51)
    git, rel = positions(bp, 51)
    a = bdk.RunDynamik(bp)
    a.set_positions(git, rel)
    a.run_steps(37)
    path = str(tmp_path / "state.bin")
    a.save_state(path)
    a.run_steps(29)
    b = bdk.RunDynamik(bp)
    b.load_state(path)
    assert b.step_counter() == 37
    b.run_steps(29)
    ga, ra = a.positions()
    gb, rb = b.positions()
    assert np.array_equal(ga, gb) and np.array_equal(ra, rb)
    assert b.step_counter() == a.step_counter() == 66
    # other parameters: refused
    bp2, _ = params(n1=1200, n2=800, L=60.0, Z=64, scheme=bdk.CIC, precision=prec, seed=51)
    c = bdk.RunDynamik(bp2)
    with pytest.raises(bdk.BdkError):
        c.load_state(path)


def test_startpos_round_trip(tmp_path):
    """positions written in the startpos format (6 significant digits as the
    reference's %g) read back to within float rounding of the printed values."""
    bp, _ = params(n1=300, n2=200, L=40.0, Z=64, precision=bdk.FP64, seed=52)
    git, rel = positions(bp, 52)
    B = bdk.derived(bp)["B"]
    x = git * B + rel
    f = tmp_path / "startpos.txt"
    lines = ["# Format: nr typ x y", ""]
    for k in range(500):
        sp = 1 if k < 300 else 2
        lines.append("%d \t %d \t %.7g \t %.7g" % (k if sp == 1 else k - 300, sp, x[k, 0], x[k, 1]))
    f.write_text("\n".join(lines) + "\n")
    g2, r2 = bdk.read_startpos(bp, str(f))
    assert np.abs((g2 * B + r2) - x).max() < 1e-4
    s = bdk.RunDynamik(bp)
    s.set_positions(g2, r2)
    s.run_steps(5)
