"""The FP32-position rungs of the precision ladder (BDK_FP32_POS32 and
BDK_FP32_POS32K): FP32 forces with the in-cell positions stored in FP32,
plain or with Kahan-compensated increments.

Checked: every stored in-cell position is an FP32 value; the trajectory
stays within a small distance of the same run with FP64 positions (same
noise); the compensated rung keeps that distance no larger than the plain
one at kT = 0; a batch member follows its trajectory alone bit for bit; a
saved state (with the compensation) resumes bit for bit."""
import os
import tempfile

import numpy as np
import pytest

import bdk
from common import params, positions

pytestmark = pytest.mark.gpu


def run(prec, T, steps, git, rel, **kw):
    bp, _ = params(precision=prec, T=T, **kw)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(steps)
    g, r = s.positions()
    B = s.B
    s.close()
    return g * B + r, r


def dist(a, b, L):
    d = a - b
    d -= L * np.round(d / L)
    return np.sqrt((d * d).sum(1))


SPEC = dict(n1=1500, n2=1500, L=80.0, Z=128, scheme=bdk.TSC, lam=20.0, dt_max=1e-4, seed=17)


@pytest.mark.parametrize("T", [0.0, 1.0])
def test_fp32_position_rungs_track_fp64_positions(T):
    print("case: position rungs T=%g" % T, flush=True)
    bp, _ = params(**SPEC)
    git, rel = positions(bp, 17, 0.9)
    ref, _ = run(bdk.FP32, T, 300, git, rel, **SPEC)
    out = {}
    for prec in (bdk.FP32_POS32, bdk.FP32_POS32K):
        x, r = run(prec, T, 300, git, rel, **SPEC)
        assert np.array_equal(r.astype(np.float32).astype(np.float64), r), "stored rel not FP32"
        d = dist(x, ref, SPEC["L"])
        out[prec] = (float(np.sqrt((d * d).mean())), float(d.max()))
        print("precision %d: drift rms %.3e max %.3e" % (prec, *out[prec]), flush=True)
        assert out[prec][1] < 1e-3
    if T == 0.0:
        assert out[bdk.FP32_POS32K][0] <= out[bdk.FP32_POS32][0]


def test_kahan_rung_batch_and_resume_bit_exact():
    print("case: POS32K batch + resume", flush=True)
    sp = dict(SPEC, n1=600, n2=400, L=50.0, Z=64)
    bp, _ = params(precision=bdk.FP32_POS32K, T=1.0, **sp)
    git, rel = positions(bp, 5, 0.9)
    alone, member = bdk.RunDynamik(bp), bdk.RunDynamik(bp)
    for s in (alone, member):
        s.set_positions(git, rel)
    b = bdk.Batch([member])
    b.run_steps(40)
    b.sync()
    alone.run_steps(40)
    ga, ra = alone.positions()
    gm, rm = member.positions()
    assert np.array_equal(ga, gm) and np.array_equal(ra, rm)
    b.close()
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "state.bin")
        alone.save_state(path)
        other = bdk.RunDynamik(bp)
        other.load_state(path)
        alone.run_steps(25)
        other.run_steps(25)
        g1, r1 = alone.positions()
        g2, r2 = other.positions()
        assert np.array_equal(g1, g2) and np.array_equal(r1, r2)
        other.close()
    alone.close()
    member.close()
