"""GPU parity on meshes whose size is not a power of two (even Z = 2^a 3^b
5^c 7^d), among them the reference's default FFT_samples = 384
(parameter.h:24), against the FP64 oracle (whose DFT of these lengths is
pinned against numpy in test_oracle.py::test_fft2d_lengths_not_powers_of_two).

Tolerances as in test_gpu_parity.py: FP64 forces and meshes within 1e-12
relative RMS, FP32 within 1e-5; FP64 trajectories keep identical cells and
positions within 1e-10.
"""
import numpy as np
import pytest

import bdk
from common import pair
from helpers import rel_rms

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def say(*a):
    print(*a, flush=True)


def check_forces(g, o, tol):
    fg = g.berechne_kapkraefte()
    fo = o.kap()
    assert rel_rms(fg, fo) < tol, rel_rms(fg, fo)
    assert rel_rms(g.mesh(bdk.MESH_RHO), o.mesh(0)) < tol
    assert rel_rms(g.mesh(bdk.MESH_FX), o.mesh(1)) < tol
    assert rel_rms(g.mesh(bdk.MESH_FY), o.mesh(2)) < tol


# 12 = 4*3 and 18 = 2*9 have no radix-8 pass; 24 = 8*3; 96 = 32*3 (radix 8,
# 4, 3); 144 = 16*9 (two radix-3 passes); 384 = 128*3 (8, 8, 2, 3); 10 = 2*5
# and 14 = 2*7; 50 = 2*25 (two radix-5 passes); 98 = 2*49; 210 = 2*3*5*7
# (every odd radix); 250 = 2*125
@pytest.mark.parametrize("scheme", [bdk.NGP, bdk.CIC, bdk.TSC])
@pytest.mark.parametrize("prec,tol", [(bdk.FP64, TOL64), (bdk.FP32, TOL32)])
@pytest.mark.parametrize("Z,L", [(12, 10.0), (18, 12.0), (24, 16.0), (96, 60.0), (144, 90.0), (384, 200.0),
                                 (10, 8.0), (14, 10.0), (50, 30.0), (98, 60.0), (210, 120.0), (250, 150.0)])
def test_capillary_parity_factor3(scheme, prec, tol, Z, L):
    say("case: kap Z=%d scheme=%d prec=%d" % (Z, scheme, prec))
    n = int(0.2 * L * L)
    g, o, git, rel = pair(n1=n, n2=n, L=L, Z=Z, scheme=scheme, precision=prec, seed=31, lam=6.0, kap=1.3,
                          f2_f1=-0.7)
    check_forces(g, o, tol)


@pytest.mark.parametrize("prec", [bdk.FP64, bdk.FP32])
def test_reference_default_setup(prec):
    """The reference's own parameter.h setup (N1 = N2 = 2000, L = 300, TSC on
    384^2, lambda = 25, kap = 1.111111111, T = 24.3055555, dt_max = 0.005,
    max_reisedistanz = 0.24, parameter.h:9-44): forces against the oracle,
    then 8 adaptive-dt steps (dt = optimaler_zeitschritt, dm:1420-1491) in
    FP64 with cells and positions against the oracle's."""
    say("case: reference default setup prec=%d" % prec)
    g, o, git, rel = pair(n1=2000, n2=2000, L=300.0, Z=384, scheme=bdk.TSC, lam=25.0, kap=1.111111111,
                          T=24.3055555, dt_max=0.005, max_travel=0.24, precision=prec, seed=37)
    check_forces(g, o, TOL64 if prec == bdk.FP64 else TOL32)
    if prec != bdk.FP64:
        return
    for _ in range(8):
        dg = g.zeitschritt(1.0)
        do = o.step(1.0)
        assert dg == pytest.approx(do, rel=1e-10)
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.array_equal(gg, og)
    assert np.abs(gr - orl).max() < 1e-10


def test_factor3_graph_replay_and_batch():
    """A 384^2 simulation: graph-replayed steps equal single steps bit for bit,
    and the same simulation as a batch member (beside a power-of-two member
    and an adaptive-dt member on a 120^2 mesh, so every mesh stage and the
    adaptive gather run once per mesh kind) equals its run alone, as does the
    adaptive member."""
    say("case: 384^2 graph replay and batch")
    kw = dict(n1=3000, n2=3000, L=150.0, Z=384, scheme=bdk.TSC, precision=bdk.FP32, lam=20.0, seed=41)
    a, _, git, rel = pair(**kw)
    b = bdk.RunDynamik(a.p)
    b.set_positions(git, rel)
    a.run_steps(20)
    for _ in range(20):
        b.zeitschritt(1.0)
    ga, ra = a.positions()
    gb, rb = b.positions()
    assert np.array_equal(ga, gb) and np.array_equal(ra, rb)
    c = bdk.RunDynamik(a.p)
    c.set_positions(git, rel)
    d, _, _, _ = pair(n1=800, n2=800, L=50.0, Z=64, scheme=bdk.CIC, precision=bdk.FP32, seed=43)
    # an adaptive-dt member on a 3 * 5 * 2^3 mesh, alone and in the batch
    kw_e = dict(n1=700, n2=500, L=60.0, Z=120, scheme=bdk.TSC, precision=bdk.FP32, seed=47, max_travel=0.05,
                dt_max=5e-3, T=2.0)
    e_alone, _, git_e, rel_e = pair(**kw_e)
    e = bdk.RunDynamik(e_alone.p)
    e.set_positions(git_e, rel_e)
    batch = bdk.Batch([c, d, e])
    batch.run_steps(20)
    batch.sync()
    gc, rc = c.positions()
    assert np.array_equal(gc, ga) and np.array_equal(rc, ra)
    e_alone.run_steps(20)
    ge, re_ = e.positions()
    gw, rw = e_alone.positions()
    assert np.array_equal(ge, gw) and np.array_equal(re_, rw)
    batch.close()
    for s in (a, b, c, d, e, e_alone):
        s.close()


def test_fp64_batch_other_mesh_sizes():
    """FP64 members on meshes of 96^2 (2^5 * 3) and 50^2 (2 * 5^2) beside a
    power-of-two member, the 96^2 one with the reference's travel cap: 15
    batched steps equal each member's steps alone bit for bit, and the
    adaptive member's trajectory follows the oracle's (cells identical,
    positions within 1e-10)."""
    say("case: FP64 batch with meshes 96^2, 50^2 and 64^2")
    kws = [dict(n1=600, n2=400, L=50.0, Z=96, scheme=bdk.TSC, precision=bdk.FP64, seed=51, max_travel=0.05,
                dt_max=5e-3, T=2.0),
           dict(n1=300, n2=300, L=30.0, Z=50, scheme=bdk.CIC, precision=bdk.FP64, seed=53),
           dict(n1=500, n2=500, L=40.0, Z=64, scheme=bdk.NGP, precision=bdk.FP64, seed=55)]
    alone, members, oracles = [], [], []
    for kw in kws:
        a, o, git, rel = pair(**kw)
        m = bdk.RunDynamik(a.p)
        m.set_positions(git, rel)
        alone.append(a)
        members.append(m)
        oracles.append(o)
    batch = bdk.Batch(members)
    batch.run_steps(15)
    batch.sync()
    for a, m in zip(alone, members):
        a.run_steps(15)
        ga, ra = a.positions()
        gm, rm = m.positions()
        assert np.array_equal(ga, gm) and np.array_equal(ra, rm)
    o = oracles[0]
    for _ in range(15):
        o.step(1.0)
    og, orl = o.all_positions()
    gm, rm = members[0].positions()
    assert np.array_equal(gm, og)
    assert np.abs(rm - orl).max() < 1e-10
    batch.close()
    for s in alone + members + oracles:
        s.close()
