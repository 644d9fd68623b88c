"""GPU parity of the HIP path against the FP64 CPU oracle, through the C-ABI.

Tolerances (written here, per the force-precision modes):
  FP64 mode: forces within 1e-12 relative RMS; positions within 1e-12 (abs)
             over short trajectories; cell membership bit-exact.
  FP32 mode: forces within 1e-5 relative RMS (FP32 forces, FP64 positions).
"""
import math
import os

import numpy as np
import pytest

import bdk
import oracle as O
from common import pair, params, positions
from helpers import rel_rms, to_git_rel

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def say(*a):
    print(*a, flush=True)


def membership_ok(g, o, git):
    start, ids = g.cells()
    nc = g.nc
    # per-cell content from the device equals the brute-force git binning
    cell = git[:, 0] * nc + git[:, 1]
    want = np.argsort(cell, kind="stable")
    assert np.array_equal(np.diff(start), np.bincount(cell, minlength=nc * nc))
    assert np.array_equal(ids, want)
    # and the oracle's extended lists (dm:459-526) are the 3x3 unions of it
    n1 = g.p.n1
    rng = np.random.default_rng(0)
    for c in rng.choice(nc * nc, size=min(200, nc * nc), replace=False):
        cx, cy = divmod(int(c), nc)
        union = []
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                cc = ((cx + a) % nc) * nc + (cy + b) % nc
                union += list(ids[start[cc]:start[cc + 1]])
        union = np.array(sorted(union), dtype=np.int64)
        s1 = union[union < n1]
        s2 = union[union >= n1] - n1
        assert sorted(o.neighbour_list(1, cx, cy)) == list(s1)
        assert sorted(o.neighbour_list(2, cx, cy)) == list(s2)
    return True


def test_cell_membership_bit_exact():
    say("case: membership RSA N=2000 L=64")
    g, o, git, rel = pair(n1=1000, n2=1000, L=64.0, Z=128, seed=3)
    assert membership_ok(g, o, git)
    say("case: membership MT init (P7 quirk)")
    bp, op = params(n1=700, n2=600, L=64.0, Z=128, noise_mode=bdk.NOISE_UPLOAD)
    mt = bdk.MT(42)
    git, rel = mt.init_positions(bp)
    g = bdk.RunDynamik(bp)
    g.set_positions(git, rel)
    op.noise_mode = 3
    o = O.Oracle(op)
    o.set_all(git, rel)
    assert membership_ok(g, o, git)


@pytest.mark.parametrize("prec,tol", [(bdk.FP64, TOL64), (bdk.FP32, TOL32)])
@pytest.mark.parametrize("ratios", [(1.0, 1.0, 1.0, 1.0), (0.8, 0.9, 1.3, 0.7)])
def test_wca_parity(prec, tol, ratios):
    s22, s12, e22, e12 = ratios
    say("case: wca prec=%d ratios=%s" % (prec, ratios))
    g, o, git, rel = pair(n1=1500, n2=1500, L=70.0, Z=64, seed=5, dmin=0.8, precision=prec,
                          sigma11_22=s22, sigma11_12=s12, eps22_11=e22, eps12_11=e12)
    fg = g.berechne_WCA()
    fo = o.wca()
    assert rel_rms(fg, fo) < tol, rel_rms(fg, fo)


@pytest.mark.parametrize("prec,tol", [(bdk.FP64, TOL64), (bdk.FP32, TOL32)])
@pytest.mark.parametrize("L", [5.0, 8.0, 12.0])
def test_wca_parity_small_boxes(prec, tol, L):
    """3 to 10 list cells per side: wrapped neighbourhoods and images."""
    n = int(0.3 * L * L)
    g, o, git, rel = pair(n1=n, n2=n, L=L, Z=8, seed=7, dmin=0.85, precision=prec)
    assert rel_rms(g.berechne_WCA(), o.wca()) < tol


def test_wca_parity_p7_initial_state():
    """Start positions of init_zufallspos with rel up to 2B (dm:1655)."""
    bp, op = params(n1=1000, n2=1000, L=64.0, Z=128, noise_mode=bdk.NOISE_UPLOAD)
    git, rel = bdk.MT(7).init_positions(bp)
    op.noise_mode = 3
    o = O.Oracle(op)
    o.set_all(git, rel)
    g = bdk.RunDynamik(bp)
    g.set_positions(git, rel)
    say("case: wca P7")
    assert rel_rms(g.berechne_WCA(), o.wca()) < TOL64
    say("case: kap P7")
    assert rel_rms(g.berechne_kapkraefte(), o.kap()) < TOL64


@pytest.mark.parametrize("scheme", [bdk.NGP, bdk.CIC, bdk.TSC])
@pytest.mark.parametrize("prec,tol", [(bdk.FP64, TOL64), (bdk.FP32, TOL32)])
# 1024: the smallest mesh whose FP32 column and c2r passes are radix 16
# (passes 16, 16, 4; 2048 and 4096 are covered at their full workloads)
@pytest.mark.parametrize("Z,L", [(8, 8.0), (16, 12.0), (64, 40.0), (128, 64.0), (256, 90.0), (1024, 300.0)])
def test_capillary_parity(scheme, prec, tol, Z, L):
    say("case: kap scheme=%d prec=%d Z=%d" % (scheme, prec, Z))
    n = int(0.25 * L * L)
    g, o, git, rel = pair(n1=n, n2=n, L=L, Z=Z, scheme=scheme, precision=prec, seed=11, lam=6.0, kap=1.3,
                          f2_f1=-0.7)
    fg = g.berechne_kapkraefte()
    fo = o.kap()
    assert rel_rms(fg, fo) < tol, rel_rms(fg, fo)
    assert rel_rms(g.mesh(bdk.MESH_RHO), o.mesh(0)) < tol
    assert rel_rms(g.mesh(bdk.MESH_FX), o.mesh(1)) < tol
    assert rel_rms(g.mesh(bdk.MESH_FY), o.mesh(2)) < tol


def test_capillary_ngp_fix_and_continuous_green():
    g, o, git, rel = pair(n1=400, n2=400, L=40.0, Z=64, scheme=bdk.NGP, ngp_fix=1)
    assert rel_rms(g.berechne_kapkraefte(), o.kap()) < TOL64
    # continuous symbol: i k/(k^2 + lambda^-2); against numpy on the same mesh
    bp, _ = params(n1=400, n2=400, L=40.0, Z=64, scheme=bdk.TSC, green_mode=bdk.GREEN_CONTINUOUS, lam=5.0)
    g2 = bdk.RunDynamik(bp)
    g2.set_positions(git, rel)
    g2.berechne_kapkraefte()
    rho = g2.mesh(bdk.MESH_RHO)
    Z, L = 64, 40.0
    j = np.arange(Z)
    q = 2 * np.pi / L * np.where(j < Z // 2, j, j - Z)
    G = 1.0 / (q[:, None] ** 2 + q[None, :] ** 2 + 1.0 / 25.0)
    qg = np.where(j == Z // 2, 0.0, q)
    rk = np.fft.fft2(rho)
    fx = np.real(np.fft.ifft2(1j * qg[:, None] * G * rk))
    assert rel_rms(g2.mesh(bdk.MESH_FX), fx) < 1e-12


def test_philox_noise_matches_oracle():
    bp, op = params(n1=300, n2=200, L=40.0, Z=64, seed=99)
    g = bdk.RunDynamik(bp)
    git, rel = positions(bp, 1)
    g.set_positions(git, rel)
    n = g.berechne_zufallskraefte()
    want = np.array([O.philox_noise(99, i, 0) for i in range(500)])
    assert np.allclose(n, want, rtol=0, atol=1e-14)
    bp2, _ = params(n1=300, n2=200, L=40.0, Z=64, seed=99, noise_mode=bdk.NOISE_GAUSS, noise_var=2.0)
    g2 = bdk.RunDynamik(bp2)
    g2.set_positions(git, rel)
    n2 = g2.berechne_zufallskraefte()
    want2 = np.array([O.philox_noise(99, i, 0, True, 2.0) for i in range(500)])
    assert np.allclose(n2, want2, rtol=0, atol=1e-13)


@pytest.mark.parametrize("scheme", [bdk.CIC, bdk.TSC])
def test_trajectory_T0_fp64(scheme):
    """Noiseless trajectory (kT = 0), 300 steps, FP64 mode."""
    say("case: T0 trajectory scheme=%d" % scheme)
    g, o, git, rel = pair(n1=600, n2=600, L=50.0, Z=64, scheme=scheme, T=0.0, dt_max=1e-4, seed=21)
    for _ in range(300):
        o.step(1.0)
    g.run_steps(300)
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.array_equal(gg, og)
    assert np.abs(gr - orl).max() < 1e-10


def test_trajectory_philox_noise_fp64():
    say("case: philox trajectory")
    g, o, git, rel = pair(n1=500, n2=500, L=50.0, Z=64, scheme=bdk.TSC, seed=22)
    for _ in range(200):
        o.step(1.0)
    g.run_steps(200)
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.array_equal(gg, og)
    assert np.abs(gr - orl).max() < 1e-11


def test_mt_upload_reproduces_reference_stream():
    """Validation mode: the host MT stream (zufall.cpp) uploaded each step
    reproduces the reference's noisy trajectory step for step (FP64), from
    a non-overlapping start configuration."""
    say("case: MT upload")
    bp, op = params(n1=500, n2=500, L=50.0, Z=64, scheme=bdk.CIC, noise_mode=bdk.NOISE_UPLOAD, seed=31337)
    op.noise_mode = 0
    git, rel = positions(bp, 41)
    mt = bdk.MT(31337)
    o = O.Oracle(op)
    o.set_all(git, rel)
    g = bdk.RunDynamik(bp)
    g.set_positions(git, rel)
    worst = 0.0
    for step in range(100):
        g.set_noise(mt.disc_noise(500, 500))
        dt_g = g.zeitschritt(1.0)
        dt_o = o.step(1.0)
        assert dt_g == dt_o
        gg, gr = g.positions()
        og, orl = o.all_positions()
        err = float(np.abs((gg * g.B + gr) - (og * o.B + orl)).max())
        worst = max(worst, err)
        assert np.array_equal(gg, og), "cells differ at step %d (max pos err %.3e)" % (step, err)
        assert err < 1e-11, "step %d: %.3e" % (step, err)
    say("MT upload: worst position error over 100 steps %.3e" % worst)


def test_mt_upload_from_reference_random_start():
    """The reference's own random start (init_zufallspos, overlapping pairs
    and rel up to 2B) for the first steps: identical cells, positions to
    1e-9 (capped WCA forces of overlapping pairs amplify rounding)."""
    bp, op = params(n1=500, n2=500, L=50.0, Z=64, scheme=bdk.TSC, noise_mode=bdk.NOISE_UPLOAD, seed=4242)
    op.noise_mode = 0
    mt = bdk.MT(4242)
    git, rel = mt.init_positions(bp)
    o = O.Oracle(op)
    o.init_zufallspos()
    og, orl = o.all_positions()
    assert np.array_equal(git, og) and np.array_equal(rel, orl)
    g = bdk.RunDynamik(bp)
    g.set_positions(git, rel)
    for step in range(3):
        g.set_noise(mt.disc_noise(500, 500))
        g.zeitschritt(1.0)
        o.step(1.0)
        gg, gr = g.positions()
        og, orl = o.all_positions()
        err = float(np.abs((gg * g.B + gr) - (og * o.B + orl)).max())
        say("reference start, step %d: max pos err %.3e" % (step, err))
        assert np.array_equal(gg, og) and err < 1e-9


def test_adaptive_dt_parity():
    """finite max_reisedistanz: dt = optimaler_zeitschritt (dm:1420-1491)."""
    say("case: adaptive dt")
    g, o, git, rel = pair(n1=400, n2=400, L=40.0, Z=64, max_travel=0.05, dt_max=5e-3, T=2.0, seed=23)
    o.wca(); o.kap(); o.noise()
    assert g.optimaler_zeitschritt() == pytest.approx(o.optimal_dt(), rel=1e-12)
    for _ in range(30):
        dg = g.zeitschritt(1.0)
        do = o.step(1.0)
        assert dg == pytest.approx(do, rel=1e-10)
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.array_equal(gg, og)
    assert np.abs(gr - orl).max() < 1e-10


def test_checkpoint_interval_and_pos_writer(tmp_path):
    say("case: interval + writer")
    g, o, git, rel = pair(n1=200, n2=150, L=30.0, Z=32, seed=24)
    p1, p2 = str(tmp_path / "g1.txt"), str(tmp_path / "g2.txt")
    q1, q2 = str(tmp_path / "o1.txt"), str(tmp_path / "o2.txt")
    g.write_pos(p1, p2, 0, 0.01)
    o.write_pos_header(q1); o.write_pos_header(q2)
    o.append_pos(1, q1, 0, 0.01); o.append_pos(2, q2, 0, 0.01)
    assert open(p1).read() == open(q1).read() and open(p2).read() == open(q2).read()
    # long double clock: 0.01 / 1e-4 takes 100 steps (0.1 / 1e-4 takes 1001)
    assert g.zeitschritte_bis_obs(0.01) == o.steps_until_obs(0.01) == O.steps_for_interval(0.01, 1e-4)
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.abs((gg * g.B + gr) - (og * o.B + orl)).max() < 1e-10


def test_refresh_bit_exact():
    g, o, git, rel = pair(n1=300, n2=300, L=40.0, Z=32, seed=25)
    rng = np.random.default_rng(2)
    rel2 = rel + rng.uniform(-0.6, 0.6, rel.shape)
    o.set_positions(1, git[:300], rel2[:300]); o.set_positions(2, git[300:], rel2[300:])
    o.L.bdo_wca_init(o.h)
    o.refresh()
    g.set_positions(git, rel2)
    g.refresh_erwNachbar()
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.array_equal(gg, og) and np.array_equal(gr, orl)


def test_graph_replay_equals_single_steps_and_determinism():
    say("case: graph replay")
    bp, _ = params(n1=800, n2=800, L=60.0, Z=128, precision=bdk.FP32, scheme=bdk.TSC)
    git, rel = positions(bp, 26)
    a = bdk.RunDynamik(bp); a.set_positions(git, rel)
    b = bdk.RunDynamik(bp); b.set_positions(git, rel)
    c = bdk.RunDynamik(bp); c.set_positions(git, rel)
    a.run_steps(50)
    for _ in range(50):
        b.zeitschritt(1.0)
    c.run_steps(50)
    pa, pb, pc = a.positions(), b.positions(), c.positions()
    assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])
    assert np.array_equal(pa[1], pc[1])


def test_fp32_trajectory_drift_stated():
    """FP32 forces with FP64 positions vs the FP64 oracle over 200 noisy
    steps (same Philox noise): drift stays below 1e-6 sigma."""
    say("case: fp32 drift")
    g, o, git, rel = pair(n1=500, n2=500, L=50.0, Z=64, scheme=bdk.TSC, precision=bdk.FP32, seed=27)
    for _ in range(200):
        o.step(1.0)
    g.run_steps(200)
    drift = np.abs(g.absolute() - o.absolute())
    drift = np.minimum(drift, 50.0 - drift)
    assert drift.max() < 1e-6


def test_error_paths_on_device():
    bp, _ = params(n1=100, n2=100, L=30.0, Z=32)
    g = bdk.RunDynamik(bp)
    with pytest.raises(bdk.BdkError) as e:
        g.berechne_WCA()
    assert e.value.code == bdk.ERR_STATE
    git, rel = positions(bp, 1)
    git_bad = git.copy(); git_bad[5, 0] = 10 ** 6
    g.set_positions(git_bad, rel)
    with pytest.raises(bdk.BdkError) as e:
        g.berechne_WCA()
    assert e.value.code == bdk.ERR_ARG
    rel_bad = rel.copy(); rel_bad[3, 1] = float("nan")
    g.set_positions(git, rel_bad)
    with pytest.raises(bdk.BdkError) as e:
        g.berechne_WCA()
    assert e.value.code == bdk.ERR_NONFINITE
    g.set_positions(git, rel)
    assert g.zeitschritt(-1.0) == 1e10 and g.step_counter() == 0


@pytest.mark.parametrize("prec", [bdk.FP64, bdk.FP32])
def test_single_species_run(prec, tmp_path):
    """N2 = 0 (the reference then opens no pos2 file, au:57-69): forces
    against the oracle, an FP64 trajectory, and the writer."""
    say("case: single species N2=0 prec=%d" % prec)
    g, o, git, rel = pair(n1=800, n2=0, L=40.0, Z=64, scheme=bdk.TSC, seed=81, precision=prec)
    tol = 1e-12 if prec == bdk.FP64 else 1e-5
    assert rel_rms(g.berechne_WCA(), o.wca()) < tol
    assert rel_rms(g.berechne_kapkraefte(), o.kap()) < tol
    if prec == bdk.FP64:
        for _ in range(50):
            o.step(1.0)
        g.run_steps(50)
        gg, gr = g.positions()
        og, orl = o.all_positions()
        assert np.array_equal(gg, og) and np.abs(gr - orl).max() < 1e-10
    p1, p2 = str(tmp_path / "p1.txt"), str(tmp_path / "p2.txt")
    g.write_pos(p1, p2, 0, 0.1)
    assert os.path.exists(p1) and not os.path.exists(p2)
    assert len(open(p1).read().split("\n")) == 2 + 800 + 2


def test_two_particles():
    """One particle of each species 1.05 apart: the WCA pair (Newton's third
    law) and the capillary pair force against the oracle, and 20 steps."""
    say("case: two particles")
    bp, _ = params(n1=1, n2=1, L=20.0, Z=32)
    der = bdk.derived(bp)
    x = np.array([[5.0, 5.0], [5.9, 5.54]])
    git, rel = to_git_rel(x, der["B"], der["nc"])
    g, o, _, _ = pair(git=git, rel=rel, n1=1, n2=1, L=20.0, Z=32, scheme=bdk.CIC)
    fw, ow = g.berechne_WCA(), o.wca()
    assert np.abs(fw - ow).max() <= 1e-12 * np.abs(ow).max()
    assert np.abs(fw.sum(0)).max() <= 1e-12 * np.abs(fw).max()
    fk, ok = g.berechne_kapkraefte(), o.kap()
    assert np.abs(fk - ok).max() <= 1e-12 * np.abs(ok).max()
    for _ in range(20):
        o.step(1.0)
    g.run_steps(20)
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.array_equal(gg, og) and np.abs(gr - orl).max() < 1e-10


@pytest.mark.parametrize("prec", [bdk.FP64, bdk.FP32])
def test_largest_mesh_forces(prec):
    """The largest mesh the boundary accepts (Z = 4096, h = 0.049): capillary
    and WCA forces against the oracle, TSC."""
    say("case: Z=4096 prec=%d" % prec)
    g, o, git, rel = pair(n1=1500, n2=1500, L=200.0, Z=4096, scheme=bdk.TSC, lam=20.0, seed=91, precision=prec)
    tol = 1e-12 if prec == bdk.FP64 else 1e-5
    assert rel_rms(g.berechne_kapkraefte(), o.kap()) < tol
    assert rel_rms(g.berechne_WCA(), o.wca()) < tol
    g.run_steps(3)
    gg, gr = g.positions()
    assert np.all(np.isfinite(gr))


def test_restrict_radial_trajectory():
    """restrictRadial (dm:138-155): species 2 kept on the fixed angle
    2 pi i / N2 around the box centre at its own distance, against the oracle
    over 30 noisy FP64 steps.  Species 2 starts on a ring of radius 15 at
    those angles (spacing 3 sigma) and species 1 away from it, so that no
    projection lands on another particle (a landing inside the WCA core
    would make the run chaotic and its rounding-level differences grow)."""
    say("case: restrictRadial")
    from helpers import rsa_positions
    L, n2 = 40.0, 30
    bp, _ = params(n1=300, n2=n2, L=L, restrict_radial=1)
    der = bdk.derived(bp)
    ang = 2 * np.pi * np.arange(n2) / n2
    ring = np.stack([0.5 * L + 15.0 * np.cos(ang), 0.5 * L + 15.0 * np.sin(ang)], 1)
    x1 = rsa_positions(600, L, 0.9, 95)
    d = np.sqrt(((x1[:, None, :] - ring[None, :, :]) ** 2).sum(2)).min(1)
    x1 = x1[d > 2.0][:300]
    x = np.concatenate([x1, ring])
    git, rel = to_git_rel(x, der["B"], der["nc"])
    g, o, _, _ = pair(git=git, rel=rel, n1=len(x1), n2=n2, L=L, restrict_radial=1, seed=95)
    for _ in range(30):
        o.step(1.0)
    g.run_steps(30)
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.array_equal(gg, og)
    assert np.abs(gr - orl).max() < 1e-10
    xs = gg[len(x1):] * g.B + gr[len(x1):] - 0.5 * L
    assert np.abs(np.arctan2(xs[:, 1], xs[:, 0]) % (2 * np.pi) - ang).max() < 1e-9


@pytest.mark.parametrize("switch", [dict(no_wca=1), dict(idgas_only=1), dict(max_pair_wca=50.0)])
def test_parameter_switches_trajectory(switch):
    """parameter.h switches against the oracle over 30 noisy FP64 steps:
    noWCA (which switches the noise off, dm:90-101), idgas_only (no pair or
    mesh forces), and a maxPairWCA cap that binds (P6: |F|/r capped, pairs
    as close as 0.8 sigma)."""
    say("case: switch %s" % switch)
    dmin = 0.8 if "max_pair_wca" in switch else 0.9
    g, o, git, rel = pair(n1=400, n2=300, L=40.0, Z=64, scheme=bdk.CIC, seed=95, dmin=dmin, **switch)
    if "max_pair_wca" in switch:
        assert rel_rms(g.berechne_WCA(), o.wca()) < 1e-12
    for _ in range(30):
        o.step(1.0)
    g.run_steps(30)
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.array_equal(gg, og)
    assert np.abs(gr - orl).max() < 1e-10
