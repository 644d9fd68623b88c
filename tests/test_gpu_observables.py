"""Device observables: radial density profiles of a blob
(observablen_auswertung.cpp:140-330) against the same formulas evaluated on
the CPU from the device positions, plus the normalisation the definition
fixes (sum over bins of density x ring area = L^2 per species)."""
import numpy as np
import pytest

import bdk
from common import params, positions
from helpers import radial_profile_ref, rho_k_ref

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [bdk.FP64, bdk.FP32])
def test_radial_profile_matches_reference_formula(prec):
    print("case: radial profile prec=%d" % prec, flush=True)
    bp, _ = params(n1=2500, n2=1500, L=90.0, Z=128, scheme=bdk.TSC, lam=20.0, precision=prec, seed=41)
    git, rel, _ = bdk.init_blob(bp, 41, 90.0 / 6, 0.1, 0.95)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(40)
    nbins, dr = 60, 1.0
    cen, prof = s.radial_profile(nbins, dr)
    x = s.absolute()
    cref, pref = radial_profile_ref(x[:2500], x[2500:], 90.0, nbins, dr)
    assert np.allclose(cen, cref, rtol=0, atol=1e-9)
    assert np.allclose(prof, pref, rtol=1e-12, atol=0)
    # normalisation: sum_b rho_b * pi dr^2 (2b+1) = L^2 for every profile
    ring = np.pi * dr * dr * (2 * np.arange(nbins) + 1)
    assert np.allclose((prof * ring).sum(1), 90.0 ** 2, rtol=1e-12)
    # the blob centre sits at the box centre
    assert np.abs(cen - 45.0).max() < 1.0


def test_radial_profile_of_a_ring():
    """particles on a circle of radius 10.5 around the box centre: every
    particle in bin 10 of the profile around its own centre."""
    bp, _ = params(n1=64, n2=64, L=40.0, Z=64, precision=bdk.FP64, seed=3)
    der = bdk.derived(bp)
    th = 2 * np.pi * (np.arange(128) + 0.25) / 128
    x = np.stack([20.0 + 10.5 * np.cos(th), 20.0 + 10.5 * np.sin(th)], 1)
    order = np.r_[np.arange(0, 128, 2), np.arange(1, 128, 2)]  # both species on the whole ring
    x = x[order]
    git = np.floor(x / der["B"]).astype(np.int32)
    rel = x - git * der["B"]
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    cen, prof = s.radial_profile(20, 1.0)
    assert np.abs(cen - 20.0).max() < 1e-9
    for k in range(6):
        nz = np.nonzero(prof[k])[0]
        assert list(nz) == [10], (k, nz)


@pytest.mark.parametrize("scheme", [bdk.NGP, bdk.CIC, bdk.TSC])
def test_rho_k_matches_reference_formula(scheme):
    """rho(k) per species against numpy.fft of the textbook deposit; shell 0
    is the DC term alone, N_s exactly."""
    print("case: rho(k) scheme=%d" % scheme, flush=True)
    L, Z = 90.0, 128
    bp, _ = params(n1=2500, n2=1500, L=L, Z=Z, scheme=scheme, lam=20.0, precision=bdk.FP64, seed=42)
    git, rel, _ = bdk.init_blob(bp, 42, L / 6, 0.1, 0.95)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(25)
    x = s.absolute()
    nb = int(0.5 * Z * 0.5)
    for sp, xs in ((1, x[:2500]), (2, x[2500:])):
        re, im = s.rho_k(sp, 0.5, nb)
        want = rho_k_ref(xs, L, Z, scheme, 0.5, nb)
        assert re[0] == pytest.approx(len(xs), rel=1e-12)
        assert np.abs(re - want).max() < 1e-10 * len(xs)
        assert not im.any()
    # the step after an observable is unchanged (the observable only used scratch buffers)
    t = bdk.RunDynamik(bp)
    t.set_positions(git, rel)
    t.run_steps(26)
    s.run_steps(1)
    assert np.array_equal(s.positions()[1], t.positions()[1])


@pytest.mark.parametrize("nbins,dr", [(40, 0.25), (30, 1.0)])
def test_pair_correlation_matches_brute_force(nbins, dr):
    """g11, g12, g22 against all pairs by brute force (the reference's loops);
    (30, 1.0) reaches past half the box, so every pair is seen once through
    the full cell range.  Far beyond the exclusion an ideal-gas-like start
    gives g near 1."""
    from helpers import pair_correlation_ref
    print("case: g(r) nbins=%d dr=%g" % (nbins, dr), flush=True)
    L = 50.0
    bp, _ = params(n1=700, n2=500, L=L, Z=64, scheme=bdk.CIC, precision=bdk.FP64, seed=61)
    git, rel = positions(bp, 61)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(10)
    g = s.pair_correlation(nbins, dr)
    x = s.absolute()
    want = pair_correlation_ref(x[:700], x[700:], L, nbins, dr)
    for got, ref in zip(g, want):
        assert np.allclose(got, ref, rtol=1e-12, atol=0)
    if dr == 0.25:
        assert g[0][:3].max() == 0.0  # nothing closer than 0.75 sigma
