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


@pytest.mark.parametrize("m", [None, 37])
def test_snapshot_density_grid_and_stats(m):
    """The snapshot grid (main_auswertung.cpp:214-320) against the numpy
    restatement on the device positions, at the reference's grid size
    m = 2 (int)(L/(2 * 0.5) + 0.5) and at an odd m; each species integrates
    to its particle count; the histograms and threshold fractions
    (main_auswertung.cpp:324-375) equal the restatement's on the device grid."""
    from helpers import density_stats_ref, tsc_snapshot_ref
    L = 60.0
    m = m or 2 * int(L / (2 * 0.5) + 0.5)
    print("case: snapshot grid m=%d" % m, flush=True)
    bp, _ = params(n1=1800, n2=1200, L=L, Z=64, scheme=bdk.TSC, lam=20.0, precision=bdk.FP32, seed=71)
    git, rel, _ = bdk.init_blob(bp, 71, L / 5, 0.2, 0.95)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(30)
    r1, r2 = s.density_grid(m)
    x = s.absolute()
    w1, w2 = tsc_snapshot_ref(x[:1800], x[1800:], L, m)
    # each of a cell's contributions is rounded to 2^-41 of a weight (fixed
    # point), times 1/w^2; the FP64 sum adds rounding of its own
    w = L / m
    for got, ref, xs in ((r1, w1, x[:1800]), (r2, w2, x[1800:])):
        k = np.floor(xs / w).astype(np.int64)
        cnt = np.zeros((m, m))
        for a in (-1, 0, 1):
            for b in (-1, 0, 1):
                np.add.at(cnt, ((k[:, 0] + a) % m, (k[:, 1] + b) % m), 1.0)
        bound = cnt * (2.0 ** -41 + 1e-15) / (w * w)
        assert np.all(np.abs(got - ref) <= bound)
    assert abs(r1.sum() * w * w - 1800) < 1e-8 and abs(r2.sum() * w * w - 1200) < 1e-8
    th, dth = (0.1, 0.2, 0.5), (0.1, 0.2, 0.5, 0.99)
    hist, over = s.density_stats(m, 100, 0.01, th, dth)
    hr, orf = density_stats_ref(r1, r2, 100, 0.01, th, dth)
    assert np.array_equal(hist, hr) and np.array_equal(over, orf)
    assert np.allclose(hist.sum(1), 1.0)
    s.close()


def test_ftrho_sums_match_direct_formula():
    """Direct rho(k) sums (record_ftrho_unkorrigiert, au:307-360) over the
    reference's (i, j) pairs against numpy on the device positions; for
    species 1 y = git_y B + rel_x, as au:333 reads; pair (0, 0) counts the
    species exactly."""
    print("case: ftrho sums", flush=True)
    L, qb = 50.0, 12
    bp, _ = params(n1=900, n2=600, L=L, Z=64, scheme=bdk.CIC, precision=bdk.FP64, seed=73)
    git, rel = positions(bp, 73)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(5)
    g, r = s.positions()
    B, dq = s.B, 4.0 * np.pi / L
    pairs = [(i, j) for i in range(qb) for j in range(qb) if i * i + j * j <= qb * qb]
    # the loop breaks at the first j beyond, which keeps exactly these pairs
    for sp, sl in ((1, slice(0, 900)), (2, slice(900, 1500))):
        re, im = s.ftrho_sums(sp, qb, dq)
        assert len(re) == len(pairs)
        x = g[sl, 0] * B + r[sl, 0]
        y = g[sl, 1] * B + (r[sl, 0] if sp == 1 else r[sl, 1])
        for k, (i, j) in enumerate(pairs):
            a = dq * (x * i + y * j)
            assert abs(re[k] - np.cos(a).sum()) < 1e-9 and abs(im[k] - np.sin(a).sum()) < 1e-9
        assert re[0] == (900 if sp == 1 else 600) and im[0] == 0.0
    s.close()
    # ftrho_fix: species 1 takes its own rel_y
    bf, _ = params(n1=900, n2=600, L=L, Z=64, scheme=bdk.CIC, precision=bdk.FP64, seed=73, ftrho_fix=1)
    f = bdk.RunDynamik(bf)
    f.set_positions(g, r)
    re, im = f.ftrho_sums(1, qb, dq)
    x = g[:900, 0] * B + r[:900, 0]
    y = g[:900, 1] * B + r[:900, 1]
    for k, (i, j) in enumerate(pairs):
        a = dq * (x * i + y * j)
        assert abs(re[k] - np.cos(a).sum()) < 1e-9 and abs(im[k] - np.sin(a).sum()) < 1e-9
    f.close()


def test_ftrho_sums_many_particles():
    """300k particles: more particles than 64 chunks of the minimum size, so
    the chunks grow instead of multiplying (the partial sums stay at 64 per
    pair); against numpy."""
    print("case: ftrho sums at 300k particles", flush=True)
    L, qb = 800.0, 5
    bp, _ = params(n1=150000, n2=150000, L=L, Z=256, scheme=bdk.CIC, precision=bdk.FP64, seed=79)
    git, rel = bdk.init_uniform(bp, 79, 0.9)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    g, r = s.positions()
    B, dq = s.B, 2.0 * np.pi / L
    pairs = [(i, j) for i in range(qb) for j in range(qb) if i * i + j * j <= qb * qb]
    re, im = s.ftrho_sums(2, qb, dq)
    x = g[150000:, 0] * B + r[150000:, 0]
    y = g[150000:, 1] * B + r[150000:, 1]
    for k, (i, j) in enumerate(pairs):
        a = dq * (x * i + y * j)
        assert abs(re[k] - np.cos(a).sum()) < 1e-7 and abs(im[k] - np.sin(a).sum()) < 1e-7
    s.close()
