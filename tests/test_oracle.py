"""Pins the FP64 CPU oracle (oracle/bd_oracle.c) against checks that do not
come from the oracle's own code: the C++ standard library's mt19937_64
stream, Random123's Philox known answers, numpy.fft and a brute-force DFT,
brute-force all-pair WCA, closed-form weights, symbols and pair forces
(scipy K1), and integer/format invariants.  CPU only."""
import json
import math
import os

import numpy as np
import pytest
import scipy.special

import oracle as O
from helpers import (deposit_numpy, derived, gather_numpy, rel_rms, rsa_positions, symbols,
                     to_git_rel, wca_bruteforce)

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module", autouse=True)
def _build():
    O.build()


def make_state(n1, n2, L, Z, seed=1, dmin=0.9, **kw):
    p = O.make_params(n1=n1, n2=n2, L=L, Z=Z, **kw)
    o = O.Oracle(p)
    x = rsa_positions(n1 + n2, L, dmin, seed)
    git, rel = to_git_rel(x, o.B, o.nc)
    o.set_all(git, rel)
    return o, x


# ---------------------------------------------------------------- RNG
def test_mt19937_64_matches_libstdcxx():
    g = json.load(open(os.path.join(GOLD, "mt19937_64_libstdcxx.json")))
    o = O.Oracle(O.make_params(n1=1, n2=1))
    o.mt_seed(5489)
    for _ in range(9999):
        o.mt_next()
    assert o.mt_next() == int(g["default_seed_10000th"]) == 9981545732273789042
    for s in g["streams"]:
        o.mt_seed(int(s["seed"]))
        assert [o.mt_next() for _ in range(len(s["raw"]))] == [int(v) for v in s["raw"]]
        o.mt_seed(int(s["seed"]))
        got = [o.mt_canonical() for _ in range(len(s["canonical"]))]
        assert got == [float.fromhex(v) for v in s["canonical"]]


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
        ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
        ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
         [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
    ]
    for ctr, key, want in kat:
        assert list(O.philox4x32(ctr, key)) == want


def test_philox_noise_distributions():
    n = 40000
    disc = np.array([O.philox_noise(7, i, 3) for i in range(n)])
    r2 = (disc ** 2).sum(1)
    assert r2.max() <= 5.0 + 1e-12
    # uniform on the disc of radius sqrt(5): variance 5/4 per component (dm:1321-1325)
    assert abs(disc.var(0).mean() - 1.25) < 0.02
    assert abs(np.mean(r2 < 1.25) - 0.25) < 0.01
    g = np.array([O.philox_noise(7, i, 3, True, 2.0) for i in range(n)])
    assert abs(g.var(0).mean() - 2.0) < 0.05
    import scipy.stats
    assert abs(scipy.stats.kurtosis(g[:, 0])) < 0.1  # Gaussian: excess kurtosis 0
    # different steps / ids give different samples
    assert not np.allclose(O.philox_noise(7, 1, 3), O.philox_noise(7, 1, 4))


def test_mt_disc_noise_consumption_order():
    """dm:1317-1339: per particle (S1 then S2) draw x then y until inside
    the disc; check against the libstdc++ stream."""
    g = json.load(open(os.path.join(GOLD, "mt19937_64_libstdcxx.json")))
    s = g["streams"][0]
    u = [float.fromhex(v) for v in s["canonical"]]
    p = O.make_params(n1=5, n2=4, L=16.0, Z=16, noise_mode=0, seed=int(s["seed"]))
    o = O.Oracle(p)
    got = o.noise()
    want = []
    k = 0
    w5 = math.sqrt(5.0)
    for _ in range(9):
        while True:
            x = w5 * (-1.0 + 2.0 * u[k])
            y = w5 * (-1.0 + 2.0 * u[k + 1])
            k += 2
            if x * x + y * y <= 5.0:
                break
        want.append((x, y))
    assert np.array_equal(got, np.array(want))


def test_init_zufallspos_precedence_quirk():
    """dm:1655: git = ((int)x)/B, not (int)(x/B); rel = x - git*B."""
    g = json.load(open(os.path.join(GOLD, "mt19937_64_libstdcxx.json")))
    s = g["streams"][1]
    u = [float.fromhex(v) for v in s["canonical"]]
    L = 64.0
    p = O.make_params(n1=30, n2=20, L=L, Z=16, noise_mode=0, seed=int(s["seed"]))
    o = O.Oracle(p)
    o.init_zufallspos()
    git, rel = o.all_positions()
    xs = np.array([L * u[2 * i] for i in range(50)])
    ys = np.array([L * u[2 * i + 1] for i in range(50)])
    gx = np.array([int(int(v) / o.B) for v in xs])
    gy = np.array([int(int(v) / o.B) for v in ys])
    assert np.array_equal(git[:, 0], gx) and np.array_equal(git[:, 1], gy)
    assert np.array_equal(rel[:, 0], xs - gx * o.B)
    assert np.array_equal(rel[:, 1], ys - gy * o.B)
    # the quirk puts some particles one cell low, rel up to about 2B
    assert rel.max() > o.B and rel.max() < 2 * o.B + 1e-9


# ---------------------------------------------------------------- FFT
def test_fft2d_bruteforce_dft():
    rng = np.random.default_rng(3)
    for n in (4, 8, 16):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        j = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(j, j) / n)
        want = W @ a @ W.T
        assert np.allclose(O.fft2d(a, -1), want, rtol=0, atol=1e-12 * np.abs(want).max())
        assert np.allclose(O.fft2d(a, +1), np.conj(W) @ a @ np.conj(W).T, atol=1e-12 * np.abs(want).max())


def test_fft2d_numpy():
    rng = np.random.default_rng(4)
    for n in (32, 128, 256):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        want = np.fft.fft2(a)
        assert rel_rms(O.fft2d(a, -1), want) < 1e-14
        assert rel_rms(O.fft2d(a, +1), np.fft.ifft2(a) * n * n) < 1e-14


def test_fft2d_lengths_not_powers_of_two():
    """The direct DFT of even lengths with factors 3, 5, 7 (the reference's
    default mesh is 384): against the matrix definition at small n and
    numpy.fft up to 384.  A wrong sign or exponent (jk instead of jk mod n is
    harmless, j + k is not) fails either check."""
    rng = np.random.default_rng(5)
    for n in (6, 10, 12, 14, 18):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        j = np.arange(n)
        W = np.exp(-2j * np.pi * np.outer(j, j) / n)
        want = W @ a @ W.T
        assert np.allclose(O.fft2d(a, -1), want, rtol=0, atol=1e-12 * np.abs(want).max())
    for n in (24, 50, 96, 98, 384):
        a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
        assert rel_rms(O.fft2d(a, -1), np.fft.fft2(a)) < 1e-13
        assert rel_rms(O.fft2d(a, +1), np.fft.ifft2(a) * n * n) < 1e-13


# ---------------------------------------------------------------- cell lists
def test_extended_lists_membership_bruteforce():
    o, _ = make_state(150, 120, 24.0, 32, seed=5)
    git, _ = o.all_positions()
    nc = o.nc
    n1 = o.p.n1
    for cx in range(nc):
        for cy in range(nc):
            for sp, sl in ((1, slice(0, n1)), (2, slice(n1, None))):
                g = git[sl]
                dx = (g[:, 0] - cx) % nc
                dy = (g[:, 1] - cy) % nc
                want = np.nonzero(((dx <= 1) | (dx == nc - 1)) & ((dy <= 1) | (dy == nc - 1)))[0]
                got = o.neighbour_list(sp, cx, cy)
                assert sorted(got) == list(want)


def test_refresh_formula_and_lists():
    o, _ = make_state(80, 80, 16.0, 16, seed=6)
    git0, rel0 = o.all_positions()
    rng = np.random.default_rng(1)
    rel = rel0 + rng.uniform(-0.5, 0.5, rel0.shape)
    o.set_positions(1, git0[:80], rel[:80])
    o.set_positions(2, git0[80:], rel[80:])
    o.refresh()
    git1, rel1 = o.all_positions()
    B, nc = o.B, o.nc
    out = (rel < 0) | (rel > B)
    moved = out.any(1)
    for i in range(160):
        if moved[i]:
            ic = np.floor((git0[i] * B + rel[i]) / B).astype(int)
            assert np.array_equal(git1[i], (ic + nc) % nc)
            assert np.array_equal(rel1[i], np.fmod(rel[i] + B, B))
        else:
            assert np.array_equal(git1[i], git0[i]) and np.array_equal(rel1[i], rel[i])
    # lists follow the new cells
    git = git1
    for cx in range(nc):
        for cy in range(nc):
            g = git[:80]
            dx = (g[:, 0] - cx) % nc
            dy = (g[:, 1] - cy) % nc
            want = np.nonzero(((dx <= 1) | (dx == nc - 1)) & ((dy <= 1) | (dy == nc - 1)))[0]
            assert sorted(o.neighbour_list(1, cx, cy)) == list(want)


# ---------------------------------------------------------------- WCA
def test_wca_two_particle_closed_form():
    """F = 24(2 r^-13 - r^-7) along the separation (the derivative of the
    WCA potential, plot_kapillar_und_WCA.plt:8), capped as |F|/r <= maxPairWCA."""
    L = 20.0
    for r in (0.6, 0.8, 0.95, 1.05, 1.12, 1.2):
        p = O.make_params(n1=2, n2=0, L=L, Z=16)
        o = O.Oracle(p)
        x = np.array([[5.3, 7.1], [5.3, 7.1 + r]])
        git, rel = to_git_rel(x, o.B, o.nc)
        o.set_all(git, rel)
        F = o.wca()
        want = 24.0 * (2 * r ** -13 - r ** -7) if r < 2 ** (1 / 6) else 0.0
        want = min(want, 1e5 * r)
        assert F[1, 1] == pytest.approx(want, rel=1e-12, abs=1e-12)
        assert F[0, 1] == pytest.approx(-want, rel=1e-12, abs=1e-12)
        assert F[1, 0] == 0.0
    # the cap acts on |F|/r (P6): at r = 0.4, |F|/r = 24(2 r^-14 - r^-8) > 1e5
    p = O.make_params(n1=2, n2=0, L=L, Z=16)
    o = O.Oracle(p)
    x = np.array([[5.3, 7.1], [5.3, 7.5]])
    git, rel = to_git_rel(x, o.B, o.nc)
    o.set_all(git, rel)
    assert o.wca()[1, 1] == pytest.approx(1e5 * 0.4, rel=1e-9)


@pytest.mark.parametrize("ratios", [(1.0, 1.0, 1.0, 1.0), (0.8, 0.9, 1.3, 0.7)])
def test_wca_bruteforce_all_pairs(ratios):
    s22, s12, e22, e12 = ratios
    o, x = make_state(200, 150, 20.0, 32, seed=9, dmin=0.75, s11_22=s22, s11_12=s12,
                      e22_11=e22, e12_11=e12)
    F = o.wca()
    git, _ = o.all_positions()
    want = wca_bruteforce(x, 200, 20.0, s22, s12, e22, e12, git=git, nc=o.nc)
    assert rel_rms(F, want) < 1e-12
    if ratios[0] == 1.0:
        assert np.abs(F.sum(0)).max() < 1e-9 * np.abs(F).max()


# ---------------------------------------------------------------- mesh
@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_deposit_textbook_weights_and_charge(scheme):
    o, x = make_state(300, 250, 32.0, 32, seed=11, scheme=scheme, f2_f1=-0.7)
    o.kap()
    rho = o.mesh(0)
    h = o.hmesh
    q = np.concatenate([np.ones(300), -0.7 * np.ones(250)])
    want = deposit_numpy(x, q, 32, h, scheme)
    assert np.allclose(rho, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert rho.sum() * h * h == pytest.approx(300 - 0.7 * 250, rel=1e-12)


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_spectral_pipeline_symbols(scheme):
    """Fx = Re IFFT(i gx FFT(rho)), gx = v sin(qh)/h G / Z^2 with the
    discrete-Laplacian G (dm:585-593, dm:1384-1394, dm:1206-1222)."""
    Z, L, lam = 32, 24.0, 5.0
    o, x = make_state(200, 200, L, Z, seed=12, scheme=scheme, lam=lam, kap=1.3)
    o.kap()
    rho = o.mesh(0)
    gx, gy = symbols(Z, L, lam, 1.3)
    rk = np.fft.fft2(rho)
    fx = np.real(np.fft.ifft2(1j * gx * rk)) * Z * Z
    fy = np.real(np.fft.ifft2(1j * gy * rk)) * Z * Z
    assert rel_rms(o.mesh(1), fx) < 1e-13
    assert rel_rms(o.mesh(2), fy) < 1e-13


@pytest.mark.parametrize("scheme", [1, 2])
def test_gather_textbook_and_adjoint(scheme):
    o, x = make_state(250, 250, 32.0, 32, seed=13, scheme=scheme, f2_f1=0.5)
    F = o.kap()
    Fx, Fy = o.mesh(1), o.mesh(2)
    h = o.hmesh
    q = np.concatenate([np.ones(250), 0.5 * np.ones(250)])
    want = np.stack([gather_numpy(x, Fx, h, scheme), gather_numpy(x, Fy, h, scheme)], 1) * q[:, None]
    assert rel_rms(F, want) < 1e-13
    # adjoint identity <deposit(q), phi> h^2 = sum_i q_i gather(phi)_i
    phi = np.random.default_rng(2).normal(size=(32, 32))
    lhs = (deposit_numpy(x, q, 32, h, scheme) * phi).sum() * h * h
    rhs = (q * gather_numpy(x, phi, h, scheme)).sum()
    assert lhs == pytest.approx(rhs, rel=1e-12)
    # momentum: sum_i q_i F_i = 0 for an antisymmetric kernel (CIC/TSC)
    assert np.abs(F.sum(0)).max() < 1e-10 * np.abs(F).max() * len(F)


def test_ngp_species2_y_bug():
    """gr:123 reads Fx[..][1] (imaginary part, ~0) as the species-2 y force."""
    o, x = make_state(100, 100, 24.0, 32, seed=14, scheme=0)
    F = o.kap()
    assert np.abs(F[100:, 1]).max() < 1e-13 * np.abs(F[:100, 1]).max()
    p = O.make_params(n1=100, n2=100, L=24.0, Z=32, scheme=0, ngp_fix=1)
    o2 = O.Oracle(p)
    git, rel = o.all_positions()
    o2.set_all(git, rel)
    F2 = o2.kap()
    assert np.abs(F2[100:, 1]).max() > 1e-3 * np.abs(F2[:100, 1]).max()


@pytest.mark.parametrize("scheme", [1, 2])
def test_two_particle_capillary_force_vs_K1(scheme):
    """Pair force vs kap K1(r/lambda)/(2 pi lambda) (plot_kapill_vs_k.plt:35),
    valid for h << r << L; the self force is removed by a one-particle run."""
    L, Z, lam, kap = 128.0, 256, 6.0, 1.0
    h = L / Z
    base = np.array([64.0 + 0.5 * h, 64.0 + 0.5 * h])
    for r in (3.0, 6.0, 9.0):
        p = O.make_params(n1=2, n2=0, L=L, Z=Z, scheme=scheme, lam=lam, kap=kap)
        o = O.Oracle(p)
        x = np.array([base, base + [0.0, r]])
        git, rel = to_git_rel(x, o.B, o.nc)
        o.set_all(git, rel)
        F = o.kap()
        p1 = O.make_params(n1=1, n2=0, L=L, Z=Z, scheme=scheme, lam=lam, kap=kap)
        o1 = O.Oracle(p1)
        g1, r1 = to_git_rel(x[1:], o1.B, o1.nc)
        o1.set_all(g1, r1)
        Fs = o1.kap()
        pair = F[1, 1] - Fs[0, 1]
        want = -kap * scipy.special.k1(r / lam) / (2 * np.pi * lam)
        assert pair == pytest.approx(want, rel=0.02)


# ---------------------------------------------------------------- dt, update, loop
def test_optimal_dt_closed_form():
    rng = np.random.default_rng(5)
    o, x = make_state(60, 60, 16.0, 16, seed=15, mrd=0.05, dt_max=0.01, T=2.0)
    F = o.wca() + o.kap()
    nz = o.noise()
    T, mrd = 2.0, 0.05
    cand = np.where(np.abs(F) < 1e-5, mrd ** 2 / (2 * T * nz ** 2),
                    0.0)
    u = nz * math.sqrt(0.5 * T) / np.where(np.abs(F) < 1e-5, 1.0, F)
    w = np.sqrt(u * u + mrd / np.abs(np.where(np.abs(F) < 1e-5, 1.0, F)))
    cand = np.where(np.abs(F) < 1e-5, cand, (w - np.abs(u)) ** 2)
    assert o.optimal_dt() == pytest.approx(min(0.01, cand.min()), rel=1e-14)
    # displacement bound: |F dt + sqrt(2T dt) n| <= mrd per component
    dt = o.optimal_dt()
    assert np.abs(F * dt + math.sqrt(2 * T * dt) * nz).max() <= mrd * (1 + 1e-9)
    # fixed-dt configuration: mrd = inf gives dt_max exactly
    o2, _ = make_state(60, 60, 16.0, 16, seed=15)
    o2.wca(); o2.kap(); o2.noise()
    assert o2.optimal_dt() == 1e-4


def test_step_update_closed_form():
    o, x = make_state(100, 80, 20.0, 32, seed=16, noise_mode=3)
    nz = np.random.default_rng(3).uniform(-1, 1, (180, 2))
    o.set_noise(nz)
    g0, r0 = o.all_positions()
    Fw = o.wca()
    Fk = o.kap()
    dt = o.step(1.0)
    assert dt == 1e-4
    want = r0 + dt * (Fw + Fk) + math.sqrt(2.0 * 1.0 * dt) * nz
    g1, r1 = o.all_positions()
    B = o.B
    assert np.allclose(g1 * B + r1, g0 * B + want, rtol=0, atol=1e-12)
    assert o.step_counter() == 1
    assert o.step(-1.0) == 1e10 and o.step_counter() == 1


def test_checkpoint_step_count_long_double():
    """zeitschritte_bis_obs accumulates t in long double (dm:444-455)."""
    assert O.steps_for_interval(0.1, 1e-4) == 1001
    assert O.steps_for_interval(0.2, 1e-4) == 2001
    assert O.steps_for_interval(1.0, 1e-4) == 10001
    o, _ = make_state(20, 20, 16.0, 16, seed=17)
    assert o.steps_until_obs(0.003) == O.steps_for_interval(0.003, 1e-4)


def test_pos_writer_format(tmp_path):
    o, x = make_state(3, 2, 16.0, 16, seed=18)
    p1 = str(tmp_path / "pos1_1.txt")
    o.write_pos_header(p1)
    o.append_pos(1, p1, 0, 0.2)
    o.append_pos(1, p1, 3, 0.2)
    txt = open(p1).read()
    g, r = o.positions(1)
    xs = g * o.B + r
    lines = ["# Format: t TAB x TAB y ", ""]
    for nr in (0, 3):
        for i in range(3):
            lines.append("%g \t %g \t %g " % (nr * 0.2, xs[i, 0], xs[i, 1]))
        lines.append("")
    assert txt == "\n".join(lines) + "\n"


def test_restrict_radial_invariants():
    """restrictRadial (dm:138-155): after a step, species-2 particle i lies
    on the ray at angle 2 pi i / N2 about the box centre, at the distance
    from the centre that its move without the switch reaches; species 1 moves
    as without the switch.  Species 2 starts on its rays (random distances)
    so that a step moves it by much less than a list cell: the reference's
    refresh re-bins with a single fmod (dm:216-325), which misplaces moves of
    more than a cell, as the projection from an arbitrary start would make.
    Checked from three states (the start, after 5 and after 10 steps)
    against a second oracle that steps the same state (same forces, same
    Philox draws) without the switch."""
    # L = 31: the centre lies inside a list cell (13.5 B), so the rays at
    # multiples of pi/2 do not run along cell edges
    n1, n2, L = 200, 40, 31.0
    kw = dict(n1=n1, n2=n2, L=L, Z=32, scheme=2, T=1.0, dt_max=1e-4, noise_mode=1, seed=77)
    m = 0.5 * L
    want = 2 * np.pi * np.arange(n2) / n2
    rad = np.random.default_rng(4).uniform(8.0, 13.0, n2)  # neighbouring rays >= 1.25 apart
    x2 = np.stack([m + rad * np.cos(want), m + rad * np.sin(want)], 1)
    x1 = rsa_positions(600, L, 0.9, 95)
    d = np.sqrt(((x1[:, None, :] - x2[None, :, :]) ** 2).sum(2)).min(1)
    x1 = x1[d > 1.5][:n1]
    base = O.Oracle(O.make_params(restrict_radial=1, **kw))
    git, rel = to_git_rel(np.concatenate([x1, x2]), base.B, base.nc)
    base.set_all(git, rel)
    for warm in (0, 5, 10):
        for _ in range(0 if warm == 0 else 5):
            base.step(1.0)
        git, rel = base.all_positions()
        on, off = O.Oracle(O.make_params(restrict_radial=1, **kw)), O.Oracle(O.make_params(**kw))
        for o in (on, off):  # both count steps from 0: the same Philox draws
            o.set_all(git, rel)
        on.step(1.0)
        off.step(1.0)
        xa, xb = on.absolute(), off.absolute()
        assert np.array_equal(xa[:n1], xb[:n1])
        ra = np.hypot(xa[n1:, 0] - m, xa[n1:, 1] - m)
        rb = np.hypot(xb[n1:, 0] - m, xb[n1:, 1] - m)
        assert np.allclose(ra, rb, rtol=0, atol=1e-9)
        ang = np.arctan2(xa[n1:, 1] - m, xa[n1:, 0] - m)
        assert np.abs(np.angle(np.exp(1j * (ang - want)))).max() < 1e-9
        # the switch does move species 2 (off its free move, onto the ray)
        assert np.abs(xa[n1:] - xb[n1:]).max() > 1e-6
        on.close()
        off.close()
