"""Shared test helpers: small seeded particle configurations and brute-force
checks written from the mathematics (not from the oracle's code)."""
import math

import numpy as np


def derived(L, Z):
    zh16 = 2.0 ** (1.0 / 6.0)
    nc = int(math.floor(L / zh16))
    B = L / nc
    return zh16, nc, B, L / Z


def rsa_positions(n, L, dmin, seed, max_tries=200000):
    """Uniform random sequential insertion with a hard exclusion dmin
    (periodic).  Grid-accelerated, small n only."""
    rng = np.random.default_rng(seed)
    ncell = max(1, int(L // dmin))
    cw = L / ncell
    grid = {}
    pts = []
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("RSA jammed")
        p = rng.uniform(0.0, L, 2)
        cx, cy = int(p[0] / cw) % ncell, int(p[1] / cw) % ncell
        ok = True
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for q in grid.get(((cx + dx) % ncell, (cy + dy) % ncell), ()):
                    d = p - q
                    d -= L * np.round(d / L)
                    if d @ d < dmin * dmin:
                        ok = False
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            grid.setdefault((cx, cy), []).append(p)
            pts.append(p)
    return np.array(pts)


def to_git_rel(x, B, nc):
    """(git, rel) with git = floor(x/B) and rel = x - git*B in [0, B)."""
    git = np.floor(x / B).astype(np.int64)
    git = np.clip(git, 0, nc - 1)
    rel = x - git * B
    return git.astype(np.int32), rel


def min_image(d, L):
    return d - L * np.round(d / L)


def wca_bruteforce(x, n1, L, s22=1.0, s12=1.0, e22=1.0, e12=1.0, cap=1.0e5, git=None, nc=None):
    """All-pair WCA forces from the closed form F = 24 eps/sigma ... evaluated
    exactly as the reference scales it: for a pair of type t with sigma
    ratio s and eps ratio e, r2' = r2 s^2, d' = d s, |F|/r = e s 24 (2 r'^-14
    - r'^-8) capped at `cap`, F = (|F|/r) d'.  The 11 pair has s = e = 1."""
    n = len(x)
    F = np.zeros_like(x)
    cut2 = 2.0 ** (1.0 / 3.0)
    for i in range(n):
        d = min_image(x[i] - x, L)
        r2 = (d * d).sum(1)
        for j in np.nonzero(r2 < 4.0)[0]:
            if j == i:
                continue
            if git is not None:
                dc = (git[i] - git[j]) % nc
                if not np.all((dc <= 1) | (dc == nc - 1)):
                    continue
            si, sj = i >= n1, j >= n1
            if not si and not sj:
                s, e = 1.0, None
            elif si and sj:
                s, e = s22, e22
            else:
                s, e = s12, e12
            rr = r2[j] * s * s
            if rr > cut2:
                continue
            c = 24.0 * (2.0 * rr ** -7 - rr ** -4)
            if e is not None:
                c *= e * s
            c = min(c, cap)
            F[i] += c * d[j] * s
    return F


def rel_rms(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.sqrt((np.abs(b) ** 2).sum())
    if nb == 0.0:
        return float(np.sqrt((np.abs(a) ** 2).sum()))
    return float(np.sqrt((np.abs(a - b) ** 2).sum()) / nb)


def textbook_weights(x, h, scheme):
    """Mesh index and 1D weights of NGP/CIC/TSC (Hockney & Eastwood,
    Computer Simulation Using Particles, ch. 5): cell k = floor(x/h), centre
    (k+1/2)h, d = (x - centre)/h; CIC (0 if d>=0 else -d, 1-|d|, d if d>=0);
    TSC ((1/2)(1/2-d)^2, 3/4-d^2, (1/2)(1/2+d)^2)."""
    k = np.floor(x / h).astype(np.int64)
    d = x / h - (k + 0.5)
    if scheme == 0:
        w = np.stack([np.zeros_like(d), np.ones_like(d), np.zeros_like(d)], -1)
    elif scheme == 1:
        w = np.stack([np.where(d < 0, -d, 0.0), 1.0 - np.abs(d), np.where(d >= 0, d, 0.0)], -1)
    else:
        w = np.stack([0.5 * (0.5 - d) ** 2, 0.75 - d * d, 0.5 * (0.5 + d) ** 2], -1)
    return k, w


def deposit_numpy(x, q, Z, h, scheme):
    rho = np.zeros((Z, Z))
    kx, wx = textbook_weights(x[:, 0], h, scheme)
    ky, wy = textbook_weights(x[:, 1], h, scheme)
    for a in range(3):
        for b in range(3):
            np.add.at(rho, ((kx + a - 1) % Z, (ky + b - 1) % Z), q * wx[:, a] * wy[:, b] / (h * h))
    return rho


def gather_numpy(x, field, h, scheme):
    Z = field.shape[0]
    kx, wx = textbook_weights(x[:, 0], h, scheme)
    ky, wy = textbook_weights(x[:, 1], h, scheme)
    out = np.zeros(len(x))
    for a in range(3):
        for b in range(3):
            out += wx[:, a] * wy[:, b] * field[(kx + a - 1) % Z, (ky + b - 1) % Z]
    return out


def symbols(Z, L, lam, kap):
    """Closed-form spectral symbols: discrete-Laplacian Green's function and
    central-difference gradient, FFT index order."""
    h = L / Z
    j = np.arange(Z)
    q = 2 * np.pi / L * np.where(j < Z // 2, j, j - Z)
    lap = 4 * np.sin(0.5 * q * h) ** 2 / h ** 2
    G = 1.0 / (lap[:, None] + lap[None, :] + 1.0 / lam ** 2)
    if lam > 0.2 * L:
        G[0, 0] = 0.0
    gx = kap * np.sin(q * h)[:, None] / h * G / Z / Z
    gy = kap * np.sin(q * h)[None, :] / h * G / Z / Z
    return gx, gy


def radial_profile_ref(x1, x2, L, nbins, dr):
    """Radial density profiles as observablen_auswertung.cpp computes them:
    centres by circular means (berechne_mittelpunkte, :164-217), then per
    particle vorfaktor/(2 bin + 1) into the bin of its distance from a centre
    (berechne_dichteprofile, :262-322).  x1, x2: absolute positions [n][2]."""
    k = 2 * np.pi / L

    def centre(x):
        s = []
        for a in range(2):
            c, sn = np.cos(x[:, a] * k).sum(), np.sin(x[:, a] * k).sum()
            s.append((np.arctan2(-sn, -c) + np.pi) / k)
        return np.array(s)

    S1, S2 = centre(x1), centre(x2)
    Sg = 0.5 * (S1 + S2)

    def prof(x, S, n):
        out = np.zeros(nbins)
        v = L * L / (n * dr * dr * np.pi)
        r = np.sqrt(((x - S) ** 2).sum(1))
        b = (r / dr).astype(np.int64)
        b[(b < 0) | (b >= nbins)] = nbins - 1
        np.add.at(out, b, v / (2 * b + 1))
        return out

    n1, n2 = len(x1), len(x2)
    profiles = np.array([prof(x1, Sg, n1), prof(x2, S2, n2), prof(x1, S1, n1), prof(x1, S2, n1),
                         prof(x2, S1, n2), prof(x2, S2, n2)])
    return np.concatenate([S1, S2, Sg]), profiles


def rho_k_ref(x, L, Z, scheme, dqr, nbins):
    """record_rhoFFTW (auswertung.cpp:529-626) on one species' absolute
    positions: density by the force deposit's weights, unnormalised forward
    2D FFT, (-1)^(i+j), shell sums of the real parts over the full plane
    divided by the shell's lattice-point count, times L^2/Z^2."""
    h = L / Z
    rho = deposit_numpy(x, np.ones(len(x)), Z, h, scheme)
    F = np.fft.fft2(rho)
    i = np.arange(Z)
    F = F * np.where((i[:, None] + i[None, :]) % 2 == 1, -1.0, 1.0)
    fq = ((i + Z // 2) % Z) - Z // 2
    b = (dqr * np.sqrt(fq[:, None] ** 2 + fq[None, :] ** 2) + 0.5).astype(np.int64)
    ok = b < nbins
    re = np.bincount(b[ok], weights=F.real[ok], minlength=nbins)[:nbins]
    cnt = np.bincount(b[ok], minlength=nbins)[:nbins]
    return np.where(cnt > 0, re * L * L / (Z * Z) / np.maximum(cnt, 1), 0.0)


def pair_correlation_ref(x1, x2, L, nbins, dr):
    """record_korrelationsfunktion11/12/22 (auswertung.cpp:110-188) by brute
    force over all pairs: abstand2 (am:100-166) per axis min over d, d+L, d-L."""
    def d2(a, b):
        d = a[:, None, :] - b[None, :, :]
        m = np.minimum(np.minimum(d * d, (d + L) ** 2), (d - L) ** 2)
        return m.sum(2)

    def hist(a2, v, mask):
        idx = np.floor(np.sqrt(a2[mask]) / dr).astype(np.int64)
        idx = idx[idx < nbins]
        return np.bincount(idx, minlength=nbins)[:nbins] * (v / (2 * np.arange(nbins) + 1))

    n1, n2 = len(x1), len(x2)
    v11 = 2.0 / ((n1 - 1) * np.pi * (n1 / L / L) * dr * dr)
    v12 = 1.0 / ((n1 - 1) * np.pi * (n2 / L / L) * dr * dr)
    v22 = 2.0 / ((n2 - 1) * np.pi * (n2 / L / L) * dr * dr)
    a11, a12, a22 = d2(x1, x1), d2(x1, x2), d2(x2, x2)
    tri1 = np.tril(np.ones((n1, n1), bool), -1)
    tri2 = np.tril(np.ones((n2, n2), bool), -1)
    return hist(a11, v11, tri1), hist(a12, v12, np.ones((n1, n2), bool)), hist(a22, v22, tri2)


def tsc_snapshot_ref(x1, x2, L, m):
    """Snapshot grid (main_auswertung.cpp:214-300), vectorised: per particle
    k = (int)(x/w) mod m, d = x - (k+1/2) w, centre 3/4 - d^2/w^2, sides
    (3/2 - |d +- w|/w)^2 / 2, times 1/w^2 into the 3x3 cells (x index
    first).  Returns rho1, rho2 [m, m] (FP64 sums)."""
    w = L / m
    out = []
    for x in (x1, x2):
        g = np.zeros((m, m))
        if len(x):
            k0 = np.floor(x / w).astype(np.int64)  # x >= 0: (int) truncation
            d = x - (k0 + 0.5) * w
            t_l, t_r = np.abs((d + w) / w), np.abs((d - w) / w)
            wt = np.stack([0.5 * (1.5 - t_l) ** 2, 0.75 - d * d / (w * w), 0.5 * (1.5 - t_r) ** 2])  # [3, n, 2]
            for a in range(3):
                for b in range(3):
                    np.add.at(g, ((k0[:, 0] - 1 + a) % m, (k0[:, 1] - 1 + b) % m),
                              wt[a, :, 0] * wt[b, :, 1] / (w * w))
        out.append(g)
    return out[0], out[1]


def tsc_snapshot_loops(x1, x2, L, m):
    """The same grid with the reference's per-particle loops, for tiny inputs."""
    w = L / m
    grids = []
    for x in (x1, x2):
        g = [[0.0] * m for _ in range(m)]
        for px, py in x:
            k, l = int(px / w) % m, int(py / w) % m
            ws = []
            for c, kk in ((px, k), (py, l)):
                d = c - (kk + 0.5) * w
                ws.append((0.5 * (1.5 - abs((d + w) / w)) ** 2, 0.75 - d * d / (w * w),
                           0.5 * (1.5 - abs((d - w) / w)) ** 2))
            for a, kk in enumerate(((k - 1 + m) % m, k, (k + 1) % m)):
                for b, ll in enumerate(((l - 1 + m) % m, l, (l + 1) % m)):
                    g[kk][ll] += ws[0][a] * ws[1][b] / (w * w)
        grids.append(np.array(g))
    return grids[0], grids[1]


def density_stats_ref(r1, r2, nbins, dbin, rho_thresh, drho_thresh):
    """Histograms and threshold fractions of a snapshot grid
    (main_auswertung.cpp:324-375): rows rho1, rho2, rho1+rho2, rho1-rho2;
    bin (int)(x/dbin - 0.5) clamped to [0, nbins-1], the difference shifted
    by nbins // 2; over = fractions with rho > t and |drho|/rho > t."""
    rs, dr = r1 + r2, r1 - r2
    pre = 1.0 / r1.size

    def bins(t):
        b = np.trunc(np.clip(t, -1.0, nbins)).astype(np.int64)
        return np.clip(b, 0, nbins - 1)

    ts = [r1 / dbin - 0.5, r2 / dbin - 0.5, rs / dbin - 0.5, dr / dbin - 0.5 + nbins // 2]
    hist = np.stack([np.bincount(bins(t).ravel(), minlength=nbins) for t in ts]) * pre
    with np.errstate(invalid="ignore", divide="ignore"):
        f = np.abs(dr) / rs
    over = [np.count_nonzero(rs > t) * pre for t in rho_thresh]
    over += [np.count_nonzero((rs > 0) & (f > t)) * pre for t in drho_thresh]
    return hist, np.array(over)
