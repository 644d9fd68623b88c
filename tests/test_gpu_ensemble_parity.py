"""Ensemble parity of the noisy path: the device's Philox disc noise and the
reference's own noise (the FP64 oracle drawing from mt19937_64 with the
rejection sampler of dm:1317-1339, zufall.cpp:15-39) give the same
statistics of a collapsing blob.

48 replicas of one blob start (300 particles, CIC, lambda = 10, kT = 1) run
400 steps on the device (one batch, a Philox seed per replica) and on the
host oracle (an mt19937_64 seed per replica).  Per replica: the mean squared
displacement of each species, the mean squared distance of each species from
its circular-mean centre (berechne_mittelpunkte, observablen_auswertung.cpp:
140-223), and the first bins of the species-1 radial profile around the
common centre (berechne_dichteprofile, :227-330), and shells 1-3 of the
species-1 rho(k) (record_rhoFFTW, auswertung.cpp:529-626: the device
observable on the device ensemble, its numpy restatement on the host
ensemble).  Tolerance: every ensemble
mean agrees within 4.5 combined standard errors.  A noise variance of 1
instead of 5/4 moves the displacement means by about 15 standard errors.
"""
import numpy as np
import pytest

import bdk
import oracle as O
from common import params
from helpers import min_image, radial_profile_ref, rho_k_ref

pytestmark = pytest.mark.gpu

M, STEPS, NB, DR = 48, 400, 4, 1.0
DQR, NK = 1.0, 4  # rho(k) shells 1-3 of species 1 (record_rhoFFTW, auswertung.cpp:529-626)
SPEC = dict(n1=150, n2=150, L=30.0, Z=64, scheme=bdk.CIC, lam=10.0, kap=3.0, T=1.0, dt_max=2e-4,
            precision=bdk.FP64)


def centre(x, L):
    k = 2 * np.pi / L
    return np.array([(np.arctan2(-np.sin(x[:, a] * k).sum(), -np.cos(x[:, a] * k).sum()) + np.pi) / k
                     for a in range(2)])


def stats(x0, x, n1, L, rhok):
    d = min_image(x - x0, L)
    out = [np.mean((d[:n1] ** 2).sum(1)), np.mean((d[n1:] ** 2).sum(1))]
    for sp in (x[:n1], x[n1:]):
        r = min_image(sp - centre(sp, L), L)
        out.append(np.mean((r ** 2).sum(1)))
    _, prof = radial_profile_ref(x[:n1], x[n1:], L, NB, DR)
    out += list(prof[0][:3])
    out += list(rhok)
    return np.array(out)


def test_philox_ensemble_matches_reference_mt_ensemble():
    print("case: ensemble parity, %d replicas x %d steps" % (M, STEPS), flush=True)
    bp0, _ = params(seed=1, **SPEC)
    git, rel, _ = bdk.init_blob(bp0, 5, SPEC["L"] / 4, 0.1, 0.95)
    B = bdk.derived(bp0)["B"]
    x0 = git * B + rel
    L, n1 = SPEC["L"], SPEC["n1"]
    sims = []
    for r in range(M):
        bp, _ = params(seed=1000 + r, **SPEC)
        s = bdk.RunDynamik(bp)
        s.set_positions(git, rel)
        sims.append(s)
    b = bdk.Batch(sims)
    b.run_steps(STEPS)
    b.sync()
    dev = []
    for s in sims:
        g, rr = s.positions()
        re, _ = s.rho_k(1, DQR, NK)  # the device observable (record_rhoFFTW)
        dev.append(stats(x0, g * B + rr, n1, L, re[1:NK]))
        s.close()
    b.close()
    host = []
    for r in range(M):
        _, op = params(seed=5000 + r, **SPEC)
        op.noise_mode = 0  # the reference's mt19937_64 disc sampler
        o = O.Oracle(op)
        o.set_all(git, rel)
        for _ in range(STEPS):
            o.step(1.0)
        g, rr = o.all_positions()
        xh = g * o.B + rr
        rk = rho_k_ref(xh[:n1], L, SPEC["Z"], SPEC["scheme"], DQR, NK)
        host.append(stats(x0, xh, n1, L, rk[1:NK]))
        o.close()
    dev, host = np.array(dev), np.array(host)
    names = ["msd1", "msd2", "rg2_1", "rg2_2", "rho1_b0", "rho1_b1", "rho1_b2", "rhok1_s1", "rhok1_s2",
             "rhok1_s3"]
    for k, nm in enumerate(names):
        md, mh = dev[:, k].mean(), host[:, k].mean()
        se = np.sqrt(dev[:, k].var(ddof=1) / M + host[:, k].var(ddof=1) / M)
        print("%-8s device %.5g  reference-noise %.5g  diff/se %.2f" % (nm, md, mh, (md - mh) / se), flush=True)
        assert abs(md - mh) <= 4.5 * se, nm
    # the displacement is noise-dominated here: sanity of the scale
    # (2 dt kT E|n|^2 per step, E|n|^2 = 5/2 for the disc sampler)
    free = STEPS * 2 * SPEC["dt_max"] * SPEC["T"] * 2.5
    assert 0.5 * free < host[:, 0].mean() < 2.0 * free
