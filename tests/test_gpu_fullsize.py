"""GPU checks at the benchmark size (2M particles, L = 1414, TSC 2048^2):
size-independent properties, plus force parity against the FP64 oracle on
the full initial state (the oracle needs a few seconds for one evaluation)."""
import numpy as np
import pytest

import bdk
import oracle as O
from helpers import rel_rms

pytestmark = pytest.mark.gpu

N1 = 10 ** 6
L, Z = 1414.0, 2048


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def start():
    p = bdk.make_params(n1=N1, n2=N1, L=L, Z=Z, scheme=bdk.TSC, lam=50.0, kap=1.0, f2_f1=0.5,
                        precision=bdk.FP32, seed=1)
    git, rel, clipped = bdk.init_blob(p, 1, L / 6, 0.1, 0.95)
    return git, rel


def sim(git, rel, prec=bdk.FP32, seed=1):
    p = bdk.make_params(n1=N1, n2=N1, L=L, Z=Z, scheme=bdk.TSC, lam=50.0, kap=1.0, f2_f1=0.5, precision=prec,
                        seed=seed)
    s = bdk.RunDynamik(p)
    s.set_positions(git, rel)
    return s


def test_fullsize_forces_vs_oracle(start):
    git, rel = start
    say("case: full-size forces vs oracle")
    s = sim(git, rel)
    fw, fk = s.berechne_WCA(), s.berechne_kapkraefte()
    op = O.make_params(n1=N1, n2=N1, L=L, Z=Z, scheme=2, lam=50.0, kap=1.0, f2_f1=0.5, noise_mode=1, seed=1)
    o = O.Oracle(op)
    o.set_all(git, rel)
    assert rel_rms(fw, o.wca()) < 1e-5
    assert rel_rms(fk, o.kap()) < 1e-5
    s64 = sim(git, rel, bdk.FP64)
    assert rel_rms(fw, s64.berechne_WCA()) < 1e-5
    assert rel_rms(fk, s64.berechne_kapkraefte()) < 1e-5


def test_fullsize_invariants_after_steps(start):
    git, rel = start
    say("case: full-size invariants after 20 steps")
    s = sim(git, rel)
    s.run_steps(20)
    g, r = s.positions()
    start_, ids = s.cells()
    cell = g[:, 0] * s.nc + g[:, 1]
    assert np.array_equal(np.diff(start_), np.bincount(cell, minlength=s.nc * s.nc))
    assert np.array_equal(ids, np.argsort(cell, kind="stable"))
    assert np.all((r >= 0) & (r <= s.B))
    fw = s.berechne_WCA()
    # Newton 3 over all pairs (sigma and eps ratios 1)
    assert np.abs(fw.sum(0)).max() < 1e-4 * np.abs(fw).sum()
    s.berechne_kapkraefte()
    rho = s.mesh(bdk.MESH_RHO)
    h = L / Z
    assert rho.sum() * h * h == pytest.approx(N1 + 0.5 * N1, rel=1e-6)


def test_fullsize_replay_is_bit_reproducible(start):
    git, rel = start
    say("case: full-size reproducibility")
    a, b = sim(git, rel), sim(git, rel)
    a.run_steps(13)
    for _ in range(13):
        b.zeitschritt(1.0)
    pa, pb = a.positions(), b.positions()
    assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])


def test_fullsize_stable_over_300_steps(start):
    git, rel = start
    say("case: full-size 300 steps")
    s = sim(git, rel)
    s.run_steps(300)
    assert np.abs(s.berechne_WCA()).max() < 1e4
