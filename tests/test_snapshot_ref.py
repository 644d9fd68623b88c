"""Pins the numpy restatement of the snapshot density grid and its
statistics (main_auswertung.cpp:214-375) that the device observable is
tested against: the vectorised grid equals the reference's per-particle
loops, a particle at a cell centre gives the closed-form TSC weights, every
species' grid integrates to its particle count, and the histogram and
threshold counts of a hand-made grid are known.  CPU only."""
import numpy as np

from helpers import density_stats_ref, tsc_snapshot_loops, tsc_snapshot_ref


def test_vectorised_grid_equals_loops():
    rng = np.random.default_rng(3)
    L, m = 9.0, 12
    x1, x2 = rng.uniform(0, L, (25, 2)), rng.uniform(0, L, (17, 2))
    x1[0] = (0.0, L - 1e-12)  # both edges of the box
    a1, a2 = tsc_snapshot_ref(x1, x2, L, m)
    b1, b2 = tsc_snapshot_loops(x1, x2, L, m)
    assert np.abs(a1 - b1).max() < 1e-12 and np.abs(a2 - b2).max() < 1e-12


def test_particle_at_cell_centre_closed_form():
    L, m = 10.0, 10
    w = L / m
    r1, r2 = tsc_snapshot_ref(np.array([[3.5, 6.5]]), np.zeros((0, 2)), L, m)
    want = np.zeros((m, m))
    side = [0.125, 0.75, 0.125]  # 3/4 - 0, (3/2 - 1)^2 / 2
    for a in range(3):
        for b in range(3):
            want[2 + a, 5 + b] = side[a] * side[b] / (w * w)
    assert np.allclose(r1, want, rtol=0, atol=1e-15) and not r2.any()


def test_grid_integrates_to_particle_count():
    rng = np.random.default_rng(4)
    L, m = 30.0, 64
    x1, x2 = rng.uniform(0, L, (500, 2)), rng.uniform(0, L, (300, 2))
    r1, r2 = tsc_snapshot_ref(x1, x2, L, m)
    w = L / m
    assert abs(r1.sum() * w * w - 500) < 1e-9 and abs(r2.sum() * w * w - 300) < 1e-9
    assert r1.min() >= 0 and r2.min() >= 0


def test_stats_of_hand_made_grid():
    # four cells: rho1 = [0, 0.05, 0.3, 2.0], rho2 = [0, 0.05, 0.0, 0.1]
    r1 = np.array([[0.0, 0.05], [0.3, 2.0]])
    r2 = np.array([[0.0, 0.05], [0.0, 0.1]])
    hist, over = density_stats_ref(r1, r2, 100, 0.01, (0.1, 0.2, 0.5), (0.1, 0.2, 0.5, 0.99))
    q = 0.25
    # rho1 bins: (int)(x/0.01 - 0.5): 0 -> -0.5 -> 0; 0.05 -> 4.5 -> 4; 0.3 -> 29; 2.0 -> 199.5 -> 99
    assert hist[0, 0] == q and hist[0, 4] == q and hist[0, 29] == q and hist[0, 99] == q
    # rho = [0, 0.1, 0.3, 2.1]: 0.1/0.01 - 0.5 = 9.5 -> 9
    assert hist[2, 0] == q and hist[2, 9] == q and hist[2, 29] == q and hist[2, 99] == q
    # drho = [0, 0, 0.3, 1.9] shifted by 50: 49.5 -> 49 (twice), 79.5 -> 79, 239.5 -> 99
    assert hist[3, 49] == 2 * q and hist[3, 79] == q and hist[3, 99] == q
    # rho > 0.1: 0.3, 2.1 (0.1 is not > 0.1); > 0.2: same; > 0.5: 2.1
    # |drho|/rho: cell 0 is 0/0 (counts nowhere), 0, 1, 0.905
    assert np.allclose(over, [0.5, 0.5, 0.25, 0.5, 0.5, 0.5, 0.25])
    assert np.allclose(hist.sum(1), 1.0)
