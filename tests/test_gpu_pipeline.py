"""The deposit that runs beside the list rebuild: after a step, the charge
mesh it left (deposited from the slot arrays, grouped by the previous cells,
halo widened by the update's displacement bound) equals bit for bit the mesh
that the list-order deposit computes for the same state, and so do the mesh
forces.  Cases: FP32 (stencil window) and FP64 (bounds-checked path), a box
small enough that a tile stages whole axes, and hot runs whose particles move
more than a cell per step (bound beyond the window's capacity, or within
it)."""
import numpy as np
import pytest

import bdk
from common import pair, params, positions
from helpers import rel_rms

pytestmark = pytest.mark.gpu

CASES = [
    # n1, n2, L, Z, scheme, prec, T, dt, steps
    (2000, 1500, 80.0, 128, bdk.TSC, bdk.FP32, 1.0, 1e-4, 30),
    (2000, 1500, 80.0, 128, bdk.CIC, bdk.FP64, 1.0, 1e-4, 30),
    (150, 100, 22.0, 32, bdk.TSC, bdk.FP32, 1.0, 1e-4, 20),
    (800, 800, 60.0, 64, bdk.NGP, bdk.FP32, 1.0, 1e-4, 20),
    (1500, 1000, 70.0, 128, bdk.TSC, bdk.FP32, 400.0, 2e-3, 10),
    (1500, 1000, 70.0, 128, bdk.CIC, bdk.FP64, 400.0, 2e-3, 10),
    # moves of about a cell: the window stays in use while particles cross
    # more than a cell downwards (their stored offset stays negative, fmod
    # keeps the sign, dm:216-325)
    (1500, 1000, 70.0, 128, bdk.TSC, bdk.FP32, 100.0, 2e-3, 10),
    (1500, 1000, 70.0, 128, bdk.CIC, bdk.FP64, 100.0, 2e-3, 10),
    # a coarse mesh (tiles of 52 x 52 list cells): one tile stages about
    # 1800 particles, more than the segment table holds (1536), and finds the
    # segments by binary search; the others use the table
    (2500, 2500, 100.0, 64, bdk.TSC, bdk.FP32, 1.0, 1e-4, 10),
]


@pytest.mark.parametrize("case", CASES)
def test_rebuild_side_deposit_equals_list_deposit(case):
    n1, n2, L, Z, scheme, prec, T, dt, steps = case
    print("case: pipeline deposit", case, flush=True)
    bp, _ = params(n1=n1, n2=n2, L=L, Z=Z, scheme=scheme, precision=prec, T=T, dt_max=dt, lam=20.0, seed=9)
    git, rel = positions(bp, 9, 0.9)
    s = bdk.RunDynamik(bp)
    s.set_positions(git, rel)
    s.run_steps(steps)
    rho_side = s.mesh(bdk.MESH_RHO)
    fx_side, fy_side = s.mesh(bdk.MESH_FX), s.mesh(bdk.MESH_FY)
    s.berechne_kapkraefte()  # list-order deposit, FFT passes, gather
    assert np.array_equal(rho_side, s.mesh(bdk.MESH_RHO))
    assert np.array_equal(fx_side, s.mesh(bdk.MESH_FX)) and np.array_equal(fy_side, s.mesh(bdk.MESH_FY))
    # the run stays consistent after the recomputation
    s.run_steps(3)
    g, r = s.positions()
    assert np.all(np.isfinite(r))
    s.close()


@pytest.mark.parametrize("scheme", [bdk.CIC, bdk.TSC])
def test_segment_search_and_table_vs_oracle(scheme):
    """The same coarse-mesh state (L = 100, Z = 64): the FP32 charge mesh and
    the capillary forces, whose tiles find their staged particles' row
    segments by binary search (the dense tile) or by the segment table,
    against the FP64 oracle within 1e-5 relative RMS."""
    print("case: coarse-mesh deposit vs oracle scheme=%d" % scheme, flush=True)
    g, o, git, rel = pair(n1=2500, n2=2500, L=100.0, Z=64, scheme=scheme, precision=bdk.FP32, lam=20.0, seed=13)
    fg = g.berechne_kapkraefte()
    fo = o.kap()  # fills the oracle's meshes
    assert rel_rms(g.mesh(bdk.MESH_RHO), o.mesh(0)) < 1e-5
    assert rel_rms(fg, fo) < 1e-5
    g.close()
    o.close()
