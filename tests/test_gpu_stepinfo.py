"""Step counters (bdk_step_info, the TimestepInfo of signaturen2.h:71-145):
largest forces per kind with their particles, candidate pair count, dt and
cell changes, against the FP64 oracle's forces and step on the same state;
the step itself equals an ordinary step."""
import numpy as np
import pytest

import bdk
from common import pair

pytestmark = pytest.mark.gpu


def argmax2(F):
    v = (F * F).sum(1)
    return int(np.argmax(v)), F[int(np.argmax(v))]


def test_step_info_against_oracle():
    print("case: step info", flush=True)
    n1 = 900
    g, o, git, rel = pair(n1=n1, n2=700, L=50.0, Z=64, scheme=bdk.TSC, seed=71, dmin=0.85)
    fw, fk, fn = o.wca(), o.kap(), o.noise()
    info = g.zeitschritt_debug(1.0)
    for (F, typ, idx, vec) in ((fw, info.max_wca_type, info.max_wca_index, info.max_wca),
                               (fk, info.max_kap_type, info.max_kap_index, info.max_kap),
                               (fn, info.max_rnd_type, info.max_rnd_index, info.max_rnd)):
        k, v = argmax2(F)
        assert (typ, idx) == ((1, k) if k < n1 else (2, k - n1))
        assert np.allclose(list(vec), v, rtol=1e-10, atol=1e-12)
    # candidate pairs as the reference's loops visit them (dm:704-1078):
    # species 1 against lists 1 and 2, species 2 against list 2
    nc = g.nc
    def around(sel):
        cnt = np.bincount(git[sel, 0] * nc + git[sel, 1], minlength=nc * nc).reshape(nc, nc)
        return sum(np.roll(np.roll(cnt, a, 0), b, 1) for a in (-1, 0, 1) for b in (-1, 0, 1))
    s1, s2 = np.arange(len(git)) < n1, np.arange(len(git)) >= n1
    a_all, a_2 = around(slice(None)), around(s2)
    want = (a_all[git[s1, 0], git[s1, 1]] - 1).sum() + (a_2[git[s2, 0], git[s2, 1]] - 1).sum()
    assert info.num_wca == int(want)
    k, _ = argmax2(fw)
    cx, cy = git[k]
    cells = [((cx + a) % nc, (cy + b) % nc) for a in (-1, 0, 1) for b in (-1, 0, 1)]
    inside = np.array([(gx, gy) in cells for gx, gy in map(tuple, git)])
    assert (info.max_wca_nlist[0], info.max_wca_nlist[1]) == (int(inside[:n1].sum()), int(inside[n1:].sum()))
    assert info.dt == 1e-4
    o.step(1.0)
    og, _ = o.all_positions()
    moved = np.any(og != git, axis=1)
    assert (info.cell_changes[0], info.cell_changes[1]) == (int(moved[:n1].sum()), int(moved[n1:].sum()))
    gg, gr = g.positions()
    og, orl = o.all_positions()
    assert np.array_equal(gg, og) and np.abs(gr - orl).max() < 1e-12
