"""CPU tests of the product library's host side: the C-ABI library loads and
exports every symbol include/bdk.h declares; host-side pieces that need no
GPU (parameter validation and errors, derived constants, MT stream and
init_zufallspos, disc noise, initial-state generators, step counting) are
checked against the oracle and against the mathematics.  No compute call
touches a GPU here."""
import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

import bdk
import oracle as O
from common import params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(bdk.LIB_PATH):
        subprocess.check_call(["python", os.path.join(ROOT, "bd-simulations_amd", "build.py")])
    O.build()


def header_functions():
    txt = open(os.path.join(ROOT, "include", "bdk.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b(bdk_[a-z_0-9]+)\s*\(", txt)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    declared = header_functions()
    assert len(declared) >= 30
    out = subprocess.check_output(["nm", "-D", "--defined-only", bdk.LIB_PATH]).decode()
    exported = set(re.findall(r" T (bdk_\w+)", out))
    missing = [n for n in declared if n not in exported]
    assert not missing, missing
    assert sorted(bdk.exported_symbols()) == declared
    L = bdk.lib()
    for n in declared:
        assert getattr(L, n) is not None


def test_derived_constants_match_oracle():
    for L_, Z in ((64.0, 128), (1414.0, 2048), (2000.0, 4096), (300.0, 384 * 0 + 256)):
        bp, op = params(L=L_, Z=Z)
        d = bdk.derived(bp)
        o = O.Oracle(op)
        assert d["nc"] == o.nc and d["B"] == o.B and d["h"] == o.hmesh and d["dq"] == o.dq
        assert d["drho"] == o.drho and d["zh16"] == o.zh16
    bp, _ = params(L=1414.0, Z=2048)
    assert bdk.derived(bp)["nc"] == 1259


@pytest.mark.parametrize("bad,msg", [
    (dict(Z=22), "FFT_samples"), (dict(Z=8192), "FFT_samples"), (dict(Z=26), "FFT_samples"),
    (dict(Z=6), "FFT_samples"), (dict(Z=105), "FFT_samples"), (dict(scheme=3), "Schema"),
    (dict(L=3.0), "3 list cells"), (dict(n1=0, n2=0), "n1+n2"), (dict(lam=0.0), "lambda"),
    (dict(dt_max=0.0), "dt_max"), (dict(precision=5), "precision"), (dict(noise_mode=9), "noise_mode"),
    (dict(max_travel=0.0), "max_travel"),
])
def test_create_rejects_invalid_parameters(bad, msg):
    bp, _ = params(**bad)
    with pytest.raises(bdk.BdkError) as e:
        bdk.RunDynamik(bp)
    assert e.value.code == bdk.ERR_ARG and msg in str(e.value)
    with pytest.raises(bdk.BdkError):
        bdk.footprint(bp)


@pytest.mark.parametrize("Z", [8, 10, 12, 14, 18, 24, 36, 54, 96, 100, 250, 384, 576, 1000, 1400, 3072, 4096])
def test_even_mesh_sizes_are_accepted(Z):
    """Z even and 2^a 3^b 5^c 7^d: the reference's default FFT_samples = 384
    (parameter.h:24) included; footprint() validates without a GPU."""
    bp, _ = params(Z=Z, L=max(40.0, 0.6 * Z))
    assert bdk.footprint(bp) > 0


def test_footprint_scales():
    small, _ = params(n1=1000, n2=1000, L=64.0, Z=128)
    big, _ = params(n1=10 ** 6, n2=10 ** 6, L=1414.0, Z=2048, precision=bdk.FP32)
    fs, fb = bdk.footprint(small), bdk.footprint(big)
    assert fs < 2 ** 22 < fb < 2 ** 30


def test_steps_for_interval_long_double():
    assert bdk.steps_for_interval(0.1, 1e-4) == 1001 == O.steps_for_interval(0.1, 1e-4)
    assert bdk.steps_for_interval(0.2, 1e-4) == 2001
    assert bdk.steps_for_interval(1.0, 1e-4) == 10001
    for obs, dt in ((0.2, 0.005), (0.3, 7e-3), (1e-3, 1e-4)):
        assert bdk.steps_for_interval(obs, dt) == O.steps_for_interval(obs, dt)


def test_mt_init_and_noise_match_oracle_bit_exact():
    """init_zufallspos (dm:1647-1678) and the disc noise (dm:1317-1339) from
    the product's host MT stream equal the oracle's restatement."""
    bp, op = params(n1=300, n2=200, L=64.0, Z=64, noise_mode=bdk.NOISE_UPLOAD)
    op.noise_mode = 0
    op.seed = 987654321
    mt = bdk.MT(987654321)
    git, rel = mt.init_positions(bp)
    o = O.Oracle(op)
    o.init_zufallspos()
    og, orl = o.all_positions()
    assert np.array_equal(git, og) and np.array_equal(rel, orl)
    for _ in range(3):
        n_host = mt.disc_noise(300, 200)
        n_or = o.noise()
        assert np.array_equal(n_host, n_or)


def test_init_uniform_exclusion_and_cells():
    bp, _ = params(n1=1000, n2=1000, L=64.0)
    git, rel = bdk.init_uniform(bp, 7, 0.9)
    d = bdk.derived(bp)
    x = git * d["B"] + rel
    assert np.all((rel >= 0) & (rel < d["B"]))
    assert np.all((git >= 0) & (git < d["nc"]))
    from scipy.spatial import cKDTree
    t = cKDTree(x, boxsize=bp.L)
    assert t.query_pairs(0.9 - 1e-12) == set()
    git2, rel2 = bdk.init_uniform(bp, 7, 0.9)
    assert np.array_equal(rel, rel2)


def test_init_blob_profile_clip_and_exclusion():
    bp, _ = params(n1=20000, n2=20000, L=283.0, Z=512)
    git, rel, clipped = bdk.init_blob(bp, 3, width=283.0 / 6, background=0.1, dmin=0.9)
    d = bdk.derived(bp)
    x = git * d["B"] + rel
    from scipy.spatial import cKDTree
    t = cKDTree(x, boxsize=bp.L)
    assert t.query_pairs(0.9 - 1e-9) == set()
    r = np.hypot(x[:, 0] - bp.L / 2, x[:, 1] - bp.L / 2)
    inner = np.mean(r < bp.L / 6)
    uniform_inner = math.pi / 36.0
    assert inner > 2 * uniform_inner
    # the core is clipped at the hexagonal-lattice density 2/(sqrt3 dmin^2)
    core = np.sum(r < bp.L / 12) / (math.pi * (bp.L / 12) ** 2)
    lattice = 2.0 / (math.sqrt(3.0) * 0.81)
    assert 0.9 * lattice < core <= 1.02 * lattice
    assert 0.0 < clipped < 1.0


def test_read_startpos_reference_format(tmp_path):
    """init_pos_aus_datei's format (dm:1576-1643): two header lines, species 1
    then 2 as 'nr typ x y', float-parsed; git = (int)(x/B), rel = x - git B
    (the reference's P8 computes rel from an uninitialised value)."""
    p = bdk.make_params(n1=3, n2=2, L=20.0, Z=16)
    B = bdk.derived(p)["B"]
    pts = [(1, 0, 1.25, 3.5), (1, 1, 19.75, 0.0), (1, 2, 10.0, 20.0), (2, 0, 5.5, 5.5), (2, 1, 0.3, 17.9)]
    f = tmp_path / "startpos.txt"
    f.write_text("# header\n\n" + "".join("%d \t %d \t %g \t %g\n" % (nr, t, x, y) for t, nr, x, y in pts))
    git, rel = bdk.read_startpos(p, str(f))
    for k, (t, nr, x, y) in enumerate(pts):
        xf, yf = float(np.float32(x)), float(np.float32(y))
        assert tuple(git[k]) == (int(xf / B), int(yf / B))
        assert rel[k, 0] == xf - git[k, 0] * B and rel[k, 1] == yf - git[k, 1] * B
    bad = tmp_path / "bad.txt"
    bad.write_text("#\n\n" + "".join("%d \t %d \t %g \t %g\n" % (nr, 2 if k == 1 else t, x, y)
                                      for k, (t, nr, x, y) in enumerate(pts)))
    with pytest.raises(bdk.BdkError):
        bdk.read_startpos(p, str(bad))
    out = tmp_path / "out.txt"
    out.write_text("#\n\n" + "".join("%d \t %d \t %g \t %g\n" % (nr, t, x + (30 if k == 3 else 0), y)
                                      for k, (t, nr, x, y) in enumerate(pts)))
    with pytest.raises(bdk.BdkError):
        bdk.read_startpos(p, str(out))
    with pytest.raises(bdk.BdkError):
        bdk.read_startpos(p, str(tmp_path / "missing.txt"))
