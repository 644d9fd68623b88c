"""Drop-in check of the C++ host shim: a driver that uses only the
reference's calls and parameter.h globals (tests/drop_in/
reference_calls_main.cpp: `RunZustand theRun;`, ausgabe_jeansgroessen,
init, obs_point, zeitschritte_bis_obs, schritte_seit_obs, obs_nr,
schreibe_obs, dyn.zeitschritt_debug with TimestepInfo, the force scan that
edits r1_git/r1_rel and calls dyn.berechne_kapkraefte / berechne_WCA11 /
getFKap / getF, as main.cpp:130-300 does) compiles against
bd-simulations_amd/host unchanged, and on the GPU writes the same
pos1_1.txt / pos2_1.txt bytes as the build's own driver bdk_sim."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "bd-simulations_amd")
SRC = os.path.join(HERE, "drop_in", "reference_calls_main.cpp")
SEED = 24680

sys.path.insert(0, PKG)
import build as B  # noqa: E402


def test_reference_call_driver_compiles(tmp_path):
    """Compiles (and links, when libbdk_hip.so is built) with -Wall -Werror."""
    exe = str(tmp_path / "refcalls")
    cmd = B.driver_cmd(SRC, exe, extra=["-Werror"])
    if not os.path.exists(B.LIB):
        # compile only: every translation unit, no link
        cmd = [c for c in cmd if not c.startswith(("-L", "-l", "-Wl", "-o")) and c != exe] + ["-fsyntax-only"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr


REF_PARAM = "/root/reference/parameter.h"


@pytest.mark.skipif(not os.path.exists(REF_PARAM), reason="the reference tree is not on this machine")
def test_reference_call_driver_compiles_against_reference_parameter_h(tmp_path):
    """The same driver against the reference's own parameter.h as it lies in
    the reference tree (N1 = N2 = 2000, L = 300, TSC, FFT_samples = 384,
    adaptive dt): the header a reference user already has configures the
    build, and its mesh size is one the library accepts."""
    exe = str(tmp_path / "refcalls_ref")
    cmd = B.driver_cmd(SRC, exe, param_header=REF_PARAM)
    if not os.path.exists(B.LIB):
        cmd = [c for c in cmd if not c.startswith(("-L", "-l", "-Wl", "-o")) and c != exe] + ["-fsyntax-only"]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr


@pytest.mark.gpu
def test_reference_call_driver_equals_bdk_sim(tmp_path):
    print("case: reference-call driver vs bdk_sim", flush=True)
    exes = [os.path.join(PKG, "bdk_refcalls"), os.path.join(PKG, "bdk_sim")]
    for e in exes:
        assert os.path.exists(e), "%s is not built (python bd-simulations_amd/build.py)" % e
    env = dict(os.environ, BDK_SEED=str(SEED))
    files = []
    for k, e in enumerate(exes):
        d = tmp_path / ("run%d" % k)
        d.mkdir()
        out = subprocess.run([e], cwd=str(d), env=env, capture_output=True, text=True, timeout=180)
        assert out.returncode == 0, out.stdout + out.stderr
        assert "Jeans-Zeit" in out.stdout
        files.append([(d / n).read_bytes() for n in ("pos1_1.txt", "pos2_1.txt")])
    assert files[0] == files[1]
    # 11 checkpoints of 2000 particles (and a blank line each) after the header
    assert files[0][0].count(b"\n") == 2 + 11 * (1000 + 1)
