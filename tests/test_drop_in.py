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
REF_DEFAULT = os.path.join(HERE, "drop_in", "refdefault", "parameter.h")


def test_param_header_reaches_every_translation_unit():
    """Every translation unit of the driver that includes parameter.h picks
    the header given to driver_cmd (a quoted #include searches the including
    file's own directory first, so host/ must hold no parameter.h)."""
    cmd = B.driver_cmd(SRC, "unused", param_header=REF_DEFAULT)
    flags = [c for c in cmd[1:] if c.startswith(("-I", "-std"))]
    srcs = [c for c in cmd if c.endswith(".cpp")]
    seen = 0
    for src in srcs:
        out = subprocess.run(["g++", "-MM"] + flags + [src], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        deps = [os.path.abspath(d) for d in out.stdout.replace("\\\n", " ").split()[1:]]
        hdrs = [d for d in deps if os.path.basename(d) == "parameter.h"]
        assert all(h == os.path.abspath(REF_DEFAULT) for h in hdrs), (src, hdrs)
        seen += len(hdrs)
    assert seen >= 2  # params_from_header.cpp and the driver itself


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


@pytest.mark.gpu
def test_reference_default_driver_equals_binding(tmp_path):
    """The reference-call driver built against the reference's default values
    (tests/drop_in/refdefault/parameter.h: N1 = N2 = 2000, L = 300, TSC on
    384^2, lambda 25, q2/q1 = -1, kT = 24.3055555, adaptive dt with
    max_reisedistanz 0.24, 12 checkpoints of 0.2) writes the same
    pos1_1.txt / pos2_1.txt bytes as the same run through the Python binding:
    the reference's random start from the seeded mt19937_64, then per
    checkpoint zeitschritte_bis_obs and schreibe_pos."""
    import bdk
    print("case: reference defaults through the drop-in driver", flush=True)
    exe = os.path.join(PKG, "bdk_refdefault")
    assert os.path.exists(exe), "bdk_refdefault is not built (python bd-simulations_amd/build.py)"
    run_dir = tmp_path / "driver"
    run_dir.mkdir()
    env = dict(os.environ, BDK_SEED=str(SEED))
    out = subprocess.run([exe], cwd=str(run_dir), env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    p = bdk.make_params(n1=2000, n2=2000, L=300.0, Z=384, scheme=bdk.TSC, lam=25.0, kap=1.111111111,
                        T=24.3055555, dt_max=0.005, max_travel=0.24, f2_f1=-1.0, precision=bdk.FP32,
                        noise_mode=bdk.NOISE_DISC, seed=SEED)
    git, rel = bdk.MT(SEED).init_positions(p)  # run 0 draws from seed + run
    s = bdk.RunDynamik(p)
    s.set_positions(git, rel)
    q1, q2 = str(tmp_path / "q1.txt"), str(tmp_path / "q2.txt")
    s.write_pos(q1, q2, 0, 0.2)
    for nr in range(1, 12):
        s.zeitschritte_bis_obs(0.2)
        s.write_pos(q1, q2, nr, 0.2)
    s.close()
    got = [(run_dir / n).read_bytes() for n in ("pos1_1.txt", "pos2_1.txt")]
    assert got[0] == open(q1, "rb").read() and got[1] == open(q2, "rb").read()
    assert got[0].count(b"\n") == 2 + 12 * (2000 + 1)
