"""The C++ drop-in driver (bd-simulations_amd/bdk_sim: main.cpp:130-300's run
loop over RunZustand, built against host/small_run/parameter.h: N1 = N2 = 1000, L = 64,
CIC 128^2, lambda 8, dt_max 1e-4, obs_dt 0.1, 11 checkpoints) writes the same
pos1_1.txt / pos2_1.txt bytes as the same run driven through the Python
binding: reference start positions from the seeded mt19937_64 (init_zufallspos,
dm:1647-1678), then per checkpoint zeitschritte_bis_obs and schreibe_pos
(au:631-648).  Fixed dt, and adaptive dt with a travel cap (BDK_MAX_TRAVEL),
whose interval runs from the step graphs with the long-double clock."""
import os
import subprocess

import pytest

import bdk

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "bd-simulations_amd", "bdk_sim")
SEED = 12345


def read(path):
    with open(path, "rb") as f:
        return f.read()


@pytest.mark.parametrize("travel", [None, 0.02])
def test_driver_files_equal_binding_run(tmp_path, travel):
    print("case: C++ driver, max_travel=%s" % travel, flush=True)
    assert os.path.exists(EXE), "bdk_sim is not built (python bd-simulations_amd/build.py)"
    env = dict(os.environ, BDK_SEED=str(SEED))
    if travel is not None:
        env["BDK_MAX_TRAVEL"] = repr(travel)
    run_dir = tmp_path / "driver"
    run_dir.mkdir()
    out = subprocess.run([EXE], cwd=str(run_dir), env=env, capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stdout + out.stderr
    kw = dict(n1=1000, n2=1000, L=64.0, Z=128, scheme=bdk.CIC, lam=8.0, kap=1.0, T=1.0, dt_max=1e-4, f2_f1=0.5,
              precision=bdk.FP32, noise_mode=bdk.NOISE_DISC, seed=SEED)
    if travel is not None:
        kw["max_travel"] = travel
    p = bdk.make_params(**kw)
    git, rel = bdk.MT(SEED + 0).init_positions(p)  # run 0 draws from seed + run
    s = bdk.RunDynamik(p)
    s.set_positions(git, rel)
    q1, q2 = str(tmp_path / "q1.txt"), str(tmp_path / "q2.txt")
    s.write_pos(q1, q2, 0, 0.1)
    steps = []
    for nr in range(1, 11):
        steps.append(s.zeitschritte_bis_obs(0.1))
        s.write_pos(q1, q2, nr, 0.1)
    s.close()
    if travel is None:
        assert steps == [1001] * 10
    assert read(str(run_dir / "pos1_1.txt")) == read(q1)
    assert read(str(run_dir / "pos2_1.txt")) == read(q2)


def test_driver_force_scan_equals_binding(tmp_path):
    """krafttestmode (main.cpp:220-248): kapkraft.txt of the driver equals the
    same scan through the binding (particle 1 moved to y = x_0 + t, wrapped
    into the box; F1kap[1][1] and F1kap[1][1] + F1_WCA[1][1] written)."""
    import math
    print("case: C++ driver, krafttestmode", flush=True)
    assert os.path.exists(EXE), "bdk_sim is not built (python bd-simulations_amd/build.py)"
    env = dict(os.environ, BDK_SEED=str(SEED), BDK_MODE="krafttest")
    out = subprocess.run([EXE], cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stdout + out.stderr
    p = bdk.make_params(n1=1000, n2=1000, L=64.0, Z=128, scheme=bdk.CIC, lam=8.0, kap=1.0, T=1.0, dt_max=1e-4,
                        f2_f1=0.5, precision=bdk.FP32, noise_mode=bdk.NOISE_DISC, seed=SEED)
    B = bdk.derived(p)["B"]
    git, rel = bdk.MT(SEED).init_positions(p)
    s = bdk.RunDynamik(p)
    lines = ["#Format: r Fkap Ftotal", ""]
    t = 0.001
    while t < 5.0:
        origin = git[0][0] * B + rel[0][0]
        y = math.fmod(origin + t, 64.0)
        git[1][1] = int(y / B)
        rel[1][1] = y - git[1][1] * B
        s.set_positions(git, rel)
        fk = s.berechne_kapkraefte()
        fw = s.berechne_WCA()
        lines.append("%g \t %g \t %g " % (t, fk[1][1], fk[1][1] + fw[1][1]))
        t += 0.01
    s.close()
    got = (tmp_path / "kapkraft.txt").read_text().split("\n")
    assert got[:-1] == lines
    assert len(lines) == 2 + 500


def test_driver_mrd_sweep(tmp_path):
    """debugmode (main.cpp:156-216): for mrd = 1, 0.95, ... > 0.01 the mean
    and largest |F_WCA|max and the mean dt over 50 steps early and late; dt
    stays within dt_max and shrinks as the travel cap does.  The sweep's
    fixed-dt steps from overlapping starts (forces near maxPairWCA) move
    particles by many cells; it stopped with a deposit-window bound error
    while the rebuild-side bound ignored the extra cell of a negative stored
    offset (fmod in the re-binning, dm:216-325) and runs through since."""
    print("case: C++ driver, debugmode", flush=True)
    assert os.path.exists(EXE), "bdk_sim is not built (python bd-simulations_amd/build.py)"
    env = dict(os.environ, BDK_SEED=str(SEED), BDK_MODE="debug")
    out = subprocess.run([EXE], cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr
    mrds = []
    m = 1.0
    while m > 0.01:
        mrds.append(m)
        m *= 0.95
    for name in ("mrd_data1.txt", "mrd_data2.txt"):
        rows = [[float(v) for v in ln.split()] for ln in (tmp_path / name).read_text().split("\n")
                if ln.strip() and not ln.startswith("#")]
        assert len(rows) == len(mrds)
        for (mrd, f_mean, f_high, dt), want in zip(rows, mrds):
            assert abs(mrd - want) <= 1e-5 * want
            assert 0.0 < dt <= 1e-4 * (1 + 1e-5)
            assert 0.0 < f_mean <= f_high * (1 + 1e-5)
        assert rows[-1][3] < rows[0][3]  # a small cap shortens the steps


def test_driver_pair_correlation_file(tmp_path):
    """auswerten_korrfunk(_mixed) (BDK_KORRFUNK=mixed): obs_point records
    g11, g12, g22 at every checkpoint (au:81-83, au:110-188) and the run ends
    with g_run0.txt (au:953-975); equal to the same records through the
    binding."""
    print("case: C++ driver, g(r) file", flush=True)
    assert os.path.exists(EXE), "bdk_sim is not built (python bd-simulations_amd/build.py)"
    env = dict(os.environ, BDK_SEED=str(SEED), BDK_KORRFUNK="mixed")
    out = subprocess.run([EXE], cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stdout + out.stderr
    p = bdk.make_params(n1=1000, n2=1000, L=64.0, Z=128, scheme=bdk.CIC, lam=8.0, kap=1.0, T=1.0, dt_max=1e-4,
                        f2_f1=0.5, precision=bdk.FP32, noise_mode=bdk.NOISE_DISC, seed=SEED)
    git, rel = bdk.MT(SEED).init_positions(p)
    s = bdk.RunDynamik(p)
    s.set_positions(git, rel)
    nb, dr = int(32.0 / 0.2) + 1, 0.2
    lines = ["# Format: t TAB r TAB g11 TAB g12 TAB g22", ""]
    for o in range(11):
        if o:
            s.zeitschritte_bis_obs(0.1)
        g11, g12, g22 = s.pair_correlation(nb, dr)
        lines += ["%g \t %g \t %g \t %g \t %g " % (o * 0.1, b * dr, g11[b], g12[b], g22[b]) for b in range(nb)]
        lines.append("")
    s.close()
    got = (tmp_path / "g_run0.txt").read_text().split("\n")
    assert got[:-1] == lines


def test_driver_rho_k_file(tmp_path):
    """auswerten_rhoviaFFTW (BDK_RHOFFTW=1): obs_point records the rho(k)
    shells of both species (record_rhoFFTW, au:527-626) with the reference's
    dq_rhoFFTW = 2 dq and rhoFFTW_bins = Z/4, and the run ends with
    rhoFFT_run0.txt (au:994-1014); equal to the same records through the
    binding."""
    import math
    print("case: C++ driver, rho(k) file", flush=True)
    assert os.path.exists(EXE), "bdk_sim is not built (python bd-simulations_amd/build.py)"
    env = dict(os.environ, BDK_SEED=str(SEED), BDK_RHOFFTW="1")
    out = subprocess.run([EXE], cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stdout + out.stderr
    p = bdk.make_params(n1=1000, n2=1000, L=64.0, Z=128, scheme=bdk.CIC, lam=8.0, kap=1.0, T=1.0, dt_max=1e-4,
                        f2_f1=0.5, precision=bdk.FP32, noise_mode=bdk.NOISE_DISC, seed=SEED)
    git, rel = bdk.MT(SEED).init_positions(p)
    s = bdk.RunDynamik(p)
    s.set_positions(git, rel)
    dq = 2.0 * math.pi / 64.0
    dq_r = 2.0 * dq
    nb = int(0.5 * 128 * dq / dq_r)
    lines = ["# Format: t TAB q TAB rho1_re TAB rho1_im TAB rho2_re TAB rho2_im ", ""]
    for o in range(11):
        if o:
            s.zeitschritte_bis_obs(0.1)
        r1, i1 = s.rho_k(1, dq / dq_r, nb)
        r2, i2 = s.rho_k(2, dq / dq_r, nb)
        lines += ["%g \t %g \t %g \t %g \t %g \t %g " % (o * 0.1, dq_r * i, r1[i], i1[i], r2[i], i2[i])
                  for i in range(nb)]
        lines.append("")
    s.close()
    got = (tmp_path / "rhoFFT_run0.txt").read_text().split("\n")
    assert got[:-1] == lines


def test_driver_abstand_file(tmp_path):
    """auswerten_abstand (BDK_ABSTAND=1): the mean distance of the species-2
    particles from the box centre at every checkpoint (schreibe_abstaende,
    au:651-670), equal to the same means of the binding's positions."""
    import math
    print("case: C++ driver, abstaende2.txt", flush=True)
    env = dict(os.environ, BDK_SEED=str(SEED), BDK_ABSTAND="1")
    out = subprocess.run([EXE], cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stdout + out.stderr
    p = bdk.make_params(n1=1000, n2=1000, L=64.0, Z=128, scheme=bdk.CIC, lam=8.0, kap=1.0, T=1.0, dt_max=1e-4,
                        f2_f1=0.5, precision=bdk.FP32, noise_mode=bdk.NOISE_DISC, seed=SEED)
    git, rel = bdk.MT(SEED).init_positions(p)
    s = bdk.RunDynamik(p)
    s.set_positions(git, rel)
    lines = []
    for o in range(11):
        if o:
            s.zeitschritte_bis_obs(0.1)
        g, r = s.positions()
        mean = 0.0
        for i in range(1000, 2000):  # the reference's loop order and arithmetic
            dx = r[i][0] + s.B * g[i][0] - 32.0
            dy = r[i][1] + s.B * g[i][1] - 32.0
            mean += math.sqrt(dx * dx + dy * dy)
        lines.append("%g \t %g " % (o * 0.1, mean / 1000))
    s.close()
    got = [ln for ln in (tmp_path / "abstaende2.txt").read_text().split("\n") if ln and not ln.startswith("#")]
    assert got == lines


def test_driver_ftrho_file(tmp_path):
    """auswerten_rhovonk (BDK_RHOVONK=1): direct rho(k) sums at every
    checkpoint, shells by i^2 + j^2 in order of first appearance (init_ftrho,
    au:262-300), written to ftrho_run0.txt (au:977-992, q = dq ftrho_dq
    sqrt(i^2 + j^2) as the reference writes it); equal to the same records
    through the binding."""
    import math
    print("case: C++ driver, ftrho file", flush=True)
    env = dict(os.environ, BDK_SEED=str(SEED), BDK_RHOVONK="1")
    out = subprocess.run([EXE], cwd=str(tmp_path), env=env, capture_output=True, text=True, timeout=180)
    assert out.returncode == 0, out.stdout + out.stderr
    p = bdk.make_params(n1=1000, n2=1000, L=64.0, Z=128, scheme=bdk.CIC, lam=8.0, kap=1.0, T=1.0, dt_max=1e-4,
                        f2_f1=0.5, precision=bdk.FP32, noise_mode=bdk.NOISE_DISC, seed=SEED)
    git, rel = bdk.MT(SEED).init_positions(p)
    s = bdk.RunDynamik(p)
    s.set_positions(git, rel)
    qb, dq, fdq = 30, 2.0 * math.pi / 64.0, 4.0 * math.pi / 64.0
    order, qabs2 = [], []
    for i in range(qb):
        for j in range(qb):
            ind2 = i * i + j * j
            if ind2 > qb * qb:
                break
            if ind2 not in qabs2:
                qabs2.append(ind2)
            order.append(qabs2.index(ind2))
    lines = ["# Format: t TAB q TAB rho1_re TAB rho1_im TAB rho2_re TAB rho2_im ", ""]
    for o in range(11):
        if o:
            s.zeitschritte_bis_obs(0.1)
        acc = []
        for sp in (1, 2):
            re, im = s.ftrho_sums(sp, qb, fdq)
            R, I = [0.0] * len(qabs2), [0.0] * len(qabs2)
            for k, b in enumerate(order):
                R[b] += re[k]
                I[b] += im[k]
            acc.append((R, I))
        for i in range(len(qabs2)):
            lines.append("%g \t %g \t %g \t %g \t %g \t %g " % (o * 0.1, dq * fdq * math.sqrt(qabs2[i]), acc[0][0][i],
                                                            acc[0][1][i], acc[1][0][i], acc[1][1][i]))
        lines.append("")
    s.close()
    got = (tmp_path / "ftrho_run0.txt").read_text().split("\n")
    assert got[:-1] == lines
