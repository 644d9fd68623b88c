"""Python host binding of libbdk_hip.so (the C-ABI of include/bdk.h).

`RunDynamik` mirrors RunZustand::RunDynamik (signaturen2.h:33-252): same
entry points (init, zeitschritt, berechne_WCA*, berechne_kapkraefte,
berechne_zufallskraefte, optimaler_zeitschritt, refresh_erwNachbar, getFKap,
getF), argument meaning and error behaviour, over the HIP library.  The HIP
library is required: loading fails loudly if it is missing or no HIP device
is present; there is no CPU fallback.
"""
import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbdk_hip.so")

NGP, CIC, TSC = 0, 1, 2
FP64, FP32, FP32_POS32, FP32_POS32K = 0, 1, 2, 3
NOISE_DISC, NOISE_GAUSS, NOISE_UPLOAD = 0, 1, 2
GREEN_DISCRETE, GREEN_CONTINUOUS = 0, 1
MESH_RHO, MESH_FX, MESH_FY = 0, 1, 2

OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_NOMEM, ERR_NONFINITE = range(6)


class BdkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("bdk error %d: %s" % (code, msg))
        self.code = code


class Params(C.Structure):
    """bd_params of include/bdk.h (parameter.h:14-99 user constants)."""
    _fields_ = [
        ("n1", C.c_int32), ("n2", C.c_int32), ("scheme", C.c_int32), ("Z", C.c_int32),
        ("L", C.c_double), ("lam", C.c_double), ("kap", C.c_double), ("T", C.c_double),
        ("dt_max", C.c_double), ("max_travel", C.c_double),
        ("sigma11_22", C.c_double), ("sigma11_12", C.c_double), ("eps22_11", C.c_double),
        ("eps12_11", C.c_double), ("f2_f1", C.c_double), ("max_pair_wca", C.c_double),
        ("idgas_only", C.c_int32), ("no_wca", C.c_int32), ("restrict_radial", C.c_int32),
        ("ngp_fix", C.c_int32), ("precision", C.c_int32), ("noise_mode", C.c_int32),
        ("green_mode", C.c_int32), ("device", C.c_int32),
        ("noise_var", C.c_double), ("seed", C.c_uint64), ("ftrho_fix", C.c_int32),
    ]


class StepInfo(C.Structure):
    """bdk_step_info_t (TimestepInfo, signaturen2.h:71-145)."""
    _fields_ = [("num_wca", C.c_int64), ("max_wca", C.c_double * 2), ("max_wca_type", C.c_int32),
                ("max_wca_index", C.c_int32), ("max_wca_nlist", C.c_int32 * 2), ("max_kap", C.c_double * 2),
                ("max_kap_type", C.c_int32), ("max_kap_index", C.c_int32), ("max_rnd", C.c_double * 2),
                ("max_rnd_type", C.c_int32), ("max_rnd_index", C.c_int32), ("dt", C.c_double),
                ("cell_changes", C.c_int32 * 2)]


_lib = None


def lib():
    """Loads libbdk_hip.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libbdk_hip.so not built: run `python bd-simulations_amd/build.py`")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, dbl, u64 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint64
        sig = {
            "bdk_params_default": (C.c_int, [vp]),
            "bdk_params_derived": (C.c_int, [vp, vp]),
            "bdk_create": (C.c_int, [vp, C.POINTER(vp)]),
            "bdk_destroy": (None, [vp]),
            "bdk_last_error": (C.c_char_p, [vp]),
            "bdk_set_positions": (C.c_int, [vp, C.c_int, i32, vp, vp]),
            "bdk_get_positions": (C.c_int, [vp, C.c_int, i32, vp, vp]),
            "bdk_compute_wca": (C.c_int, [vp, vp, vp]),
            "bdk_compute_capillary": (C.c_int, [vp, vp, vp]),
            "bdk_set_noise": (C.c_int, [vp, vp, vp]),
            "bdk_get_noise": (C.c_int, [vp, vp, vp]),
            "bdk_optimal_dt": (C.c_int, [vp, vp]),
            "bdk_step": (C.c_int, [vp, dbl, vp]),
            "bdk_run_steps": (C.c_int, [vp, i64, vp]),
            "bdk_run_interval": (C.c_int, [vp, dbl, vp]),
            "bdk_run_interval_clock": (C.c_int, [vp, dbl, vp, vp, vp]),
            "bdk_set_max_travel": (C.c_int, [vp, dbl]),
            "bdk_steps_for_interval": (C.c_int, [dbl, dbl, vp]),
            "bdk_refresh": (C.c_int, [vp]),
            "bdk_sync": (C.c_int, [vp]),
            "bdk_prepare": (C.c_int, [vp]),
            "bdk_step_counter": (i64, [vp]),
            "bdk_get_cells": (C.c_int, [vp, vp, vp]),
            "bdk_get_mesh": (C.c_int, [vp, C.c_int, vp]),
            "bdk_footprint": (C.c_int, [vp, vp]),
            "bdk_profile_steps": (C.c_int, [vp, i64, i32, vp, C.c_char_p, i32, vp]),
            "bdk_profile_kernel": (C.c_int, [vp, C.c_int, i32, vp]),
            "bdk_mt_create": (C.c_int, [u64, C.POINTER(vp)]),
            "bdk_mt_destroy": (None, [vp]),
            "bdk_mt_uniform": (dbl, [vp, dbl, dbl]),
            "bdk_mt_init_positions": (C.c_int, [vp, vp, vp, vp, vp, vp]),
            "bdk_mt_disc_noise": (C.c_int, [vp, i32, i32, vp, vp]),
            "bdk_init_uniform": (C.c_int, [vp, u64, dbl, vp, vp]),
            "bdk_init_blob": (C.c_int, [vp, u64, dbl, dbl, dbl, vp, vp, vp]),
            "bdk_write_pos": (C.c_int, [vp, C.c_char_p, C.c_char_p, i32, dbl]),
            "bdk_write_pos_async": (C.c_int, [vp, C.c_char_p, C.c_char_p, i32, dbl]),
            "bdk_density_grid": (C.c_int, [vp, i32, vp, vp]),
            "bdk_ftrho_sums": (C.c_int, [vp, i32, i32, dbl, vp, vp, vp]),
            "bdk_density_stats": (C.c_int, [vp, i32, i32, dbl, vp, vp, vp, vp]),
            "bdk_write_wait": (C.c_int, [vp]),
            "bdk_step_info": (C.c_int, [vp, dbl, vp]),
            "bdk_read_startpos": (C.c_int, [vp, C.c_char_p, vp, vp]),
            "bdk_save_state": (C.c_int, [vp, C.c_char_p]),
            "bdk_load_state": (C.c_int, [vp, C.c_char_p]),
            "bdk_radial_profile": (C.c_int, [vp, i32, dbl, vp, vp]),
            "bdk_rho_k": (C.c_int, [vp, i32, dbl, i32, vp, vp]),
            "bdk_pair_correlation": (C.c_int, [vp, i32, dbl, vp, vp, vp]),
            "bdk_pair_correlation_noalloc": (C.c_int, [vp, i32, dbl, vp, vp, vp]),
            "bdk_batch_create": (C.c_int, [vp, i32, C.POINTER(vp)]),
            "bdk_batch_destroy": (None, [vp]),
            "bdk_batch_last_error": (C.c_char_p, [vp]),
            "bdk_batch_run_steps": (C.c_int, [vp, i64]),
            "bdk_batch_sync": (C.c_int, [vp]),
            "bdk_batch_status": (C.c_int, [vp, i32, vp]),
            "bdk_batch_tune": (C.c_int, [vp, i32, vp]),
            "bdk_batch_set_tuning": (C.c_int, [vp, vp]),
            "bdk_batch_profile": (C.c_int, [vp, i32, i32, vp]),
            "bdk_serve_create": (C.c_int, [i32, u64, i32, i64, C.POINTER(vp)]),
            "bdk_serve_destroy": (None, [vp]),
            "bdk_serve_last_error": (C.c_char_p, [vp]),
            "bdk_serve_submit": (C.c_int, [vp, vp, vp, vp, i64, vp]),
            "bdk_serve_step": (C.c_int, [vp, vp, vp, vp]),
            "bdk_serve_result": (C.c_int, [vp, i64, vp, vp, vp, vp]),
            "bdk_serve_stats": (C.c_int, [vp, vp, vp, vp]),
            "bdk_serve_timing": (C.c_int, [vp, vp, vp, vp, vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return ["bdk_params_default", "bdk_params_derived", "bdk_create", "bdk_destroy", "bdk_last_error",
            "bdk_set_positions", "bdk_get_positions", "bdk_compute_wca", "bdk_compute_capillary",
            "bdk_set_noise", "bdk_get_noise", "bdk_optimal_dt", "bdk_step", "bdk_run_steps",
            "bdk_run_interval", "bdk_run_interval_clock", "bdk_set_max_travel", "bdk_steps_for_interval", "bdk_refresh", "bdk_prepare", "bdk_sync", "bdk_step_counter",
            "bdk_get_cells", "bdk_get_mesh", "bdk_footprint", "bdk_profile_steps", "bdk_profile_kernel",
            "bdk_mt_create",
            "bdk_mt_destroy", "bdk_mt_uniform", "bdk_mt_init_positions", "bdk_mt_disc_noise",
            "bdk_init_uniform", "bdk_init_blob", "bdk_write_pos", "bdk_write_pos_async", "bdk_write_wait",
            "bdk_step_info", "bdk_read_startpos", "bdk_save_state", "bdk_load_state",
            "bdk_radial_profile", "bdk_rho_k", "bdk_pair_correlation", "bdk_pair_correlation_noalloc", "bdk_density_grid", "bdk_density_stats", "bdk_ftrho_sums", "bdk_batch_create", "bdk_batch_destroy", "bdk_batch_last_error", "bdk_batch_run_steps",
            "bdk_batch_sync", "bdk_batch_status", "bdk_batch_tune", "bdk_batch_set_tuning", "bdk_batch_profile",
            "bdk_serve_create", "bdk_serve_destroy", "bdk_serve_last_error", "bdk_serve_submit",
            "bdk_serve_step", "bdk_serve_result", "bdk_serve_stats", "bdk_serve_timing"]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_params(**kw):
    p = Params()
    lib().bdk_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def derived(p):
    out = np.zeros(6)
    lib().bdk_params_derived(C.byref(p), _p(out))
    return dict(zh16=out[0], nc=int(out[1]), B=out[2], h=out[3], dq=out[4], drho=out[5])


def footprint(p):
    b = C.c_uint64(0)
    rc = lib().bdk_footprint(C.byref(p), C.byref(b))
    if rc:
        raise BdkError(rc, lib().bdk_last_error(None).decode())
    return b.value


def steps_for_interval(obs_dt, dt):
    n = C.c_int64(0)
    lib().bdk_steps_for_interval(obs_dt, dt, C.byref(n))
    return n.value


def init_uniform(p, seed, dmin=0.9):
    N = p.n1 + p.n2
    git = np.zeros((N, 2), np.int32)
    rel = np.zeros((N, 2))
    rc = lib().bdk_init_uniform(C.byref(p), seed, dmin, _p(git), _p(rel))
    if rc:
        raise BdkError(rc, lib().bdk_last_error(None).decode())
    return git, rel


def read_startpos(p, path):
    """init_pos_aus_datei (dm:1576-1643) with P8 fixed; (git, rel), species 1 first."""
    N = p.n1 + p.n2
    git = np.zeros((N, 2), np.int32)
    rel = np.zeros((N, 2))
    rc = lib().bdk_read_startpos(C.byref(p), path.encode(), _p(git), _p(rel))
    if rc:
        raise BdkError(rc, lib().bdk_last_error(None).decode())
    return git, rel


def init_blob(p, seed, width, background=0.1, dmin=0.9):
    N = p.n1 + p.n2
    git = np.zeros((N, 2), np.int32)
    rel = np.zeros((N, 2))
    clipped = C.c_double(0.0)
    rc = lib().bdk_init_blob(C.byref(p), seed, width, background, dmin, _p(git), _p(rel), C.byref(clipped))
    if rc:
        raise BdkError(rc, lib().bdk_last_error(None).decode())
    return git, rel, clipped.value


class MT:
    """The host mt19937_64 stream of zufall.cpp:18-39."""

    def __init__(self, seed):
        h = C.c_void_p()
        lib().bdk_mt_create(seed, C.byref(h))
        self.h = h

    def __del__(self):
        try:
            lib().bdk_mt_destroy(self.h)
        except Exception:
            pass

    def uniform(self, lo, hi):
        return lib().bdk_mt_uniform(self.h, lo, hi)

    def init_positions(self, p):
        g1 = np.zeros((p.n1, 2), np.int32)
        r1 = np.zeros((p.n1, 2))
        g2 = np.zeros((max(p.n2, 1), 2), np.int32)
        r2 = np.zeros((max(p.n2, 1), 2))
        lib().bdk_mt_init_positions(self.h, C.byref(p), _p(g1), _p(r1), _p(g2), _p(r2))
        return np.concatenate([g1, g2[: p.n2]]), np.concatenate([r1, r2[: p.n2]])

    def disc_noise(self, n1, n2):
        a = np.zeros((max(n1, 1), 2))
        b = np.zeros((max(n2, 1), 2))
        lib().bdk_mt_disc_noise(self.h, n1, n2, _p(a), _p(b))
        return np.concatenate([a[:n1], b[:n2]])


class RunDynamik:
    """One simulation on the GPU, with the entry points of
    RunZustand::RunDynamik (signaturen2.h:68-188)."""

    def __init__(self, params):
        self.p = params
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.bdk_create(C.byref(params), C.byref(h))
        if rc:
            raise BdkError(rc, self.L.bdk_last_error(None).decode())
        self.h = h
        d = derived(params)
        self.nc, self.B, self.hmesh = d["nc"], d["B"], d["h"]
        self._F = None

    def close(self):
        if getattr(self, "h", None):
            self.L.bdk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc:
            raise BdkError(rc, self.L.bdk_last_error(self.h).decode())

    @property
    def N(self):
        return self.p.n1 + self.p.n2

    # --- positions (the aliases of am:305-308), species 1 first
    def set_positions(self, git, rel):
        git = np.ascontiguousarray(git, dtype=np.int32).reshape(-1, 2)
        rel = np.ascontiguousarray(rel, dtype=np.float64).reshape(-1, 2)
        n1 = self.p.n1
        g1, r1 = np.ascontiguousarray(git[:n1]), np.ascontiguousarray(rel[:n1])
        g2, r2 = np.ascontiguousarray(git[n1:]), np.ascontiguousarray(rel[n1:])
        self._chk(self.L.bdk_set_positions(self.h, 1, n1, _p(g1), _p(r1)))
        self._chk(self.L.bdk_set_positions(self.h, 2, self.p.n2, _p(g2), _p(r2)))

    init = set_positions  # RunDynamik::init builds lists from the current positions (am:361-364)

    def positions(self):
        out_g, out_r = [], []
        for sp, n in ((1, self.p.n1), (2, self.p.n2)):
            g = np.zeros((max(n, 1), 2), np.int32)
            r = np.zeros((max(n, 1), 2))
            self._chk(self.L.bdk_get_positions(self.h, sp, n, _p(g), _p(r)))
            out_g.append(g[:n])
            out_r.append(r[:n])
        return np.concatenate(out_g), np.concatenate(out_r)

    def absolute(self):
        g, r = self.positions()
        return r + self.B * g

    def _two(self, fn):
        a = np.zeros((max(self.p.n1, 1), 2))
        b = np.zeros((max(self.p.n2, 1), 2))
        self._chk(fn(self.h, _p(a), _p(b)))
        return np.concatenate([a[: self.p.n1], b[: self.p.n2]])

    # --- force entry points
    def berechne_WCA(self):
        """berechne_WCA11 + berechne_WCA22 + addiere_WCA12 (dm:617-1059)."""
        return self._two(self.L.bdk_compute_wca)

    def berechne_kapkraefte(self):
        """dm:1180-1233."""
        return self._two(self.L.bdk_compute_capillary)

    def set_noise(self, noise):
        noise = np.ascontiguousarray(noise, dtype=np.float64).reshape(-1, 2)
        a = np.ascontiguousarray(noise[: self.p.n1])
        b = np.ascontiguousarray(noise[self.p.n1:])
        self._chk(self.L.bdk_set_noise(self.h, _p(a), _p(b)))

    def berechne_zufallskraefte(self):
        """Noise the next step uses (dm:1317-1339)."""
        return self._two(self.L.bdk_get_noise)

    def optimaler_zeitschritt(self):
        dt = C.c_double(0.0)
        self._chk(self.L.bdk_optimal_dt(self.h, C.byref(dt)))
        return dt.value

    def zeitschritt(self, tmax=1.0):
        """dm:74-163; tmax <= 0 returns 1e10 without stepping (dm:80-83)."""
        dt = C.c_double(0.0)
        rc = self.L.bdk_step(self.h, tmax, C.byref(dt))
        if rc == ERR_ARG and not tmax > 0.0:
            return dt.value
        self._chk(rc)
        return dt.value

    def run_steps(self, k):
        t = C.c_double(0.0)
        self._chk(self.L.bdk_run_steps(self.h, k, C.byref(t)))
        return t.value

    def zeitschritte_bis_obs(self, obs_dt):
        n = C.c_int64(0)
        self._chk(self.L.bdk_run_interval(self.h, obs_dt, C.byref(n)))
        return n.value

    def set_max_travel(self, max_travel):
        """Travel cap of the following steps (+inf: fixed dt)."""
        self._chk(self.L.bdk_set_max_travel(self.h, max_travel))

    def run_interval_clock(self, obs_dt):
        """(steps, clock) of zeitschritte_bis_obs; the clock as numpy.longdouble
        (exact on x86, where it is the x87 extended type)."""
        n, hi, lo = C.c_int64(0), C.c_double(0.0), C.c_double(0.0)
        self._chk(self.L.bdk_run_interval_clock(self.h, obs_dt, C.byref(n), C.byref(hi), C.byref(lo)))
        return n.value, np.longdouble(hi.value) + np.longdouble(lo.value)

    def refresh_erwNachbar(self):
        self._chk(self.L.bdk_refresh(self.h))

    def prepare(self):
        """bdk_prepare: upload, lists and forces enqueued now."""
        self._chk(self.L.bdk_prepare(self.h))

    def sync(self):
        self._chk(self.L.bdk_sync(self.h))

    def step_counter(self):
        return self.L.bdk_step_counter(self.h)

    def cells(self):
        ncell = self.nc * self.nc
        start = np.zeros(ncell + 1, np.int32)
        ids = np.zeros(max(self.N, 1), np.int32)
        self._chk(self.L.bdk_get_cells(self.h, _p(start), _p(ids)))
        return start, ids[: self.N]

    def mesh(self, which):
        Z = self.p.Z
        out = np.zeros((Z, Z))
        self._chk(self.L.bdk_get_mesh(self.h, which, _p(out)))
        return out

    def getFKap(self):
        """y capillary force on species-1 particle 1 (s2:182-185)."""
        return self.berechne_kapkraefte()[1, 1]

    def getF(self):
        """capillary + WCA, y, species-1 particle 1 (s2:186-188)."""
        return self.berechne_kapkraefte()[1, 1] + self.berechne_WCA()[1, 1]

    def profile(self, k):
        ms = np.zeros(64)
        names = C.create_string_buffer(4096)
        n = C.c_int32(0)
        self._chk(self.L.bdk_profile_steps(self.h, k, 64, _p(ms), names, 4096, C.byref(n)))
        labels = names.value.decode().split("\n")
        return list(zip(labels, ms[: n.value].tolist()))

    KERNELS = {"k_wca": 0, "k_deposit": 1, "k_r2c_rows": 2, "k_col": 3, "k_c2r_rows": 4, "k_gather": 5,
               "step": 6, "k_update": 7}

    def profile_kernel(self, name, reps=20):
        """ms per launch of one kernel, reps back-to-back launches between
        two HIP events on the simulation's stream."""
        ms = C.c_double(0.0)
        self._chk(self.L.bdk_profile_kernel(self.h, self.KERNELS[name], reps, C.byref(ms)))
        return ms.value

    def zeitschritt_debug(self, tmax=1.0):
        """One step with the TimestepInfo counters (dm:168-211)."""
        info = StepInfo()
        self._chk(self.L.bdk_step_info(self.h, tmax, C.byref(info)))
        return info

    def save_state(self, path):
        self._chk(self.L.bdk_save_state(self.h, path.encode()))

    def load_state(self, path):
        self._chk(self.L.bdk_load_state(self.h, path.encode()))

    def radial_profile(self, nbins, dr):
        """centres [6] and profiles [6, nbins] (rho1ring, rho2ring, rho1zen1,
        rho1zen2, rho2zen1, rho2zen2), observablen_auswertung.cpp:140-330."""
        cen = np.zeros(6)
        prof = np.zeros((6, nbins))
        self._chk(self.L.bdk_radial_profile(self.h, nbins, dr, _p(cen), _p(prof)))
        return cen, prof

    def ftrho_sums(self, species, qbins, dq_ft):
        """(cos sums, sin sums) over the reference's (i, j) pairs (au:307-360)."""
        re, im = np.zeros(qbins * qbins), np.zeros(qbins * qbins)
        n = C.c_int32(0)
        self._chk(self.L.bdk_ftrho_sums(self.h, species, qbins, dq_ft, _p(re), _p(im), C.byref(n)))
        return re[: n.value], im[: n.value]

    def density_grid(self, m):
        """rho1, rho2 [m, m] of the snapshot grid (main_auswertung.cpp:214-320)."""
        r1, r2 = np.zeros((m, m)), np.zeros((m, m))
        self._chk(self.L.bdk_density_grid(self.h, m, _p(r1), _p(r2)))
        return r1, r2

    def density_stats(self, m, nbins=100, dbin=0.01, rho_thresh=(0.1, 0.2, 0.5), drho_thresh=(0.1, 0.2, 0.5, 0.99)):
        """hist [4, nbins] (rho1, rho2, rho1+rho2, rho1-rho2) and over [7]
        (main_auswertung.cpp:324-385), as fractions of the m*m cells."""
        hist, over = np.zeros((4, nbins)), np.zeros(7)
        rt = np.asarray(rho_thresh, dtype=np.float64)
        dt = np.asarray(drho_thresh, dtype=np.float64)
        self._chk(self.L.bdk_density_stats(self.h, m, nbins, dbin, _p(rt), _p(dt), _p(hist), _p(over)))
        return hist, over

    def pair_correlation(self, nbins, dr):
        """g11, g12, g22 (auswertung.cpp:110-188)."""
        out = [np.zeros(nbins) for _ in range(3)]
        self._chk(self.L.bdk_pair_correlation(self.h, nbins, dr, _p(out[0]), _p(out[1]), _p(out[2])))
        return out

    def pair_correlation_noalloc(self, nbins, dr):
        """The same through bdk_pair_correlation_noalloc (no device allocation)."""
        out = [np.zeros(nbins) for _ in range(3)]
        self._chk(self.L.bdk_pair_correlation_noalloc(self.h, nbins, dr, _p(out[0]), _p(out[1]), _p(out[2])))
        return out

    def rho_k(self, species, dq_ratio=0.5, nbins=None):
        """record_rhoFFTW (auswertung.cpp:529-626): shell averages of the
        centred FFT of one species' density; returns (re, im)."""
        if nbins is None:
            nbins = int(0.5 * self.p.Z * dq_ratio)
        re = np.zeros(nbins)
        im = np.zeros(nbins)
        self._chk(self.L.bdk_rho_k(self.h, species, dq_ratio, nbins, _p(re), _p(im)))
        return re, im

    def write_pos(self, path1, path2, nr, obs_dt):
        self._chk(self.L.bdk_write_pos(self.h, path1.encode(), path2.encode() if path2 else None, nr, obs_dt))

    def write_pos_async(self, path1, path2, nr, obs_dt):
        """The checkpoint of write_pos, formatted and written by a host thread
        while later steps run; write_wait() returns its status."""
        self._chk(self.L.bdk_write_pos_async(self.h, path1.encode(), path2.encode() if path2 else None, nr,
                                             obs_dt))

    def write_wait(self):
        self._chk(self.L.bdk_write_wait(self.h))


class Batch:
    """Several RunDynamik simulations advanced with one launch per stage
    (bdk_batch_*).  The batch keeps references to the members; their state
    stays readable through each member after sync()."""

    STAGES = {"k_wca": 0, "k_deposit": 1, "k_r2c_rows": 2, "k_col": 3, "k_c2r_rows": 4, "step": 7}

    def __init__(self, sims):
        self.sims = list(sims)
        self.L = lib()
        arr = (C.c_void_p * len(self.sims))(*[s.h for s in self.sims])
        h = C.c_void_p()
        rc = self.L.bdk_batch_create(arr, len(self.sims), C.byref(h))
        if rc:
            raise BdkError(rc, self.L.bdk_batch_last_error(None).decode())
        self.h = h

    def _chk(self, rc):
        if rc:
            raise BdkError(rc, self.L.bdk_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.bdk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_steps(self, k):
        self._chk(self.L.bdk_batch_run_steps(self.h, k))

    def sync(self):
        self._chk(self.L.bdk_batch_sync(self.h))

    def status(self):
        """Per-member status after the enqueued steps (OK, ERR_NONFINITE, ...)."""
        out = np.zeros(len(self.sims), np.int32)
        self._chk(self.L.bdk_batch_status(self.h, len(self.sims), _p(out)))
        return out

    TUNE_KEYS = ("ept_r2c", "ept_col", "ept_c2r", "wca_block", "deposit_tile", "persist")

    def tune(self, reps=10):
        """The autotuner's choice (bdk_batch_tune) as a dict."""
        out = (C.c_int32 * len(self.TUNE_KEYS))()
        self._chk(self.L.bdk_batch_tune(self.h, reps, C.cast(out, C.c_void_p)))
        return dict(zip(self.TUNE_KEYS, list(out)))

    def set_tuning(self, **choice):
        """bdk_batch_set_tuning; unnamed knobs keep the defaults (8, 8, 8, 256, 38, 0)."""
        d = dict(ept_r2c=8, ept_col=8, ept_c2r=8, wca_block=256, deposit_tile=38, persist=0)
        d.update(choice)
        arr = (C.c_int32 * len(self.TUNE_KEYS))(*[d[k] for k in self.TUNE_KEYS])
        self._chk(self.L.bdk_batch_set_tuning(self.h, C.cast(arr, C.c_void_p)))

    def profile(self, stage, reps=20):
        ms = C.c_double(0.0)
        self._chk(self.L.bdk_batch_profile(self.h, self.STAGES[stage], reps, C.byref(ms)))
        return ms.value


class Server:
    """Job server with continuous batching (bdk_serve_*)."""

    def __init__(self, hbm_budget, max_jobs=256, tick_steps=256, device=0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.bdk_serve_create(device, int(hbm_budget), max_jobs, tick_steps, C.byref(h))
        if rc:
            raise BdkError(rc, "bdk_serve_create failed")
        self.h = h
        self.sizes = {}

    def _chk(self, rc):
        if rc:
            raise BdkError(rc, self.L.bdk_serve_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.bdk_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, params, git, rel, steps):
        git = np.ascontiguousarray(git, dtype=np.int32).reshape(-1, 2)
        rel = np.ascontiguousarray(rel, dtype=np.float64).reshape(-1, 2)
        jid = C.c_int64(0)
        self._chk(self.L.bdk_serve_submit(self.h, C.byref(params), _p(git), _p(rel), steps, C.byref(jid)))
        self.sizes[jid.value] = len(git)
        return jid.value

    def step(self):
        a, q, d = C.c_int32(0), C.c_int32(0), C.c_int32(0)
        self._chk(self.L.bdk_serve_step(self.h, C.byref(a), C.byref(q), C.byref(d)))
        return a.value, q.value, d.value

    def result(self, jid):
        """(git, rel, t_admit, t_done) of a finished job; raises BdkError with
        the job's error code if it failed."""
        n = self.sizes.get(jid, 0)  # unknown ids: the library reports them
        git = np.zeros((n, 2), np.int32)
        rel = np.zeros((n, 2))
        ta, td = C.c_double(0.0), C.c_double(0.0)
        self._chk(self.L.bdk_serve_result(self.h, jid, _p(git), _p(rel), C.byref(ta), C.byref(td)))
        return git, rel, ta.value, td.value

    def stats(self):
        ps, busy, used = C.c_int64(0), C.c_double(0.0), C.c_uint64(0)
        self._chk(self.L.bdk_serve_stats(self.h, C.byref(ps), C.byref(busy), C.byref(used)))
        return ps.value, busy.value, used.value

    def timing(self):
        """Host seconds in admission, batch rebuilds, stepping and retirement."""
        v = [C.c_double(0.0) for _ in range(4)]
        self._chk(self.L.bdk_serve_timing(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("admit_s", "batch_s", "run_s", "retire_s"), [x.value for x in v]))


def has_gpu():
    try:
        import subprocess
        out = subprocess.run(["rocm-smi", "--showid"], capture_output=True, timeout=20).stdout.decode()
        return "GPU" in out
    except Exception:
        return os.path.exists("/dev/kfd")
