"""ctypes wrapper of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
CPU-baseline leg of the benchmark, as the reference result.  The product (the HIP
library and its host code) never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class Params(C.Structure):
    """Mirror of bdo_params (oracle/bd_oracle.h), parameter.h:14-99."""
    _fields_ = [
        ("n1", C.c_int32), ("n2", C.c_int32), ("scheme", C.c_int32), ("Z", C.c_int32),
        ("L", C.c_double), ("lam", C.c_double), ("kap", C.c_double), ("T", C.c_double),
        ("dt_max", C.c_double), ("mrd", C.c_double),
        ("s11_22", C.c_double), ("s11_12", C.c_double), ("e22_11", C.c_double),
        ("e12_11", C.c_double), ("f2_f1", C.c_double), ("max_pair_wca", C.c_double),
        ("idgas_only", C.c_int32), ("no_wca", C.c_int32), ("restrict_radial", C.c_int32),
        ("ngp_fix", C.c_int32), ("noise_mode", C.c_int32), ("pad0", C.c_int32),
        ("noise_var", C.c_double), ("seed", C.c_uint64),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        L.bdo_create.restype = C.c_void_p
        L.bdo_create.argtypes = [P(Params)]
        L.bdo_destroy.argtypes = [C.c_void_p]
        L.bdo_derived.argtypes = [C.c_void_p, C.c_void_p]
        for f in ("bdo_set_positions", "bdo_get_positions"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
            getattr(L, f).restype = C.c_int
        for f in ("bdo_wca_init", "bdo_init_zufallspos", "bdo_berechne_wca11", "bdo_berechne_wca22",
                  "bdo_addiere_wca12", "bdo_berechne_kapkraefte", "bdo_berechne_zufallskraefte",
                  "bdo_refresh_erwnachbar"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = None
        L.bdo_optimaler_zeitschritt.argtypes = [C.c_void_p]
        L.bdo_optimaler_zeitschritt.restype = C.c_double
        L.bdo_zeitschritt.argtypes = [C.c_void_p, C.c_double]
        L.bdo_zeitschritt.restype = C.c_double
        L.bdo_zeitschritte_bis_obs.argtypes = [C.c_void_p, C.c_double]
        L.bdo_zeitschritte_bis_obs.restype = C.c_int64
        L.bdo_steps_for_interval.argtypes = [C.c_double, C.c_double]
        L.bdo_steps_for_interval.restype = C.c_int64
        L.bdo_get_forces.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.bdo_get_forces.restype = C.c_int
        L.bdo_set_noise.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.bdo_get_mesh.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.bdo_get_list.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int]
        L.bdo_get_list.restype = C.c_int
        L.bdo_step_counter.argtypes = [C.c_void_p]
        L.bdo_step_counter.restype = C.c_int64
        L.bdo_write_pos_header.argtypes = [C.c_char_p]
        L.bdo_append_pos.argtypes = [C.c_void_p, C.c_int, C.c_char_p, C.c_int, C.c_double]
        L.bdo_mt_seed.argtypes = [C.c_void_p, C.c_uint64]
        L.bdo_mt_next.argtypes = [C.c_void_p]
        L.bdo_mt_next.restype = C.c_uint64
        L.bdo_mt_canonical.argtypes = [C.c_void_p]
        L.bdo_mt_canonical.restype = C.c_double
        L.bdo_philox4x32.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.bdo_philox_noise.argtypes = [C.c_uint64, C.c_uint32, C.c_uint64, C.c_int, C.c_double,
                                       C.c_void_p]
        L.bdo_fft2d.argtypes = [C.c_void_p, C.c_int, C.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def make_params(**kw):
    """Defaults follow parameter.h's structure with the benchmark configs'
    conventions: sigma/eps ratios 1, maxPairWCA 1e5, fixed dt (mrd = inf)."""
    d = dict(n1=100, n2=100, scheme=2, Z=64, L=32.0, lam=8.0, kap=1.0, T=1.0, dt_max=1e-4,
             mrd=float("inf"), s11_22=1.0, s11_12=1.0, e22_11=1.0, e12_11=1.0, f2_f1=0.5,
             max_pair_wca=1.0e5, idgas_only=0, no_wca=0, restrict_radial=0, ngp_fix=0,
             noise_mode=1, pad0=0, noise_var=1.0, seed=1234)
    d.update(kw)
    return Params(**d)


class Oracle:
    """One reference run state (RunZustand::RunDynamik) on the CPU."""

    def __init__(self, params):
        self.p = params
        self.L = lib()
        self.h = self.L.bdo_create(C.byref(params))
        d = np.zeros(6)
        self.L.bdo_derived(self.h, _ptr(d))
        self.zh16, self.nc, self.B, self.hmesh, self.dq, self.drho = d
        self.nc = int(self.nc)

    def close(self):
        if self.h:
            self.L.bdo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n(self, sp):
        return self.p.n1 if sp == 1 else self.p.n2

    def set_positions(self, sp, git, rel):
        git = np.ascontiguousarray(git, dtype=np.int32).reshape(-1, 2)
        rel = np.ascontiguousarray(rel, dtype=np.float64).reshape(-1, 2)
        assert self.L.bdo_set_positions(self.h, sp, len(git), _ptr(git), _ptr(rel)) == 0

    def set_all(self, git, rel):
        """git/rel over all particles, species 1 first."""
        n1 = self.p.n1
        self.set_positions(1, git[:n1], rel[:n1])
        self.set_positions(2, git[n1:], rel[n1:])
        self.L.bdo_wca_init(self.h)

    def positions(self, sp):
        n = self.n(sp)
        git = np.zeros((n, 2), np.int32)
        rel = np.zeros((n, 2))
        self.L.bdo_get_positions(self.h, sp, n, _ptr(git), _ptr(rel))
        return git, rel

    def all_positions(self):
        g1, r1 = self.positions(1)
        g2, r2 = self.positions(2)
        return np.concatenate([g1, g2]), np.concatenate([r1, r2])

    def absolute(self):
        g, r = self.all_positions()
        return g * self.B + r

    def forces(self, which, sp):
        out = np.zeros((self.n(sp), 2))
        self.L.bdo_get_forces(self.h, which, sp, _ptr(out))
        return out

    def all_forces(self, which):
        return np.concatenate([self.forces(which, 1), self.forces(which, 2)])

    def wca(self):
        self.L.bdo_berechne_wca11(self.h)
        self.L.bdo_berechne_wca22(self.h)
        self.L.bdo_addiere_wca12(self.h)
        return self.all_forces(0)

    def kap(self):
        self.L.bdo_berechne_kapkraefte(self.h)
        return self.all_forces(1)

    def noise(self):
        self.L.bdo_berechne_zufallskraefte(self.h)
        return self.all_forces(2)

    def set_noise(self, n_all):
        n_all = np.ascontiguousarray(n_all, dtype=np.float64).reshape(-1, 2)
        a = np.ascontiguousarray(n_all[: self.p.n1])
        b = np.ascontiguousarray(n_all[self.p.n1:])
        self.L.bdo_set_noise(self.h, _ptr(a), _ptr(b))

    def optimal_dt(self):
        return self.L.bdo_optimaler_zeitschritt(self.h)

    def step(self, tmax=1.0):
        return self.L.bdo_zeitschritt(self.h, tmax)

    def refresh(self):
        self.L.bdo_refresh_erwnachbar(self.h)

    def steps_until_obs(self, obs_dt):
        return self.L.bdo_zeitschritte_bis_obs(self.h, obs_dt)

    def mesh(self, which):
        Z = self.p.Z
        out = np.zeros((Z, Z))
        self.L.bdo_get_mesh(self.h, which, _ptr(out))
        return out

    def neighbour_list(self, sp, cx, cy):
        cap = 4096
        out = np.zeros(cap, np.int32)
        n = self.L.bdo_get_list(self.h, sp, cx, cy, _ptr(out), cap)
        return out[:n].copy()

    def init_zufallspos(self):
        self.L.bdo_init_zufallspos(self.h)

    def mt_seed(self, seed):
        self.L.bdo_mt_seed(self.h, seed)

    def mt_next(self):
        return self.L.bdo_mt_next(self.h)

    def mt_canonical(self):
        return self.L.bdo_mt_canonical(self.h)

    def step_counter(self):
        return self.L.bdo_step_counter(self.h)

    def write_pos_header(self, path):
        self.L.bdo_write_pos_header(path.encode())

    def append_pos(self, sp, path, nr, obs_dt):
        self.L.bdo_append_pos(self.h, sp, path.encode(), nr, obs_dt)


def philox4x32(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    lib().bdo_philox4x32(_ptr(c), _ptr(k), _ptr(out))
    return out


def philox_noise(seed, pid, step, gauss=False, var=1.0):
    out = np.zeros(2)
    lib().bdo_philox_noise(seed, pid, step, int(gauss), var, _ptr(out))
    return out


def fft2d(a, sign):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    n = a.shape[0]
    buf = a.view(np.float64).copy()
    lib().bdo_fft2d(_ptr(buf), n, sign)
    return buf.view(np.complex128).reshape(n, n)


def steps_for_interval(obs_dt, dt):
    return lib().bdo_steps_for_interval(obs_dt, dt)
