"""Precision ladder: FP32 forces and FP32 positions against FP64 on a dense
system.

L = 2000, N1 = N2 = 2,000,000 (density 1.0), Z = 4096, for NGP, CIC and
TSC: per-component relative RMS and maximum error of the WCA and the
capillary forces in FP32 mode against FP64 mode (whose arithmetic is the
reference's, pinned against the FP64 CPU oracle by the parity tests) on
identical seeded positions, and the trajectory drift against FP64 over 500
steps at kT = 0 and kT = 1 (the same Philox noise in all) of three rungs:
FP32 forces with FP64 positions, FP32 forces with FP32 in-cell positions,
and FP32 forces with FP32 positions whose increments are Kahan-compensated.

python bd-simulations_amd/precision_ladder.py [--steps 500] [--n 2000000] [--json out.json]
"""
import argparse
import json
import sys
import time

import numpy as np

import bdk

SCHEMES = {"NGP": bdk.NGP, "CIC": bdk.CIC, "TSC": bdk.TSC}
RUNGS = {"fp64": bdk.FP64, "fp32": bdk.FP32, "fp32_pos32": bdk.FP32_POS32, "fp32_pos32k": bdk.FP32_POS32K}


def rel_err(a, b):
    """per component: relative RMS and max |a-b| / max |b|."""
    out = {}
    for k, name in enumerate("xy"):
        d = a[:, k] - b[:, k]
        out[name] = {"rel_rms": float(np.sqrt(np.sum(d * d) / max(np.sum(b[:, k] ** 2), 1e-300))),
                     "rel_max": float(np.abs(d).max() / max(np.abs(b[:, k]).max(), 1e-300))}
    return out


def drift(s32, s64, L):
    g32, r32 = s32.positions()
    g64, r64 = s64.positions()
    d = (g32 * s32.B + r32) - (g64 * s64.B + r64)
    d -= L * np.round(d / L)
    a = np.sqrt((d * d).sum(1))
    return {"max": float(a.max()), "rms": float(np.sqrt((a * a).mean()))}


def run(n=2_000_000, L=2000.0, Z=4096, steps=500, seed=7, log=print):
    base = dict(n1=n, n2=n, L=L, Z=Z, lam=50.0, kap=1.0, f2_f1=0.5, dt_max=1e-4, seed=seed)
    p0 = bdk.make_params(precision=bdk.FP64, scheme=bdk.TSC, **base)
    log("ladder: uniform start, %d particles (thinned 0.95 lattice)" % (2 * n))
    git, rel, _ = bdk.init_blob(p0, seed, L, 1.0, 0.95)
    res = {"n_particles": 2 * n, "L": L, "Z": Z, "steps": steps, "schemes": {}}
    for name, sch in SCHEMES.items():
        r = {}
        sims = {}
        for prec in (bdk.FP64, bdk.FP32):
            s = bdk.RunDynamik(bdk.make_params(precision=prec, scheme=sch, **base))
            s.set_positions(git, rel)
            sims[prec] = s
        r["wca"] = rel_err(sims[bdk.FP32].berechne_WCA(), sims[bdk.FP64].berechne_WCA())
        r["capillary"] = rel_err(sims[bdk.FP32].berechne_kapkraefte(), sims[bdk.FP64].berechne_kapkraefte())
        for s in sims.values():
            s.close()
        for T in (0.0, 1.0):
            tr = {}
            for rung, prec in RUNGS.items():
                s = bdk.RunDynamik(bdk.make_params(precision=prec, scheme=sch, T=T, **base))
                s.set_positions(git, rel)
                s.run_steps(2)  # forces, graphs
                s.sync()
                t0 = time.perf_counter()
                s.run_steps(steps)
                s.sync()
                tr[rung] = (s, time.perf_counter() - t0)
            r["drift_T%g" % T] = {rung: drift(tr[rung][0], tr["fp64"][0], L) for rung in RUNGS if rung != "fp64"}
            r["ms_per_step_T%g" % T] = {rung: 1e3 * t / steps for rung, (_, t) in tr.items()}
            for s, _ in tr.values():
                s.close()
        log("ladder %s: %s" % (name, json.dumps(r)))
        res["schemes"][name] = r
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=2_000_000, help="particles per species")
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--json", default=None)
    a = ap.parse_args()
    res = run(n=a.n, steps=a.steps, log=lambda *x: print(*x, file=sys.stderr, flush=True))
    txt = json.dumps(res)
    if a.json:
        open(a.json, "w").write(txt)
    print(txt, flush=True)


if __name__ == "__main__":
    main()
