"""Builds matched oracle / GPU simulations from the same parameters and the
same (git, rel) positions."""
import numpy as np

import bdk
import oracle as O
from helpers import rsa_positions, to_git_rel

ORACLE_NOISE = {bdk.NOISE_DISC: 1, bdk.NOISE_GAUSS: 2, bdk.NOISE_UPLOAD: 3}


def params(**kw):
    """Returns (bdk params, oracle params) for the same physics."""
    d = dict(n1=500, n2=500, scheme=bdk.CIC, Z=64, L=40.0, lam=8.0, kap=1.0, T=1.0, dt_max=1e-4,
             max_travel=float("inf"), sigma11_22=1.0, sigma11_12=1.0, eps22_11=1.0, eps12_11=1.0,
             f2_f1=0.5, max_pair_wca=1.0e5, idgas_only=0, no_wca=0, restrict_radial=0, ngp_fix=0,
             precision=bdk.FP64, noise_mode=bdk.NOISE_DISC, green_mode=0, device=0, noise_var=1.0,
             seed=1234)
    d.update(kw)
    bp = bdk.make_params(**d)
    op = O.make_params(n1=d["n1"], n2=d["n2"], scheme=d["scheme"], Z=d["Z"], L=d["L"], lam=d["lam"],
                       kap=d["kap"], T=d["T"], dt_max=d["dt_max"], mrd=d["max_travel"],
                       s11_22=d["sigma11_22"], s11_12=d["sigma11_12"], e22_11=d["eps22_11"],
                       e12_11=d["eps12_11"], f2_f1=d["f2_f1"], max_pair_wca=d["max_pair_wca"],
                       idgas_only=d["idgas_only"], no_wca=d["no_wca"],
                       restrict_radial=d["restrict_radial"], ngp_fix=d["ngp_fix"],
                       noise_mode=ORACLE_NOISE.get(d["noise_mode"], 1), noise_var=d["noise_var"], seed=d["seed"])
    return bp, op


def positions(bp, seed=1, dmin=0.9):
    der = bdk.derived(bp)
    x = rsa_positions(bp.n1 + bp.n2, bp.L, dmin, seed)
    return to_git_rel(x, der["B"], der["nc"])


def pair(seed=1, dmin=0.9, git=None, rel=None, **kw):
    bp, op = params(**kw)
    if git is None:
        git, rel = positions(bp, seed, dmin)
    o = O.Oracle(op)
    o.set_all(git, rel)
    g = bdk.RunDynamik(bp)
    g.set_positions(git, rel)
    return g, o, git, rel
