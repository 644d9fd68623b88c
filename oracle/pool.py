"""All-core CPU baseline of an ensemble: independent replicas of the FP64
oracle, one per worker process, stepped for a fixed wall-clock budget.

TEST INFRASTRUCTURE ONLY: used by the CPU-baseline leg of the benchmark, as
the reported host figure for many independent runs (the reference runs its
`runs` one after another in one process, main.cpp:126-300; a host user with
many cores would run several such processes side by side).
"""
import time

import numpy as np


def run_replica(args):
    """args = (params dict for oracle.make_params, git, rel, budget_s).
    Returns (particle-steps, seconds) of the timed steps after one warm-up."""
    import oracle as O

    kw, git, rel, budget = args
    o = O.Oracle(O.make_params(**kw))
    o.set_all(np.asarray(git), np.asarray(rel))
    o.step(1.0)
    n = kw["n1"] + kw["n2"]
    steps = 0
    t0 = time.perf_counter()
    while True:
        o.step(1.0)
        steps += 1
        dt = time.perf_counter() - t0
        if dt >= budget:
            break
    o.close()
    return n * steps, dt


def forces(args):
    """args = (params dict for oracle.make_params, git, rel).  Returns the
    oracle's (WCA, capillary) forces [N][2] of all particles, species 1
    first, on these positions (dm:617-1059, dm:1180-1233)."""
    import oracle as O

    kw, git, rel = args
    o = O.Oracle(O.make_params(**kw))
    o.set_all(np.asarray(git), np.asarray(rel))
    w = o.wca()
    k = o.kap()
    o.close()
    return w, k
