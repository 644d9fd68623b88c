"""The checkpoint loop with position output (main.cpp:266 over
zeitschritte_bis_obs dm:444-455 and schreibe_pos au:631-648) on the 2M
particle workload (TSC 2048^2, L = 1414, lambda = 50, blob start, FP32,
dt = 1e-4, obs_dt = 0.1: 1001 steps per checkpoint).

Modes: no output; bdk_write_pos after each interval (the reference's order:
the device waits while the host formats the text); bdk_write_pos_async after
each interval (the text of checkpoint k is formatted while the steps of
interval k+1 run).  The synchronous writer is also timed alone with one
formatting thread and with the default thread count.  Output goes to a
temporary directory (under --dir if given) that is removed afterwards.

Each mode runs on a fresh handle from the same start: at this fixed dt the
run stays stable for about 4400 steps (then a noise-driven close pair makes
the explicit step unstable, in FP64 as in FP32), so --checkpoints stays <= 4.

python bd-simulations_amd/checkpoint_loop.py [--checkpoints 4] [--dir DIR] [--json out.json]
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

import bdk

WORKLOAD = dict(n1=1_000_000, n2=1_000_000, L=1414.0, Z=2048, lam=50.0, kap=1.0, f2_f1=0.5, T=1.0, dt=1e-4,
                scheme=bdk.TSC, blob_width=1414.0 / 6.0, background=0.1, dmin=0.95, seed=1, obs_dt=0.1)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--checkpoints", type=int, default=4)
    ap.add_argument("--json", default=None)
    ap.add_argument("--dir", default=None, help="where the temporary output directory is made")
    args = ap.parse_args()
    if not 1 <= args.checkpoints <= 4:
        sys.exit("checkpoint_loop.py: --checkpoints must be 1..4 (the run is stable for about 4400 steps)")
    w = WORKLOAD
    p = bdk.make_params(n1=w["n1"], n2=w["n2"], L=w["L"], Z=w["Z"], scheme=w["scheme"], lam=w["lam"],
                        kap=w["kap"], f2_f1=w["f2_f1"], T=w["T"], dt_max=w["dt"], precision=bdk.FP32,
                        noise_mode=bdk.NOISE_DISC, seed=w["seed"])
    git, rel, _ = bdk.init_blob(p, w["seed"], w["blob_width"], w["background"], w["dmin"])
    N = w["n1"] + w["n2"]
    tmp = tempfile.mkdtemp(prefix="bdk_ckpt_", dir=args.dir)
    res = {"workload": "2M particles, TSC 2048^2, L=1414, lambda=50, blob start, FP32",
           "obs_dt": w["obs_dt"], "checkpoints": args.checkpoints, "dir": os.path.dirname(tmp)}
    try:
        sim = bdk.RunDynamik(p)
        sim.set_positions(git, rel)
        sim.run_steps(20)
        paths = (os.path.join(tmp, "pos1_1.txt"), os.path.join(tmp, "pos2_1.txt"))
        # the writer alone, same state, one formatting thread and the default
        for name, env in (("write_pos_1thread_s", "1"), ("write_pos_s", None)):
            if env:
                os.environ["BDK_WRITER_THREADS"] = env
            else:
                os.environ.pop("BDK_WRITER_THREADS", None)
            t0 = time.perf_counter()
            sim.write_pos(*paths, 0, w["obs_dt"])
            res[name] = time.perf_counter() - t0
            print("%s %.3f" % (name, res[name]), file=sys.stderr, flush=True)
        os.environ.pop("BDK_WRITER_THREADS", None)
        res["bytes_per_checkpoint"] = os.path.getsize(paths[0]) + os.path.getsize(paths[1])
        sim.close()
        # each mode from the same start on a fresh handle (at this fixed dt the
        # blob run stays stable for about 4400 steps)
        for mode in ("none", "sync", "async"):
            sim = bdk.RunDynamik(p)
            sim.set_positions(git, rel)
            sim.run_steps(20)
            steps = 0
            t_write = 0.0
            t0 = time.perf_counter()
            for nr in range(args.checkpoints):
                steps += sim.zeitschritte_bis_obs(w["obs_dt"])
                t1 = time.perf_counter()
                if mode == "sync":
                    sim.write_pos(*paths, nr, w["obs_dt"])
                elif mode == "async":
                    sim.write_pos_async(*paths, nr, w["obs_dt"])
                t_write += time.perf_counter() - t1
            t1 = time.perf_counter()
            sim.write_wait()
            t_write += time.perf_counter() - t1
            dt = time.perf_counter() - t0
            res[mode] = {"s_per_checkpoint": dt / args.checkpoints, "steps_per_checkpoint": steps // args.checkpoints,
                         "particle_steps_per_s": N * steps / dt,
                         "s_in_write_calls_per_checkpoint": t_write / args.checkpoints}
            print("%-5s %.3f s/checkpoint (%.3f in write calls), %.3e particle-steps/s"
                  % (mode, dt / args.checkpoints, t_write / args.checkpoints, N * steps / dt),
                  file=sys.stderr, flush=True)
            sim.close()
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    line = json.dumps(res)
    print(line)
    if args.json:
        with open(args.json, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
