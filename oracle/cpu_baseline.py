"""CPU baseline runner: the oracle's step loop timed on the host cores.

TEST INFRASTRUCTURE ONLY: started by the cpu_baseline leg of bench.py as a child process
(so OpenMP reads OMP_NUM_THREADS / OMP_PROC_BIND / OMP_PLACES before it starts, and no GPU
is touched).  It reads the input arrays from an .npz file and prints one JSON line.

Timing follows the plan for the baseline: wall clock of the step loop only (input set-up
and the first force evaluation excluded, neighbour rebuilds every `every` steps included),
one discarded warm-up run, then the median of `repeats` runs, each from the same start state.
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle.oracle import Oracle, num_threads  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("inputs")
    ap.add_argument("--steps", type=int, required=True)
    ap.add_argument("--warmup-steps", type=int, default=2)
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--every", type=int, default=10)
    ap.add_argument("--dt", type=float, default=0.005)
    ap.add_argument("--rcut", type=float, default=2.5)
    ap.add_argument("--skin", type=float, default=0.3)
    a = ap.parse_args()
    d = np.load(a.inputs)
    r, v, ids = d["r"], d["v"], d["id"]
    bmin, bmax, dims, cs = d["bounds_min"], d["bounds_max"], d["grid_dims"], float(d["cell_size"])

    def fresh():
        o = Oracle(bmin, bmax, dims, cs, ghost_dist=a.rcut + a.skin)
        o.set_particles(r, v, ids)
        o.init(1.0, 1.0, a.rcut, a.skin)
        return o

    if a.warmup_steps > 0:
        o = fresh()
        o.run(a.warmup_steps, a.dt, 1.0, 1.0, a.rcut, a.skin, rebuild_every=a.every)
        del o
    walls = []
    for _ in range(a.repeats):
        o = fresh()
        t0 = time.perf_counter()
        o.run(a.steps, a.dt, 1.0, 1.0, a.rcut, a.skin, rebuild_every=a.every)
        walls.append(time.perf_counter() - t0)
        del o
    med = statistics.median(walls)
    print(json.dumps(dict(atom_steps_per_s=r.shape[0] * a.steps / med, walls_s=walls, median_s=med,
                          threads=num_threads(), atoms=int(r.shape[0]), steps=a.steps)), flush=True)


if __name__ == "__main__":
    main()
