#!/usr/bin/env python3
"""QAOA edge-evaluation throughput on one MI355X.

Workload (the metric's configuration): QAOA p=1 MaxCut on a seeded random
3-regular graph with 30 vertices (45 edges; tests/golden/graphs/3reg_n30_s2024.dgf).
Each <Z_i Z_j> is its own light-cone network (qTorch's F_p, src/maxcut.cpp:162-204).
One step = one energy evaluation of the whole graph (45 edge networks).

Prints one JSON line.  --gpus N with N > 1 is not supported (one GPU).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import random
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

GRAPH = os.path.join(REPO, "tests", "golden", "graphs", "3reg_n30_s2024.dgf")


def angles(seed: int = 77):
    """gamma ~ U[0, pi], beta ~ U[0, pi/2]; exp(-i gamma ZZ) and RX(2 beta) in the
    reference's angle vector [beta, gamma_ref] with gamma_ref = -2 gamma."""
    rng = random.Random(seed)
    gamma = rng.uniform(0, math.pi)
    beta = rng.uniform(0, math.pi / 2)
    return [beta, -2.0 * gamma]


def cpu_baseline(dgf: str, bg, target_s: float = 12.0):
    """The reference itself (oracle/_ref/ref_driver_O2, built from the qTorch sources)
    contracting every edge network stochastically, as F_p does, on this host."""
    import qtorch_amd as qa
    drv = os.path.join(REPO, "oracle", "_ref", "ref_driver_O2")
    if not os.path.exists(drv):
        return None
    ne, _ = qa.qaoa_num_edges(dgf)
    tmp = tempfile.mkdtemp(prefix="qt_cpu_")
    lst = os.path.join(tmp, "jobs.txt")
    with open(lst, "w") as f:
        for e in range(ne):
            q, m = qa.qaoa_edge_text(dgf, 1, True, e, bg)
            qp, mp = os.path.join(tmp, f"e{e}.qasm"), os.path.join(tmp, f"e{e}.meas")
            open(qp, "w").write(q)
            open(mp, "w").write(m)
            f.write(f"{qp} {mp}\n")
    out = subprocess.run([drv, "time", lst, "1"], capture_output=True, text=True, timeout=600).stdout
    t1 = float(out.split("TIME")[1].split()[0])
    reps = max(1, int(target_s / max(t1, 1e-3)))
    out = subprocess.run([drv, "time", lst, str(reps)], capture_output=True, text=True, timeout=900).stdout
    t = float(out.split("TIME")[1].split()[0])
    jobs = int(out.split("JOBS")[1].split()[0])
    return {"value": jobs / t, "unit": "edge-evals/s", "cores": 8, "kind": "reference",
            "sample": f"{jobs} edge networks ({reps} x 45 edges) through the reference's "
                      f"ContractionTools::Contract(Stochastic) built -O2 from its sources "
                      f"(up to 8 std::threads: 2 partition threads, 8 for output rank >= 8), {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine measures one GPU; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np

    import qtorch_amd as qa

    dgf = open(GRAPH).read()
    bg = angles()
    ctx = qa.Context(0)
    ne, _ = qa.qaoa_num_edges(dgf)
    for _ in range(args.warmup):
        qa.qaoa_energy_plain(ctx, dgf, 1, True, bg, seed=0)
    ctx.sync()
    lat = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        qa.qaoa_energy_plain(ctx, dgf, 1, True, bg, seed=s)
        lat.append(time.perf_counter() - ts)
    ctx.sync()
    wall = time.perf_counter() - t0
    value = ne * args.steps / wall
    res = {
        "metric": "QAOA edge-evals/s (p99 latency per energy evaluation), complex128 plain path",
        "value": value, "unit": "edge-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": "qaoa_p1_maxcut_3regular_n30", "edges": ne, "form": "plain_c128",
                   "order": "reference stochastic (seeded)"},
        "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(dgf, bg)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
