#!/usr/bin/env python3
"""QAOA edge-evaluation throughput on one MI355X.

Workload: QAOA p=1 MaxCut on a seeded random 3-regular graph with 30 vertices
(45 edges; tests/golden/graphs/3reg_n30_s2024.dgf).  Each <Z_i Z_j> is its own
light-cone network in qTorch's Liouville form, exactly the networks F_p builds
(src/maxcut.cpp:162-204, reference qubit map).  One step = one energy evaluation
of the graph for each of B angle sets (gamma ~ U[0, pi], beta ~ U[0, pi/2], seeded)
= 45 * B edge networks (B = 131,072 by default), inputs resident on the device.

value: edge-evals/s of the complex128 specialised executor (generated straight-line
code per edge program, conjugate symmetry, energies only: each step launches one
kernel that also sums the previous step's partial energies), every edge within
1e-12 of complex128 table-executor values and, for the reference's angle set, of the
values the reference printed (tests/golden/cases.json).  Beside it: the complex64
forms (4M, 3M, bf16x3) with their per-edge error, the drop-in complex128 path
(reference stochastic order, one pairwise contraction per launch) on the same GPU,
and the reference itself on this host's CPU cores.

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
METRIC = "QAOA edge-evals/s & p99 latency, FP32-MFMA vs FP64 plain, rel err\u22641e-5"
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 (vector and matrix)
# MI355X vector FMA rate without packed FP32 instructions (never emitted here): FP64
# and FP32 both 78.6 TFLOP/s.  The specialised kernel is plain VALU FMA chains.
VALU_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# HBM bytes per edge network measured by rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# runs of this benchmark (each scaled by what the counter reads for a 1 GiB device
# copy in the same process), keyed by (kernel, form, angle sets per launch); the
# profile file that holds the counts is named beside the figure.
PMC_BYTES_PER_NETWORK = {
    ("tables", "c64_4m", 8192): ((2.0 * 14451.6 + 6544.0) * 1024 / 368640, "profiles/k_batch_qaoa_p1_3reg_n30.md"),
    ("specialised", "c64_4m", 8192): ((2.0 * 14490.2 + 5824.0) * 1024 / 368640,
                                      "profiles/qt_batch_specialised_qaoa_p1_3reg_n30.md"),
    ("specialised_sums", "c128", 131072): ((2.0 * 41664.9 + 17408.0) * 1024 / (131072 * 45),
                                           "profiles/qt_batch_specialised_sums_qaoa_p1_3reg_n30.md"),
}


def host_info():
    """CPU model and the cores this process may use (for the CPU baseline)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0))}


def angle_sets(n: int, seed: int = 2024):
    """Reference angle vectors [beta, gamma_ref], gamma_ref = -2 gamma, so that the
    emitted CNOT-Rz-CNOT is exp(-i gamma ZZ) and the mixer RX(2 beta)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    gamma = rng.uniform(0, math.pi, n)
    beta = rng.uniform(0, math.pi / 2, n)
    return np.stack([beta, -2.0 * gamma], axis=1)


def cpu_baseline(dgf: str, bg, target_s: float = 12.0):
    """The reference (oracle/_ref/ref_driver_O2, built from the qTorch sources) runs
    F_p's per-edge ContractionTools::Contract(Stochastic) on this host."""
    import qtorch_amd as qa
    drv = os.path.join(REPO, "oracle", "_ref", "ref_driver_O2")
    if not os.path.exists(drv):
        return None
    ne, _ = qa.qaoa_num_edges(dgf)
    tmp = tempfile.mkdtemp(prefix="qt_cpu_")
    lst = os.path.join(tmp, "jobs.txt")
    with open(lst, "w") as f:
        for e in range(ne):
            q, m = qa.qaoa_edge_text(dgf, 1, True, e, list(bg))
            qp, mp = os.path.join(tmp, f"e{e}.qasm"), os.path.join(tmp, f"e{e}.meas")
            with open(qp, "w") as g:
                g.write(q)
            with open(mp, "w") as g:
                g.write(m)
            f.write(f"{qp} {mp}\n")
    out = subprocess.run([drv, "time", lst, "1"], capture_output=True, text=True, timeout=600).stdout
    t1 = float(out.split("TIME")[1].split()[0])
    reps = max(1, int(target_s / max(t1, 1e-3)))
    out = subprocess.run([drv, "time", lst, str(reps)], capture_output=True, text=True, timeout=900).stdout
    t = float(out.split("TIME")[1].split()[0])
    jobs = int(out.split("JOBS")[1].split()[0])
    # ParallelContract runs its two partitions on two std::threads, then one; the
    # per-MAC fan-out to 8 threads needs an output rank >= 8, which these light
    # cones never reach: at most 2 threads run at once
    res = {"value": jobs / t, "unit": "edge-evals/s", "cores": 2, "kind": "reference",
           "sample": f"{jobs} edge networks ({reps} energy evaluations of the 45-edge graph) through the "
                     f"reference's ContractionTools::Contract(Stochastic), F_p's per-edge loop run serially; "
                     f"{t:.1f} s",
           "threads": "2 (ParallelContract's partition threads; no output rank reaches the 8-thread fan-out)",
           "compiler": "g++ -O2 -g -std=c++11 -pthread (the reference's Makefile builds without -O)",
           **host_info()}
    # upper bound for the host: independent copies on every core (edges spread over cores)
    cores = min(len(os.sched_getaffinity(0)), 16)
    per = max(1, int(0.25 * target_s / max(t1, 1e-3)))  # about a quarter of the sample's time per copy
    t0 = time.perf_counter()
    procs = [subprocess.Popen([drv, "time", lst, str(per)], stdout=subprocess.PIPE, text=True)
             for _ in range(cores)]
    done = sum(int(p.communicate(timeout=900)[0].split("JOBS")[1].split()[0]) for p in procs)
    wall = time.perf_counter() - t0
    res["all_cores"] = {"value": done / wall, "unit": "edge-evals/s", "cores": cores, "processes": cores,
                        "sample": f"{done} edge networks in {cores} concurrent copies, {wall:.1f} s"}
    return res


def golden_reference(golden_path: str):
    """complex128 values the reference printed for the golden angle set."""
    with open(golden_path) as f:
        cases = json.load(f)["cases"]
    ref = sorted((c for c in cases if c["name"].startswith("qaoa3_ref_")), key=lambda c: c["edge"])
    import numpy as np
    return np.array([c["value"][0] for c in ref]), ref[0]["betas_gammas"]


# ---- the other configurations (--workload) ----------------------------------------------

def _form_name(qa, prec):
    return {qa.QT_PLAIN_C128: "c128", qa.QT_C64_4M: "c64_4m", qa.QT_C64_3M: "c64_3m", qa.QT_C64_BF16X3: "c64_bf16x3"}[prec]


def run_qft8(args, qa, np):
    """QFT on |0^8>: one amplitude's probability in the line-graph (min-fill) order on the
    drop-in complex128 path, and the full 256-entry probability vector as one batch of
    256 measured networks; every entry is 1/256."""
    from qtorch_amd import workloads as wl
    ctx = qa.Context(0)
    q = wl.qft_qasm(8)
    bits = wl.bitstring(8, args.seed)
    net = qa.Network(q, wl.measure_text(bits), ctx)
    order = net.linegraph_minfill()
    t0 = time.perf_counter()
    p_one = net.linegraph_contract(order)
    t_one = time.perf_counter() - t0
    strings = [format(i, "08b")[::-1] for i in range(256)]
    nb = qa.NetBatch(ctx, [q] * 256, [wl.measure_text(b) for b in strings], precision=qa.QT_PLAIN_C128)
    nb.autotune()
    nb.launch()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nb.launch()
    ctx.sync()
    t_vec = (time.perf_counter() - t0) / args.steps
    probs = nb.results().real
    return {"metric": "QFT8 probability vector (256 entries) per second", "value": 1.0 / t_vec,
            "unit": "vectors/s", "n_gpus": 1, "steps": args.steps, "warmup": 1, "ms_per_step": 1e3 * t_vec,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c128", "data": "synthetic",
            "config": {"workload": "qft8_probability_vector", "networks_per_step": 256, "measure_bits": bits},
            "one_amplitude": {"order": "line-graph min-fill", "probability": p_one.real,
                              "abs_err_vs_1_256": abs(p_one - 1 / 256), "ms": 1e3 * t_one},
            "vector": {"max_abs_err_vs_1_256": float(np.max(np.abs(probs - 1 / 256))), "sum": float(probs.sum())}}


def run_grid(args, qa, np):
    """Precision stress: one amplitude of seeded grid random circuits (RX/RY/RZ layers,
    CZ on alternating grid-edge patterns) in the amplitude picture, sliced to at most
    --max-bits index bits per tensor; every form against complex128.  6x8 depth 8 and
    4x6 depth 14 in the line-graph order, every slice; 6x8 depth 14 in the stochastic
    order sliced to 28 bits (2^27 slices of about 2^37 MACs each, beyond reach whole):
    the first --grid-sample slices that are not zero by structure, whose sum each form
    computes; and 6x8 depth 14 whole, in the circuit sweep order with CZ split into
    bonded halves: unsliced (29-bit peak: one 4 GiB complex64 tensor) and sliced to 28
    bits as configured (64 slices)."""
    from qtorch_amd import workloads as wl
    ctx = qa.Context(0)
    cases = [(6, 8, 8, qa.ORDER_LINEGRAPH, args.max_bits, None), (4, 6, 14, qa.ORDER_LINEGRAPH, args.max_bits, None),
             (6, 8, 14, qa.ORDER_STOCHASTIC, 28, args.grid_sample),
             (6, 8, 14, qa.ORDER_SWEEP, 29, None), (6, 8, 14, qa.ORDER_SWEEP, 28, None)]
    out = []
    chosen = [int(x) for x in args.grid_cases.split(",")] if args.grid_cases else range(len(cases))
    for i, (r, c, d, order, max_bits, sample) in enumerate(cases):
        if sample == 0 or i not in chosen:
            continue
        q = wl.grid_random_circuit(r, c, d, seed=2024)
        bits = wl.bitstring(r * c, args.seed)
        t0 = time.perf_counter()
        sn = qa.SlicedNetwork(q, bits, dim_log2=1, order=order, seed=0 if order == qa.ORDER_SWEEP else 2,
                              max_bits=max_bits)
        plan_s = time.perf_counter() - t0
        plan = sn.plan()
        print(f"grid {r}x{c} depth {d}: planned in {plan_s:.1f} s, {plan['slices']} slices", file=sys.stderr,
              flush=True)
        n_slices = plan["slices"]
        if sample is not None:
            # the first slices that are not zero by the structure alone (diagonal
            # gates tie sliced wires; |0> states and output bits fix them)
            picked = sn.nonzero_slices(sample)
            n_slices = len(picked)
            sn.set_slices(picked)
        res = {}
        ref = None
        forms = (qa.QT_PLAIN_C128, qa.QT_C64_4M, qa.QT_C64_3M, qa.QT_C64_BF16X3, qa.QT_C64_4M | qa.QT_RESCALE)
        n_forms = args.grid_forms if plan["slices"] * plan["cmacs_per_slice"] < 2 ** 40 else min(args.grid_forms, 2)
        for prec in forms[:n_forms]:
            sn.run(ctx, prec)  # warm the allocator
            v, ms = sn.run(ctx, prec)
            if prec == qa.QT_PLAIN_C128:
                ref = v
            flops = 8.0 * plan["cmacs_per_slice"] * n_slices
            name = _form_name(qa, prec & ~qa.QT_RESCALE) + ("_rescaled" if prec & qa.QT_RESCALE else "")
            res[name] = {"amplitude": [v.real, v.imag], "ms": ms, "tflops": flops / (ms * 1e-3) / 1e12,
                         "rel_err_vs_c128": abs(v - ref) / abs(ref) if ref != 0 else None}
            print(f"  {name}: {ms:.1f} ms", file=sys.stderr, flush=True)
        case = {"circuit": f"grid_{r}x{c}_depth{d}", "qubits": r * c, "bits": bits, "plan": plan,
                "order": {qa.ORDER_STOCHASTIC: "stochastic (seed 2)", qa.ORDER_SWEEP: "circuit sweep, CZ split"}.get(
                    order, "line-graph min-fill"),
                "plan_s": plan_s, "forms": res}
        if sample is not None:
            case["slices_contracted"] = picked
            case["value_is"] = f"sum of {n_slices} of {plan['slices']} slices (the first ones not zero by structure)"
        out.append(case)
        sn.close()
    best = max(out[0]["forms"].items(), key=lambda kv: kv[1]["tflops"])
    return {"metric": "grid amplitude contraction TFLOP/s by form (error vs complex128)",
            "value": best[1]["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": best[1]["ms"], "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
            "dtype": best[0], "data": "synthetic",
            "config": {"workload": "grid_random_circuit_amplitudes", "max_bits": args.max_bits,
                       "cases": [o["circuit"] for o in out]},
            "cases": out}


def run_uneven(args, qa, np):
    """Uneven batching as configured: 512 seeded 4-regular graphs of 20-120 vertices
    (log-uniform), QAOA p = 2 with seeded angles per graph, every edge's light cone
    (the correct qubit map), FP64 energy per graph.  Light cones whose tensors fit a
    device program run in graph-captured persistent table-executor launches, the rest
    on the sliced path with MFMA GEMMs on several streams; complex64 4M.  Compared with
    complex128 one pairwise contraction per launch, one network at a time, on a
    seeded sample of the networks (and every edge of the smallest graphs)."""
    from qtorch_amd import uneven, workloads as wl
    rng = random.Random(args.seed)
    graphs = []
    for g in range(args.graphs):
        n = int(round(math.exp(rng.uniform(math.log(20), math.log(120)))))
        n += n % 2
        graphs.append((wl.dgf_text(wl.regular_graph(n, 4, seed=7000 + g)), 2, wl.qaoa_angles(2, 9000 + g)))
    ctx = qa.Context(0)
    u = uneven.UnevenEnergies(ctx, graphs, precision=qa.QT_C64_4M, streams=args.streams)
    print(f"uneven: {len(u.nets)} light cones of {len(graphs)} graphs planned in {u.planning_s:.1f} s "
          f"({len(u.small)} batched in {len(u.batches)} batches, {len(u.big)} large)", file=sys.stderr, flush=True)
    lanes, wgs = u.configure()
    energies, values, t_small, t_big = u.run()
    # complex128, one pairwise contraction per launch, one network at a time: a seeded
    # sample of the networks, and every edge of the three smallest graphs
    sizes = [int(qa.qaoa_num_edges(d)[1]) for d, _, _ in graphs]
    smallest = sorted(range(len(graphs)), key=lambda g: sizes[g])[:3]
    sample = sorted(set(rng.sample(range(len(u.nets)), min(args.networks, len(u.nets)))) |
                    {i for i, x in enumerate(u.nets) if x[0] in smallest})
    t0 = time.perf_counter()
    ref = {}
    for i in sample:
        _, _, q, m = u.nets[i]
        net = qa.Network(q, m, ctx)
        net.order(qa.ORDER_LINEGRAPH, 0)
        ref[i] = net.value()
        net.close()
    t_one = (time.perf_counter() - t0) / len(sample)
    err = max(abs(values[i] - ref[i]) / max(abs(ref[i]), 1e-3) for i in sample)
    e_err = 0.0
    for g in smallest:
        e_ref = sum(0.5 * (1.0 - ref[i].real) for i, x in enumerate(u.nets) if x[0] == g)
        e_err = max(e_err, abs(energies[g] - e_ref) / abs(e_ref))
    total = t_small + t_big
    n_nets = len(u.nets)
    u.close()
    return {"metric": "QAOA p=2 light-cone networks per second, 512 graphs, every edge (batched + GEMM path)",
            "value": n_nets / total, "unit": "networks/s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": 1e3 * total, "higher_is_better": True, "scaling": "strong", "vs_baseline": None,
            "dtype": "c64", "data": "synthetic",
            "config": {"workload": "qaoa_p2_4regular_n20_120_512graphs_every_edge", "graphs": len(graphs),
                       "networks": n_nets, "batched": len(u.small), "batches": len(u.batches),
                       "gemm_path": len(u.big), "executor": [lanes, wgs], "gemm_path_streams": args.streams},
            "batched_ms": 1e3 * t_small, "gemm_path_ms": 1e3 * t_big, "host_planning_s": u.planning_s,
            "graph_energies_per_s": len(graphs) / total,
            "one_at_a_time_c128_ms_per_network": 1e3 * t_one,
            "one_at_a_time_c128_ms_extrapolated": 1e3 * t_one * n_nets,
            "speedup_vs_one_at_a_time": t_one * n_nets / total,
            "reference_sample": len(sample), "max_rel_err_vs_c128": err, "floor": "|d| / max(|x|, 1e-3)",
            "max_energy_rel_err_vs_c128": e_err, "energy_checked_graphs": [sizes[g] for g in smallest],
            "energies_head": [float(x) for x in energies[:8]]}


def run_serving(args, qa, np):
    """Poisson arrivals at a fixed rate, from a pool of structures planned before the
    timed region; requests run in the continuous-batching server (complex64 4M for
    the batched routes, complex128 on the large-network route unless --eager-form
    says otherwise).  Latency = submission to result.
    * mix "light": 70% QAOA energies (3-regular graphs with p = 1 or 2, 4-regular with
      p = 1; 20-200 vertices) and 30% single amplitudes of grid random circuits (16-24
      qubits, depth 8): every request fits a batched program.
    * mix "config": 70% QAOA energies on 3- and 4-regular graphs of 20-200 vertices with
      p = 1, 2 (45 %, 45 %) and 3 (10 %), 30% amplitudes of grid circuits of 24-48
      qubits at depth 8 and 14; light cones and circuits beyond a program's tensors run
      on the large-network route, and requests whose planned work exceeds the
      per-request limit (p = 3 on these graphs, 30+ qubits at depth 14) are refused."""
    from qtorch_amd import workloads as wl
    rng = random.Random(args.seed)
    graphs = []
    for i in range(12):
        d = 3 if i % 3 else 4
        n = int(round(math.exp(rng.uniform(math.log(20), math.log(200)))))
        n += (n * d) % 2
        graphs.append((wl.dgf_text(wl.regular_graph(n, d, seed=500 + i)), d, n))
    if args.mix == "config":
        shapes = [(4, 6, 8), (5, 6, 8), (6, 6, 8), (6, 8, 8), (4, 6, 14), (6, 8, 14)]
    else:
        shapes = [(4, 4, 8), (4, 5, 8), (4, 6, 8), (3, 6, 8)]
    circuits = [(wl.grid_random_circuit(r, c, dep, seed=600 + i), r * c, f"amplitude {r}x{c} d{dep}")
                for i, (r, c, dep) in enumerate(shapes)]
    reqs = []
    n_req = int(args.qps * args.duration)
    for _ in range(n_req):
        if rng.random() < 0.7:
            dgf, d, n = rng.choice(graphs)
            if args.mix == "config":
                p = rng.choices([1, 2, 3], weights=[45, 45, 10])[0]
            else:
                p = rng.choice([1, 2]) if d == 3 else 1
            reqs.append(("qaoa", dgf, p, wl.qaoa_angles(p, rng.randrange(1 << 30)), f"qaoa {d}-regular p={p}"))
        else:
            q, n, label = rng.choice(circuits)
            reqs.append(("amp", q, wl.bitstring(n, rng.randrange(1 << 30)), None, label))
    # 64 GiB of the 288 GiB HBM for the requests admitted at once: large light cones
    # (8-9 GiB of live tensors per edge network) run two edge networks side by side, two
    # such requests at a time
    eager = {"c128": qa.QT_PLAIN_C128, "c64_4m": qa.QT_C64_4M}[args.eager_form]
    srv = qa.Server(slots=2, precision=qa.QT_C64_4M, hbm_budget=64 << 30, eager_precision=eager)

    def submit(r):
        return srv.submit_qaoa(r[1], r[2], r[3]) if r[0] == "qaoa" else srv.submit_amplitude(r[1], r[2])

    # plan every structure once (outside the timed region)
    seen = set()
    for r in reqs:
        key = (r[0], r[1], r[2] if r[0] == "qaoa" else 0)
        if key not in seen:
            seen.add(key)
            t = time.perf_counter()
            w = srv.wait(submit(r))
            print(f"structure {len(seen)}: {r[4]} route {w['route']} status {w['status']} "
                  f"{1e3 * (time.perf_counter() - t):.1f} ms", file=sys.stderr, flush=True)
    base = srv.stats()
    arrivals = np.cumsum(np.random.default_rng(args.seed).exponential(1.0 / args.qps, n_req))
    ids = [0] * n_req
    t0 = time.perf_counter()
    for k, t in enumerate(arrivals):
        delay = t0 + t - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        ids[k] = submit(reqs[k])
    print(f"submitted {n_req} requests in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    res = []
    for k, i in enumerate(ids):
        res.append(srv.wait(i))
        if (k + 1) % max(n_req // 10, 1) == 0:
            print(f"{k + 1} results at {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array([r["latency_ms"] for r in res if r["status"] == 0])
    ok = sum(1 for r in res if r["status"] == 0)
    routes = {k: sum(1 for r in res if r["route"] == k and r["status"] == 0) for k in ("small", "medium", "eager")}
    classes = {}
    for r, q in zip(res, reqs):
        c = classes.setdefault(q[4], {"requests": 0, "refused": 0, "lat": []})
        c["requests"] += 1
        if r["status"] == 0:
            c["lat"].append(r["latency_ms"])
        else:
            c["refused"] += 1
    for c in classes.values():
        lt = c.pop("lat")
        c["p50_ms"] = float(np.percentile(lt, 50)) if lt else None
        c["p99_ms"] = float(np.percentile(lt, 99)) if lt else None
    # the FP64 plain path on a sample of the same requests
    ctx = qa.Context(0)
    sample = [k for k in range(n_req) if reqs[k][0] == "qaoa" and res[k]["status"] == 0 and reqs[k][2] == 1][:10]
    t1 = time.perf_counter()
    errs = []
    for k in sample:
        e, _ = qa.qaoa_energy_plain(ctx, reqs[k][1], reqs[k][2], False, list(reqs[k][3]), seed=k)
        errs.append(abs(res[k]["value"].real - e) / abs(e))
    plain_ms = 1e3 * (time.perf_counter() - t1) / max(len(sample), 1)
    # amplitudes served on the large-network route against complex128 (one request of
    # each circuit, up to three), when that route runs a complex64 form
    amp_errs = []
    if args.eager_form != "c128":
        done_circ = set()
        for k in range(n_req):
            r = reqs[k]
            if r[0] != "amp" or res[k]["status"] != 0 or res[k]["route"] != "eager" or r[4] in done_circ:
                continue
            done_circ.add(r[4])
            sn = qa.SlicedNetwork(r[1], r[2], dim_log2=1, order=qa.ORDER_SWEEP, max_bits=30)
            ref, _ = sn.run(ctx, qa.QT_PLAIN_C128)
            sn.close()
            amp_errs.append(abs(res[k]["value"] - ref) / abs(ref))
            if len(done_circ) == 3:
                break
    edge_evals = st["edge_evals"] - base["edge_evals"]
    return {"metric": "serving requests/s (p50/p99 latency) under Poisson arrivals", "value": ok / wall,
            "unit": "requests/s", "n_gpus": 1, "steps": n_req, "warmup": len(seen), "ms_per_step": 1e3 * wall / n_req,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64", "data": "synthetic",
            "config": {"workload": "serving_mix_qaoa70_amplitude30", "mix": args.mix, "qps": args.qps,
                       "duration_s": args.duration, "slots": 2, "hbm_budget_gib": 64, "structures": len(seen),
                       "large_network_form": args.eager_form},
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "completed": ok, "refused": n_req - ok, "routes": routes, "classes": classes,
            "edge_evals_per_s": edge_evals / wall,
            "launches": st["launches"] - base["launches"], "overlap_ms": st["overlap_ms"] - base["overlap_ms"],
            "busy_ms": st["busy_ms"] - base["busy_ms"], "peak_planned_bytes": st["peak_planned_bytes"],
            "plain_c128_ms_per_qaoa_request": plain_ms, "max_energy_rel_err_vs_plain_c128": max(errs) if errs else None,
            "large_route_amplitude_rel_err_vs_c128": max(amp_errs) if amp_errs else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--sets", type=int, default=131072, help="angle sets per step")
    ap.add_argument("--form", default="c128", choices=["c64_4m", "c64_3m", "c64_bf16x3", "c128"])
    ap.add_argument("--per-edge", action="store_true",
                    help="timed launches also write every edge's <ZZ> (default: energies only)")
    ap.add_argument("--latency-steps", type=int, default=2000,
                    help="steps launched and waited for one at a time for p50 / p99")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the other forms' accuracy and speed")
    ap.add_argument("--exec", default="", help="table executor shape 'lanes,wgs_per_cu' instead of the autotuner's")
    ap.add_argument("--no-hermitian", action="store_true",
                    help="specialised executor without conjugate symmetry (bit-identical to the table executor)")
    ap.add_argument("--executor", default="specialised", choices=["specialised", "tables"],
                    help="qaoa_p1: generated straight-line code (complex64 4M and complex128) or the table executor")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run plain device copies of this many MiB (calibrates HBM byte counters)")
    ap.add_argument("--workload", default="qaoa_p1", choices=["qaoa_p1", "qft8", "grid", "uneven", "serving"])
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--max-bits", type=int, default=22, help="grid: index bits per tensor after slicing")
    ap.add_argument("--grid-sample", type=int, default=8, help="grid: slices of the 6x8 depth-14 circuit (0: skip)")
    ap.add_argument("--grid-cases", default="", help="grid: comma-separated case indices (0: 6x8 d8, 1: 4x6 d14, "
                    "2: 6x8 d14 sample); default all")
    ap.add_argument("--grid-forms", type=int, default=5, help="grid: the first N of c128, 4M, 3M, bf16x3, 4M rescaled")
    ap.add_argument("--networks", type=int, default=256, help="uneven: sampled networks checked one at a time")
    ap.add_argument("--graphs", type=int, default=512, help="uneven: graphs")
    ap.add_argument("--streams", type=int, default=8, help="uneven: host threads / streams for the large networks")
    ap.add_argument("--qps", type=float, default=400.0, help="serving: arrival rate")
    ap.add_argument("--mix", default="light", choices=["light", "config"], help="serving: request mix")
    ap.add_argument("--duration", type=float, default=10.0, help="serving: seconds of arrivals")
    ap.add_argument("--eager-form", default="c128", choices=["c128", "c64_4m"],
                    help="serving: arithmetic form of the large-network route")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine measures one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.copy_mib:
        import torch
        src = torch.ones(args.copy_mib << 20, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": args.copy_mib << 20, "copies": 3}), flush=True)
        del src, dst
    if args.workload != "qaoa_p1":
        import numpy as np

        import qtorch_amd as qa
        run = {"qft8": run_qft8, "grid": run_grid, "uneven": run_uneven, "serving": run_serving}[args.workload]
        print(json.dumps(run(args, qa, np)))
        return
    import numpy as np

    import qtorch_amd as qa

    forms = {"c64_4m": qa.QT_C64_4M, "c64_3m": qa.QT_C64_3M, "c64_bf16x3": qa.QT_C64_BF16X3,
             "c128": qa.QT_PLAIN_C128}
    dgf = open(GRAPH).read()
    ctx = qa.Context(0)
    zz_ref, bg_ref = golden_reference(os.path.join(REPO, "tests", "golden", "cases.json"))
    sets = angle_sets(args.sets)
    sets[0] = bg_ref  # set 0 is the angle set whose reference values are known

    # ---- headline: batched form ----
    b = qa.QaoaBatch(ctx, dgf, 1, ref_map=True, precision=forms[args.form], order=qa.ORDER_GREEDY)
    b.set_angles(sets)
    if args.exec:
        lanes, wgs = (int(x) for x in args.exec.split(","))
        b.config(lanes, wgs)
    else:
        lanes, wgs, _ = b.autotune()
    # the table executor's time on the same launches, for comparison
    b.launch()
    ctx.sync()
    tables_ms = b.time(max(5, args.steps // 5))
    specialised = args.executor == "specialised" and args.form in ("c64_4m", "c128")
    # the light cones' leaves are gate superoperators U (x) U* and Hermitian measurement
    # vectors: conjugate-symmetric, so the generated code computes half of each tensor
    compile_ms, compile_error = None, None
    if specialised:
        try:
            compile_ms = b.compile(hermitian=not args.no_hermitian)
        except (qa.DeviceFailure, qa.InvalidFunctionInput) as exc:  # the table executor runs
            specialised, compile_error = False, str(exc)
    if specialised:
        lanes, wgs = 1, 0  # one lane per network, one wavefront per 64 networks
    # the timed steps need the energies only: per-edge values are not written
    # (group-sums-only launches, bit-identical energies); accuracy below is checked on
    # a launch that writes every edge's value
    sums_only = specialised and not args.per_edge and b.sums_only(True)
    # direct launches or a captured HIP graph, whichever runs a step faster here
    b.launch()
    ctx.sync()
    direct_ms = b.time(20)
    try:
        b.capture()
        b.launch()
        ctx.sync()
        graph_ms = b.time(20)
    except qa.DeviceFailure:
        graph_ms = None
    graph = graph_ms is not None and graph_ms < direct_ms
    if graph_ms is not None and not graph:
        b.uncapture()
    ne = b.n_edges
    # an optimiser supplies new angles for every evaluation: each step first rewrites
    # the sets' Rz / Rx blocks on the device from a fresh array of angles already in
    # HBM (qth_qaoa_batch_update_angles), then launches.  The arrays are distinct per
    # step; the latency run cycles through them.
    import torch
    fresh_host = np.stack([angle_sets(args.sets, seed=10_000 + i) for i in range(args.steps)])
    fresh = torch.from_numpy(fresh_host).to("cuda:0")
    torch.cuda.synchronize()
    ptrs = [fresh[i].data_ptr() for i in range(args.steps)]
    b.update_angles(device_ptr=ptrs[0], n_sets=args.sets)  # the first call checks the generated kernel's literals
    b.launch()
    b.flush()
    ctx.sync()
    # per-step latency (new angles, launch, completion with its energy sums), then the
    # throughput run; with direct launches of the specialised kernel each launch also
    # sums the previous launch's energies, and the last launch's sums run at the flush
    lat = []
    for i in range(args.latency_steps):
        t0 = time.perf_counter()
        b.update_angles(device_ptr=ptrs[i % args.steps], n_sets=args.sets)
        b.launch()
        b.flush()
        ctx.sync()
        lat.append(time.perf_counter() - t0)
    ctx.sync()
    # warm-up steps right before the timed ones (the device clocks settle under load)
    for i in range(args.warmup):
        b.update_angles(device_ptr=ptrs[i % args.steps], n_sets=args.sets)
        b.launch()
    b.flush()
    ctx.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        b.update_angles(device_ptr=ptrs[i], n_sets=args.sets)
        b.launch()
    b.flush()
    ctx.sync()
    wall = time.perf_counter() - t0
    e_fresh, _ = b.results()  # the last step's angles
    # the same steps with the angles left as they are (the same sets every step)
    for _ in range(args.warmup):
        b.launch()
    b.flush()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.launch()
    b.flush()
    ctx.sync()
    wall_fixed = time.perf_counter() - t0
    # the parameter kernel alone: updates queued back to back
    t0 = time.perf_counter()
    for i in range(args.steps):
        b.update_angles(device_ptr=ptrs[i], n_sets=args.sets)
    ctx.sync()
    update_ms = 1e3 * (time.perf_counter() - t0) / args.steps
    # the last fresh step's energies against networks rebuilt from the gates' text on
    # the host (complex128 table executor)
    bchk = qa.QaoaBatch(ctx, dgf, 1, ref_map=True, precision=qa.QT_PLAIN_C128, order=qa.ORDER_GREEDY)
    bchk.set_angles(fresh_host[-1])
    bchk.launch()
    e_chk, _ = bchk.results()
    bchk.close()
    fresh_err = float(np.max(np.abs(e_fresh - e_chk) / np.abs(e_chk)))
    b.update_angles(sets)  # back to the sets the accuracy checks below use
    del fresh
    # HIP events on the engine's stream around whole launches, then around the
    # contraction kernel alone, each after launches that bring the clocks up
    b.time(args.warmup)
    launch_ms = b.time(max(args.steps, 200))
    kernel_ms = b.time_kernel(max(args.steps, 200))
    energies, _ = b.results()
    if sums_only:  # one launch that writes every edge's value, for the accuracy checks
        b.sums_only(False)
        if graph:
            b.uncapture()
        b.launch()
        energies_v, zz = b.results()
        sums_only_bitwise = bool(np.array_equal(energies, energies_v))
    else:
        energies_v, zz = b.results()
        sums_only_bitwise = None
    energies = energies_v
    info = b.info()
    value = ne * args.sets * args.steps / wall
    value_fixed = ne * args.sets * args.steps / wall_fixed

    # ---- accuracy against complex128 on the same plan, and against the reference ----
    b64 = qa.QaoaBatch(ctx, dgf, 1, ref_map=True, precision=qa.QT_PLAIN_C128, order=qa.ORDER_GREEDY)
    b64.set_angles(sets)
    b64.launch()
    e64, zz64 = b64.results()
    rel = np.abs(zz - zz64) / np.maximum(np.abs(zz64), 1e-3)
    edge_err = float(np.max(rel))
    within = float(np.mean(rel <= 1e-5))
    energy_err = float(np.max(np.abs(energies - e64) / np.abs(e64)))
    ref_err = float(np.max(np.abs(zz[0] - zz_ref) / np.maximum(np.abs(zz_ref), 1e-3)))
    c128_vs_ref = float(np.max(np.abs(zz64[0] - zz_ref) / np.maximum(np.abs(zz_ref), 1e-15)))

    # ---- every batched form on the same sets: error against complex128 and kernel time ----
    by_form = {}
    if not args.no_forms:
        for name, prec in forms.items():
            if name == args.form:
                continue
            if prec == qa.QT_PLAIN_C128:
                bf, e_f, zz_f = b64, e64, zz64
            else:
                bf = qa.QaoaBatch(ctx, dgf, 1, ref_map=True, precision=prec, order=qa.ORDER_GREEDY)
                bf.set_angles(sets)
                bf.launch()
                e_f, zz_f = bf.results()
            l_f, w_f, _ = bf.autotune()
            ms_f = bf.time(5)
            spec_ms = None
            if prec == qa.QT_PLAIN_C128 and specialised:  # complex128 also has generated code
                bf.compile(hermitian=not args.no_hermitian)
                bf.launch()
                ctx.sync()
                spec_ms = bf.time(10)
                bf.config(l_f, w_f)
            rel_f = np.abs(zz_f - zz64) / np.maximum(np.abs(zz64), 1e-3)
            by_form[name] = {
                "edge_evals_per_s": ne * args.sets / (ms_f * 1e-3), "kernel_ms": ms_f, "executor": [l_f, w_f],
                "max_edge_rel_err_vs_c128": float(np.max(rel_f)), "edges_within_1e-5": float(np.mean(rel_f <= 1e-5)),
                "max_energy_rel_err_vs_c128": float(np.max(np.abs(e_f - e64) / np.abs(e64)))}
            if spec_ms is not None:
                by_form[name]["specialised_launch_ms"] = spec_ms
                by_form[name]["specialised_edge_evals_per_s"] = ne * args.sets / (spec_ms * 1e-3)
            if bf is not b64:
                bf.close()
    b64.close()

    # ---- drop-in complex128 path, one energy evaluation per step ----
    plain = None
    if not args.no_plain:
        n_plain = 5
        qa.qaoa_energy_plain(ctx, dgf, 1, True, list(bg_ref), seed=0)
        ctx.sync()
        t0 = time.perf_counter()
        for s in range(n_plain):
            qa.qaoa_energy_plain(ctx, dgf, 1, True, list(sets[s]), seed=s)
        ctx.sync()
        tp = time.perf_counter() - t0
        plain = {"value": ne * n_plain / tp, "unit": "edge-evals/s", "ms_per_energy": 1e3 * tp / n_plain,
                 "order": "reference stochastic (seeded), one pairwise contraction per launch"}

    flops_per_launch = 8.0 * info["cmacs"]
    algorithmic_tflops = flops_per_launch / (kernel_ms * 1e-3) / 1e12
    n_networks = ne * args.sets
    elem = 16 if args.form == "c128" else 8
    if specialised:
        # the generated code drops terms with exact-zero constant factors and values
        # that never reach a result: its executed FLOPs (2 per FMA) face the VALU peak
        _, fmas = b.specialised()
        executed_flops = 2.0 * fmas
        achieved = executed_flops / (kernel_ms * 1e-3) / 1e12
        bound, peak = "valu", VALU_PEAK_TFLOPS
        kernel_name = ("qt_batch_specialised_sums (generated straight-line FMA chains, one lane per angle set and "
                       "slice of edges, energies only)" if sums_only else
                       "qt_batch_specialised (generated straight-line FMA chains, one lane per network)")
    else:
        executed_flops = flops_per_launch
        achieved = algorithmic_tflops
        bound, peak = "valu", VALU_PEAK_TFLOPS if args.form == "c128" else FP32_PEAK_TFLOPS
        kernel_name = "k_batch_grouped (persistent table executor)" if lanes < 64 else "k_batch (persistent table executor)"
    # algorithmic HBM bytes of one launch: every instance's parameter block read once
    # per angle set, one result written per network, the previous launch's results
    # read by the deferred energy sums, one FP64 energy per angle set
    param_bytes = args.sets * 32 * elem  # the 32 p parameter elements of an angle set (p = 1)
    if sums_only:  # 16 partial sums per set written, read back by the next launch's sums
        alg_bytes = param_bytes + args.sets * 16 * 8 * 2 + args.sets * 8
    else:
        alg_bytes = param_bytes + n_networks * 16 * (2 if specialised else 1) + args.sets * 8
    pmc = PMC_BYTES_PER_NETWORK.get(("specialised_sums" if sums_only else "specialised" if specialised else "tables",
                                     args.form, args.sets))
    res = {
        "metric": METRIC,
        "value": value, "unit": "edge-evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64" if args.form != "c128" else "c128", "data": "synthetic",
        "config": {"workload": "qaoa_p1_maxcut_3regular_n30_edges45", "angle_sets_per_step": args.sets,
                   "edge_networks_per_step": n_networks, "form": args.form,
                   "order": "greedy", "executor": "specialised" if specialised else "tables",
                   "conjugate_symmetry": bool(specialised and not args.no_hermitian),
                   "energies_only": bool(sums_only),
                   "executor_lanes": lanes, "wgs_per_cu": wgs,
                   "global_scratch": info["global_scratch"] and not specialised,
                   "staged_leaves": info["staged_leaves"] and not specialised},
        "compile_ms": compile_ms, "compile_error": compile_error, "table_executor_ms_per_launch": tables_ms,
        "hip_graph": graph,
        "launch_ms_direct": direct_ms, "launch_ms_graph": graph_ms,
        "angles": "fresh every step: Rz / Rx blocks rebuilt on the device from a new array of angle sets "
                  "in HBM before each launch (one array per step)",
        "fresh_angles": {"param_update_ms": update_ms, "max_energy_rel_err_vs_host_rebuild_c128": fresh_err,
                         "compared": f"the last step's {args.sets} energies against networks rebuilt from the gates' "
                                     "text on the host (complex128 table executor)"},
        "fixed_angles": {"value": value_fixed, "ms_per_step": 1e3 * wall_fixed / args.steps,
                         "note": "the same angle sets every step (no parameter update)"},
        "latency_steps": len(lat),
        "p50_step_ms": 1e3 * float(np.percentile(lat, 50)), "p99_step_ms": 1e3 * float(np.percentile(lat, 99)),
        "accuracy": {"max_edge_rel_err_vs_c128": edge_err, "edges_within_1e-5": within,
                     "max_energy_rel_err_vs_c128": energy_err,
                     "max_edge_rel_err_vs_reference": ref_err, "c128_batched_vs_reference": c128_vs_ref,
                     "floor": "|d| / max(|x|, 1e-3)",
                     "energies_only_equal_summed_values": sums_only_bitwise,
                     "compared": f"{args.sets} angle sets x {ne} edges against the complex128 table executor; "
                                 f"set 0 against the reference's own values"},
        "other_forms": by_form,
        "plain_c128_gpu": plain,
        "speedup_vs_plain_c128": (value / plain["value"]) if plain else None,
        "roofline": {"bound": bound, "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": pmc[0] * n_networks if pmc else None,
                     "traffic_source": pmc[1] if pmc else None,
                     "kernel": kernel_name, "kernel_ms_per_launch": kernel_ms,
                     "launch_ms_with_group_sums": launch_ms,
                     "executed_flops_per_launch": executed_flops,
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "hbm_achieved_gbs": alg_bytes / (kernel_ms * 1e-3) / 1e9,
                     "hbm_frac": alg_bytes / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "peak_note": "VALU FMA rate without packed FP32 (FP64 = FP32 = 78.6 TFLOP/s on MI355X); "
                                  "the kernel issues no MFMA",
                     "dense_equivalent_flops_per_launch": flops_per_launch,
                     "dense_equivalent_tflops": algorithmic_tflops},
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(dgf, bg_ref)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
