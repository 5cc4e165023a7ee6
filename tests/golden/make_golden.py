#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ by running the qTorch reference
compiled from its own sources (oracle/_ref/ref_driver, built by
oracle/build_ref.sh) in this container.

Every case records the input files (QASM + measurement text; the reference's own
sample files are copied as data), the contraction sequence the reference chose
(its `mCreatedFrom` record, src/preprocess.h:41-46) or the line-graph order it was
given, and the value it printed at %.17g.  Only this script reads the reference
(where its sources are); the tests read the committed fixtures.

The QAOA light cones and the per-edge QASM text F_p writes come from the
reference's own ExtraData and emission functions (src/maxcut.h:31-233) through the
driver's `cones` command: `cones.json` holds, per graph and p, every edge's
qubitsNeeded and realIterations, and per angle set the SHA-256 of every edge's
QASM text (the full text for a few edges).

Usage:  python tests/golden/make_golden.py           (needs oracle/_ref built)
        python tests/golden/make_golden.py --cones   (rewrites cones.json only)
"""
from __future__ import annotations

import hashlib
import json
import math
import os
import random
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("QTORCH_REFERENCE", "/root/reference")
DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_driver_O2")
DRIVER_O0 = os.path.join(REPO, "oracle", "_ref", "ref_driver")
SAMPLES = os.path.join(HERE, "samples")

sys.path.insert(0, REPO)
from oracle import qtoracle as qo  # noqa: E402


def run(args, cwd):
    out = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError(f"{args}: rc={out.returncode}\n{out.stdout}\n{out.stderr}")
    res = {"seq": []}
    for line in out.stdout.splitlines():
        parts = line.split()
        if parts[0] == "VALUE":
            res["value"] = [float(parts[1]), float(parts[2])]
        elif parts[0] == "FLOPS":
            res["flops"] = int(parts[1])
        elif parts[0] == "SEQ":
            res["seq"].append([int(parts[1]), int(parts[2])])
        elif parts[0] == "NWIRES":
            res["nwires"] = int(parts[1])
        elif parts[0] == "REDUCED_NODES":
            res["reduced_nodes"] = int(parts[1])
    return res


def write(path, text):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write(text)


def min_fill_order(nw, edges):
    """Min-fill elimination order of the line graph (vertices = wires), ties by id.
    Used only to give the reference's LGContract a sensible order to replay."""
    adj = [set() for _ in range(nw)]
    for a, b in edges:
        if a != b:
            adj[a].add(b)
            adj[b].add(a)
    alive = set(range(nw))
    order = []
    while alive:
        best, best_fill = None, None
        for v in sorted(alive):
            nb = list(adj[v])
            fill = 0
            for i in range(len(nb)):
                for j in range(i + 1, len(nb)):
                    if nb[j] not in adj[nb[i]]:
                        fill += 1
            if best_fill is None or fill < best_fill:
                best, best_fill = v, fill
        nb = list(adj[best])
        for i in range(len(nb)):
            for j in range(i + 1, len(nb)):
                adj[nb[i]].add(nb[j])
                adj[nb[j]].add(nb[i])
        for u in nb:
            adj[u].discard(best)
        alive.discard(best)
        order.append(best + 1)
    return order


def user_sequence_from_record(record, n_orig):
    """Turns a recorded mCreatedFrom list into original-node pairs the way the
    reference's own testUserDefinedSequence does (src/tests.cpp:797-825)."""
    rep = {}
    added = set()
    out = []
    for i, (a, b) in enumerate(record):
        node_id = n_orig + i
        t1, t2 = a, b
        while t1 in rep or t2 in rep:
            t1 = rep.get(t1, t1)
            t2 = rep.get(t2, t2)
        if (t1, t2) in added or t1 >= n_orig or t2 >= n_orig:
            continue
        out.append((t1, t2))
        added.add((t1, t2))
        added.add((t2, t1))
        rep[node_id] = a
    return out


def qbb_out_text(order):
    """The qbb.out shape LineGraph::LGContract parses (src/LineGraph.h:318-333)."""
    return (" Line graph min-fill ordering written by tests/golden/make_golden.py\n"
            " The optimal ordering is \n" + " ".join(str(o) for o in order) + " \n")


def qft8_qasm():
    """Textbook 8-qubit QFT: H, CPHASE(2 pi / 2^k) for k = 2..8 (control j+k-1 on
    target j), then the bit-reversal swaps.  CPHASE angles are parsed with stod."""
    lines = ["8"]
    for j in range(8):
        lines.append(f"H {j}")
        for k in range(2, 8 - j + 1):
            lines.append(f"CPHASE {2.0 * math.pi / 2 ** k:.17g} {j + k - 1} {j}")
    for j in range(4):
        lines.append(f"SWAP {j} {7 - j}")
    return "\n".join(lines) + "\n"


def regular_graph(n, d, seed):
    """Seeded random d-regular simple graph by the pairing model with restarts."""
    rng = random.Random(seed)
    while True:
        stubs = [v for v in range(n) for _ in range(d)]
        rng.shuffle(stubs)
        edges = set()
        ok = True
        for i in range(0, len(stubs), 2):
            a, b = stubs[i], stubs[i + 1]
            if a == b or (min(a, b), max(a, b)) in edges:
                ok = False
                break
            edges.add((min(a, b), max(a, b)))
        if ok:
            e = sorted(edges)
            rng.shuffle(e)
            return [(a, b) if rng.random() < 0.5 else (b, a) for a, b in e]


def ref_cones(dgf_path, p, bg):
    """Runs the reference's ExtraData(p, dgf) and its emission for one angle set:
    [(qubits_needed, real_iterations, qasm_text)] per edge."""
    out = subprocess.run([DRIVER, "cones", dgf_path, str(p)] + ["%.17g" % x for x in bg],
                         capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        raise RuntimeError(f"cones {dgf_path} p={p}: rc={out.returncode}\n{out.stdout}\n{out.stderr}")
    cones, cur, qasm = [], None, None
    for line in out.stdout.splitlines(keepends=True):
        if qasm is not None:
            if line == "QASM_END\n":
                cur[2] = "".join(qasm)
                qasm = None
            else:
                qasm.append(line)
            continue
        parts = line.split()
        if parts[0] == "CONE":
            assert int(parts[1]) == len(cones)
            cur = [int(parts[2]), [], None]
            cones.append(cur)
        elif parts[0] == "IT":
            cur[1].append([int(parts[1]), int(parts[2])])
        elif parts[0] == "QASM_BEGIN":
            qasm = []
    return [tuple(c) for c in cones]


# graphs and angle sets whose light cones and emitted text are pinned; the angle sets
# exercise ostream's 6-significant-digit printing (exponent forms, rounding, signs)
CONE_GRAPHS = ["graphs/3reg_n30_s2024.dgf", "samples/3regRand30Node50.dgf",
               "samples/4regRand30Node26.dgf"]
CONE_ANGLES = {
    1: [[0.3, -1.1], [1.2345678e-5, -123456.789], [0.1234565, 2.5e-7], [-0.0, 3.14159265358979]],
    2: [[0.392699, 0.392699, 0.785399, 0.785399], [0.3, 0.2, -0.7, 1.1],
        [1e-7, -2.0000005, 98765.4321, -1e-12]],
    3: [[0.1, 0.2, 0.3, -0.4, -0.5, 0.6]],
}


def golden_qaoa_angles():
    """The angle set of the qaoa3_* cases: gamma ~ U[0, pi], beta ~ U[0, pi/2] seeded,
    as [beta, -2 gamma] (exp(-i gamma ZZ), RX(2 beta) in the reference's terms)."""
    arng = random.Random(77)
    gamma = arng.uniform(0, math.pi)
    beta = arng.uniform(0, math.pi / 2)
    return gamma, beta, [beta, -2.0 * gamma]


def write_cones():
    doc = {"generator": "tests/golden/make_golden.py --cones",
           "reference": "ExtraData + outputInitialPlusStateToFile + applyU_CsThenU_Bs "
                        "(src/maxcut.h:31-233) through oracle/_ref/ref_driver_O2 cones",
           "graphs": []}
    for rel in CONE_GRAPHS:
        path = os.path.join(HERE, rel)
        for p in (1, 2, 3):
            if p == 3 and not rel.startswith("graphs/"):
                continue
            entry = {"dgf": rel, "p": p, "sets": []}
            sets = list(CONE_ANGLES[p])
            if rel == "graphs/3reg_n30_s2024.dgf" and p == 1:
                sets.insert(0, golden_qaoa_angles()[2])
            for bg in sets:
                cones = ref_cones(path, p, bg)
                if "qubits" not in entry:
                    entry["qubits"] = [c[0] for c in cones]
                    entry["iterations"] = [c[1] for c in cones]
                assert entry["qubits"] == [c[0] for c in cones]
                assert entry["iterations"] == [c[1] for c in cones]
                entry["sets"].append({
                    "betas_gammas": bg,
                    "sha256": [hashlib.sha256(c[2].encode()).hexdigest()[:32] for c in cones],
                    "text": {str(e): cones[e][2] for e in (0, len(cones) - 1)},
                })
            doc["graphs"].append(entry)
            print(f"cones {rel} p={p}: {len(entry['qubits'])} edges, qubits "
                  f"{min(entry['qubits'])}-{max(entry['qubits'])}", flush=True)
    with open(os.path.join(HERE, "cones.json"), "w") as f:
        json.dump(doc, f, indent=0)


def main():
    if not os.path.exists(DRIVER):
        sys.exit("build oracle/_ref first: oracle/build_ref.sh")
    if "--cones" in sys.argv:
        write_cones()
        return
        sys.exit("build oracle/_ref first: oracle/build_ref.sh")
    os.makedirs(SAMPLES, exist_ok=True)
    # reference sample data used by the reference's own tests
    for name in ["test_JW.qasm", "4regRand20Node1-p1.qasm", "4regRand20Node5-p1.qasm",
                 "measure125.txt", "rand-nq6-cn2-d10_rx.qasm", "rand-nq6-cn2-d10_rxz.qasm",
                 "rand-nq6-cn2-d10_rxyz.qasm", "bell_pair.qasm", "tofolli.qasm", "s.gate",
                 "t.gate", "tdagger.gate", "catStateEightQubits.qasm", "catState100Qubits.qasm",
                 "qft8.qasm", "qft4.qasm", "teleportation.qasm", "3regRand30Node50.dgf",
                 "4regRand30Node26.dgf", "cz.gate", "zee.gate", "arbitraryGateExample.qasm"]:
        shutil.copy(os.path.join(REF, "Samples", name), os.path.join(SAMPLES, name))
    os.makedirs(os.path.join(SAMPLES, "arbitrary2qubitmatrixset"), exist_ok=True)
    for i in range(0, 100, 10):
        shutil.copy(os.path.join(REF, "Samples", "arbitrary2qubitmatrixset", f"matrix{i}.gate"),
                    os.path.join(SAMPLES, "arbitrary2qubitmatrixset", f"matrix{i}.gate"))

    cases = []
    work = tempfile.mkdtemp(prefix="golden_")

    def add(name, qasm_rel, meas, mode="stoch", extra=None, qasm_text=None):
        qpath = os.path.join(HERE, qasm_rel)
        if qasm_text is not None:
            write(qpath, qasm_text)
        mpath = os.path.join(work, name + ".meas")
        write(mpath, meas)
        case = {"name": name, "qasm": qasm_rel, "measure": meas, "mode": mode}
        if mode == "stoch":
            r = run([DRIVER, "stoch", qpath, mpath], cwd=HERE)
        elif mode == "rstoch":
            r = run([DRIVER, "rstoch", qpath, mpath], cwd=HERE)
        elif mode in ("lg", "rlg"):
            reduce_flag = "1" if mode == "rlg" else "0"
            nw = run([DRIVER, "nwires", qpath, mpath, reduce_flag], cwd=HERE)["nwires"]
            net = qo.Network.from_files(qpath, mpath)
            net.base_dir = HERE
            if mode == "rlg":
                net.reduce_circuit()
            wires, edges = net.line_graph_wires()
            assert len(wires) == nw, (name, len(wires), nw)
            order = min_fill_order(nw, [(a.wire_id, b.wire_id) for a, b in edges])
            qbb = os.path.join(work, name + ".qbb")
            write(qbb, qbb_out_text(order))
            r = run([DRIVER, "lg", qpath, mpath, qbb, reduce_flag], cwd=HERE)
            case["qbb_order"] = order
            case["cnf"] = qo.lg_cnf_text(wires, edges)
        elif mode == "user":
            rec = run([DRIVER, "stoch", qpath, mpath], cwd=HERE)
            n_orig = len(qo.Network.from_files(qpath, mpath).all_nodes)
            user = user_sequence_from_record(rec["seq"], n_orig)
            seqf = os.path.join(work, name + ".seq")
            write(seqf, "\n".join(f"{a} {b}" for a, b in user))
            r = run([DRIVER, "user", qpath, mpath, seqf], cwd=HERE)
            r["seq"] = user
        else:
            raise ValueError(mode)
        case.update(r)
        if extra:
            case.update(extra)
        cases.append(case)
        print(f"{name:40s} {mode:6s} value={case['value'][0]:.17g} flops={case.get('flops')}",
              flush=True)

    # --- src/tests.cpp known-answer circuits -------------------------------------------
    for m, tag in [("Z T T T", "ZTTT"), ("T Z T T", "TZTT"), ("Y X X Y", "YXXY")]:
        add(f"jw_{tag}", "samples/test_JW.qasm", m, "stoch")
        add(f"jw_{tag}_reduced", "samples/test_JW.qasm", m, "rstoch")
    add("jw_ZTTT_lg_reduced", "samples/test_JW.qasm", "Z T T T", "rlg")
    add("jw_ZTTT_lg", "samples/test_JW.qasm", "Z T T T", "lg")
    add("jw_T_user", "samples/test_JW.qasm", "T", "user")
    with open(os.path.join(SAMPLES, "measure125.txt")) as f:
        m125 = f.read()
    add("large_4reg20_1", "samples/4regRand20Node1-p1.qasm", m125, "stoch",
        {"known": 0.0035757, "tol": 1e-5})
    add("large_4reg20_5", "samples/4regRand20Node5-p1.qasm", m125, "stoch",
        {"known": 0.00255591, "tol": 1e-5})
    add("large_4reg20_5_lg_reduced", "samples/4regRand20Node5-p1.qasm", m125, "rlg",
        {"known": 0.00255591, "tol": 1e-5})
    rot = [("Ry 0.1 0", "X", 0.0998334), ("Ry 0.1 0", "Y", 0.0), ("Ry 0.1 0", "Z", 0.995004),
           ("Ry -0.1 0", "X", -0.0998334166468), ("Ry -0.1 0", "Y", 0.0),
           ("Ry -0.1 0", "Z", 0.995004165278),
           ("H 0\nRz 0.1 0", "X", 0.995004165278), ("H 0\nRz 0.1 0", "Y", 0.0998334166468),
           ("H 0\nRz 0.1 0", "Z", 0.0), ("H 0\nRz -0.1 0", "X", 0.995004165278),
           ("H 0\nRz -0.1 0", "Y", -0.0998334166468), ("H 0\nRz -0.1 0", "Z", 0.0),
           ("Rx 0.1 0", "X", 0.0), ("Rx 0.1 0", "Y", -0.0998334166468),
           ("Rx 0.1 0", "Z", 0.995004165278), ("Rx -0.1 0", "X", 0.0),
           ("Rx -0.1 0", "Y", 0.0998334166468), ("Rx -0.1 0", "Z", 0.995004165278)]
    for i, (body, m, known) in enumerate(rot):
        add(f"rotation_{i}", f"circuits/rotation_{i}.qasm", m, "stoch",
            {"known": known, "tol": 1e-5}, qasm_text="1\n" + body + "\n")
    write(os.path.join(HERE, "circuits", "hadtest.gate"), "0.707107 0.707107 0.707107 -0.707107\n")
    add("had_def1_0", "circuits/had_def1.qasm", "0", "stoch", {"known": 0.5, "tol": 1e-5},
        qasm_text="1\ndef1 HadamardTester circuits/hadtest.gate\nHadamardTester 0\n")
    add("had_def1_1", "circuits/had_def1.qasm", "1", "stoch", {"known": 0.5, "tol": 1e-5})
    for kind in ["rx", "rxz", "rxyz"]:
        for q in [0, 2, 4]:
            for m in ["X", "Y", "Z"]:
                meas = " ".join(m if i == q else "T" for i in range(6))
                add(f"rand_{kind}_{m}{q}", f"samples/rand-nq6-cn2-d10_{kind}.qasm", meas, "stoch")
    add("rand_rxyz_Z0_lg", "samples/rand-nq6-cn2-d10_rxyz.qasm", "Z T T T T T", "lg")
    for inp in ["00", "01", "10", "11"]:
        body = "".join(f"X {i}\n" for i, b in enumerate(inp) if b == "1")
        with open(os.path.join(SAMPLES, "bell_pair.qasm")) as f:
            bell = f.read().split("\n", 1)
        for m in ["0 0", "0 1", "1 0", "1 1"]:
            add(f"bell_{inp}_{m.replace(' ', '')}", f"circuits/bell_{inp}.qasm", m, "stoch",
                qasm_text=bell[0] + "\n" + body + bell[1])
    for m in ["0 " * 8, "1 " * 8]:
        add(f"cat8_{m[0]}", "samples/catStateEightQubits.qasm", m, "stoch",
            {"known": 0.5, "tol": 1e-5})
    add("cat100_0", "samples/catState100Qubits.qasm", "0 " * 100, "stoch", {"known": 0.5, "tol": 1e-5})
    with open(os.path.join(SAMPLES, "tofolli.qasm")) as f:
        tof = f.read().split("\n", 1)
    tof_body = tof[1].replace("Samples/", "samples/")
    for inp in ["000", "001", "010", "011", "100", "101", "110", "111"]:
        outb = inp[:2] + (str(1 - int(inp[2])) if inp[:2] == "11" else inp[2])
        x = "".join(f"X {i}\n" for i, b in enumerate(inp) if b == "1")
        add(f"toffoli_{inp}", f"circuits/toffoli_{inp}.qasm", " ".join(outb), "stoch",
            {"known": 1.0, "tol": 1e-5}, qasm_text=tof[0] + "\n" + x + tof_body)
    for i in range(0, 100, 10):
        for inp in ["00", "11"]:
            x = "".join(f"X {q}\n" for q, b in enumerate(inp) if b == "1")
            for m in ["0 0", "1 1", "Y Y", "Y T", "T Y"]:
                add(f"arb2q_{i}_{inp}_{m.replace(' ', '')}", f"circuits/arb2q_{i}_{inp}.qasm", m,
                    "stoch", qasm_text=f"2\n{x}def2 TEST samples/arbitrary2qubitmatrixset/matrix{i}.gate\nTEST 0 1\n")
    add("disconnected_a", "circuits/disconnected_a.qasm", "0 0", "stoch", {"known": 1.0, "tol": 1e-8},
        qasm_text="2\n")
    add("disconnected_b", "circuits/disconnected_b.qasm", "0 0", "stoch", {"known": 0.25, "tol": 1e-8},
        qasm_text="2\nH 0\nH 1\n")
    add("qft8_sample_lg", "samples/qft8.qasm", "0 0 0 0 0 0 0 0", "lg")

    # --- configuration 1: textbook QFT8 on |0^8>, LineGraph (min-fill) order -------------
    rng = random.Random(1234)
    bits = "".join(rng.choice("01") for _ in range(8))
    add("qft8_bits", "circuits/qft8_textbook.qasm", " ".join(bits), "lg",
        {"known": 1.0 / 256, "bits": bits}, qasm_text=qft8_qasm())
    add("qft8_bits_stoch", "circuits/qft8_textbook.qasm", " ".join(bits), "stoch",
        {"known": 1.0 / 256, "bits": bits})
    for x in range(0, 256, 17):
        b = format(x, "08b")[::-1]
        add(f"qft8_all_{x}", "circuits/qft8_textbook.qasm", " ".join(b), "lg",
            {"known": 1.0 / 256, "bits": b})

    # --- configuration 2: QAOA p=1 MaxCut, 3-regular 30 nodes, per-edge light cones -------
    edges = regular_graph(30, 3, seed=2024)
    write(os.path.join(HERE, "graphs", "3reg_n30_s2024.dgf"),
          "c seeded random 3-regular graph, 30 vertices (tests/golden/make_golden.py)\n" +
          "".join(f"e {a} {b}\n" for a, b in edges))
    gamma, beta, bg = golden_qaoa_angles()
    dgf3 = os.path.join(HERE, "graphs", "3reg_n30_s2024.dgf")
    for compat in (True, False):
        # the reference map's text is the reference's own emission; the corrected map
        # (pair.second -> 1 as well) has no reference counterpart and is the oracle's
        cones = ref_cones(dgf3, 1, bg) if compat else qo.light_cones(edges, 1, reference_qubit_map=False)
        tag = "ref" if compat else "fix"
        for e, cone in enumerate(cones):
            nq, real = cone[0], cone[1]
            q = cone[2] if compat else qo.qaoa_edge_qasm(nq, real, 1, bg)
            add(f"qaoa3_{tag}_e{e}", f"circuits/qaoa3_{tag}_e{e}.qasm", qo.qaoa_edge_measure(nq),
                "stoch", {"edge": e, "gamma": gamma, "beta": beta, "betas_gammas": bg,
                          "qubit_map": tag}, qasm_text=q)

    # --- a sample graph from the reference, p=2, reference map ---------------------------
    cones2 = ref_cones(os.path.join(SAMPLES, "3regRand30Node50.dgf"), 2,
                       [0.392699, 0.392699, 0.785399, 0.785399])
    for e in range(0, len(cones2), 9):
        nq, q = cones2[e][0], cones2[e][2]
        add(f"qaoa_sample_p2_e{e}", f"circuits/qaoa_sample_p2_e{e}.qasm",
            qo.qaoa_edge_measure(nq), "stoch", {"edge": e}, qasm_text=q)

    # the -O2 driver and the reference's own -O0 flags give the same bits
    for c in cases[:3]:
        seqf = os.path.join(work, "chk.seq")
        write(seqf, "\n".join(f"{a} {b}" for a, b in c["seq"]))
        mpath = os.path.join(work, "chk.meas")
        write(mpath, c["measure"])
        r0 = run([DRIVER_O0, "given", os.path.join(HERE, c["qasm"]), mpath, seqf], cwd=HERE)
        r2 = run([DRIVER, "given", os.path.join(HERE, c["qasm"]), mpath, seqf], cwd=HERE)
        assert r0["value"] == r2["value"] == c["value"], (c["name"], r0, r2, c["value"])

    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump({"generator": "tests/golden/make_golden.py",
                   "reference_build": "oracle/build_ref.sh (ref_driver_O2)",
                   "cases": cases}, f, indent=0)
    shutil.rmtree(work)
    print(f"wrote {len(cases)} cases")
    write_cones()


if __name__ == "__main__":
    main()
