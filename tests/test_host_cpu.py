"""CPU tests of the product's host side (no GPU): the C-ABI libraries load and export
every declared symbol, and the C++ host logic (parsing, ContractNodes bookkeeping,
ReduceCircuit, orderings, LineGraph, maxcut light cones) agrees with the oracle and
with the reference's recorded runs.  Plan-only networks run the ordering logic and
record every step without computing values."""
import ctypes
import math
import os

import numpy as np
import pytest

import qtorch_amd as qa
from oracle import qtoracle as qo
from tests import statevector as sv

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


def _qasm(case):
    with open(os.path.join(GOLDEN_DIR, case["qasm"])) as f:
        return f.read()


@pytest.mark.parametrize("header,lib", [("qtorch_hip.h", qa._native.hip), ("qtorch_host.h", qa._native.host)])
def test_every_declared_symbol_is_exported(header, lib):
    names = qa.declared_symbols(header)
    assert len(names) >= 10
    for n in names:
        assert hasattr(lib, n), n


def test_no_gpu_context_fails_loudly():
    """Without a GPU the engine must refuse to create a context (no CPU fallback)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(qa.DeviceFailure):
        qa.Context(0)


def _plan_only_replay(case):
    net = qa.Network(_qasm(case), case["measure"], None, base_dir=GOLDEN_DIR)
    mode = case["mode"]
    if mode == "stoch":
        net.contract_given(case["seq"])
    elif mode == "rstoch":
        net.reduce()
        net.contract_given(case["seq"])
    elif mode in ("lg", "rlg"):
        if mode == "rlg":
            net.reduce()
        assert net.linegraph_cnf() == case["cnf"]
        net.linegraph_contract(case["qbb_order"])
    elif mode == "user":
        net.contract_user(case["seq"])
    return net


def test_plan_only_replay_matches_reference_counters(golden):
    """Replaying the reference's own sequences, the host reproduces its complex-MAC
    counter exactly (every step's ranks are right) and, for given sequences, the
    very same mCreatedFrom record (node ids and pairing)."""
    for case in golden:
        net = _plan_only_replay(case)
        assert net.flops() == case["flops"], case["name"]
        assert net.counts()["done"], case["name"]
        if case["mode"] == "stoch":
            assert net.record() == [tuple(p) for p in case["seq"]], case["name"]


def test_stochastic_order_is_valid_and_seeded(golden):
    """The seeded stochastic ordering (ParallelContract) completes, is reproducible
    for a seed, and its record replays in the oracle to the same counter and to the
    state-vector value."""
    names = {"jw_ZTTT", "rand_rxyz_Z0", "bell_00_00", "qaoa3_ref_e3", "cat8_0", "toffoli_110",
             "qft8_bits_stoch", "arb2q_30_11_YY"}
    seen = 0
    for case in golden:
        if case["name"] not in names:
            continue
        seen += 1
        q = _qasm(case)
        a = qa.Network(q, case["measure"], None, base_dir=GOLDEN_DIR)
        a.contract_stochastic(1234)
        b = qa.Network(q, case["measure"], None, base_dir=GOLDEN_DIR)
        b.contract_stochastic(1234)
        assert a.record() == b.record()
        onet = qo.Network(q, case["measure"], base_dir=GOLDEN_DIR)
        val = onet.contract_given_sequence(a.record())
        assert onet.flops == a.flops()
        psi, n = sv.simulate(q, base_dir=GOLDEN_DIR)
        ref = sv.expectation(psi, n, case["measure"])
        assert abs(val - ref) < 1e-12 * max(1, abs(ref)), case["name"]
    assert seen == len(names)


def test_reduce_then_stochastic(golden):
    for case in golden:
        if case["mode"] != "rstoch":
            continue
        net = qa.Network(_qasm(case), case["measure"], None, base_dir=GOLDEN_DIR)
        net.reduce()
        assert net.counts()["uncontracted"] == case["reduced_nodes"]
        net.contract_stochastic(7)
        onet = qo.Network(_qasm(case), case["measure"], base_dir=GOLDEN_DIR)
        onet.reduce_circuit()
        k = _n_reduce_steps(onet)
        rec = net.record()
        assert rec[:k] == [n.created_from for n in onet.all_nodes if n.created_from != (0, 0)]
        val = onet.contract_given_sequence(rec[k:])
        assert abs(val - complex(*case["value"])) < 1e-12


def _n_reduce_steps(onet):
    return sum(1 for n in onet.all_nodes if n.created_from != (0, 0))


def test_linegraph_minfill_completes(golden):
    for case in golden:
        if case["mode"] not in ("lg", "rlg"):
            continue
        net = qa.Network(_qasm(case), case["measure"], None, base_dir=GOLDEN_DIR)
        if case["mode"] == "rlg":
            net.reduce()
        order = net.linegraph_minfill()
        assert sorted(order) == list(range(1, len(order) + 1))
        net.linegraph_contract(order)
        assert net.counts()["uncontracted"] == 1


def test_qaoa_edge_text_byte_identical_to_reference(ref_cones):
    """The product's per-edge F_p text (reference qubit map) equals the text the
    reference's ExtraData + emission wrote, for every edge, graph, p and angle set
    recorded in cones.json (src/maxcut.h:31-233, src/maxcut.cpp:172-189)."""
    import hashlib
    n = 0
    for g in ref_cones:
        with open(os.path.join(GOLDEN_DIR, g["dgf"])) as f:
            dgf = f.read()
        for s in g["sets"]:
            for e, want in enumerate(s["sha256"]):
                q, m = qa.qaoa_edge_text(dgf, g["p"], True, e, s["betas_gammas"])
                if str(e) in s["text"]:
                    assert q == s["text"][str(e)], (g["dgf"], g["p"], e)
                assert hashlib.sha256(q.encode()).hexdigest()[:32] == want, (g["dgf"], g["p"], e)
                assert m == "".join("Z " if i < 2 else "T " for i in range(g["qubits"][e]))
                n += 1
    assert n > 1000


def test_qaoa_edge_text_matches_golden_cases(golden):
    with open(os.path.join(GOLDEN_DIR, "graphs", "3reg_n30_s2024.dgf")) as f:
        dgf = f.read()
    for case in golden:
        if not case["name"].startswith("qaoa3_"):
            continue
        ref_map = case["qubit_map"] == "ref"
        q, m = qa.qaoa_edge_text(dgf, 1, ref_map, case["edge"], case["betas_gammas"])
        assert q == _qasm(case), case["name"]
        assert m == case["measure"]


def test_qaoa_light_cones_match_oracle():
    with open(os.path.join(GOLDEN_DIR, "samples", "4regRand30Node26.dgf")) as f:
        dgf = f.read()
    edges = qo.read_dgf(dgf)
    bg = [0.3, 0.2, -0.7, 1.1]
    for ref_map in (True, False):
        cones = qo.light_cones(edges, 2, reference_qubit_map=ref_map)
        for e in range(0, len(edges), 7):
            nq, real, _ = cones[e]
            q, m = qa.qaoa_edge_text(dgf, 2, ref_map, e, bg)
            assert q == qo.qaoa_edge_qasm(nq, real, 2, bg)
            assert m == qo.qaoa_edge_measure(nq)


def test_error_behaviour():
    with pytest.raises(qa.InvalidFileFormat):
        qa.Network("1\nFOO 0\n", "Z", None)
    with pytest.raises(qa.InvalidFileFormat):
        qa.Network("2\nCNOT 1 1\n", "Z Z", None)
    with pytest.raises(qa.InvalidFileFormat):
        qa.Network("1\nRx abc 0\n", "Z", None)
    with pytest.raises(qa.InvalidFile):
        qa.Network("1\ndef1 G /nonexistent/g.gate\nG 0\n", "Z", None)
    net = qa.Network("2\nH 0\nCNOT 0 1\n", "0 0", None)
    with pytest.raises(qa.InvalidUserContractionSequence):
        net.contract_user([(0, 2)])
    net = qa.Network("2\nH 0\nCNOT 0 1\n", "0 0", None)
    with pytest.raises(qa.ContractionFailure):
        net.contract_given([(0, 2)])
    # a rejected pair (not connected) returns -1 like the reference's nullptr
    net = qa.Network("2\nH 0\nCNOT 0 1\n", "0 0", None)
    assert net.contract_nodes(0, 1, 100) == -1


def test_disconnected_network_final_value_rule():
    """Two unentangled qubits: components contract to scalars, then scalar x scalar
    (src/Network.h:772, 961-969)."""
    net = qa.Network("2\nH 0\nH 1\n", "0 0", None)
    net.contract_stochastic(3)
    assert net.counts()["done"]
    rec = net.record()
    onet = qo.Network("2\nH 0\nH 1\n", "0 0")
    assert abs(onet.contract_given_sequence(rec) - 0.25) < 1e-15


def test_program_schedule_keeps_scratch_small():
    """Compiling a plan into a device program schedules its steps (same tree, same
    per-step arithmetic) to keep few intermediates live: the executor's scratch is
    at least the live set of the worst step under any order and, for the p = 1
    light cones, at most a few of the largest intermediates (peak rank 3 = 64
    elements); the MAC count is the plan's."""
    with open(os.path.join(GOLDEN_DIR, "graphs", "3reg_n30_s2024.dgf")) as f:
        dgf = f.read()
    ne, _ = qa.qaoa_num_edges(dgf)
    for e in range(ne):
        q, m = qa.qaoa_edge_text(dgf, 1, True, e, [0.3, 0.7])
        net = qa.Network(q, m, None)
        n_orig = net.counts()["original"]
        net.order(qa.ORDER_GREEDY, 0)
        plan = net.plan()
        st = net.program_stats()
        assert st["cmacs"] == net.flops() and st["n_steps"] == len(plan)
        lower = 0
        for i, (a, b, c, ra, rb, k, rc, _) in enumerate(plan):
            live = (4 ** rc if i + 1 < len(plan) else 0) + sum(4 ** r for op, r in ((a, ra), (b, rb)) if op >= n_orig)
            lower = max(lower, live)
        assert lower <= st["scratch_elems"] <= 256, (e, lower, st)
        net.close()


def test_serving_plan_routes_and_bytes():
    """Serving plans each request on the host: small light cones run in lane-grouped
    waves, larger programs in workgroup executors, networks with tensors beyond a
    program's 2^16 elements on the large-network path, whose plan is the one with the
    least work among the line-graph order and the circuit sweep (sliced to 28 to 30
    bits), whose planned bytes cover at least that plan's largest intermediate
    (complex128) and whose MAC count is that plan's."""
    from qtorch_amd import workloads as wl
    with open(os.path.join(GOLDEN_DIR, "graphs", "3reg_n30_s2024.dgf")) as f:
        dgf = f.read()
    p1 = qa.serve_plan(qa.SERVE_QAOA_ENERGY, dgf, 1)
    assert p1["route"] == "small" and p1["cmacs"] == 45 * 3840
    assert qa.serve_plan(qa.SERVE_QAOA_ENERGY, dgf, 2)["route"] == "medium"
    assert qa.serve_plan(qa.SERVE_AMPLITUDE, wl.qft_qasm(8), bits="01100101")["route"] == "small"
    q = wl.grid_random_circuit(4, 5, 14, seed=1)
    bits = wl.bitstring(20, 3)
    big = qa.serve_plan(qa.SERVE_AMPLITUDE, q, bits=bits)
    net = qa.Network(q, bits, None, dim_log2=1)
    net.order(qa.ORDER_LINEGRAPH, 0)
    assert int(net.plan()[:, 6].max()) > 16 and big["route"] == "eager"
    plans = [qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=28).plan()]
    plans += [qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_SWEEP, max_bits=mb).plan() for mb in (28, 29, 30)]
    work = [pl["cmacs_per_slice"] * pl["slices"] for pl in plans]
    chosen = plans[work.index(min(work))]
    assert big["cmacs"] == min(work) <= net.flops()
    assert big["planned_bytes"] >= 16 * 2 ** chosen["peak_bits"]
    with pytest.raises(qa.InvalidFunctionInput):
        qa.serve_plan(7, dgf, 1)
    with pytest.raises(qa.InvalidFunctionInput):
        qa.serve_plan(qa.SERVE_AMPLITUDE, wl.qft_qasm(4), bits="01x1")


def test_serving_plans_the_48_qubit_amplitude_by_sweep():
    """48 qubits at depth 14: in the line-graph order sliced to 28 bits, slices times
    MACs per slice is beyond 2^63 (the planner's count saturates rather than wrapping
    to a negative number the work limit would admit); the serving planner takes the
    circuit sweep instead (CZ split, sliced to fit half the default 16 GiB budget),
    whose work is positive and within the per-request limit of 2^44 complex MACs."""
    from qtorch_amd import workloads as wl
    q = wl.grid_random_circuit(6, 8, 14, seed=605)
    bits = wl.bitstring(48, 7)
    lg = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=28).plan()
    assert lg["cmacs_per_slice"] * lg["slices"] > 2 ** 63
    plan = qa.serve_plan(qa.SERVE_AMPLITUDE, q, bits=bits)
    sweep = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_SWEEP, max_bits=28).plan()
    assert plan["route"] == "eager" and 0 < plan["cmacs"] <= 2 ** 44
    assert plan["cmacs"] == sweep["cmacs_per_slice"] * sweep["slices"]


def test_nonzero_slices_listing():
    """nonzero_slices lists slices in increasing order, within range, a shorter
    request giving a prefix of a longer one; a grid circuit's CZ gates tie sliced
    wires, so some slices are left out."""
    from qtorch_amd import workloads as wl
    q = wl.grid_random_circuit(4, 4, 10, seed=3)
    sn = qa.SlicedNetwork(q, wl.bitstring(16, 5), dim_log2=1, order=qa.ORDER_GREEDY, max_bits=8)
    n = sn.plan()["slices"]
    full = sn.nonzero_slices(n)
    assert full == sorted(set(full)) and all(0 <= s < n for s in full)
    assert 0 < len(full) < n
    assert sn.nonzero_slices(3) == full[:3]
    with pytest.raises(qa.InvalidFunctionInput):
        sn.set_slices([n])
    with pytest.raises(qa.InvalidFunctionInput):
        sn.set_range(n, 1)
    sn.close()


def test_sliced_plan_bounds():
    """Slicing keeps every tensor of a slice within max_bits index bits; D^s slices of
    one shape; without slicing the per-slice MAC count is the plan's own."""
    from qtorch_amd import workloads as wl
    q = wl.grid_random_circuit(4, 5, 8, seed=2)
    bits = wl.bitstring(20, 8)
    whole = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=40).plan()
    net = qa.Network(q, bits, None, dim_log2=1)
    net.order(qa.ORDER_LINEGRAPH, 0)
    assert whole["sliced_wires"] == 0 and whole["cmacs_per_slice"] == net.flops()
    assert whole["peak_bits"] == int(net.plan()[:, 6].max()) or whole["peak_bits"] >= int(net.plan()[:, 6].max())
    for mb in (whole["peak_bits"] - 2, 10, 8):
        p = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=mb).plan()
        assert p["peak_bits"] <= mb and p["slices"] == 2 ** p["sliced_wires"] and p["sliced_wires"] >= 1
        assert p["cmacs_per_slice"] * p["slices"] >= whole["cmacs_per_slice"] // 4
    with pytest.raises(qa.InvalidFunctionInput):
        qa.SlicedNetwork(q, bits, dim_log2=3, max_bits=8)


def test_sliced_fusion_plan_counts():
    """The fused subtrees of an unsliced network, recounted from the plan: a step is
    small when its result has <= 12 index bits, its operands <= 14 and its complex
    MACs <= 2^16, and it is fused when its operands are leaves or fused steps; a
    subtree's root is a fused step whose consumer is not fused (or the last step)."""
    from qtorch_amd import workloads as wl
    q = wl.grid_random_circuit(4, 5, 8, seed=2)
    bits = wl.bitstring(20, 8)
    got = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=40).plan()
    net = qa.Network(q, bits, None, dim_log2=1)
    net.order(qa.ORDER_LINEGRAPH, 0)
    plan = net.plan()
    step_of = {int(r[2]): i for i, r in enumerate(plan) if int(r[2]) >= 0}
    fused, consumer = [], {}
    for i, (a, b, c, ra, rb, k, rc, _) in enumerate(plan.tolist()):
        ok = rc <= 12 and ra <= 14 and rb <= 14 and rc + k <= 16
        for op in (a, b):
            if op in step_of:
                consumer[step_of[op]] = i
                ok = ok and fused[step_of[op]]
        fused.append(ok)
    roots = [i for i in range(len(plan)) if fused[i] and (i not in consumer or not fused[consumer[i]])]
    assert got["fused_steps"] == sum(fused) > 0
    assert got["fused_subtrees"] == len(roots) >= 1


def test_max_time_stops_contractions_with_contraction_failure():
    """totTimer / maxTime (src/Network.h:50-51): once the started timer passes maxTime,
    the stochastic contraction stops before its next step and raises
    ContractionFailure (src/ContractionTools.h:271,332,380-386); a stopped timer never
    times out, and the default limit is the reference's 60 s."""
    q = _ghz_qasm(12)
    assert qa.get_max_time() == 60.0
    try:
        qa.set_max_time(0.0)
        qa.timer_start()
        net = qa.Network(q, None, None)
        with pytest.raises(qa.ContractionFailure):
            net.contract_stochastic(3)
        assert net.counts()["uncontracted"] == net.counts()["original"]  # stopped before the first step
        # ContractNodes itself checks between steps: a user sequence stops too
        net2 = qa.Network(q, None, None)
        with pytest.raises(qa.ContractionFailure):
            net2.contract_given([(0, 12)])
        qa.timer_reset()  # not started: no limit
        net3 = qa.Network(q, None, None)
        net3.contract_stochastic(3)
        assert net3.counts()["done"]
    finally:
        qa.timer_reset()
        qa.set_max_time(60.0)


def _ghz_qasm(n):
    return f"{n}\nH 0\n" + "".join(f"CNOT {i} {i + 1}\n" for i in range(n - 1)) + "".join(
        f"Rz 0.{i + 1} {i}\n" for i in range(n))


def test_preprocess_finds_a_sequence_in_time():
    """preProcess (src/preprocess.h:28-52): the first seeded stochastic contraction
    that finishes within the threshold gives its mCreatedFrom record, which replays
    in the oracle to the state-vector value; with no time at all it raises
    ContractionFailure; maxTime keeps the threshold afterwards, as in the reference."""
    q = _ghz_qasm(10)
    meas = "Z Z 0 1 T T T T T Z"
    try:
        found, seq = qa.preprocess(q, meas, 30.0, seed=11)
        assert found and len(seq) > 0
        assert qa.get_max_time() == 30.0
        assert qa.timer_elapsed() == 0.0  # reset after success
        onet = qo.Network(q, meas)
        val = onet.contract_given_sequence(seq)
        psi, n = sv.simulate(q)
        ref = sv.expectation(psi, n, meas)
        assert abs(val - ref) < 1e-12 * max(1.0, abs(ref))
        # the same seed gives the same record
        assert qa.preprocess(q, meas, 30.0, seed=11)[1] == seq
        with pytest.raises(qa.ContractionFailure):
            qa.preprocess(q, meas, 0.0, seed=11)
    finally:
        qa.timer_reset()
        qa.set_max_time(60.0)


def test_contract_nodes_from_two_threads():
    """ContractNodes guards its bookkeeping like the reference's mLocker
    (src/Network.h:718-862), so two threads can contract disjoint pairs of one network
    (src/ContractionTools.h:359-368): two independent registers are contracted
    concurrently, one thread each, following each register's own recorded order; the
    node ids stay unique and dense, the counter is the sum of both, and the merged
    record replays to the same counter."""
    import threading
    n, layers = 24, 3

    def reg_gates(q0):
        out = []
        for l in range(layers):
            out += [f"H {q0 + i}" for i in range(n)]
            out += [f"CNOT {q0 + i} {q0 + i + 1}" for i in range(l % 2, n - 1, 2)]
            out += [f"Rz 0.{l + 3} {q0 + i}" for i in range(n)]
        return out

    ga, gb = reg_gates(0), reg_gates(n)
    assert len(ga) == len(gb)
    g = len(ga)
    # each register alone: its recorded order (local ids)
    solo_q = f"{n}\n" + "\n".join(reg_gates(0)) + "\n"
    solo = qa.Network(solo_q, None, None)
    solo.contract_stochastic(5)
    local_seq, solo_flops = solo.record(), solo.flops()
    both_q = f"{2 * n}\n" + "\n".join(ga + gb) + "\n"

    def to_global(reg, lid, created):
        if lid < n:
            return reg * n + lid                      # initial states
        if lid < n + g:
            return 2 * n + reg * g + (lid - n)        # gates
        if lid < n + g + n:
            return 2 * n + 2 * g + reg * n + (lid - n - g)  # trace nodes
        return created[lid]

    for rep in range(5):
        net = qa.Network(both_q, None, None)
        finals = [None, None]
        errors = []

        def run(reg):
            try:
                created = {}
                nxt = n + g + n
                for a, b in local_seq:
                    rid = net.contract_nodes(to_global(reg, a, created), to_global(reg, b, created), 100)
                    assert rid >= 0
                    created[nxt] = rid
                    nxt += 1
                finals[reg] = rid
            except Exception as exc:  # reported below
                errors.append(exc)

        ts = [threading.Thread(target=run, args=(r,)) for r in (0, 1)]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        assert not errors, errors
        assert net.contract_nodes(finals[0], finals[1], 100) == -1  # scalar x scalar ends the network
        c = net.counts()
        assert c["done"] and c["all"] == c["original"] + 2 * len(local_seq) + 1
        assert net.flops() == 2 * solo_flops + 1
        rec = net.record()
        assert len(rec) == 2 * len(local_seq) + 1
        replay = qa.Network(both_q, None, None)
        replay.contract_given(rec)
        assert replay.flops() == net.flops() and replay.counts()["done"]


def test_sliced_planned_bytes_cover_step_scratch():
    """A slice's planned device bytes cover, at every GEMM-shaped step, the live
    operands and result plus the permuted copies of both operands that the GEMM may
    make (and the GEMMs' split-K partial products or the split reduction's partial
    sums: 64 MiB at most), not only the live tensors."""
    from qtorch_amd import workloads as wl
    q = wl.grid_random_circuit(5, 6, 10, seed=4)
    bits = wl.bitstring(30, 2)
    plan = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=40).plan()
    net = qa.Network(q, bits, None, dim_log2=1)
    net.order(qa.ORDER_LINEGRAPH, 0)
    steps = net.plan()
    need = 0
    for _, _, _, ra, rb, k, rc, _ in steps:
        if (ra - k) >= 6 and (rb - k) >= 6 and k >= 3:  # dimension-2 wires: one bit per axis
            need = max(need, 16 * (2 * (2 ** ra) + 2 * (2 ** rb) + 2 ** rc))
    assert plan["gemm_cmacs_per_slice"] > 0 and need > 0
    assert plan["peak_bytes"] >= need + (64 << 20), (plan["peak_bytes"], need)


def test_nonzero_slices_report_a_truncated_search():
    """qth_sliced_nonzero_slices_ex: a budget too small to finish says so (complete =
    0); a sufficient one lists every nonzero slice and says it is complete; the plain
    entry point returns an error instead of a silently short list."""
    from qtorch_amd import workloads as wl
    q = wl.grid_random_circuit(4, 4, 8, seed=11)
    sn = qa.SlicedNetwork(q, wl.bitstring(16, 5), dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=8)
    full, done = sn.nonzero_slices_ex(1 << 20, budget=1 << 24)
    assert done and 0 < len(full) <= sn.plan()["slices"]
    part, done = sn.nonzero_slices_ex(1 << 20, budget=3)
    assert not done and part == full[:len(part)]
    capped, done = sn.nonzero_slices_ex(1, budget=1 << 24)
    assert done and capped == full[:1]
