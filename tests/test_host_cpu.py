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


def test_qaoa_edge_text_matches_reference_emission(golden):
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
