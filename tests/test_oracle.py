"""Pins the oracle (oracle/qtoracle.py over oracle/contract_ref.c) before anything
is compared against it.

* every golden case recorded from the reference compiled from its own sources is
  replayed in the reference's own order and must agree BIT-FOR-BIT (same summation
  order, no FMA), and the reference's complex-MAC counter must agree exactly;
* the reference's known answers (src/tests.cpp) hold at the reference's tolerances;
* an independent state-vector simulation (tests/statevector.py) and the published
  QAOA p=1 closed form agree to 1e-12;
* the pairwise arithmetic equals numpy.einsum on random tensors (index layout,
  operand order and digit significance), and a planted transposition fails it.
"""
import hashlib
import math
import os

import numpy as np
import pytest

from oracle import qtoracle as qo
from tests import statevector as sv

GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


def replay(case):
    qpath = os.path.join(GOLDEN_DIR, case["qasm"])
    with open(qpath) as f:
        qasm = f.read()
    net = qo.Network(qasm, case["measure"], base_dir=GOLDEN_DIR)
    mode = case["mode"]
    if mode == "stoch":
        val = net.contract_given_sequence(case["seq"])
    elif mode == "rstoch":
        net.reduce_circuit()
        val = net.contract_given_sequence(case["seq"])
    elif mode in ("lg", "rlg"):
        if mode == "rlg":
            net.reduce_circuit()
        wires, edges = net.line_graph_wires()
        assert qo.lg_cnf_text(wires, edges) == case["cnf"]
        val = net.lg_contract(wires, case["qbb_order"])
    elif mode == "user":
        val = net.contract_user_sequence(case["seq"])
    else:
        raise ValueError(mode)
    return val, net


def _small(case):
    return case["flops"] < 5_000_000


def test_golden_replay_bit_exact(golden):
    n = 0
    for case in golden:
        if not _small(case):
            continue
        val, net = replay(case)
        assert (val.real, val.imag) == tuple(case["value"]), case["name"]
        assert net.flops == case["flops"], case["name"]
        n += 1
    assert n > 250


@pytest.mark.slow
def test_golden_replay_large(golden):
    for case in golden:
        if _small(case):
            continue
        val, net = replay(case)
        assert (val.real, val.imag) == tuple(case["value"]), case["name"]
        assert net.flops == case["flops"], case["name"]


def test_known_answers(golden):
    """Values the reference's own tests assert (src/tests.cpp:254-2004)."""
    for case in golden:
        if "known" in case:
            assert abs(case["value"][0] - case["known"]) <= case.get("tol", 1e-12), case["name"]
            assert abs(case["value"][1]) <= 1e-5


def test_statevector_agrees(golden):
    """Whole circuits against an independent simulation (gate tables, wire order,
    measurement operators, angle parsing)."""
    checked = 0
    for case in golden:
        if case["mode"] not in ("stoch", "lg") or case["flops"] > 5_000_000:
            continue
        with open(os.path.join(GOLDEN_DIR, case["qasm"])) as f:
            qasm = f.read()
        nq = int(qasm.split("\n")[0].strip())
        if nq > 14:
            continue
        psi, n = sv.simulate(qasm, base_dir=GOLDEN_DIR)
        ref = sv.expectation(psi, n, case["measure"])
        got = complex(*case["value"])
        assert abs(got - ref) <= 1e-12 * max(1.0, abs(ref)), (case["name"], got, ref)
        checked += 1
    assert checked > 200


def _qaoa_p1_closed_form(du, dv, lam, gamma_w, beta):
    """<C_uv> for QAOA p=1 MaxCut (Wang, Hadfield, Jiang, Rieffel, PRA 97, 022304 (2018),
    Thm 1): d = other-neighbour counts, lam = triangles on the edge."""
    g, b = gamma_w, beta
    t1 = 0.25 * math.sin(4 * b) * math.sin(g) * (math.cos(g) ** du + math.cos(g) ** dv)
    t2 = 0.25 * math.sin(2 * b) ** 2 * math.cos(g) ** (du + dv - 2 * lam) * (1 - math.cos(2 * g) ** lam)
    return 0.5 + t1 - t2


def test_qaoa_p1_closed_form(golden):
    """Per-edge <Z_u Z_v> of the correctly mapped light cones against the closed form."""
    with open(os.path.join(GOLDEN_DIR, "graphs", "3reg_n30_s2024.dgf")) as f:
        edges = qo.read_dgf(f.read())
    adj = {}
    for u, v in edges:
        adj.setdefault(u, set()).add(v)
        adj.setdefault(v, set()).add(u)
    n = 0
    for case in golden:
        if not case["name"].startswith("qaoa3_fix_"):
            continue
        u, v = edges[case["edge"]]
        lam = len(adj[u] & adj[v])
        g_text = qo.fmt_g(-case["betas_gammas"][1])
        b_text = qo.fmt_g(case["betas_gammas"][0] * 2.0)
        gamma_w = -qo.stof(g_text)     # CNOT Rz(t) CNOT = exp(-i t/2 ZZ); C_uv = (1 - ZZ)/2
        beta = qo.stof(b_text) / 2.0   # Rx(2 beta) = exp(-i beta X)
        c_uv = _qaoa_p1_closed_form(len(adj[u]) - 1, len(adj[v]) - 1, lam, gamma_w, beta)
        zz = case["value"][0]
        assert abs((1 - zz) / 2 - c_uv) < 1e-12, (case["name"], (1 - zz) / 2, c_uv)
        n += 1
    assert n == len(edges)


def _einsum_reference(a, ra, b, rb, free, pairs, dlog2):
    """C via numpy.einsum over the documented layout: axis i of a rank-r tensor is
    flat digit i (little-endian), so the C-order reshape lists axes in reverse."""
    d = 1 << dlog2
    A = a.reshape([d] * ra)[tuple([slice(None)] * ra)]
    B = b.reshape([d] * rb)
    letters = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ"
    la = [None] * ra
    lb = [None] * rb
    it = iter(letters)
    for pa, pb in pairs:
        l = next(it)
        la[pa] = l
        lb[pb] = l
    out = []
    for from_a, ax in free:
        l = next(it)
        (la if from_a else lb)[ax] = l
        out.append(l)
    # numpy axis j of the reshaped array is digit r-1-j
    sa = "".join(reversed(la))
    sb = "".join(reversed(lb))
    so = "".join(reversed(out))
    c = np.einsum(f"{sa},{sb}->{so}", A, B)
    return c.reshape(-1)


@pytest.mark.parametrize("dlog2", [1, 2])
def test_contract_indices_vs_einsum(dlog2):
    rng = np.random.default_rng(7 + dlog2)
    for trial in range(60):
        ra = int(rng.integers(0, 5 if dlog2 == 2 else 8))
        rb = int(rng.integers(0, 5 if dlog2 == 2 else 8))
        k = int(rng.integers(0, min(ra, rb) + 1))
        pa = list(rng.permutation(ra)[:k])
        pb = list(rng.permutation(rb)[:k])
        free = [(True, i) for i in range(ra) if i not in pa] + [(False, j) for j in range(rb) if j not in pb]
        order = rng.permutation(len(free))
        free = [free[i] for i in order]
        d = 1 << dlog2
        a = rng.standard_normal(d ** ra) + 1j * rng.standard_normal(d ** ra)
        b = rng.standard_normal(d ** rb) + 1j * rng.standard_normal(d ** rb)
        pairs = list(zip(pa, pb))
        got = qo.contract_indices(a, ra, b, rb, free, pairs, dlog2)
        ref = _einsum_reference(a, ra, b, rb, free, pairs, dlog2)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def test_einsum_check_catches_planted_transpose():
    """A planted mistake (free axes of C swapped) must fail the einsum check."""
    rng = np.random.default_rng(3)
    a = rng.standard_normal(16) + 1j * rng.standard_normal(16)
    b = rng.standard_normal(16) + 1j * rng.standard_normal(16)
    good = [(True, 0), (False, 1)]
    bad = [(False, 1), (True, 0)]
    ref = _einsum_reference(a, 2, b, 2, good, [(1, 0)], 2)
    wrong = qo.contract_indices(a, 2, b, 2, bad, [(1, 0)], 2)
    assert not np.allclose(wrong, ref)


def test_parse_tokens_like_std_regex():
    """Single-space split, empty final piece dropped, '#' only at column 0
    (src/Network.h:976-992; behaviour of std::sregex_token_iterator)."""
    assert qo.parse_tokens("H 0") == ["H", "0"]
    assert qo.parse_tokens("H 0 ") == ["H", "0"]
    assert qo.parse_tokens(" H 0") == ["", "H", "0"]
    assert qo.parse_tokens("H  0") == ["H", "", "0"]
    assert qo.parse_tokens("# comment") == []
    assert qo.parse_tokens("CNOT 0 1 # c") == ["CNOT", "0", "1", "#", "c"]


def test_unknown_gate_raises():
    with pytest.raises(qo.InvalidFileFormat):
        qo.Network("1\nFOO 0\n", "Z")
    with pytest.raises(qo.InvalidFileFormat):
        qo.Network("2\nCNOT 0 0\n", "Z Z")
    with pytest.raises(qo.InvalidFileFormat):
        qo.Network("1\nH 3\n", "Z")


def test_qft8_closed_form(golden):
    """QFT on |0^8> gives probability 1/256 for every output string."""
    for case in golden:
        if case["name"].startswith("qft8_"):
            assert abs(case["value"][0] - 1 / 256) < 1e-14


def _sha(text):
    return hashlib.sha256(text.encode()).hexdigest()[:32]


def test_light_cones_match_reference_extradata(ref_cones):
    """The oracle's light cones in the reference's qubit map equal ExtraData's
    qubitsNeeded and realIterations (src/maxcut.h:105-186) for every edge of three
    graphs at p = 1, 2 (and 3 on one graph), and its per-edge QASM text is
    byte-identical to what outputInitialPlusStateToFile + applyU_CsThenU_Bs wrote
    (src/maxcut.h:191-233) for angle sets that exercise 6-digit printing."""
    n = 0
    for g in ref_cones:
        with open(os.path.join(GOLDEN_DIR, g["dgf"])) as f:
            edges = qo.read_dgf(f.read())
        cones = qo.light_cones(edges, g["p"], reference_qubit_map=True)
        assert [c[0] for c in cones] == g["qubits"], g["dgf"]
        assert [[list(t) for t in c[1]] for c in cones] == g["iterations"], g["dgf"]
        for s in g["sets"]:
            for e, (nq, real, _) in enumerate(cones):
                text = qo.qaoa_edge_qasm(nq, real, g["p"], s["betas_gammas"])
                if str(e) in s["text"]:
                    assert text == s["text"][str(e)], (g["dgf"], g["p"], e)
                assert _sha(text) == s["sha256"][e], (g["dgf"], g["p"], e, s["betas_gammas"])
                n += 1
    assert n > 1000


def test_qaoa_fixture_circuits_are_reference_emission(ref_cones, golden):
    """The qaoa3_ref_* and qaoa_sample_p2_* circuits the golden cases contract are the
    reference's own per-edge texts."""
    by_key = {(g["dgf"], g["p"]): g for g in ref_cones}
    g1 = by_key[("graphs/3reg_n30_s2024.dgf", 1)]
    g2 = by_key[("samples/3regRand30Node50.dgf", 2)]
    n = 0
    for case in golden:
        if case["name"].startswith("qaoa3_ref_"):
            s = [s for s in g1["sets"] if s["betas_gammas"] == case["betas_gammas"]][0]
            want = s["sha256"][case["edge"]]
        elif case["name"].startswith("qaoa_sample_p2_"):
            want = g2["sets"][0]["sha256"][case["edge"]]
        else:
            continue
        with open(os.path.join(GOLDEN_DIR, case["qasm"])) as f:
            assert _sha(f.read()) == want, case["name"]
        n += 1
    assert n == 50


def test_light_cone_check_catches_planted_mistake(ref_cones):
    """A plausible restatement error (the reference map with pair.second also fixed to
    1, i.e. the corrected map) fails the ExtraData comparison on these graphs."""
    g = [g for g in ref_cones if g["dgf"] == "samples/3regRand30Node50.dgf" and g["p"] == 1][0]
    with open(os.path.join(GOLDEN_DIR, g["dgf"])) as f:
        edges = qo.read_dgf(f.read())
    fixed = qo.light_cones(edges, 1, reference_qubit_map=False)
    assert [[list(t) for t in c[1]] for c in fixed] != g["iterations"]
