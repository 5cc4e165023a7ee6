"""GPU tests of sliced execution of one large network (qth_sliced_*): every form
against a state-vector simulation, sliced against unsliced, and a Liouville-picture
network with a closed-form value."""
import numpy as np
import pytest

import qtorch_amd as qa
from qtorch_amd import workloads as wl
from tests import statevector as sv

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = qa.Context(0)
    yield c
    c.close()


def _amp_index(bits):
    return sum(int(b) << q for q, b in enumerate(bits))


@pytest.mark.parametrize("prec,tol", [(qa.QT_PLAIN_C128, 1e-10), (qa.QT_C64_4M, 2e-5), (qa.QT_C64_3M, 4e-5),
                                      (qa.QT_C64_BF16X3, 2e-3), (qa.QT_C64_4M | qa.QT_RESCALE, 2e-5)])
def test_sliced_amplitude_matches_statevector(ctx, prec, tol):
    q = wl.grid_random_circuit(4, 4, 8, seed=11)
    bits = wl.bitstring(16, 5)
    sn = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=8)
    plan = sn.plan()
    assert plan["sliced_wires"] > 0 and plan["peak_bits"] <= 8 and plan["slices"] == 2 ** plan["sliced_wires"]
    v, ms = sn.run(ctx, prec)
    psi, n = sv.simulate(q)
    ref = psi[_amp_index(bits)]
    assert abs(v - ref) <= tol * max(abs(ref), 2.0 ** (-n / 2)), (v, ref)


def test_sliced_equals_unsliced(ctx):
    q = wl.grid_random_circuit(4, 5, 8, seed=2)
    bits = wl.bitstring(20, 8)
    whole = qa.SlicedNetwork(q, bits, dim_log2=1, max_bits=40)
    assert whole.plan()["sliced_wires"] == 0
    cut = qa.SlicedNetwork(q, bits, dim_log2=1, max_bits=whole.plan()["peak_bits"] - 3)
    assert cut.plan()["sliced_wires"] >= 1
    v0, _ = whole.run(ctx)
    v1, _ = cut.run(ctx)
    assert abs(v1 - v0) <= 1e-12 * max(abs(v0), 2.0 ** -10)


def test_sliced_liouville_qft(ctx):
    """qTorch's own picture (dimension-4 wires): P(0101) of QFT4 on |0000> = 1/16."""
    sn = qa.SlicedNetwork(wl.qft_qasm(4), "0 1 0 1", dim_log2=2, order=qa.ORDER_GREEDY, max_bits=6)
    assert sn.plan()["sliced_wires"] > 0
    v, _ = sn.run(ctx)
    assert abs(v - 1 / 16) <= 1e-13


@pytest.mark.parametrize("prec,tol", [(qa.QT_PLAIN_C128, 1e-12), (qa.QT_C64_4M, 1e-5),
                                      (qa.QT_C64_4M | qa.QT_RESCALE, 1e-5)])
def test_fused_subtrees_match_step_by_step(ctx, prec, tol):
    """Small subtrees run as device programs (one batch launch per chunk of slices,
    results written to tensors that the remaining steps read): the same value as
    running every step through qt_contract_pair, with all slices in one launch and
    with three slices per launch, and the state vector's amplitude."""
    q = wl.grid_random_circuit(4, 5, 8, seed=4)
    bits = wl.bitstring(20, 6)
    sn = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=14)
    plan = sn.plan()
    assert plan["fused_subtrees"] >= 2 and plan["fused_steps"] > 0 and plan["slices"] >= 4
    sn.set_fusion(0)
    v_step, _ = sn.run(ctx, prec)
    sn.set_fusion(-1)
    v_all, _ = sn.run(ctx, prec)
    sn.set_fusion(3)
    v_three, _ = sn.run(ctx, prec)
    psi, n = sv.simulate(q)
    ref = psi[_amp_index(bits)]
    scale = max(abs(ref), 2.0 ** (-n / 2))
    for v in (v_step, v_all, v_three):
        assert abs(v - ref) <= max(tol, 1e-10) * scale, (v, ref)
    assert abs(v_all - v_step) <= tol * scale
    assert v_three == v_all  # per-slice values do not depend on the chunking


def test_fused_subtrees_liouville(ctx):
    """Dimension-4 wires: a measured QFT5 network, fused and step by step, equals the
    closed form |<0..0|QFT|x>|^2 = 1/32."""
    sn = qa.SlicedNetwork(wl.qft_qasm(5), "1 0 1 1 0", dim_log2=2, order=qa.ORDER_GREEDY, max_bits=8)
    assert sn.plan()["fused_subtrees"] >= 1
    v_fused, _ = sn.run(ctx)
    sn.set_fusion(0)
    v_step, _ = sn.run(ctx)
    assert abs(v_fused - 1 / 32) <= 1e-13 and abs(v_step - 1 / 32) <= 1e-13


@pytest.mark.parametrize("fusion", [-1, 0])
def test_slice_ranges_sum_to_the_whole(ctx, fusion):
    """Slices [0, k) plus slices [k, end) sum to the value of every slice (the
    state-vector amplitude, pinned by the test above); a count past the end stops at
    the last slice; the full range is the default."""
    q = wl.grid_random_circuit(4, 4, 8, seed=11)
    bits = wl.bitstring(16, 5)
    sn = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_LINEGRAPH, max_bits=8)
    sn.set_fusion(fusion)
    n = sn.plan()["slices"]
    assert n >= 4
    whole, _ = sn.run(ctx)
    k = n // 3
    sn.set_range(0, k)
    a, _ = sn.run(ctx)
    sn.set_range(k, -1)
    b, _ = sn.run(ctx)
    sn.set_range(k, n)  # a count past the end stops at the last slice
    b2, _ = sn.run(ctx)
    assert b2 == b
    assert abs((a + b) - whole) <= 1e-12 * abs(whole), (a, b, whole)
    assert abs(a) > 0 and abs(b) > 0
    sn.set_range(0, -1)
    again, _ = sn.run(ctx)
    assert again == whole
    with pytest.raises(qa.InvalidFunctionInput):
        sn.set_range(n, 1)
    with pytest.raises(qa.InvalidFunctionInput):
        sn.set_range(0, 0)


def test_nonzero_slices_cover_the_amplitude(ctx):
    """Slices left out by nonzero_slices contract to exactly zero, and the listed
    slices sum to the amplitude of every slice."""
    q = wl.grid_random_circuit(4, 4, 10, seed=3)
    bits = wl.bitstring(16, 5)
    sn = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_GREEDY, max_bits=8)
    n = sn.plan()["slices"]
    assert n <= 512
    whole, _ = sn.run(ctx)
    listed = sn.nonzero_slices(n)
    assert 0 < len(listed) < n
    skipped = sorted(set(range(n)) - set(listed))
    sn.set_slices(skipped)
    zero, _ = sn.run(ctx)
    assert zero == 0
    for s in skipped[:8]:
        sn.set_range(s, 1)
        assert sn.run(ctx)[0] == 0
    sn.set_slices(listed)
    part, _ = sn.run(ctx)
    assert abs(part - whole) <= 1e-12 * abs(whole), (part, whole)


@pytest.mark.parametrize("rows,cols,depth,max_bits", [(4, 4, 8, 6), (4, 5, 10, 10), (4, 5, 10, 40), (3, 6, 12, 9)])
def test_sweep_order_with_split_cz_matches_statevector(ctx, rows, cols, depth, max_bits):
    """The circuit sweep order with CZ split into two bonded rank-3 halves, unsliced and
    sliced, in complex128 and 4M, against the state vector's amplitude."""
    q = wl.grid_random_circuit(rows, cols, depth, seed=7)
    n = rows * cols
    bits = wl.bitstring(n, 3)
    sn = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_SWEEP, seed=0, max_bits=max_bits)
    plan = sn.plan()
    assert plan["peak_bits"] <= max_bits
    psi, _ = sv.simulate(q)
    ref = psi[_amp_index(bits)]
    scale = max(abs(ref), 2.0 ** (-n / 2))
    for prec, tol in ((qa.QT_PLAIN_C128, 1e-12), (qa.QT_C64_4M, 2e-5)):
        v, _ = sn.run(ctx, prec)
        assert abs(v - ref) <= tol * scale, (prec, v, ref)
    sn.close()


def test_split_diagonal_gates_match_statevector(ctx):
    """CZ, CPHASE and CRk split into bonded halves (the sweep order) on a circuit that
    uses all three, against the state vector."""
    lines = ["6"] + [f"H {i}" for i in range(6)]
    lines += ["CZ 0 1", "CPHASE 0.7 1 2", "CRk 2 3", "Rx 0.4 3", "CZ 3 4", "CPHASE -1.3 5 0", "CRk 1 5",
              "Ry 1.1 2", "CZ 2 5", "Rz 0.9 4", "CPHASE 2.2 4 1"]
    q = "\n".join(lines) + "\n"
    psi, n = sv.simulate(q)
    for bits in ("010110", "111000", "000001"):
        sn = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_SWEEP, seed=1, max_bits=40)
        v, _ = sn.run(ctx, qa.QT_PLAIN_C128)
        ref = psi[_amp_index(bits)]
        assert abs(v - ref) <= 1e-13, (bits, v, ref)
        sn.close()


def test_grid_6x8_depth14_whole_amplitude(ctx):
    """The precision-stress configuration computed whole: 48-qubit 6x8 depth-14 grid
    circuit, one seeded output amplitude, circuit sweep order (29-bit peak, 2.5e11
    complex MACs, unsliced), complex128 against 4M and 4M with per-step rescaling
    (the sliced 28-bit plan, 64 slices, is compared with it in the benchmark)."""
    q = wl.grid_random_circuit(6, 8, 14, seed=2024)
    bits = wl.bitstring(48, 1)
    sn = qa.SlicedNetwork(q, bits, dim_log2=1, order=qa.ORDER_SWEEP, seed=0, max_bits=29)
    plan = sn.plan()
    assert plan["slices"] == 1 and plan["peak_bits"] == 29 and plan["cmacs_per_slice"] < 2 ** 38
    v128, _ = sn.run(ctx, qa.QT_PLAIN_C128)
    assert 1e-9 < abs(v128) < 1e-6  # |amp|^2 ~ 2^-48 for a random circuit
    for prec in (qa.QT_C64_4M, qa.QT_C64_4M | qa.QT_RESCALE):
        v, _ = sn.run(ctx, prec)
        assert abs(v - v128) <= 1e-5 * abs(v128), (prec, v, v128)
    sn.close()
