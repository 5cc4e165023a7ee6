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
