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
