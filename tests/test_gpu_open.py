"""GPU tests of networks with open output wires (qth_network_create_open): the
contraction ends in a tensor over the open qubits instead of a scalar.  In qTorch's
Liouville picture that tensor is the reduced output density matrix, so one
contraction gives the whole probability vector that the reference gets from one
measured network per output string (src/Network.h:199-240); in the amplitude
picture it is the output state vector."""
import itertools

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


def _density_from_tensor(vals, axq, qubits):
    """rho[ket, bra] over `qubits` (bit j of ket/bra = qubits[j]) from the flat
    tensor: digit of axis i = 2 ket + bra of qubit axq[i]."""
    n = len(qubits)
    rho = np.zeros((1 << n, 1 << n), dtype=np.complex128)
    for ket in range(1 << n):
        for bra in range(1 << n):
            idx = 0
            for axis, q in enumerate(axq):
                j = qubits.index(q)
                idx += (2 * ((ket >> j) & 1) + ((bra >> j) & 1)) << (2 * axis)
            rho[ket, bra] = vals[idx]
    return rho


def _reduced_density(psi, n, qubits):
    """Tr over the other qubits of |psi><psi| (bit q of psi's index = qubit q)."""
    t = psi.reshape([2] * n)  # axis n-1-q is qubit q
    keep = [n - 1 - q for q in qubits][::-1]  # most significant first: qubits[-1] ... qubits[0]
    rest = [a for a in range(n) if a not in keep]
    m = np.transpose(t, keep + rest).reshape(1 << len(qubits), -1)
    return m @ m.conj().T


def test_qft8_probability_vector_in_one_contraction(ctx):
    """QFT8 on |0^8> with all eight outputs open: one contraction (greedy order) gives
    the 256-entry probability vector, every entry 1/256 (closed form), each equal to
    the measured network of that string (256 networks in one batch), and the whole
    output density matrix equals |psi><psi| of an independent state vector."""
    q = wl.qft_qasm(8)
    net = qa.Network(q, None, ctx, open_qubits=list(range(8)))
    net.order(qa.ORDER_GREEDY)
    vals, axq = net.result_tensor()
    assert sorted(axq) == list(range(8)) and len(vals) == 4 ** 8
    probs = net.probabilities()
    assert probs.shape == (256,)
    assert np.max(np.abs(probs - 1 / 256)) <= 1e-14
    strings = [format(i, "08b")[::-1] for i in range(256)]  # string position q = bit q of i
    nb = qa.NetBatch(ctx, [q] * 256, [wl.measure_text(b) for b in strings], precision=qa.QT_PLAIN_C128)
    nb.launch()
    measured = nb.results().real
    nb.close()
    assert np.max(np.abs(probs - measured)) <= 1e-14
    psi, n = sv.simulate(q)
    rho_ref = np.outer(psi, psi.conj())
    rho = _density_from_tensor(vals, axq, list(range(8)))
    assert np.max(np.abs(rho - rho_ref)) <= 1e-13, np.max(np.abs(rho - rho_ref))


def test_marginals_with_some_outputs_open(ctx):
    """A random circuit with qubits {1, 4, 5} open and the rest traced: the reduced
    density matrix against the state vector's partial trace; the marginal
    probabilities against measured networks ('0' / '1' on the open qubits, 'T'
    elsewhere)."""
    q = wl.grid_random_circuit(2, 4, 5, seed=3)
    open_q = [1, 4, 5]
    net = qa.Network(q, None, ctx, open_qubits=open_q)
    net.order(qa.ORDER_GREEDY)
    vals, axq = net.result_tensor()
    rho = _density_from_tensor(vals, axq, open_q)
    psi, n = sv.simulate(q)
    rho_ref = _reduced_density(psi, n, open_q)
    assert np.max(np.abs(rho - rho_ref)) <= 1e-13
    probs = net.probabilities()
    for x in range(8):
        meas = ["T"] * n
        for j, qq in enumerate(open_q):
            meas[qq] = str((x >> j) & 1)
        m = qa.Network(q, " ".join(meas), ctx)
        v = m.contract_stochastic(x)
        assert abs(probs[x] - v.real) <= 1e-13, (x, probs[x], v)


def test_amplitude_picture_output_state(ctx):
    """Amplitude picture, every output open: the output state vector of a 4x4 grid
    circuit (split CZ halves are not used here) against the state vector."""
    q = wl.grid_random_circuit(4, 4, 6, seed=9)
    net = qa.Network(q, None, ctx, dim_log2=1, open_qubits=list(range(16)))
    net.order(qa.ORDER_GREEDY)
    vals, axq = net.result_tensor()
    psi, n = sv.simulate(q)
    got = np.zeros(1 << 16, dtype=np.complex128)
    x = np.arange(1 << 16)
    idx = np.zeros(1 << 16, dtype=np.int64)
    for axis, qq in enumerate(axq):
        idx += ((x >> qq) & 1) << axis
    got[:] = vals[idx]
    assert np.max(np.abs(got - psi)) <= 1e-13
    assert np.allclose(net.probabilities(), np.abs(psi) ** 2, atol=1e-14)
