"""GPU tests of the drivers: the qtorch CLI (src/main.cpp counterpart) on .inp scripts,
the QAOA outer loop and the greedy final cut (src/maxcut.cpp)."""
import os
import re
import subprocess

import numpy as np
import pytest

import qtorch_amd as qa

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN_DIR = os.path.join(REPO, "tests", "golden")
CLI = os.path.join(REPO, "qtorch_amd", "bin", "qtorch")


@pytest.fixture(scope="module")
def ctx():
    c = qa.Context(0)
    yield c
    c.close()


def _run_cli(tmp_path, method, extra=""):
    meas = tmp_path / "m.txt"
    meas.write_text("Z T T T")
    out = tmp_path / "o.txt"
    inp = tmp_path / "s.inp"
    inp.write_text(f">string qasm {GOLDEN_DIR}/samples/test_JW.qasm\n>string measurement {meas}\n"
                   f">string contractmethod {method}\n>string outputpath {out}\n>int seed 3\n{extra}")
    r = subprocess.run([CLI, str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    text = out.read_text()
    m = re.search(r"Result of Contraction: \(([-0-9.e+]+),([-0-9.e+]+)\)", text)
    assert m, text
    assert "Number of floating point ops in full contraction" in text
    return complex(float(m.group(1)), float(m.group(2)))


@pytest.mark.parametrize("method", ["simple-stoch", "linegraph-qbb"])
def test_cli_known_answer(tmp_path, method):
    """test_JW.qasm with Z T T T: -0.620705 (src/tests.cpp:266), after ReduceCircuit."""
    v = _run_cli(tmp_path, method)
    assert abs(v.real + 0.62070542274223) < 1e-12 and abs(v.imag) < 1e-12


def _graph():
    with open(os.path.join(GOLDEN_DIR, "graphs", "3reg_n30_s2024.dgf")) as f:
        return f.read()


def test_qaoa_optimize_and_final_cut(ctx):
    dgf = _graph()
    start = [0.392699, 0.785399]  # the reference's starting point (src/maxcut.cpp:155-157)
    e_start, _ = qa.qaoa_energy_plain(ctx, dgf, 1, True, start, seed=0)
    ang, e_best, n = qa.qaoa_optimize(ctx, dgf, 1, qa.QT_C64_4M, n_random=128, max_evals=600, seed=1)
    assert n > 128 and e_best >= e_start - 1e-6
    e_check, _ = qa.qaoa_energy_plain(ctx, dgf, 1, True, list(ang), seed=2)
    assert abs(e_check - e_best) <= 1e-5 * abs(e_check)
    bits, cut = qa.maxcut_final_string(ctx, dgf, 1, list(ang), seed=0)
    edges = [tuple(map(int, l.split()[1:3])) for l in dgf.splitlines() if l.startswith("e")]
    assert cut == sum(1 for a, b in edges if bits[a] != bits[b])
    assert cut >= len(edges) // 2


@pytest.mark.parametrize("n,p,seed", [(10, 1, 4), (8, 2, 6)])
def test_final_cut_against_state_vector(ctx, n, p, seed):
    """maxcutGetFinalString (src/maxcut.cpp:29-140) on small 3-regular graphs: every
    conditional probability P(bit i = 0 | the bits before) that the device computes
    from one full-circuit network per qubit, and the bitstring they decide, against a
    state vector of the same circuit text (H on every qubit; p x [CNOT, Rz(-gamma),
    CNOT per edge; Rx(2 beta) per qubit], angles printed at 6 significant digits)."""
    from qtorch_amd import workloads as wl
    from tests import statevector as sv
    edges = wl.regular_graph(n, 3, seed)
    dgf = wl.dgf_text(edges)
    bg = wl.qaoa_angles(p, seed)
    bits, cut, probs = qa.maxcut_final_string(ctx, dgf, p, bg, seed=0, with_probs=True)
    lines = [str(n)] + [f"H {i}" for i in range(n)]
    for i in range(p):
        for a, b in edges:
            lines += [f"CNOT {a} {b}", f"Rz {-bg[i + p]:g} {b}", f"CNOT {a} {b}"]
        lines += [f"Rx {bg[i] * 2.0:g} {j}" for j in range(n)]
    psi, nq = sv.simulate("\n".join(lines) + "\n")
    prob = np.abs(psi) ** 2
    idx = np.arange(1 << n)
    chosen, current = [], 1.0
    for i in range(n):
        mask = np.ones(1 << n, dtype=bool)
        for j, b in enumerate(chosen):
            mask &= ((idx >> j) & 1) == b
        mask &= ((idx >> i) & 1) == 0
        cond = prob[mask].sum() / current
        assert abs(probs[i] - cond) <= 1e-10, (i, probs[i], cond)
        if abs(cond - 0.5) <= 1e-9:
            # a tie up to rounding (the first bit always is: the QAOA state is symmetric
            # under flipping every bit): the device's choice stands
            b = bits[i]
        else:
            b = 0 if cond > 0.5 else 1
        chosen.append(b)
        current *= cond if b == 0 else 1.0 - cond
    assert bits == chosen
    assert cut == sum(1 for a, b in edges if bits[a] != bits[b])
