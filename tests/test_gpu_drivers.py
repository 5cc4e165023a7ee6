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
