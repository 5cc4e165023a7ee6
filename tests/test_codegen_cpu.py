"""The specialised executor's generated code, run on the CPU: the same source that
qt_batch_compile hands to the device compiler is built with g++ (a few macros stand in
for the HIP keywords; explicit FMAs keep their single rounding) and evaluated for the
golden angle set of the 30-vertex 3-regular graph; every edge's <ZZ> must match the
values the reference printed (complex128 to 1e-12, complex64 4M to 1e-5 with the 1e-3
floor).  This pins the generator (operand roles, zero dropping, liveness, the final
scalar) independently of the GPU."""
import ctypes
import json
import os
import subprocess

import numpy as np
import pytest

import qtorch_amd as qa

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PRELUDE = r"""
#define __device__
#define __global__
#define __launch_bounds__(x)
struct qt_index3 { unsigned x, y, z; };
static qt_index3 blockIdx, threadIdx, gridDim;
#define __builtin_amdgcn_readfirstlane(x) (x)
#define __shared__ static
#define __syncthreads() ((void)0)
"""


def _golden():
    with open(os.path.join(GOLDEN, "cases.json")) as f:
        cases = json.load(f)["cases"]
    ref = sorted((c for c in cases if c["name"].startswith("qaoa3_ref_")), key=lambda c: c["edge"])
    return np.array([c["value"][0] for c in ref]), ref[0]["betas_gammas"]


def _build(src, n, real, tmp_path, name):
    calls = "".join(f"    case {i}: prog{i}(P, out); break;\n" for i in range(n))
    text = (PRELUDE + src + f'\nextern "C" void run(int prog, const {real}* P, double* out) {{\n'
            f"  switch (prog) {{\n{calls}    default: break;\n  }}\n}}\n")
    cpp = tmp_path / f"{name}.cpp"
    so = tmp_path / f"{name}.so"
    cpp.write_text(text)
    subprocess.run(["g++", "-O1", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", "-w", str(cpp), "-o",
                    str(so)], check=True, timeout=300)
    return ctypes.CDLL(str(so))


@pytest.mark.parametrize("hermitian", [False, True])
@pytest.mark.parametrize("precision", [qa.QT_PLAIN_C128, qa.QT_C64_4M])
def test_generated_code_matches_reference(tmp_path, precision, hermitian):
    """With hermitian, only one element of each conjugate pair of every intermediate is
    computed (the other is its conjugate) and self-mirror elements are real."""
    dgf = open(os.path.join(GOLDEN, "graphs", "3reg_n30_s2024.dgf")).read()
    zz_ref, bg = _golden()
    first, n = 0, 6
    src = qa.qaoa_source(dgf, 1, first, n, precision, hermitian=hermitian)
    assert "qt_batch_specialised" in src and src.count("prog") >= n
    c128 = precision == qa.QT_PLAIN_C128
    real = "double" if c128 else "float"
    lib = _build(src, n, real, tmp_path, ("c128" if c128 else "c64") + ("h" if hermitian else ""))
    block = qa.qaoa_param_block(1, bg)
    P = np.ascontiguousarray(block.view(np.float64) if c128 else block.view(np.float64).astype(np.float32))
    out = np.zeros(2)
    for e in range(first, first + n):
        lib.run(ctypes.c_int(e - first), P.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
        got, ref = out[0], zz_ref[e]
        if c128:
            assert abs(got - ref) <= 1e-12 * max(abs(ref), 1e-3), (e, got, ref)
        else:
            assert abs(got - ref) <= 1e-5 * max(abs(ref), 1e-3), (e, got, ref)
        assert abs(out[1]) <= (1e-12 if c128 else 1e-6)


def test_generated_code_drops_zero_terms():
    """Constant gates in the Liouville picture are sparse: the generated code runs far
    fewer FMAs than 4 per complex MAC of the plan's device steps; conjugate symmetry
    roughly halves them again."""
    dgf = open(os.path.join(GOLDEN, "graphs", "3reg_n30_s2024.dgf")).read()
    fmas = qa.qaoa_source(dgf, 1, 0, 1, qa.QT_C64_4M).count("__builtin_fmaf(")
    herm = qa.qaoa_source(dgf, 1, 0, 1, qa.QT_C64_4M, hermitian=True).count("__builtin_fmaf(")
    assert 0 < fmas < 4 * 2473
    assert 0 < herm < 0.6 * fmas


def test_generated_code_compiles_without_packed_fp32(tmp_path):
    """The product's own compile path (hipRTC, the options qt_batch_compile uses) turns
    generated code into gfx950 code with scalar FP32 FMAs only: no packed FP32
    instructions, and as many FMA instructions as the generator counts (within the
    compiler's own exact simplifications of multiply-adds by +-1)."""
    import re
    import shutil
    objdump = shutil.which("llvm-objdump") or "/opt/rocm/llvm/bin/llvm-objdump"
    if not os.path.exists(objdump):
        pytest.skip("llvm-objdump not available")
    dgf = open(os.path.join(GOLDEN, "graphs", "3reg_n30_s2024.dgf")).read()
    src = qa.qaoa_source(dgf, 1, 0, 2, qa.QT_C64_4M)
    co = tmp_path / "k.co"
    co.write_bytes(qa.program_compile(src, "gfx950"))
    asm = subprocess.run([objdump, "-d", str(co)], capture_output=True, text=True, check=True, timeout=120).stdout
    assert "qt_batch_specialised" in asm
    assert not re.search(r"\bv_pk_(fma|mul|add)_f32\b", asm)
    n_fma = len(re.findall(r"\bv_fma(c|mk|ak)?_f32", asm))
    generated = src.count("__builtin_fmaf(")
    assert 0.9 * generated <= n_fma <= generated
