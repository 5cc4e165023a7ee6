"""GPU tests of qt_contract_pair_noalloc, the entry point that allocates nothing: its
scratch is a fixed 64 MiB device array per GPU.  Through the C ABI, against
qt_contract_pair on the same tensors (the same kernels: identical bits), two calls
in a row with the second smaller, in the GEMM regime (operand permutes in the
array) and the split reduction (partial sums in the array); steps whose scratch
exceeds the array, and QT_RESCALE, are refused before anything runs."""
import ctypes

import numpy as np
import pytest

import qtorch_amd as qa
from oracle import qtoracle as qo

pytestmark = pytest.mark.gpu
hip = qa._native.hip


@pytest.fixture(scope="module")
def ctx():
    c = qa.Context(0)
    yield c
    c.close()


def _rand(rng, n, dtype):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)


def _case(rng, ra, rb, k):
    pa = list(rng.permutation(ra)[:k])
    pb = list(rng.permutation(rb)[:k])
    free = [(True, i) for i in range(ra) if i not in pa] + [(False, j) for j in range(rb) if j not in pb]
    return list(zip(pa, pb)), free


def _noalloc(ctx, a, b, pairs, free, precision):
    c = qa.Tensor(ctx, len(free), a.dim_log2, a.dtype)
    pa = qa._native._i32([p[0] for p in pairs])
    pb = qa._native._i32([p[1] for p in pairs])
    fs = qa._native._i32([(ax | qa.QT_FROM_A) if fa else ax for fa, ax in free])
    macs = ctypes.c_int64(0)
    ctx.check(hip.qt_contract_pair_noalloc(ctx.handle, a.handle, a.rank, b.handle, b.rank, len(pairs), pa, pb,
                                           len(free), fs, c.handle, precision, ctypes.byref(macs)),
              "qt_contract_pair_noalloc")
    return c, macs.value


# (rank_a, rank_b, k, dim_log2, dtype): two calls in a row per entry, the second smaller
SEQUENCES = [
    [(9, 9, 4, 2, qa.QT_DTYPE_C128), (6, 6, 3, 2, qa.QT_DTYPE_C128)],   # GEMM: 4 MiB operand permutes, then 64 KB
    [(8, 8, 7, 2, qa.QT_DTYPE_C128), (7, 7, 7, 2, qa.QT_DTYPE_C128)],   # split reduction partials
    [(20, 20, 8, 1, qa.QT_DTYPE_C64), (12, 12, 5, 1, qa.QT_DTYPE_C64)],  # complex64 4M GEMM
    [(6, 4, 2, 2, qa.QT_DTYPE_C128), (4, 2, 1, 2, qa.QT_DTYPE_C128)],   # direct: no scratch
]


@pytest.mark.parametrize("seq", SEQUENCES)
def test_noalloc_equals_contract_pair(ctx, seq):
    rng = np.random.default_rng(len(seq) + seq[0][0])
    for ra, rb, k, d, dt in seq:
        npdt = np.complex128 if dt == qa.QT_DTYPE_C128 else np.complex64
        prec = qa.QT_PLAIN_C128 if dt == qa.QT_DTYPE_C128 else qa.QT_C64_4M
        pairs, free = _case(rng, ra, rb, k)
        a_h = _rand(rng, 1 << (d * ra), npdt)
        b_h = _rand(rng, 1 << (d * rb), npdt)
        a = qa.Tensor(ctx, ra, d, dt).upload(a_h)
        b = qa.Tensor(ctx, rb, d, dt).upload(b_h)
        print(f"noalloc case ranks {ra} {rb} k {k} dim_log2 {d} dtype {dt}", flush=True)
        c_ref, m_ref = qa.contract_pair(ctx, a, b, pairs, free, prec)
        c_na, m_na = _noalloc(ctx, a, b, pairs, free, prec)
        assert m_na == m_ref == 1 << (d * (len(free) + k))
        got, ref = c_na.download(), c_ref.download()
        assert np.array_equal(got, ref), (ra, rb, k)
        if dt == qa.QT_DTYPE_C128 and ra + rb <= 16:  # and the oracle, where it is quick
            o = qo.contract_indices(a_h, ra, b_h, rb, free, pairs, dim_log2=d)
            assert np.allclose(got, o, rtol=1e-12, atol=1e-12 * np.abs(o).max())


def test_noalloc_refuses_what_does_not_fit(ctx):
    """A step whose operand permutes exceed 64 MiB (A: 256 MiB of complex128) and a
    rescaled step are refused with InvalidFunctionInput; the context stays usable."""
    rng = np.random.default_rng(9)
    pairs = [(0, 5), (1, 4), (2, 3)]  # A's summed axes first: A must be permuted
    free = [(True, i) for i in range(3, 12)] + [(False, j) for j in range(3)]
    a = qa.Tensor(ctx, 12, 2, qa.QT_DTYPE_C128)
    b = qa.Tensor(ctx, 6, 2, qa.QT_DTYPE_C128).upload(_rand(rng, 1 << 12, np.complex128))
    with pytest.raises(qa.InvalidFunctionInput):
        _noalloc(ctx, a, b, pairs, free, qa.QT_PLAIN_C128)
    x = qa.Tensor(ctx, 2, 2, qa.QT_DTYPE_C64).upload(_rand(rng, 16, np.complex64))
    y = qa.Tensor(ctx, 2, 2, qa.QT_DTYPE_C64).upload(_rand(rng, 16, np.complex64))
    with pytest.raises(qa.InvalidFunctionInput):
        _noalloc(ctx, x, y, [(0, 1)], [(True, 1), (False, 0)], qa.QT_C64_4M | qa.QT_RESCALE)
    c1, _ = _noalloc(ctx, x, y, [(0, 1)], [(True, 1), (False, 0)], qa.QT_C64_4M)
    c2, _ = qa.contract_pair(ctx, x, y, [(0, 1)], [(True, 1), (False, 0)], qa.QT_C64_4M)
    assert np.array_equal(c1.download(), c2.download())
