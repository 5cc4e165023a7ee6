"""GPU parity of the plain complex128 path against the oracle and the reference.

* qt_contract_pair on random tensors in every kernel regime (direct, split
  reduction, permute + MFMA GEMM) against the oracle's restatement of the
  reference loop: bit-identical where the direct kernel runs (same summation
  order, no FMA), 1e-12 relative elsewhere;
* whole networks replayed in the reference's own recorded order through the C++
  host on the device, against the values the reference printed (1e-12 relative,
  1e-15 absolute floor);
* seeded stochastic orders on the device against the state-vector oracle;
* F_p of the QAOA p=1 configuration (30-node 3-regular graph, 45 edges) against
  the reference's per-edge values.
"""
import os

import numpy as np
import pytest

import qtorch_amd as qa
from oracle import qtoracle as qo
from tests import statevector as sv

pytestmark = pytest.mark.gpu
GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ctx():
    c = qa.Context(0)
    yield c
    c.close()


def _rand(rng, n, dtype=np.complex128):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)


def _case(rng, ra, rb, k, interleave=False):
    pa = list(rng.permutation(ra)[:k])
    pb = list(rng.permutation(rb)[:k])
    free = [(True, i) for i in range(ra) if i not in pa] + [(False, j) for j in range(rb) if j not in pb]
    if interleave:
        free = [free[i] for i in rng.permutation(len(free))]
    return list(zip(pa, pb)), free


# (rank_a, rank_b, k, dim_log2): each regime of the dispatcher
SHAPES = [
    (1, 1, 1, 2), (2, 1, 1, 2), (2, 2, 1, 2), (4, 2, 1, 2), (4, 4, 2, 2), (6, 2, 1, 2), (3, 3, 0, 2),
    (0, 0, 0, 2), (5, 1, 1, 2), (8, 4, 2, 2),          # direct
    (7, 7, 7, 2), (8, 7, 7, 2), (7, 8, 7, 2),          # split reduction (small output, large sum)
    (5, 5, 2, 2), (6, 6, 3, 2), (7, 6, 3, 2), (6, 7, 4, 2),  # permute + MFMA GEMM
    (10, 10, 5, 1), (12, 9, 4, 1), (14, 14, 14, 1), (6, 4, 2, 1),  # dimension-2 tensors
    (8, 2, 1, 2), (2, 8, 1, 2), (7, 4, 2, 2), (10, 6, 5, 2),  # skinny: <= 5 free bits on one side
    (12, 3, 2, 1), (3, 12, 1, 1), (14, 8, 3, 1), (8, 14, 3, 1), (12, 2, 2, 1),
    (15, 15, 3, 1),  # GEMM with a summed extent of 8 (zero-filled k-tile)
]


def _regime(free, k, d):
    """Which kernel the dispatcher picks (qt_contract.hip: contract_pair)."""
    na = sum(1 for f in free if f[0])
    split = all(f[0] for f in free[:na])
    nc, ns = len(free) * d, k * d
    if split and na * d >= 6 and (len(free) - na) * d >= 6 and (ns >= 4 or (ns >= 3 and nc >= 24)):
        return "gemm"
    if nc <= 6 and ns >= 13:
        return "reduce"
    if split and min(na * d, nc - na * d) <= 5 and max(na * d, nc - na * d) >= 10:
        return "skinny"
    return "direct"


@pytest.mark.parametrize("ra,rb,k,d", SHAPES)
def test_contract_pair_plain_vs_oracle(ctx, ra, rb, k, d):
    rng = np.random.default_rng(100 * ra + 10 * rb + k + d)
    for interleave in (False, True):
        pairs, free = _case(rng, ra, rb, k, interleave)
        a = _rand(rng, 1 << (d * ra))
        b = _rand(rng, 1 << (d * rb))
        ta = qa.Tensor(ctx, ra, d).upload(a)
        tb = qa.Tensor(ctx, rb, d).upload(b)
        tc, macs = qa.contract_pair(ctx, ta, tb, pairs, free)
        got = tc.download()
        ref = qo.contract_indices(a, ra, b, rb, free, pairs, d)
        assert macs == 1 << (d * (len(free) + k))
        if _regime(free, k, d) in ("direct", "skinny"):  # the reference's summation order
            assert np.array_equal(got, ref), (ra, rb, k, d, interleave)
        else:
            np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


# (big rank, gate rank, big axes summed (in the gate's axis order), dim_log2)
GATE_STEPS = [(20, 4, (0, 1), 1), (20, 4, (7, 2), 1), (20, 4, (18, 19), 1), (20, 3, (0,), 1), (20, 3, (19,), 1),
              (20, 5, (3, 11, 6), 1), (21, 6, (1, 4, 9, 20), 1), (10, 2, (0,), 2), (10, 2, (9,), 2),
              (10, 1, (4,), 2), (10, 4, (2, 7), 2)]


@pytest.mark.parametrize("prec", [qa.QT_PLAIN_C128, qa.QT_C64_4M])
@pytest.mark.parametrize("big_first", [True, False])
@pytest.mark.parametrize("rbig,rg,pos,d", GATE_STEPS)
def test_gate_on_large_tensor(ctx, rbig, rg, pos, d, big_first, prec):
    """A gate (or a thin tensor) against a 2^20-element tensor whose free axes keep
    their order, as every step of a circuit sweep: the large side's offset comes from
    inserting zero bits at its summed positions (k_thin).  complex128 is bit-identical
    to the oracle's loop in the reference's order (src/Network.h:876-971); 4M within
    FP32 rounding.  Summed positions at the bottom, the top and unsorted."""
    rng = np.random.default_rng(rbig * 100 + rg * 10 + sum(pos) + d)
    dt = np.complex128 if prec == qa.QT_PLAIN_C128 else np.complex64
    dtype = qa.QT_DTYPE_C128 if prec == qa.QT_PLAIN_C128 else qa.QT_DTYPE_C64
    big = _rand(rng, 1 << (d * rbig), dt)
    gate = _rand(rng, 1 << (d * rg), dt)
    k = len(pos)
    big_free = [a for a in range(rbig) if a not in pos]
    if big_first:
        ra, rb, a, b = rbig, rg, big, gate
        pairs = [(p, i) for i, p in enumerate(pos)]
        free = [(True, x) for x in big_free] + [(False, g) for g in range(k, rg)]
    else:
        ra, rb, a, b = rg, rbig, gate, big
        pairs = [(i, p) for i, p in enumerate(pos)]
        free = [(True, g) for g in range(k, rg)] + [(False, x) for x in big_free]
    assert _regime(free, k, d) == "skinny"
    ta = qa.Tensor(ctx, ra, d, dtype).upload(a)
    tb = qa.Tensor(ctx, rb, d, dtype).upload(b)
    tc, _ = qa.contract_pair(ctx, ta, tb, pairs, free, prec)
    got = tc.download()
    ref = qo.contract_indices(a.astype(np.complex128), ra, b.astype(np.complex128), rb, free, pairs, d)
    if prec == qa.QT_PLAIN_C128:
        assert np.array_equal(got, ref)
    else:
        assert np.abs(got.astype(np.complex128) - ref).max() <= 2e-6 * np.sqrt(1 << (d * k)) * np.abs(ref).max()


@pytest.mark.parametrize("prec", [qa.QT_PLAIN_C128, qa.QT_C64_4M])
def test_gemm_with_many_permute_tiles(ctx, prec):
    """A 22-bit operand whose summed axes are its lowest: the GEMM path permutes it in
    2^12 or more LDS tiles (a two-dimensional grid of workgroups)."""
    rng = np.random.default_rng(5)
    ra, rb, d = 11, 5, 2
    pairs = [(0, 1), (1, 0)]
    free = [(True, i) for i in range(2, ra)] + [(False, j) for j in range(2, rb)]
    dt = np.complex128 if prec == qa.QT_PLAIN_C128 else np.complex64
    a = _rand(rng, 1 << (d * ra), dt)
    b = _rand(rng, 1 << (d * rb), dt)
    assert _regime(free, len(pairs), d) == "gemm"
    ta = qa.Tensor(ctx, ra, d, qa.QT_DTYPE_C128 if dt == np.complex128 else qa.QT_DTYPE_C64).upload(a)
    tb = qa.Tensor(ctx, rb, d, qa.QT_DTYPE_C128 if dt == np.complex128 else qa.QT_DTYPE_C64).upload(b)
    tc, _ = qa.contract_pair(ctx, ta, tb, pairs, free, prec)
    got = tc.download().astype(np.complex128)
    ref = qo.contract_indices(a.astype(np.complex128), ra, b.astype(np.complex128), rb, free, pairs, d)
    tol = 1e-12 if prec == qa.QT_PLAIN_C128 else 2e-6 * np.sqrt(16)
    assert np.abs(got - ref).max() <= tol * np.abs(ref).max()


_C64_REF = {}  # oracle results by shape: the inputs depend only on the shape's seed


@pytest.mark.parametrize("prec,tol", [(qa.QT_C64_4M, 2e-6), (qa.QT_C64_3M, 4e-6), (qa.QT_C64_BF16X3, 1e-4)])
@pytest.mark.parametrize("ra,rb,k,d", [(4, 4, 2, 2), (7, 7, 7, 2), (6, 6, 3, 2), (10, 10, 5, 1),
                                      (12, 12, 6, 1), (14, 13, 7, 1), (8, 8, 3, 2), (9, 9, 4, 2),
                                      (8, 2, 1, 2), (3, 12, 1, 1), (10, 6, 5, 2), (11, 11, 3, 1), (12, 12, 4, 1),
                                      (19, 19, 12, 1)])
def test_contract_pair_c64_forms(ctx, prec, tol, ra, rb, k, d):
    """complex64 forms against the FP64 oracle: error relative to the largest |C|.
    Shapes with 64+ free elements on each side and 32+ summed run on the MFMA GEMM
    (FP32 MFMA for 4M / 3M, bf16 MFMA for bf16x3), the rest on the vector kernels.
    The last (a 128 x 128 output, 2^12 summed) splits the summed range of the GEMM
    into 16 partial products in every form; its oracle result (2^26 MACs in the
    reference's loop nest) is computed once for the three forms."""
    rng = np.random.default_rng(7 + ra + rb)
    pairs, free = _case(rng, ra, rb, k)
    a = _rand(rng, 1 << (d * ra), np.complex64)
    b = _rand(rng, 1 << (d * rb), np.complex64)
    ta = qa.Tensor(ctx, ra, d, qa.QT_DTYPE_C64).upload(a)
    tb = qa.Tensor(ctx, rb, d, qa.QT_DTYPE_C64).upload(b)
    tc, _ = qa.contract_pair(ctx, ta, tb, pairs, free, prec)
    got = tc.download().astype(np.complex128)
    key = (ra, rb, k, d)
    if key not in _C64_REF:
        _C64_REF[key] = qo.contract_indices(a.astype(np.complex128), ra, b.astype(np.complex128), rb, free, pairs, d)
    ref = _C64_REF[key]
    scale = np.abs(ref).max() * np.sqrt(1 << (d * k))
    assert np.abs(got - ref).max() <= tol * scale


def test_gemm_split_k_c128(ctx):
    """A complex128 GEMM-shaped step with one 128 x 128 output (four 64 x 64 tiles) and
    2^12 summed indices: the summed range splits into 16 partial products added in
    order; against the oracle's loop nest (src/Network.h:899-933)."""
    rng = np.random.default_rng(19)
    ra = rb = 19
    pairs, free = _case(rng, ra, rb, 12)
    a = _rand(rng, 1 << ra)
    b = _rand(rng, 1 << rb)
    ta = qa.Tensor(ctx, ra, 1).upload(a)
    tb = qa.Tensor(ctx, rb, 1).upload(b)
    tc, _ = qa.contract_pair(ctx, ta, tb, pairs, free)
    got = tc.download()
    ref = qo.contract_indices(a, ra, b, rb, free, pairs, 1)
    assert _regime(free, 12, 1) == "gemm"
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


def test_rescaled_chain_does_not_underflow(ctx):
    """A chain of contractions whose values shrink by 2^-30 per factor: plain complex64
    underflows to zero, while per-step power-of-two rescaling keeps the FP64 value to
    FP32 accuracy (exponents carried on the device, applied when the scalar is read)."""
    rng = np.random.default_rng(3)
    scale = 2.0 ** -30
    v0 = (_rand(rng, 2) * scale).astype(np.complex64)
    mats = [(_rand(rng, 4) * scale).astype(np.complex64) for _ in range(6)]
    u = (_rand(rng, 2) * scale).astype(np.complex64)
    ref = v0.astype(np.complex128)
    for m in mats:  # t'[j] = sum_i t[i] M[i + 2 j]
        ref = ref @ m.astype(np.complex128).reshape(2, 2, order="F")
    ref = complex(ref @ u.astype(np.complex128))
    assert abs(ref) < 2.0 ** -200
    for rescale in (True, False):
        prec = qa.QT_C64_4M | (qa.QT_RESCALE if rescale else 0)
        t = qa.Tensor(ctx, 1, 1, qa.QT_DTYPE_C64).upload(v0)
        for m in mats:
            tm = qa.Tensor(ctx, 2, 1, qa.QT_DTYPE_C64).upload(m)
            t, _ = qa.contract_pair(ctx, t, tm, [(0, 0)], [(False, 1)], prec)
        tu = qa.Tensor(ctx, 1, 1, qa.QT_DTYPE_C64).upload(u)
        s_, _ = qa.contract_pair(ctx, t, tu, [(0, 0)], [], prec)
        got = s_.scalar()
        if rescale:
            assert abs(got - ref) <= 1e-5 * abs(ref), (got, ref)
            assert s_.exponent() < -200
        else:
            assert got == 0


def test_plain_step_after_rescaled_steps_keeps_exponents(ctx):
    """A step without QT_RESCALE whose operands carry exponents from rescaled steps
    gives C the sum of their exponents (its data are then the plain product of the
    scaled operands): the value read back equals the all-rescaled chain's."""
    rng = np.random.default_rng(5)
    scale = 2.0 ** -40
    v0 = (_rand(rng, 2) * scale).astype(np.complex64)
    mats = [(_rand(rng, 4) * scale).astype(np.complex64) for _ in range(3)]
    ref = v0.astype(np.complex128)
    for m in mats:
        ref = ref @ m.astype(np.complex128).reshape(2, 2, order="F")
    t = qa.Tensor(ctx, 1, 1, qa.QT_DTYPE_C64).upload(v0)
    for i, m in enumerate(mats):
        prec = qa.QT_C64_4M | (qa.QT_RESCALE if i < 2 else 0)  # the last step is plain
        tm = qa.Tensor(ctx, 2, 1, qa.QT_DTYPE_C64).upload(m)
        t, _ = qa.contract_pair(ctx, t, tm, [(0, 0)], [(False, 1)], prec)
    e = t.exponent()
    assert e < -60
    got = t.download().astype(np.complex128) * 2.0 ** e
    assert np.allclose(got, ref, rtol=1e-5, atol=0), (got, ref)
    # two rescaled operands into one plain step: both exponents are carried
    tu = qa.Tensor(ctx, 1, 1, qa.QT_DTYPE_C64).upload((_rand(rng, 2) * scale).astype(np.complex64))
    tv = qa.Tensor(ctx, 2, 1, qa.QT_DTYPE_C64).upload((_rand(rng, 4) * scale).astype(np.complex64))
    su, _ = qa.contract_pair(ctx, tu, tv, [(0, 0)], [(False, 1)], qa.QT_C64_4M | qa.QT_RESCALE)
    s_, _ = qa.contract_pair(ctx, t, su, [(0, 0)], [], qa.QT_C64_4M)
    assert s_.exponent() == t.exponent() + su.exponent()
    ref_s = complex(np.sum(got * (su.download().astype(np.complex128) * 2.0 ** su.exponent())))
    assert abs(s_.scalar() - ref_s) <= 1e-5 * abs(ref_s)


def test_context_time_limit_returns_timeout(ctx):
    """qt_ctx_set_time_limit: once the limit has passed, qt_contract_pair and batch
    launches return QT_TIMEOUT (ContractionFailure) before queueing anything; removing
    the limit restores them."""
    import time
    rng = np.random.default_rng(6)
    a = qa.Tensor(ctx, 2, 2).upload(_rand(rng, 16))
    b = qa.Tensor(ctx, 2, 2).upload(_rand(rng, 16))
    c0, _ = qa.contract_pair(ctx, a, b, [(0, 1)], [(True, 1), (False, 0)])
    ctx.set_time_limit(1e-6)
    time.sleep(0.01)
    try:
        with pytest.raises(qa.ContractionFailure):
            qa.contract_pair(ctx, a, b, [(0, 1)], [(True, 1), (False, 0)])
    finally:
        ctx.set_time_limit(0)
    c1, _ = qa.contract_pair(ctx, a, b, [(0, 1)], [(True, 1), (False, 0)])
    assert np.array_equal(c1.download(), c0.download())
    # a generous limit does not get in the way
    ctx.set_time_limit(3600)
    try:
        c2, _ = qa.contract_pair(ctx, a, b, [(0, 1)], [(True, 1), (False, 0)])
        assert np.array_equal(c2.download(), c0.download())
    finally:
        ctx.set_time_limit(0)


def test_device_fp64_accumulation(ctx):
    """Rank-0 results of any form are summed in FP64 on the device."""
    rng = np.random.default_rng(4)
    acc = qa.Tensor(ctx, 0, 2, qa.QT_DTYPE_C128)
    total = 0j
    for i in range(5):
        a = _rand(rng, 16, np.complex64)
        b = _rand(rng, 16, np.complex64)
        ta = qa.Tensor(ctx, 2, 2, qa.QT_DTYPE_C64).upload(a)
        tb = qa.Tensor(ctx, 2, 2, qa.QT_DTYPE_C64).upload(b)
        ts, _ = qa.contract_pair(ctx, ta, tb, [(0, 1), (1, 0)], [], qa.QT_C64_4M | qa.QT_RESCALE)
        ts.accumulate_into(acc)
        total += ts.scalar()
    assert abs(acc.scalar() - total) <= 1e-12 * abs(total)


def test_invalid_arguments_are_rejected(ctx):
    ta = qa.Tensor(ctx, 2)
    tb = qa.Tensor(ctx, 2)
    with pytest.raises(qa.InvalidFunctionInput):
        qa.contract_pair(ctx, ta, tb, [(0, 0), (0, 1)], [(True, 1), (False, 1)][:0])
    with pytest.raises(qa.InvalidFunctionInput):
        qa.contract_pair(ctx, ta, tb, [(5, 0)], [(True, 1), (False, 1)])
    with pytest.raises(qa.InvalidFunctionInput):  # precision does not fit the dtype
        qa.contract_pair(ctx, ta, tb, [(0, 0)], [(True, 1), (False, 1)], qa.QT_C64_4M)


def _qasm(case):
    with open(os.path.join(GOLDEN_DIR, case["qasm"])) as f:
        return f.read()


def _close(got, ref):
    return abs(got - ref) <= 1e-12 * abs(ref) + 1e-15


def test_golden_networks_on_device(ctx, golden):
    """Every golden case, replayed in the reference's own order on the device."""
    n_exact = 0
    for case in golden:
        print("golden", case["name"], flush=True)
        net = qa.Network(_qasm(case), case["measure"], ctx, base_dir=GOLDEN_DIR)
        mode = case["mode"]
        if mode == "stoch":
            got = net.contract_given(case["seq"])
        elif mode == "rstoch":
            net.reduce()
            got = net.contract_given(case["seq"])
        elif mode in ("lg", "rlg"):
            if mode == "rlg":
                net.reduce()
            got = net.linegraph_contract(case["qbb_order"])
        else:
            got = net.contract_user(case["seq"])
        ref = complex(*case["value"])
        assert _close(got, ref), (case["name"], got, ref)
        assert net.flops() == case["flops"]
        n_exact += got == ref
        net.close()
    print(f"bit-identical to the reference: {n_exact}/{len(golden)}")


def test_seeded_stochastic_on_device_vs_statevector(ctx, golden):
    names = ["jw_YXXY", "rand_rxz_Y2", "toffoli_011", "arb2q_70_00_YT", "qft8_bits_stoch", "cat8_1"]
    for case in golden:
        if case["name"] not in names:
            continue
        q = _qasm(case)
        for seed in (1, 2):
            net = qa.Network(q, case["measure"], ctx, base_dir=GOLDEN_DIR)
            got = net.contract_stochastic(seed)
            psi, n = sv.simulate(q, base_dir=GOLDEN_DIR)
            ref = sv.expectation(psi, n, case["measure"])
            assert _close(got, ref) or abs(got - ref) < 1e-13, (case["name"], got, ref)


def test_qaoa_p1_energy_plain_vs_reference(ctx, golden):
    with open(os.path.join(GOLDEN_DIR, "graphs", "3reg_n30_s2024.dgf")) as f:
        dgf = f.read()
    ref_cases = sorted((c for c in golden if c["name"].startswith("qaoa3_ref_")), key=lambda c: c["edge"])
    bg = ref_cases[0]["betas_gammas"]
    energy, zz = qa.qaoa_energy_plain(ctx, dgf, 1, True, bg, seed=11)
    ref_zz = np.array([c["value"][0] for c in ref_cases])
    assert np.all(np.abs(zz - ref_zz) <= 1e-12 * np.abs(ref_zz) + 1e-15)
    assert abs(energy - np.sum(0.5 * (1 - ref_zz))) < 1e-12 * abs(energy)


def test_teleportation_known_answer(ctx, tmp_path):
    """The reference's teleportation test (src/tests.cpp:1433-1526): a one-qubit gate
    file [[alpha, 0], [beta, 0]] prepares alpha|0> + beta|1> on qubit 0, then H 1,
    CNOT 1 2, CNOT 0 1, H 0; the projectors 0 0 0 / 0 0 1 / 1 0 0 / 1 0 1 give
    |alpha/2|^2, |beta/2|^2, |alpha/2|^2, |beta/2|^2.  Ten seeded states, written as
    the reference writes them (std::complex at the default 6 digits), contracted on
    the device in seeded stochastic orders; the reference's tolerance is 1e-5, here
    1e-12 against the closed form of the printed values."""
    rng = np.random.default_rng(1433)
    for i in range(10):
        t = rng.uniform(-1, 1, 4)
        t /= np.sqrt(np.sum(t * t))
        alpha, beta = complex(t[0], t[1]), complex(t[2], t[3])
        gate = tmp_path / "testTeleportGate.gate"
        gate.write_text(f"({alpha.real:g},{alpha.imag:g}) 0 ({beta.real:g},{beta.imag:g}) 0")
        a_w = complex(float(f"{alpha.real:g}"), float(f"{alpha.imag:g}"))
        b_w = complex(float(f"{beta.real:g}"), float(f"{beta.imag:g}"))
        qasm = "3\ndef1 AB testTeleportGate.gate\nAB 0\nH 1\nCNOT 1 2\nCNOT 0 1\nH 0\n"
        for meas, amp in (("0 0 0", a_w), ("0 0 1", b_w), ("1 0 0", a_w), ("1 0 1", b_w)):
            net = qa.Network(qasm, meas, ctx, base_dir=str(tmp_path))
            v = net.contract_stochastic(10 * i + len(meas))
            assert abs(v.real - abs(amp / 2) ** 2) <= 1e-12 and abs(v.imag) <= 1e-12, (i, meas, v, amp)
