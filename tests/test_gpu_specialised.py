"""GPU tests of the specialised executor (qt_batch_compile): generated straight-line
code per program, compiled at run time for the device.  Per output element it runs
the table executors' FMAs in their order (terms with exact-zero constant factors
dropped), so its results equal theirs bit for bit; it follows new parameter blocks,
survives graph capture, and gives way to the table executor when asked to or when
the constants it embeds change."""
import math
import os

import numpy as np
import pytest

import qtorch_amd as qa

pytestmark = pytest.mark.gpu
GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def ctx():
    c = qa.Context(0)
    yield c
    c.close()


def _graph():
    with open(os.path.join(GOLDEN_DIR, "graphs", "3reg_n30_s2024.dgf")) as f:
        return f.read()


def _sets(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0, math.pi / 2, n), -2 * rng.uniform(0, math.pi, n)], axis=1)


@pytest.mark.parametrize("precision,ref_map", [(qa.QT_C64_4M, True), (qa.QT_C64_4M, False),
                                               (qa.QT_PLAIN_C128, True)])
def test_specialised_equals_tables_bitwise(ctx, precision, ref_map):
    """200 angle sets x 45 edges (partial wavefront chunks at every edge's end)."""
    b = qa.QaoaBatch(ctx, _graph(), 1, ref_map=ref_map, precision=precision)
    b.set_angles(_sets(200, 3))
    b.launch()
    e0, z0 = b.results()
    ms = b.compile()
    active, fmas = b.specialised()
    assert active and ms > 0
    assert 0 < fmas < 4 * b.info()["cmacs"]  # fewer FMAs than 4 per complex MAC of the plan
    b.launch()
    e1, z1 = b.results()
    assert np.array_equal(z1, z0) and np.array_equal(e1, e0)
    # new angle sets rewrite the pool but not the constants: the generated code stays
    sets2 = _sets(77, 4)
    b.set_angles(sets2)
    assert b.specialised()[0]
    b.launch()
    e2, z2 = b.results()
    # replays of a captured graph
    b.capture()
    for _ in range(3):
        b.launch()
        e3, z3 = b.results()
        assert np.array_equal(z3, z2) and np.array_equal(e3, e2)
        ctx.sync()
    # an explicit table-executor shape drops the generated code; same values
    b.config(8, 2)
    assert not b.specialised()[0]
    b.launch()
    e4, z4 = b.results()
    assert np.array_equal(z4, z2) and np.array_equal(e4, e2)
    b.close()


def test_specialised_against_reference_values(ctx, golden):
    """The golden angle set: complex128 generated code against the reference's per-edge
    values (1e-12), complex64 4M within the metric's tolerance."""
    ref = sorted((c for c in golden if c["name"].startswith("qaoa3_ref_")), key=lambda c: c["edge"])
    zz_ref = np.array([c["value"][0] for c in ref])
    for precision, tol in ((qa.QT_PLAIN_C128, 1e-12), (qa.QT_C64_4M, 1e-5)):
        b = qa.QaoaBatch(ctx, _graph(), 1, ref_map=True, precision=precision)
        b.set_angles(np.array([ref[0]["betas_gammas"]]))
        b.compile()
        b.launch()
        _, zz = b.results()
        err = np.abs(zz[0] - zz_ref) / np.maximum(np.abs(zz_ref), 1e-3)
        assert err.max() <= tol, (precision, err.max())
        b.close()


def test_specialise_refuses_what_it_cannot_run(ctx):
    """Forms without generated code, programs too large to specialise: the error is
    reported and the table executor keeps running."""
    b = qa.QaoaBatch(ctx, _graph(), 1, ref_map=True, precision=qa.QT_C64_3M)
    b.set_angles(_sets(16, 5))
    with pytest.raises(qa.InvalidFunctionInput):
        b.compile()
    assert not b.specialised()[0]
    b.launch()
    e, _ = b.results()
    assert np.all(np.isfinite(e))
    b.close()
    with open(os.path.join(GOLDEN_DIR, "samples", "3regRand30Node50.dgf")) as f:
        dgf = f.read()
    b2 = qa.QaoaBatch(ctx, dgf, 2, ref_map=False, precision=qa.QT_C64_4M)
    b2.set_angles(np.array([[0.392699, 0.3, 0.785399, -0.6]]))
    with pytest.raises(qa.InvalidFunctionInput):
        b2.compile()
    b2.launch()
    e2, _ = b2.results()
    assert np.isfinite(e2[0])
    b2.close()
