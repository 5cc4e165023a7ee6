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


@pytest.mark.parametrize("precision,ref_map,n_sets", [(qa.QT_C64_4M, True, 200), (qa.QT_C64_4M, False, 200),
                                                      (qa.QT_PLAIN_C128, True, 200), (qa.QT_C64_4M, True, 32768),
                                                      (qa.QT_PLAIN_C128, False, 1024)])
def test_specialised_equals_tables_bitwise(ctx, precision, ref_map, n_sets):
    """200 angle sets x 45 edges (partial wavefront chunks at every edge's end); 1024
    sets form a grid of full chunks; 32768 sets' parameter blocks (8 MiB) exceed an
    XCD's L2 and run in the XCD-tiled order."""
    b = qa.QaoaBatch(ctx, _graph(), 1, ref_map=ref_map, precision=precision)
    b.set_angles(_sets(n_sets, 3))
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


def test_headline_form_on_random_angle_sets(ctx):
    """The form and executor the benchmark reports (complex128, specialised, conjugate
    symmetry, greedy order) on 1024 seeded random angle sets x 45 edges: every edge
    within 1e-12 of an independent FP64 state vector of the reference's own edge
    circuit (|d| <= 1e-12 max(|x|, 1e-3), far inside the metric's 1e-5), and every
    energy F = sum 0.5 (1 - <ZZ>) within 1e-12 relative."""
    from tests import statevector as sv
    dgf = _graph()
    sets = _sets(1024, 2024)
    b = qa.QaoaBatch(ctx, dgf, 1, ref_map=True, precision=qa.QT_PLAIN_C128, order=qa.ORDER_GREEDY)
    b.set_angles(sets)
    b.compile(hermitian=True)
    assert b.specialised()[0]
    b.launch()
    energies, zz = b.results()
    ref = np.empty_like(zz)
    for e in range(b.n_edges):
        texts = [qa.qaoa_edge_text(dgf, 1, True, e, list(bg)) for bg in sets]
        psi, n = sv.simulate_batch([t[0] for t in texts])
        ref[:, e] = sv.expectation_batch(psi, n, texts[0][1]).real
    err = np.abs(zz - ref) / np.maximum(np.abs(ref), 1e-3)
    assert err.max() <= 1e-12, err.max()
    e_ref = np.sum(0.5 * (1.0 - ref), axis=1)
    assert np.max(np.abs(energies - e_ref) / np.abs(e_ref)) <= 1e-12
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


@pytest.mark.parametrize("precision,tol", [(qa.QT_C64_4M, 2e-4), (qa.QT_PLAIN_C128, 1e-12)])
def test_conjugate_symmetric_specialisation(ctx, precision, tol):
    """QT_SPECIALISE_HERMITIAN: one element of each conjugate pair computed; values
    within rounding of the complex128 table executor (per edge with the 1e-3 floor;
    energies 1e-5 / 1e-12 relative), for several sets of angles and after new ones."""
    ref = qa.QaoaBatch(ctx, _graph(), 1, ref_map=True, precision=qa.QT_PLAIN_C128)
    b = qa.QaoaBatch(ctx, _graph(), 1, ref_map=True, precision=precision)
    for k, sets in enumerate((_sets(130, 7), _sets(64, 8))):
        ref.set_angles(sets)
        ref.launch()
        e_ref, z_ref = ref.results()
        b.set_angles(sets)
        if k == 0:
            b.compile(hermitian=True)
        assert b.specialised()[0]
        b.launch()
        e, z = b.results()
        err = np.abs(z - z_ref) / np.maximum(np.abs(z_ref), 1e-3)
        assert err.max() <= tol, err.max()
        assert np.max(np.abs(e - e_ref) / np.abs(e_ref)) <= (1e-5 if precision == qa.QT_C64_4M else 1e-12)
    b.close()
    ref.close()


def test_conjugate_symmetry_is_checked(ctx):
    """Amplitude-picture programs (one bit per axis) cannot use the symmetry: refused,
    and the table executor keeps running."""
    from qtorch_amd import workloads as wl
    q = wl.grid_random_circuit(2, 3, 4, seed=1)
    nb = qa.NetBatch(ctx, [q, q], [wl.bitstring(6, 1), wl.bitstring(6, 2)], precision=qa.QT_C64_4M, dim_log2=1)
    nb.launch()
    v0 = nb.results()
    with pytest.raises(qa.InvalidFunctionInput):
        nb.compile(hermitian=True)
    nb.compile()
    nb.launch()
    v1 = nb.results()
    assert np.array_equal(v0, v1)
    nb.close()


def test_parameter_literals_are_guarded(ctx):
    """Parameter elements equal in every instance at compile time become literals (with
    one angle set, all of them).  New parameter blocks that break that drop the
    generated kernel: the table executor runs and the values stay right; blocks that
    keep it (the same set again) leave the kernel in place."""
    one = _sets(1, 11)
    b = qa.QaoaBatch(ctx, _graph(), 1, ref_map=True, precision=qa.QT_C64_4M)
    b.set_angles(one)
    b.launch()
    e_tab, z_tab = b.results()
    b.compile()
    b.launch()
    e1, z1 = b.results()
    assert b.specialised()[0]
    assert np.array_equal(z1, z_tab) and np.array_equal(e1, e_tab)
    b.set_angles(one)  # the same block: literals still hold
    assert b.specialised()[0]
    many = _sets(50, 12)
    b.set_angles(many)  # different angles: the literals no longer hold
    assert not b.specialised()[0]
    b.launch()
    e2, z2 = b.results()
    ref = qa.QaoaBatch(ctx, _graph(), 1, ref_map=True, precision=qa.QT_C64_4M)
    ref.set_angles(many)
    ref.launch()
    e3, z3 = ref.results()
    assert np.array_equal(z2, z3) and np.array_equal(e2, e3)
    ref.close()
    b.close()


def test_deferred_group_sums(ctx):
    """Direct launches of the specialised kernel sum the previous launch's energies in
    their own first workgroups; results, poll and flush launch the last one's.  Energies
    after launches with different angle sets equal the table executor's bit for bit."""
    import time
    graph = _graph()
    b = qa.QaoaBatch(ctx, graph, 1, ref_map=True, precision=qa.QT_C64_4M)
    t = qa.QaoaBatch(ctx, graph, 1, ref_map=True, precision=qa.QT_C64_4M)
    b.set_angles(_sets(256, 21))
    b.compile()
    for seed in (22, 23):
        sets = _sets(256, seed)
        b.set_angles(sets)
        t.set_angles(sets)
        for _ in range(3):  # consecutive launches: each sums the previous one
            b.launch()
        e, z = b.results()
        t.launch()
        e_t, z_t = t.results()
        assert np.array_equal(z, z_t) and np.array_equal(e, e_t)
    b.launch()
    b.launch()
    done = 0
    t0 = time.time()
    while not done and time.time() - t0 < 30:
        done = b.poll()
    assert done
    e2, _ = b.results()
    assert np.array_equal(e2, e_t)
    b.launch()
    b.flush()
    ctx.sync()
    e3, _ = b.results()
    assert np.array_equal(e3, e_t)
    b.close()
    t.close()


@pytest.mark.parametrize("precision,n_sets,hermitian", [(qa.QT_PLAIN_C128, 200, True), (qa.QT_C64_4M, 200, False),
                                                        (qa.QT_PLAIN_C128, 32768, True)])
def test_group_sums_only_equal_the_summed_values(ctx, precision, n_sets, hermitian):
    """qt_batch_set_sums_only: the generated sums kernel runs 64 angle sets through the
    edges of one slice (edges s, s + 16, s + 32) and writes 16 partial sums per set,
    summed in the next launch; the energies equal, bit for bit, those of launches that
    write every edge's value and sum them with the strided group sums, over
    consecutive launches, captured-graph replays, new angle sets and back."""
    b = qa.QaoaBatch(ctx, _graph(), 1, ref_map=True, precision=precision)
    b.set_angles(_sets(n_sets, 21))
    b.compile(hermitian=hermitian)
    b.launch()
    e_vals, zz = b.results()
    assert zz is not None
    assert b.sums_only(True)
    for _ in range(3):  # consecutive launches: each sums the previous one's partial sums
        b.launch()
    e1, z1 = b.results()
    assert z1 is None and np.array_equal(e1, e_vals)
    sets2 = _sets(n_sets, 22)
    b.set_angles(sets2)
    assert b.sums_only(True)  # still active for the new instances
    b.launch()
    e2, _ = b.results()
    b.capture()
    for _ in range(3):
        b.launch()
        e3, _ = b.results()
        assert np.array_equal(e3, e2)
        ctx.sync()
    assert not b.sums_only(False)
    b.launch()
    e4, z4 = b.results()
    assert np.array_equal(e4, e2) and z4 is not None
    assert np.array_equal(e4, np.sum(0.5 * (1.0 - z4), axis=1)) or np.max(np.abs(e4 - np.sum(0.5 * (1.0 - z4), axis=1))) <= 1e-12 * np.max(np.abs(e4))
    b.close()


def _energies(ctx, precision, sets, p=1):
    r = qa.QaoaBatch(ctx, _graph(), p, precision=precision)
    r.set_angles(sets)
    r.launch()
    e, z = r.results()
    r.close()
    return e, z


@pytest.mark.parametrize("precision,p,compiled", [(qa.QT_PLAIN_C128, 1, False), (qa.QT_PLAIN_C128, 1, True),
                                                  (qa.QT_C64_4M, 1, True), (qa.QT_PLAIN_C128, 2, False),
                                                  (qa.QT_C64_4M, 2, False)])
def test_update_angles_matches_rebuilt_networks(ctx, precision, p, compiled):
    """New angles written on the device (qth_qaoa_batch_update_angles ->
    qt_batch_rotation_params) against networks rebuilt from the gates' text on the
    host (QaoaBatch.set_angles): the parsed float angle is the same bit for bit, the
    device's cos / sin differ from glibc's by at most a few units in the last place,
    so complex128 per-edge values agree to 1e-13 and complex64 4M ones to its own
    rounding.  Host- and device-resident angle arrays; the generated kernel stays."""
    n = 300
    b = qa.QaoaBatch(ctx, _graph(), p, precision=precision)
    b.set_angles(_sets_p(n, p, 11))
    if compiled:
        b.compile(hermitian=True)
    for it, seed in enumerate((12, 13)):
        new = _sets_p(n, p, seed)
        if it == 0:
            b.update_angles(new)
        else:
            import torch
            dev = torch.tensor(new, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            b.update_angles(device_ptr=dev.data_ptr(), n_sets=n)
        b.launch()
        e, z = b.results()
        assert b.specialised()[0] == compiled
        e_ref, z_ref = _energies(ctx, precision, new, p)
        tol = 1e-13 if precision == qa.QT_PLAIN_C128 else 2e-6
        assert np.abs(z - z_ref).max() <= tol, (it, np.abs(z - z_ref).max())
        assert np.abs(e - e_ref).max() <= tol * np.abs(e_ref).max() * 4
    b.close()


def _sets_p(n, p, seed):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.uniform(0, math.pi / 2, (n, p)), rng.uniform(-math.pi, math.pi, (n, p))], axis=1)


def test_update_angles_sums_only_and_literals(ctx):
    """Energies-only launches follow device-written angles; a kernel generated from a
    single angle set (every parameter element a literal) is dropped by the first
    update that could change its literals, and the table executor's results follow."""
    b = qa.QaoaBatch(ctx, _graph(), 1, precision=qa.QT_PLAIN_C128)
    b.set_angles(_sets_p(512, 1, 21))
    b.compile(hermitian=True)
    assert b.sums_only(True)
    new = _sets_p(512, 1, 22)
    b.update_angles(new)
    b.launch()
    e, _ = b.results()
    e_ref, _ = _energies(ctx, qa.QT_PLAIN_C128, new)
    assert np.abs(e - e_ref).max() <= 1e-12 * np.abs(e_ref).max()
    b.close()
    one = qa.QaoaBatch(ctx, _graph(), 1, precision=qa.QT_PLAIN_C128)
    one.set_angles(_sets_p(1, 1, 23))
    one.compile()
    assert one.specialised()[0]
    new1 = _sets_p(1, 1, 24)
    one.update_angles(new1)
    assert not one.specialised()[0]
    one.launch()
    e1, z1 = one.results()
    e1_ref, z1_ref = _energies(ctx, qa.QT_PLAIN_C128, new1)
    assert np.abs(z1 - z1_ref).max() <= 1e-13
    one.close()


def test_update_angles_out_of_range_reports(ctx):
    """An angle outside the exact text round trip's range (here 1e300) poisons its
    set's blocks and makes the next results call fail; valid angles recover."""
    b = qa.QaoaBatch(ctx, _graph(), 1, precision=qa.QT_PLAIN_C128)
    sets = _sets_p(8, 1, 31)
    b.set_angles(sets)
    bad = sets.copy()
    bad[3, 1] = 1e300
    b.update_angles(bad)
    b.launch()
    with pytest.raises(qa.EngineError):
        b.results()
    b.update_angles(sets)
    b.launch()
    e, _ = b.results()
    e_ref, _ = _energies(ctx, qa.QT_PLAIN_C128, sets)
    assert np.abs(e - e_ref).max() <= 1e-12 * np.abs(e_ref).max()
    b.close()
