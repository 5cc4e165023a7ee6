"""GPU tests of the serving loop (qth_server_*): continuous batching of QAOA-energy
and single-amplitude requests.  Results must not depend on what else shares a
launch: each request's value in a mixed stream is compared bit for bit with the
same request served alone, and against independent references (the drop-in
complex128 path, a state-vector simulation)."""
import math
import os
import random
import threading
import time

import numpy as np
import pytest

import qtorch_amd as qa
from qtorch_amd import workloads as wl
from tests import statevector as sv

pytestmark = pytest.mark.gpu
GOLDEN_DIR = os.path.join(os.path.dirname(__file__), "golden")


def _graph():
    with open(os.path.join(GOLDEN_DIR, "graphs", "3reg_n30_s2024.dgf")) as f:
        return f.read()


def _amp_index(bits):
    return sum(int(b) << q for q, b in enumerate(bits))


@pytest.fixture(scope="module")
def ctx():
    c = qa.Context(0)
    yield c
    c.close()


def _mix(seed):
    """A seeded mix: QAOA energies (3- and 4-regular graphs, p = 1, 2) and amplitudes
    of small grid circuits and QFTs; every request appears in arrival order."""
    rng = random.Random(seed)
    reqs = []
    graphs = [(wl.dgf_text(wl.regular_graph(n, d, seed=100 + n + d)), d) for n, d in ((20, 3), (24, 4), (30, 3))]
    circuits = [(wl.grid_random_circuit(3, 4, 8, seed=5), 12), (wl.qft_qasm(8), 8),
                (wl.grid_random_circuit(4, 4, 8, seed=6), 16)]
    for i in range(48):
        if rng.random() < 0.7:
            g, d = rng.choice(graphs)
            p = rng.choice([1, 1, 2]) if d == 3 else 1
            ang = wl.qaoa_angles(p, rng.randrange(1 << 20))
            reqs.append(("qaoa", g, p, ang))
        else:
            q, n = rng.choice(circuits)
            reqs.append(("amp", q, wl.bitstring(n, rng.randrange(1 << 20))))
    return reqs


def _submit(srv, r):
    if r[0] == "qaoa":
        return srv.submit_qaoa(r[1], r[2], r[3])
    return srv.submit_amplitude(r[1], r[2])


def test_request_alone_equals_request_in_mix():
    """Every request of a mixed stream (submitted from another thread while launches
    are in flight) returns exactly the value it returns when served alone."""
    reqs = _mix(7)
    alone = []
    srv = qa.Server(slots=2, precision=qa.QT_C64_4M)
    try:
        for r in reqs:
            res = srv.wait(_submit(srv, r))
            assert res["status"] == 0, res
            alone.append(res["value"])
    finally:
        srv.close()
    srv = qa.Server(slots=2, precision=qa.QT_C64_4M)
    try:
        ids = [None] * len(reqs)
        order = list(range(len(reqs)))
        random.Random(3).shuffle(order)

        def producer():
            for k in order:
                ids[k] = _submit(srv, reqs[k])
                time.sleep(0.0005)

        t = threading.Thread(target=producer)
        t.start()
        t.join()
        got = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    for k, (a, g) in enumerate(zip(alone, got)):
        assert g["status"] == 0, g
        assert g["value"] == a, (k, reqs[k][0], a, g)
    assert st["completed"] == len(reqs) and st["launches"] >= 2
    assert st["p99_ms"] >= st["p50_ms"] > 0


def test_served_values_match_references(ctx):
    """QAOA energies against the drop-in complex128 path (1e-5 relative); amplitudes
    against a state-vector simulation (1e-5 of the state's scale)."""
    dgf = _graph()
    srv = qa.Server(slots=2, precision=qa.QT_C64_4M)
    try:
        checks = []
        for s in range(4):
            ang = wl.qaoa_angles(1, 40 + s)
            checks.append(("qaoa", srv.submit_qaoa(dgf, 1, ang), ang))
        for s, (r, c) in enumerate([(3, 4), (4, 4)]):
            q = wl.grid_random_circuit(r, c, 8, seed=20 + s)
            bits = wl.bitstring(r * c, 9 + s)
            checks.append(("amp", srv.submit_amplitude(q, bits), (q, bits)))
        for kind, rid, arg in checks:
            res = srv.wait(rid)
            assert res["status"] == 0, res
            if kind == "qaoa":
                e, _ = qa.qaoa_energy_plain(ctx, dgf, 1, False, list(arg), seed=1)
                assert abs(res["value"].real - e) <= 1e-5 * abs(e)
            else:
                q, bits = arg
                psi, n = sv.simulate(q)
                ref = psi[_amp_index(bits)]
                assert abs(res["value"] - ref) <= 1e-5 * max(abs(ref), 2.0 ** (-n / 2)), (res, ref)
    finally:
        srv.close()


def test_large_network_takes_the_large_network_path():
    """A 20-qubit depth-14 grid amplitude has tensors beyond a program's limit: it runs
    on the complex128 large-network path (sliced, fused small subtrees; matching a
    state-vector simulation to 1e-10) while small requests keep flowing, and its
    planned bytes are counted at admission."""
    q = wl.grid_random_circuit(4, 5, 14, seed=1)
    bits = wl.bitstring(20, 3)
    plan = qa.serve_plan(qa.SERVE_AMPLITUDE, q, bits=bits)
    assert plan["route"] == "eager"
    srv = qa.Server(slots=2)
    try:
        big = srv.submit_amplitude(q, bits)
        small = [srv.submit_qaoa(_graph(), 1, wl.qaoa_angles(1, s)) for s in range(8)]
        res = srv.wait(big)
        for i in small:
            assert srv.wait(i)["status"] == 0
        st = srv.stats()
    finally:
        srv.close()
    assert res["status"] == 0 and res["route"] == "eager"
    psi, n = sv.simulate(q)
    ref = psi[_amp_index(bits)]
    assert abs(res["value"] - ref) <= 1e-10 * max(abs(ref), 2.0 ** (-n / 2))
    assert st["peak_planned_bytes"] >= plan["planned_bytes"]


def test_large_network_path_in_complex64():
    """The large-network route in the complex64 4M form (qth_server_create_ex): the
    20-qubit depth-14 grid amplitude and a large light cone's energy within that
    form's error of a state-vector simulation (1e-5 relative)."""
    q = wl.grid_random_circuit(4, 5, 14, seed=1)
    bits = wl.bitstring(20, 3)
    srv = qa.Server(slots=2, eager_precision=qa.QT_C64_4M)
    try:
        res = srv.wait(srv.submit_amplitude(q, bits))
    finally:
        srv.close()
    assert res["status"] == 0 and res["route"] == "eager"
    psi, n = sv.simulate(q)
    ref = psi[_amp_index(bits)]
    assert abs(res["value"] - ref) <= 1e-5 * max(abs(ref), 2.0 ** (-n / 2))


def test_admission_rejects_what_cannot_fit():
    """A request whose planned device bytes exceed the whole budget is refused
    (QTH_OTHER) without touching the device; others still run."""
    q = wl.grid_random_circuit(4, 5, 14, seed=1)
    bits = wl.bitstring(20, 3)
    plan = qa.serve_plan(qa.SERVE_AMPLITUDE, q, bits=bits)
    srv = qa.Server(slots=1, hbm_budget=plan["planned_bytes"] // 2)
    try:
        r1 = srv.wait(srv.submit_amplitude(q, bits))
        r2 = srv.wait(srv.submit_qaoa(_graph(), 1, wl.qaoa_angles(1, 1)))
        bad = srv.wait(srv.submit_amplitude(q, bits[:-1] + "2"))
        st = srv.stats()
    finally:
        srv.close()
    assert r1["status"] == 9 and r2["status"] == 0
    assert bad["status"] != 0
    assert st["rejected"] == 2


@pytest.mark.parametrize("n,d,p", [(8, 3, 3), (10, 4, 2)])
def test_large_light_cones_served_exactly(n, d, p):
    """p = 3 and 4-regular p = 2 light cones exceed a program's tensors: they run on the
    large-network path (planned once per structure, each request's angles set into the
    plan) in complex128; the energy matches a state-vector simulation of every edge's
    light-cone circuit to 1e-10, for two angle sets in a row."""
    dgf = wl.dgf_text(wl.regular_graph(n, d, seed=31))
    assert qa.serve_plan(qa.SERVE_QAOA_ENERGY, dgf, p)["route"] == "eager"
    ne, _ = qa.qaoa_num_edges(dgf)
    srv = qa.Server(slots=1)
    try:
        for s in (1, 2):
            ang = wl.qaoa_angles(p, s)
            res = srv.wait(srv.submit_qaoa(dgf, p, ang))
            assert res["status"] == 0 and res["route"] == "eager", res
            ref = 0.0
            for e in range(ne):
                q, m = qa.qaoa_edge_text(dgf, p, False, e, ang)
                psi, nq = sv.simulate(q)
                ref += 0.5 * (1.0 - sv.expectation(psi, nq, m).real)
            assert abs(res["value"].real - ref) <= 1e-10 * abs(ref), (res, ref)
    finally:
        srv.close()


def test_admission_rejects_too_much_work():
    """A request whose planned work exceeds the per-request limit (here p = 3 light
    cones of a 20-vertex graph, about 1e18 complex MACs) is refused without running."""
    dgf = wl.dgf_text(wl.regular_graph(20, 3, seed=920))
    srv = qa.Server(slots=1)
    try:
        res = srv.wait(srv.submit_qaoa(dgf, 3, wl.qaoa_angles(3, 1)))
        st = srv.stats()
    finally:
        srv.close()
    assert res["status"] == 9 and st["rejected"] == 1


def test_large_route_requests_in_flight_together():
    """Three large-network requests of one structure with different angles, submitted
    before any is waited for: two run at once (each on its own streams, sharing the
    plan, each with its own leaf values set from its angles); every energy matches a
    state vector of every edge's light cone to 1e-10, and the planned bytes admitted
    at once never exceed the budget."""
    dgf = wl.dgf_text(wl.regular_graph(10, 4, seed=31))
    p = 2
    assert qa.serve_plan(qa.SERVE_QAOA_ENERGY, dgf, p)["route"] == "eager"
    ne, _ = qa.qaoa_num_edges(dgf)
    budget = 16 << 30
    srv = qa.Server(slots=1, hbm_budget=budget)
    try:
        angs = [wl.qaoa_angles(p, s) for s in (11, 12, 13)]
        ids = [srv.submit_qaoa(dgf, p, a) for a in angs]
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    for ang, r in zip(angs, res):
        assert r["status"] == 0 and r["route"] == "eager", r
        ref = 0.0
        for e in range(ne):
            q, m = qa.qaoa_edge_text(dgf, p, False, e, ang)
            psi, nq = sv.simulate(q)
            ref += 0.5 * (1.0 - sv.expectation(psi, nq, m).real)
        assert abs(r["value"].real - ref) <= 1e-10 * abs(ref), (r, ref)
    assert 0 < st["peak_planned_bytes"] <= budget


def test_configured_mix_at_a_low_rate():
    """The configured request mix at a low rate: QAOA energies on 3- and 4-regular
    graphs (p = 1, 2, 3) and amplitudes of grid circuits from 4x6 depth 8 to the
    48-qubit 6x8 depth 14, submitted with Poisson arrivals.  Every request either
    completes with the value computed on its own (QAOA: the drop-in complex128 path,
    1e-5; amplitudes: the state vector to 24 qubits, 1e-5 on the complex64 batched
    routes and 1e-10 on the complex128 large-network route, else the same plan
    contracted directly, 1e-10) or is refused for a stated reason (planned work);
    the 48-qubit depth-14 amplitude is served (circuit sweep plan); per-class p99
    latencies come back."""
    rng = random.Random(17)
    ctx = qa.Context(0)
    graphs = [(wl.dgf_text(wl.regular_graph(n, d, seed=500 + n)), d) for n, d in ((12, 3), (14, 4), (16, 3))]
    circuits = [(wl.grid_random_circuit(r, c, dep, seed=600 + r * c + dep), r * c, f"{r}x{c} d{dep}")
                for r, c, dep in ((4, 6, 8), (4, 6, 14), (6, 8, 14))]
    reqs = []
    for k in range(24):
        if k % 3:
            dgf, d = graphs[k % len(graphs)]
            p = (1, 2, 1, 2, 3)[k % 5]
            reqs.append(("qaoa", dgf, p, wl.qaoa_angles(p, 40 + k), f"qaoa {d}-regular p={p}"))
        else:
            q, n, label = circuits[(k // 3) % len(circuits)]
            reqs.append(("amp", q, wl.bitstring(n, 70 + k), n, label))
    srv = qa.Server(slots=2, precision=qa.QT_C64_4M, hbm_budget=64 << 30)
    try:
        ids = []
        for r in reqs:
            time.sleep(rng.expovariate(200.0))
            ids.append(srv.submit_qaoa(r[1], r[2], r[3]) if r[0] == "qaoa" else srv.submit_amplitude(r[1], r[2]))
        res = [srv.wait(i) for i in ids]
        st = srv.stats()
    finally:
        srv.close()
    lat = {}
    served_48 = 0
    for r, out in zip(reqs, res):
        if out["status"] != 0:
            assert r[0] == "qaoa" and r[2] == 3, (r[4], out)  # only p = 3 exceeds the per-request work limit
            continue
        lat.setdefault(r[4], []).append(out["latency_ms"])
        if r[0] == "qaoa":
            ref, _ = qa.qaoa_energy_plain(ctx, r[1], r[2], False, list(r[3]), seed=3)
            assert abs(out["value"].real - ref) <= 1e-5 * abs(ref), (r[4], out, ref)
        elif r[3] <= 24:
            psi, _ = sv.simulate(r[1])
            ref = psi[_amp_index(r[2])]
            tol = 1e-10 if out["route"] == "eager" else 1e-5  # the batched routes run complex64 4M
            assert abs(out["value"] - ref) <= tol * max(abs(ref), 2.0 ** (-r[3] / 2)), (r[4], out, ref)
        else:
            served_48 += 1
            sn = qa.SlicedNetwork(r[1], r[2], dim_log2=1, order=qa.ORDER_SWEEP, seed=0, max_bits=29)
            ref, _ = sn.run(ctx, qa.QT_PLAIN_C128)
            sn.close()
            assert abs(out["value"] - ref) <= 1e-10 * abs(ref), (r[4], out, ref)
    assert served_48 >= 1
    p99 = {k: float(np.percentile(v, 99)) for k, v in lat.items()}
    assert all(v > 0 for v in p99.values()) and st["completed"] == sum(len(v) for v in lat.values())
    assert st["completed"] + st["rejected"] == len(reqs)
    ctx.close()
