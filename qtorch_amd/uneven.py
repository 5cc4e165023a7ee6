"""Uneven batching of many QAOA energies (F_p, src/maxcut.cpp:162-204): every edge
light cone of every graph, each graph with its own angles, FP64 energy per graph.

Light cones whose tensors fit a device program (line-graph min-fill order, every
tensor of at most 2^16 elements) run in batches of persistent table-executor launches, each
captured in a HIP graph; the others run on the sliced large-network path (MFMA
GEMMs for the large steps) on several streams at once.  Host planning runs on a
thread pool (the native calls release the interpreter lock).  This module only
orchestrates the native engine (libqtorch_host.so / libqtorch_hip.so)."""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from typing import List, Sequence, Tuple

import numpy as np

from . import _native as qa

# a program's tensors are limited to 2^16 elements: D = 4 wires, rank 8
PROGRAM_MAX_RANK = 8


def edge_networks(graphs: Sequence[Tuple[str, int, Sequence[float]]], ref_map: bool = False):
    """(graph, edge, qasm, measure) of every edge light cone of every (dgf, p, angles)."""
    out = []
    for g, (dgf, p, ang) in enumerate(graphs):
        ne, _ = qa.qaoa_num_edges(dgf)
        for e in range(ne):
            q, m = qa.qaoa_edge_text(dgf, p, ref_map, e, list(ang))
            out.append((g, e, q, m))
    return out


def _plan(qasm: str, measure: str):
    """(largest tensor rank, mCreatedFrom record) of the network in the line-graph
    min-fill order (plan only)."""
    net = qa.Network(qasm, measure, None)
    net.order(qa.ORDER_LINEGRAPH, 0)
    r = int(net.plan()[:, 6].max())
    rec = net.record()
    net.close()
    return r, rec


class UnevenEnergies:
    """Plans the networks once (split into batched and large ones), then run() contracts
    all of them and returns the FP64 energy of every graph."""

    def __init__(self, ctx: qa.Context, graphs, precision: int = qa.QT_C64_4M, ref_map: bool = False,
                 threads: int = 16, streams: int = 8, chunk: int = 4096, max_bits: int = 28,
                 program_max_rank: int = PROGRAM_MAX_RANK):
        t0 = time.perf_counter()
        self.ctx, self.precision, self.n_graphs = ctx, precision, len(graphs)
        self.nets = edge_networks(graphs, ref_map)
        with ThreadPoolExecutor(threads) as pool:
            plans = list(pool.map(lambda x: _plan(x[2], x[3]), self.nets, chunksize=64))
        self.small = [i for i, (r, _) in enumerate(plans) if r <= program_max_rank]
        self.big = [i for i, (r, _) in enumerate(plans) if r > program_max_rank]
        chunks = [self.small[k:k + chunk] for k in range(0, len(self.small), chunk)]

        def make_batch(ids):  # the recorded orders replayed (ContractGivenSequence)
            return qa.NetBatch(ctx, [self.nets[i][2] for i in ids], [self.nets[i][3] for i in ids],
                               precision=precision, order=qa.ORDER_GIVEN, given=[plans[i][1] for i in ids])

        def make_sliced(i):
            return qa.SlicedNetwork(self.nets[i][2], self.nets[i][3], dim_log2=2, order=qa.ORDER_LINEGRAPH,
                                    max_bits=max_bits)

        with ThreadPoolExecutor(threads) as pool:
            self.batches = list(pool.map(make_batch, chunks))
            sliced = list(pool.map(make_sliced, self.big))
        self.batch_ids = chunks
        self.sliced = dict(zip(self.big, sliced))
        self.streams = [qa.Context(getattr(ctx, "device", 0)) for _ in range(max(1, streams))]
        # the largest networks first, dealt round-robin to the streams
        order = sorted(self.big, key=lambda i: -self.sliced[i].plan()["cmacs_per_slice"])
        self.shards = [order[k::len(self.streams)] for k in range(len(self.streams))]
        self.planning_s = time.perf_counter() - t0
        self.configured = False

    def configure(self) -> Tuple[int, int]:
        """Executor shape from the autotuner on the largest batch, used by all of them;
        each batch launched once and captured in a HIP graph."""
        lanes, wgs = 8, 4
        if self.batches:
            lanes, wgs, _ = self.batches[0].autotune()
        for b in self.batches:
            b.config(lanes, wgs)
            b.launch()
            b.capture()
        for k, shard in enumerate(self.shards):  # warm each stream's allocator
            if shard:
                self.sliced[shard[0]].run(self.streams[k], self.precision)
        self.ctx.sync()
        self.configured = True
        return lanes, wgs

    def run(self):
        """(energies per graph, values per network, batched seconds, large-path seconds)."""
        if not self.configured:
            self.configure()
        values = np.zeros(len(self.nets), dtype=np.complex128)
        t0 = time.perf_counter()
        for b in self.batches:
            b.launch()
        self.ctx.sync()
        t_small = time.perf_counter() - t0
        for b, ids in zip(self.batches, self.batch_ids):
            values[ids] = b.results()

        def work(k):
            return [(i, self.sliced[i].run(self.streams[k], self.precision)[0]) for i in self.shards[k]]

        t0 = time.perf_counter()
        with ThreadPoolExecutor(len(self.streams)) as pool:
            for part in pool.map(work, range(len(self.streams))):
                for i, v in part:
                    values[i] = v
        t_big = time.perf_counter() - t0
        # F_p of every graph: its edges' 0.5 (1 - <ZZ>) added in edge order, in FP64
        energies = np.zeros(self.n_graphs)
        for i, (g, _, _, _) in enumerate(self.nets):
            energies[g] += 0.5 * (1.0 - values[i].real)
        return energies, values, t_small, t_big

    def close(self):
        for b in self.batches:
            b.close()
        for s in self.sliced.values():
            s.close()
        for c in self.streams:
            c.close()

