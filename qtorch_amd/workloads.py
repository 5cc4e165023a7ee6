"""Seeded synthetic workloads in qTorch's input formats (QASM text, measurement
strings, .dgf graphs; formats of Examples/howToWriteQASMfiles.txt and
src/maxcut.h:56-103)."""
from __future__ import annotations

import math
import random
from typing import List, Sequence, Tuple


def qft_qasm(n: int) -> str:
    """Textbook n-qubit QFT: H, controlled phases CPHASE(2 pi / 2^k) (k = 2..n-j, control
    j+k-1, target j), then the bit-reversal swaps.  On |0^n> every output string has
    probability 2^-n."""
    lines = [str(n)]
    for j in range(n):
        lines.append(f"H {j}")
        for k in range(2, n - j + 1):
            lines.append(f"CPHASE {2.0 * math.pi / 2 ** k:.17g} {j + k - 1} {j}")
    for j in range(n // 2):
        lines.append(f"SWAP {j} {n - 1 - j}")
    return "\n".join(lines) + "\n"


def bitstring(n: int, seed: int) -> str:
    rng = random.Random(seed)
    return "".join(rng.choice("01") for _ in range(n))


def measure_text(bits: str) -> str:
    return " ".join(bits)


def grid_random_circuit(rows: int, cols: int, depth: int, seed: int) -> str:
    """Random circuit on a rows x cols grid: every layer applies a seeded RX, RY or RZ
    (angle U[0, 2 pi)) to each qubit, then CZ on one of four grid-edge patterns
    (horizontal even/odd columns, vertical even/odd rows) in turn."""
    rng = random.Random(seed)
    n = rows * cols
    q = lambda r, c: r * cols + c  # noqa: E731
    patterns: List[List[Tuple[int, int]]] = [[], [], [], []]
    for r in range(rows):
        for c in range(cols - 1):
            patterns[c % 2].append((q(r, c), q(r, c + 1)))
    for r in range(rows - 1):
        for c in range(cols):
            patterns[2 + r % 2].append((q(r, c), q(r + 1, c)))
    lines = [str(n)]
    for layer in range(depth):
        for i in range(n):
            g = rng.choice(["RX", "RY", "RZ"])
            lines.append(f"{g} {rng.uniform(0, 2 * math.pi):.9g} {i}")
        for a, b in patterns[layer % 4]:
            lines.append(f"CZ {a} {b}")
    return "\n".join(lines) + "\n"


def regular_graph(n: int, d: int, seed: int) -> List[Tuple[int, int]]:
    """Seeded random d-regular simple graph (pairing model with restarts)."""
    if (n * d) % 2:
        raise ValueError("n * d must be even")
    rng = random.Random(seed)
    while True:
        stubs = [v for v in range(n) for _ in range(d)]
        rng.shuffle(stubs)
        edges = set()
        ok = True
        for i in range(0, len(stubs), 2):
            a, b = stubs[i], stubs[i + 1]
            if a == b or (min(a, b), max(a, b)) in edges:
                ok = False
                break
            edges.add((min(a, b), max(a, b)))
        if ok:
            e = sorted(edges)
            rng.shuffle(e)
            return [(a, b) if rng.random() < 0.5 else (b, a) for a, b in e]


def dgf_text(edges: Sequence[Tuple[int, int]], comment: str = "seeded random regular graph") -> str:
    return f"c {comment}\n" + "".join(f"e {a} {b}\n" for a, b in edges)


def qaoa_angles(p: int, seed: int) -> List[float]:
    """Reference angle vector [betas..., gammas_ref...] for exp(-i gamma ZZ) and RX(2 beta)
    with gamma ~ U[0, pi], beta ~ U[0, pi/2] (gamma_ref = -2 gamma, since the emitted
    CNOT-Rz(-gamma_ref)-CNOT is exp(+i gamma_ref/2 ZZ))."""
    rng = random.Random(seed)
    betas = [rng.uniform(0, math.pi / 2) for _ in range(p)]
    gammas = [-2.0 * rng.uniform(0, math.pi) for _ in range(p)]
    return betas + gammas
