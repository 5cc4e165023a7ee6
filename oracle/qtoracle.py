"""ORACLE — test infrastructure only.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker.  The product path
(``qtorch_amd``) never imports it.

A CPU restatement of qTorch's contraction path in Python (graph bookkeeping) over
the plain-C restatement of ``Network::ContractIndices`` in ``contract_ref.c``
(arithmetic).  Every function cites the reference file:line it follows
(paths relative to the qTorch repository).

Pinned by ``tests/test_oracle.py``: replayed contraction sequences recorded from
the reference compiled from its own sources (``oracle/_ref``) agree bit-for-bit
or to 1e-15; the reference's known-answer tests (``src/tests.cpp``); closed forms
(QFT, Bell, cat, single-qubit rotations); ``numpy.einsum`` for the index layout;
an independent state-vector simulation for whole circuits.
"""
from __future__ import annotations

import ctypes
import math
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

PI = 3.14159265358979323846  # src/Node.h:45


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        lib = ctypes.CDLL(path)
        lib.oracle_contract_indices.restype = ctypes.c_int
        lib.oracle_contract_indices.argtypes = [
            ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_int, ctypes.c_void_p]
        lib.oracle_stof.restype = ctypes.c_double
        lib.oracle_stof.argtypes = [ctypes.c_char_p]
        lib.oracle_stod.restype = ctypes.c_double
        lib.oracle_stod.argtypes = [ctypes.c_char_p]
        _LIB = lib
    return _LIB


def stof(text: str) -> float:
    """``std::stof`` then widening to double (src/Network.h:342)."""
    return _lib().oracle_stof(text.encode())


def stod(text: str) -> float:
    """``std::stod`` (src/Network.h:615)."""
    return _lib().oracle_stod(text.encode())


class InvalidFile(Exception):
    """src/Exceptions.h:24"""


class InvalidFileFormat(Exception):
    """src/Exceptions.h:28"""


class ContractionFailure(Exception):
    """src/Exceptions.h:36"""


class InvalidUserContractionSequence(Exception):
    """src/Exceptions.h:40"""


class InvalidFunctionInput(Exception):
    """src/Exceptions.h:44"""


# ---------------------------------------------------------------------------
# Pairwise contraction arithmetic (src/Network.h:876-971)
# ---------------------------------------------------------------------------

def contract_indices(a: np.ndarray, rank_a: int, b: np.ndarray, rank_b: int,
                     free: Sequence[Tuple[bool, int]], pairs: Sequence[Tuple[int, int]],
                     dim_log2: int = 2) -> np.ndarray:
    """C = sum over pairs of A*B with the reference's loop nest (src/Network.h:899-933).

    ``free`` is ``toNotSumOn`` (``(from_a, axis)`` in C's digit order) and
    ``pairs`` is ``toSumOn`` (``(axis_a, axis_b)``).
    """
    a = np.ascontiguousarray(a, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    if a.size == 0 or b.size == 0:
        raise InvalidFunctionInput()  # src/Network.h:938-940
    nfree, k = len(free), len(pairs)
    c = np.zeros(1 << (dim_log2 * nfree), dtype=np.complex128)
    ffa = (ctypes.c_int * max(nfree, 1))(*[1 if f[0] else 0 for f in free])
    fax = (ctypes.c_int * max(nfree, 1))(*[f[1] for f in free])
    pa = (ctypes.c_int * max(k, 1))(*[p[0] for p in pairs])
    pb = (ctypes.c_int * max(k, 1))(*[p[1] for p in pairs])
    rc = _lib().oracle_contract_indices(
        a.ctypes.data, rank_a, b.ctypes.data, rank_b, nfree, ffa, fax, k, pa, pb,
        dim_log2, c.ctypes.data)
    if rc != 0:
        raise InvalidFunctionInput()
    return c


# ---------------------------------------------------------------------------
# Node / Wire (src/Node.h:108-194, src/Wire.h:26-60)
# ---------------------------------------------------------------------------

INITSTATE, MEASURETRACE, INTERMEDIATE, GATE = "INITSTATE", "MEASURETRACE", "INTERMEDIATESTATE", "GATE"


class Wire:
    """src/Wire.h:26-60: two endpoints, the qubit number, a line-graph id."""
    __slots__ = ("node_a", "node_b", "qubit", "wire_id", "contracted")

    def __init__(self, node_a, node_b, qubit):
        self.node_a = node_a
        self.node_b = node_b
        self.qubit = qubit
        self.wire_id = -1
        self.contracted = False


class Node:
    """src/Node.h:108-166.  ``vals`` has 4**rank complex entries, flat index
    sum_i d_i 4**i over the wire order (src/Node.h:178-186)."""

    def __init__(self, rank: int, kind: str = INTERMEDIATE, name: str = "INTERMEDIATESTATE"):
        self.rank = rank
        self.vals: Optional[np.ndarray] = np.zeros(4 ** rank, dtype=np.complex128)
        self.wires: List[Wire] = []
        self.wire_numbers: List[int] = []
        self.id = -1
        self.contracted = False
        self.created_from = (0, 0)
        self.kind = kind
        self.name = name

    def set(self, digits: Sequence[int], value) -> None:
        """Node::Index(vector) (src/Node.h:178-186)."""
        idx = 0
        mult = 1
        for d in digits:
            idx += d * mult
            mult *= 4
        self.vals[idx] = value


# Gate tensors (src/Node.h:197-898).  Each constructor states the same entries as
# the reference's table, evaluated with the same libm calls (Python's math module
# calls the C library), so the values are bit-identical.

def cnot_node() -> Node:  # src/Node.h:197-220
    n = Node(4, GATE, "CNOT")
    for d in [(0, 0, 0, 0), (0, 1, 0, 1), (0, 2, 0, 2), (0, 3, 0, 3),
              (1, 0, 1, 1), (1, 1, 1, 0), (1, 2, 1, 3), (1, 3, 1, 2),
              (2, 0, 2, 2), (2, 1, 2, 3), (2, 2, 2, 0), (2, 3, 2, 1),
              (3, 0, 3, 3), (3, 1, 3, 2), (3, 2, 3, 1), (3, 3, 3, 0)]:
        n.set(d, 1.0)
    return n


def rx_node(t: float) -> Node:  # src/Node.h:222-244
    n = Node(2, GATE, "Rx")
    c2 = math.pow(math.cos(t / 2.0), 2)
    s2 = math.pow(math.sin(t / 2.0), 2)
    h = math.sin(t) / 2.0
    nh = -1.0 * (math.sin(t) / 2.0)
    n.set((0, 0), c2); n.set((0, 1), complex(0.0, h)); n.set((0, 2), complex(0.0, nh)); n.set((0, 3), s2)
    n.set((1, 0), complex(0.0, h)); n.set((1, 1), c2); n.set((1, 2), s2); n.set((1, 3), complex(0.0, nh))
    n.set((2, 0), complex(0.0, nh)); n.set((2, 1), s2); n.set((2, 2), c2); n.set((2, 3), complex(0.0, h))
    n.set((3, 0), s2); n.set((3, 1), complex(0.0, nh)); n.set((3, 2), complex(0.0, h)); n.set((3, 3), c2)
    return n


def ry_node(t: float) -> Node:  # src/Node.h:246-269
    n = Node(2, GATE, "Ry")
    c2 = math.pow(math.cos(t / 2.0), 2)
    s2 = math.pow(math.sin(t / 2.0), 2)
    h = math.sin(t) / 2.0
    nh = -1.0 * math.sin(t) / 2.0
    ns2 = -1.0 * math.pow(math.sin(t / 2), 2)
    n.set((0, 0), c2); n.set((0, 1), h); n.set((0, 2), h); n.set((0, 3), s2)
    n.set((1, 0), nh); n.set((1, 1), c2); n.set((1, 2), ns2); n.set((1, 3), h)
    n.set((2, 0), nh); n.set((2, 1), ns2); n.set((2, 2), c2); n.set((2, 3), h)
    n.set((3, 0), s2); n.set((3, 1), nh); n.set((3, 2), nh); n.set((3, 3), c2)
    return n


def _phase_diag(n: Node, t: float) -> Node:
    n.set((0, 0), 1.0)
    n.set((1, 1), complex(math.cos(t), -1.0 * math.sin(t)))
    n.set((2, 2), complex(math.cos(t), math.sin(t)))
    n.set((3, 3), 1.0)
    return n


def rz_node(t: float) -> Node:  # src/Node.h:271-281
    return _phase_diag(Node(2, GATE, "Rz"), t)


def phase_node(t: float) -> Node:  # src/Node.h:284-294
    return _phase_diag(Node(2, GATE, "Phase"), t)


def h_node() -> Node:  # src/Node.h:296-318
    n = Node(2, GATE, "H")
    signs = {(0, 0): 1, (0, 1): 1, (0, 2): 1, (0, 3): 1, (1, 0): 1, (1, 1): -1, (1, 2): 1, (1, 3): -1,
             (2, 0): 1, (2, 1): 1, (2, 2): -1, (2, 3): -1, (3, 0): 1, (3, 1): -1, (3, 2): -1, (3, 3): 1}
    for d, s in signs.items():
        n.set(d, s / 2.0)
    return n


def x_node() -> Node:  # src/Node.h:320-330
    n = Node(2, GATE, "X")
    for d in [(0, 3), (1, 2), (2, 1), (3, 0)]:
        n.set(d, 1.0)
    return n


def y_node() -> Node:  # src/Node.h:332-342
    n = Node(2, GATE, "Y")
    n.set((0, 3), 1.0); n.set((1, 2), -1.0); n.set((2, 1), -1.0); n.set((3, 0), 1.0)
    return n


def z_node() -> Node:  # src/Node.h:344-354
    n = Node(2, GATE, "Z")
    n.set((0, 0), 1.0); n.set((1, 1), -1.0); n.set((2, 2), -1.0); n.set((3, 3), 1.0)
    return n


def zero_state_node() -> Node:  # src/Node.h:356-363
    n = Node(1, INITSTATE, "|0><0|")
    n.set((0,), 1.0)
    return n


def measure_node(ch: str) -> Node:
    """Trace / X / Y / Z measure and the two projectors (src/Node.h:365-421),
    chosen as in AddMeasurementsOrTrace (src/Network.h:199-228)."""
    n = Node(1, MEASURETRACE, "measure")
    if ch == "X":
        n.set((1,), 1.0); n.set((2,), 1.0)
    elif ch == "Y":
        n.set((1,), complex(0, 1.0)); n.set((2,), complex(0, -1.0))
    elif ch == "Z":
        n.set((0,), 1.0); n.set((3,), -1.0)
    elif ch == "0":
        n.set((0,), 1.0)
    elif ch == "1":
        n.set((3,), 1.0)
    else:
        n.set((0,), 1.0); n.set((3,), 1.0)
    return n


def crk_node(control: int) -> Node:  # src/Node.h:424-447
    n = Node(4, GATE, "CRk")
    y = 2.0 * PI * -1.0 / math.pow(2, control + 1.0)
    em = complex(math.cos(y), math.sin(y))            # exp(2 pi (-i) / 2^(c+1))
    ep = complex(math.cos(-y), math.sin(-y))          # exp(2 pi (+i) / 2^(c+1))
    one = [(0, 0, 0, 0), (0, 1, 0, 1), (0, 2, 0, 2), (0, 3, 0, 3), (1, 0, 1, 0), (1, 2, 1, 2),
           (2, 0, 2, 0), (2, 1, 2, 1), (3, 0, 3, 0), (3, 3, 3, 3)]
    for d in one:
        n.set(d, 1.0)
    for d in [(1, 1, 1, 1), (1, 3, 1, 3), (3, 1, 3, 1)]:
        n.set(d, em)
    for d in [(2, 2, 2, 2), (2, 3, 2, 3), (3, 2, 3, 2)]:
        n.set(d, ep)
    return n


def cz_node() -> Node:  # src/Node.h:449-473
    n = Node(4, GATE, "CZ")
    for d in [(0, 0, 0, 0), (0, 1, 0, 1), (0, 2, 0, 2), (0, 3, 0, 3), (1, 0, 1, 0), (1, 2, 1, 2),
              (2, 0, 2, 0), (2, 1, 2, 1), (3, 0, 3, 0), (3, 3, 3, 3)]:
        n.set(d, 1.0)
    for d in [(1, 1, 1, 1), (1, 3, 1, 3), (2, 2, 2, 2), (2, 3, 2, 3), (3, 1, 3, 1), (3, 2, 3, 2)]:
        n.set(d, -1.0)
    return n


def cphase_node(t: float) -> Node:  # src/Node.h:475-498
    n = Node(4, GATE, "CPhase")
    em = complex(math.cos(t), -1.0 * math.sin(t))
    ep = complex(math.cos(t), math.sin(t))
    for d in [(0, 0, 0, 0), (0, 1, 0, 1), (0, 2, 0, 2), (0, 3, 0, 3), (1, 0, 1, 0), (1, 2, 1, 2),
              (2, 0, 2, 0), (2, 1, 2, 1), (3, 0, 3, 0), (3, 3, 3, 3)]:
        n.set(d, 1.0)
    for d in [(1, 1, 1, 1), (1, 3, 1, 3), (3, 1, 3, 1)]:
        n.set(d, em)
    for d in [(2, 2, 2, 2), (2, 3, 2, 3), (3, 2, 3, 2)]:
        n.set(d, ep)
    return n


def swap_node() -> Node:  # src/Node.h:515-542
    n = Node(4, GATE, "SWAP")
    for d in [(0, 0, 0, 0), (0, 1, 1, 0), (1, 0, 0, 1), (1, 1, 1, 1), (0, 2, 2, 0), (0, 3, 3, 0),
              (1, 2, 2, 1), (1, 3, 3, 1), (2, 0, 0, 2), (2, 1, 1, 2), (3, 0, 0, 3), (3, 1, 1, 3),
              (2, 2, 2, 2), (2, 3, 3, 2), (3, 2, 2, 3), (3, 3, 3, 3)]:
        n.set(d, 1.0)
    return n


def _read_complex_tokens(path: str, count: int) -> List[complex]:
    """``input >> std::complex<double>`` repeated (src/Node.h:563-569, 607-613):
    a token is ``(re,im)``, ``(re)`` or a bare real."""
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        raise InvalidFile()
    out: List[complex] = []
    i, n = 0, len(text)
    while len(out) < count:
        while i < n and text[i].isspace():
            i += 1
        if i >= n:
            raise InvalidFileFormat()
        if text[i] == "(":
            j = text.index(")", i)
            parts = text[i + 1:j].split(",")
            re_ = float(parts[0])
            im_ = float(parts[1]) if len(parts) > 1 else 0.0
            out.append(complex(re_, im_))
            i = j + 1
        else:
            j = i
            while j < n and not text[j].isspace():
                j += 1
            out.append(complex(float(text[i:j]), 0.0))
            i = j
    return out


def arbitrary_one_qubit_node(path: str, name: str) -> Node:
    """src/Node.h:544-586: entry (in, out) = U[ko,ki] * conj(U[bo,bi]), d = 2 ket + bra."""
    u = _read_complex_tokens(path, 4)
    n = Node(2, GATE, name)
    for di in range(4):
        ki, bi = di >> 1, di & 1
        for do in range(4):
            ko, bo = do >> 1, do & 1
            n.set((di, do), u[2 * ko + ki] * u[2 * bo + bi].conjugate())
    return n


def arbitrary_two_qubit_node(path: str, name: str) -> Node:
    """src/Node.h:588-898: 4x4 U row-major over basis 2*q1 + q2, wires
    [in q1, in q2, out q1, out q2]; entry = U[ket_out, ket_in] * conj(U[bra_out, bra_in])."""
    u = _read_complex_tokens(path, 16)
    n = Node(4, GATE, name)
    for d0 in range(4):
        for d1 in range(4):
            for d2 in range(4):
                for d3 in range(4):
                    ket_in = 2 * (d0 >> 1) + (d1 >> 1)
                    bra_in = 2 * (d0 & 1) + (d1 & 1)
                    ket_out = 2 * (d2 >> 1) + (d3 >> 1)
                    bra_out = 2 * (d2 & 1) + (d3 & 1)
                    n.set((d0, d1, d2, d3),
                          u[4 * ket_out + ket_in] * u[4 * bra_out + bra_in].conjugate())
    return n


# ---------------------------------------------------------------------------
# Network (src/Network.h:54-1282)
# ---------------------------------------------------------------------------

def parse_tokens(line: str) -> List[str]:
    """Network::ParseTokens (src/Network.h:976-992): '#' only at column 0,
    split on single spaces, an empty final piece dropped (std::regex split)."""
    if len(line) == 0 or line.find("#") == 0:
        return []
    parts = line.split(" ")
    if parts and parts[-1] == "":
        parts.pop()
    return parts


def _stoi(tok: str) -> int:
    s = tok.lstrip()
    j = 0
    if j < len(s) and s[j] in "+-":
        j += 1
    k = j
    while k < len(s) and s[k].isdigit():
        k += 1
    if k == j:
        raise InvalidFileFormat()
    return int(s[:k])


class Network:
    """qtorch::Network (src/Network.h:54-140)."""

    def __init__(self, qasm_text: str, measure_text: Optional[str], base_dir: str = "."):
        self.all_nodes: List[Node] = []
        self.uncontracted: List[Node] = []
        self.nodes_by_wire: List[List[Node]] = []
        self.final_val = complex(0.0, 0.0)
        self.done = False
        self.flops = 0
        self.base_dir = base_dir
        self._one_q: Dict[str, str] = {}
        self._two_q: Dict[str, str] = {}
        self._parse(qasm_text, measure_text)

    @classmethod
    def from_files(cls, qasm_path: str, measure_path: Optional[str]) -> "Network":
        try:
            with open(qasm_path) as f:
                q = f.read()
        except OSError:
            raise InvalidFile()
        m = None
        if measure_path is not None and os.path.exists(measure_path):
            with open(measure_path) as f:
                m = f.read()
        return cls(q, m)

    # ParseNetwork (src/Network.h:246-304)
    def _parse(self, qasm_text: str, measure_text: Optional[str]) -> None:
        lines = qasm_text.split("\n")
        self.nq = _stoi(lines[0])
        self.nodes_by_wire = [[] for _ in range(self.nq)]
        self._hanging: List[Wire] = []
        # CreateInitialStates (src/Network.h:177-194)
        for i in range(self.nq):
            n = zero_state_node()
            w = Wire(n, None, i)
            n.wires.append(w)
            self._hanging.append(w)
            self.all_nodes.append(n)
            n.wire_numbers.append(i)
            self.nodes_by_wire[i].append(n)
            n.id = len(self.all_nodes) - 1
        # getline loop: a final newline yields one more (empty) line
        for line in lines[1:]:
            self._parse_node(line)
        # measurement characters via `stream >> char` (src/Network.h:282-298)
        chars = [c for c in (measure_text or "") if not c.isspace()]
        meas = [chars[i] if i < len(chars) else "T" for i in range(self.nq)]
        # AddMeasurementsOrTrace (src/Network.h:199-240)
        for i in range(self.nq):
            n = measure_node(meas[i])
            self._hanging[i].node_b = n
            n.wires.append(self._hanging[i])
            self.all_nodes.append(n)
            n.wire_numbers.append(i)
            self.nodes_by_wire[i].append(n)
            n.id = len(self.all_nodes) - 1
        self.uncontracted = list(self.all_nodes)

    def _attach1(self, node: Node, q: int) -> None:
        if q > self.nq - 1:
            raise InvalidFileFormat()
        node.wires.append(self._hanging[q])
        self._hanging[q].node_b = node
        w = Wire(node, None, q)
        self._hanging[q] = w
        node.wires.append(w)
        node.wire_numbers.append(q)
        self.nodes_by_wire[q].append(node)

    def _attach2(self, node: Node, q1: int, q2: int) -> None:
        if q1 > self.nq - 1 or q2 > self.nq - 1 or q1 == q2:
            raise InvalidFileFormat()
        node.wires.append(self._hanging[q1])
        self._hanging[q1].node_b = node
        node.wires.append(self._hanging[q2])
        self._hanging[q2].node_b = node
        w1 = Wire(node, None, q1)
        self._hanging[q1] = w1
        node.wires.append(w1)
        w2 = Wire(node, None, q2)
        self._hanging[q2] = w2
        node.wires.append(w2)
        node.wire_numbers += [q1, q2]
        self.nodes_by_wire[q1].append(node)
        self.nodes_by_wire[q2].append(node)

    # ParseNode (src/Network.h:326-707)
    def _parse_node(self, line: str) -> None:
        t = parse_tokens(line)
        if not t:
            return
        g = t[0]
        if g in ("Rx", "RX", "Ry", "RY", "Rz", "RZ", "PHASE"):
            ang = stof(t[1])
            if g in ("Rx", "RX"):
                node = rx_node(ang)
            elif g in ("Ry", "RY"):
                node = ry_node(ang)
            elif g in ("Rz", "RZ"):
                node = rz_node(ang)
            else:
                node = phase_node(ang)
            self._attach1(node, _stoi(t[2]))
        elif g in ("H", "X", "Y", "Z"):
            node = {"H": h_node, "X": x_node, "Y": y_node, "Z": z_node}[g]()
            self._attach1(node, _stoi(t[1]))
        elif g in ("CNOT", "SWAP", "CZ"):
            node = {"CNOT": cnot_node, "SWAP": swap_node, "CZ": cz_node}[g]()
            self._attach2(node, _stoi(t[1]), _stoi(t[2]))
        elif g == "CRk":
            q1, q2 = _stoi(t[1]), _stoi(t[2])
            if q1 > self.nq - 1 or q2 > self.nq - 1 or q1 == q2:
                raise InvalidFileFormat()
            node = crk_node(q1)
            self._attach2(node, q1, q2)
        elif g == "CPHASE":
            node = cphase_node(stod(t[1]))
            self._attach2(node, _stoi(t[2]), _stoi(t[3]))
        elif g == "def1":
            self._one_q[t[1]] = t[2]
            return
        elif g == "def2":
            self._two_q[t[1]] = t[2]
            return
        elif g in self._one_q:
            node = arbitrary_one_qubit_node(self._path(self._one_q[g]), g)
            self._attach1(node, _stoi(t[1]))
        elif g in self._two_q:
            node = arbitrary_two_qubit_node(self._path(self._two_q[g]), g)
            self._attach2(node, _stoi(t[1]), _stoi(t[2]))
        else:
            raise InvalidFileFormat()
        self.all_nodes.append(node)
        node.id = len(self.all_nodes) - 1

    def _path(self, p: str) -> str:
        return p if os.path.isabs(p) else os.path.join(self.base_dir, p)

    # ContractNodes (src/Network.h:715-864)
    def contract_nodes(self, a: Node, b: Node, threshold: int) -> Optional[Node]:
        if a.contracted or b.contracted:
            return None
        ia: List[int] = []
        ib: List[int] = []
        ic: List[Tuple[bool, int]] = []
        connected: List[Wire] = []
        remaining: List[Wire] = []
        for i, w in enumerate(a.wires):
            if w.node_b is b or w.node_a is b:
                ia.append(i)
                connected.append(w)
            else:
                ic.append((True, i))
                remaining.append(w)
        for cw in connected:
            for kk, wb in enumerate(b.wires):
                if cw is wb:
                    ib.append(kk)
        for j, w in enumerate(b.wires):
            if not (w.node_b is a or w.node_a is a):
                ic.append((False, j))
                remaining.append(w)
        if (len(ia) == 0 and len(ic) > 0) or len(remaining) > max(a.rank, b.rank) + threshold:
            return None
        pairs = list(zip(ia, ib))
        c = Node(len(ic))
        for w in remaining:
            c.wires.append(w)
            if w.node_a is a or w.node_a is b:
                w.node_a = c
            elif w.node_b is a or w.node_b is b:
                w.node_b = c
        b.contracted = True
        a.contracted = True
        for w in connected:
            w.contracted = True
        self._contract_indices(ic, pairs, a, b, c)
        if self.done:
            t = Node(0)
            t.id = len(self.all_nodes)
            t.created_from = (a.id, b.id)
            self.all_nodes.append(t)
            self._remove(a)
            self._replace(b, c)
            a.vals = None
            b.vals = None
            return None
        c.id = len(self.all_nodes)
        c.created_from = (a.id, b.id)
        self.all_nodes.append(c)
        a.vals = None
        b.vals = None
        self._remove(a)
        self._replace(b, c)
        return c

    # ContractIndices (src/Network.h:876-971)
    def _contract_indices(self, free, pairs, a: Node, b: Node, c: Node) -> None:
        self.flops += 4 ** (len(free) + len(pairs))  # src/Network.h:884-885
        if a.vals is None or b.vals is None:
            raise InvalidFunctionInput()
        c.vals = contract_indices(a.vals, a.rank, b.vals, b.rank, free, pairs)
        if len(free) == 0:
            fv = self.final_val
            if (abs(fv.real) <= 1.0e-30 and abs(fv.imag) <= 1.0e-30) or (a.rank == 0 and b.rank == 0):
                self.final_val = complex(c.vals[0])
            if len(self.uncontracted) == 2:
                self.done = True

    def _remove(self, n: Node) -> None:  # FindAndRemove (src/Network.h:1269-1282)
        for i, v in enumerate(self.uncontracted):
            if v is n:
                del self.uncontracted[i]
                return

    def _replace(self, old: Node, new: Node) -> None:  # FindAndReplace 1-D (:1253-1265)
        for i, v in enumerate(self.uncontracted):
            if v is old:
                self.uncontracted[i] = new
                return

    @staticmethod
    def _replace_2d(table: List[List[Node]], old: Node, new: Node) -> None:
        """FindAndReplace 2-D (src/Network.h:1233-1249), including its quirk: after two
        replacements every later row is inspected at its first entry only."""
        count = 0
        for row in table:
            for i in range(len(row)):
                if row[i] is old:
                    row[i] = new
                    count += 1
                if count >= 2:
                    break

    # ReduceCircuit (src/Network.h:1015-1118)
    def reduce_circuit(self) -> None:
        nq = self.nq
        place: List[List[Node]] = [[] for _ in range(nq)]
        for i in range(nq):
            row = self.nodes_by_wire[i]
            idx = 0
            while idx < len(row):
                node = row[idx]
                if node.rank == 2:
                    prev = place[i][-1]
                    tmp = self.contract_nodes(prev, node, 0)
                    src = node.wires[0].node_a
                    tmp.wire_numbers.append(src.wire_numbers[0])
                    if src.rank > 2:
                        tmp.wire_numbers.append(src.wire_numbers[1])
                    if prev.kind == INITSTATE:
                        tmp.kind = INITSTATE
                    self._replace_2d(self.nodes_by_wire, prev, tmp)
                    self._replace_2d(place, prev, tmp)
                else:
                    place[i].append(node)
                idx += 1
        self.nodes_by_wire = place

        tmp_by_wire: List[List[Optional[Node]]] = [[] for _ in range(nq)]
        cur: List[Optional[Node]] = [None] * nq
        cidx = [0] * nq
        updated = [False] * nq
        flag = False
        found = False
        nbw = self.nodes_by_wire
        while not flag:
            for j in range(nq):
                if cidx[j] >= len(nbw[j]) or updated[j]:
                    continue
                t = nbw[j][cidx[j]]
                if t.rank == 1:
                    cur[j] = t
                    updated[j] = True
                    cidx[j] += 1
                    continue
                w0, w1 = t.wire_numbers[0], t.wire_numbers[1]
                if not updated[w1] and not updated[w0]:
                    if nbw[w0][cidx[w0]] is nbw[w1][cidx[w1]]:
                        if cur[w0] is cur[w1]:
                            cur[w0] = self.contract_nodes(t, cur[w0], 0)
                            cur[w1] = cur[w0]
                            cur[w0].wire_numbers.append(w0)
                            cur[w0].wire_numbers.append(w1)
                            found = True
                        else:
                            cur[w0] = t
                            cur[w1] = t
                        cidx[w0] += 1
                        cidx[w1] += 1
                        updated[w0] = True
                        updated[w1] = True
            for j in range(nq):
                if updated[j]:
                    if found:
                        tmp_by_wire[j][-1] = cur[j]
                    else:
                        tmp_by_wire[j].append(cur[j])
                elif cidx[j] < len(nbw[j]) and not found:
                    tmp_by_wire[j].append(None)
            if all(not u for u in updated):
                flag = True
            updated = [False] * nq
            found = False
        self.nodes_by_wire = tmp_by_wire

    # --- orderings that replay recorded or user-given sequences ---

    def contract_given_sequence(self, seq: Sequence[Tuple[int, int]]) -> complex:
        """ContractionTools::ContractGivenSequence (src/ContractionTools.h:392-411)."""
        for a, b in seq:
            self.contract_nodes(self.all_nodes[a], self.all_nodes[b], 100)
        if not self.done:
            raise ContractionFailure()
        return self.final_val

    def contract_user_sequence(self, seq: Sequence[Tuple[int, int]]) -> complex:
        """ContractUserDefinedSequenceOfWires (src/ContractionTools.h:154-214)."""
        n_all = len(self.all_nodes)
        for a, b in seq:
            if a < 0 or a > n_all or b < 0 or b > n_all:
                raise InvalidFileFormat()
        mp = list(range(n_all))
        for a, b in seq:
            if self.all_nodes[mp[a]].contracted or self.all_nodes[mp[b]].contracted:
                continue
            self.contract_nodes(self.all_nodes[mp[a]], self.all_nodes[mp[b]], 10000)
            mp[a] = len(self.all_nodes) - 1
            mp[b] = len(self.all_nodes) - 1
        if not self.done:
            raise InvalidUserContractionSequence()
        return self.final_val

    def line_graph_wires(self) -> Tuple[List[Wire], List[Tuple[Wire, Wire]]]:
        """LineGraph constructor (src/LineGraph.h:92-171): wire ids in uncontracted-node x
        wire order; one line-graph edge per wire pair of a node."""
        wires: List[Wire] = []
        edges: List[Tuple[Wire, Wire]] = []
        seen = set()
        for node in self.uncontracted:
            ws = list(node.wires)
            for wid, w in enumerate(ws):
                if id(w) not in seen:
                    w.wire_id = len(wires)
                    wires.append(w)
                    seen.add(id(w))
                for wid2 in range(wid):
                    edges.append((ws[wid2], w))
        return wires, edges

    def lg_contract(self, wires: List[Wire], order_1based: Sequence[int]) -> complex:
        """LineGraph::LGContract (src/LineGraph.h:357-399)."""
        for o in order_1based:
            w = wires[o - 1]
            if not w.contracted:
                self.contract_nodes(w.node_a, w.node_b, 100)
        if len(self.uncontracted) != 1 or self.uncontracted[0].vals is None \
                or self.uncontracted[0].vals.size != 1:
            raise ContractionFailure()
        return complex(self.uncontracted[0].vals[0])


def lg_cnf_text(wires: List[Wire], edges: List[Tuple[Wire, Wire]]) -> str:
    """CNF text written by LineGraph::runQuickBB (src/LineGraph.h:251-261)."""
    out = [f"p cnf {len(wires)} {len(edges)}\n"]
    for a, b in edges:
        out.append(f"{a.wire_id + 1} {b.wire_id + 1} 0\n")
    return "".join(out)


# ---------------------------------------------------------------------------
# QAOA MaxCut (src/maxcut.h:31-233, src/maxcut.cpp:162-204)
# ---------------------------------------------------------------------------

def read_dgf(text: str) -> List[Tuple[int, int]]:
    """ExtraData::ReadInData (src/maxcut.h:56-103): 'c' comment lines, 'e u v' edges."""
    pairs = []
    toks = text.split("\n")
    for line in toks:
        s = line.strip()
        if not s:
            continue
        if s[0] == "c":
            continue
        if s[0] != "e":
            continue
        _, u, v = s.split()[:3]
        pairs.append((int(u), int(v)))
    return pairs


def light_cones(pairs: Sequence[Tuple[int, int]], p: int, reference_qubit_map: bool = True):
    """ExtraData::PopulateIterations + Helper (src/maxcut.h:105-186).

    Returns, per edge, ``(qubits_needed, real_iterations, measured_pair)``.  With
    ``reference_qubit_map`` the reference's map is reproduced exactly: only
    ``pair.first -> 0`` is fixed (src/maxcut.h:119-120) and Z is measured on qubits
    0 and 1 (src/maxcut.cpp:179-188), whatever index BFS gave ``pair.second``.
    Otherwise ``pair.second -> 1`` is fixed as well, so the measured pair is the edge.
    """
    nq = max(max(u, v) for u, v in pairs) + 1
    adj: List[List[int]] = [[] for _ in range(nq)]
    for u, v in pairs:
        adj[u].append(v)
        adj[v].append(u)
    out = []
    for first, second in pairs:
        mp = [-1] * nq
        mp[first] = 0
        needed = 1
        if not reference_qubit_map:
            mp[second] = 1
            needed = 2
        iters: List[Tuple[int, int]] = []
        working = [first, second]
        checked = [False] * nq
        in_working = [False] * nq
        in_working[first] = True
        in_working[second] = True
        new_working = [False] * nq
        for _ in range(p):
            for vtx in working:
                for nb in adj[vtx]:
                    if not checked[nb]:
                        if mp[nb] == -1:
                            mp[nb] = needed
                            needed += 1
                        iters.append((vtx, nb))
                        if not in_working[nb]:
                            new_working[nb] = True
                checked[vtx] = True
            working = [i for i, f in enumerate(new_working) if f]
            in_working = list(new_working)
            new_working = [False] * nq
        real = [(mp[a], mp[b]) for a, b in iters]
        out.append((needed, real, (0, 1)))
    return out


def fmt_g(x: float) -> str:
    """``std::ostream << double`` at the default precision 6 (C's %g)."""
    return "%g" % x


def qaoa_edge_qasm(num_qubits: int, real_iters: Sequence[Tuple[int, int]], p: int,
                   betas_gammas: Sequence[float]) -> str:
    """QASM text F_p writes per edge (src/maxcut.cpp:172-178; src/maxcut.h:194-233)."""
    lines = [str(num_qubits)]
    for i in range(num_qubits):
        lines.append(f"H {i}")
    for i in range(p):
        gamma = betas_gammas[i + p]
        beta = betas_gammas[i]
        for a, b in real_iters:
            lines.append(f"CNOT {a} {b}")
            lines.append(f"Rz {fmt_g(-gamma)} {b}")
            lines.append(f"CNOT {a} {b}")
        for j in range(num_qubits):
            lines.append(f"Rx {fmt_g(beta * 2.0)} {j}")
    return "\n".join(lines) + "\n"


def qaoa_edge_measure(num_qubits: int) -> str:
    """Measure file F_p writes (src/maxcut.cpp:179-188): Z on qubits 0 and 1."""
    return "".join("Z " if i in (0, 1) else "T " for i in range(num_qubits))


def network_flops_and_value(net: Network) -> Tuple[complex, int]:
    return net.final_val, net.flops
