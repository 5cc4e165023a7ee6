"""Independent state-vector simulator used as a second oracle for whole circuits.

It shares no code with oracle/qtoracle.py: gates are built as 2x2 / 4x4 unitaries
from their textbook definitions, in the conventions the reference's superoperator
tables encode (src/Node.h:197-898, with d = 2*ket + bra):
  Rx(t) = exp(-i t X/2), Ry(t) = [[c,-s],[s,c]], Rz(t) = PHASE(t) = diag(e^{-it/2}, e^{it/2}),
  CPHASE(t) = diag(1,1,1,e^{it}), CRk c t = CPHASE(2 pi / 2^(c+1)), CZ, CNOT(control first), SWAP.
Angles of Rx/Ry/Rz/PHASE go through float32 (std::stof, src/Network.h:342).
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_I2 = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.diag([1, -1]).astype(np.complex128)
_H = np.array([[1, 1], [1, -1]], dtype=np.complex128) / math.sqrt(2.0)


_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]


def _f32(text: str) -> float:
    """C strtof (what std::stof calls), widened to double."""
    return float(_libc.strtof(text.encode(), None))


def _rx(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -1j * s], [-1j * s, c]], dtype=np.complex128)


def _ry(t):
    c, s = math.cos(t / 2), math.sin(t / 2)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def _rz(t):
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


def _read_complex(path, n):
    toks = open(path).read().split()
    out = []
    for t in toks:
        if t.startswith("("):
            re_, im_ = t.strip("()").split(",")
            out.append(complex(float(re_), float(im_)))
        else:
            out.append(complex(float(t), 0.0))
    return np.array(out[:n], dtype=np.complex128)


def apply_1q(psi, u, q, n):
    psi = psi.reshape([2] * n)  # axis n-1-q is qubit q (qubit q = bit q of the index)
    ax = n - 1 - q
    psi = np.moveaxis(np.tensordot(u, psi, axes=([1], [ax])), 0, ax)
    return psi.reshape(-1)


def apply_2q(psi, u, q1, q2, n):
    """u acts on basis 2*b(q1) + b(q2)."""
    psi = psi.reshape([2] * n)
    a1, a2 = n - 1 - q1, n - 1 - q2
    u4 = u.reshape(2, 2, 2, 2)  # [o1, o2, i1, i2]
    psi = np.tensordot(u4, psi, axes=([2, 3], [a1, a2]))
    psi = np.moveaxis(psi, [0, 1], [a1, a2])
    return psi.reshape(-1)


def simulate(qasm_text: str, base_dir: str = "."):
    lines = qasm_text.split("\n")
    n = int(lines[0].strip().split()[0])
    psi = np.zeros(2 ** n, dtype=np.complex128)
    psi[0] = 1.0
    one_q, two_q = {}, {}
    for line in lines[1:]:
        if not line or line.startswith("#"):
            continue
        t = [x for x in line.split(" ") if x != ""]
        g = t[0]
        if g in ("Rx", "RX"):
            psi = apply_1q(psi, _rx(_f32(t[1])), int(t[2]), n)
        elif g in ("Ry", "RY"):
            psi = apply_1q(psi, _ry(_f32(t[1])), int(t[2]), n)
        elif g in ("Rz", "RZ", "PHASE"):
            psi = apply_1q(psi, _rz(_f32(t[1])), int(t[2]), n)
        elif g in ("H", "X", "Y", "Z"):
            psi = apply_1q(psi, {"H": _H, "X": _X, "Y": _Y, "Z": _Z}[g], int(t[1]), n)
        elif g == "CNOT":
            u = np.eye(4, dtype=np.complex128)[[0, 1, 3, 2]]
            psi = apply_2q(psi, u, int(t[1]), int(t[2]), n)
        elif g == "SWAP":
            u = np.eye(4, dtype=np.complex128)[[0, 2, 1, 3]]
            psi = apply_2q(psi, u, int(t[1]), int(t[2]), n)
        elif g == "CZ":
            psi = apply_2q(psi, np.diag([1, 1, 1, -1]).astype(np.complex128), int(t[1]), int(t[2]), n)
        elif g == "CPHASE":
            th = float(t[1])
            psi = apply_2q(psi, np.diag([1, 1, 1, np.exp(1j * th)]), int(t[2]), int(t[3]), n)
        elif g == "CRk":
            c = int(t[1])
            th = 2 * math.pi / 2 ** (c + 1)
            psi = apply_2q(psi, np.diag([1, 1, 1, np.exp(1j * th)]), c, int(t[2]), n)
        elif g == "def1":
            one_q[t[1]] = os.path.join(base_dir, t[2])
        elif g == "def2":
            two_q[t[1]] = os.path.join(base_dir, t[2])
        elif g in one_q:
            psi = apply_1q(psi, _read_complex(one_q[g], 4).reshape(2, 2), int(t[1]), n)
        elif g in two_q:
            psi = apply_2q(psi, _read_complex(two_q[g], 16).reshape(4, 4), int(t[1]), int(t[2]), n)
        else:
            raise ValueError(g)
    return psi, n


def expectation(psi, n, measure_text: str) -> complex:
    """Tr(M rho) for the product of per-qubit measurement operators
    (X, Y, Z, 0 = |0><0|, 1 = |1><1|, anything else = identity)."""
    chars = [c for c in measure_text if not c.isspace()]
    ops = {"X": _X, "Y": _Y, "Z": _Z, "0": np.diag([1, 0]).astype(np.complex128),
           "1": np.diag([0, 1]).astype(np.complex128)}
    phi = psi.copy()
    for q in range(n):
        ch = chars[q] if q < len(chars) else "T"
        if ch in ops:
            phi = apply_1q(phi, ops[ch], q, n)
    return complex(np.vdot(psi, phi))


def _apply_1q_batch(psi, u, q, n):
    """psi: (S, 2^n); u: (S, 2, 2) or (2, 2)."""
    s = psi.shape[0]
    t = np.moveaxis(psi.reshape([s] + [2] * n), 1 + n - 1 - q, -1)
    shape = t.shape
    t = t.reshape(s, -1, 2)
    t = np.einsum("sab,skb->ska", u, t) if u.ndim == 3 else t @ u.T
    return np.moveaxis(t.reshape(shape), -1, 1 + n - 1 - q).reshape(s, -1)


def _apply_2q_batch(psi, u, q1, q2, n):
    s = psi.shape[0]
    t = psi.reshape([s] + [2] * n)
    t = np.moveaxis(t, [1 + n - 1 - q1, 1 + n - 1 - q2], [-2, -1])
    shape = t.shape
    t = t.reshape(s, -1, 4) @ u.T
    return np.moveaxis(t.reshape(shape), [-2, -1], [1 + n - 1 - q1, 1 + n - 1 - q2]).reshape(s, -1)


def simulate_batch(qasm_texts):
    """State vectors of circuits that share their gate sequence and differ only in
    the angles of Rx / Ry / Rz / PHASE gates (the edge light cones of one graph for
    many angle sets): (S, 2^n) and n.  Same conventions as simulate()."""
    split = [t.split("\n") for t in qasm_texts]
    n = int(split[0][0].strip().split()[0])
    s = len(split)
    psi = np.zeros((s, 2 ** n), dtype=np.complex128)
    psi[:, 0] = 1.0
    rot = {"Rx": _rx, "RX": _rx, "Ry": _ry, "RY": _ry, "Rz": _rz, "RZ": _rz, "PHASE": _rz}
    for li in range(1, len(split[0])):
        line = split[0][li]
        if not line or line.startswith("#"):
            continue
        t = [x for x in line.split(" ") if x != ""]
        g = t[0]
        if g in rot:
            us = np.empty((s, 2, 2), dtype=np.complex128)
            cache = {}
            for k in range(s):
                tok = split[k][li].split(" ")[1]
                if tok not in cache:
                    cache[tok] = rot[g](_f32(tok))
                us[k] = cache[tok]
            psi = _apply_1q_batch(psi, us, int(t[2]), n)
        elif g in ("H", "X", "Y", "Z"):
            psi = _apply_1q_batch(psi, {"H": _H, "X": _X, "Y": _Y, "Z": _Z}[g], int(t[1]), n)
        elif g == "CNOT":
            psi = _apply_2q_batch(psi, np.eye(4, dtype=np.complex128)[[0, 1, 3, 2]], int(t[1]), int(t[2]), n)
        elif g == "CZ":
            psi = _apply_2q_batch(psi, np.diag([1, 1, 1, -1]).astype(np.complex128), int(t[1]), int(t[2]), n)
        else:
            raise ValueError(g)
    return psi, n


def expectation_batch(psi, n, measure_text: str):
    """Tr(M rho) per row of a batch of state vectors (see expectation())."""
    chars = [c for c in measure_text if not c.isspace()]
    ops = {"X": _X, "Y": _Y, "Z": _Z, "0": np.diag([1, 0]).astype(np.complex128),
           "1": np.diag([0, 1]).astype(np.complex128)}
    phi = psi.copy()
    for q in range(n):
        ch = chars[q] if q < len(chars) else "T"
        if ch in ops:
            phi = _apply_1q_batch(phi, ops[ch], q, n)
    return np.einsum("sk,sk->s", psi.conj(), phi)
