"""CPU tests of the text round trip that device-side rotation blocks reproduce
(qt_angle_text_float, the host instantiation of the routine the device runs): the
reference writes an angle with `ostream <<` at the default precision
(src/maxcut.h:211,217) and parses it with std::stof (src/Network.h:342).  Checked
against the C library's own snprintf("%g") and strtof on random magnitudes across
the supported range, exact decimal ties, dyadic values and the edges."""
import ctypes
import math
import struct

import numpy as np
import pytest

import qtorch_amd as qa

_libc = ctypes.CDLL(None)
_libc.strtof.restype = ctypes.c_float
_libc.strtof.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
_libc.snprintf.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p, ctypes.c_double]


def _text_float(x: float) -> float:
    buf = ctypes.create_string_buffer(64)
    _libc.snprintf(buf, 64, b"%g", ctypes.c_double(x))
    return _libc.strtof(buf.value, None)


def _bits(f: float) -> bytes:
    return struct.pack("<f", f)


def _check(xs):
    bad = [x for x in xs if _bits(qa.angle_text_float(float(x))) != _bits(_text_float(float(x)))]
    assert not bad, [(x, "%g" % x, qa.angle_text_float(x), _text_float(x)) for x in bad[:5]]


def test_random_magnitudes():
    rng = np.random.default_rng(0)
    xs = np.sign(rng.standard_normal(100000)) * 10.0 ** rng.uniform(-16, 14.99, 100000)
    _check(xs)
    _check(rng.uniform(-2 * math.pi, 2 * math.pi, 100000))  # the QAOA range


def test_decimal_ties_and_dyadics():
    """Seven-digit values ending in 5 that are exact doubles (ties broken to even),
    their neighbours, and short dyadic fractions."""
    rng = np.random.default_rng(1)
    ties = (rng.integers(100000, 999999, 20000) * 10 + 5).astype(np.float64)
    xs = []
    for t in ties:
        for sc in (1.0, 0.25, 2.0 ** -10, 1e-3):
            x = t * sc
            xs += [x, np.nextafter(x, 0), np.nextafter(x, np.inf)]
    xs += [12345.25, 12345.75, 0.125, 999999.5, 9999995.0, 123456.5]
    xs += list(rng.integers(1, 1 << 20, 20000) / 2.0 ** rng.integers(0, 30, 20000))
    _check(xs)


def test_edges_and_range():
    _check([1e-16, 1.0000001e-16, 9.9999999e14, 1.0, 10.0, 100000.0, 999999.0, 1e6, 0.1, 0.3, math.pi, -math.pi])
    assert _bits(qa.angle_text_float(0.0)) == _bits(0.0)
    assert _bits(qa.angle_text_float(-0.0)) == _bits(-0.0)  # "-0"
    for x in (1e15, -3e20, 1e-17, float("inf"), float("nan")):
        with pytest.raises(qa.InvalidFunctionInput):
            qa.angle_text_float(x)
