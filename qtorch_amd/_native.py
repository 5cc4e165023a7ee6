"""ctypes bindings of libqtorch_hip.so and libqtorch_host.so (in-tree builds).

There is no CPU fallback: if a library is missing the import fails, and creating
a device context without a GPU raises.
"""
from __future__ import annotations

import ctypes
import os
import re
from typing import List, Optional, Sequence, Tuple

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_PKG, "lib")
INCLUDE_DIR = os.path.join(os.path.dirname(_PKG), "include")
HIP_LIB_PATH = os.path.join(LIB_DIR, "libqtorch_hip.so")
HOST_LIB_PATH = os.path.join(LIB_DIR, "libqtorch_host.so")

QT_OK, QT_INVALID_INPUT, QT_OOM, QT_DEVICE, QT_TIMEOUT, QT_UNSUPPORTED = range(6)
QT_DTYPE_C128, QT_DTYPE_C64 = 0, 1
QT_PLAIN_C128, QT_C64_4M, QT_C64_3M, QT_C64_BF16X3 = 0, 1, 2, 3
QT_SCRATCH_AUTO, QT_SCRATCH_LDS, QT_SCRATCH_GLOBAL, QT_SCRATCH_STAGE_LEAVES = 0, 1, 2, 4
QT_SPECIALISE_HERMITIAN = 1
QT_RESCALE = 1 << 8  # or-ed into a complex64 form: per-step power-of-two rescaling
QT_FROM_A = 1 << 30

_c = ctypes
_vp = _c.c_void_p
_i32p = _c.POINTER(_c.c_int32)
_i64p = _c.POINTER(_c.c_int64)
_dp = _c.POINTER(_c.c_double)


def _load(path: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python qtorch_amd/build.py`")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


hip = _load(HIP_LIB_PATH)
host = _load(HOST_LIB_PATH)

_HIP_SIGS = {
    "qt_ctx_create": (_c.c_int, [_c.c_int, _c.POINTER(_vp)]),
    "qt_ctx_destroy": (_c.c_int, [_vp]),
    "qt_sync": (_c.c_int, [_vp]),
    "qt_ctx_set_time_limit": (_c.c_int, [_vp, _c.c_double]),
    "qt_last_error": (_c.c_char_p, [_vp]),
    "qt_mem_reserved": (_c.c_int64, [_vp]),
    "qt_ctx_trim": (_c.c_int, [_vp]),
    "qt_batch_resident_bytes": (_c.c_int64, [_vp]),
    "qt_tensor_alloc": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_int, _c.POINTER(_vp)]),
    "qt_tensor_alloc_empty": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_int, _c.POINTER(_vp)]),
    "qt_tensor_upload": (_c.c_int, [_vp, _vp, _vp, _c.c_int64]),
    "qt_tensor_download": (_c.c_int, [_vp, _vp, _vp, _c.c_int64]),
    "qt_tensor_create_many": (_c.c_int, [_vp, _c.c_int, _i32p, _c.c_int, _c.c_int, _c.POINTER(_vp),
                                         _c.POINTER(_vp)]),
    "qt_tensor_free": (_c.c_int, [_vp, _vp]),
    "qt_tensor_numel": (_c.c_int64, [_vp]),
    "qt_tensor_rank": (_c.c_int, [_vp]),
    "qt_contract_pair": (_c.c_int, [_vp, _vp, _c.c_int, _vp, _c.c_int, _c.c_int, _i32p, _i32p, _c.c_int,
                                    _i32p, _vp, _c.c_int, _i64p]),
    "qt_contract_pair_noalloc": (_c.c_int, [_vp, _vp, _c.c_int, _vp, _c.c_int, _c.c_int, _i32p, _i32p, _c.c_int,
                                            _i32p, _vp, _c.c_int, _i64p]),
    "qt_tensor_read_scalar": (_c.c_int, [_vp, _vp, _dp, _dp]),
    "qt_tensor_accumulate": (_c.c_int, [_vp, _vp, _vp]),
    "qt_tensor_exponent": (_c.c_int, [_vp, _vp, _c.POINTER(_c.c_int32)]),
    "qt_batch_add_programs": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_vp), _i64p, _i32p]),
    "qt_batch_write_leaves": (_c.c_int, [_vp, _c.c_int64, _dp, _c.c_int64]),
    "qt_batch_poll": (_c.c_int, [_vp, _c.POINTER(_c.c_int)]),
    "qt_batch_set_sums_only": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_c.c_int)]),
    "qt_batch_rotation_params": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp, _c.c_int, _c.c_int, _vp, _c.c_int,
                                            _c.c_int64, _c.c_int64]),
    "qt_angle_text_float": (_c.c_int, [_c.c_double, _c.POINTER(_c.c_float)]),
    "qt_program_create": (_c.c_int, [_vp, _c.c_int, _c.c_int, _i32p, _i64p, _c.c_int, _i32p, _i32p,
                                     _c.POINTER(_vp)]),
    "qt_program_stats": (_c.c_int, [_vp, _i64p, _i64p, _c.POINTER(_c.c_int)]),
    "qt_program_destroy": (_c.c_int, [_vp]),
    "qt_batch_create": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_vp), _c.c_int, _c.POINTER(_vp)]),
    "qt_batch_set_leaves": (_c.c_int, [_vp, _dp, _c.c_int64]),
    "qt_batch_set_instances": (_c.c_int, [_vp, _c.c_int, _i32p, _i64p, _c.c_int, _i32p, _c.c_double,
                                          _c.c_double]),
    "qt_batch_set_outputs": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_vp)]),
    "qt_batch_compile": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_c.c_float)]),
    "qt_batch_time_kernel": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_c.c_float)]),
    "qt_batch_uncapture": (_c.c_int, [_vp]),
    "qt_batch_flush": (_c.c_int, [_vp]),
    "qt_program_compile": (_c.c_int, [_c.c_char_p, _c.c_char_p, _vp, _c.c_size_t, _c.POINTER(_c.c_size_t)]),
    "qt_batch_specialised_info": (_c.c_int, [_vp, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int64)]),
    "qt_program_source": (_c.c_int, [_c.c_int, _c.POINTER(_vp), _i64p, _dp, _c.c_int64, _c.c_int, _c.c_int, _c.c_char_p,
                                     _c.c_size_t, _c.POINTER(_c.c_size_t)]),
    "qt_batch_config": (_c.c_int, [_vp, _c.c_int, _c.c_int]),
    "qt_batch_config_ex": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_int]),
    "qt_batch_autotune": (_c.c_int, [_vp, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int), _c.POINTER(_c.c_float)]),
    "qt_batch_launch": (_c.c_int, [_vp]),
    "qt_batch_capture": (_c.c_int, [_vp]),
    "qt_batch_results": (_c.c_int, [_vp, _dp, _dp]),
    "qt_batch_time": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_c.c_float)]),
    "qt_batch_info": (_c.c_int, [_vp, _i64p, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int), _c.POINTER(_c.c_int),
                                 _c.POINTER(_c.c_int)]),
    "qt_batch_destroy": (_c.c_int, [_vp]),
}
_HOST_SIGS = {
    "qth_last_error": (_c.c_char_p, []),
    "qth_set_max_time": (_c.c_int, [_c.c_double]),
    "qth_get_max_time": (_c.c_double, []),
    "qth_timer_start": (_c.c_int, []),
    "qth_timer_reset": (_c.c_int, []),
    "qth_timer_elapsed": (_c.c_double, []),
    "qth_preprocess": (_c.c_int, [_vp, _c.c_char_p, _c.c_char_p, _c.c_double, _c.c_uint64, _i32p, _c.c_int,
                                  _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "qth_network_create": (_c.c_int, [_vp, _c.c_char_p, _c.c_char_p, _c.c_char_p, _c.POINTER(_vp)]),
    "qth_network_create_ex": (_c.c_int, [_vp, _c.c_char_p, _c.c_char_p, _c.c_char_p, _c.c_int, _c.POINTER(_vp)]),
    "qth_network_create_open": (_c.c_int, [_vp, _c.c_char_p, _c.c_char_p, _c.c_char_p, _c.c_int, _i32p, _c.c_int,
                                           _c.POINTER(_vp)]),
    "qth_network_result_tensor": (_c.c_int, [_vp, _dp, _c.c_int64, _i32p, _c.POINTER(_c.c_int)]),
    "qth_network_destroy": (_c.c_int, [_vp]),
    "qth_network_reduce": (_c.c_int, [_vp]),
    "qth_contract_nodes": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.c_int, _c.POINTER(_c.c_int)]),
    "qth_contract_stochastic": (_c.c_int, [_vp, _c.c_uint64, _dp, _dp]),
    "qth_contract_given": (_c.c_int, [_vp, _i32p, _c.c_int, _dp, _dp]),
    "qth_contract_user": (_c.c_int, [_vp, _i32p, _c.c_int, _dp, _dp]),
    "qth_linegraph_cnf": (_c.c_int, [_vp, _c.c_char_p, _c.c_int64, _i64p]),
    "qth_linegraph_minfill": (_c.c_int, [_vp, _i32p, _c.c_int, _c.POINTER(_c.c_int)]),
    "qth_linegraph_contract": (_c.c_int, [_vp, _i32p, _c.c_int, _dp, _dp]),
    "qth_network_order": (_c.c_int, [_vp, _c.c_int, _c.c_uint64]),
    "qth_record": (_c.c_int, [_vp, _i32p, _c.c_int, _c.POINTER(_c.c_int)]),
    "qth_flops": (_c.c_int64, [_vp]),
    "qth_network_value": (_c.c_int, [_vp, _dp, _dp]),
    "qth_counts": (_c.c_int, [_vp] + [_c.POINTER(_c.c_int)] * 4),
    "qth_plan": (_c.c_int, [_vp, _i32p, _c.c_int, _c.POINTER(_c.c_int)]),
    "qth_program_stats": (_c.c_int, [_vp, _c.POINTER(_c.c_int64), _c.POINTER(_c.c_int64), _c.POINTER(_c.c_int)]),
    "qth_qaoa_num_edges": (_c.c_int, [_c.c_char_p, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int)]),
    "qth_qaoa_edge_text": (_c.c_int, [_c.c_char_p, _c.c_int, _c.c_int, _c.c_int, _dp, _c.c_char_p, _c.c_int64,
                                      _c.c_char_p, _c.c_int64]),
    "qth_qaoa_energy_plain": (_c.c_int, [_vp, _c.c_char_p, _c.c_int, _c.c_int, _dp, _c.c_uint64, _dp, _dp]),
    "qth_qaoa_batch_create": (_c.c_int, [_vp, _c.c_char_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_uint64,
                                         _c.POINTER(_vp)]),
    "qth_qaoa_batch_set_angles": (_c.c_int, [_vp, _c.c_int, _dp]),
    "qth_qaoa_batch_results": (_c.c_int, [_vp, _dp, _dp]),
    "qth_qaoa_batch_update_angles": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_int]),
    "qth_qaoa_batch_device": (_vp, [_vp]),
    "qth_qaoa_batch_cmacs": (_c.c_int64, [_vp]),
    "qth_qaoa_batch_destroy": (_c.c_int, [_vp]),
    "qth_netbatch_create": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_c.c_char_p), _c.POINTER(_c.c_char_p),
                                       _c.c_char_p, _c.c_int, _c.c_int, _c.c_uint64, _i32p, _i32p, _c.c_int,
                                       _c.POINTER(_vp)]),
    "qth_qaoa_optimize": (_c.c_int, [_vp, _c.c_char_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_uint64, _dp,
                                     _dp, _c.POINTER(_c.c_int)]),
    "qth_maxcut_final_string": (_c.c_int, [_vp, _c.c_char_p, _c.c_int, _dp, _c.c_uint64, _c.POINTER(_c.c_int),
                                           _c.POINTER(_c.c_int)]),
    "qth_maxcut_final_string_ex": (_c.c_int, [_vp, _c.c_char_p, _c.c_int, _dp, _c.c_uint64, _c.POINTER(_c.c_int),
                                              _c.POINTER(_c.c_int), _dp]),
    "qth_netbatch_device": (_vp, [_vp]),
    "qth_netbatch_results": (_c.c_int, [_vp, _dp]),
    "qth_netbatch_cmacs": (_c.c_int, [_vp, _i64p]),
    "qth_netbatch_destroy": (_c.c_int, [_vp]),
    "qth_server_create": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.c_int64, _c.c_int, _c.POINTER(_vp)]),
    "qth_server_create_ex": (_c.c_int, [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int64, _c.c_int,
                                        _c.POINTER(_vp)]),
    "qth_server_destroy": (_c.c_int, [_vp]),
    "qth_server_submit_qaoa": (_c.c_int, [_vp, _c.c_char_p, _c.c_int, _c.c_int, _c.POINTER(_c.c_double),
                                          _c.POINTER(_c.c_int64)]),
    "qth_server_submit_amplitude": (_c.c_int, [_vp, _c.c_char_p, _c.c_char_p, _c.POINTER(_c.c_int64)]),
    "qth_server_poll": (_c.c_int, [_vp, _c.c_int64, _c.POINTER(_c.c_int), _c.POINTER(_c.c_int),
                                   _c.POINTER(_c.c_double), _c.POINTER(_c.c_double), _c.POINTER(_c.c_double),
                                   _c.POINTER(_c.c_int)]),
    "qth_server_wait": (_c.c_int, [_vp, _c.c_int64, _c.POINTER(_c.c_double), _c.POINTER(_c.c_double),
                                   _c.POINTER(_c.c_double), _c.POINTER(_c.c_int)]),
    "qth_server_drain": (_c.c_int, [_vp]),
    "qth_server_stats": (_c.c_int, [_vp, _c.POINTER(_c.c_int64), _c.POINTER(_c.c_double), _c.POINTER(_c.c_int64)]),
    "qth_sliced_create": (_c.c_int, [_c.c_char_p, _c.c_char_p, _c.c_int, _c.c_int, _c.c_uint64, _c.c_int,
                                     _c.POINTER(_vp)]),
    "qth_sliced_plan": (_c.c_int, [_vp, _c.POINTER(_c.c_int64)]),
    "qth_sliced_run": (_c.c_int, [_vp, _vp, _c.c_int, _c.POINTER(_c.c_double), _c.POINTER(_c.c_double),
                                  _c.POINTER(_c.c_double)]),
    "qth_sliced_destroy": (_c.c_int, [_vp]),
    "qth_sliced_set_fusion": (_c.c_int, [_vp, _c.c_int64]),
    "qth_sliced_set_range": (_c.c_int, [_vp, _c.c_int64, _c.c_int64]),
    "qth_sliced_set_slices": (_c.c_int, [_vp, _c.POINTER(_c.c_int64), _c.c_int64]),
    "qth_sliced_nonzero_slices": (_c.c_int, [_vp, _c.c_int64, _c.POINTER(_c.c_int64), _c.POINTER(_c.c_int64)]),
    "qth_sliced_nonzero_slices_ex": (_c.c_int, [_vp, _c.c_int64, _c.c_int64, _c.POINTER(_c.c_int64),
                                                _c.POINTER(_c.c_int64), _c.POINTER(_c.c_int)]),
    "qth_server_memory": (_c.c_int, [_vp, _c.POINTER(_c.c_int64), _c.POINTER(_c.c_int64)]),
    "qth_qaoa_source": (_c.c_int, [_c.c_char_p, _c.c_int, _c.c_int, _c.c_int, _c.c_uint64, _c.c_int, _c.c_int,
                                   _c.c_int, _c.c_int, _c.c_char_p, _c.c_size_t, _c.POINTER(_c.c_size_t)]),
    "qth_qaoa_param_block": (_c.c_int, [_c.c_int, _dp, _dp]),
    "qth_serve_plan": (_c.c_int, [_c.c_int, _c.c_char_p, _c.c_int, _c.c_int, _c.c_char_p, _c.c_int,
                                  _c.POINTER(_c.c_int), _c.POINTER(_c.c_int64), _c.POINTER(_c.c_int64)]),
}

ORDER_GIVEN, ORDER_GREEDY, ORDER_STOCHASTIC, ORDER_LINEGRAPH, ORDER_SWEEP = 0, 1, 2, 3, 4


def _missing(name: str):
    def call(*_args, **_kw):
        raise ImportError(f"{name} is not exported by the loaded library: rebuild with `python qtorch_amd/build.py`")
    return call


def _bind(lib, sigs):
    for name, (res, args) in sigs.items():
        try:
            f = getattr(lib, name)
        except AttributeError:
            setattr(lib, name, _missing(name))
            continue
        f.restype = res
        f.argtypes = args


_bind(hip, _HIP_SIGS)
_bind(host, _HOST_SIGS)


def declared_symbols(header: str) -> List[str]:
    """Function names declared in include/<header>."""
    with open(os.path.join(INCLUDE_DIR, header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qth?_[a-z0-9_]+)\s*\(", text)))


class EngineError(RuntimeError):
    pass


# exception classes matching the reference's (src/Exceptions.h)
class InvalidFunctionInput(EngineError):
    pass


class InvalidFile(EngineError):
    pass


class InvalidFileFormat(EngineError):
    pass


class ContractionFailure(EngineError):
    pass


class InvalidUserContractionSequence(EngineError):
    pass


class InvalidContractionMethod(EngineError):
    pass


class QbbFailure(EngineError):
    pass


class DeviceFailure(EngineError):
    pass


_HOST_ERRORS = {1: InvalidFunctionInput, 2: InvalidFile, 3: InvalidFileFormat, 4: ContractionFailure,
                5: InvalidUserContractionSequence, 6: InvalidContractionMethod, 7: QbbFailure, 8: DeviceFailure,
                9: EngineError}


def _host_check(rc: int) -> None:
    if rc != 0:
        msg = host.qth_last_error().decode(errors="replace")
        raise _HOST_ERRORS.get(rc, EngineError)(msg)


class Context:
    """A device context: one HIP stream and a caching allocator (qt_ctx)."""

    def __init__(self, device: int = 0):
        h = _vp()
        rc = hip.qt_ctx_create(device, ctypes.byref(h))
        if rc != QT_OK:
            raise DeviceFailure(f"qt_ctx_create(device={device}) failed with status {rc}: no usable HIP device")
        self.handle = h
        self.device = device

    def check(self, rc: int, what: str = "") -> None:
        if rc == QT_OK:
            return
        msg = hip.qt_last_error(self.handle).decode(errors="replace")
        if rc == QT_INVALID_INPUT:
            raise InvalidFunctionInput(f"{what}: {msg}")
        if rc == QT_TIMEOUT:
            raise ContractionFailure(f"{what}: {msg}")
        raise DeviceFailure(f"{what}: status {rc}: {msg}")

    def sync(self) -> None:
        self.check(hip.qt_sync(self.handle), "qt_sync")

    def set_time_limit(self, seconds: float) -> None:
        """Work requested more than `seconds` from now fails with ContractionFailure
        (QT_TIMEOUT); seconds <= 0 removes the limit."""
        self.check(hip.qt_ctx_set_time_limit(self.handle, float(seconds)), "qt_ctx_set_time_limit")

    def mem_reserved(self) -> int:
        return int(hip.qt_mem_reserved(self.handle))

    def close(self) -> None:
        """Destroys the context unless tensors or batches are still alive on it (then
        it stays until they are gone and the object itself is collected)."""
        if self.handle and hip.qt_ctx_destroy(self.handle) == QT_OK:
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tensor:
    """A device tensor (qt_tensor): D^rank interleaved complex values."""

    def __init__(self, ctx: Context, rank: int, dim_log2: int = 2, dtype: int = QT_DTYPE_C128):
        self.ctx = ctx
        self.rank = rank
        self.dim_log2 = dim_log2
        self.dtype = dtype
        h = _vp()
        ctx.check(hip.qt_tensor_alloc(ctx.handle, rank, dim_log2, dtype, ctypes.byref(h)), "qt_tensor_alloc")
        self.handle = h

    @property
    def np_dtype(self):
        return np.complex128 if self.dtype == QT_DTYPE_C128 else np.complex64

    @property
    def numel(self) -> int:
        return 1 << (self.rank * self.dim_log2)

    def upload(self, values: np.ndarray) -> "Tensor":
        v = np.ascontiguousarray(values, dtype=self.np_dtype).reshape(-1)
        self.ctx.check(hip.qt_tensor_upload(self.ctx.handle, self.handle, v.ctypes.data, v.nbytes), "upload")
        return self

    def download(self) -> np.ndarray:
        v = np.empty(self.numel, dtype=self.np_dtype)
        self.ctx.check(hip.qt_tensor_download(self.ctx.handle, self.handle, v.ctypes.data, v.nbytes), "download")
        return v

    def exponent(self) -> int:
        """Power-of-two exponent of a rescaled complex64 tensor (values = data * 2^e)."""
        e = ctypes.c_int32()
        self.ctx.check(hip.qt_tensor_exponent(self.ctx.handle, self.handle, ctypes.byref(e)), "qt_tensor_exponent")
        return e.value

    def scalar(self) -> complex:
        """Value of a rank-0 tensor (exponent applied in FP64)."""
        re, im = ctypes.c_double(), ctypes.c_double()
        self.ctx.check(hip.qt_tensor_read_scalar(self.ctx.handle, self.handle, ctypes.byref(re), ctypes.byref(im)),
                       "qt_tensor_read_scalar")
        return complex(re.value, im.value)

    def accumulate_into(self, acc: "Tensor") -> None:
        """acc += self (rank-0 tensors, acc complex128), in FP64 on the device."""
        self.ctx.check(hip.qt_tensor_accumulate(self.ctx.handle, self.handle, acc.handle), "qt_tensor_accumulate")

    def free(self) -> None:
        if self.handle:
            hip.qt_tensor_free(self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                self.free()
        except Exception:
            pass


def _i32(vals: Sequence[int]):
    arr = (ctypes.c_int32 * max(len(vals), 1))(*vals)
    return arr


def contract_pair(ctx: Context, a: Tensor, b: Tensor, pairs: Sequence[Tuple[int, int]],
                  free: Sequence[Tuple[bool, int]], precision: int = QT_PLAIN_C128) -> Tuple[Tensor, int]:
    """qt_contract_pair: C = contraction of A and B (Network::ContractIndices semantics)."""
    c = Tensor(ctx, len(free), a.dim_log2, a.dtype)
    pa = _i32([p[0] for p in pairs])
    pb = _i32([p[1] for p in pairs])
    fs = _i32([(ax | QT_FROM_A) if from_a else ax for from_a, ax in free])
    macs = ctypes.c_int64(0)
    ctx.check(hip.qt_contract_pair(ctx.handle, a.handle, a.rank, b.handle, b.rank, len(pairs), pa, pb,
                                   len(free), fs, c.handle, precision, ctypes.byref(macs)), "qt_contract_pair")
    return c, macs.value


class Network:
    """qTorch Network built by the C++ host library (libqtorch_host.so).

    ``ctx=None`` builds a plan-only network (ordering logic and step records, no
    values)."""

    def __init__(self, qasm: str, measure: Optional[str], ctx: Optional[Context] = None, base_dir: str = "",
                 dim_log2: int = 2, open_qubits: Optional[Sequence[int]] = None):
        h = _vp()
        m = measure.encode() if measure is not None else None
        if open_qubits:
            oq = list(open_qubits)
            _host_check(host.qth_network_create_open(ctx.handle if ctx else None, qasm.encode(), m,
                                                     base_dir.encode(), dim_log2, _i32(oq), len(oq), ctypes.byref(h)))
        else:
            _host_check(host.qth_network_create_ex(ctx.handle if ctx else None, qasm.encode(), m,
                                                   base_dir.encode(), dim_log2, ctypes.byref(h)))
        self.handle = h
        self.ctx = ctx
        self.dim_log2 = dim_log2

    def result_tensor(self) -> Tuple[np.ndarray, List[int]]:
        """(values, axis qubits) of the tensor a finished network with open qubits ends
        in (a scalar and [] for a measured network)."""
        rank = ctypes.c_int(0)
        axes = (ctypes.c_int32 * 64)()
        _host_check(host.qth_network_result_tensor(self.handle, None, 0, axes, ctypes.byref(rank)))
        n = 1 << (self.dim_log2 * rank.value)
        out = np.zeros(2 * n, dtype=np.float64)
        _host_check(host.qth_network_result_tensor(self.handle, out.ctypes.data_as(_dp), n, axes, ctypes.byref(rank)))
        return out.view(np.complex128), list(axes[:rank.value])

    def probabilities(self) -> np.ndarray:
        """P(x) for every output string x of the open qubits (x's bit j = the j-th open
        qubit in increasing qubit order): the diagonal of the output density matrix
        (digits 3 x_q, Liouville picture) or |amplitude|^2 (amplitude picture)."""
        vals, axq = self.result_tensor()
        qs = sorted(axq)
        n = len(qs)
        x = np.arange(1 << n)
        idx = np.zeros(1 << n, dtype=np.int64)
        per = 3 if self.dim_log2 == 2 else 1
        for axis, q in enumerate(axq):
            bit = (x >> qs.index(q)) & 1
            idx += per * bit << (self.dim_log2 * axis)
        v = vals[idx]
        return v.real if self.dim_log2 == 2 else np.abs(v) ** 2

    def close(self):
        if self.handle:
            host.qth_network_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reduce(self) -> None:
        _host_check(host.qth_network_reduce(self.handle))

    def contract_nodes(self, a: int, b: int, threshold: int) -> int:
        r = ctypes.c_int(-1)
        _host_check(host.qth_contract_nodes(self.handle, a, b, threshold, ctypes.byref(r)))
        return r.value

    def _val(self, fn, *args) -> complex:
        re, im = ctypes.c_double(), ctypes.c_double()
        _host_check(fn(self.handle, *args, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def contract_stochastic(self, seed: int) -> complex:
        return self._val(host.qth_contract_stochastic, ctypes.c_uint64(seed))

    def contract_given(self, seq: Sequence[Tuple[int, int]]) -> complex:
        flat = [x for p in seq for x in p]
        return self._val(host.qth_contract_given, _i32(flat), len(seq))

    def contract_user(self, seq: Sequence[Tuple[int, int]]) -> complex:
        flat = [x for p in seq for x in p]
        return self._val(host.qth_contract_user, _i32(flat), len(seq))

    def order(self, order: int, seed: int = 0) -> None:
        _host_check(host.qth_network_order(self.handle, order, ctypes.c_uint64(seed)))

    def linegraph_cnf(self) -> str:
        n = ctypes.c_int64(0)
        _host_check(host.qth_linegraph_cnf(self.handle, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value + 1)
        _host_check(host.qth_linegraph_cnf(self.handle, buf, n.value + 1, ctypes.byref(n)))
        return buf.value.decode()

    def linegraph_minfill(self) -> List[int]:
        n = ctypes.c_int(0)
        _host_check(host.qth_linegraph_minfill(self.handle, None, 0, ctypes.byref(n)))
        out = (ctypes.c_int32 * max(n.value, 1))()
        _host_check(host.qth_linegraph_minfill(self.handle, out, n.value, ctypes.byref(n)))
        return list(out[:n.value])

    def linegraph_contract(self, order: Sequence[int]) -> complex:
        return self._val(host.qth_linegraph_contract, _i32(list(order)), len(order))

    def record(self) -> List[Tuple[int, int]]:
        n = ctypes.c_int(0)
        _host_check(host.qth_record(self.handle, None, 0, ctypes.byref(n)))
        out = (ctypes.c_int32 * max(2 * n.value, 1))()
        _host_check(host.qth_record(self.handle, out, n.value, ctypes.byref(n)))
        return [(out[2 * i], out[2 * i + 1]) for i in range(n.value)]

    def value(self) -> complex:
        """Final value of a contracted network (mFinalVal)."""
        return self._val(host.qth_network_value)

    def flops(self) -> int:
        return int(host.qth_flops(self.handle))

    def counts(self):
        v = [ctypes.c_int() for _ in range(4)]
        _host_check(host.qth_counts(self.handle, *[ctypes.byref(x) for x in v]))
        return {"all": v[0].value, "uncontracted": v[1].value, "original": v[2].value, "done": bool(v[3].value)}

    def program_stats(self):
        """complex MACs, scratch elements per executor and steps of the device program
        compiled from the recorded plan (no device needed)."""
        cm, se, ns = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        _host_check(host.qth_program_stats(self.handle, ctypes.byref(cm), ctypes.byref(se), ctypes.byref(ns)))
        return {"cmacs": cm.value, "scratch_elems": se.value, "n_steps": ns.value}

    def plan(self) -> np.ndarray:
        n = ctypes.c_int(0)
        _host_check(host.qth_plan(self.handle, None, 0, ctypes.byref(n)))
        out = np.zeros((max(n.value, 1), 8), dtype=np.int32)
        _host_check(host.qth_plan(self.handle, out.ctypes.data_as(_i32p), n.value, ctypes.byref(n)))
        return out[:n.value]


def set_max_time(seconds: float) -> None:
    """The reference's maxTime (src/Network.h:51): contractions stop with
    ContractionFailure once the started timer passes it."""
    _host_check(host.qth_set_max_time(float(seconds)))


def get_max_time() -> float:
    return float(host.qth_get_max_time())


def timer_start() -> None:
    """totTimer.start() (src/Timer.h)."""
    _host_check(host.qth_timer_start())


def timer_reset() -> None:
    _host_check(host.qth_timer_reset())


def timer_elapsed() -> float:
    return float(host.qth_timer_elapsed())


def preprocess(qasm: str, measure: Optional[str], time_threshold: float, seed: int = 0,
               ctx: Optional[Context] = None) -> Tuple[bool, List[Tuple[int, int]]]:
    """preProcess (src/preprocess.h:28-52): (found, mCreatedFrom record) of the first of
    up to 100 seeded stochastic contractions that finishes within time_threshold s."""
    n, found = ctypes.c_int(0), ctypes.c_int(0)
    h = ctx.handle if ctx is not None else None
    m = measure.encode() if measure is not None else None
    cap = 1 << 16  # steps: a network with more nodes cannot be recorded here
    out = (ctypes.c_int32 * (2 * cap))()
    _host_check(host.qth_preprocess(h, qasm.encode(), m, float(time_threshold), ctypes.c_uint64(seed), out, cap,
                                    ctypes.byref(n), ctypes.byref(found)))
    if n.value > cap:
        raise InvalidFunctionInput(f"preprocess: {n.value} steps exceed the {cap} this binding records")
    return bool(found.value), [(out[2 * i], out[2 * i + 1]) for i in range(n.value)]


def qaoa_edge_text(dgf: str, p: int, ref_map: bool, e: int, betas_gammas: Sequence[float]) -> Tuple[str, str]:
    bg = (ctypes.c_double * len(betas_gammas))(*betas_gammas)
    q = ctypes.create_string_buffer(1 << 20)
    m = ctypes.create_string_buffer(1 << 16)
    _host_check(host.qth_qaoa_edge_text(dgf.encode(), p, 1 if ref_map else 0, e, bg, q, 1 << 20, m, 1 << 16))
    return q.value.decode(), m.value.decode()


def qaoa_num_edges(dgf: str) -> Tuple[int, int]:
    ne, nv = ctypes.c_int(), ctypes.c_int()
    _host_check(host.qth_qaoa_num_edges(dgf.encode(), ctypes.byref(ne), ctypes.byref(nv)))
    return ne.value, nv.value


def qaoa_energy_plain(ctx: Context, dgf: str, p: int, ref_map: bool, betas_gammas: Sequence[float],
                      seed: int = 0) -> Tuple[float, np.ndarray]:
    ne, _ = qaoa_num_edges(dgf)
    bg = (ctypes.c_double * len(betas_gammas))(*betas_gammas)
    zz = np.zeros(ne, dtype=np.float64)
    energy = ctypes.c_double()
    _host_check(host.qth_qaoa_energy_plain(ctx.handle, dgf.encode(), p, 1 if ref_map else 0, bg,
                                           ctypes.c_uint64(seed), zz.ctypes.data_as(_dp), ctypes.byref(energy)))
    return energy.value, zz


class _DeviceBatch:
    """Launch / capture / timing of a qt_batch owned by a host-side batch object."""

    ctx: Context
    dev: ctypes.c_void_p

    def launch(self) -> None:
        self.ctx.check(hip.qt_batch_launch(self.dev), "qt_batch_launch")

    def capture(self) -> None:
        self.ctx.check(hip.qt_batch_capture(self.dev), "qt_batch_capture")

    def poll(self) -> bool:
        """True once the last launch (with its group sums) has finished; never blocks."""
        d = ctypes.c_int()
        self.ctx.check(hip.qt_batch_poll(self.dev, ctypes.byref(d)), "qt_batch_poll")
        return bool(d.value)

    def flush(self) -> None:
        """Launch the group sums still pending for the last launch (asynchronous)."""
        self.ctx.check(hip.qt_batch_flush(self.dev), "qt_batch_flush")

    def uncapture(self) -> None:
        self.ctx.check(hip.qt_batch_uncapture(self.dev), "qt_batch_uncapture")

    def config(self, lanes: int, wgs_per_cu: int, scratch: int = 0) -> None:
        """Executor shape; scratch: 0 automatic, 1 LDS, 2 device memory."""
        self.ctx.check(hip.qt_batch_config_ex(self.dev, lanes, wgs_per_cu, scratch), "qt_batch_config_ex")

    def autotune(self):
        l, w, ms = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        self.ctx.check(hip.qt_batch_autotune(self.dev, ctypes.byref(l), ctypes.byref(w), ctypes.byref(ms)),
                       "qt_batch_autotune")
        return l.value, w.value, ms.value

    def time(self, reps: int) -> float:
        ms = ctypes.c_float()
        self.ctx.check(hip.qt_batch_time(self.dev, reps, ctypes.byref(ms)), "qt_batch_time")
        return ms.value

    def time_kernel(self, reps: int) -> float:
        """Mean ms of the contraction kernel alone (no group sums)."""
        ms = ctypes.c_float()
        self.ctx.check(hip.qt_batch_time_kernel(self.dev, reps, ctypes.byref(ms)), "qt_batch_time_kernel")
        return ms.value

    def compile(self, hermitian: bool = False) -> float:
        """Generate, compile (hipRTC) and load straight-line code for this batch's
        programs; later launches run it.  hermitian: the leaves are conjugate-symmetric
        density-matrix-picture tensors (QT_SPECIALISE_HERMITIAN).  Returns the
        milliseconds it took."""
        ms = ctypes.c_float()
        self.ctx.check(hip.qt_batch_compile(self.dev, QT_SPECIALISE_HERMITIAN if hermitian else 0, ctypes.byref(ms)),
                       "qt_batch_compile")
        return ms.value

    def sums_only(self, on: bool = True) -> bool:
        """Group sums only (qt_batch_set_sums_only): launches of the compiled batch write
        no per-instance values; returns whether launches now take that path."""
        a = ctypes.c_int()
        self.ctx.check(hip.qt_batch_set_sums_only(self.dev, 1 if on else 0, ctypes.byref(a)), "qt_batch_set_sums_only")
        self._sums_only = bool(on)
        return bool(a.value)

    def specialised(self):
        """(active, FMAs per launch) of the generated executor."""
        a, f = ctypes.c_int(), ctypes.c_int64()
        self.ctx.check(hip.qt_batch_specialised_info(self.dev, ctypes.byref(a), ctypes.byref(f)),
                       "qt_batch_specialised_info")
        return bool(a.value), f.value

    def info(self):
        cm = ctypes.c_int64()
        l, w, g, se = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.ctx.check(hip.qt_batch_info(self.dev, ctypes.byref(cm), ctypes.byref(l), ctypes.byref(w),
                                         ctypes.byref(g), ctypes.byref(se)), "qt_batch_info")
        return {"cmacs": cm.value, "lanes": l.value, "wgs_per_cu": w.value, "global_scratch": bool(g.value & 1),
                "staged_leaves": bool(g.value & 2), "scratch_elems": se.value}


class QaoaBatch(_DeviceBatch):
    """F_p for many angle sets per launch (qth_qaoa_batch_*)."""

    def __init__(self, ctx: Context, dgf: str, p: int, ref_map: bool = True, precision: int = QT_C64_4M,
                 order: int = ORDER_GREEDY, seed: int = 0):
        h = _vp()
        _host_check(host.qth_qaoa_batch_create(ctx.handle, dgf.encode(), p, 1 if ref_map else 0, precision, order,
                                               ctypes.c_uint64(seed), ctypes.byref(h)))
        self.handle = h
        self.ctx = ctx
        self.p = p
        self.dev = host.qth_qaoa_batch_device(h)
        self.n_edges, _ = qaoa_num_edges(dgf)
        self.n_sets = 0

    def set_angles(self, bg: np.ndarray) -> None:
        bg = np.ascontiguousarray(bg, dtype=np.float64).reshape(-1, 2 * self.p)
        _host_check(host.qth_qaoa_batch_set_angles(self.handle, bg.shape[0], bg.ctypes.data_as(_dp)))
        self.n_sets = bg.shape[0]

    def update_angles(self, bg=None, device_ptr: int = 0, n_sets: int = 0) -> None:
        """New angles for the current sets, written on the device (qth_qaoa_batch_update_angles):
        from a host array ``bg`` (n_sets x 2p), or from ``device_ptr``, the address of n_sets x 2p
        doubles in device memory (e.g. a torch tensor's data_ptr(), its writes completed)."""
        if bg is not None:
            bg = np.ascontiguousarray(bg, dtype=np.float64).reshape(-1, 2 * self.p)
            _host_check(host.qth_qaoa_batch_update_angles(self.handle, bg.shape[0], bg.ctypes.data, 1))
        else:
            _host_check(host.qth_qaoa_batch_update_angles(self.handle, n_sets, _c.c_void_p(device_ptr), 0))

    def results(self):
        """(energies, per-edge <ZZ>); the second is None after group-sums-only launches."""
        e = np.zeros(self.n_sets)
        if getattr(self, "_sums_only", False):
            _host_check(host.qth_qaoa_batch_results(self.handle, e.ctypes.data_as(_dp), None))
            return e, None
        zz = np.zeros(self.n_sets * self.n_edges)
        _host_check(host.qth_qaoa_batch_results(self.handle, e.ctypes.data_as(_dp), zz.ctypes.data_as(_dp)))
        return e, zz.reshape(self.n_sets, self.n_edges)

    def cmacs_per_energy(self) -> int:
        return int(host.qth_qaoa_batch_cmacs(self.handle))

    def close(self):
        if self.handle:
            host.qth_qaoa_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def qaoa_source(dgf: str, p: int, first_edge: int, n_edges: int, precision: int = QT_C64_4M, ref_map: bool = True,
                order: int = ORDER_GREEDY, seed: int = 0, hermitian: bool = False) -> str:
    """Specialised-executor source of some QAOA edge programs (host only)."""
    n = ctypes.c_size_t()
    args = (dgf.encode(), p, 1 if ref_map else 0, order, ctypes.c_uint64(seed), precision,
            QT_SPECIALISE_HERMITIAN if hermitian else 0, first_edge, n_edges)
    _host_check(host.qth_qaoa_source(*args, None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _host_check(host.qth_qaoa_source(*args, buf, n.value + 1, ctypes.byref(n)))
    return buf.value.decode()


def program_compile(source: str, arch: str = "gfx950") -> bytes:
    """Code object of generated source for an architecture (host only)."""
    n = ctypes.c_size_t()
    rc = hip.qt_program_compile(source.encode(), arch.encode(), None, 0, ctypes.byref(n))
    if rc != QT_OK:
        raise DeviceFailure(f"qt_program_compile: status {rc}")
    buf = ctypes.create_string_buffer(n.value)
    rc = hip.qt_program_compile(source.encode(), arch.encode(), buf, n.value, ctypes.byref(n))
    if rc != QT_OK:
        raise DeviceFailure(f"qt_program_compile: status {rc}")
    return buf.raw[:n.value]


def angle_text_float(x: float) -> float:
    """qt_angle_text_float: the float that printing x at the default ostream precision
    and parsing it with std::stof gives (the routine the device runs for rotation blocks)."""
    f = ctypes.c_float(0.0)
    st = hip.qt_angle_text_float(ctypes.c_double(x), ctypes.byref(f))
    if st != 0:
        raise InvalidFunctionInput(f"angle {x!r} outside the supported range")
    return f.value


def qaoa_param_block(p: int, betas_gammas: Sequence[float]) -> np.ndarray:
    """The 32 p complex parameter elements of one angle set (complex128 array)."""
    bg = np.ascontiguousarray(betas_gammas, dtype=np.float64)
    out = np.zeros(64 * p)
    _host_check(host.qth_qaoa_param_block(p, bg.ctypes.data_as(_dp), out.ctypes.data_as(_dp)))
    return out.view(np.complex128)


class NetBatch(_DeviceBatch):
    """Many networks (QASM + measurement) in one batch (qth_netbatch_*)."""

    def __init__(self, ctx: Context, qasms: Sequence[str], measures: Sequence[Optional[str]], base_dir: str = "",
                 order: int = ORDER_GREEDY, seed: int = 0, given: Optional[Sequence[Sequence[Tuple[int, int]]]] = None,
                 precision: int = QT_C64_4M, dim_log2: int = 2):
        n = len(qasms)
        q = (ctypes.c_char_p * n)(*[x.encode() for x in qasms])
        m = (ctypes.c_char_p * n)(*[(x.encode() if x is not None else None) for x in measures])
        g = gl = None
        if order == ORDER_GIVEN:
            flat = [v for seq in given for pr in seq for v in pr]
            g = _i32(flat)
            gl = _i32([len(seq) for seq in given])
        h = _vp()
        _host_check(host.qth_netbatch_create(ctx.handle, n, q, m, base_dir.encode(), dim_log2, order,
                                             ctypes.c_uint64(seed),
                                             g, gl, precision, ctypes.byref(h)))
        self.handle = h
        self.ctx = ctx
        self.n = n
        self.dev = host.qth_netbatch_device(h)

    def results(self) -> np.ndarray:
        v = np.zeros(2 * self.n)
        _host_check(host.qth_netbatch_results(self.handle, v.ctypes.data_as(_dp)))
        return v[0::2] + 1j * v[1::2]

    def cmacs(self) -> np.ndarray:
        c = np.zeros(self.n, dtype=np.int64)
        _host_check(host.qth_netbatch_cmacs(self.handle, c.ctypes.data_as(_i64p)))
        return c

    def close(self):
        if self.handle:
            host.qth_netbatch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def qaoa_optimize(ctx: Context, dgf: str, p: int, precision: int = QT_C64_4M, n_random: int = 256,
                  max_evals: int = 2000, seed: int = 0):
    """Maximise F_p (src/maxcut.cpp:145-222) on the device; returns (angles, energy, evaluations)."""
    ang = np.zeros(2 * p)
    e = ctypes.c_double()
    n = ctypes.c_int()
    _host_check(host.qth_qaoa_optimize(ctx.handle, dgf.encode(), p, precision, n_random, max_evals,
                                       ctypes.c_uint64(seed), ang.ctypes.data_as(_dp), ctypes.byref(e),
                                       ctypes.byref(n)))
    return ang, e.value, n.value


def maxcut_final_string(ctx: Context, dgf: str, p: int, betas_gammas: Sequence[float], seed: int = 0,
                        with_probs: bool = False):
    """Greedy final cut (src/maxcut.cpp:29-140); returns (bits, cut edges), and the
    conditional probability that decided each bit when with_probs."""
    _, nv = qaoa_num_edges(dgf)
    bg = (ctypes.c_double * len(betas_gammas))(*betas_gammas)
    bits = (ctypes.c_int * nv)()
    cut = ctypes.c_int()
    probs = (ctypes.c_double * nv)()
    _host_check(host.qth_maxcut_final_string_ex(ctx.handle, dgf.encode(), p, bg, ctypes.c_uint64(seed), bits,
                                                ctypes.byref(cut), probs))
    if with_probs:
        return list(bits), cut.value, list(probs)
    return list(bits), cut.value


# ---- serving -----------------------------------------------------------------------

SERVE_QAOA_ENERGY, SERVE_AMPLITUDE = 0, 1
ROUTE_SMALL, ROUTE_MEDIUM, ROUTE_EAGER = 0, 1, 2
_ROUTES = {ROUTE_SMALL: "small", ROUTE_MEDIUM: "medium", ROUTE_EAGER: "eager"}


def serve_plan(kind: int, text: str, p: int = 1, ref_map: bool = False, bits: Optional[str] = None,
               precision: int = QT_C64_4M):
    """Route, planned device bytes and complex MACs of a serving request (host only)."""
    route, planned, cm = ctypes.c_int(), ctypes.c_int64(), ctypes.c_int64()
    _host_check(host.qth_serve_plan(kind, text.encode(), p, 1 if ref_map else 0,
                                    bits.encode() if bits is not None else None, precision, ctypes.byref(route),
                                    ctypes.byref(planned), ctypes.byref(cm)))
    return {"route": _ROUTES[route.value], "planned_bytes": planned.value, "cmacs": cm.value}


class Server:
    """Continuous-batching request loop (include/qtorch_host.h, qth_server_*): QAOA
    energies and single amplitudes submitted from any thread, results polled or
    awaited by id."""

    def __init__(self, device: int = 0, slots: int = 2, precision: int = QT_C64_4M, hbm_budget: int = 16 << 30,
                 max_requests_per_launch: int = 512, eager_precision: int = QT_PLAIN_C128):
        h = _vp()
        _host_check(host.qth_server_create_ex(device, slots, precision, eager_precision, hbm_budget,
                                              max_requests_per_launch, ctypes.byref(h)))
        self.handle = h

    def submit_qaoa(self, dgf: str, p: int, betas_gammas: Sequence[float], ref_map: bool = False) -> int:
        bg = (ctypes.c_double * len(betas_gammas))(*betas_gammas)
        i = ctypes.c_int64()
        _host_check(host.qth_server_submit_qaoa(self.handle, dgf.encode(), p, 1 if ref_map else 0, bg,
                                                ctypes.byref(i)))
        return i.value

    def submit_amplitude(self, qasm: str, bits: str) -> int:
        i = ctypes.c_int64()
        _host_check(host.qth_server_submit_amplitude(self.handle, qasm.encode(), bits.encode(), ctypes.byref(i)))
        return i.value

    @staticmethod
    def _result(status, re, im, lat, route):
        return {"status": status, "value": complex(re, im), "latency_ms": lat, "route": _ROUTES.get(route, None)}

    def poll(self, rid: int):
        done, st, route = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        re, im, lat = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _host_check(host.qth_server_poll(self.handle, rid, ctypes.byref(done), ctypes.byref(st), ctypes.byref(re),
                                         ctypes.byref(im), ctypes.byref(lat), ctypes.byref(route)))
        return self._result(st.value, re.value, im.value, lat.value, route.value) if done.value else None

    def wait(self, rid: int):
        route = ctypes.c_int()
        re, im, lat = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        st = host.qth_server_wait(self.handle, rid, ctypes.byref(re), ctypes.byref(im), ctypes.byref(lat),
                                  ctypes.byref(route))
        return self._result(st, re.value, im.value, lat.value, route.value)

    def drain(self) -> None:
        _host_check(host.qth_server_drain(self.handle))

    def stats(self):
        c = (ctypes.c_int64 * 5)()
        m = (ctypes.c_double * 5)()
        peak = ctypes.c_int64()
        _host_check(host.qth_server_stats(self.handle, c, m, ctypes.byref(peak)))
        res_b, resets = ctypes.c_int64(), ctypes.c_int64()
        _host_check(host.qth_server_memory(self.handle, ctypes.byref(res_b), ctypes.byref(resets)))
        return {"submitted": c[0], "completed": c[1], "rejected": c[2], "launches": c[3], "edge_evals": c[4],
                "p50_ms": m[0], "p99_ms": m[1], "mean_ms": m[2], "overlap_ms": m[3], "busy_ms": m[4],
                "peak_planned_bytes": peak.value, "resident_bytes": res_b.value, "slot_resets": resets.value}

    def close(self) -> None:
        if self.handle:
            host.qth_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- one large network, sliced -----------------------------------------------------

class SlicedNetwork:
    """A network planned once (QTH_ORDER_* order), wires sliced until every tensor
    has at most max_bits index bits; run() contracts every slice in the given form and
    sums the slices in FP64 on the device."""

    def __init__(self, qasm: str, measure: str, dim_log2: int = 1, order: int = ORDER_LINEGRAPH, seed: int = 0,
                 max_bits: int = 28):
        h = _vp()
        _host_check(host.qth_sliced_create(qasm.encode(), measure.encode(), dim_log2, order, ctypes.c_uint64(seed),
                                           max_bits, ctypes.byref(h)))
        self.handle = h

    def plan(self):
        v = (ctypes.c_int64 * 8)()
        _host_check(host.qth_sliced_plan(self.handle, v))
        return {"sliced_wires": v[0], "peak_bits": v[1], "slices": v[2], "cmacs_per_slice": v[3],
                "gemm_cmacs_per_slice": v[4], "peak_bytes": v[5], "fused_subtrees": v[6], "fused_steps": v[7]}

    def set_fusion(self, slices_per_launch: int = -1) -> None:
        """-1 (default): small subtrees run as device programs, one batch launch per
        chunk of slices sized by memory; n > 0: at most n slices per launch; 0: every
        step through qt_contract_pair."""
        _host_check(host.qth_sliced_set_fusion(self.handle, int(slices_per_launch)))

    def set_range(self, first: int, count: int = -1) -> None:
        """Later runs contract only slices [first, first + count) (count < 0: to the
        last) and return their sum."""
        _host_check(host.qth_sliced_set_range(self.handle, int(first), int(count)))

    def set_slices(self, slices: Sequence[int]) -> None:
        """Later runs contract exactly these slices and return their sum."""
        arr = (ctypes.c_int64 * len(slices))(*[int(x) for x in slices])
        _host_check(host.qth_sliced_set_slices(self.handle, arr, len(slices)))

    def nonzero_slices(self, max_count: int) -> List[int]:
        """Up to max_count slices, in increasing order, in which no leaf's sub-tensor
        is all zero (every other slice is exactly zero).  Host only; raises when the
        search budget runs out first."""
        out = (ctypes.c_int64 * max(int(max_count), 1))()
        n = ctypes.c_int64()
        _host_check(host.qth_sliced_nonzero_slices(self.handle, int(max_count), out, ctypes.byref(n)))
        return list(out[:n.value])

    def nonzero_slices_ex(self, max_count: int, budget: int = 1 << 24) -> Tuple[List[int], bool]:
        """(slices, complete): complete is False when the budget of partial assignments
        ran out before the list was full or every slice was visited."""
        out = (ctypes.c_int64 * max(int(max_count), 1))()
        n = ctypes.c_int64()
        done = ctypes.c_int()
        _host_check(host.qth_sliced_nonzero_slices_ex(self.handle, int(max_count), int(budget), out, ctypes.byref(n),
                                                      ctypes.byref(done)))
        return list(out[:n.value]), bool(done.value)

    def run(self, ctx: Context, precision: int = QT_PLAIN_C128):
        re, im, ms = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _host_check(host.qth_sliced_run(self.handle, ctx.handle, precision, ctypes.byref(re), ctypes.byref(im),
                                        ctypes.byref(ms)))
        return complex(re.value, im.value), ms.value

    def close(self) -> None:
        if self.handle:
            host.qth_sliced_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
