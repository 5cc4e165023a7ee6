"""qtorch_amd — qTorch's tensor-network contraction path on AMD Instinct MI355X (gfx950).

The product is native code: libqtorch_hip.so (HIP kernels + C ABI, include/qtorch_hip.h)
and libqtorch_host.so (qTorch's C++ host surface + C ABI, include/qtorch_host.h).
This package only binds them for Python callers (tests, benchmarks, services).
"""
from ._native import (  # noqa: F401
    QT_C64_3M, QT_C64_4M, QT_C64_BF16X3, QT_DTYPE_C64, QT_DTYPE_C128, QT_FROM_A, QT_PLAIN_C128,
    QT_SCRATCH_AUTO, QT_SCRATCH_GLOBAL, QT_SCRATCH_LDS, QT_SCRATCH_STAGE_LEAVES, QT_RESCALE, QT_SPECIALISE_HERMITIAN, SlicedNetwork,
    ORDER_GIVEN, ORDER_GREEDY, ORDER_LINEGRAPH, ORDER_STOCHASTIC, ORDER_SWEEP, ContractionFailure, Context, NetBatch,
    QaoaBatch, DeviceFailure, EngineError, InvalidContractionMethod, InvalidFile,
    InvalidFileFormat, InvalidFunctionInput, InvalidUserContractionSequence, Network, QbbFailure, Tensor,
    contract_pair, declared_symbols, maxcut_final_string, qaoa_optimize, qaoa_edge_text, qaoa_energy_plain, qaoa_num_edges,
    qaoa_param_block, qaoa_source, program_compile, set_max_time, get_max_time, timer_start, timer_reset,
    timer_elapsed, preprocess, angle_text_float,
    ROUTE_EAGER, ROUTE_MEDIUM, ROUTE_SMALL, SERVE_AMPLITUDE, SERVE_QAOA_ENERGY, Server, serve_plan,
)

__all__ = [n for n in dir() if not n.startswith("_")]
