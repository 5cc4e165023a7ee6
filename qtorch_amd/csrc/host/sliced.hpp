// Sliced execution of one large network with a chosen arithmetic form (see sliced.cpp).
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "qtorch.hpp"

namespace qtorch {

struct SliceStats {
    int sliced_wires = 0;
    int64_t slices = 0;
    int peak_bits = 0;             // largest tensor of a slice, in index bits
    int64_t cmacs_per_slice = 0;   // complex MACs of one slice (the reference's counter)
    int64_t gemm_cmacs_per_slice = 0;  // of those, in steps shaped for the MFMA GEMM
    int64_t peak_bytes = 0;        // device bytes of one slice's live tensors
    double ms = 0;                 // wall time of Run (device work included)
};

// A network (QASM + measurement text, either picture) planned once in a given
// order, with wires sliced (fixed to each of their D values in turn) until every
// tensor of a slice has at most max_bits index bits.  The slices are independent
// networks of one shape; their scalars are summed in FP64 on the device.
class SlicedNetwork {
public:
    SlicedNetwork(const std::string& qasm, const std::string& measure, int dim_log2, int order, uint64_t seed,
                  int max_bits);
    // precision: QT_PLAIN_C128 (complex128), or a complex64 form, optionally or-ed
    // with QT_RESCALE (every step rescaled by a power of two: no FP32 underflow for
    // any depth, at two extra passes per step).  Contracts every slice on ctx's stream.
    std::complex<double> Run(qt_ctx* ctx, int precision, SliceStats* stats) const;
    const SliceStats& Plan() const { return plan_; }

private:
    struct Step {
        int a, b;  // operand ids: leaves 0..n_leaves-1, then step results
        int rank_a, rank_b, k, rank_c;
        std::vector<int32_t> pair_a, pair_b, free_src;
    };
    int dl_ = 2;
    int n_leaves_ = 0;
    std::vector<std::vector<std::complex<double>>> leaf_vals_;  // unsliced
    std::vector<std::vector<int>> leaf_wires_;                  // wire id of each axis
    std::vector<int> sliced_;                                   // sliced wire ids
    std::vector<Step> steps_;                                   // with the sliced axes removed
    SliceStats plan_;
};

}  // namespace qtorch
