// Sliced execution of one large network with a chosen arithmetic form (see sliced.cpp).
#pragma once

#include <complex>
#include <memory>
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
    int fused_subtrees = 0;        // subtrees of small steps run as device programs
    int fused_steps = 0;           // steps inside them (of a slice)
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
    // precision: QT_PLAIN_C128 (complex128), or a complex64 form (optionally with
    // QT_RESCALE: large steps rescaled by powers of two).  Contracts every slice on
    // ctx's stream.
    std::complex<double> Run(qt_ctx* ctx, int precision, SliceStats* stats) const;
    // Leaf values of (qasm, measure) for this network's structure (parsed as in the
    // constructor), held apart from the network: several threads can run the same
    // planned network with their own values at once.
    struct LeafState {
        std::vector<std::vector<std::complex<double>>> vals;
        std::vector<double> const_pool;  // the fused programs' constant leaves, FP64 pairs
    };
    LeafState MakeLeafState(const std::string& qasm, const std::string& measure) const;
    // New values for one leaf of a leaf state (the fused programs' constants follow).
    void SetLeaf(LeafState& ls, int leaf, const std::vector<std::complex<double>>& values) const;
    // Run with those leaf values instead of the network's own.
    std::complex<double> Run(qt_ctx* ctx, int precision, SliceStats* stats, const LeafState& leaves) const;
    const SliceStats& Plan() const { return plan_; }
    // slices_per_launch = 0: every step through qt_contract_pair; < 0 (default): fused
    // subtrees, slices per launch chosen by memory; > 0: at most that many per launch.
    void SetFusion(int64_t slices_per_launch) { fuse_slices_ = slices_per_launch; }
    // Run contracts only slices [first, first + count) of the plan's slices (count < 0:
    // to the last one) and returns their sum; the default is every slice.
    void SetSliceRange(int64_t first, int64_t count);
    // Run contracts exactly these slices (each in [0, slices)) and returns their sum.
    void SetSliceList(const std::vector<int64_t>& slices);
    // The first max_count slices, in increasing order, that are not zero for one of
    // these reasons: a leaf's sub-tensor is all zero, or two sliced wires tied by a
    // chain of diagonal leaves (CZ, RZ, ...) take different values, or a wire takes a
    // value its chain rules out (a |0> state, an output bit).  Every slice left out
    // contributes exactly zero.  Stops early after visiting `budget` partial
    // assignments of the sliced wires; *complete (if given) is then false, and true
    // when the list is full or every slice was visited.
    std::vector<int64_t> NonzeroSlices(int64_t max_count, int64_t budget = int64_t(1) << 24,
                                       bool* complete = nullptr) const;
    // New leaf values for the same network structure (the plan, slicing and fused
    // programs depend on the structure only): the nodes of (qasm, measure) parsed as
    // in the constructor must have the ranks of this network's leaves.
    void SetLeafValues(const std::string& qasm, const std::string& measure);

private:
    struct Step {
        int a, b;  // operand ids: leaves 0..n_leaves-1, then step results
        int rank_a, rank_b, k, rank_c;
        std::vector<int32_t> pair_a, pair_b, free_src;
    };
    void PlanFusion();
    // leaf i's sub-tensor for `slice` as interleaved FP64 (re, im) at dst
    void Extract(int leaf, int64_t slice, double* dst) const { Extract(leaf_vals_, leaf, slice, dst); }
    void Extract(const std::vector<std::vector<std::complex<double>>>& vals, int leaf, int64_t slice,
                 double* dst) const;
    std::complex<double> RunWith(qt_ctx* ctx, int precision, SliceStats* stats,
                                 const std::vector<std::vector<std::complex<double>>>& vals,
                                 const std::vector<double>& const_pool) const;

    int dl_ = 2;
    // the sweep order in the amplitude picture parses diagonal two-qubit gates as two
    // rank-3 halves (NetworkOptions::split_diagonal)
    bool split_diagonal_ = false;
    int n_leaves_ = 0;
    std::vector<std::vector<std::complex<double>>> leaf_vals_;  // unsliced
    std::vector<std::vector<int>> leaf_wires_;                  // wire id of each axis
    std::vector<int> sliced_;                                   // sliced wire ids
    std::vector<std::vector<std::pair<int, int>>> cut_;         // per leaf: (axis, sliced-wire index)
    std::vector<int> leaf_rank_;                                // per leaf: rank without the sliced axes
    std::vector<Step> steps_;                                   // with the sliced axes removed
    SliceStats plan_;
    // fused subtrees: the maximal subtrees of small steps, one device program each
    int64_t fuse_slices_ = -1;
    int64_t slice_first_ = 0, slice_count_ = -1;
    std::vector<int64_t> slice_list_;  // non-empty: the slices Run contracts
    std::vector<char> step_fused_;                  // per step
    std::vector<char> leaf_direct_;                 // per leaf: an operand of an unfused step
    std::vector<int> roots_;                        // root step of each subtree
    std::vector<std::shared_ptr<qt_program>> progs_;  // one per root
    std::vector<double> const_pool_;                // FP64 values of their uncut leaves
    std::vector<std::pair<int, int64_t>> const_leaves_;  // (uncut leaf, its offset in const_pool_)
    std::vector<std::pair<int, int64_t>> param_leaves_;  // (cut leaf, offset in a slice's block)
    int64_t block_elems_ = 0;                       // leaf elements of one slice's block
    int64_t root_bytes_c128_ = 0;                   // subtree results of one slice, complex128
};

}  // namespace qtorch
