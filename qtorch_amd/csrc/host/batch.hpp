// Batched evaluation of many networks on the device (see batch.cpp).
#pragma once

#include <complex>
#include <string>
#include <unordered_map>
#include <vector>

#include "qtorch.hpp"

namespace qtorch {

enum { kOrderGiven = 0, kOrderGreedy = 1, kOrderStochastic = 2, kOrderLineGraph = 3, kOrderSweep = 4 };

// Greedy cost order on a (plan-only or device) network.
void GreedyContract(Network& net);
// Circuit sweep order: local merges, then one growing tensor absorbing the rest in
// the order of a line through the qubits (see batch.cpp).  variant 0 takes the line
// of the interaction graph's Fiedler vector; others perturb it (seeded).
void SweepContract(Network& net, uint64_t variant);
// Contracts a network in the requested order (records its plan).
void OrderNetwork(Network& net, int order, uint64_t seed, const std::vector<std::pair<int, int>>* given);

// Leaf tensors, stored once per distinct value set.
class LeafPool {
public:
    int64_t Add(const std::vector<std::complex<double>>& values);
    const std::vector<std::complex<double>>& Data() const { return data_; }

private:
    std::vector<std::complex<double>> data_;
    std::unordered_map<std::string, int64_t> index_;
};

// Engine status -> qtorch exception (DeviceFailure); ctx may be NULL.
void CheckQt(qt_ctx* ctx, qt_status s, const char* what);

// Device program of a contracted (recorded) network; leaf i is original node i.
// ctx may be NULL (compilation is host work).
qt_program* CompileProgram(qt_ctx* ctx, const Network& net, const std::vector<int64_t>& leaf_refs);

// The same with constant folding: leaf i is a parameter when param_refs[i] >= 0
// (that reference, QT_LEAF_PARAM | offset), else a constant with values[i].  Every
// step whose two operands are constants (and whose result has at most
// max_fold_elems elements) is evaluated here in FP64 and its result becomes a
// constant leaf (added to pool); the last step always runs on the device.
qt_program* CompileProgramFolded(const Network& net, const std::vector<int64_t>& param_refs,
                                 const std::vector<std::vector<std::complex<double>>>& values, LeafPool& pool,
                                 int64_t max_fold_elems, int64_t* folded_cmacs);

// F_p's edge light cones (src/maxcut.cpp:171-198) compiled once: one program per
// edge, constant leaves added to `pool` (references from 0), the Rz / Rx gates of
// layer i read parameter elements 32 i and 32 i + 16; the angle-free parts of each
// light cone are contracted here once (constant folding).  Throws DeviceFailure
// when a light cone's tensors exceed a program's limit.
struct CompiledQaoa {
    CompiledQaoa() = default;
    CompiledQaoa(CompiledQaoa&& o) noexcept
        : progs(std::move(o.progs)), edge_cmacs(std::move(o.edge_cmacs)), cmacs(o.cmacs),
          folded_cmacs(o.folded_cmacs), max_scratch(o.max_scratch) {
        o.progs.clear();
    }
    CompiledQaoa(const CompiledQaoa&) = delete;
    CompiledQaoa& operator=(const CompiledQaoa&) = delete;
    ~CompiledQaoa();
    std::vector<qt_program*> progs;
    std::vector<int64_t> edge_cmacs;  // the reference's counter per edge network
    int64_t cmacs = 0;
    int64_t folded_cmacs = 0;         // of those, evaluated once here (angle-free subtrees)
    int64_t max_scratch = 0;
};
CompiledQaoa CompileQaoa(const ExtraData& data, int p, int order, uint64_t seed, LeafPool& pool);
// The 32 p parameter elements of one angle set [betas..., gammas...]: Rz(-gamma_i)
// at 32 i and Rx(2 beta_i) at 32 i + 16, angles printed as the reference prints them.
std::vector<std::complex<double>> QaoaParamBlock(int p, const double* betas_gammas);

// F_p for many angle sets at once: one program per edge light cone, the Rz / Rx
// tensors of each angle set in a parameter block, FP64 energy sums on the device.
class QaoaBatch {
public:
    QaoaBatch(qt_ctx* ctx, const std::string& dgf, int p, bool ref_map, int precision, int order, uint64_t seed);
    ~QaoaBatch();
    QaoaBatch(const QaoaBatch&) = delete;
    QaoaBatch& operator=(const QaoaBatch&) = delete;
    // bg: n_sets x 2p angles, each set laid out as [betas..., gammas...] (src/maxcut.cpp:155-157)
    void SetAngles(int n_sets, const double* bg);
    // New angles for the n_sets sets already set (same layout as SetAngles): the
    // Rz / Rx blocks are rebuilt on the device (qt_batch_rotation_params); bg is in
    // host memory (on_host) or device memory.  Asynchronous.
    void UpdateAngles(int n_sets, const double* bg, bool on_host);
    void Results(double* energies, double* zz);
    qt_batch* Batch() const { return batch_; }
    int NumEdges() const { return int(data_.pairs.size()); }
    int64_t CmacsPerEnergy() const;

private:
    qt_ctx* ctx_;
    ExtraData data_;
    int p_;
    LeafPool pool_;
    int64_t const_elems_ = 0;
    int n_sets_ = 0;
    std::vector<int64_t> edge_cmacs_;
    qt_batch* batch_ = nullptr;
};

// A list of networks (QASM + measurement text) evaluated in one batch.
class NetBatch {
public:
    NetBatch(qt_ctx* ctx, const std::vector<std::string>& qasm, const std::vector<const std::string*>& meas,
             const std::string& base_dir, int dim_log2, int order, uint64_t seed,
             const std::vector<std::vector<std::pair<int, int>>>* given, int precision);
    ~NetBatch();
    NetBatch(const NetBatch&) = delete;
    NetBatch& operator=(const NetBatch&) = delete;
    void Results(std::complex<double>* values);
    qt_batch* Batch() const { return batch_; }
    const std::vector<long long>& Cmacs() const { return cmacs_; }

private:
    qt_ctx* ctx_;
    LeafPool pool_;
    std::vector<long long> cmacs_;
    qt_batch* batch_ = nullptr;
    int n_ = 0;
};

}  // namespace qtorch
