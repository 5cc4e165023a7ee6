// Batched evaluation of many networks on the device (see batch.cpp).
#pragma once

#include <string>
#include <unordered_map>
#include <vector>

#include "qtorch.hpp"

namespace qtorch {

enum { kOrderGiven = 0, kOrderGreedy = 1, kOrderStochastic = 2, kOrderLineGraph = 3 };

// Greedy cost order on a (plan-only or device) network.
void GreedyContract(Network& net);
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

// Device program of a contracted (recorded) network; leaf i is original node i.
qt_program* CompileProgram(qt_ctx* ctx, const Network& net, const std::vector<int64_t>& leaf_refs);

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
