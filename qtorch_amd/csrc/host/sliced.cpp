// Sliced execution of one large network.
//
// The network is planned once (an ordering of the host, e.g. the line-graph
// min-fill order or the greedy order), then wires are chosen for slicing: while a
// tensor of the plan has more than max_bits index bits, the wire that occurs in the
// most of the largest tensors is fixed.  Fixing a wire to one of its D values
// removes that axis from every tensor that carries it (leaves are cut to the
// sub-tensor, steps lose the pair or the free axis), so all D^s slices are the same
// smaller network with different leaf values; sum_v over a wire's values is the
// contraction the wire stood for.  Each slice runs step by step through
// qt_contract_pair (MFMA GEMMs for large steps) and its scalar is added in FP64 on
// the device.
#include "sliced.hpp"

#include <algorithm>
#include <chrono>
#include <unordered_map>

#include "batch.hpp"

namespace qtorch {

SlicedNetwork::SlicedNetwork(const std::string& qasm, const std::string& measure, int dim_log2, int order,
                             uint64_t seed, int max_bits)
    : dl_(dim_log2) {
    if ((dim_log2 != 1 && dim_log2 != 2) || max_bits < 1 || max_bits > 40) throw InvalidFunctionInput();
    NetworkOptions o;
    o.dim_log2 = dim_log2;
    Network net(qasm, &measure, o);
    n_leaves_ = net.GetNumOriginalNodes();
    std::unordered_map<const Wire*, int> wire_id;
    for (int i = 0; i < n_leaves_; ++i) {
        const auto& node = net.GetAllNodes()[size_t(i)];
        leaf_vals_.push_back(node->mHostVals);
        std::vector<int> w;
        for (const auto& wp : node->GetWires()) w.push_back(wire_id.emplace(wp.get(), int(wire_id.size())).first->second);
        if (int(w.size()) != node->mRank) throw InvalidFunctionInput();
        leaf_wires_.push_back(w);
    }
    OrderNetwork(net, order, seed, nullptr);
    const auto& plan = net.GetPlan();
    if (plan.empty()) throw InvalidFunctionInput();
    // wires of every operand (leaves, then step results) and the unsliced steps
    std::vector<std::vector<int>> wires(leaf_wires_);
    std::unordered_map<int, int> operand_of_node;
    for (int i = 0; i < n_leaves_; ++i) operand_of_node[i] = i;
    std::vector<Step> full;
    for (const auto& ps : plan) {
        Step st;
        st.a = operand_of_node.at(ps.a_id);
        st.b = operand_of_node.at(ps.b_id);
        st.rank_a = ps.rank_a;
        st.rank_b = ps.rank_b;
        st.k = ps.k;
        st.rank_c = ps.rank_c;
        st.pair_a = ps.pair_a;
        st.pair_b = ps.pair_b;
        st.free_src = ps.free_src;
        const auto& wa = wires[size_t(st.a)];
        const auto& wb = wires[size_t(st.b)];
        if (int(wa.size()) != st.rank_a || int(wb.size()) != st.rank_b) throw ContractionFailure();
        std::vector<int> wc;
        for (int32_t f : st.free_src)
            wc.push_back((f & QT_FROM_A) ? wa[size_t(f & ~QT_FROM_A)] : wb[size_t(f)]);
        const int id = n_leaves_ + int(full.size());
        if (ps.c_id >= 0) operand_of_node[ps.c_id] = id;
        wires.push_back(wc);
        full.push_back(st);
    }
    // slicing: fix the wire that occurs in most of the tensors above the limit
    std::vector<char> is_sliced(wire_id.size(), 0);
    auto bits_of = [&](const std::vector<int>& w) {
        int n = 0;
        for (int x : w) n += is_sliced[size_t(x)] ? 0 : 1;
        return n * dl_;
    };
    while (true) {
        int worst = 0;
        for (const auto& w : wires) worst = std::max(worst, bits_of(w));
        if (worst <= max_bits) break;
        std::vector<int> count(wire_id.size(), 0);
        for (const auto& w : wires)
            if (bits_of(w) == worst)
                for (int x : w)
                    if (!is_sliced[size_t(x)]) ++count[size_t(x)];
        const int pick = int(std::max_element(count.begin(), count.end()) - count.begin());
        is_sliced[size_t(pick)] = 1;
        sliced_.push_back(pick);
        if (int(sliced_.size()) * dl_ > 40) throw InvalidFunctionInput();  // more than 2^40 slices
    }
    // the sliced steps: axes renumbered without the sliced wires
    auto kept = [&](const std::vector<int>& w) {
        std::vector<int> idx(w.size(), -1);
        int n = 0;
        for (size_t j = 0; j < w.size(); ++j)
            if (!is_sliced[size_t(w[j])]) idx[j] = n++;
        return idx;
    };
    for (size_t s = 0; s < full.size(); ++s) {
        const Step& f = full[s];
        const auto& wa = wires[size_t(f.a)];
        const auto& wb = wires[size_t(f.b)];
        const auto ka = kept(wa), kb = kept(wb);
        Step st;
        st.a = f.a;
        st.b = f.b;
        for (size_t i = 0; i < f.pair_a.size(); ++i) {
            if (wa[size_t(f.pair_a[i])] != wb[size_t(f.pair_b[i])]) throw ContractionFailure();
            if (ka[size_t(f.pair_a[i])] < 0) continue;
            st.pair_a.push_back(ka[size_t(f.pair_a[i])]);
            st.pair_b.push_back(kb[size_t(f.pair_b[i])]);
        }
        for (int32_t fs : f.free_src) {
            const bool from_a = (fs & QT_FROM_A) != 0;
            const int ax = fs & ~QT_FROM_A;
            const int n = from_a ? ka[size_t(ax)] : kb[size_t(ax)];
            if (n >= 0) st.free_src.push_back(from_a ? (n | QT_FROM_A) : n);
        }
        st.rank_a = bits_of(wa) / dl_;
        st.rank_b = bits_of(wb) / dl_;
        st.k = int(st.pair_a.size());
        st.rank_c = int(st.free_src.size());
        steps_.push_back(st);
        // statistics of one slice
        const int64_t macs = int64_t(1) << ((st.rank_c + st.k) * dl_);
        plan_.cmacs_per_slice += macs;
        int from_a = 0;
        while (from_a < st.rank_c && (st.free_src[size_t(from_a)] & QT_FROM_A)) ++from_a;
        bool split = true;
        for (int i = from_a; i < st.rank_c; ++i)
            if (st.free_src[size_t(i)] & QT_FROM_A) split = false;
        if (split && from_a * dl_ >= 6 && (st.rank_c - from_a) * dl_ >= 6 && st.k * dl_ >= 5)
            plan_.gemm_cmacs_per_slice += macs;
    }
    plan_.sliced_wires = int(sliced_.size());
    plan_.slices = int64_t(1) << (int(sliced_.size()) * dl_);
    for (const auto& w : wires) plan_.peak_bits = std::max(plan_.peak_bits, bits_of(w));
    // live bytes of one slice (complex128): leaves, then intermediates until consumed
    int64_t leaves = 0, live = 0, peak = 0;
    for (int i = 0; i < n_leaves_; ++i) leaves += int64_t(16) << bits_of(leaf_wires_[size_t(i)]);
    std::vector<int64_t> bytes(wires.size(), 0);
    for (size_t s = 0; s < steps_.size(); ++s) {
        const int id = n_leaves_ + int(s);
        bytes[size_t(id)] = int64_t(16) << (steps_[s].rank_c * dl_);
        live += bytes[size_t(id)];
        peak = std::max(peak, live);
        for (int op : {steps_[s].a, steps_[s].b})
            if (op >= n_leaves_) live -= bytes[size_t(op)];
    }
    plan_.peak_bytes = leaves + peak;
}

std::complex<double> SlicedNetwork::Run(qt_ctx* ctx, int precision, SliceStats* stats) const {
    if (!ctx) throw DeviceFailure("SlicedNetwork::Run needs a device context");
    const int base = precision & ~QT_RESCALE;
    const bool c128 = base == QT_PLAIN_C128;
    if ((c128 && base != precision) ||
        (!c128 && base != QT_C64_4M && base != QT_C64_3M && base != QT_C64_BF16X3))
        throw InvalidFunctionInput();
    const int dtype = c128 ? QT_DTYPE_C128 : QT_DTYPE_C64;
    const int form = precision;
    const int D = 1 << dl_;
    const auto t0 = std::chrono::steady_clock::now();
    qt_tensor* acc = nullptr;
    CheckQt(ctx, qt_tensor_alloc(ctx, 0, dl_, QT_DTYPE_C128, &acc), "qt_tensor_alloc");
    std::vector<qt_tensor*> ops(size_t(n_leaves_) + steps_.size(), nullptr);
    auto release = [&] {
        for (auto*& t : ops)
            if (t) {
                qt_tensor_free(ctx, t);
                t = nullptr;
            }
    };
    try {
        // sliced axes of every leaf
        std::vector<std::vector<std::pair<int, int>>> cut(static_cast<size_t>(n_leaves_));  // (axis, slice digit)
        std::vector<int> ranks(static_cast<size_t>(n_leaves_));
        for (int i = 0; i < n_leaves_; ++i) {
            const auto& w = leaf_wires_[size_t(i)];
            int r = 0;
            for (size_t j = 0; j < w.size(); ++j) {
                const auto it = std::find(sliced_.begin(), sliced_.end(), w[j]);
                if (it != sliced_.end()) cut[size_t(i)].push_back({int(j), int(it - sliced_.begin())});
                else ++r;
            }
            ranks[size_t(i)] = r;
        }
        std::vector<std::vector<double>> host(static_cast<size_t>(n_leaves_));
        std::vector<std::vector<float>> host32(static_cast<size_t>(n_leaves_));
        std::vector<const void*> ptrs(static_cast<size_t>(n_leaves_));
        for (int64_t slice = 0; slice < plan_.slices; ++slice) {
            for (int i = 0; i < n_leaves_; ++i) {
                const auto& src = leaf_vals_[size_t(i)];
                const int rank = int(leaf_wires_[size_t(i)].size());
                const int64_t n = int64_t(1) << (ranks[size_t(i)] * dl_);
                auto& h = host[size_t(i)];
                h.assign(size_t(2 * n), 0.0);
                // fixed part of the source index, then the kept axes' digits
                int64_t base = 0;
                for (const auto& [axis, digit] : cut[size_t(i)]) {
                    const int64_t v = (slice >> (digit * dl_)) & (D - 1);
                    base += v << (axis * dl_);
                }
                for (int64_t e = 0; e < n; ++e) {
                    int64_t idx = base;
                    int kept = 0;
                    for (int ax = 0; ax < rank; ++ax) {
                        bool is_cut = false;
                        for (const auto& c : cut[size_t(i)]) is_cut |= c.first == ax;
                        if (is_cut) continue;
                        idx += ((e >> (kept * dl_)) & (D - 1)) << (ax * dl_);
                        ++kept;
                    }
                    h[size_t(2 * e)] = src[size_t(idx)].real();
                    h[size_t(2 * e + 1)] = src[size_t(idx)].imag();
                }
                if (c128) {
                    ptrs[size_t(i)] = h.data();
                } else {
                    host32[size_t(i)].assign(h.begin(), h.end());
                    ptrs[size_t(i)] = host32[size_t(i)].data();
                }
            }
            CheckQt(ctx, qt_tensor_create_many(ctx, n_leaves_, ranks.data(), dl_, dtype, ptrs.data(), ops.data()),
                    "qt_tensor_create_many");
            for (size_t s = 0; s < steps_.size(); ++s) {
                const Step& st = steps_[s];
                qt_tensor* c = nullptr;
                CheckQt(ctx, qt_tensor_alloc_empty(ctx, st.rank_c, dl_, dtype, &c), "qt_tensor_alloc_empty");
                ops[size_t(n_leaves_) + s] = c;
                CheckQt(ctx,
                        qt_contract_pair(ctx, ops[size_t(st.a)], st.rank_a, ops[size_t(st.b)], st.rank_b, st.k,
                                         st.pair_a.data(), st.pair_b.data(), st.rank_c, st.free_src.data(), c, form,
                                         nullptr),
                        "qt_contract_pair");
                for (int op : {st.a, st.b}) {
                    qt_tensor_free(ctx, ops[size_t(op)]);
                    ops[size_t(op)] = nullptr;
                }
            }
            CheckQt(ctx, qt_tensor_accumulate(ctx, ops.back(), acc), "qt_tensor_accumulate");
            release();
        }
        double re = 0, im = 0;
        CheckQt(ctx, qt_tensor_read_scalar(ctx, acc, &re, &im), "qt_tensor_read_scalar");
        qt_tensor_free(ctx, acc);
        if (stats) {
            *stats = plan_;
            stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return {re, im};
    } catch (...) {
        release();
        qt_tensor_free(ctx, acc);
        throw;
    }
}

}  // namespace qtorch
