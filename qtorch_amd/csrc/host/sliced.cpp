// Sliced execution of one large network.
//
// The network is planned once (an ordering of the host, e.g. the line-graph
// min-fill order or the greedy order), then wires are chosen for slicing: while a
// tensor of the plan has more than max_bits index bits, the wire that occurs in the
// most of the largest tensors is fixed.  Fixing a wire to one of its D values
// removes that axis from every tensor that carries it (leaves are cut to the
// sub-tensor, steps lose the pair or the free axis), so all D^s slices are the same
// smaller network with different leaf values; sum_v over a wire's values is the
// contraction the wire stood for.  The small subtrees of every slice run as device
// programs in one batch launch per chunk of slices; the remaining steps of each
// slice run through qt_contract_pair (MFMA GEMMs for large steps) and its scalar is
// added in FP64 on the device.
#include "sliced.hpp"

#include <algorithm>
#include <chrono>
#include <unordered_map>

#include "batch.hpp"

namespace qtorch {

namespace {
// lines through the qubits the sweep order tries (the plan with the least work after
// slicing is kept)
constexpr uint64_t kSweepVariants = 16;
}  // namespace

SlicedNetwork::SlicedNetwork(const std::string& qasm, const std::string& measure, int dim_log2, int order,
                             uint64_t seed, int max_bits)
    : dl_(dim_log2), split_diagonal_(order == kOrderSweep && dim_log2 == 1) {
    if ((dim_log2 != 1 && dim_log2 != 2) || max_bits < 1 || max_bits > 40) throw InvalidFunctionInput();
    NetworkOptions o;
    o.dim_log2 = dim_log2;
    o.split_diagonal = split_diagonal_;
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
    // One candidate plan: the unsliced steps in the order, the wires of every operand
    // (leaves, then step results) and the wires the slicer fixes.
    struct Candidate {
        std::vector<Step> full;
        std::vector<std::vector<int>> wires;
        std::vector<char> is_sliced;
        std::vector<int> sliced;
        long double work = 0;  // slices x complex MACs per slice
    };
    auto plan_candidate = [&](Network& cn) {
        Candidate cd;
        const auto& plan = cn.GetPlan();
        if (plan.empty()) throw InvalidFunctionInput();
        cd.wires = leaf_wires_;
        std::unordered_map<int, int> operand_of_node;
        for (int i = 0; i < n_leaves_; ++i) operand_of_node[i] = i;
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
            const auto& wa = cd.wires[size_t(st.a)];
            const auto& wb = cd.wires[size_t(st.b)];
            if (int(wa.size()) != st.rank_a || int(wb.size()) != st.rank_b) throw ContractionFailure();
            std::vector<int> wc;
            for (int32_t f : st.free_src)
                wc.push_back((f & QT_FROM_A) ? wa[size_t(f & ~QT_FROM_A)] : wb[size_t(f)]);
            const int id = n_leaves_ + int(cd.full.size());
            if (ps.c_id >= 0) operand_of_node[ps.c_id] = id;
            cd.wires.push_back(wc);
            cd.full.push_back(st);
        }
        // slicing: fix the wire that occurs in most of the tensors above the limit
        // (weighted by how far above it they are); among equals, the one that leaves
        // the least work per slice
        cd.is_sliced.assign(wire_id.size(), 0);
        auto bits_of = [&](const std::vector<int>& w) {
            int n = 0;
            for (int x : w) n += cd.is_sliced[size_t(x)] ? 0 : 1;
            return n * dl_;
        };
        auto slice_macs = [&]() {
            long double m = 0;
            for (size_t s2 = 0; s2 < cd.full.size(); ++s2) {
                const Step& f = cd.full[s2];
                int bits = bits_of(cd.wires[size_t(n_leaves_) + s2]);
                for (int32_t pa : f.pair_a)
                    if (!cd.is_sliced[size_t(cd.wires[size_t(f.a)][size_t(pa)])]) bits += dl_;
                m += std::ldexp(1.0L, bits);
            }
            return m;
        };
        while (true) {
            int worst = 0;
            for (const auto& w : cd.wires) worst = std::max(worst, bits_of(w));
            if (worst <= max_bits) break;
            std::vector<int64_t> score(wire_id.size(), 0);
            for (const auto& w : cd.wires) {
                const int over = bits_of(w) - max_bits;
                if (over > 0)
                    for (int x : w)
                        if (!cd.is_sliced[size_t(x)]) score[size_t(x)] += int64_t(1) << std::min(over, 20);
            }
            const int64_t top = *std::max_element(score.begin(), score.end());
            int pick = -1;
            long double best_macs = 0;
            for (size_t x = 0; x < score.size(); ++x) {
                if (score[x] != top || top == 0) continue;
                cd.is_sliced[x] = 1;
                const long double m = slice_macs();
                cd.is_sliced[x] = 0;
                if (pick < 0 || m < best_macs) {
                    pick = int(x);
                    best_macs = m;
                }
            }
            if (pick < 0) throw InvalidFunctionInput();
            cd.is_sliced[size_t(pick)] = 1;
            cd.sliced.push_back(pick);
            if (int(cd.sliced.size()) * dl_ > 40) throw InvalidFunctionInput();  // more than 2^40 slices
        }
        cd.work = std::ldexp(slice_macs(), int(cd.sliced.size()) * dl_);
        return cd;
    };
    OrderNetwork(net, order, seed, nullptr);
    Candidate best;
    bool have = false;
    try {
        best = plan_candidate(net);
        have = true;
    } catch (const InvalidFunctionInput&) {  // more than 2^40 slices
        if (order != kOrderSweep) throw;
    }
    if (order == kOrderSweep) {
        // other lines through the qubits: the sweep's width depends on it
        for (uint64_t v = 1; v < kSweepVariants; ++v) {
            Network cn(qasm, &measure, o);
            SweepContract(cn, seed * kSweepVariants + v);
            try {
                Candidate c = plan_candidate(cn);
                if (!have || c.work < best.work) best = std::move(c);
                have = true;
            } catch (const InvalidFunctionInput&) {
            }
        }
        if (!have) throw InvalidFunctionInput();
    }
    std::vector<Step> full = std::move(best.full);
    std::vector<std::vector<int>> wires = std::move(best.wires);
    std::vector<char> is_sliced = std::move(best.is_sliced);
    sliced_ = std::move(best.sliced);
    auto bits_of = [&](const std::vector<int>& w) {
        int n = 0;
        for (int x : w) n += is_sliced[size_t(x)] ? 0 : 1;
        return n * dl_;
    };
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
        // saturates at INT64_MAX rather than wrapping (a work bound compares against it)
        if (__builtin_add_overflow(plan_.cmacs_per_slice, macs, &plan_.cmacs_per_slice))
            plan_.cmacs_per_slice = INT64_MAX;
        int from_a = 0;
        while (from_a < st.rank_c && (st.free_src[size_t(from_a)] & QT_FROM_A)) ++from_a;
        bool split = true;
        for (int i = from_a; i < st.rank_c; ++i)
            if (st.free_src[size_t(i)] & QT_FROM_A) split = false;
        if (split && from_a * dl_ >= 6 && (st.rank_c - from_a) * dl_ >= 6 && st.k * dl_ >= 5)
            if (__builtin_add_overflow(plan_.gemm_cmacs_per_slice, macs, &plan_.gemm_cmacs_per_slice))
                plan_.gemm_cmacs_per_slice = INT64_MAX;
    }
    plan_.sliced_wires = int(sliced_.size());
    plan_.slices = int64_t(1) << (int(sliced_.size()) * dl_);
    for (const auto& w : wires) plan_.peak_bits = std::max(plan_.peak_bits, bits_of(w));
    // device bytes of one slice (complex128): leaves, then intermediates until consumed,
    // plus each step's scratch while it runs (a GEMM-shaped step may permute copies of
    // both operands); the GEMMs' split-K partial products (at most 64 MiB), or the split
    // reduction's partial sums (at most 32 MiB), and the fused subtrees' outputs
    // (PlanFusion) are added on top
    int64_t leaves = 0, live = 0, peak = 0;
    for (int i = 0; i < n_leaves_; ++i) leaves += int64_t(16) << bits_of(leaf_wires_[size_t(i)]);
    std::vector<int64_t> bytes(wires.size(), 0);
    for (size_t s = 0; s < steps_.size(); ++s) {
        const Step& st = steps_[s];
        const int id = n_leaves_ + int(s);
        bytes[size_t(id)] = int64_t(16) << (st.rank_c * dl_);
        live += bytes[size_t(id)];
        int from_a = 0;
        while (from_a < st.rank_c && (st.free_src[size_t(from_a)] & QT_FROM_A)) ++from_a;
        const bool gemm_shape = from_a * dl_ >= 6 && (st.rank_c - from_a) * dl_ >= 6 && st.k * dl_ >= 3;
        const int64_t scratch =
            gemm_shape ? (int64_t(16) << (st.rank_a * dl_)) + (int64_t(16) << (st.rank_b * dl_)) : 0;
        peak = std::max(peak, live + scratch);
        for (int op : {st.a, st.b})
            if (op >= n_leaves_) live -= bytes[size_t(op)];
    }
    plan_.peak_bytes = leaves + peak + (int64_t(64) << 20);
    // sliced axes of every leaf
    cut_.resize(size_t(n_leaves_));
    leaf_rank_.resize(size_t(n_leaves_));
    for (int i = 0; i < n_leaves_; ++i) {
        const auto& w = leaf_wires_[size_t(i)];
        int r = 0;
        for (size_t j = 0; j < w.size(); ++j) {
            const auto it = std::find(sliced_.begin(), sliced_.end(), w[j]);
            if (it != sliced_.end()) cut_[size_t(i)].push_back({int(j), int(it - sliced_.begin())});
            else ++r;
        }
        leaf_rank_[size_t(i)] = r;
    }
    PlanFusion();
}

// Small steps (result <= 2^12 elements, operands <= 2^14, <= 2^16 complex MACs) whose
// operands are leaves or small steps form subtrees; each maximal one becomes a device
// program whose last step writes its result tensor.  One batch launch runs every
// subtree of a chunk of slices (instances = subtree x slice; a slice's cut leaves
// are its instance's parameter block), replacing one launch per small step.
void SlicedNetwork::PlanFusion() {
    constexpr int kOutBits = 12, kOperandBits = 14, kMacBits = 16;
    const int n = int(steps_.size());
    step_fused_.assign(size_t(n), 0);
    std::vector<int> consumer(size_t(n), -1);
    int fused = 0;
    for (int s = 0; s < n; ++s) {
        const Step& st = steps_[size_t(s)];
        bool ok = st.rank_c * dl_ <= kOutBits && st.rank_a * dl_ <= kOperandBits && st.rank_b * dl_ <= kOperandBits &&
                  (st.rank_c + st.k) * dl_ <= kMacBits;
        for (int op : {st.a, st.b})
            if (op >= n_leaves_) {
                consumer[size_t(op - n_leaves_)] = s;
                ok = ok && step_fused_[size_t(op - n_leaves_)];
            }
        step_fused_[size_t(s)] = ok ? 1 : 0;
        if (ok) ++fused;
    }
    leaf_direct_.assign(size_t(n_leaves_), 0);
    for (int s = 0; s < n; ++s)
        if (!step_fused_[size_t(s)])
            for (int op : {steps_[size_t(s)].a, steps_[size_t(s)].b})
                if (op < n_leaves_) leaf_direct_[size_t(op)] = 1;
    for (int s = 0; s < n; ++s)
        if (step_fused_[size_t(s)] && (consumer[size_t(s)] < 0 || !step_fused_[size_t(consumer[size_t(s)])]))
            roots_.push_back(s);
    for (int root : roots_) {
        // the subtree's steps in plan order (a topological order) and its leaves
        std::vector<int> sub;
        std::vector<int> stack{root};
        while (!stack.empty()) {
            const int s = stack.back();
            stack.pop_back();
            sub.push_back(s);
            for (int op : {steps_[size_t(s)].a, steps_[size_t(s)].b})
                if (op >= n_leaves_) stack.push_back(op - n_leaves_);
        }
        std::sort(sub.begin(), sub.end());
        std::unordered_map<int, int> local;  // operand id -> program operand id
        std::vector<int32_t> leaf_ranks, desc, axes;
        std::vector<int64_t> leaf_refs;
        auto operand = [&](int op) {
            if (op >= n_leaves_) return local.at(op);
            auto it = local.find(op);
            if (it != local.end()) return it->second;
            const int64_t elems = int64_t(1) << (leaf_rank_[size_t(op)] * dl_);
            if (cut_[size_t(op)].empty()) {
                const_leaves_.push_back({op, int64_t(const_pool_.size() / 2)});
                leaf_refs.push_back(int64_t(const_pool_.size() / 2));
                for (const auto& v : leaf_vals_[size_t(op)]) {
                    const_pool_.push_back(v.real());
                    const_pool_.push_back(v.imag());
                }
            } else {
                leaf_refs.push_back(QT_LEAF_PARAM | block_elems_);
                param_leaves_.push_back({op, block_elems_});
                block_elems_ += elems;
            }
            leaf_ranks.push_back(leaf_rank_[size_t(op)]);
            const int id = int(leaf_ranks.size()) - 1;
            local[op] = id;
            return id;
        };
        // leaves first (program ids: leaves, then steps)
        for (int s : sub) {
            if (steps_[size_t(s)].a < n_leaves_) operand(steps_[size_t(s)].a);
            if (steps_[size_t(s)].b < n_leaves_) operand(steps_[size_t(s)].b);
        }
        const int nl = int(leaf_ranks.size());
        for (size_t j = 0; j < sub.size(); ++j) {
            const Step& st = steps_[size_t(sub[j])];
            desc.insert(desc.end(), {operand(st.a), operand(st.b), st.rank_a, st.rank_b, st.k, st.rank_c});
            axes.insert(axes.end(), st.pair_a.begin(), st.pair_a.end());
            axes.insert(axes.end(), st.pair_b.begin(), st.pair_b.end());
            axes.insert(axes.end(), st.free_src.begin(), st.free_src.end());
            local[n_leaves_ + sub[j]] = nl + int(j);
        }
        qt_program* p = nullptr;
        if (qt_program_create(nullptr, dl_, nl, leaf_ranks.data(), leaf_refs.data(), int(sub.size()), desc.data(),
                              axes.data(), &p) != QT_OK)
            throw ContractionFailure();
        progs_.emplace_back(p, [](qt_program* q) { qt_program_destroy(q); });
        root_bytes_c128_ += int64_t(16) << (steps_[size_t(root)].rank_c * dl_);
    }
    plan_.fused_subtrees = int(roots_.size());
    plan_.fused_steps = fused;
    // a chunk of slices' subtree results (Run: at most 256 MiB, or one slice's), their
    // parameter blocks and the batch's constants
    if (!roots_.empty()) {
        const int64_t chunk = std::max<int64_t>(
            1, std::min<int64_t>({plan_.slices, 4096, (int64_t(256) << 20) / std::max<int64_t>(1, root_bytes_c128_)}));
        plan_.peak_bytes += chunk * (root_bytes_c128_ + 16 * block_elems_) + int64_t(const_pool_.size()) * 8;
    }
}

SlicedNetwork::LeafState SlicedNetwork::MakeLeafState(const std::string& qasm, const std::string& measure) const {
    NetworkOptions o;
    o.dim_log2 = dl_;
    o.split_diagonal = split_diagonal_;
    Network net(qasm, &measure, o);
    if (net.GetNumOriginalNodes() != n_leaves_) throw InvalidFunctionInput();
    LeafState ls;
    ls.vals.reserve(size_t(n_leaves_));
    for (int i = 0; i < n_leaves_; ++i) {
        const auto& node = net.GetAllNodes()[size_t(i)];
        if (node->mRank != int(leaf_wires_[size_t(i)].size()) ||
            node->mHostVals.size() != leaf_vals_[size_t(i)].size())
            throw InvalidFunctionInput();
        ls.vals.push_back(node->mHostVals);
    }
    ls.const_pool = const_pool_;
    for (const auto& [leaf, off] : const_leaves_) {
        const auto& v = ls.vals[size_t(leaf)];
        for (size_t e = 0; e < v.size(); ++e) {
            ls.const_pool[size_t(2 * (off + int64_t(e)))] = v[e].real();
            ls.const_pool[size_t(2 * (off + int64_t(e)) + 1)] = v[e].imag();
        }
    }
    return ls;
}

void SlicedNetwork::SetLeaf(LeafState& ls, int leaf, const std::vector<std::complex<double>>& v) const {
    if (leaf < 0 || leaf >= n_leaves_ || ls.vals.size() != size_t(n_leaves_) ||
        v.size() != leaf_vals_[size_t(leaf)].size())
        throw InvalidFunctionInput();
    ls.vals[size_t(leaf)] = v;
    for (const auto& [cl, off] : const_leaves_)
        if (cl == leaf)
            for (size_t e = 0; e < v.size(); ++e) {
                ls.const_pool[size_t(2 * (off + int64_t(e)))] = v[e].real();
                ls.const_pool[size_t(2 * (off + int64_t(e)) + 1)] = v[e].imag();
            }
}

void SlicedNetwork::SetLeafValues(const std::string& qasm, const std::string& measure) {
    LeafState ls = MakeLeafState(qasm, measure);
    leaf_vals_ = std::move(ls.vals);
    const_pool_ = std::move(ls.const_pool);
}

void SlicedNetwork::Extract(const std::vector<std::vector<std::complex<double>>>& vals, int leaf, int64_t slice,
                            double* dst) const {
    const auto& src = vals[size_t(leaf)];
    const int rank = int(leaf_wires_[size_t(leaf)].size());
    const int64_t D = int64_t(1) << dl_;
    const int64_t n = int64_t(1) << (leaf_rank_[size_t(leaf)] * dl_);
    // fixed part of the source index, then the kept axes' digits
    int64_t base = 0;
    std::vector<int> kept;
    for (int ax = 0; ax < rank; ++ax) {
        bool is_cut = false;
        for (const auto& [axis, digit] : cut_[size_t(leaf)])
            if (axis == ax) {
                is_cut = true;
                base += ((slice >> (digit * dl_)) & (D - 1)) << (ax * dl_);
            }
        if (!is_cut) kept.push_back(ax);
    }
    for (int64_t e = 0; e < n; ++e) {
        int64_t idx = base;
        for (size_t k = 0; k < kept.size(); ++k) idx += ((e >> (int(k) * dl_)) & (D - 1)) << (kept[k] * dl_);
        dst[2 * e] = src[size_t(idx)].real();
        dst[2 * e + 1] = src[size_t(idx)].imag();
    }
}

void SlicedNetwork::SetSliceRange(int64_t first, int64_t count) {
    if (first < 0 || first >= plan_.slices || count == 0) throw InvalidFunctionInput();
    slice_first_ = first;
    slice_count_ = count;
    slice_list_.clear();
}

void SlicedNetwork::SetSliceList(const std::vector<int64_t>& slices) {
    if (slices.empty()) throw InvalidFunctionInput();
    for (int64_t s : slices)
        if (s < 0 || s >= plan_.slices) throw InvalidFunctionInput();
    slice_list_ = slices;
}

std::vector<int64_t> SlicedNetwork::NonzeroSlices(int64_t max_count, int64_t budget, bool* complete) const {
    std::vector<int64_t> found;
    const int n_digits = int(sliced_.size());
    if (complete) *complete = true;
    if (max_count <= 0) return found;
    if (n_digits == 0) return {0};
    const int64_t D = int64_t(1) << dl_;
    // leaves checked once their lowest cut digit is set (digits are set from the
    // highest down, so found slices come in increasing order)
    std::vector<std::vector<int>> check_at(static_cast<size_t>(n_digits));
    for (int leaf = 0; leaf < n_leaves_; ++leaf) {
        int lo = n_digits;
        for (const auto& [axis, digit] : cut_[size_t(leaf)]) lo = std::min(lo, digit);
        if (lo < n_digits) check_at[size_t(lo)].push_back(leaf);
    }
    // wires tied by diagonal leaves (nonzero only where two axes agree: CZ, RZ, ...)
    // carry one value; a leaf that is zero for some value of an axis rules that
    // value out for the axis's class (a |0> state, an output bit)
    int n_wires = 0;
    for (const auto& w : leaf_wires_)
        for (int x : w) n_wires = std::max(n_wires, x + 1);
    std::vector<int> parent(static_cast<size_t>(n_wires));
    for (int i = 0; i < n_wires; ++i) parent[size_t(i)] = i;
    auto find = [&](int x) {
        while (parent[size_t(x)] != x) x = parent[size_t(x)] = parent[size_t(parent[size_t(x)])];
        return x;
    };
    std::vector<uint32_t> allowed(static_cast<size_t>(n_wires), (1u << D) - 1);
    for (int leaf = 0; leaf < n_leaves_; ++leaf) {
        const auto& w = leaf_wires_[size_t(leaf)];
        const auto& vals = leaf_vals_[size_t(leaf)];
        const int r = int(w.size());
        std::vector<uint32_t> mask(size_t(r), 0);
        std::vector<char> untied(size_t(r * r), 0);
        bool any = false;
        for (size_t idx = 0; idx < vals.size(); ++idx) {
            if (vals[idx] == std::complex<double>(0.0, 0.0)) continue;
            any = true;
            for (int a = 0; a < r; ++a) {
                const int64_t da = (int64_t(idx) >> (a * dl_)) & (D - 1);
                mask[size_t(a)] |= 1u << da;
                for (int b = a + 1; b < r; ++b)
                    if (da != ((int64_t(idx) >> (b * dl_)) & (D - 1))) untied[size_t(a * r + b)] = 1;
            }
        }
        for (int a = 0; a < r; ++a) allowed[size_t(w[size_t(a)])] &= mask[size_t(a)];
        if (any)
            for (int a = 0; a < r; ++a)
                for (int b = a + 1; b < r; ++b)
                    if (!untied[size_t(a * r + b)]) parent[size_t(find(w[size_t(a)]))] = find(w[size_t(b)]);
    }
    std::vector<uint32_t> class_allowed(static_cast<size_t>(n_wires), (1u << D) - 1);
    for (int x = 0; x < n_wires; ++x) class_allowed[size_t(find(x))] &= allowed[size_t(x)];
    // for each digit, the next higher digit whose wire is in the same class (or -1)
    std::vector<int> tied_above(static_cast<size_t>(n_digits), -1);
    for (int d = 0; d < n_digits; ++d)
        for (int e = d + 1; e < n_digits && tied_above[size_t(d)] < 0; ++e)
            if (find(sliced_[size_t(d)]) == find(sliced_[size_t(e)])) tied_above[size_t(d)] = e;
    std::vector<double> sub;
    auto leaf_nonzero = [&](int leaf, int64_t slice) {
        sub.assign(size_t(2) << (leaf_rank_[size_t(leaf)] * dl_), 0.0);
        Extract(leaf, slice, sub.data());
        for (double v : sub)
            if (v != 0.0) return true;
        return false;
    };
    int64_t visits = 0;
    // depth-first over digits n_digits-1 .. 0
    std::vector<int64_t> value(size_t(n_digits) + 1, 0);
    int digit = n_digits - 1;
    value[size_t(digit)] = -1;
    int64_t slice = 0;
    while (digit < n_digits && int64_t(found.size()) < max_count && visits < budget) {
        const int64_t shift = int64_t(digit) * dl_;
        slice &= ~((D - 1) << shift);
        if (++value[size_t(digit)] == D) {  // digit exhausted: back up
            ++digit;
            continue;
        }
        ++visits;
        slice |= value[size_t(digit)] << shift;
        if (!((class_allowed[size_t(find(sliced_[size_t(digit)]))] >> value[size_t(digit)]) & 1u)) continue;
        if (tied_above[size_t(digit)] >= 0 && value[size_t(tied_above[size_t(digit)])] != value[size_t(digit)]) continue;
        bool ok = true;
        for (int leaf : check_at[size_t(digit)])
            if (!leaf_nonzero(leaf, slice)) {
                ok = false;
                break;
            }
        if (!ok) continue;
        if (digit == 0) {
            found.push_back(slice);
            continue;
        }
        --digit;
        value[size_t(digit)] = -1;
    }
    // complete: every slice visited, or the list is full; otherwise the budget ran out
    if (complete) *complete = digit >= n_digits || int64_t(found.size()) >= max_count;
    return found;
}

std::complex<double> SlicedNetwork::Run(qt_ctx* ctx, int precision, SliceStats* stats) const {
    return RunWith(ctx, precision, stats, leaf_vals_, const_pool_);
}

std::complex<double> SlicedNetwork::Run(qt_ctx* ctx, int precision, SliceStats* stats, const LeafState& leaves) const {
    if (leaves.vals.size() != leaf_vals_.size() || leaves.const_pool.size() != const_pool_.size())
        throw InvalidFunctionInput();
    return RunWith(ctx, precision, stats, leaves.vals, leaves.const_pool);
}

std::complex<double> SlicedNetwork::RunWith(qt_ctx* ctx, int precision, SliceStats* stats,
                                            const std::vector<std::vector<std::complex<double>>>& vals,
                                            const std::vector<double>& const_pool) const {
    if (!ctx) throw DeviceFailure("SlicedNetwork::Run needs a device context");
    const int base = precision & ~QT_RESCALE;
    const bool c128 = base == QT_PLAIN_C128;
    if ((c128 && base != precision) ||
        (!c128 && base != QT_C64_4M && base != QT_C64_3M && base != QT_C64_BF16X3))
        throw InvalidFunctionInput();
    const int dtype = c128 ? QT_DTYPE_C128 : QT_DTYPE_C64;
    const auto t0 = std::chrono::steady_clock::now();
    const bool fuse = fuse_slices_ != 0 && !roots_.empty();
    const int n_roots = int(roots_.size());
    qt_tensor* acc = nullptr;
    qt_batch* batch = nullptr;
    std::vector<qt_tensor*> ops(size_t(n_leaves_) + steps_.size(), nullptr), outs;
    auto release = [&] {
        for (auto*& t : ops)
            if (t) {
                qt_tensor_free(ctx, t);
                t = nullptr;
            }
    };
    auto release_outs = [&] {
        for (auto*& t : outs)
            if (t) {
                qt_tensor_free(ctx, t);
                t = nullptr;
            }
    };
    try {
        CheckQt(ctx, qt_tensor_alloc(ctx, 0, dl_, QT_DTYPE_C128, &acc), "qt_tensor_alloc");
        int64_t const_elems = 0, chunk = 1;
        if (fuse) {
            std::vector<qt_program*> pp;
            for (const auto& p : progs_) pp.push_back(p.get());
            CheckQt(ctx, qt_batch_create(ctx, n_roots, pp.data(), base, &batch), "qt_batch_create");
            const_elems = std::max<int64_t>(1, int64_t(const_pool.size() / 2));
            std::vector<double> cp(const_pool);
            cp.resize(size_t(2 * const_elems), 0.0);
            CheckQt(ctx, qt_batch_write_leaves(batch, 0, cp.data(), const_elems), "qt_batch_write_leaves");
            // slices per launch: their subtree results stay within 256 MiB
            const int64_t per_slice = std::max<int64_t>(1, root_bytes_c128_ / (c128 ? 1 : 2));
            chunk = std::max<int64_t>(1, std::min<int64_t>({plan_.slices, 4096, (int64_t(256) << 20) / per_slice}));
            if (fuse_slices_ > 0) chunk = std::min(chunk, fuse_slices_);
        }
        // leaves that unfused steps read directly, uploaded per slice
        std::vector<int> direct;
        for (int i = 0; i < n_leaves_; ++i)
            if (!fuse || leaf_direct_[size_t(i)]) direct.push_back(i);
        std::vector<int> ranks;
        for (int i : direct) ranks.push_back(leaf_rank_[size_t(i)]);
        std::vector<std::vector<double>> host(direct.size());
        std::vector<std::vector<float>> host32(direct.size());
        std::vector<const void*> ptrs(direct.size());
        std::vector<qt_tensor*> made(direct.size());
        std::vector<double> block;
        // the slices contracted: the list, else the range (every slice by default)
        const bool listed = !slice_list_.empty();
        const int64_t s_begin = listed ? 0 : slice_first_;
        const int64_t s_end = listed ? int64_t(slice_list_.size())
                              : slice_count_ < 0 || slice_count_ > plan_.slices - slice_first_
                                  ? plan_.slices
                                  : slice_first_ + slice_count_;
        auto slice_id = [&](int64_t i) { return listed ? slice_list_[size_t(i)] : i; };
        for (int64_t s0 = s_begin; s0 < s_end; s0 += chunk) {
            const int64_t n_sl = std::min(chunk, s_end - s0);
            if (fuse) {
                // every subtree of every slice of the chunk in one launch
                if (block_elems_ > 0) {
                    block.assign(size_t(2 * n_sl * block_elems_), 0.0);
                    for (int64_t j = 0; j < n_sl; ++j)
                        for (const auto& [leaf, off] : param_leaves_)
                            Extract(vals, leaf, slice_id(s0 + j), block.data() + 2 * (j * block_elems_ + off));
                    CheckQt(ctx, qt_batch_write_leaves(batch, const_elems, block.data(), n_sl * block_elems_),
                            "qt_batch_write_leaves");
                }
                const int n_inst = int(n_sl) * n_roots;
                std::vector<int32_t> prog(static_cast<size_t>(n_inst));
                std::vector<int64_t> pbase(static_cast<size_t>(n_inst));
                outs.assign(size_t(n_inst), nullptr);
                for (int64_t j = 0; j < n_sl; ++j)
                    for (int r = 0; r < n_roots; ++r) {
                        const size_t i = size_t(j * n_roots + r);
                        prog[i] = r;
                        pbase[i] = const_elems + j * block_elems_;
                        CheckQt(ctx, qt_tensor_alloc_empty(ctx, steps_[size_t(roots_[size_t(r)])].rank_c, dl_, dtype,
                                                           &outs[i]),
                                "qt_tensor_alloc_empty");
                    }
                CheckQt(ctx, qt_batch_set_instances(batch, n_inst, prog.data(), pbase.data(), 0, nullptr, 0.0, 1.0),
                        "qt_batch_set_instances");
                CheckQt(ctx, qt_batch_set_outputs(batch, n_inst, outs.data()), "qt_batch_set_outputs");
                CheckQt(ctx, qt_batch_launch(batch), "qt_batch_launch");
            }
            for (int64_t j = 0; j < n_sl; ++j) {
                const int64_t slice = slice_id(s0 + j);
                for (size_t d = 0; d < direct.size(); ++d) {
                    const int leaf = direct[d];
                    const int64_t n = int64_t(1) << (leaf_rank_[size_t(leaf)] * dl_);
                    host[d].assign(size_t(2 * n), 0.0);
                    Extract(vals, leaf, slice, host[d].data());
                    if (c128) {
                        ptrs[d] = host[d].data();
                    } else {
                        host32[d].assign(host[d].begin(), host[d].end());
                        ptrs[d] = host32[d].data();
                    }
                }
                if (!direct.empty()) {
                    CheckQt(ctx, qt_tensor_create_many(ctx, int(direct.size()), ranks.data(), dl_, dtype, ptrs.data(),
                                                       made.data()),
                            "qt_tensor_create_many");
                    for (size_t d = 0; d < direct.size(); ++d) ops[size_t(direct[d])] = made[d];
                }
                if (fuse)
                    for (int r = 0; r < n_roots; ++r) {
                        auto*& t = outs[size_t(j * n_roots + r)];
                        ops[size_t(n_leaves_ + roots_[size_t(r)])] = t;
                        t = nullptr;
                    }
                for (size_t s = 0; s < steps_.size(); ++s) {
                    if (fuse && step_fused_[s]) continue;
                    const Step& st = steps_[s];
                    qt_tensor* c = nullptr;
                    CheckQt(ctx, qt_tensor_alloc_empty(ctx, st.rank_c, dl_, dtype, &c), "qt_tensor_alloc_empty");
                    ops[size_t(n_leaves_) + s] = c;
                    CheckQt(ctx,
                            qt_contract_pair(ctx, ops[size_t(st.a)], st.rank_a, ops[size_t(st.b)], st.rank_b, st.k,
                                             st.pair_a.data(), st.pair_b.data(), st.rank_c, st.free_src.data(), c,
                                             precision, nullptr),
                            "qt_contract_pair");
                    for (int op : {st.a, st.b}) {
                        qt_tensor_free(ctx, ops[size_t(op)]);
                        ops[size_t(op)] = nullptr;
                    }
                }
                CheckQt(ctx, qt_tensor_accumulate(ctx, ops.back(), acc), "qt_tensor_accumulate");
                release();
            }
        }
        double re = 0, im = 0;
        CheckQt(ctx, qt_tensor_read_scalar(ctx, acc, &re, &im), "qt_tensor_read_scalar");
        qt_tensor_free(ctx, acc);
        acc = nullptr;
        if (batch) qt_batch_destroy(batch);
        batch = nullptr;
        if (stats) {
            *stats = plan_;
            stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        return {re, im};
    } catch (...) {
        release();
        release_outs();
        if (acc) qt_tensor_free(ctx, acc);
        if (batch) qt_batch_destroy(batch);
        throw;
    }
}

}  // namespace qtorch
