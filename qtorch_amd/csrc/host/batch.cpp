// Batched evaluation on the device: networks are ordered on the host (the
// reference's orderings, or a greedy cost order), compiled into device programs
// and run many at a time by the persistent batch kernel (include/qtorch_hip.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <random>
#include <sstream>
#include <unordered_map>

#include "batch.hpp"

namespace qtorch {

void CheckQt(qt_ctx* ctx, qt_status s, const char* what) {
    if (s == QT_OK) return;
    if (s == QT_INVALID_INPUT) throw DeviceFailure(std::string(what) + ": invalid input: " + qt_last_error(ctx));
    if (s == QT_TIMEOUT) throw ContractionFailure();
    throw DeviceFailure(std::string(what) + ": " + qt_last_error(ctx));
}

namespace {

int SharedWires(const std::shared_ptr<Node>& a, const std::shared_ptr<Node>& b) {
    int n = 0;
    for (const auto& w : a->GetWires())
        if (w->GetNodeA().lock() == b || w->GetNodeB().lock() == b) ++n;
    return n;
}

}  // namespace

// Greedy order: repeatedly contract the connected pair whose result grows the
// stored size least (D^rC - D^rA - D^rB), then the one with fewer MACs; ties go to
// the lower node ids.  Scalars of disconnected components are multiplied last.
void GreedyContract(Network& net) {
    const int dl = net.GetDimLog2();
    while (!net.IsDone()) {
        const auto un = net.GetUncontractedNodes();
        std::shared_ptr<Node> best_a, best_b;
        long double best_cost = 0, best_flops = 0;
        for (const auto& a : un) {
            for (const auto& w : a->GetWires()) {
                if (w->IsContracted()) continue;
                auto b = w->GetNodeA().lock() == a ? w->GetNodeB().lock() : w->GetNodeA().lock();
                if (!b || b == a || b->mContracted) continue;
                const int k = SharedWires(a, b);
                const int rc = a->mRank + b->mRank - 2 * k;
                const long double cost = std::ldexp(1.0L, dl * rc) - std::ldexp(1.0L, dl * a->mRank) -
                                         std::ldexp(1.0L, dl * b->mRank);
                const long double flops = std::ldexp(1.0L, dl * (rc + k));
                const bool better = !best_a || cost < best_cost || (cost == best_cost && flops < best_flops) ||
                                    (cost == best_cost && flops == best_flops &&
                                     std::make_pair(a->mID, b->mID) < std::make_pair(best_a->mID, best_b->mID));
                if (better) {
                    best_a = a;
                    best_b = b;
                    best_cost = cost;
                    best_flops = flops;
                }
            }
        }
        if (!best_a) {  // only scalars of separate components are left
            std::vector<std::shared_ptr<Node>> sc;
            for (const auto& n : un)
                if (n->mRank == 0) sc.push_back(n);
            if (sc.size() < 2) throw ContractionFailure();
            best_a = sc[0];
            best_b = sc[1];
        }
        net.ContractNodes(best_a, best_b, 1 << 20);
    }
}

namespace {

// Second eigenvector of the weighted Laplacian L = D - W of an n-vertex graph (the
// Fiedler vector) by power iteration on c I - L with the constant vector projected
// out; c bounds L's spectrum (Gershgorin).  Deterministic start.
std::vector<double> FiedlerVector(int n, const std::vector<std::vector<std::pair<int, double>>>& adj) {
    std::vector<double> deg(static_cast<size_t>(n), 0.0), x(static_cast<size_t>(n)), y(static_cast<size_t>(n));
    double c = 1.0;
    for (int i = 0; i < n; ++i) {
        for (const auto& [j, w] : adj[size_t(i)]) deg[size_t(i)] += w;
        c = std::max(c, 2.0 * deg[size_t(i)]);
    }
    for (int i = 0; i < n; ++i) x[size_t(i)] = double(i) - 0.5 * double(n - 1) + 0.01 * std::sin(double(i));
    for (int it = 0; it < 20000; ++it) {
        double mean = 0.0;
        for (double v : x) mean += v;
        mean /= n;
        for (auto& v : x) v -= mean;
        for (int i = 0; i < n; ++i) {
            double lx = deg[size_t(i)] * x[size_t(i)];
            for (const auto& [j, w] : adj[size_t(i)]) lx -= w * x[size_t(j)];
            y[size_t(i)] = c * x[size_t(i)] - lx;
        }
        double norm = 0.0;
        for (double v : y) norm += v * v;
        norm = std::sqrt(norm);
        if (norm == 0.0) break;
        for (int i = 0; i < n; ++i) x[size_t(i)] = y[size_t(i)] / norm;
    }
    return x;
}

}  // namespace

// Circuit sweep order (no reference counterpart; the reference's orderings are
// ParallelContract and LGContract).  The qubits are arranged on a line by the
// Fiedler vector of their interaction graph (qubits joined by multi-qubit tensors),
// which puts few interactions across every cut of the line; each leaf gets the key
// (its qubit's place on that line, its place along the qubit's wire).  First every
// pair of connected tensors whose result is no larger than its larger operand and
// has at most 4 axes is merged (single-qubit gates into their neighbours; a cluster
// keeps its smallest key); then the clusters are absorbed one by one, in key order,
// into one growing tensor whose axes are the wires cut by the sweep line (a cluster
// not yet connected to it waits until it is).
void SweepContract(Network& net, uint64_t variant) {
    const auto& all = net.GetAllNodes();
    const int n_q = net.GetNumQubits(), n_orig = net.GetNumOriginalNodes();
    if (n_q < 1) throw InvalidFunctionInput();
    std::unordered_map<const Node*, int> orig_of;
    for (int i = 0; i < n_orig; ++i) orig_of[all[size_t(i)].get()] = i;
    auto qubit_of = [&](const Node* n) {
        const auto& wn = n->GetWireNumber();
        return wn.empty() ? 0 : wn[0];
    };
    // interaction graph: qubits of the two ends of every wire between leaves on
    // different qubits, and the qubits of each multi-qubit leaf
    std::vector<std::map<int, double>> w(static_cast<size_t>(n_q));
    auto link = [&](int a, int b) {
        if (a == b || a < 0 || b < 0 || a >= n_q || b >= n_q) return;
        w[size_t(a)][b] += 1.0;
        w[size_t(b)][a] += 1.0;
    };
    for (int i = 0; i < n_orig; ++i) {
        const Node* n = all[size_t(i)].get();
        const auto& wn = n->GetWireNumber();
        for (size_t j = 1; j < wn.size(); ++j) link(wn[0], wn[j]);
        for (const auto& wp : const_cast<Node*>(n)->GetWires()) {
            const auto a = wp->GetNodeA().lock(), b = wp->GetNodeB().lock();
            if (a && b) link(qubit_of(a.get()), qubit_of(b.get()));
        }
    }
    std::vector<std::vector<std::pair<int, double>>> adj(static_cast<size_t>(n_q));
    for (int q = 0; q < n_q; ++q)
        for (const auto& [r, x] : w[size_t(q)]) adj[size_t(q)].push_back({r, x});
    std::vector<double> fv = FiedlerVector(n_q, adj);
    if (variant > 0) {
        // another line: the Fiedler values perturbed by a seeded fraction of their
        // spread (neighbours on the line may swap), possibly reversed
        std::mt19937_64 rng(variant);
        std::normal_distribution<double> noise(0.0, 1.0);
        const auto [lo, hi] = std::minmax_element(fv.begin(), fv.end());
        const double scale = (*hi - *lo) / double(n_q) * (0.5 + double(variant % 4));
        const double sign = (variant / 4) % 2 ? -1.0 : 1.0;
        for (auto& v : fv) v = sign * v + scale * noise(rng);
    }
    std::vector<int> line(static_cast<size_t>(n_q));
    for (int q = 0; q < n_q; ++q) line[size_t(q)] = q;
    std::stable_sort(line.begin(), line.end(), [&](int a, int b) { return fv[size_t(a)] < fv[size_t(b)]; });
    std::vector<int> place(static_cast<size_t>(n_q));
    for (int i = 0; i < n_q; ++i) place[size_t(line[size_t(i)])] = i;
    // keys of the leaves: (place of the qubit, place along its wire)
    std::unordered_map<const Node*, std::pair<int, int>> key;
    const auto& by_wire = net.GetNodesByWire();
    for (int q = 0; q < n_q; ++q)
        for (size_t t = 0; t < by_wire[size_t(q)].size(); ++t) {
            const Node* n = by_wire[size_t(q)][t].get();
            const std::pair<int, int> k{place[size_t(q)], int(t)};
            auto it = key.find(n);
            if (it == key.end() || k < it->second) key[n] = k;
        }
    for (int i = 0; i < n_orig; ++i) key.emplace(all[size_t(i)].get(), std::make_pair(n_q, i));
    auto neighbours = [](const std::shared_ptr<Node>& a) {
        std::vector<std::shared_ptr<Node>> out;
        for (const auto& wp : a->GetWires()) {
            if (wp->IsContracted()) continue;
            auto b = wp->GetNodeA().lock() == a ? wp->GetNodeB().lock() : wp->GetNodeA().lock();
            if (b && b != a && !b->mContracted && std::find(out.begin(), out.end(), b) == out.end()) out.push_back(b);
        }
        return out;
    };
    auto contract = [&](const std::shared_ptr<Node>& a, const std::shared_ptr<Node>& b) {
        const auto k = std::min(key.at(a.get()), key.at(b.get()));
        auto c = net.ContractNodes(a, b, 1 << 20);
        if (c) key[c.get()] = k;
        return c;
    };
    // phase 1: merges that do not grow a tensor beyond its larger operand or 4 axes
    bool changed = true;
    while (changed && !net.IsDone()) {
        changed = false;
        auto un = net.GetUncontractedNodes();
        std::sort(un.begin(), un.end(), [&](const auto& x, const auto& y) { return key.at(x.get()) < key.at(y.get()); });
        for (const auto& a : un) {
            if (a->mContracted) continue;
            for (const auto& b : neighbours(a)) {
                const int k = SharedWires(a, b), rc = a->mRank + b->mRank - 2 * k;
                if (rc <= std::max(a->mRank, b->mRank) && rc <= 4) {
                    contract(a, b);
                    changed = true;
                    break;
                }
            }
            if (net.IsDone()) break;
        }
    }
    if (net.IsDone()) return;
    // phase 2: the sweep
    auto rest = net.GetUncontractedNodes();
    std::sort(rest.begin(), rest.end(), [&](const auto& x, const auto& y) { return key.at(x.get()) < key.at(y.get()); });
    std::shared_ptr<Node> acc = rest.front();
    std::vector<std::shared_ptr<Node>> waiting(rest.begin() + 1, rest.end());  // in key order
    while (!net.IsDone()) {
        bool progressed = false;
        for (size_t i = 0; i < waiting.size(); ++i) {
            if (SharedWires(acc, waiting[i]) == 0) continue;
            auto b = waiting[i];
            waiting.erase(waiting.begin() + long(i));
            auto c = contract(acc, b);
            if (net.IsDone()) return;
            acc = c;
            progressed = true;
            break;
        }
        if (progressed) continue;
        // only disconnected parts are left: contract them on their own, then the scalars
        if (waiting.empty()) throw ContractionFailure();
        std::shared_ptr<Node> other;
        for (const auto& x : waiting)
            if (x->mRank == 0) other = x;
        if (acc->mRank == 0 && other) {
            waiting.erase(std::find(waiting.begin(), waiting.end(), other));
            auto c = contract(acc, other);
            if (net.IsDone()) return;
            acc = c;
            continue;
        }
        // contract the first waiting cluster with its neighbours until it is a scalar
        auto x = waiting.front();
        auto nb = neighbours(x);
        auto it = std::find_if(waiting.begin() + 1, waiting.end(),
                               [&](const auto& y) { return std::find(nb.begin(), nb.end(), y) != nb.end(); });
        if (it == waiting.end()) throw ContractionFailure();
        auto y = *it;
        waiting.erase(it);
        waiting.erase(waiting.begin());
        auto c = contract(x, y);
        if (net.IsDone()) return;
        waiting.insert(waiting.begin(), c);
    }
}

int64_t LeafPool::Add(const std::vector<std::complex<double>>& v) {
    if (v.empty()) throw InvalidFunctionInput();  // a leaf without values (already contracted)
    std::string key(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(std::complex<double>));
    auto it = index_.find(key);
    if (it != index_.end()) return it->second;
    const int64_t off = int64_t(data_.size());
    data_.insert(data_.end(), v.begin(), v.end());
    index_.emplace(std::move(key), off);
    return off;
}

qt_program* CompileProgram(qt_ctx* ctx, const Network& net, const std::vector<int64_t>& leaf_refs) {
    const int n_leaves = net.GetNumOriginalNodes();
    const auto& all = net.GetAllNodes();
    const auto& plan = net.GetPlan();
    if (plan.empty() || int(leaf_refs.size()) != n_leaves) throw InvalidFunctionInput();
    std::vector<int32_t> ranks(static_cast<size_t>(n_leaves));
    for (int i = 0; i < n_leaves; ++i) ranks[size_t(i)] = all[size_t(i)]->mRank;
    std::vector<int32_t> desc, axes;
    for (const auto& st : plan) {
        desc.insert(desc.end(), {st.a_id, st.b_id, st.rank_a, st.rank_b, st.k, st.rank_c});
        axes.insert(axes.end(), st.pair_a.begin(), st.pair_a.end());
        axes.insert(axes.end(), st.pair_b.begin(), st.pair_b.end());
        axes.insert(axes.end(), st.free_src.begin(), st.free_src.end());
    }
    qt_program* p = nullptr;
    CheckQt(ctx, qt_program_create(ctx, net.GetDimLog2(), n_leaves, ranks.data(), leaf_refs.data(), int(plan.size()), desc.data(),
                                   axes.data(), &p),
            "qt_program_create");
    return p;
}

namespace {

// One pairwise contraction on the host in FP64 (the qt_contract_pair conventions:
// digit i of C is free axis free_src[i], digit j of the summed index is pair k-1-j).
std::vector<std::complex<double>> HostContract(const std::vector<std::complex<double>>& A,
                                               const std::vector<std::complex<double>>& B, const PlanStep& st,
                                               int dl) {
    const int k = st.k, rc = st.rank_c;
    std::vector<int> c_src, s_a, s_b;  // bit positions (B's bits + 64)
    for (int i = 0; i < rc; ++i) {
        const int32_t f = st.free_src[size_t(i)];
        const int ax = f & ~QT_FROM_A;
        for (int t = 0; t < dl; ++t) c_src.push_back((f & QT_FROM_A) ? ax * dl + t : 64 + ax * dl + t);
    }
    s_a.resize(size_t(k * dl));
    s_b.resize(size_t(k * dl));
    for (int i = 0; i < k; ++i)
        for (int t = 0; t < dl; ++t) {
            s_a[size_t((k - 1 - i) * dl + t)] = st.pair_a[size_t(i)] * dl + t;
            s_b[size_t((k - 1 - i) * dl + t)] = st.pair_b[size_t(i)] * dl + t;
        }
    const int64_t nC = int64_t(1) << (rc * dl), nS = int64_t(1) << (k * dl);
    std::vector<int64_t> ga(static_cast<size_t>(nS)), gb(static_cast<size_t>(nS));
    for (int64_t s = 0; s < nS; ++s) {
        int64_t a = 0, b = 0;
        for (int j = 0; j < k * dl; ++j)
            if ((s >> j) & 1) {
                a += int64_t(1) << s_a[size_t(j)];
                b += int64_t(1) << s_b[size_t(j)];
            }
        ga[size_t(s)] = a;
        gb[size_t(s)] = b;
    }
    std::vector<std::complex<double>> C(static_cast<size_t>(nC));
    for (int64_t c = 0; c < nC; ++c) {
        int64_t a = 0, b = 0;
        for (int j = 0; j < rc * dl; ++j)
            if ((c >> j) & 1) {
                if (c_src[size_t(j)] < 64) a += int64_t(1) << c_src[size_t(j)];
                else b += int64_t(1) << (c_src[size_t(j)] - 64);
            }
        std::complex<double> acc = 0.0;
        for (int64_t s = 0; s < nS; ++s) acc += A[size_t(a + ga[size_t(s)])] * B[size_t(b + gb[size_t(s)])];
        C[size_t(c)] = acc;
    }
    return C;
}

}  // namespace

qt_program* CompileProgramFolded(const Network& net, const std::vector<int64_t>& param_refs,
                                 const std::vector<std::vector<std::complex<double>>>& values, LeafPool& pool,
                                 int64_t max_fold_elems, int64_t* folded_cmacs) {
    const int n_leaves = net.GetNumOriginalNodes();
    const auto& all = net.GetAllNodes();
    const auto& plan = net.GetPlan();
    const int dl = net.GetDimLog2();
    if (plan.empty() || int(param_refs.size()) != n_leaves || int(values.size()) != n_leaves)
        throw InvalidFunctionInput();
    const int n_ops = n_leaves + int(plan.size());
    std::vector<char> cst(static_cast<size_t>(n_ops), 0);
    std::vector<std::vector<std::complex<double>>> val(static_cast<size_t>(n_ops));
    std::vector<int> rank(static_cast<size_t>(n_ops), 0);
    for (int i = 0; i < n_leaves; ++i) {
        cst[size_t(i)] = param_refs[size_t(i)] < 0;
        if (cst[size_t(i)]) val[size_t(i)] = values[size_t(i)];
        rank[size_t(i)] = all[size_t(i)]->mRank;
    }
    std::unordered_map<int, int> op_of_node;
    for (int i = 0; i < n_leaves; ++i) op_of_node[i] = i;
    std::vector<std::pair<int, int>> ops(plan.size());
    std::vector<char> kept(plan.size(), 1);
    int64_t folded = 0;
    for (size_t s = 0; s < plan.size(); ++s) {
        const auto& st = plan[s];
        const int a = op_of_node.at(st.a_id), b = op_of_node.at(st.b_id), c = n_leaves + int(s);
        ops[s] = {a, b};
        if (st.c_id >= 0) op_of_node[st.c_id] = c;
        rank[size_t(c)] = st.rank_c;
        const int64_t elems = int64_t(1) << (st.rank_c * dl);
        // constant folding: both operands known here, result small enough for a leaf
        if (s + 1 < plan.size() && cst[size_t(a)] && cst[size_t(b)] && elems <= max_fold_elems) {
            val[size_t(c)] = HostContract(val[size_t(a)], val[size_t(b)], st, dl);
            cst[size_t(c)] = 1;
            kept[s] = 0;
            folded += int64_t(1) << ((st.rank_c + st.k) * dl);
            val[size_t(a)].clear();
            val[size_t(b)].clear();
        }
    }
    // device leaves: parameter leaves and constants consumed by device steps
    std::vector<int> dev_id(static_cast<size_t>(n_ops), -1);
    std::vector<int32_t> ranks;
    std::vector<int64_t> refs;
    for (size_t s = 0; s < plan.size(); ++s) {
        if (!kept[s]) continue;
        for (int o : {ops[s].first, ops[s].second}) {
            const bool leaf_like = o < n_leaves || cst[size_t(o)];
            if (!leaf_like || dev_id[size_t(o)] >= 0) continue;
            dev_id[size_t(o)] = int(refs.size());
            ranks.push_back(rank[size_t(o)]);
            refs.push_back(cst[size_t(o)] ? pool.Add(val[size_t(o)]) : param_refs[size_t(o)]);
        }
    }
    const int n_dev_leaves = int(refs.size());
    std::vector<int32_t> desc, axes;
    int n_steps = 0;
    for (size_t s = 0; s < plan.size(); ++s) {
        if (!kept[s]) continue;
        dev_id[size_t(n_leaves) + s] = n_dev_leaves + n_steps++;
        const auto& st = plan[s];
        desc.insert(desc.end(), {dev_id[size_t(ops[s].first)], dev_id[size_t(ops[s].second)], st.rank_a, st.rank_b,
                                 st.k, st.rank_c});
        axes.insert(axes.end(), st.pair_a.begin(), st.pair_a.end());
        axes.insert(axes.end(), st.pair_b.begin(), st.pair_b.end());
        axes.insert(axes.end(), st.free_src.begin(), st.free_src.end());
    }
    if (folded_cmacs) *folded_cmacs = folded;
    qt_program* p = nullptr;
    CheckQt(nullptr,
            qt_program_create(nullptr, dl, n_dev_leaves, ranks.data(), refs.data(), n_steps, desc.data(), axes.data(), &p),
            "qt_program_create");
    return p;
}

void OrderNetwork(Network& net, int order, uint64_t seed, const std::vector<std::pair<int, int>>* given) {
    if (order == kOrderGreedy) {
        GreedyContract(net);
    } else if (order == kOrderStochastic) {
        std::shared_ptr<Network> alias(&net, [](Network*) {});
        ContractionTools ct(alias);
        ct.SetSeed(seed);
        ct.Contract(Stochastic);
    } else if (order == kOrderGiven) {
        if (!given) throw InvalidFunctionInput();
        for (const auto& p : *given) {
            const auto& all = net.GetAllNodes();
            if (p.first < 0 || p.second < 0 || size_t(p.first) >= all.size() || size_t(p.second) >= all.size())
                throw InvalidFunctionInput();
            net.ContractNodes(all[size_t(p.first)], all[size_t(p.second)], 100);
        }
        if (!net.IsDone()) throw ContractionFailure();
    } else if (order == kOrderSweep) {
        SweepContract(net, 0);
    } else if (order == kOrderLineGraph) {
        std::shared_ptr<Network> alias(&net, [](Network*) {});
        LineGraph lg(alias);
        lg.ContractInOrder(lg.MinFillOrder());
    } else {
        throw InvalidContractionMethod();
    }
}

// ---- QAOA batch --------------------------------------------------------------------

namespace {

std::vector<std::complex<double>> GateValues(const std::string& line) {
    NetworkOptions opt;  // plan-only: leaves keep their host values
    const std::string meas;
    Network n("1\n" + line + "\n", &meas, opt);
    return n.GetAllNodes()[1]->mHostVals;
}

std::string Fmt(double x) {
    std::ostringstream os;
    os << x;  // default ostream precision, as the reference writes angles (src/maxcut.h:211,217)
    return os.str();
}

}  // namespace

CompiledQaoa::~CompiledQaoa() {
    for (auto* p : progs) qt_program_destroy(p);
}

CompiledQaoa CompileQaoa(const ExtraData& data, int p, int order, uint64_t seed, LeafPool& pool) {
    CompiledQaoa out;
    const std::vector<double> placeholder(size_t(2 * p), 0.5);  // structure only: angles are parameters
    for (size_t e = 0; e < data.pairs.size(); ++e) {
        const std::string q = QaoaEdgeQasm(data, int(e), placeholder), m = QaoaEdgeMeasure(data, int(e));
        NetworkOptions opt;
        Network net(q, &m, opt);
        // leaves (before ordering: contraction clears a node's data): angle-dependent
        // Rz / Rx gates read the instance's parameter block, the rest the shared pool
        std::vector<int64_t> refs;
        std::vector<std::vector<std::complex<double>>> vals;
        int n_rz = 0, n_rx = 0;
        const int n_iter = int(data.realIterations[e].size()), nq = data.qubitsNeeded[e];
        for (int i = 0; i < net.GetNumOriginalNodes(); ++i) {
            const auto& node = net.GetAllNodes()[size_t(i)];
            vals.push_back(node->mHostVals);
            if (node->GetTypeOfNode() == GateType::RZ) {
                const int layer = n_iter ? n_rz++ / n_iter : 0;
                refs.push_back(QT_LEAF_PARAM | int64_t(32 * layer));
            } else if (node->GetTypeOfNode() == GateType::RX) {
                const int layer = n_rx++ / nq;
                refs.push_back(QT_LEAF_PARAM | int64_t(32 * layer + 16));
            } else {
                refs.push_back(-1);  // angle-free: folded with its neighbours where possible
            }
        }
        OrderNetwork(net, order, seed + e, nullptr);
        out.edge_cmacs.push_back(net.getNumFloatOps());
        out.cmacs += net.getNumFloatOps();
        int64_t folded = 0;
        out.progs.push_back(CompileProgramFolded(net, refs, vals, pool, int64_t(1) << 12, &folded));
        out.folded_cmacs += folded;
        int64_t scratch = 0;
        (void)qt_program_stats(out.progs.back(), nullptr, &scratch, nullptr);
        out.max_scratch = std::max(out.max_scratch, scratch);
    }
    return out;
}

std::vector<std::complex<double>> QaoaParamBlock(int p, const double* bg) {
    std::vector<std::complex<double>> block(size_t(32 * p));
    for (int i = 0; i < p; ++i) {
        const double beta = bg[i], gamma = bg[i + p];
        const auto rz = GateValues("Rz " + Fmt(-gamma) + " 0");     // src/maxcut.h:211
        const auto rx = GateValues("Rx " + Fmt(beta * 2.0) + " 0");  // src/maxcut.h:217
        std::copy(rz.begin(), rz.end(), block.begin() + 32 * i);
        std::copy(rx.begin(), rx.end(), block.begin() + 32 * i + 16);
    }
    return block;
}

QaoaBatch::QaoaBatch(qt_ctx* ctx, const std::string& dgf, int p, bool ref_map, int precision, int order,
                     uint64_t seed)
    : ctx_(ctx), data_(ExtraData::FromText(p, dgf, ref_map)), p_(p) {
    if (!ctx) throw DeviceFailure("QaoaBatch needs a device context");
    CompiledQaoa c = CompileQaoa(data_, p, order, seed, pool_);
    edge_cmacs_ = c.edge_cmacs;
    const_elems_ = int64_t(pool_.Data().size());
    CheckQt(ctx, qt_batch_create(ctx, int(c.progs.size()), c.progs.data(), precision, &batch_), "qt_batch_create");
}

QaoaBatch::~QaoaBatch() {
    if (batch_) qt_batch_destroy(batch_);
}

void QaoaBatch::SetAngles(int n_sets, const double* bg) {
    if (n_sets < 1 || !bg) throw InvalidFunctionInput();
    n_sets_ = n_sets;
    std::vector<std::complex<double>> pool = pool_.Data();
    const int64_t block = 32 * int64_t(p_);
    pool.reserve(size_t(const_elems_ + block * n_sets));
    for (int a = 0; a < n_sets; ++a) {
        const auto b = QaoaParamBlock(p_, bg + size_t(2 * p_) * size_t(a));
        pool.insert(pool.end(), b.begin(), b.end());
    }
    CheckQt(ctx_, qt_batch_set_leaves(batch_, reinterpret_cast<const double*>(pool.data()), int64_t(pool.size())),
            "qt_batch_set_leaves");
    const int ne = int(data_.pairs.size());
    std::vector<int32_t> prog, group;
    std::vector<int64_t> base;
    // edge-major: consecutive instances share a program, so a wavefront can run
    // several angle sets of one edge network side by side
    for (int e = 0; e < ne; ++e)
        for (int a = 0; a < n_sets; ++a) {
            prog.push_back(e);
            base.push_back(const_elems_ + block * a);
            group.push_back(a);
        }
    CheckQt(ctx_, qt_batch_set_instances(batch_, int(prog.size()), prog.data(), base.data(), n_sets, group.data(),
                                         0.5, -0.5),
            "qt_batch_set_instances");
}

void QaoaBatch::UpdateAngles(int n_sets, const double* bg, bool on_host) {
    if (n_sets != n_sets_ || n_sets < 1 || !bg) throw InvalidFunctionInput();
    // the parameter block of each set: layer i holds Rz(-gamma_i) at 32 i and
    // Rx(2 beta_i) at 32 i + 16, as QaoaParamBlock writes it
    std::vector<qt_rotation> rots;
    for (int i = 0; i < p_; ++i) {
        rots.push_back(qt_rotation{QT_ROT_RZ, p_ + i, -1.0, int64_t(32 * i)});
        rots.push_back(qt_rotation{QT_ROT_RX, i, 2.0, int64_t(32 * i + 16)});
    }
    CheckQt(ctx_, qt_batch_rotation_params(batch_, 2, int(rots.size()), rots.data(), n_sets, 2 * p_, bg,
                                           on_host ? 1 : 0, const_elems_, 32 * int64_t(p_)),
            "qt_batch_rotation_params");
}

void QaoaBatch::Results(double* energies, double* zz) {
    if (!zz) {  // energies only (also after group-sums-only launches)
        CheckQt(ctx_, qt_batch_results(batch_, nullptr, energies), "qt_batch_results");
        return;
    }
    std::vector<double> vals(size_t(2 * n_sets_) * data_.pairs.size());
    CheckQt(ctx_, qt_batch_results(batch_, vals.data(), energies), "qt_batch_results");
    const size_t ne = data_.pairs.size(), ns = size_t(n_sets_);
    if (zz)
        for (size_t e = 0; e < ne; ++e)
            for (size_t a = 0; a < ns; ++a) zz[a * ne + e] = vals[2 * (e * ns + a)];
}

int64_t QaoaBatch::CmacsPerEnergy() const {
    int64_t s = 0;
    for (auto c : edge_cmacs_) s += c;
    return s;
}

// ---- generic batch of networks -----------------------------------------------------

NetBatch::NetBatch(qt_ctx* ctx, const std::vector<std::string>& qasm, const std::vector<const std::string*>& meas,
                   const std::string& base_dir, int dim_log2, int order, uint64_t seed,
                   const std::vector<std::vector<std::pair<int, int>>>* given, int precision)
    : ctx_(ctx) {
    if (!ctx || qasm.size() != meas.size() || qasm.empty()) throw InvalidFunctionInput();
    std::vector<qt_program*> progs;
    try {
        for (size_t i = 0; i < qasm.size(); ++i) {
            NetworkOptions opt;
            opt.base_dir = base_dir;
            opt.dim_log2 = dim_log2;
            Network net(qasm[i], meas[i], opt);
            std::vector<int64_t> refs;  // before ordering: contraction clears a node's data
            for (int l = 0; l < net.GetNumOriginalNodes(); ++l) refs.push_back(pool_.Add(net.GetAllNodes()[size_t(l)]->mHostVals));
            OrderNetwork(net, order, seed + i, given ? &(*given)[i] : nullptr);
            cmacs_.push_back(net.getNumFloatOps());
            progs.push_back(CompileProgram(ctx, net, refs));
        }
    } catch (...) {
        for (auto* p : progs) qt_program_destroy(p);
        throw;
    }
    qt_status s = qt_batch_create(ctx, int(progs.size()), progs.data(), precision, &batch_);
    for (auto* p : progs) qt_program_destroy(p);
    CheckQt(ctx, s, "qt_batch_create");
    CheckQt(ctx, qt_batch_set_leaves(batch_, reinterpret_cast<const double*>(pool_.Data().data()),
                                     int64_t(pool_.Data().size())),
            "qt_batch_set_leaves");
    const int n = int(qasm.size());
    std::vector<int32_t> prog(static_cast<size_t>(n)), group(static_cast<size_t>(n), 0);
    std::vector<int64_t> base(size_t(n), 0);
    for (int i = 0; i < n; ++i) prog[size_t(i)] = i;
    CheckQt(ctx, qt_batch_set_instances(batch_, n, prog.data(), base.data(), 0, nullptr, 0.0, 1.0),
            "qt_batch_set_instances");
    n_ = n;
}

NetBatch::~NetBatch() {
    if (batch_) qt_batch_destroy(batch_);
}

void NetBatch::Results(std::complex<double>* values) {
    CheckQt(ctx_, qt_batch_results(batch_, reinterpret_cast<double*>(values), nullptr), "qt_batch_results");
}

}  // namespace qtorch
