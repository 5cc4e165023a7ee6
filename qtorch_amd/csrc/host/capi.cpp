// C ABI of the host library (include/qtorch_host.h).  Exceptions never cross it.
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

#include "batch.hpp"
#include "maxcut_driver.hpp"
#include "serve.hpp"
#include "sliced.hpp"
#include "qtorch.hpp"
#include "qtorch_host.h"

using namespace qtorch;

struct qth_network {
    std::shared_ptr<Network> net;
    std::unique_ptr<LineGraph> lg;
};
struct qth_qaoa_batch {
    std::unique_ptr<QaoaBatch> b;
};
struct qth_netbatch {
    std::unique_ptr<NetBatch> b;
};

namespace {

thread_local std::string g_err;

template <class F>
int Guard(F&& f) {
    try {
        f();
        g_err.clear();
        return QTH_OK;
    } catch (const InvalidFunctionInput& e) {
        g_err = e.what();
        return QTH_INVALID_FUNCTION_INPUT;
    } catch (const InvalidFile& e) {
        g_err = e.what();
        return QTH_INVALID_FILE;
    } catch (const InvalidFileFormat& e) {
        g_err = e.what();
        return QTH_INVALID_FILE_FORMAT;
    } catch (const ContractionFailure& e) {
        g_err = e.what();
        return QTH_CONTRACTION_FAILURE;
    } catch (const InvalidUserContractionSequence& e) {
        g_err = e.what();
        return QTH_INVALID_USER_SEQUENCE;
    } catch (const InvalidContractionMethod& e) {
        g_err = e.what();
        return QTH_INVALID_METHOD;
    } catch (const QbbFailure& e) {
        g_err = e.what();
        return QTH_QBB_FAILURE;
    } catch (const DeviceFailure& e) {
        g_err = e.what();
        return QTH_DEVICE;
    } catch (const std::invalid_argument& e) {  // std::stoi / std::stof on a bad token
        g_err = std::string("invalid number in input: ") + e.what();
        return QTH_INVALID_FILE_FORMAT;
    } catch (const std::out_of_range& e) {
        g_err = std::string("number out of range in input: ") + e.what();
        return QTH_INVALID_FILE_FORMAT;
    } catch (const std::exception& e) {
        g_err = e.what();
        return QTH_OTHER;
    }
}

void Out(const std::complex<double>& v, double* re, double* im) {
    if (re) *re = v.real();
    if (im) *im = v.imag();
}

std::vector<std::pair<int, int>> Pairs(const int32_t* p, int n) {
    std::vector<std::pair<int, int>> v;
    for (int i = 0; i < n; ++i) v.push_back({p[2 * i], p[2 * i + 1]});
    return v;
}

LineGraph& Lg(qth_network* h) {
    if (!h->lg) h->lg = std::make_unique<LineGraph>(h->net);
    return *h->lg;
}

}  // namespace

extern "C" {

const char* qth_last_error(void) { return g_err.c_str(); }

int qth_set_max_time(double seconds) {
    if (!(seconds == seconds)) return QTH_INVALID_FUNCTION_INPUT;
    maxTime = seconds;
    return QTH_OK;
}

double qth_get_max_time(void) { return maxTime; }

int qth_timer_start(void) {
    totTimer = Timer();
    totTimer.start();
    return QTH_OK;
}

int qth_timer_reset(void) {
    totTimer.reset();
    return QTH_OK;
}

double qth_timer_elapsed(void) { return totTimer.getElapsed(); }

int qth_preprocess(qt_ctx* ctx, const char* qasm, const char* measure, double time_threshold, uint64_t seed,
                   int32_t* pairs, int cap, int* n, int* found) {
    if (!qasm || !n || !found || cap < 0 || (cap > 0 && !pairs)) return QTH_INVALID_FUNCTION_INPUT;
    *n = 0;
    *found = 0;
    return Guard([&] {
        std::vector<std::pair<int, int>> seq;
        const std::string m = measure ? measure : "";
        *found = preProcess(std::string(qasm), measure ? &m : nullptr, ctx, seed, seq, time_threshold) ? 1 : 0;
        *n = int(seq.size());
        for (int i = 0; i < *n && i < cap; ++i) {
            pairs[2 * i] = seq[size_t(i)].first;
            pairs[2 * i + 1] = seq[size_t(i)].second;
        }
    });
}

int qth_network_create(qt_ctx* ctx, const char* qasm, const char* measure, const char* base_dir,
                       qth_network** out) {
    if (!qasm || !out) return QTH_INVALID_FUNCTION_INPUT;
    *out = nullptr;
    return Guard([&] {
        NetworkOptions opt;
        opt.ctx = ctx;
        opt.base_dir = base_dir ? base_dir : "";
        std::string m = measure ? measure : "";
        auto h = std::make_unique<qth_network>();
        h->net = std::make_shared<Network>(std::string(qasm), measure ? &m : nullptr, opt);
        *out = h.release();
    });
}

int qth_network_create_ex(qt_ctx* ctx, const char* qasm, const char* measure, const char* base_dir,
                          int dim_log2, qth_network** out) {
    if (!qasm || !out || (dim_log2 != 1 && dim_log2 != 2)) return QTH_INVALID_FUNCTION_INPUT;
    *out = nullptr;
    return Guard([&] {
        NetworkOptions opt;
        opt.ctx = ctx;
        opt.base_dir = base_dir ? base_dir : "";
        opt.dim_log2 = dim_log2;
        std::string m = measure ? measure : "";
        auto h = std::make_unique<qth_network>();
        h->net = std::make_shared<Network>(std::string(qasm), measure ? &m : nullptr, opt);
        *out = h.release();
    });
}

int qth_network_create_open(qt_ctx* ctx, const char* qasm, const char* measure, const char* base_dir, int dim_log2,
                            const int32_t* open_qubits, int n_open, qth_network** out) {
    if (!qasm || !out || (dim_log2 != 1 && dim_log2 != 2) || n_open < 0 || (n_open > 0 && !open_qubits))
        return QTH_INVALID_FUNCTION_INPUT;
    *out = nullptr;
    return Guard([&] {
        NetworkOptions opt;
        opt.ctx = ctx;
        opt.base_dir = base_dir ? base_dir : "";
        opt.dim_log2 = dim_log2;
        opt.open_qubits.assign(open_qubits, open_qubits + n_open);
        std::string m = measure ? measure : "";
        auto h = std::make_unique<qth_network>();
        h->net = std::make_shared<Network>(std::string(qasm), measure ? &m : nullptr, opt);
        *out = h.release();
    });
}

int qth_network_result_tensor(qth_network* h, double* values, int64_t cap_elems, int32_t* axis_qubits, int* rank) {
    if (!h || !rank || cap_elems < 0 || (cap_elems > 0 && !values)) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        std::vector<int> q;
        const auto v = h->net->GetFinalTensor(&q);
        *rank = int(q.size());
        if (axis_qubits)
            for (size_t i = 0; i < q.size(); ++i) axis_qubits[i] = q[i];
        for (int64_t i = 0; i < cap_elems && i < int64_t(v.size()); ++i) {
            values[2 * i] = v[size_t(i)].real();
            values[2 * i + 1] = v[size_t(i)].imag();
        }
    });
}

int qth_network_destroy(qth_network* h) {
    delete h;
    return QTH_OK;
}

int qth_network_reduce(qth_network* h) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { h->net->ReduceCircuit(); });
}

int qth_contract_nodes(qth_network* h, int a, int b, int threshold, int* result_id) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const auto& all = h->net->GetAllNodes();
        if (a < 0 || b < 0 || size_t(a) >= all.size() || size_t(b) >= all.size()) throw InvalidFunctionInput();
        auto r = h->net->ContractNodes(all[size_t(a)], all[size_t(b)], threshold);
        if (result_id) *result_id = r ? r->mID : -1;
    });
}

int qth_contract_stochastic(qth_network* h, uint64_t seed, double* re, double* im) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        ContractionTools ct(h->net);
        ct.SetSeed(seed);
        ct.Contract(Stochastic);
        Out(ct.GetFinalVal(), re, im);
    });
}

int qth_contract_given(qth_network* h, const int32_t* pairs, int n, double* re, double* im) {
    if (!h || (n > 0 && !pairs)) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        ContractionTools ct(h->net);
        ct.ContractGivenSequence(Pairs(pairs, n));
        Out(ct.GetFinalVal(), re, im);
    });
}

int qth_contract_user(qth_network* h, const int32_t* pairs, int n, double* re, double* im) {
    if (!h || (n > 0 && !pairs)) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        ContractionTools ct(h->net);
        ct.ContractUserDefinedSequence(Pairs(pairs, n));
        Out(ct.GetFinalVal(), re, im);
    });
}

int qth_network_order(qth_network* h, int order, uint64_t seed) {
    if (!h || order == QTH_ORDER_GIVEN) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { OrderNetwork(*h->net, order, seed, nullptr); });
}

int qth_linegraph_cnf(qth_network* h, char* buf, int64_t cap, int64_t* len) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const std::string s = Lg(h).CnfText();
        if (len) *len = int64_t(s.size());
        if (buf && cap > 0) {
            const size_t n = std::min(size_t(cap - 1), s.size());
            std::memcpy(buf, s.data(), n);
            buf[n] = 0;
        }
    });
}

int qth_linegraph_minfill(qth_network* h, int32_t* order, int cap, int* n) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const auto o = Lg(h).MinFillOrder();
        if (n) *n = int(o.size());
        for (int i = 0; i < int(o.size()) && i < cap; ++i) order[i] = o[size_t(i)];
    });
}

int qth_linegraph_contract(qth_network* h, const int32_t* order, int n, double* re, double* im) {
    if (!h || (n > 0 && !order)) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        std::vector<int> o(order, order + n);
        Out(Lg(h).ContractInOrder(o), re, im);
    });
}

int qth_record(qth_network* h, int32_t* pairs, int cap, int* n) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const auto r = RecordSequence(*h->net);
        if (n) *n = int(r.size());
        for (int i = 0; i < int(r.size()) && i < cap; ++i) {
            pairs[2 * i] = r[size_t(i)].first;
            pairs[2 * i + 1] = r[size_t(i)].second;
        }
    });
}

int64_t qth_flops(qth_network* h) { return h ? h->net->getNumFloatOps() : -1; }

int qth_network_value(qth_network* h, double* re, double* im) {
    if (!h || !re || !im) return QTH_INVALID_FUNCTION_INPUT;
    const auto v = h->net->GetFinalValue();
    *re = v.real();
    *im = v.imag();
    return QTH_OK;
}

int qth_counts(qth_network* h, int* all_nodes, int* uncontracted, int* original, int* done) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    if (all_nodes) *all_nodes = int(h->net->GetAllNodes().size());
    if (uncontracted) *uncontracted = int(h->net->GetUncontractedNodes().size());
    if (original) *original = h->net->GetNumOriginalNodes();
    if (done) *done = h->net->IsDone() ? 1 : 0;
    return QTH_OK;
}

int qth_plan(qth_network* h, int32_t* steps, int cap, int* n) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    const auto& plan = h->net->GetPlan();
    if (n) *n = int(plan.size());
    for (int i = 0; i < int(plan.size()) && i < cap; ++i) {
        const auto& s = plan[size_t(i)];
        int32_t* o = steps + 8 * i;
        o[0] = s.a_id;
        o[1] = s.b_id;
        o[2] = s.c_id;
        o[3] = s.rank_a;
        o[4] = s.rank_b;
        o[5] = s.k;
        o[6] = s.rank_c;
        o[7] = 0;
    }
    return QTH_OK;
}

int qth_program_stats(qth_network* h, int64_t* cmacs, int64_t* scratch_elems, int* n_steps) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const std::vector<int64_t> refs(static_cast<size_t>(h->net->GetNumOriginalNodes()), 0);
        qt_program* p = CompileProgram(nullptr, *h->net, refs);
        const qt_status s = qt_program_stats(p, cmacs, scratch_elems, n_steps);
        qt_program_destroy(p);
        if (s != QT_OK) throw InvalidFunctionInput();
    });
}

int qth_qaoa_num_edges(const char* dgf, int* n_edges, int* n_vertices) {
    if (!dgf) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        auto d = ExtraData::FromText(1, dgf, true);
        if (n_edges) *n_edges = int(d.pairs.size());
        if (n_vertices) *n_vertices = d.numQubits;
    });
}

int qth_qaoa_edge_text(const char* dgf, int p, int ref_map, int e, const double* bg, char* qasm,
                       int64_t qasm_cap, char* measure, int64_t measure_cap) {
    if (!dgf || !bg || p < 1) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        auto d = ExtraData::FromText(p, dgf, ref_map != 0);
        if (e < 0 || size_t(e) >= d.pairs.size()) throw InvalidFunctionInput();
        std::vector<double> ang(bg, bg + 2 * p);
        const std::string q = QaoaEdgeQasm(d, e, ang), m = QaoaEdgeMeasure(d, e);
        if (int64_t(q.size()) >= qasm_cap || int64_t(m.size()) >= measure_cap) throw InvalidFunctionInput();
        std::memcpy(qasm, q.c_str(), q.size() + 1);
        std::memcpy(measure, m.c_str(), m.size() + 1);
    });
}

int qth_qaoa_energy_plain(qt_ctx* ctx, const char* dgf, int p, int ref_map, const double* bg, uint64_t seed,
                          double* zz, double* energy) {
    if (!ctx || !dgf || !bg || p < 1) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        auto d = ExtraData::FromText(p, dgf, ref_map != 0);
        std::vector<double> ang(bg, bg + 2 * p), v;
        const double f = QaoaEnergyPlain(ctx, d, ang, seed, &v);
        if (zz) std::memcpy(zz, v.data(), v.size() * sizeof(double));
        if (energy) *energy = f;
    });
}

int qth_qaoa_batch_create(qt_ctx* ctx, const char* dgf, int p, int ref_map, int precision, int order,
                          uint64_t seed, qth_qaoa_batch** out) {
    if (!ctx || !dgf || !out || p < 1) return QTH_INVALID_FUNCTION_INPUT;
    *out = nullptr;
    return Guard([&] {
        auto h = std::make_unique<qth_qaoa_batch>();
        h->b = std::make_unique<QaoaBatch>(ctx, dgf, p, ref_map != 0, precision, order, seed);
        *out = h.release();
    });
}

int qth_qaoa_batch_set_angles(qth_qaoa_batch* h, int n_sets, const double* bg) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { h->b->SetAngles(n_sets, bg); });
}

int qth_qaoa_batch_update_angles(qth_qaoa_batch* h, int n_sets, const double* bg, int on_host) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { h->b->UpdateAngles(n_sets, bg, on_host != 0); });
}

int qth_qaoa_batch_results(qth_qaoa_batch* h, double* energies, double* zz) {
    if (!h) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { h->b->Results(energies, zz); });
}

qt_batch* qth_qaoa_batch_device(qth_qaoa_batch* h) { return h ? h->b->Batch() : nullptr; }

int64_t qth_qaoa_batch_cmacs(qth_qaoa_batch* h) { return h ? h->b->CmacsPerEnergy() : -1; }

int qth_qaoa_batch_destroy(qth_qaoa_batch* h) {
    delete h;
    return QTH_OK;
}

int qth_qaoa_source(const char* dgf, int p, int ref_map, int order, uint64_t seed, int precision, int flags,
                    int first_edge, int n_edges, char* buf, size_t cap, size_t* len) {
    if (!dgf || p < 1 || !len || first_edge < 0 || n_edges < 1) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const ExtraData data = ExtraData::FromText(p, dgf, ref_map != 0);
        LeafPool pool;
        CompiledQaoa c = CompileQaoa(data, p, order, seed, pool);
        if (first_edge + n_edges > int(c.progs.size())) throw InvalidFunctionInput();
        const std::vector<int64_t> base(size_t(n_edges), 0);
        const auto& d = pool.Data();
        CheckQt(nullptr,
                qt_program_source(n_edges, c.progs.data() + first_edge, base.data(),
                                  reinterpret_cast<const double*>(d.data()), int64_t(d.size()), precision, flags, buf,
                                  cap, len),
                "qt_program_source");
    });
}

int qth_qaoa_param_block(int p, const double* bg, double* out) {
    if (p < 1 || !bg || !out) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const auto b = QaoaParamBlock(p, bg);
        std::memcpy(out, b.data(), b.size() * sizeof(b[0]));
    });
}

int qth_qaoa_optimize(qt_ctx* ctx, const char* dgf, int p, int precision, int n_random, int max_evals,
                      uint64_t seed, double* angles, double* energy, int* evaluations) {
    if (!ctx || !dgf || !angles) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        auto r = MaxcutOptimalAngles(ctx, dgf, p, precision, n_random, max_evals, seed);
        std::memcpy(angles, r.angles.data(), r.angles.size() * sizeof(double));
        if (energy) *energy = r.energy;
        if (evaluations) *evaluations = r.evaluations;
    });
}

int qth_maxcut_final_string(qt_ctx* ctx, const char* dgf, int p, const double* bg, uint64_t seed, int* bits,
                            int* cut_edges) {
    if (!ctx || !dgf || !bg || !bits || p < 1) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const auto b = MaxcutFinalString(ctx, dgf, p, std::vector<double>(bg, bg + 2 * p), seed, cut_edges);
        std::memcpy(bits, b.data(), b.size() * sizeof(int));
    });
}

int qth_maxcut_final_string_ex(qt_ctx* ctx, const char* dgf, int p, const double* bg, uint64_t seed, int* bits,
                               int* cut_edges, double* cond_probs) {
    if (!ctx || !dgf || !bg || !bits || p < 1) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        std::vector<double> cp;
        const auto b = MaxcutFinalString(ctx, dgf, p, std::vector<double>(bg, bg + 2 * p), seed, cut_edges, &cp);
        std::memcpy(bits, b.data(), b.size() * sizeof(int));
        if (cond_probs) std::memcpy(cond_probs, cp.data(), cp.size() * sizeof(double));
    });
}

int qth_netbatch_create(qt_ctx* ctx, int n, const char* const* qasm, const char* const* measure,
                        const char* base_dir, int dim_log2, int order, uint64_t seed, const int32_t* given,
                        const int32_t* given_len, int precision, qth_netbatch** out) {
    if (!ctx || n < 1 || !qasm || !out || (dim_log2 != 1 && dim_log2 != 2)) return QTH_INVALID_FUNCTION_INPUT;
    *out = nullptr;
    return Guard([&] {
        std::vector<std::string> q, mstore(static_cast<size_t>(n));
        std::vector<const std::string*> m;
        for (int i = 0; i < n; ++i) {
            q.emplace_back(qasm[i]);
            if (measure && measure[i]) {
                mstore[size_t(i)] = measure[i];
                m.push_back(&mstore[size_t(i)]);
            } else {
                m.push_back(nullptr);
            }
        }
        std::vector<std::vector<std::pair<int, int>>> seqs;
        if (order == QTH_ORDER_GIVEN) {
            if (!given || !given_len) throw InvalidFunctionInput();
            size_t pos = 0;
            for (int i = 0; i < n; ++i) {
                seqs.push_back(Pairs(given + pos, given_len[i]));
                pos += size_t(2 * given_len[i]);
            }
        }
        auto h = std::make_unique<qth_netbatch>();
        h->b = std::make_unique<NetBatch>(ctx, q, m, base_dir ? base_dir : "", dim_log2, order, seed,
                                          order == QTH_ORDER_GIVEN ? &seqs : nullptr, precision);
        *out = h.release();
    });
}

qt_batch* qth_netbatch_device(qth_netbatch* h) { return h ? h->b->Batch() : nullptr; }

int qth_netbatch_results(qth_netbatch* h, double* values) {
    if (!h || !values) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { h->b->Results(reinterpret_cast<std::complex<double>*>(values)); });
}

int qth_netbatch_cmacs(qth_netbatch* h, int64_t* cmacs) {
    if (!h || !cmacs) return QTH_INVALID_FUNCTION_INPUT;
    const auto& c = h->b->Cmacs();
    for (size_t i = 0; i < c.size(); ++i) cmacs[i] = c[i];
    return QTH_OK;
}

int qth_netbatch_destroy(qth_netbatch* h) {
    delete h;
    return QTH_OK;
}

// ---- serving ---------------------------------------------------------------------

struct qth_server {
    std::unique_ptr<Server> s;
};

int qth_server_create(int device, int slots, int precision, int64_t hbm_budget, int max_requests_per_launch,
                      qth_server** out) {
    if (!out) return QTH_INVALID_FUNCTION_INPUT;
    *out = nullptr;
    return Guard([&] {
        ServeOptions o;
        o.device = device;
        o.slots = slots;
        o.precision = precision;
        o.hbm_budget = hbm_budget;
        o.max_requests_per_launch = max_requests_per_launch;
        auto h = std::make_unique<qth_server>();
        h->s = std::make_unique<Server>(o);
        *out = h.release();
    });
}

int qth_server_create_ex(int device, int slots, int precision, int eager_precision, int64_t hbm_budget,
                         int max_requests_per_launch, qth_server** out) {
    if (!out || (eager_precision != QT_PLAIN_C128 && eager_precision != QT_C64_4M && eager_precision != QT_C64_3M &&
                 eager_precision != QT_C64_BF16X3))
        return QTH_INVALID_FUNCTION_INPUT;
    *out = nullptr;
    return Guard([&] {
        ServeOptions o;
        o.device = device;
        o.slots = slots;
        o.precision = precision;
        o.eager_precision = eager_precision;
        o.hbm_budget = hbm_budget;
        o.max_requests_per_launch = max_requests_per_launch;
        auto h = std::make_unique<qth_server>();
        h->s = std::make_unique<Server>(o);
        *out = h.release();
    });
}

int qth_server_destroy(qth_server* s) {
    if (!s) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { delete s; });
}

int qth_server_submit_qaoa(qth_server* s, const char* dgf, int p, int ref_map, const double* bg, int64_t* id) {
    if (!s || !dgf || !bg || !id || p < 1) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        ServeRequest r;
        r.kind = kServeQaoaEnergy;
        r.text = dgf;
        r.p = p;
        r.ref_map = ref_map != 0;
        r.angles.assign(bg, bg + 2 * p);
        *id = s->s->Submit(std::move(r));
    });
}

int qth_server_submit_amplitude(qth_server* s, const char* qasm, const char* bits, int64_t* id) {
    if (!s || !qasm || !bits || !id) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        ServeRequest r;
        r.kind = kServeAmplitude;
        r.text = qasm;
        r.bits = bits;
        *id = s->s->Submit(std::move(r));
    });
}

int qth_server_poll(qth_server* s, int64_t id, int* done, int* status, double* re, double* im, double* latency_ms,
                    int* route) {
    if (!s || !done) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        ServeResult r;
        *done = s->s->TryGet(id, &r) ? 1 : 0;
        if (!*done) return;
        if (status) *status = r.status;
        if (re) *re = r.re;
        if (im) *im = r.im;
        if (latency_ms) *latency_ms = r.latency_ms;
        if (route) *route = r.route;
    });
}

int qth_server_wait(qth_server* s, int64_t id, double* re, double* im, double* latency_ms, int* route) {
    if (!s) return QTH_INVALID_FUNCTION_INPUT;
    ServeResult r;
    const int rc = Guard([&] { r = s->s->Wait(id); });
    if (rc != QTH_OK) return rc;
    if (re) *re = r.re;
    if (im) *im = r.im;
    if (latency_ms) *latency_ms = r.latency_ms;
    if (route) *route = r.route;
    return r.status;
}

int qth_server_drain(qth_server* s) {
    if (!s) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { s->s->Drain(); });
}

int qth_server_memory(qth_server* s, int64_t* resident_bytes, int64_t* slot_resets) {
    if (!s) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const ServeStats st = s->s->Stats();
        if (resident_bytes) *resident_bytes = st.resident_bytes;
        if (slot_resets) *slot_resets = st.slot_resets;
    });
}

int qth_server_stats(qth_server* s, int64_t* counts, double* ms, int64_t* peak) {
    if (!s) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        const ServeStats st = s->s->Stats();
        if (counts) {
            counts[0] = st.submitted;
            counts[1] = st.completed;
            counts[2] = st.rejected;
            counts[3] = st.launches;
            counts[4] = st.edge_evals;
        }
        if (ms) {
            ms[0] = st.p50_ms;
            ms[1] = st.p99_ms;
            ms[2] = st.mean_ms;
            ms[3] = st.overlap_ms;
            ms[4] = st.busy_ms;
        }
        if (peak) *peak = st.peak_planned_bytes;
    });
}

int qth_serve_plan(int kind, const char* text, int p, int ref_map, const char* bits, int precision, int* route,
                   int64_t* planned_bytes, int64_t* cmacs) {
    if (!text) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        ServeRequest r;
        r.kind = kind;
        r.text = text;
        r.p = p;
        r.ref_map = ref_map != 0;
        if (bits) r.bits = bits;
        ServeOptions o;
        o.precision = precision;
        const ServePlan pl = PlanRequest(r, o);
        if (route) *route = pl.route;
        if (planned_bytes) *planned_bytes = pl.planned_bytes;
        if (cmacs) *cmacs = pl.cmacs;
    });
}

// ---- sliced networks ------------------------------------------------------------

struct qth_sliced {
    std::unique_ptr<SlicedNetwork> n;
};

int qth_sliced_create(const char* qasm, const char* measure, int dim_log2, int order, uint64_t seed, int max_bits,
                      qth_sliced** out) {
    if (!qasm || !measure || !out) return QTH_INVALID_FUNCTION_INPUT;
    *out = nullptr;
    return Guard([&] {
        auto h = std::make_unique<qth_sliced>();
        h->n = std::make_unique<SlicedNetwork>(qasm, measure, dim_log2, order, seed, max_bits);
        *out = h.release();
    });
}

int qth_sliced_plan(qth_sliced* s, int64_t* v) {
    if (!s || !v) return QTH_INVALID_FUNCTION_INPUT;
    const SliceStats& p = s->n->Plan();
    v[0] = p.sliced_wires;
    v[1] = p.peak_bits;
    v[2] = p.slices;
    v[3] = p.cmacs_per_slice;
    v[4] = p.gemm_cmacs_per_slice;
    v[5] = p.peak_bytes;
    v[6] = p.fused_subtrees;
    v[7] = p.fused_steps;
    return QTH_OK;
}

int qth_sliced_set_fusion(qth_sliced* s, int64_t slices_per_launch) {
    if (!s) return QTH_INVALID_FUNCTION_INPUT;
    s->n->SetFusion(slices_per_launch);
    return QTH_OK;
}

int qth_sliced_set_range(qth_sliced* s, int64_t first, int64_t count) {
    if (!s) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { s->n->SetSliceRange(first, count); });
}

int qth_sliced_set_slices(qth_sliced* s, const int64_t* slices, int64_t n) {
    if (!s || !slices || n < 1) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] { s->n->SetSliceList(std::vector<int64_t>(slices, slices + n)); });
}

int qth_sliced_nonzero_slices_ex(qth_sliced* s, int64_t max_count, int64_t budget, int64_t* out, int64_t* n_out,
                                 int* complete) {
    if (!s || !out || !n_out || max_count < 1 || budget < 1) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        bool done = true;
        const auto v = s->n->NonzeroSlices(max_count, budget, &done);
        std::copy(v.begin(), v.end(), out);
        *n_out = int64_t(v.size());
        if (complete) *complete = done ? 1 : 0;
    });
}

int qth_sliced_nonzero_slices(qth_sliced* s, int64_t max_count, int64_t* out, int64_t* n_out) {
    int complete = 1;
    const int rc = qth_sliced_nonzero_slices_ex(s, max_count, int64_t(1) << 24, out, n_out, &complete);
    if (rc == QTH_OK && !complete) {
        g_err = "qth_sliced_nonzero_slices: the search budget ran out before max_count slices were found; the list "
                "written is incomplete (qth_sliced_nonzero_slices_ex takes a larger budget)";
        return QTH_OTHER;
    }
    return rc;
}

int qth_sliced_run(qth_sliced* s, qt_ctx* ctx, int precision, double* re, double* im, double* ms) {
    if (!s || !ctx) return QTH_INVALID_FUNCTION_INPUT;
    return Guard([&] {
        SliceStats st;
        const auto v = s->n->Run(ctx, precision, &st);
        if (re) *re = v.real();
        if (im) *im = v.imag();
        if (ms) *ms = st.ms;
    });
}

int qth_sliced_destroy(qth_sliced* s) {
    if (!s) return QTH_INVALID_FUNCTION_INPUT;
    delete s;
    return QTH_OK;
}

}  // extern "C"
