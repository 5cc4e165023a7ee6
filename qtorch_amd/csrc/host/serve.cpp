// Serving: continuous batching of QAOA-energy and single-amplitude requests.
//
// Requests arrive on any thread (Submit).  Their networks are planned once per
// structure (graph + depth, or circuit text) and kept compiled; a request only adds
// a small parameter block (its angles, or its output bitstring).  A loop thread
// drives `slots` HIP streams, each with two device batches (qt_batch): one for
// small programs (lane-grouped waves, intermediates in LDS) and one for programs
// with larger intermediates (workgroup executors, device scratch).  Whenever a
// stream is idle the loop admits the waiting requests, in arrival order, while
// their planned device bytes fit the budget, writes their parameters, and launches
// them; requests that arrive meanwhile go to the next idle stream, so the work of
// consecutive launches overlaps on the device.  Networks too large for a program
// (tensors above 2^16 elements) run on a separate thread and context through the
// large-network path in complex128 (sliced to a bound on index bits, small subtrees
// fused into one batch launch, MFMA GEMMs for the large steps).  Every request's
// result is computed by the same program, in the same order, whatever else shares
// its launch.
#include "serve.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <sstream>
#include <system_error>

#include "batch.hpp"
#include "qtorch_host.h"
#include "sliced.hpp"

namespace qtorch {

struct Server::Structure {
    int kind = 0, route = kRouteSmall, dim_log2 = 2, p = 1;
    ExtraData data;                       // QAOA light cones
    std::string text;                     // circuit (amplitude)
    int n_qubits = 0;
    std::vector<qt_program*> progs;       // batched routes: constant refs from 0
    std::vector<std::complex<double>> consts;
    int64_t param_elems = 0;              // per request
    int64_t scratch_elems = 0;            // largest program's
    int64_t cmacs = 0;                    // per request
    int64_t planned_bytes = 0;            // per request
    // large-network route: each network planned once (line-graph order, sliced, small
    // subtrees fused); a request runs it with its own leaf values
    std::vector<std::shared_ptr<SlicedNetwork>> eager_nets;
    // QAOA on that route: per edge network its leaf values for placeholder angles and
    // its angle-dependent leaves (leaf, 32 layer + 0 for Rz / 16 for Rx: the offset of
    // the gate's values in QaoaParamBlock); a request overwrites only those
    std::vector<SlicedNetwork::LeafState> eager_base;
    std::vector<std::vector<std::pair<int, int>>> eager_param_leaves;
    int eager_workers = 1;                // edge networks of a request run side by side
    std::string error;                    // set when the request cannot run
    ~Structure() {
        for (auto* p : progs) qt_program_destroy(p);
    }
};

struct Server::Pending {
    int64_t id = 0;
    ServeRequest req;
    std::shared_ptr<Structure> st;
    Clock::time_point t_submit;
    std::vector<std::complex<double>> params;
};

namespace {

constexpr int kBatchClasses = 2;  // small, medium

int64_t ElemBytes(int precision) { return precision == QT_PLAIN_C128 ? 16 : 8; }

// a + b * c, saturating at INT64_MAX (planned work of very large networks overflows)
int64_t SatMulAdd(int64_t a, int64_t b, int64_t c) {
    int64_t bc = 0, r = 0;
    if (__builtin_mul_overflow(b, c, &bc) || __builtin_add_overflow(a, bc, &r)) return INT64_MAX;
    return r;
}

std::shared_ptr<Server::Structure> BuildStructure(const ServeRequest& r, const ServeOptions& opt) {
    auto st = std::make_shared<Server::Structure>();
    st->kind = r.kind;
    const int64_t es = ElemBytes(opt.precision);
    try {
        if (r.kind == kServeQaoaEnergy) {
            if (r.p < 1) throw InvalidFunctionInput();
            st->p = r.p;
            st->data = ExtraData::FromText(r.p, r.text, r.ref_map);
            st->param_elems = 32 * int64_t(r.p);
            LeafPool pool;
            try {
                auto c = CompileQaoa(st->data, r.p, kOrderGreedy, 0, pool);
                st->progs = std::move(c.progs);
                st->cmacs = c.cmacs;
                st->scratch_elems = c.max_scratch;
                st->consts = pool.Data();
                st->route = st->scratch_elems <= opt.small_scratch ? kRouteSmall : kRouteMedium;
                // parameters, instance records and outputs; executors' intermediates live in
                // LDS (small route) or in the batch's fixed scratch area (medium route)
                st->planned_bytes = st->param_elems * es + 32 * int64_t(st->progs.size());
            } catch (const DeviceFailure&) {  // a light cone's tensors exceed a program's limit
                // the large-network path: each edge network planned once (line-graph
                // order, sliced to opt.eager_max_bits), its small subtrees fused
                st->route = kRouteEager;
                int64_t peak = 0;
                for (size_t e = 0; e < st->data.pairs.size(); ++e) {
                    const std::string q = QaoaEdgeQasm(st->data, int(e), std::vector<double>(size_t(2 * r.p), 0.5));
                    const std::string m = QaoaEdgeMeasure(st->data, int(e));
                    auto sn = std::make_shared<SlicedNetwork>(q, m, 2, kOrderLineGraph, 0, opt.eager_max_bits);
                    peak = std::max(peak, sn->Plan().peak_bytes);
                    {  // the angle-dependent leaves, numbered as CompileQaoa numbers their layers
                        NetworkOptions po;
                        Network net(q, &m, po);
                        std::vector<std::pair<int, int>> pl;
                        const int n_iter = int(st->data.realIterations[e].size()), nq = st->data.qubitsNeeded[e];
                        int n_rz = 0, n_rx = 0;
                        for (int i = 0; i < net.GetNumOriginalNodes(); ++i) {
                            const auto ty = net.GetAllNodes()[size_t(i)]->GetTypeOfNode();
                            if (ty == GateType::RZ) pl.push_back({i, 32 * (n_iter ? n_rz++ / n_iter : 0)});
                            else if (ty == GateType::RX) pl.push_back({i, 32 * (n_rx++ / nq) + 16});
                        }
                        st->eager_param_leaves.push_back(std::move(pl));
                        st->eager_base.push_back(sn->MakeLeafState(q, m));
                    }
                    st->cmacs = SatMulAdd(st->cmacs, sn->Plan().cmacs_per_slice, sn->Plan().slices);
                    st->eager_nets.push_back(std::move(sn));
                    if (st->cmacs > opt.max_request_cmacs) break;  // refused at admission anyway
                }
                // edge networks run on up to eager_workers streams at once, as many as
                // the budget holds (at least one)
                const int64_t per_loop = std::max(1, opt.eager_workers / std::max(1, opt.eager_loops));
                const int64_t fit = peak > 0 ? opt.hbm_budget / peak : per_loop;
                st->eager_workers = int(std::max<int64_t>(
                    1, std::min<int64_t>({fit, per_loop, int64_t(st->eager_nets.size())})));
                st->planned_bytes = SatMulAdd(0, peak, st->eager_workers);
            }
        } else if (r.kind == kServeAmplitude) {
            st->text = r.text;
            st->dim_log2 = 1;
            NetworkOptions o;
            o.dim_log2 = 1;
            const std::string zeros(r.bits.size(), '0');
            Network net(r.text, &zeros, o);
            if (net.GetNumQubits() != int(r.bits.size())) throw InvalidFunctionInput();
            st->n_qubits = int(r.bits.size());
            // the output bitstring's <0| / <1| leaves are the request's parameters
            const int n_orig = net.GetNumOriginalNodes();
            std::vector<int64_t> refs;
            std::vector<std::vector<std::complex<double>>> vals;
            LeafPool pool;
            for (int i = 0; i < n_orig; ++i) {
                const int m = i - (n_orig - st->n_qubits);
                refs.push_back(m >= 0 ? (QT_LEAF_PARAM | int64_t(2 * m)) : -1);
                vals.push_back(net.GetAllNodes()[size_t(i)]->mHostVals);
            }
            st->param_elems = 2 * int64_t(st->n_qubits);
            GreedyContract(net);
            st->cmacs = net.getNumFloatOps();
            try {
                // the circuit's bitstring-free parts are contracted once, here
                qt_program* p = CompileProgramFolded(net, refs, vals, pool, int64_t(1) << 12, nullptr);
                st->progs.push_back(p);
                int64_t scratch = 0;
                (void)qt_program_stats(p, nullptr, &scratch, nullptr);
                st->scratch_elems = scratch;
                st->consts = pool.Data();
                st->route = scratch <= opt.small_scratch ? kRouteSmall : kRouteMedium;
                st->planned_bytes = st->param_elems * es + 32;
            } catch (const DeviceFailure&) {
                // the large-network path: the line-graph order sliced to eager_max_bits, or
                // the circuit sweep (CZ split into bonded halves) sliced to eager_max_bits
                // or up to two bits more when the budget holds it, whichever plans the
                // least work
                st->route = kRouteEager;
                std::shared_ptr<SlicedNetwork> best;
                int64_t best_work = INT64_MAX;
                auto consider = [&](int order, int max_bits) {
                    try {
                        auto sn = std::make_shared<SlicedNetwork>(r.text, zeros, 1, order, 0, max_bits);
                        const int64_t work = SatMulAdd(0, sn->Plan().cmacs_per_slice, sn->Plan().slices);
                        if (sn->Plan().peak_bytes <= opt.hbm_budget / 2 && (!best || work < best_work)) {
                            best = sn;
                            best_work = work;
                        }
                    } catch (const InvalidFunctionInput&) {  // more slices than a plan allows
                    }
                };
                consider(kOrderLineGraph, opt.eager_max_bits);
                for (int extra = 0; extra <= 2; ++extra) consider(kOrderSweep, opt.eager_max_bits + extra);
                if (!best) best = std::make_shared<SlicedNetwork>(r.text, zeros, 1, kOrderLineGraph, 0, opt.eager_max_bits);
                st->planned_bytes = best->Plan().peak_bytes;
                st->cmacs = SatMulAdd(0, best->Plan().cmacs_per_slice, best->Plan().slices);
                st->eager_nets.push_back(std::move(best));
            }
        } else {
            throw InvalidFunctionInput();
        }
    } catch (const std::exception& e) {
        st->error = e.what();
    }
    return st;
}

int StatusOf(const std::string& error) { return error.empty() ? QTH_OK : QTH_INVALID_FUNCTION_INPUT; }

// an output bitstring is one '0' or '1' per qubit (the structure checks the count)
bool ValidBits(const std::string& bits) {
    for (char ch : bits)
        if (ch != '0' && ch != '1') return false;
    return true;
}

}  // namespace

// One HIP stream with its two batches and its launch in flight.
struct Server::Slot {
    qt_ctx* ctx = nullptr;
    qt_batch* batch[kBatchClasses] = {nullptr, nullptr};
    // per batch: program ids and constant offset of each structure added to it
    std::map<const Structure*, int32_t> first_prog[kBatchClasses];
    std::vector<std::shared_ptr<Structure>> keep;  // structures whose programs it holds
    int64_t const_next[kBatchClasses] = {0, 0};
    bool busy = false;
    std::vector<std::shared_ptr<Pending>> running;
    int64_t running_bytes = 0;
    int64_t resident = 0;  // device bytes of the programs and constants its batches hold
    // per batch: instance -> (request index in `running`, edge)
    std::vector<std::pair<int, int>> inst_of[kBatchClasses];
    int n_groups[kBatchClasses] = {0, 0};
    std::vector<int> group_of_req[kBatchClasses];  // request index -> group (or -1)
};

Server::Server(const ServeOptions& opt) : opt_(opt), t0_(Clock::now()) {
    if (opt_.slots < 1 || opt_.max_requests_per_launch < 1 || opt_.hbm_budget < 1 || opt_.param_capacity < 64 ||
        opt_.eager_workers < 1 || opt_.eager_workers > 16 || opt_.eager_loops < 1 ||
        opt_.eager_loops > opt_.eager_workers || opt_.max_structures < 1 || opt_.resident_limit < 1)
        throw InvalidFunctionInput();
    // the slots' resident programs stay within a quarter of the budget
    opt_.resident_limit = std::min(opt_.resident_limit, std::max<int64_t>(1, opt_.hbm_budget / (4 * opt_.slots)));
    try {
        for (int i = 0; i < opt_.slots; ++i) {
            auto s = std::make_unique<Slot>();
            if (qt_ctx_create(opt_.device, &s->ctx) != QT_OK) throw DeviceFailure("serve: no device context");
            for (int c = 0; c < kBatchClasses; ++c) {
                CheckQt(s->ctx, qt_batch_create(s->ctx, 0, nullptr, opt_.precision, &s->batch[c]), "qt_batch_create");
                s->const_next[c] = opt_.param_capacity;  // [0, param_capacity): parameters of one launch
                // small: lane-grouped waves; medium: one workgroup per network, scratch in device memory
                CheckQt(s->ctx,
                        c == kRouteSmall ? qt_batch_config_ex(s->batch[c], opt_.small_lanes, opt_.small_wgs, QT_SCRATCH_AUTO)
                                         : qt_batch_config_ex(s->batch[c], 256, 1, QT_SCRATCH_GLOBAL),
                        "qt_batch_config_ex");
            }
            slots_.push_back(std::move(s));
        }
        for (int w = 0; w < opt_.eager_workers; ++w) {
            qt_ctx* c = nullptr;
            if (qt_ctx_create(opt_.device, &c) != QT_OK) throw DeviceFailure("serve: no device context");
            eager_ctxs_.push_back(c);
        }
    } catch (...) {
        for (auto& s : slots_) {
            for (auto* b : s->batch)
                if (b) qt_batch_destroy(b);
            if (s->ctx) qt_ctx_destroy(s->ctx);
        }
        for (auto* c : eager_ctxs_) qt_ctx_destroy(c);
        throw;
    }
    loop_ = std::thread([this] { Loop(); });
    for (int k = 0; k < opt_.eager_loops; ++k) eager_threads_.emplace_back([this, k] { EagerLoop(k); });
}

Server::~Server() {
    {
        std::lock_guard<std::mutex> g(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    if (loop_.joinable()) loop_.join();
    for (auto& t : eager_threads_)
        if (t.joinable()) t.join();
    for (auto& s : slots_) {
        for (auto* b : s->batch)
            if (b) qt_batch_destroy(b);
        qt_ctx_destroy(s->ctx);
    }
    for (auto* c : eager_ctxs_) qt_ctx_destroy(c);
}

std::shared_ptr<Server::Structure> Server::GetStructure(const ServeRequest& r) {
    std::ostringstream os;
    os << r.kind << '|' << r.p << '|' << r.ref_map << '|' << r.bits.size() << '|' << r.text;
    const std::string key = os.str();
    std::promise<std::shared_ptr<Structure>> built;
    std::shared_future<std::shared_ptr<Structure>> fut;
    bool build = false;
    {
        std::lock_guard<std::mutex> g(struct_mu_);
        auto it = structures_.find(key);
        if (it == structures_.end()) {
            if (int(structures_.size()) >= opt_.max_structures) {  // drop the least recently used ready one
                auto lru = structures_.end();
                for (auto jt = structures_.begin(); jt != structures_.end(); ++jt)
                    if (jt->second.fut.wait_for(std::chrono::seconds(0)) == std::future_status::ready &&
                        (lru == structures_.end() || jt->second.last_use < lru->second.last_use))
                        lru = jt;
                if (lru != structures_.end()) structures_.erase(lru);
            }
            fut = built.get_future().share();
            it = structures_.emplace(key, CachedStructure{fut, 0}).first;
            build = true;
        }
        it->second.last_use = ++struct_tick_;
        fut = it->second.fut;
    }
    if (build) {  // planning runs unlocked: cached structures stay available meanwhile
        try {
            built.set_value(BuildStructure(r, opt_));
        } catch (...) {
            built.set_exception(std::current_exception());
            std::lock_guard<std::mutex> g(struct_mu_);
            structures_.erase(key);
        }
    }
    return fut.get();
}

ServePlan PlanRequest(const ServeRequest& r, const ServeOptions& opt) {
    if (r.kind == kServeAmplitude && !ValidBits(r.bits)) throw InvalidFunctionInput();
    auto st = BuildStructure(r, opt);
    if (!st->error.empty()) throw InvalidFunctionInput();
    ServePlan p;
    p.route = st->route;
    p.planned_bytes = st->planned_bytes;
    p.cmacs = st->cmacs;
    return p;
}

int64_t Server::Submit(ServeRequest r) {
    auto p = std::make_shared<Pending>();
    p->t_submit = Clock::now();
    if (r.kind != kServeAmplitude || ValidBits(r.bits)) p->st = GetStructure(r);
    // the request's parameter block
    if (p->st && p->st->error.empty()) {
        if (r.kind == kServeQaoaEnergy) {
            if (int(r.angles.size()) != 2 * r.p) p->st = nullptr;
            else if (p->st->route != kRouteEager) p->params = QaoaParamBlock(r.p, r.angles.data());
        } else {
            for (char ch : r.bits) {
                p->params.push_back(ch == '0' ? 1.0 : 0.0);
                p->params.push_back(ch == '1' ? 1.0 : 0.0);
            }
        }
    }
    p->req = std::move(r);
    std::unique_lock<std::mutex> lk(mu_);
    p->id = next_id_++;
    ++stats_.submitted;
    ++outstanding_;
    const bool too_big = p->st && (p->st->planned_bytes > opt_.hbm_budget ||
                                   (p->st->route == kRouteEager && p->st->cmacs > opt_.max_request_cmacs));
    if (!p->st || !p->st->error.empty() || too_big) {
        ServeResult res;
        res.status = !p->st ? QTH_INVALID_FUNCTION_INPUT : !p->st->error.empty() ? StatusOf(p->st->error) : QTH_OTHER;
        res.error = !p->st                   ? "bad angles or bitstring"
                    : !p->st->error.empty()  ? p->st->error
                    : p->st->planned_bytes > opt_.hbm_budget ? "planned device memory exceeds the budget"
                                                             : "planned work exceeds the per-request limit";
        if (p->st) res.planned_bytes = p->st->planned_bytes;
        ++stats_.rejected;
        lk.unlock();
        Complete(*p, res);
        return p->id;
    }
    if (p->st->route == kRouteEager) eager_queue_.push_back(p);
    else queue_.push_back(p);
    lk.unlock();
    cv_.notify_all();
    return p->id;
}

void Server::Complete(Pending& p, ServeResult res) {
    const double ms = std::chrono::duration<double, std::milli>(Clock::now() - p.t_submit).count();
    res.latency_ms = ms;
    {
        std::lock_guard<std::mutex> g(mu_);
        if (res.status == QTH_OK) {
            latencies_.push_back(ms);
            ++stats_.completed;
            if (p.req.kind == kServeQaoaEnergy && p.st) stats_.edge_evals += int64_t(p.st->data.pairs.size());
        }
        results_[p.id] = res;
        --outstanding_;
    }
    done_cv_.notify_all();
}

bool Server::TryGet(int64_t id, ServeResult* out) {
    std::lock_guard<std::mutex> g(mu_);
    auto it = results_.find(id);
    if (it == results_.end()) return false;
    if (out) *out = it->second;
    results_.erase(it);
    return true;
}

ServeResult Server::Wait(int64_t id) {
    std::unique_lock<std::mutex> lk(mu_);
    if (id < 1 || id >= next_id_) throw InvalidFunctionInput();
    done_cv_.wait(lk, [&] { return results_.count(id) > 0; });
    ServeResult r = results_[id];
    results_.erase(id);
    return r;
}

void Server::Drain() {
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return outstanding_ == 0; });
}

ServeStats Server::Stats() {
    std::lock_guard<std::mutex> g(mu_);
    ServeStats s = stats_;
    s.resident_bytes = resident_bytes_;
    if (!latencies_.empty()) {
        std::vector<double> v = latencies_;
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            const double pos = q * double(v.size() - 1);
            const size_t lo = size_t(pos), hi = std::min(lo + 1, v.size() - 1);
            return v[lo] + (v[hi] - v[lo]) * (pos - double(lo));
        };
        s.p50_ms = pct(0.50);
        s.p99_ms = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        s.mean_ms = sum / double(v.size());
    }
    return s;
}

// ---- the loop ------------------------------------------------------------------------

// An idle slot whose resident programs and constants exceed the limit drops them all:
// its batches are created anew and structures are added again on their next use.
void Server::ResetSlot(Slot& s) {
    for (int c = 0; c < kBatchClasses; ++c) {
        if (s.batch[c]) qt_batch_destroy(s.batch[c]);
        s.batch[c] = nullptr;
        s.first_prog[c].clear();
        s.const_next[c] = opt_.param_capacity;
    }
    s.keep.clear();
    (void)qt_ctx_trim(s.ctx);
    for (int c = 0; c < kBatchClasses; ++c) {
        CheckQt(s.ctx, qt_batch_create(s.ctx, 0, nullptr, opt_.precision, &s.batch[c]), "qt_batch_create");
        CheckQt(s.ctx,
                c == kRouteSmall ? qt_batch_config_ex(s.batch[c], opt_.small_lanes, opt_.small_wgs, QT_SCRATCH_AUTO)
                                 : qt_batch_config_ex(s.batch[c], 256, 1, QT_SCRATCH_GLOBAL),
                "qt_batch_config_ex");
    }
    std::lock_guard<std::mutex> g(mu_);
    resident_bytes_ -= s.resident;
    s.resident = 0;
    ++stats_.slot_resets;
}

bool Server::StartIteration(Slot& s) {
    if (s.resident > opt_.resident_limit) ResetSlot(s);
    std::vector<std::shared_ptr<Pending>> take;
    int64_t bytes = 0;
    {
        std::lock_guard<std::mutex> g(mu_);
        int64_t params[kBatchClasses] = {0, 0};
        while (!queue_.empty() && int(take.size()) < opt_.max_requests_per_launch) {
            auto& p = queue_.front();
            const int cls = p->st->route;
            // resident programs and constants count against the budget too
            if (in_flight_bytes_ + resident_bytes_ + bytes + p->st->planned_bytes > opt_.hbm_budget) break;
            if (params[cls] + int64_t(p->params.size()) > opt_.param_capacity) break;
            params[cls] += int64_t(p->params.size());
            bytes += p->st->planned_bytes;
            take.push_back(p);
            queue_.pop_front();
        }
        if (take.empty()) return false;
        in_flight_bytes_ += bytes;
        stats_.peak_planned_bytes = std::max(stats_.peak_planned_bytes, in_flight_bytes_);
    }
    s.running = take;
    s.running_bytes = bytes;
    try {
        for (int cls = 0; cls < kBatchClasses; ++cls) {
            qt_batch* b = s.batch[cls];
            s.inst_of[cls].clear();
            s.group_of_req[cls].assign(take.size(), -1);
            s.n_groups[cls] = 0;
            std::vector<double> params;
            // (program, request, edge, parameter base) of every instance
            struct Inst {
                int32_t prog;
                int req, edge;
                int64_t base;
            };
            std::vector<Inst> inst;
            for (size_t r = 0; r < take.size(); ++r) {
                auto& p = *take[r];
                if (p.st->route != cls) continue;
                auto it = s.first_prog[cls].find(p.st.get());
                if (it == s.first_prog[cls].end()) {
                    // first request of this structure on this batch: its constants and programs
                    const int64_t base = s.const_next[cls];
                    if (!p.st->consts.empty())
                        CheckQt(s.ctx, qt_batch_write_leaves(b, base, reinterpret_cast<const double*>(p.st->consts.data()),
                                                             int64_t(p.st->consts.size())),
                                "qt_batch_write_leaves");
                    s.const_next[cls] += std::max<int64_t>(int64_t(p.st->consts.size()), 1);
                    std::vector<int64_t> bases(p.st->progs.size(), base);
                    int32_t first = 0;
                    CheckQt(s.ctx, qt_batch_add_programs(b, int(p.st->progs.size()), p.st->progs.data(), bases.data(),
                                                         &first),
                            "qt_batch_add_programs");
                    it = s.first_prog[cls].emplace(p.st.get(), first).first;
                    s.keep.push_back(p.st);
                    int64_t res = 0;
                    for (auto* bb : s.batch) res += qt_batch_resident_bytes(bb);
                    std::lock_guard<std::mutex> g(mu_);
                    resident_bytes_ += res - s.resident;
                    s.resident = res;
                }
                const int64_t pbase = int64_t(params.size() / 2);
                params.insert(params.end(), reinterpret_cast<const double*>(p.params.data()),
                              reinterpret_cast<const double*>(p.params.data()) + 2 * p.params.size());
                if (p.req.kind == kServeQaoaEnergy) s.group_of_req[cls][r] = s.n_groups[cls]++;
                for (size_t e = 0; e < p.st->progs.size(); ++e)
                    inst.push_back({it->second + int32_t(e), int(r), int(e), pbase});
            }
            if (inst.empty()) continue;  // this batch is not launched (its last launch has finished)
            // instances of one program next to each other (lane-grouped waves share them)
            std::stable_sort(inst.begin(), inst.end(), [](const Inst& a, const Inst& b) { return a.prog < b.prog; });
            std::vector<int32_t> prog, group;
            std::vector<int64_t> base;
            for (const auto& x : inst) {
                prog.push_back(x.prog);
                base.push_back(x.base);
                group.push_back(s.group_of_req[cls][size_t(x.req)]);
                s.inst_of[cls].push_back({x.req, x.edge});
            }
            if (!params.empty())
                CheckQt(s.ctx, qt_batch_write_leaves(b, 0, params.data(), int64_t(params.size() / 2)),
                        "qt_batch_write_leaves");
            CheckQt(s.ctx, qt_batch_set_instances(b, int(prog.size()), prog.data(), base.data(), s.n_groups[cls],
                                                  group.data(), 0.5, -0.5),
                    "qt_batch_set_instances");
            CheckQt(s.ctx, qt_batch_launch(b), "qt_batch_launch");
        }
    } catch (const std::exception& e) {
        for (auto& p : take) {
            ServeResult res;
            res.status = QTH_DEVICE;
            res.error = e.what();
            Complete(*p, res);
        }
        std::lock_guard<std::mutex> g(mu_);
        in_flight_bytes_ -= bytes;
        s.running.clear();
        return true;
    }
    std::lock_guard<std::mutex> g(mu_);
    const double now = std::chrono::duration<double, std::milli>(Clock::now() - t0_).count();
    if (busy_slots_ >= 1) stats_.busy_ms += now - last_event_ms_;
    if (busy_slots_ >= 2) stats_.overlap_ms += now - last_event_ms_;
    last_event_ms_ = now;
    ++busy_slots_;
    ++stats_.launches;
    s.busy = true;
    return true;
}

bool Server::FinishIteration(Slot& s, bool wait) {
    for (auto* b : s.batch) {
        int done = 0;
        do {
            if (qt_batch_poll(b, &done) != QT_OK) done = 1;  // an error surfaces in the results below
            if (!done && !wait) return false;
        } while (!done);
    }
    std::vector<ServeResult> res(s.running.size());
    for (size_t r = 0; r < s.running.size(); ++r) res[r].route = s.running[r]->st->route;
    for (int cls = 0; cls < kBatchClasses; ++cls) {
        if (s.inst_of[cls].empty()) continue;
        std::vector<double> vals(2 * s.inst_of[cls].size()), sums(size_t(std::max(s.n_groups[cls], 1)));
        const qt_status st = qt_batch_results(s.batch[cls], vals.data(), sums.data());
        for (size_t i = 0; i < s.inst_of[cls].size(); ++i) {
            const int r = s.inst_of[cls][i].first;
            if (st != QT_OK) {
                res[size_t(r)].status = QTH_DEVICE;
                res[size_t(r)].error = qt_last_error(s.ctx);
                continue;
            }
            if (s.running[size_t(r)]->req.kind == kServeAmplitude) {
                res[size_t(r)].re = vals[2 * i];
                res[size_t(r)].im = vals[2 * i + 1];
            }
        }
        if (st == QT_OK)
            for (size_t r = 0; r < s.running.size(); ++r)
                if (s.group_of_req[cls][r] >= 0) res[r].re = sums[size_t(s.group_of_req[cls][r])];
    }
    for (size_t r = 0; r < s.running.size(); ++r) {
        res[r].planned_bytes = s.running[r]->st->planned_bytes;
        Complete(*s.running[r], res[r]);
    }
    std::lock_guard<std::mutex> g(mu_);
    const double now = std::chrono::duration<double, std::milli>(Clock::now() - t0_).count();
    if (busy_slots_ >= 1) stats_.busy_ms += now - last_event_ms_;
    if (busy_slots_ >= 2) stats_.overlap_ms += now - last_event_ms_;
    last_event_ms_ = now;
    --busy_slots_;
    in_flight_bytes_ -= s.running_bytes;
    s.running.clear();
    s.busy = false;
    return true;
}

void Server::Loop() {
    while (true) {
        bool progressed = false;
        for (auto& s : slots_)
            if (s->busy) progressed |= FinishIteration(*s, false);
        for (auto& s : slots_)
            if (!s->busy) progressed |= StartIteration(*s);
        bool any_busy = false;
        for (auto& s : slots_) any_busy |= s->busy;
        std::unique_lock<std::mutex> lk(mu_);
        if (stop_ && queue_.empty() && !any_busy) break;
        if (progressed) continue;
        if (any_busy || !queue_.empty()) {  // a launch to finish, or requests waiting for memory
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        } else {
            cv_.wait(lk, [&] { return stop_ || !queue_.empty(); });
        }
    }
}

void Server::EagerLoop(int k) {
    // this loop's streams: eager_workers / eager_loops of them
    const size_t per = size_t(std::max(1, opt_.eager_workers / opt_.eager_loops));
    qt_ctx* const* ctxs = eager_ctxs_.data() + size_t(k) * per;
    while (true) {
        std::shared_ptr<Pending> p;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !eager_queue_.empty(); });
            if (eager_queue_.empty()) return;
            // admission: waits until its planned bytes fit beside the work in flight
            if (in_flight_bytes_ + resident_bytes_ + eager_queue_.front()->st->planned_bytes > opt_.hbm_budget &&
                (busy_slots_ > 0 || eager_running_ > 0)) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(50));
                continue;
            }
            p = eager_queue_.front();
            eager_queue_.pop_front();
            ++eager_running_;
            in_flight_bytes_ += p->st->planned_bytes;
            stats_.peak_planned_bytes = std::max(stats_.peak_planned_bytes, in_flight_bytes_);
        }
        ServeResult res;
        res.route = kRouteEager;
        res.planned_bytes = p->st->planned_bytes;
        try {
            // the large-network path (complex128 unless set otherwise): fused small
            // subtrees, MFMA GEMMs for the large steps, FP64 sums of slices
            if (p->req.kind == kServeQaoaEnergy) {
                // edge networks on eager_workers threads, one stream each; the per-edge
                // values are summed in edge order afterwards, so the energy does not
                // depend on which worker ran which edge
                const int n_edges = int(p->st->data.pairs.size());
                std::vector<double> zz(size_t(n_edges), 0.0);
                const auto block = QaoaParamBlock(p->st->p, p->req.angles.data());
                std::atomic<int> next{0};
                std::vector<std::exception_ptr> errs(per);
                auto work = [&](size_t w) {
                    try {
                        for (int e = next++; e < n_edges; e = next++) {
                            // the shared plan, run with this request's leaf values: the
                            // structure's own with the Rz / Rx leaves set from the angles
                            // (the values parsing the request's QASM text would give)
                            const SlicedNetwork& sn = *p->st->eager_nets[size_t(e)];
                            SlicedNetwork::LeafState leaves = p->st->eager_base[size_t(e)];
                            for (const auto& [leaf, off] : p->st->eager_param_leaves[size_t(e)])
                                sn.SetLeaf(leaves, leaf,
                                           std::vector<std::complex<double>>(block.begin() + off, block.begin() + off + 16));
                            zz[size_t(e)] = sn.Run(ctxs[w], opt_.eager_precision, nullptr, leaves).real();
                        }
                    } catch (...) {
                        errs[w] = std::current_exception();
                        next = n_edges;
                    }
                };
                std::vector<std::thread> pool;
                const size_t n_workers =
                    std::min({per, size_t(n_edges), size_t(p->st->eager_workers)});
                try {
                    for (size_t w = 1; w < n_workers; ++w) pool.emplace_back(work, w);
                } catch (const std::system_error&) {
                    // fewer threads: the ones started and this one still take every edge
                }
                work(0);
                for (auto& t : pool) t.join();
                for (auto& ep : errs)
                    if (ep) std::rethrow_exception(ep);
                double f = 0.0;
                for (double v : zz) f += 0.5 * (1.0 - v);  // src/maxcut.cpp:196
                res.re = f;
            } else {
                const SlicedNetwork& sn = *p->st->eager_nets[0];
                const auto v = sn.Run(ctxs[0], opt_.eager_precision, nullptr, sn.MakeLeafState(p->req.text, p->req.bits));
                res.re = v.real();
                res.im = v.imag();
            }
        } catch (const std::exception& e) {
            res.status = QTH_DEVICE;
            res.error = e.what();
        }
        // this loop's streams keep cached blocks for the next request; beyond their share
        // of the budget they are returned, so held memory stays near what is planned
        for (size_t w = 0; w < per; ++w)
            if (qt_mem_reserved(ctxs[w]) > opt_.hbm_budget / int64_t(2 * opt_.eager_workers)) (void)qt_ctx_trim(ctxs[w]);
        {
            std::lock_guard<std::mutex> g(mu_);
            in_flight_bytes_ -= p->st->planned_bytes;
            --eager_running_;
        }
        Complete(*p, res);
    }
}

}  // namespace qtorch
