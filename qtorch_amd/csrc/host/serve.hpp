// Serving: a continuous-batching request loop over the device engine (see serve.cpp).
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "qtorch.hpp"
#include "qtorch_hip.h"

namespace qtorch {

struct ServeOptions {
    int device = 0;
    int slots = 2;                      // streams that run batched iterations side by side
    int precision = QT_C64_4M;          // arithmetic form of the batched routes
    int eager_precision = QT_PLAIN_C128;  // arithmetic form of the large-network path
    int64_t hbm_budget = 16LL << 30;    // planned device bytes of the requests admitted at once
    int max_requests_per_launch = 512;
    int64_t small_scratch = 768;        // elements: programs up to this run in lane-grouped waves (LDS)
    int small_lanes = 8, small_wgs = 4; // executor shape of the small route
    int64_t param_capacity = 1 << 22;   // parameter elements per slot and launch
    int eager_max_bits = 28;            // large-network path: index bits per tensor after slicing
    int64_t max_request_cmacs = int64_t(1) << 44;  // large-network path: planned complex MACs of a request
    int eager_workers = 4;              // large-network path: streams (edge networks run side by side)
    int eager_loops = 2;                // large-network requests in flight at once (sharing those streams)
    int max_structures = 4096;          // planned structures kept on the host (least recently used dropped)
    // device bytes of the programs and constants resident in a slot's batches: beyond
    // this, an idle slot drops every program (structures are added again on use)
    int64_t resident_limit = 2LL << 30;
};

enum ServeKind { kServeQaoaEnergy = 0, kServeAmplitude = 1 };
enum ServeRoute { kRouteSmall = 0, kRouteMedium = 1, kRouteEager = 2 };

struct ServeRequest {
    int kind = kServeQaoaEnergy;
    std::string text;            // .dgf graph (QAOA) or QASM circuit (amplitude)
    int p = 1;                   // QAOA depth
    bool ref_map = false;        // QAOA: keep the reference's light-cone qubit map
    std::vector<double> angles;  // QAOA: [betas..., gammas...] (src/maxcut.cpp:155-157)
    std::string bits;            // amplitude: output bitstring, one 0/1 per qubit
};

struct ServeResult {
    int status = 0;  // QTH_* code of include/qtorch_host.h
    double re = 0, im = 0;  // QAOA: energy F_p (re); amplitude: <bits|C|0...0>
    double latency_ms = 0;  // submission -> result ready
    int route = -1;
    int64_t planned_bytes = 0;
    std::string error;
};

struct ServePlan {
    int route = kRouteSmall;
    int64_t planned_bytes = 0;
    int64_t cmacs = 0;  // complex MACs of one request
};
// Route and planned device bytes of a request (host planning only; throws
// InvalidFunctionInput for a request that cannot run).
ServePlan PlanRequest(const ServeRequest& r, const ServeOptions& opt);

struct ServeStats {
    int64_t submitted = 0, completed = 0, rejected = 0, launches = 0;
    double p50_ms = 0, p99_ms = 0, mean_ms = 0;
    double overlap_ms = 0;       // time with launches of two or more slots in flight
    double busy_ms = 0;          // time with any launch in flight
    int64_t peak_planned_bytes = 0;
    int64_t edge_evals = 0;      // QAOA edge networks evaluated
    int64_t resident_bytes = 0;  // programs and constants resident in the slots' batches
    int64_t slot_resets = 0;     // times a slot dropped its programs (resident limit)
};

class Server {
public:
    explicit Server(const ServeOptions& opt);
    ~Server();
    Server(const Server&) = delete;
    Server& operator=(const Server&) = delete;

    int64_t Submit(ServeRequest r);            // thread safe
    bool TryGet(int64_t id, ServeResult* out);  // removes a finished result
    ServeResult Wait(int64_t id);
    void Drain();                               // until every submitted request is finished
    ServeStats Stats();

    struct Structure;
    struct Slot;
    struct Pending;

private:
    using Clock = std::chrono::steady_clock;
    std::shared_ptr<Structure> GetStructure(const ServeRequest& r);
    void Loop();
    void EagerLoop(int k);
    bool StartIteration(Slot& s);
    void ResetSlot(Slot& s);
    bool FinishIteration(Slot& s, bool wait);
    void Complete(Pending& p, ServeResult res);

    ServeOptions opt_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::shared_ptr<Pending>> queue_, eager_queue_;
    std::unordered_map<int64_t, ServeResult> results_;
    // planned structures by key: built outside struct_mu_ (other requests for the same
    // key wait on the future, the rest go on); least recently used ones beyond
    // opt_.max_structures are dropped (requests and slots that hold one keep it)
    struct CachedStructure {
        std::shared_future<std::shared_ptr<Structure>> fut;
        uint64_t last_use = 0;
    };
    std::map<std::string, CachedStructure> structures_;
    uint64_t struct_tick_ = 0;
    std::mutex struct_mu_;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::vector<qt_ctx*> eager_ctxs_;  // one stream per large-network worker
    std::vector<double> latencies_;
    int64_t next_id_ = 1, in_flight_bytes_ = 0, outstanding_ = 0;
    int64_t resident_bytes_ = 0;  // programs and constants resident in the slots' batches
    ServeStats stats_;
    bool stop_ = false;
    Clock::time_point t0_;
    double last_event_ms_ = 0;
    int busy_slots_ = 0;
    std::thread loop_;
    std::vector<std::thread> eager_threads_;
    int eager_running_ = 0;  // large-network requests being contracted
};

}  // namespace qtorch
