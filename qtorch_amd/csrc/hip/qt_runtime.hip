// Context, caching allocator and tensor entry points of libqtorch_hip.so.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include "qt_internal.h"

namespace qt {

size_t Pool::round(size_t bytes) {
    if (bytes <= 256) return 256;
    if (bytes <= (size_t(1) << 24)) {  // power-of-two classes up to 16 MiB
        size_t r = 256;
        while (r < bytes) r <<= 1;
        return r;
    }
    const size_t g = size_t(1) << 21;  // 2 MiB granules above
    return (bytes + g - 1) / g * g;
}

void* Pool::alloc(size_t bytes, hipError_t* err) {
    const size_t r = round(bytes);
    auto it = free_.find(r);
    if (it != free_.end() && !it->second.empty()) {
        void* p = it->second.back();
        it->second.pop_back();
        *err = hipSuccess;
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, r);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, r);
    }
    *err = e;
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    sizes_[p] = r;
    reserved_ += int64_t(r);
    return p;
}

void Pool::release(void* p) {
    if (!p) return;
    auto it = sizes_.find(p);
    if (it == sizes_.end()) return;
    free_[it->second].push_back(p);
}

void Pool::trim() {
    for (auto& kv : free_) {
        for (void* p : kv.second) {
            (void)hipFree(p);
            reserved_ -= int64_t(kv.first);
            sizes_.erase(p);
        }
        kv.second.clear();
    }
}

qt_status set_error(Ctx* c, qt_status s, const std::string& msg) {
    if (c) c->err = msg;
    return s;
}

qt_status check_deadline(Ctx* c, const char* what) {
    if (!c->has_deadline || std::chrono::steady_clock::now() < c->deadline) return QT_OK;
    return set_error(c, QT_TIMEOUT, std::string(what) + ": the context's time limit has passed");
}

qt_status check_hip(Ctx* c, hipError_t e, const char* what) {
    if (e == hipSuccess) return QT_OK;
    (void)hipGetLastError();
    return set_error(c, e == hipErrorOutOfMemory ? QT_OOM : QT_DEVICE,
                     std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace qt

using qt::check_hip;
using qt::set_error;

namespace {

// Scratch of qt_contract_pair_noalloc: one array per device (each device holds its
// own copy for the life of the process).  In host code `scratch` is only a symbol:
// its device address on each device comes from hipGetSymbolAddress with that device
// current, taken once per device when the first context on it is created.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The calling thread's current device is `device`.
qt_status take_scratch_address(qt::Ctx* c, int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return QT_OK;  // the no-allocation entry refuses such devices
    if (g_scratch_ptr[device]) return QT_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return check_hip(c, e != hipSuccess ? e : hipErrorInvalidSymbol, "scratch address");
    g_scratch_ptr[device] = ptr;
    return QT_OK;
}

}  // namespace

extern "C" {

qt_status qt_ctx_create(int device, qt_ctx** out) {
    if (!out) return QT_INVALID_INPUT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        return QT_DEVICE;
    }
    if (device < 0 || device >= n) return QT_INVALID_INPUT;
    qt_ctx* c = new (std::nothrow) qt_ctx();
    if (!c) return QT_OOM;
    c->impl.device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->impl.stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete c;
        return QT_DEVICE;
    }
    if (take_scratch_address(&c->impl, device) != QT_OK) {
        (void)hipStreamDestroy(c->impl.stream);
        delete c;
        return QT_DEVICE;
    }
    *out = c;
    return QT_OK;
}

qt_status qt_ctx_destroy(qt_ctx* c) {
    if (!c) return QT_INVALID_INPUT;
    {
        std::lock_guard<std::recursive_mutex> g(c->impl.mu);
        if (c->impl.live > 0)
            return set_error(&c->impl, QT_INVALID_INPUT,
                             "qt_ctx_destroy: " + std::to_string(c->impl.live) + " tensors or batches still alive");
        (void)hipSetDevice(c->impl.device);
        (void)hipStreamSynchronize(c->impl.stream);
        c->impl.pool.trim();
        if (c->impl.partials) (void)hipFree(c->impl.partials);
        if (c->impl.d_exp) (void)hipFree(c->impl.d_exp);
        if (c->impl.d_absmax) (void)hipFree(c->impl.d_absmax);
        for (int k = 0; k < 2; ++k) {
            if (c->impl.staging[k]) (void)hipHostFree(c->impl.staging[k]);
            if (c->impl.staging_ev[k]) (void)hipEventDestroy(c->impl.staging_ev[k]);
        }
        (void)hipStreamDestroy(c->impl.stream);
    }
    delete c;
    return QT_OK;
}

qt_status qt_ctx_set_time_limit(qt_ctx* c, double seconds) {
    if (!c || !(seconds == seconds)) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(c->impl.mu);
    c->impl.has_deadline = seconds > 0;
    if (c->impl.has_deadline)
        c->impl.deadline = std::chrono::steady_clock::now() +
                           std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                               std::chrono::duration<double>(std::min(seconds, 1e9)));
    return QT_OK;
}

qt_status qt_sync(qt_ctx* c) {
    if (!c) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(c->impl.mu);
    (void)hipSetDevice(c->impl.device);
    return check_hip(&c->impl, hipStreamSynchronize(c->impl.stream), "hipStreamSynchronize");
}

const char* qt_last_error(const qt_ctx* c) { return c ? c->impl.err.c_str() : "null context"; }

int64_t qt_mem_reserved(const qt_ctx* c) { return c ? c->impl.pool.reserved() : 0; }

qt_status qt_ctx_trim(qt_ctx* c) {
    if (!c) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(c->impl.mu);
    (void)hipSetDevice(c->impl.device);
    qt_status s = check_hip(&c->impl, hipStreamSynchronize(c->impl.stream), "trim sync");
    if (s == QT_OK) c->impl.pool.trim();
    return s;
}

qt_status qt_tensor_alloc(qt_ctx* c, int rank, int dim_log2, int dtype, qt_tensor** out) {
    if (!c || !out) return QT_INVALID_INPUT;
    *out = nullptr;
    auto* ctx = &c->impl;
    if (rank < 0 || (dim_log2 != 1 && dim_log2 != 2) ||
        (dtype != QT_DTYPE_C128 && dtype != QT_DTYPE_C64))
        return set_error(ctx, QT_INVALID_INPUT, "qt_tensor_alloc: bad rank/dim/dtype");
    if (rank * dim_log2 > 40)
        return set_error(ctx, QT_OOM, "qt_tensor_alloc: tensor larger than 2^40 elements");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    qt_tensor* t = new (std::nothrow) qt_tensor();
    if (!t) return QT_OOM;
    t->impl.rank = rank;
    t->impl.dim_log2 = dim_log2;
    t->impl.dtype = dtype;
    t->impl.numel = int64_t(1) << (rank * dim_log2);
    t->impl.bytes = size_t(t->impl.numel) * qt::elem_bytes(dtype);
    hipError_t e;
    t->impl.data = ctx->pool.alloc(t->impl.bytes, &e);
    if (!t->impl.data) {
        delete t;
        return check_hip(ctx, e, "qt_tensor_alloc");
    }
    qt_status s = qt::launch_fill_zero(ctx, t->impl.data, t->impl.bytes);
    if (s != QT_OK) {
        ctx->pool.release(t->impl.data);
        delete t;
        return s;
    }
    ++ctx->live;
    *out = t;
    return QT_OK;
}

qt_status qt_tensor_create_many(qt_ctx* c, int n, const int* ranks, int dim_log2, int dtype,
                                const void* const* host, qt_tensor** out) {
    if (!c || n < 0 || (n > 0 && (!ranks || !host || !out))) return QT_INVALID_INPUT;
    auto* ctx = &c->impl;
    if ((dim_log2 != 1 && dim_log2 != 2) || (dtype != QT_DTYPE_C128 && dtype != QT_DTYPE_C64))
        return set_error(ctx, QT_INVALID_INPUT, "qt_tensor_create_many: bad dim/dtype");
    if (n == 0) return QT_OK;
    // one device block (256-byte aligned slots) and one pinned staging copy for all tensors
    std::vector<size_t> off(static_cast<size_t>(n));
    size_t total = 0;
    for (int i = 0; i < n; ++i) {
        if (ranks[i] < 0 || ranks[i] * dim_log2 > 30 || !host[i])
            return set_error(ctx, QT_INVALID_INPUT, "qt_tensor_create_many: bad rank or host pointer");
        off[size_t(i)] = total;
        const size_t bytes = (size_t(1) << (ranks[i] * dim_log2)) * qt::elem_bytes(dtype);
        total += (bytes + 255) / 256 * 256;
    }
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    const int k = ctx->staging_next;
    ctx->staging_next ^= 1;
    if (!ctx->staging_ev[k]) {
        qt_status s = check_hip(ctx, hipEventCreateWithFlags(&ctx->staging_ev[k], hipEventDisableTiming), "event");
        if (s != QT_OK) return s;
    }
    if (ctx->staging_pending[k]) {  // the copy that last read this buffer
        qt_status s = check_hip(ctx, hipEventSynchronize(ctx->staging_ev[k]), "staging reuse");
        if (s != QT_OK) return s;
        ctx->staging_pending[k] = false;
    }
    if (ctx->staging_bytes[k] < total) {
        if (ctx->staging[k]) (void)hipHostFree(ctx->staging[k]);
        ctx->staging[k] = nullptr;
        ctx->staging_bytes[k] = 0;
        qt_status s = check_hip(ctx, hipHostMalloc(&ctx->staging[k], total), "staging");
        if (s != QT_OK) return s;
        ctx->staging_bytes[k] = total;
    }
    char* const stage = static_cast<char*>(ctx->staging[k]);
    hipError_t e;
    void* block = ctx->pool.alloc(total, &e);
    if (!block) return check_hip(ctx, e, "qt_tensor_create_many");
    int* refs = new (std::nothrow) int(n);
    if (!refs) {
        ctx->pool.release(block);
        return QT_OOM;
    }
    for (int i = 0; i < n; ++i) {
        qt_tensor* t = new (std::nothrow) qt_tensor();
        if (!t) {
            for (int j = 0; j < i; ++j) delete out[j];
            delete refs;
            ctx->pool.release(block);
            return QT_OOM;
        }
        t->impl.rank = ranks[i];
        t->impl.dim_log2 = dim_log2;
        t->impl.dtype = dtype;
        t->impl.numel = int64_t(1) << (ranks[i] * dim_log2);
        t->impl.bytes = size_t(t->impl.numel) * qt::elem_bytes(dtype);
        t->impl.data = static_cast<char*>(block) + off[size_t(i)];
        t->impl.block = block;
        t->impl.block_refs = refs;
        std::memcpy(stage + off[size_t(i)], host[i], t->impl.bytes);
        out[i] = t;
    }
    // asynchronous: the staging buffer is reused once this copy's event has completed
    qt_status s = check_hip(ctx, hipMemcpyAsync(block, stage, total, hipMemcpyHostToDevice, ctx->stream),
                            "qt_tensor_create_many copy");
    if (s == QT_OK) {
        s = check_hip(ctx, hipEventRecord(ctx->staging_ev[k], ctx->stream), "event record");
        if (s == QT_OK) ctx->staging_pending[k] = true;
        else (void)hipStreamSynchronize(ctx->stream);  // no event: the copy must end before reuse
    }
    if (s != QT_OK) {
        for (int i = 0; i < n; ++i) {
            delete out[i];
            out[i] = nullptr;
        }
        delete refs;
        ctx->pool.release(block);
    } else {
        ctx->live += n;
    }
    return s;
}

qt_status qt_tensor_upload(qt_ctx* c, qt_tensor* t, const void* host, int64_t nbytes) {
    if (!c || !t || !host) return QT_INVALID_INPUT;
    auto* ctx = &c->impl;
    if (nbytes != int64_t(t->impl.bytes))
        return set_error(ctx, QT_INVALID_INPUT, "qt_tensor_upload: size mismatch");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    // host memory is pageable: the copy is synchronous with respect to the host buffer
    qt_status s = check_hip(ctx, hipMemcpyAsync(t->impl.data, host, size_t(nbytes),
                                                hipMemcpyHostToDevice, ctx->stream),
                            "qt_tensor_upload");
    if (s != QT_OK) return s;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "qt_tensor_upload sync");
}

qt_status qt_tensor_download(qt_ctx* c, const qt_tensor* t, void* host, int64_t nbytes) {
    if (!c || !t || !host) return QT_INVALID_INPUT;
    auto* ctx = &c->impl;
    if (nbytes != int64_t(t->impl.bytes))
        return set_error(ctx, QT_INVALID_INPUT, "qt_tensor_download: size mismatch");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    qt_status s = check_hip(ctx, hipMemcpyAsync(host, t->impl.data, size_t(nbytes),
                                                hipMemcpyDeviceToHost, ctx->stream),
                            "qt_tensor_download");
    if (s != QT_OK) return s;
    return check_hip(ctx, hipStreamSynchronize(ctx->stream), "qt_tensor_download sync");
}

qt_status qt_tensor_read_scalar(qt_ctx* c, const qt_tensor* t, double* re, double* im) {
    if (!c || !t || !re || !im) return QT_INVALID_INPUT;
    if (t->impl.numel != 1)
        return set_error(&c->impl, QT_INVALID_INPUT, "qt_tensor_read_scalar: tensor is not rank 0");
    if (t->impl.dtype == QT_DTYPE_C128) {
        double v[2];
        qt_status s = qt_tensor_download(c, t, v, sizeof v);
        *re = v[0];
        *im = v[1];
        return s;
    }
    float v[2];
    qt_status s = qt_tensor_download(c, t, v, sizeof v);
    int32_t e = 0;
    if (s == QT_OK) {
        std::lock_guard<std::recursive_mutex> g(c->impl.mu);
        s = qt::read_exponent(&c->impl, &t->impl, &e);
    }
    *re = std::ldexp(double(v[0]), e);
    *im = std::ldexp(double(v[1]), e);
    return s;
}

qt_status qt_tensor_accumulate(qt_ctx* c, const qt_tensor* t, qt_tensor* acc) {
    if (!c || !t || !acc) return QT_INVALID_INPUT;
    if (t->impl.numel != 1 || acc->impl.numel != 1 || acc->impl.dtype != QT_DTYPE_C128 || t == acc)
        return set_error(&c->impl, QT_INVALID_INPUT, "qt_tensor_accumulate: needs two distinct scalars, acc complex128");
    std::lock_guard<std::recursive_mutex> g(c->impl.mu);
    (void)hipSetDevice(c->impl.device);
    return qt::accumulate_scalar(&c->impl, &t->impl, &acc->impl);
}

qt_status qt_tensor_exponent(qt_ctx* c, const qt_tensor* t, int32_t* e) {
    if (!c || !t || !e) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(c->impl.mu);
    (void)hipSetDevice(c->impl.device);
    return qt::read_exponent(&c->impl, &t->impl, e);
}

qt_status qt_tensor_free(qt_ctx* c, qt_tensor* t) {
    if (!c || !t) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(c->impl.mu);
    --c->impl.live;
    qt::release_exponent(&c->impl, &t->impl);
    if (t->impl.block_refs) {
        if (--*t->impl.block_refs == 0) {
            c->impl.pool.release(t->impl.block);
            delete t->impl.block_refs;
        }
    } else {
        c->impl.pool.release(t->impl.data);
    }
    delete t;
    return QT_OK;
}

qt_status qt_tensor_alloc_empty(qt_ctx* c, int rank, int dim_log2, int dtype, qt_tensor** out) {
    if (!c || !out) return QT_INVALID_INPUT;
    *out = nullptr;
    auto* ctx = &c->impl;
    if (rank < 0 || (dim_log2 != 1 && dim_log2 != 2) || (dtype != QT_DTYPE_C128 && dtype != QT_DTYPE_C64))
        return set_error(ctx, QT_INVALID_INPUT, "qt_tensor_alloc_empty: bad rank/dim/dtype");
    if (rank * dim_log2 > 40)
        return set_error(ctx, QT_OOM, "qt_tensor_alloc_empty: tensor larger than 2^40 elements");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    qt_tensor* t = new (std::nothrow) qt_tensor();
    if (!t) return QT_OOM;
    t->impl.rank = rank;
    t->impl.dim_log2 = dim_log2;
    t->impl.dtype = dtype;
    t->impl.numel = int64_t(1) << (rank * dim_log2);
    t->impl.bytes = size_t(t->impl.numel) * qt::elem_bytes(dtype);
    hipError_t e;
    t->impl.data = ctx->pool.alloc(t->impl.bytes, &e);
    if (!t->impl.data) {
        delete t;
        return check_hip(ctx, e, "qt_tensor_alloc_empty");
    }
    ++ctx->live;
    *out = t;
    return QT_OK;
}

int64_t qt_tensor_numel(const qt_tensor* t) { return t ? t->impl.numel : 0; }
int qt_tensor_rank(const qt_tensor* t) { return t ? t->impl.rank : -1; }

}  // extern "C"

namespace {

// qt_contract_pair and qt_contract_pair_noalloc: argument checks, the bit map, the
// launch.  ws != nullptr: the step's scratch comes from the workspace and nothing
// is allocated (QT_RESCALE, whose exponent table is allocated on first use, is
// refused there).
qt_status contract_pair_entry(qt_ctx* c, const qt_tensor* A, int rank_a, const qt_tensor* B, int rank_b, int k,
                              const int32_t* pair_a, const int32_t* pair_b, int rank_c, const int32_t* free_src,
                              qt_tensor* C, int precision, int64_t* cmacs, const qt::Workspace* ws) {
    if (!c) return QT_INVALID_INPUT;
    auto* ctx = &c->impl;
    if (!A || !B || !C || (k > 0 && (!pair_a || !pair_b)) || (rank_c > 0 && !free_src))
        return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: null argument");
    const qt::Tensor& a = A->impl;
    const qt::Tensor& b = B->impl;
    qt::Tensor& cc = C->impl;
    if (a.rank != rank_a || b.rank != rank_b || cc.rank != rank_c)
        return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: rank does not match tensor");
    if (a.dtype != b.dtype || a.dtype != cc.dtype || a.dim_log2 != b.dim_log2 ||
        a.dim_log2 != cc.dim_log2)
        return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: dtype/dim mismatch");
    if (&a == &cc || &b == &cc)
        return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: C aliases an operand");
    if (k < 0 || k > rank_a || k > rank_b || rank_c != rank_a + rank_b - 2 * k)
        return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: inconsistent ranks");
    const bool c128 = a.dtype == QT_DTYPE_C128;
    const bool rescale = (precision & QT_RESCALE) != 0;
    precision &= ~QT_RESCALE;
    if (ws && rescale)
        return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair_noalloc: QT_RESCALE needs qt_contract_pair");
    if ((c128 && (precision != QT_PLAIN_C128 || rescale)) ||
        (!c128 && precision != QT_C64_4M && precision != QT_C64_3M && precision != QT_C64_BF16X3))
        return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: precision does not fit dtype");
    const int dl = a.dim_log2;
    if ((rank_c + k) * dl > 62)
        return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: index space too large");
    // every axis of A and of B is used exactly once (summed or free)
    if (rank_a > 64 || rank_b > 64) return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: rank above 64");
    int used_a[64] = {0}, used_b[64] = {0};
    qt::BitMap m;
    std::memset(&m, 0, sizeof m);
    m.nbits_c = rank_c * dl;
    m.nbits_s = k * dl;
    for (int i = 0; i < k; ++i) {
        const int pa = pair_a[i], pb = pair_b[i];
        if (pa < 0 || pa >= rank_a || pb < 0 || pb >= rank_b || used_a[pa]++ || used_b[pb]++)
            return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: bad pair");
        // digit j of s belongs to pair k-1-j (src/Network.h:912-916)
        const int j = k - 1 - i;
        for (int t = 0; t < dl; ++t) {
            m.s_a[j * dl + t] = uint8_t(pa * dl + t);
            m.s_b[j * dl + t] = uint8_t(pb * dl + t);
        }
    }
    for (int i = 0; i < rank_c; ++i) {
        const bool from_a = (free_src[i] & QT_FROM_A) != 0;
        const int ax = free_src[i] & ~QT_FROM_A;
        if (from_a ? (ax < 0 || ax >= rank_a || used_a[ax]++) : (ax < 0 || ax >= rank_b || used_b[ax]++))
            return set_error(ctx, QT_INVALID_INPUT, "qt_contract_pair: bad free axis");
        for (int t = 0; t < dl; ++t)
            m.c_src[i * dl + t] = uint8_t(from_a ? ax * dl + t : 64 + ax * dl + t);
    }
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    qt_status s = qt::check_deadline(ctx, "qt_contract_pair");
    if (s != QT_OK) return s;
    (void)hipSetDevice(ctx->device);
    s = qt::contract_pair(ctx, &a, &b, &cc, m, precision, ws);
    // without QT_RESCALE, C carries its operands' exponents (0 when they have none)
    if (s == QT_OK) s = rescale ? qt::rescale_c64(ctx, &a, &b, &cc) : qt::carry_exponent(ctx, &a, &b, &cc);
    if (s == QT_OK && cmacs) *cmacs = int64_t(1) << ((rank_c + k) * dl);
    return s;
}

}  // namespace

extern "C" {

qt_status qt_contract_pair(qt_ctx* c, const qt_tensor* A, int rank_a, const qt_tensor* B,
                           int rank_b, int k, const int32_t* pair_a, const int32_t* pair_b,
                           int rank_c, const int32_t* free_src, qt_tensor* C, int precision,
                           int64_t* cmacs) {
    return contract_pair_entry(c, A, rank_a, B, rank_b, k, pair_a, pair_b, rank_c, free_src, C, precision, cmacs,
                               nullptr);
}

qt_status qt_contract_pair_noalloc(qt_ctx* c, const qt_tensor* A, int rank_a, const qt_tensor* B, int rank_b, int k,
                                   const int32_t* pair_a, const int32_t* pair_b, int rank_c, const int32_t* free_src,
                                   qt_tensor* C, int precision, int64_t* cmacs) {
    if (!c) return QT_INVALID_INPUT;
    qt::Workspace ws;
    {
        std::lock_guard<std::mutex> g(g_scratch_mu);
        const int dev = c->impl.device;
        if (dev < 0 || dev >= kMaxDevices || !g_scratch_ptr[dev])
            return set_error(&c->impl, QT_DEVICE, "qt_contract_pair_noalloc: no scratch address for the device");
        ws.base = static_cast<char*>(g_scratch_ptr[dev]);
    }
    ws.bytes = kScratchBytes;
    return contract_pair_entry(c, A, rank_a, B, rank_b, k, pair_a, pair_b, rank_c, free_src, C, precision, cmacs,
                               &ws);
}

}  // extern "C"
