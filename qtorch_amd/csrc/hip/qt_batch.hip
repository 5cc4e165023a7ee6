// Batched execution of many small contraction programs (libqtorch_hip.so).
//
// A program is the step list of one network (e.g. one QAOA edge light cone) in a
// fixed order.  qt_program_create compiles it once on the host: every step's index
// maps become two tables of packed 16-bit offsets (one entry per output element,
// one per summed index), intermediates get slots in a per-executor scratch area by
// interval allocation, and leaves become references into a leaf pool (a reference
// may be relative to the instance's parameter block, so instances of one program
// can carry different angles).
//
// qt_batch_launch runs a persistent kernel: each executor (one wavefront, or one
// 256-thread workgroup for larger programs) takes instances from an atomic work
// counter and contracts a whole network, step after step, with its intermediates in
// LDS (or, when they do not fit, in a global scratch area), then writes the final
// scalar.  A second kernel sums per-group results in FP64 (QAOA energies).  The
// launch sequence can be captured into a HIP graph and replayed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#include "qt_batch_impl.h"

namespace qt {
namespace {

constexpr int kWg = 256;

template <class V>
struct Real;
template <>
struct Real<double2> {
    using T = double;
};
template <>
struct Real<float2> {
    using T = float;
};

__device__ __forceinline__ float bf16r(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

// complex multiply-accumulate in the batch's arithmetic form
template <class V, int MODE>
struct Acc {
    using T = typename Real<V>::T;
    T re = 0, im = 0, p3 = 0;
    __device__ __forceinline__ void mac(V a, V b) {
        if constexpr (MODE == QT_C64_3M) {
            re = fmaf(a.x, b.x, re);
            im = fmaf(a.y, b.y, im);
            p3 = fmaf(a.x + a.y, b.x + b.y, p3);
        } else if constexpr (MODE == QT_C64_BF16X3) {
            const float arh = bf16r(a.x), arl = bf16r(a.x - arh), aih = bf16r(a.y), ail = bf16r(a.y - aih);
            const float brh = bf16r(b.x), brl = bf16r(b.x - brh), bih = bf16r(b.y), bil = bf16r(b.y - bih);
            re = fmaf(arh, brh, re); re = fmaf(arh, brl, re); re = fmaf(arl, brh, re);
            re = fmaf(-aih, bih, re); re = fmaf(-aih, bil, re); re = fmaf(-ail, bih, re);
            im = fmaf(arh, bih, im); im = fmaf(arh, bil, im); im = fmaf(arl, bih, im);
            im = fmaf(aih, brh, im); im = fmaf(aih, brl, im); im = fmaf(ail, brh, im);
        } else {
            re = fma(a.x, b.x, re);
            re = fma(-a.y, b.y, re);
            im = fma(a.x, b.y, im);
            im = fma(a.y, b.x, im);
        }
    }
    __device__ __forceinline__ V get() const {
        V r;
        if constexpr (MODE == QT_C64_3M) {
            r.x = re - im;
            r.y = p3 - re - im;
        } else {
            r.x = re;
            r.y = im;
        }
        return r;
    }
};

template <int LANES, bool GLOBAL_SCRATCH>
__device__ __forceinline__ void exec_sync() {
    if constexpr (LANES == kWg) {
        __syncthreads();
    } else if constexpr (GLOBAL_SCRATCH) {
        // intermediates in global memory: wait for this wave's stores before any
        // lane reads another lane's results
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
    } else {
        // one wavefront: LDS operations of a wave complete in order; keep the
        // compiler from moving memory operations across the step boundary
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

template <class V, int MODE, int LANES, class PA, class PB>
__device__ __forceinline__ void run_step(PA A, PB B, V* C, const uint32_t* __restrict__ fc,
                                         const uint32_t* __restrict__ gs, int nC, int K, int lane,
                                         double2* out_final) {
    auto put = [&](int c, const Acc<V, MODE>& acc) {
        const V v = acc.get();
        if (out_final) *out_final = double2{double(v.x), double(v.y)};
        else C[c] = v;
    };
    if (K == 4) {
        // one summed D = 4 axis (the most common step): the four summed offsets are
        // wave-uniform and stay in scalar registers for every output of the lane
        const uint32_t g0 = gs[0], g1 = gs[1], g2 = gs[2], g3 = gs[3];
        const int a0 = int(g0 & 0xffffu), a1 = int(g1 & 0xffffu), a2 = int(g2 & 0xffffu), a3 = int(g3 & 0xffffu);
        const int b0 = int(g0 >> 16), b1 = int(g1 >> 16), b2 = int(g2 >> 16), b3 = int(g3 >> 16);
        auto one = [&](int c, uint32_t f) {
            const auto pa = A + int(f & 0xffffu);
            const auto pb = B + int(f >> 16);
            const V x0 = pa[a0], y0 = pb[b0], x1 = pa[a1], y1 = pb[b1];
            const V x2 = pa[a2], y2 = pb[b2], x3 = pa[a3], y3 = pb[b3];
            Acc<V, MODE> acc;
            acc.mac(x0, y0);
            acc.mac(x1, y1);
            acc.mac(x2, y2);
            acc.mac(x3, y3);
            put(c, acc);
        };
        // two outputs per lane per iteration (16 operand loads in flight); the next
        // pair's table entries are loaded while this pair is computed
        int c = lane;
        uint32_t f0 = c < nC ? fc[c] : 0u, f1 = c + LANES < nC ? fc[c + LANES] : 0u;
        for (; c + LANES < nC; c += 2 * LANES) {
            const uint32_t g0 = f0, g1 = f1;
            if (c + 2 * LANES < nC) f0 = fc[c + 2 * LANES];
            if (c + 3 * LANES < nC) f1 = fc[c + 3 * LANES];
            const auto pa0 = A + int(g0 & 0xffffu), pa1 = A + int(g1 & 0xffffu);
            const auto pb0 = B + int(g0 >> 16), pb1 = B + int(g1 >> 16);
            const V x0 = pa0[a0], y0 = pb0[b0], x1 = pa0[a1], y1 = pb0[b1];
            const V x2 = pa0[a2], y2 = pb0[b2], x3 = pa0[a3], y3 = pb0[b3];
            const V u0 = pa1[a0], v0 = pb1[b0], u1 = pa1[a1], v1 = pb1[b1];
            const V u2 = pa1[a2], v2 = pb1[b2], u3 = pa1[a3], v3 = pb1[b3];
            Acc<V, MODE> acc, acd;
            acc.mac(x0, y0);
            acd.mac(u0, v0);
            acc.mac(x1, y1);
            acd.mac(u1, v1);
            acc.mac(x2, y2);
            acd.mac(u2, v2);
            acc.mac(x3, y3);
            acd.mac(u3, v3);
            put(c, acc);
            put(c + LANES, acd);
        }
        if (c < nC) one(c, f0);
        return;
    }
    uint32_t f_next = lane < nC ? fc[lane] : 0u;
    for (int c = lane; c < nC; c += LANES) {
        const uint32_t f = f_next;
        if (c + LANES < nC) f_next = fc[c + LANES];
        const auto pa = A + int(f & 0xffffu);
        const auto pb = B + int(f >> 16);
        Acc<V, MODE> acc;
        if ((K & 3) == 0) {
            // summed extents are powers of D: four independent operand loads in flight
            for (int s = 0; s < K; s += 4) {
                const uint32_t g0 = gs[s], g1 = gs[s + 1], g2 = gs[s + 2], g3 = gs[s + 3];
                const V x0 = pa[int(g0 & 0xffffu)], y0 = pb[int(g0 >> 16)];
                const V x1 = pa[int(g1 & 0xffffu)], y1 = pb[int(g1 >> 16)];
                const V x2 = pa[int(g2 & 0xffffu)], y2 = pb[int(g2 >> 16)];
                const V x3 = pa[int(g3 & 0xffffu)], y3 = pb[int(g3 >> 16)];
                acc.mac(x0, y0);
                acc.mac(x1, y1);
                acc.mac(x2, y2);
                acc.mac(x3, y3);
            }
        } else {
            for (int s = 0; s < K; ++s) {
                const uint32_t g = gs[s];
                acc.mac(pa[int(g & 0xffffu)], pb[int(g >> 16)]);
            }
        }
        put(c, acc);
    }
}

// With TOUT, every instance's last step writes its whole result (a tensor, or a
// scalar) to the instance's output tensor tout[i] instead of the FP64 value array.
template <class V, int MODE, int LANES, bool GLOBAL_SCRATCH, bool TOUT>
__global__ __launch_bounds__(kWg) void k_batch(const ProgHdr* __restrict__ progs, const StepDesc* __restrict__ steps,
                                               const uint32_t* __restrict__ tabs,
                                               const int64_t* __restrict__ leaf_refs, const V* __restrict__ pool,
                                               const Instance* __restrict__ inst, int n_inst,
                                               int32_t* __restrict__ counter, double2* __restrict__ out,
                                               V* __restrict__ gscratch, int scratch_elems,
                                               V* const* __restrict__ tout) {
    extern __shared__ unsigned char lds_raw[];
    __shared__ int wg_next;
    constexpr int kExec = kWg / LANES;
    const int ex = threadIdx.x / LANES;
    const int lane = threadIdx.x % LANES;
    V* mem;
    if constexpr (GLOBAL_SCRATCH)
        mem = gscratch + (int64_t(blockIdx.x) * kExec + ex) * scratch_elems;
    else
        mem = reinterpret_cast<V*>(lds_raw) + ex * scratch_elems;
    while (true) {
        int i;
        if constexpr (LANES == kWg) {
            if (threadIdx.x == 0) wg_next = atomicAdd(counter, 1);
            __syncthreads();
            i = wg_next;
            __syncthreads();
        } else {
            int t = 0;
            if (lane == 0) t = atomicAdd(counter, 1);
            i = __shfl(t, 0, 64);
        }
        i = __builtin_amdgcn_readfirstlane(i);
        if (i >= n_inst) break;
        const Instance in = inst[i];
        const ProgHdr P = progs[in.prog];
        const int64_t* refs = leaf_refs + P.leaf_begin;
        for (int si = 0; si < P.n_steps; ++si) {
            const StepDesc st = steps[P.step_begin + si];
            const uint32_t* fc = tabs + st.tab;
            const int nC = 1 << st.nc_bits, K = 1 << st.k_bits;
            const uint32_t* gs = fc + nC;
            double2* fin = (st.flags & 1) ? out + i : nullptr;
            V* C = mem + st.c;
            if constexpr (TOUT) {
                if (st.flags & 1) {
                    fin = nullptr;
                    C = tout[i];
                }
            }
            auto leaf = [&](int code) -> const V* {
                const int64_t r = refs[-1 - code];
                const int64_t off =
                    (r & (int64_t(1) << 62)) ? in.param_base + (r & ~(int64_t(1) << 62)) : P.const_base + r;
                return pool + off;
            };
            if (st.a < 0 && st.b < 0)
                run_step<V, MODE, LANES>(leaf(st.a), leaf(st.b), C, fc, gs, nC, K, lane, fin);
            else if (st.a < 0)
                run_step<V, MODE, LANES>(leaf(st.a), (const V*)(mem + st.b), C, fc, gs, nC, K, lane, fin);
            else if (st.b < 0)
                run_step<V, MODE, LANES>((const V*)(mem + st.a), leaf(st.b), C, fc, gs, nC, K, lane, fin);
            else
                run_step<V, MODE, LANES>((const V*)(mem + st.a), (const V*)(mem + st.b), C, fc, gs, nC, K, lane,
                                         fin);
            exec_sync<LANES, GLOBAL_SCRATCH>();
        }
    }
}

// Several instances of one program per wavefront: lane l serves instance slot l / G
// with G lanes; the wave takes a chunk of up to 64 / G consecutive instances of the
// same program (built on the host) and walks the program once for all of them.
// With STAGE, a wave copies its program's distinct constant leaves to LDS once per
// chunk, and every slot its instance's parameter block, so the steps read leaves
// from LDS; without it, leaves are read from the pool through the caches.
// LDS per wave: [W slot scratch areas (LDS mode)] [W parameter blocks] [constant leaves].
template <class V, int MODE, int G, bool GLOBAL_SCRATCH, bool STAGE>
__global__ __launch_bounds__(64) void k_batch_grouped(const ProgHdr* __restrict__ progs,
                                                      const StepDesc* __restrict__ steps,
                                                      const uint32_t* __restrict__ tabs,
                                                      const int64_t* __restrict__ leaf_refs,
                                                      const int32_t* __restrict__ leaf_lds,
                                                      const CstEntry* __restrict__ csts,
                                                      const V* __restrict__ pool, const Instance* __restrict__ inst,
                                                      const int2* __restrict__ chunks, int n_chunks,
                                                      int32_t* __restrict__ counter, double2* __restrict__ out,
                                                      V* __restrict__ gscratch, int slot_stride, int pstage) {
    extern __shared__ unsigned char lds_raw[];
    constexpr int W = 64 / G;
    constexpr int64_t kParam = int64_t(1) << 62;
    const int lane = threadIdx.x, slot = lane / G, gl = lane % G;
    V* lds = reinterpret_cast<V*>(lds_raw);
    V* mem;
    if constexpr (GLOBAL_SCRATCH) {
        mem = gscratch + (int64_t(blockIdx.x) * W + slot) * slot_stride;
    } else {
        mem = lds + slot * slot_stride;
        lds += W * slot_stride;
    }
    V* const prm = lds + slot * pstage;  // this slot's parameter block
    V* const cst = lds + W * pstage;     // the program's constant leaves
    while (true) {
        int t = 0;
        if (lane == 0) t = atomicAdd(counter, 1);
        const int ci = __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
        if (ci >= n_chunks) break;
        const int2 ch = chunks[ci];
        const bool active = slot < ch.y;
        const int i = ch.x + (active ? slot : 0);
        const int64_t pbase = inst[i].param_base;
        const ProgHdr P = progs[__builtin_amdgcn_readfirstlane(inst[ch.x].prog)];
        const int64_t* refs = leaf_refs + P.leaf_begin;
        const int32_t* lofs = leaf_lds + P.leaf_begin;
        if constexpr (STAGE) {
            exec_sync<64, GLOBAL_SCRATCH>();  // the previous chunk's steps are done with the leaves
            for (int j = 0; j < P.n_cst; ++j) {
                const CstEntry e = csts[P.cst_begin + j];
                const V* src = pool + P.const_base + e.src;
                for (int x = lane; x < e.n; x += 64) cst[e.dst + x] = src[x];
            }
            if (active)
                for (int x = gl; x < P.param_elems; x += G) prm[x] = pool[pbase + x];
            exec_sync<64, GLOBAL_SCRATCH>();
        }
        for (int si = 0; si < P.n_steps; ++si) {
            const StepDesc st = steps[P.step_begin + si];
            const uint32_t* fc = tabs + st.tab;
            const int nC = 1 << st.nc_bits, K = 1 << st.k_bits;
            const uint32_t* gs = fc + nC;
            double2* fin = (st.flags & 1) ? out + i : nullptr;
            V* C = mem + st.c;
            auto leaf = [&](int code) -> const V* {
                if constexpr (STAGE) {
                    const int32_t o = lofs[-1 - code];
                    return o < 0 ? prm + (o & 0x7fffffff) : cst + o;
                } else {
                    const int64_t r = refs[-1 - code];
                    return pool + ((r & kParam) ? pbase + (r & ~kParam) : P.const_base + r);
                }
            };
            if (active) {
                if (st.a < 0 && st.b < 0)
                    run_step<V, MODE, G>(leaf(st.a), leaf(st.b), C, fc, gs, nC, K, gl, fin);
                else if (st.a < 0)
                    run_step<V, MODE, G>(leaf(st.a), (const V*)(mem + st.b), C, fc, gs, nC, K, gl, fin);
                else if (st.b < 0)
                    run_step<V, MODE, G>((const V*)(mem + st.a), leaf(st.b), C, fc, gs, nC, K, gl, fin);
                else
                    run_step<V, MODE, G>((const V*)(mem + st.a), (const V*)(mem + st.b), C, fc, gs, nC, K, gl, fin);
            }
            exec_sync<64, GLOBAL_SCRATCH>();
        }
    }
}

__global__ void k_reset_counter(int32_t* c) { *c = 0; }

// The same sums when every group has m members and member j of group g is instance
// j * n_groups + g (angle set g of edge j in an edge-major batch).  A workgroup takes
// 32 groups; its 16 slices of 32 threads each add the members j = slice, slice + 16,
// ... in order (neighbouring threads read neighbouring values, every load of a
// thread in flight at once), then one thread per group adds the 16 partial sums in
// slice order: a fixed order, whatever the timing.
constexpr int kSumGroups = 32, kSumSlices = 16;
__global__ __launch_bounds__(kSumGroups* kSumSlices) void k_group_sum_strided(const double2* __restrict__ out,
                                                                             double* __restrict__ sums, int n_groups,
                                                                             int m, double c0, double c1) {
    __shared__ double part[kSumSlices][kSumGroups];
    const int gl = int(threadIdx.x) % kSumGroups, slice = int(threadIdx.x) / kSumGroups;
    const int g = int(blockIdx.x) * kSumGroups + gl;
    double acc = 0.0;
    if (g < n_groups) {
#pragma unroll 4
        for (int j = slice; j < m; j += kSumSlices) acc += __builtin_fma(c1, out[int64_t(j) * n_groups + g].x, c0);
    }
    part[slice][gl] = acc;
    __syncthreads();
    if (slice == 0 && g < n_groups) {
        double t = part[0][gl];
        for (int k = 1; k < kSumSlices; ++k) t += part[k][gl];
        sums[g] = t;
    }
}

// FP64 group sums: group g = sum over its members of c0 + c1 * Re(value).
__global__ void k_group_sum(const double2* __restrict__ out, const int32_t* __restrict__ seg,
                            const int32_t* __restrict__ members, double* __restrict__ sums, int n_groups, double c0,
                            double c1) {
    const int g = blockIdx.x;
    if (g >= n_groups) return;
    double acc = 0.0;
    for (int j = seg[g] + int(threadIdx.x); j < seg[g + 1]; j += 64) acc += c0 + c1 * out[members[j]].x;
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_down(acc, d, 64);
    if (threadIdx.x == 0) sums[g] = acc;
}

__global__ void k_to_c64(const double2* __restrict__ in, float2* __restrict__ out, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = float2{float(in[i].x), float(in[i].y)};
}

// Device array of at least n elements, reused while it is large enough (contents
// not kept), then filled from src if src is given.
template <class T>
qt_status dev_put(Ctx* c, T** dst, size_t* cap, const T* src, size_t n) {
    if (n > *cap) {
        batch_free(c, *dst);
        *dst = nullptr;
        *cap = 0;
        qt_status s = check_hip(c, batch_malloc(c, reinterpret_cast<void**>(dst), n * sizeof(T)), "batch alloc");
        if (s != QT_OK) return s;
        *cap = n;
    }
    if (!src || n == 0) return QT_OK;
    return check_hip(c, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice), "batch upload");
}

template <class T>
qt_status dev_upload(Ctx* c, T** dst, const T* src, size_t n) {
    if (*dst) {
        batch_free(c, *dst);
        *dst = nullptr;
    }
    if (n == 0) return QT_OK;
    qt_status s = check_hip(c, batch_malloc(c, reinterpret_cast<void**>(dst), n * sizeof(T)), "batch alloc");
    if (s != QT_OK) return s;
    return check_hip(c, hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice), "batch upload");
}

// executors per workgroup and workgroup size of an executor shape
inline int execs_per_wg(int lanes) { return lanes >= 64 ? kWg / lanes : 64 / lanes; }
// LDS elements of a lane-grouped wave: slot scratch (LDS mode) and leaf stages
inline size_t grouped_lds_elems(const qt_batch* b) {
    const size_t w = size_t(64 / b->lanes);
    return (b->use_global_scratch ? 0 : w * size_t(b->slot_stride)) +
           (b->stage_leaves ? w * size_t(b->pstage) + size_t(b->cstage) : 0);
}
inline int grid_wgs(const qt_batch* b) { return b->n_cus * b->wgs_per_cu * (b->lanes >= 64 ? 1 : 4); }

template <class V, int MODE>
qt_status launch_grouped_t(qt_batch* b) {
    Ctx* c = b->ctx;
    const unsigned grid = unsigned(grid_wgs(b));
    const size_t lds = grouped_lds_elems(b) * sizeof(V);
    V* gs = static_cast<V*>(b->d_global_scratch);
    const V* pool = static_cast<const V*>(b->d_pool);
    auto d_out = reinterpret_cast<double2*>(b->d_out);
#define QT_LAUNCH_G(GL, GS)                                                                                        \
    if (b->stage_leaves) {                                                                                          \
        QT_LAUNCH_GST(GL, GS, true);                                                                               \
    } else {                                                                                                       \
        QT_LAUNCH_GST(GL, GS, false);                                                                              \
    }
#define QT_LAUNCH_GST(GL, GS, ST)                                                                                  \
    hipLaunchKernelGGL((k_batch_grouped<V, MODE, GL, GS, ST>), dim3(grid), dim3(64), lds, c->stream, b->d_progs,   \
                       b->d_steps, b->d_tabs, b->d_leaf_refs, b->d_leaf_lds, b->d_csts, pool, b->d_inst,           \
                       b->d_chunks, b->n_chunks, b->d_counter, d_out, gs, b->slot_stride, b->pstage)
#define QT_LAUNCH_GS(GS)                      \
    switch (b->lanes) {                       \
        case 4: QT_LAUNCH_G(4, GS); break;    \
        case 8: QT_LAUNCH_G(8, GS); break;    \
        case 16: QT_LAUNCH_G(16, GS); break;  \
        default: QT_LAUNCH_G(32, GS); break;  \
    }
    if (b->use_global_scratch) {
        QT_LAUNCH_GS(true);
    } else {
        QT_LAUNCH_GS(false);
    }
#undef QT_LAUNCH_GS
#undef QT_LAUNCH_G
#undef QT_LAUNCH_GST
    return check_hip(c, hipGetLastError(), "k_batch_grouped");
}

template <class V, int MODE>
qt_status launch_batch_t(qt_batch* b) {
    if (b->lanes < 64) {
        if (b->tensor_out)
            return set_error(b->ctx, QT_INVALID_INPUT, "batch: tensor outputs need 64 or 256 lanes per network");
        return launch_grouped_t<V, MODE>(b);
    }
    Ctx* c = b->ctx;
    const int kExec = kWg / b->lanes;
    const unsigned grid = unsigned(b->n_cus * b->wgs_per_cu);
    const size_t lds = b->use_global_scratch ? 0 : size_t(kExec) * size_t(b->scratch_elems) * sizeof(V);
    V* gs = static_cast<V*>(b->d_global_scratch);
    const V* pool = static_cast<const V*>(b->d_pool);
    auto d_out = reinterpret_cast<double2*>(b->d_out);
    V* const* tout = reinterpret_cast<V* const*>(b->d_tout);
#define QT_LAUNCH_T(L, G, T)                                                                                       \
    hipLaunchKernelGGL((k_batch<V, MODE, L, G, T>), dim3(grid), dim3(kWg), lds, c->stream, b->d_progs, b->d_steps, \
                       b->d_tabs, b->d_leaf_refs, pool, b->d_inst, b->n_inst, b->d_counter, d_out, gs,            \
                       b->scratch_elems, tout)
#define QT_LAUNCH(L, G)              \
    if (b->tensor_out) {             \
        QT_LAUNCH_T(L, G, true);     \
    } else {                         \
        QT_LAUNCH_T(L, G, false);    \
    }
    if (b->use_global_scratch) {
        if (b->lanes == 64) {
            QT_LAUNCH(64, true);
        } else {
            QT_LAUNCH(256, true);
        }
    } else {
        if (b->lanes == 64) {
            QT_LAUNCH(64, false);
        } else {
            QT_LAUNCH(256, false);
        }
    }
#undef QT_LAUNCH
#undef QT_LAUNCH_T
    return check_hip(c, hipGetLastError(), "k_batch");
}

qt_status launch_tables(qt_batch* b) {
    switch (b->precision) {
        case QT_PLAIN_C128: return launch_batch_t<double2, QT_PLAIN_C128>(b);
        case QT_C64_3M: return launch_batch_t<float2, QT_C64_3M>(b);
        case QT_C64_BF16X3: return launch_batch_t<float2, QT_C64_BF16X3>(b);
        default: return launch_batch_t<float2, QT_C64_4M>(b);
    }
}

qt_status launch_batch(qt_batch* b, bool capture = false) {
    b->in_use = true;  // queued work reads the batch's arrays
    Ctx* c = b->ctx;
    // a failed reconfiguration leaves scratch or chunk arrays that do not fit the shape
    if (!b->shape_ok) return set_error(c, QT_INVALID_INPUT, "batch: executor shape not configured (last change failed)");
    if (b->needs_tout && !b->tensor_out)
        return set_error(c, QT_INVALID_INPUT, "batch: programs with tensor results need qt_batch_set_outputs");
    if (b->specialised && b->sums_only && b->spec_sums_ok) return launch_specialised_sums(b, capture);
    qt_status s = QT_OK;
    if (b->sums_pending && b->pending_parts) {  // sum the partial sums before d_out is overwritten
        s = flush_sums(b);
        if (s != QT_OK) return s;
    }
    b->values_valid = true;
    if (b->specialised) {
        // direct launches leave the group sums to the next launch (or a flush)
        const bool defer = !capture && b->group_stride_m > 0 && b->d_out_prev;
        s = launch_specialised(b, defer);
        if (s != QT_OK || defer) return s;
    } else {
        hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(1), 0, c->stream, b->d_counter);
        s = check_hip(c, hipGetLastError(), "k_reset_counter");
        if (s != QT_OK) return s;
        s = launch_tables(b);
    }
    if (s != QT_OK || b->n_groups == 0) return s;
    return launch_group_sums(b);
}

size_t elem_size(int precision) { return precision == QT_PLAIN_C128 ? 16 : 8; }

qt_status configure_scratch(qt_batch* b) {
    Ctx* c = b->ctx;
    b->shape_ok = false;  // set once the arrays below match the executor shape
    const size_t es = elem_size(b->precision);
    const int kExec = execs_per_wg(b->lanes);
    // instance slots of a lane-grouped wave sit an odd number of elements apart, so
    // the slots' same-offset accesses fall into different LDS banks
    b->slot_stride = b->lanes < 64 ? b->scratch_elems + 1 : b->scratch_elems;
    // leaf stages of lane-grouped waves (LDS in either placement)
    b->stage_leaves = b->lanes < 64 && (b->scratch_place & QT_SCRATCH_STAGE_LEAVES);
    const size_t stage = b->stage_leaves ? (size_t(kExec) * size_t(b->pstage) + size_t(b->cstage)) * es : 0;
    const size_t lds_need = size_t(kExec) * size_t(b->slot_stride) * es + stage;
    const size_t lds_cap = 160 * 1024;
    const size_t wgs_on_cu = size_t(b->wgs_per_cu) * (b->lanes >= 64 ? 1 : 4);
    if (stage > 64 * 1024)
        return set_error(c, QT_INVALID_INPUT, "batch: the leaf stages do not fit a workgroup's LDS");
    const int place = b->scratch_place & 3;
    if (place == QT_SCRATCH_LDS && lds_need > 64 * 1024)
        return set_error(c, QT_INVALID_INPUT, "batch: the executors' scratch does not fit a workgroup's LDS");
    // LDS unless it does not fit or would hold fewer workgroups per CU than asked for
    // (then the intermediates go through L1/L2); an explicit placement wins
    b->use_global_scratch = place == QT_SCRATCH_GLOBAL ||
                            (place == QT_SCRATCH_AUTO && (lds_need * wgs_on_cu > lds_cap || lds_need > 64 * 1024));
    // chunks of consecutive instances of one program, at most one per instance slot
    if (b->lanes < 64) {
        const int W = 64 / b->lanes;
        std::vector<int2> ch;
        for (int i = 0; i < b->n_inst;) {
            int j = i + 1;
            while (j < b->n_inst && j - i < W && b->h_prog[size_t(j)] == b->h_prog[size_t(i)]) ++j;
            ch.push_back(int2{i, j - i});
            i = j;
        }
        qt_status s = dev_put(c, &b->d_chunks, &b->cap_chunks, ch.data(), ch.size());
        if (s != QT_OK) return s;
        b->n_chunks = int(ch.size());
    }
    if (b->use_global_scratch) {
        const size_t need = size_t(grid_wgs(b)) * size_t(kExec) * size_t(b->slot_stride) * es;
        if (need > (size_t(4) << 30))
            return set_error(c, QT_OOM, "batch: scratch of this executor shape exceeds 4 GiB");
        if (need > b->global_scratch_bytes) {
            batch_free(c, b->d_global_scratch);
            b->d_global_scratch = nullptr;
            b->global_scratch_bytes = 0;
            qt_status s = check_hip(c, batch_malloc(c, &b->d_global_scratch, need), "batch global scratch");
            if (s != QT_OK) return s;
            b->global_scratch_bytes = need;
        }
    }
    if (b->graph_exec) {
        (void)hipGraphExecDestroy(b->graph_exec);
        b->graph_exec = nullptr;
    }
    if (b->graph) {
        (void)hipGraphDestroy(b->graph);
        b->graph = nullptr;
    }
    b->shape_ok = true;
    return QT_OK;
}

// Appends programs (host copies, then the device arrays again); the constant leaf
// references of program i are offsets from const_base[i] in the pool.
qt_status append_programs(qt_batch* b, int n, qt_program* const* progs, const int64_t* const_base,
                          bool configure = true) {
    Ctx* c = b->ctx;
    for (int i = 0; i < n; ++i)
        if (!progs[i] || const_base[i] < 0) return set_error(c, QT_INVALID_INPUT, "batch: null program or base");
    if (int64_t(b->h_tabs.size()) + [&] {
            int64_t t = 0;
            for (int i = 0; i < n; ++i) t += int64_t(progs[i]->tables.size());
            return t;
        }() > int64_t(INT32_MAX))
        return set_error(c, QT_INVALID_INPUT, "batch: program tables exceed 2^31 entries");
    for (int i = 0; i < n; ++i) {
        const qt_program* p = progs[i];
        ProgHdr h{};
        h.step_begin = int32_t(b->h_steps.size());
        h.n_steps = int32_t(p->steps.size());
        h.leaf_begin = int32_t(b->h_refs.size());
        h.n_leaves = int32_t(p->leaf_refs.size());
        h.scratch_elems = p->scratch_elems;
        h.const_base = const_base[i];
        const int32_t tab0 = int32_t(b->h_tabs.size());
        for (auto st : p->steps) {
            st.tab += tab0;
            b->h_steps.push_back(st);
        }
        b->h_tabs.insert(b->h_tabs.end(), p->tables.begin(), p->tables.end());
        int64_t cend = 0, pend = 0;
        std::map<int64_t, int32_t> cst_at;  // distinct constant leaves -> offset in the LDS area
        int32_t cst_elems = 0;
        h.cst_begin = int32_t(b->h_csts.size());
        for (size_t l = 0; l < p->leaf_refs.size(); ++l) {
            const int64_t r = p->leaf_refs[l];
            const int32_t elems = int32_t(1) << (p->leaf_ranks[l] * p->dim_log2);
            const int64_t end = (r & ~QT_LEAF_PARAM) + elems;
            if (r & QT_LEAF_PARAM) {
                pend = std::max(pend, end);
                b->h_leaf_lds.push_back(int32_t(uint32_t(1) << 31 | uint32_t(r & ~QT_LEAF_PARAM)));
            } else {
                cend = std::max(cend, const_base[i] + end);
                auto it = cst_at.find(r);
                if (it == cst_at.end()) {
                    it = cst_at.emplace(r, cst_elems).first;
                    b->h_csts.push_back(CstEntry{r, elems, cst_elems});
                    cst_elems += elems;
                }
                b->h_leaf_lds.push_back(it->second);
            }
            b->h_refs.push_back(r);
        }
        h.n_cst = int32_t(b->h_csts.size()) - h.cst_begin;
        h.param_elems = int32_t(pend);
        h.pad = p->dim_log2;
        b->pstage = std::max(b->pstage, int32_t(pend));
        b->cstage = std::max(b->cstage, cst_elems);
        b->prog_const_end.push_back(cend);
        b->prog_param_end.push_back(pend);
        b->max_leaf_ref_end = std::max(b->max_leaf_ref_end, cend);
        b->max_param_ref_end = std::max(b->max_param_ref_end, pend);
        b->scratch_elems = std::max(b->scratch_elems, (p->scratch_elems + 15) / 16 * 16);
        b->prog_cmacs.push_back(p->cmacs);
        b->prog_out_elems.push_back(p->out_elems);
        b->h_hdr.push_back(h);
    }
    b->n_programs += n;
    drop_specialised(b);  // generated for the previous program set
    quiesce(b);  // a launch in flight reads the arrays
    qt_status s = dev_put(c, &b->d_progs, &b->cap_progs, b->h_hdr.data(), b->h_hdr.size());
    if (s == QT_OK) s = dev_put(c, &b->d_steps, &b->cap_steps, b->h_steps.data(), b->h_steps.size());
    if (s == QT_OK) s = dev_put(c, &b->d_tabs, &b->cap_tabs, b->h_tabs.data(), b->h_tabs.size());
    if (s == QT_OK) s = dev_put(c, &b->d_leaf_refs, &b->cap_refs, b->h_refs.data(), b->h_refs.size());
    if (s == QT_OK) s = dev_put(c, &b->d_leaf_lds, &b->cap_leaf_lds, b->h_leaf_lds.data(), b->h_leaf_lds.size());
    if (s == QT_OK) s = dev_put(c, &b->d_csts, &b->cap_csts, b->h_csts.data(), b->h_csts.size());
    if (s == QT_OK && configure) s = configure_scratch(b);
    return s;
}

// Writes n FP64 complex values at pool offset `offset` (converted for complex64
// batches), growing the pool (contents kept) when needed.  The stream is idle.
qt_status write_pool(qt_batch* b, int64_t offset, const double* values, int64_t n) {
    Ctx* c = b->ctx;
    b->rot_full_valid = false;  // rotation blocks are written in full again next time
    if (b->specialised) {
        bool keep = !specialised_consts_change(b, offset, values, n);  // its literals are the constants
        if (keep && b->mirror_stale) keep = refresh_pool_mirror(b) == QT_OK;
        if (keep) {
            // the host mirror follows the pool; the parameter blocks must keep what the
            // generated code assumes of them
            auto& m = b->h_pool_mirror;
            if (int64_t(m.size()) < 2 * (offset + n)) m.resize(size_t(2 * (offset + n)), 0.0);
            const bool c128 = b->precision == QT_PLAIN_C128;
            for (int64_t i = 0; i < 2 * n; ++i)
                m[size_t(2 * offset + i)] = c128 ? values[i] : double(float(values[i]));
            keep = specialised_params_hold(b);
        }
        if (!keep) {
            drop_specialised(b);
            if (b->graph_exec) (void)configure_scratch(b);  // the captured graph runs the dropped kernel
        }
    }
    const size_t es = elem_size(b->precision);
    const int64_t need = offset + n;
    if (need > b->pool_cap) {
        const int64_t cap = std::max(need, b->pool_cap + b->pool_cap / 2);
        void* fresh = nullptr;
        qt_status s = check_hip(c, batch_malloc(c, &fresh, size_t(cap) * es), "leaf pool");
        if (s != QT_OK) return s;
        if (b->d_pool && b->pool_elems > 0)
            s = check_hip(c, hipMemcpy(fresh, b->d_pool, size_t(b->pool_elems) * es, hipMemcpyDeviceToDevice),
                          "leaf pool grow");
        if (s != QT_OK) {
            batch_free(c, fresh);
            return s;
        }
        batch_free(c, b->d_pool);
        b->d_pool = fresh;
        b->pool_cap = cap;
        if (b->graph_exec) (void)configure_scratch(b);  // the captured graph holds the old pool
    }
    char* dst = static_cast<char*>(b->d_pool) + size_t(offset) * es;
    qt_status s;
    if (b->precision == QT_PLAIN_C128) {
        s = check_hip(c, hipMemcpy(dst, values, size_t(n) * 16, hipMemcpyHostToDevice), "leaf pool copy");
    } else {
        // rounded to complex64 on the host (the same round-to-nearest as a device
        // conversion) and copied: no kernel, so nothing to wait for afterwards
        std::vector<float> f(size_t(2 * n));
        for (int64_t i = 0; i < 2 * n; ++i) f[size_t(i)] = float(values[i]);
        s = check_hip(c, hipMemcpy(dst, f.data(), size_t(n) * 8, hipMemcpyHostToDevice), "leaf pool copy");
    }
    if (s == QT_OK) b->pool_elems = std::max(b->pool_elems, need);
    return s;
}

// A program's steps form a tree: every intermediate is consumed exactly once and
// the last step is the root.  Any post-order of that tree computes the same values
// (every step's own arithmetic is unchanged), but the scratch an executor needs
// depends on the order.  Sethi-Ullman: for each step, evaluating operand X first
// costs max(need(X), size(X) + need(Y), size(X) + size(Y) + size(result)), where
// need() is a subtree's peak and size() its result (leaves live in the pool and
// cost nothing); take the cheaper of the two operand orders.  Returns false (and
// leaves the order as given) for input that is not such a tree: the validation
// that follows reports it.
bool reorder_steps(int dl, int n_leaves, int n_steps, const int32_t* desc, const int32_t* axes,
                   std::vector<int32_t>& desc_out, std::vector<int32_t>& axes_out) {
    if (dl != 1 && dl != 2) return false;
    std::vector<int64_t> need(size_t(n_steps), 0), size(size_t(n_steps), 0);
    std::vector<int> first(size_t(n_steps), 0), used(size_t(n_steps), 0);
    std::vector<size_t> axis_begin(size_t(n_steps) + 1, 0);
    for (int s = 0; s < n_steps; ++s) {
        const int32_t* d = desc + 6 * s;
        const int a = d[0], b = d[1], k = d[4], rc = d[5];
        if (a < 0 || b < 0 || a >= n_leaves + s || b >= n_leaves + s || a == b || k < 0 || rc < 0 || rc * dl > 30)
            return false;
        axis_begin[size_t(s) + 1] = axis_begin[size_t(s)] + size_t(2 * k + rc);
        int64_t n_op[2] = {0, 0}, s_op[2] = {0, 0};
        const int ops[2] = {a, b};
        for (int t = 0; t < 2; ++t)
            if (ops[t] >= n_leaves) {
                const int j = ops[t] - n_leaves;
                if (used[size_t(j)]++) return false;
                n_op[t] = need[size_t(j)];
                s_op[t] = size[size_t(j)];
            }
        size[size_t(s)] = int64_t(1) << (rc * dl);
        const int64_t both = s_op[0] + s_op[1] + size[size_t(s)];
        const int64_t a_first = std::max({n_op[0], s_op[0] + n_op[1], both});
        const int64_t b_first = std::max({n_op[1], s_op[1] + n_op[0], both});
        first[size_t(s)] = b_first < a_first ? 1 : 0;
        need[size_t(s)] = std::min(a_first, b_first);
    }
    for (int s = 0; s + 1 < n_steps; ++s)
        if (!used[size_t(s)]) return false;
    // iterative post-order from the root
    std::vector<int> order;
    order.reserve(size_t(n_steps));
    std::vector<std::pair<int, int>> stack{{n_steps - 1, 0}};
    while (!stack.empty()) {
        auto& top = stack.back();
        const int s = top.first;
        if (top.second == 2) {
            order.push_back(s);
            stack.pop_back();
            continue;
        }
        const int which = top.second == 0 ? first[size_t(s)] : 1 - first[size_t(s)];
        ++top.second;
        const int op = desc[6 * s + which];
        if (op >= n_leaves) stack.push_back({op - n_leaves, 0});
    }
    if (int(order.size()) != n_steps) return false;
    std::vector<int> new_pos(static_cast<size_t>(n_steps));
    for (int i = 0; i < n_steps; ++i) new_pos[size_t(order[size_t(i)])] = i;
    desc_out.resize(size_t(6 * n_steps));
    axes_out.clear();
    axes_out.reserve(axis_begin[size_t(n_steps)]);
    for (int i = 0; i < n_steps; ++i) {
        const int s = order[size_t(i)];
        for (int t = 0; t < 6; ++t) desc_out[size_t(6 * i + t)] = desc[6 * s + t];
        for (int t = 0; t < 2; ++t) {
            const int op = desc[6 * s + t];
            if (op >= n_leaves) desc_out[size_t(6 * i + t)] = n_leaves + new_pos[size_t(op - n_leaves)];
        }
        axes_out.insert(axes_out.end(), axes + axis_begin[size_t(s)], axes + axis_begin[size_t(s) + 1]);
    }
    return true;
}

}  // namespace

qt_status launch_group_sums(qt_batch* b) {
    b->in_use = true;  // queued work reads the batch's arrays
    Ctx* c = b->ctx;
    qt_status s;
    if (b->group_stride_m > 0) {
        hipLaunchKernelGGL(k_group_sum_strided, dim3(unsigned((b->n_groups + kSumGroups - 1) / kSumGroups)),
                           dim3(kSumGroups * kSumSlices), 0, c->stream,
                           reinterpret_cast<const double2*>(b->d_out), b->d_group, b->n_groups, b->group_stride_m,
                           b->c0, b->c1);
        s = check_hip(c, hipGetLastError(), "k_group_sum_strided");
    } else {
        hipLaunchKernelGGL(k_group_sum, dim3(unsigned(b->n_groups)), dim3(64), 0, c->stream,
                           reinterpret_cast<const double2*>(b->d_out), b->d_seg, b->d_members, b->d_group, b->n_groups,
                           b->c0, b->c1);
        s = check_hip(c, hipGetLastError(), "k_group_sum");
    }
    return s;
}

void quiesce(qt_batch* b) {
    if (!b->in_use) return;
    (void)hipStreamSynchronize(b->ctx->stream);
    b->in_use = false;
}

qt_status after_specialised_drop(qt_batch* b) {
    return b->graph_exec ? configure_scratch(b) : QT_OK;  // the captured graph runs the dropped kernel
}

qt_status flush_sums(qt_batch* b) {
    if (!b->sums_pending) return QT_OK;
    b->sums_pending = false;
    if (b->pending_parts) {  // a group-sums-only launch's partial sums
        b->pending_parts = false;
        return launch_part_sums(b);
    }
    return launch_group_sums(b);
}

}  // namespace qt

using qt::check_hip;
using qt::set_error;

extern "C" {

qt_status qt_program_create(qt_ctx* cx, int dim_log2, int n_leaves, const int32_t* leaf_ranks,
                            const int64_t* leaf_refs, int n_steps, const int32_t* step_desc,
                            const int32_t* step_axes, qt_program** out) {
    if (!out || n_leaves < 1 || n_steps < 1 || !leaf_ranks || !leaf_refs || !step_desc || !step_axes)
        return QT_INVALID_INPUT;
    *out = nullptr;
    // run the same contraction tree in the order that keeps the least scratch live
    std::vector<int32_t> desc2, axes2;
    if (qt::reorder_steps(dim_log2, n_leaves, n_steps, step_desc, step_axes, desc2, axes2)) {
        step_desc = desc2.data();
        step_axes = axes2.data();
    }
    qt::Ctx* ctx = cx ? &cx->impl : nullptr;  // compilation touches no device memory
    if (dim_log2 != 1 && dim_log2 != 2) return set_error(ctx, QT_INVALID_INPUT, "qt_program_create: bad dim");
    const int dl = dim_log2;
    auto p = std::make_unique<qt_program>();
    p->dim_log2 = dl;
    p->leaf_ranks.assign(leaf_ranks, leaf_ranks + n_leaves);
    p->leaf_refs.assign(leaf_refs, leaf_refs + n_leaves);
    for (int i = 0; i < n_leaves; ++i)
        if (leaf_ranks[i] < 0 || leaf_ranks[i] * dl > 16 || leaf_refs[i] < 0)
            return set_error(ctx, QT_INVALID_INPUT, "qt_program_create: leaf too large or bad reference");
    // ranks of every operand id; uses of every intermediate (an id is consumed once)
    std::vector<int> rank_of(size_t(n_leaves + n_steps), -1), last_use(size_t(n_steps), -1);
    for (int i = 0; i < n_leaves; ++i) rank_of[size_t(i)] = leaf_ranks[i];
    size_t axis_pos = 0;
    std::vector<size_t> axis_begin(static_cast<size_t>(n_steps));
    for (int s = 0; s < n_steps; ++s) {
        const int32_t* d = step_desc + 6 * s;
        const int a = d[0], b = d[1], ra = d[2], rb = d[3], k = d[4], rc = d[5];
        const int lim = n_leaves + s;
        if (a < 0 || b < 0 || a >= lim || b >= lim || a == b || rank_of[size_t(a)] != ra || rank_of[size_t(b)] != rb ||
            k < 0 || k > ra || k > rb || rc != ra + rb - 2 * k || rc * dl > 16 || (rc + k) * dl > 30)
            return set_error(ctx, QT_INVALID_INPUT, "qt_program_create: inconsistent step " + std::to_string(s));
        for (int op : {a, b})
            if (op >= n_leaves) {
                const int j = op - n_leaves;
                if (last_use[size_t(j)] >= 0)
                    return set_error(ctx, QT_INVALID_INPUT, "qt_program_create: intermediate used twice");
                last_use[size_t(j)] = s;
            }
        rank_of[size_t(n_leaves + s)] = rc;
        axis_begin[size_t(s)] = axis_pos;
        axis_pos += size_t(2 * k + rc);
    }
    for (int s = 0; s + 1 < n_steps; ++s)
        if (last_use[size_t(s)] < 0)
            return set_error(ctx, QT_INVALID_INPUT, "qt_program_create: an intermediate is never consumed");

    // scratch slots: first fit over live intervals (an intermediate lives from its
    // creating step until the step that consumes it)
    struct Block {
        int32_t off, len;
        int free_after;  // step index after which it can be reused
    };
    std::vector<Block> live;
    std::vector<int32_t> slot(size_t(n_steps), -1);
    int32_t peak = 0;
    for (int s = 0; s + 1 < n_steps; ++s) {
        const int32_t need = int32_t(1) << (rank_of[size_t(n_leaves + s)] * dl);
        // blocks whose consumer ran before step s are free (the result of step s is
        // written while its own operands are still read, so they stay reserved)
        std::vector<std::pair<int32_t, int32_t>> busy;
        for (const auto& bl : live)
            if (bl.free_after >= s) busy.push_back({bl.off, bl.off + bl.len});
        std::sort(busy.begin(), busy.end());
        int32_t off = 0;
        for (const auto& iv : busy) {
            if (off + need <= iv.first) break;
            off = std::max(off, iv.second);
        }
        slot[size_t(s)] = off;
        live.push_back({off, need, last_use[size_t(s)]});
        peak = std::max(peak, off + need);
    }
    p->scratch_elems = std::max(peak, 1);

    // tables and step descriptors
    for (int s = 0; s < n_steps; ++s) {
        const int32_t* d = step_desc + 6 * s;
        const int a = d[0], b = d[1], ra = d[2], rb = d[3], k = d[4], rc = d[5];
        const int32_t* ax = step_axes + axis_begin[size_t(s)];
        std::vector<int> used_a(size_t(ra), 0), used_b(size_t(rb), 0);
        std::vector<int> c_src(size_t(rc * dl)), sa(size_t(k * dl)), sb(size_t(k * dl));
        for (int i = 0; i < k; ++i) {
            const int pa = ax[i], pb = ax[k + i];
            if (pa < 0 || pa >= ra || pb < 0 || pb >= rb || used_a[size_t(pa)]++ || used_b[size_t(pb)]++)
                return set_error(ctx, QT_INVALID_INPUT, "qt_program_create: bad pair");
            const int j = k - 1 - i;  // the last pair varies fastest (src/Network.h:912-916)
            for (int t = 0; t < dl; ++t) {
                sa[size_t(j * dl + t)] = pa * dl + t;
                sb[size_t(j * dl + t)] = pb * dl + t;
            }
        }
        for (int i = 0; i < rc; ++i) {
            const int32_t f = ax[2 * k + i];
            const bool from_a = (f & QT_FROM_A) != 0;
            const int axis = f & ~QT_FROM_A;
            if (from_a ? (axis < 0 || axis >= ra || used_a[size_t(axis)]++)
                       : (axis < 0 || axis >= rb || used_b[size_t(axis)]++))
                return set_error(ctx, QT_INVALID_INPUT, "qt_program_create: bad free axis");
            for (int t = 0; t < dl; ++t) c_src[size_t(i * dl + t)] = from_a ? axis * dl + t : 64 + axis * dl + t;
        }
        qt::StepDesc st{};
        st.nc_bits = rc * dl;
        st.k_bits = k * dl;
        st.tab = int32_t(p->tables.size());
        const int64_t nC = int64_t(1) << st.nc_bits, K = int64_t(1) << st.k_bits;
        for (int64_t cc = 0; cc < nC; ++cc) {
            uint32_t oa = 0, ob = 0;
            for (int i = 0; i < st.nc_bits; ++i)
                if ((cc >> i) & 1) {
                    if (c_src[size_t(i)] < 64) oa += 1u << c_src[size_t(i)];
                    else ob += 1u << (c_src[size_t(i)] - 64);
                }
            p->tables.push_back(oa | (ob << 16));
        }
        for (int64_t ss = 0; ss < K; ++ss) {
            uint32_t ga = 0, gb = 0;
            for (int j = 0; j < st.k_bits; ++j)
                if ((ss >> j) & 1) {
                    ga += 1u << sa[size_t(j)];
                    gb += 1u << sb[size_t(j)];
                }
            p->tables.push_back(ga | (gb << 16));
        }
        st.a = a < n_leaves ? -(1 + a) : slot[size_t(a - n_leaves)];
        st.b = b < n_leaves ? -(1 + b) : slot[size_t(b - n_leaves)];
        st.c = s + 1 < n_steps ? slot[size_t(s)] : 0;
        st.flags = s + 1 == n_steps ? 1 : 0;
        st.pad = (ra * dl) | ((rb * dl) << 8);  // element bits of the operands (leaf staging)
        p->steps.push_back(st);
        p->cmacs += int64_t(1) << ((rc + k) * dl);
        p->max_tensor_elems = std::max(p->max_tensor_elems, int32_t(1) << (rc * dl));
        if (s + 1 == n_steps) p->out_elems = int32_t(1) << (rc * dl);
    }
    *out = p.release();
    return QT_OK;
}

qt_status qt_program_stats(const qt_program* p, int64_t* cmacs, int64_t* scratch_elems, int* n_steps) {
    if (!p) return QT_INVALID_INPUT;
    if (cmacs) *cmacs = p->cmacs;
    if (scratch_elems) *scratch_elems = p->scratch_elems;
    if (n_steps) *n_steps = int(p->steps.size());
    return QT_OK;
}

qt_status qt_program_destroy(qt_program* p) {
    delete p;
    return QT_OK;
}

qt_status qt_batch_create(qt_ctx* cx, int n_programs, qt_program* const* progs, int precision, qt_batch** out) {
    if (!cx || !out || n_programs < 0 || (n_programs > 0 && !progs)) return QT_INVALID_INPUT;
    *out = nullptr;
    auto* ctx = &cx->impl;
    if (precision != QT_PLAIN_C128 && precision != QT_C64_4M && precision != QT_C64_3M &&
        precision != QT_C64_BF16X3)
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_create: bad precision");
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    auto b = std::make_unique<qt_batch>();
    b->ctx = ctx;
    b->precision = precision;
    qt_status s = check_hip(ctx, qt::batch_malloc(ctx, reinterpret_cast<void**>(&b->d_counter), sizeof(int32_t)), "counter");
    if (s == QT_OK) s = check_hip(ctx, hipDeviceGetAttribute(&b->n_cus, hipDeviceAttributeMultiprocessorCount,
                                                              ctx->device), "CU count");
    if (s == QT_OK) s = check_hip(ctx, hipEventCreate(&b->ev0), "event");
    if (s == QT_OK) s = check_hip(ctx, hipEventCreate(&b->ev1), "event");
    if (s == QT_OK) s = check_hip(ctx, hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming), "event");
    ++ctx->live;
    if (s == QT_OK && n_programs > 0) {
        std::vector<int64_t> zero(static_cast<size_t>(n_programs), 0);
        s = qt::append_programs(b.get(), n_programs, progs, zero.data(), false);  // shape chosen below
    }
    if (s == QT_OK) {
        // default executor shape from the scratch footprint (the autotuner may refine it)
        const size_t per_exec = size_t(b->scratch_elems) * qt::elem_size(precision);
        b->lanes = per_exec * 4 <= 16 * 1024 ? 64 : 256;
        b->wgs_per_cu = per_exec * 4 <= 16 * 1024 ? 4 : per_exec <= 32 * 1024 ? 2 : 1;
        s = qt::configure_scratch(b.get());
    }
    if (s != QT_OK) {
        qt_batch_destroy(b.release());
        return s;
    }
    *out = b.release();
    return QT_OK;
}

qt_status qt_batch_add_programs(qt_batch* b, int n, qt_program* const* progs, const int64_t* const_base,
                                int32_t* first_id) {
    if (!b || n < 1 || !progs || !const_base) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    if (first_id) *first_id = b->n_programs;
    return qt::append_programs(b, n, progs, const_base);
}

qt_status qt_batch_set_leaves(qt_batch* b, const double* pool, int64_t n_elems) {
    if (!b || !pool || n_elems < 1) return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    qt::quiesce(b);
    if (b->max_leaf_ref_end > n_elems)
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_set_leaves: pool smaller than the leaf references");
    b->pool_elems = 0;
    b->n_inst = 0;  // parameter bases were checked against the old pool: set instances again
    b->sums_pending = false;
    b->pending_parts = false;
    (void)qt::configure_scratch(b);
    return qt::write_pool(b, 0, pool, n_elems);
}

qt_status qt_batch_write_leaves(qt_batch* b, int64_t offset, const double* values, int64_t n_elems) {
    if (!b || offset < 0 || n_elems < 1 || !values) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    qt::quiesce(b);  // a launch in flight may read the pool
    return qt::write_pool(b, offset, values, n_elems);
}

qt_status qt_batch_set_instances(qt_batch* b, int n, const int32_t* prog, const int64_t* param_base, int n_groups,
                                 const int32_t* group, double c0, double c1) {
    if (!b || n < 0 || (n > 0 && (!prog || !param_base)) || n_groups < 0 || (n_groups > 0 && !group))
        return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    // a launch still in flight reads the instance arrays that are replaced below
    qt::quiesce(b);
    if (b->pool_elems == 0) return set_error(ctx, QT_INVALID_INPUT, "qt_batch_set_instances: set the leaves first");
    // validate everything before any state changes
    std::vector<qt::Instance> inst(static_cast<size_t>(n));
    int64_t cmacs = 0;
    bool needs_tout = false;
    for (int i = 0; i < n; ++i) {
        if (prog[i] < 0 || prog[i] >= b->n_programs || param_base[i] < 0 ||
            param_base[i] + b->prog_param_end[size_t(prog[i])] > b->pool_elems ||
            b->prog_const_end[size_t(prog[i])] > b->pool_elems)
            return set_error(ctx, QT_INVALID_INPUT, "qt_batch_set_instances: bad program or parameter block");
        if (n_groups > 0 && (group[i] < -1 || group[i] >= n_groups))
            return set_error(ctx, QT_INVALID_INPUT, "qt_batch_set_instances: bad group");
        inst[size_t(i)] = qt::Instance{prog[i], 0, param_base[i]};
        cmacs += b->prog_cmacs[size_t(prog[i])];
        needs_tout = needs_tout || b->prog_out_elems[size_t(prog[i])] > 1;
    }
    // group g's members in instance order (group -1: no group)
    std::vector<int32_t> seg(size_t(n_groups) + 1, 0), members;
    if (n_groups > 0) {
        for (int i = 0; i < n; ++i)
            if (group[i] >= 0) ++seg[size_t(group[i]) + 1];
        for (int k = 0; k < n_groups; ++k) seg[size_t(k) + 1] += seg[size_t(k)];
        members.resize(size_t(seg[size_t(n_groups)]));
        std::vector<int32_t> fill(seg.begin(), seg.end() - 1);
        for (int i = 0; i < n; ++i)
            if (group[i] >= 0) members[size_t(fill[size_t(group[i])]++)] = i;
    }
    // groups laid out as member j of group g = instance j * n_groups + g
    int stride_m = 0;
    if (n_groups > 0 && n % n_groups == 0) {
        const int m = n / n_groups;
        bool strided = int(members.size()) == n;
        for (int g = 0; g < n_groups && strided; ++g)
            for (int j = 0; j < m && strided; ++j)
                strided = seg[size_t(g)] == g * m && members[size_t(g * m + j)] == j * n_groups + g;
        if (strided) stride_m = m;
    }
    b->n_inst = 0;
    b->n_groups = 0;
    b->group_stride_m = 0;
    b->h_prog.assign(prog, prog + n);
    b->h_pbase.assign(param_base, param_base + n);
    qt_status s = qt::dev_put(ctx, &b->d_inst, &b->cap_inst, inst.data(), inst.size());
    b->sums_pending = false;  // the previous instances' results are gone
    b->pending_parts = false;
    if (s == QT_OK) s = qt::dev_put<double>(ctx, &b->d_out, &b->cap_out, nullptr, size_t(std::max(n, 1)) * 2);
    if (s == QT_OK && n_groups > 0)
        s = qt::dev_put<double>(ctx, &b->d_out_prev, &b->cap_out_prev, nullptr, size_t(std::max(n, 1)) * 2);
    if (s == QT_OK && n_groups > 0) {
        s = qt::dev_put(ctx, &b->d_seg, &b->cap_seg, seg.data(), seg.size());
        if (s == QT_OK) s = qt::dev_put(ctx, &b->d_members, &b->cap_members, members.data(), members.size());
        if (s == QT_OK) s = qt::dev_put<double>(ctx, &b->d_group, &b->cap_group, nullptr, size_t(n_groups));
    }
    if (s != QT_OK) return s;
    b->n_inst = n;
    b->n_groups = n_groups;
    b->group_stride_m = stride_m;
    b->needs_tout = needs_tout;
    b->tensor_out = false;  // output tensors belong to the instances they were set for
    b->cmacs_total = cmacs;
    b->c0 = c0;
    b->c1 = c1;
    ++b->layout_gen;
    if (b->specialised) {
        s = qt::specialised_chunks(b);
        if (s == QT_OK && b->mirror_stale) s = qt::refresh_pool_mirror(b);
        if (s != QT_OK || !qt::specialised_params_hold(b)) qt::drop_specialised(b);  // the table executor runs
        s = QT_OK;
    }
    return qt::configure_scratch(b);  // rebuilds the chunk list, drops a captured graph
}

qt_status qt_batch_set_outputs(qt_batch* b, int n, qt_tensor* const* outs) {
    if (!b || !outs) return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (n != b->n_inst || n == 0)
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_set_outputs: one output tensor per instance");
    const int dtype = b->precision == QT_PLAIN_C128 ? QT_DTYPE_C128 : QT_DTYPE_C64;
    std::vector<void*> ptrs(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        if (!outs[i]) return set_error(ctx, QT_INVALID_INPUT, "qt_batch_set_outputs: null tensor");
        const qt::Tensor& t = outs[i]->impl;
        if (t.dtype != dtype || t.numel != int64_t(b->prog_out_elems[size_t(b->h_prog[size_t(i)])]))
            return set_error(ctx, QT_INVALID_INPUT, "qt_batch_set_outputs: tensor " + std::to_string(i) +
                                                        " does not match its program's result");
        for (int j = 0; j < i; ++j)
            if (ptrs[size_t(j)] == t.data)
                return set_error(ctx, QT_INVALID_INPUT, "qt_batch_set_outputs: two instances share a tensor");
        ptrs[size_t(i)] = t.data;
    }
    qt::quiesce(b);  // a launch in flight reads the pointer array
    const size_t cap0 = b->cap_tout;
    qt_status s = qt::dev_put(ctx, &b->d_tout, &b->cap_tout, ptrs.data(), ptrs.size());
    if (s != QT_OK) {
        b->tensor_out = false;
        return s;
    }
    for (int i = 0; i < n; ++i) qt::release_exponent(ctx, &outs[i]->impl);  // results are plain values
    // the kernel variant changes with the first outputs; a grown array moves
    if (!b->tensor_out || b->cap_tout != cap0) {
        b->tensor_out = true;
        if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graph_exec = nullptr;
        b->graph = nullptr;
    }
    return QT_OK;
}

qt_status qt_batch_config(qt_batch* b, int lanes, int wgs_per_cu) {
    return qt_batch_config_ex(b, lanes, wgs_per_cu, QT_SCRATCH_AUTO);
}

qt_status qt_batch_config_ex(qt_batch* b, int lanes, int wgs_per_cu, int scratch) {
    if (!b || (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64 && lanes != 256) ||
        wgs_per_cu < 1 || wgs_per_cu > 16 || (scratch & ~(3 | QT_SCRATCH_STAGE_LEAVES)) != 0 || (scratch & 3) == 3)
        return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    qt::drop_specialised(b);  // an explicit table-executor shape replaces generated code
    const int l0 = b->lanes, w0 = b->wgs_per_cu, s0 = b->scratch_place;
    b->lanes = lanes;
    b->wgs_per_cu = wgs_per_cu;
    b->scratch_place = scratch;
    const qt_status s = qt::configure_scratch(b);
    if (s != QT_OK) {  // keep the previous, working shape
        b->lanes = l0;
        b->wgs_per_cu = w0;
        b->scratch_place = s0;
        (void)qt::configure_scratch(b);
    }
    return s;
}

qt_status qt_batch_launch(qt_batch* b) {
    b->in_use = true;  // queued work reads the batch's arrays
    if (!b) return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    qt_status s = qt::check_deadline(ctx, "qt_batch_launch");
    if (s != QT_OK) return s;
    (void)hipSetDevice(ctx->device);
    if (b->n_inst == 0) return QT_OK;
    s = b->graph_exec ? check_hip(ctx, hipGraphLaunch(b->graph_exec, ctx->stream), "hipGraphLaunch")
                                      : qt::launch_batch(b);
    // qt_batch_poll records the completion event when it is asked: an event recorded
    // after every launch costs the device about 5 us between consecutive launches
    if (s == QT_OK) b->ev_pending = true;
    return s;
}

int64_t qt_batch_resident_bytes(const qt_batch* b) {
    if (!b) return 0;
    const int64_t elem = b->precision == QT_PLAIN_C128 ? 16 : 8;
    return int64_t(b->cap_progs * sizeof(qt::ProgHdr) + b->cap_steps * sizeof(qt::StepDesc) +
                   b->cap_tabs * sizeof(uint32_t) + b->cap_refs * sizeof(int64_t) +
                   b->cap_leaf_lds * sizeof(int32_t) + b->cap_csts * sizeof(qt::CstEntry)) +
           b->pool_cap * elem;
}

qt_status qt_batch_set_sums_only(qt_batch* b, int on, int* active) {
    if (!b) return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (b->sums_only != (on != 0)) {
        b->sums_only = on != 0;
        // a captured graph holds the other kind of launch
        if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        b->graph_exec = nullptr;
        b->graph = nullptr;
    }
    if (active) *active = (b->specialised && b->sums_only && b->spec_sums_ok) ? 1 : 0;
    return QT_OK;
}

qt_status qt_batch_poll(qt_batch* b, int* done) {
    if (!b || !done) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    if (b->sums_pending || b->ev_pending) {  // the last launch is done once its group sums are
        (void)hipSetDevice(b->ctx->device);
        qt_status s = qt::flush_sums(b);
        if (s == QT_OK) s = check_hip(b->ctx, hipEventRecord(b->ev_done, b->ctx->stream), "event record");
        if (s != QT_OK) return s;
        b->ev_pending = false;
    }
    const hipError_t e = hipEventQuery(b->ev_done);
    if (e == hipErrorNotReady) {
        *done = 0;
        return QT_OK;
    }
    *done = 1;
    return check_hip(b->ctx, e, "batch poll");
}

qt_status qt_batch_capture(qt_batch* b) {
    if (!b) return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (b->graph_exec) return QT_OK;
    qt_status s = qt::flush_sums(b);
    if (s == QT_OK && b->specialised && b->sums_only && b->spec_sums_ok) s = qt::reserve_parts(b);
    if (s != QT_OK) return s;
    s = check_hip(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "begin capture");
    if (s != QT_OK) return s;
    qt_status ls = qt::launch_batch(b, true);
    hipGraph_t graph = nullptr;
    s = check_hip(ctx, hipStreamEndCapture(ctx->stream, &graph), "end capture");
    if (ls != QT_OK) s = ls;
    if (s != QT_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return s;
    }
    b->graph = graph;
    return check_hip(ctx, hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0), "graph instantiate");
}

qt_status qt_batch_uncapture(qt_batch* b) {
    if (!b) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);  // a replay in flight holds the graph
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    return QT_OK;
}

qt_status qt_batch_flush(qt_batch* b) {
    if (!b) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    (void)hipSetDevice(b->ctx->device);
    return qt::flush_sums(b);
}

qt_status qt_batch_results(qt_batch* b, double* values, double* group_sums) {
    if (!b) return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (values && !b->values_valid)
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_results: group-sums-only launches keep no per-instance values");
    qt_status s = qt::flush_sums(b);
    if (s == QT_OK) s = check_hip(ctx, hipStreamSynchronize(ctx->stream), "batch sync");
    if (s == QT_OK) b->in_use = false;
    if (s == QT_OK && b->params_unchecked) {
        int bad = 0;
        s = check_hip(ctx, hipMemcpy(&bad, b->d_param_status, sizeof bad, hipMemcpyDeviceToHost), "status");
        b->params_unchecked = false;
        if (s == QT_OK && bad) {
            s = qt::launch_fill_zero(ctx, b->d_param_status, 256);
            return set_error(ctx, QT_INVALID_INPUT,
                             "qt_batch_results: an angle given to qt_batch_rotation_params is outside the "
                             "supported range (0, or magnitude in [1e-16, 1e15))");
        }
    }
    if (s == QT_OK && values && b->n_inst > 0)
        s = check_hip(ctx, hipMemcpy(values, b->d_out, size_t(b->n_inst) * 16, hipMemcpyDeviceToHost), "results");
    if (s == QT_OK && group_sums && b->n_groups > 0)
        s = check_hip(ctx, hipMemcpy(group_sums, b->d_group, size_t(b->n_groups) * 8, hipMemcpyDeviceToHost),
                      "group sums");
    return s;
}

qt_status qt_batch_time(qt_batch* b, int reps, float* ms_per_launch) {
    if (!b || reps < 1 || !ms_per_launch) return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    qt_status s = check_hip(ctx, hipEventRecord(b->ev0, ctx->stream), "event record");
    for (int r = 0; r < reps && s == QT_OK; ++r) s = qt_batch_launch(b);
    if (s == QT_OK) s = qt::flush_sums(b);  // the last launch's sums are part of its work
    if (s == QT_OK) s = check_hip(ctx, hipEventRecord(b->ev1, ctx->stream), "event record");
    if (s == QT_OK) s = check_hip(ctx, hipEventSynchronize(b->ev1), "event sync");
    float ms = 0;
    if (s == QT_OK) s = check_hip(ctx, hipEventElapsedTime(&ms, b->ev0, b->ev1), "elapsed");
    *ms_per_launch = ms / float(reps);
    return s;
}

qt_status qt_batch_time_kernel(qt_batch* b, int reps, float* ms_per_launch) {
    b->in_use = true;  // queued work reads the batch's arrays
    if (!b || reps < 1 || !ms_per_launch) return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    *ms_per_launch = 0;
    if (b->n_inst == 0) return QT_OK;
    if (!b->shape_ok) return set_error(ctx, QT_INVALID_INPUT, "batch: executor shape not configured (last change failed)");
    if (b->needs_tout && !b->tensor_out)
        return set_error(ctx, QT_INVALID_INPUT, "batch: programs with tensor results need qt_batch_set_outputs");
    // a specialised kernel is timed as it runs in a stream of direct launches: each
    // launch also sums the previous launch's results (one untimed launch first)
    const bool sums_only = b->specialised && b->sums_only && b->spec_sums_ok;
    const bool defer = b->specialised && b->group_stride_m > 0 && b->d_out_prev;
    qt_status s = sums_only ? qt::launch_specialised_sums(b, false) : defer ? qt::launch_specialised(b, true) : QT_OK;
    if (s == QT_OK) s = check_hip(ctx, hipEventRecord(b->ev0, ctx->stream), "event record");
    for (int r = 0; r < reps && s == QT_OK; ++r) {
        if (sums_only) {
            s = qt::launch_specialised_sums(b, false);
        } else if (b->specialised) {
            s = qt::launch_specialised(b, defer);
        } else {
            hipLaunchKernelGGL(qt::k_reset_counter, dim3(1), dim3(1), 0, ctx->stream, b->d_counter);
            s = check_hip(ctx, hipGetLastError(), "k_reset_counter");
            if (s == QT_OK) s = qt::launch_tables(b);
        }
    }
    if (s == QT_OK) s = check_hip(ctx, hipEventRecord(b->ev1, ctx->stream), "event record");
    if (s == QT_OK) s = qt::flush_sums(b);  // after the timed region
    if (s == QT_OK) s = check_hip(ctx, hipEventSynchronize(b->ev1), "event sync");
    float ms = 0;
    if (s == QT_OK) s = check_hip(ctx, hipEventElapsedTime(&ms, b->ev0, b->ev1), "elapsed");
    *ms_per_launch = ms / float(reps);
    return s;
}

qt_status qt_batch_autotune(qt_batch* b, int* lanes, int* wgs_per_cu, float* ms) {
    if (!b) return QT_INVALID_INPUT;
    std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
    const int lane_opts[6] = {4, 8, 16, 32, 64, 256};
    const int wg_opts[5] = {1, 2, 4, 8, 16};
    int best_l = b->lanes, best_w = b->wgs_per_cu, best_p = b->scratch_place;
    float best = 1e30f;
    for (int l : lane_opts)
        for (int w : wg_opts)
            for (int place : {int(QT_SCRATCH_LDS), int(QT_SCRATCH_GLOBAL), QT_SCRATCH_LDS | QT_SCRATCH_STAGE_LEAVES,
                              QT_SCRATCH_GLOBAL | QT_SCRATCH_STAGE_LEAVES}) {
                if ((place & QT_SCRATCH_STAGE_LEAVES) && l >= 64) continue;
                if (qt_batch_config_ex(b, l, w, place) != QT_OK) continue;
                if (qt_batch_launch(b) != QT_OK) continue;  // warm-up
                float t = 0;
                if (qt_batch_time(b, 3, &t) != QT_OK) continue;
                if (t < best) {
                    best = t;
                    best_l = l;
                    best_w = w;
                    best_p = place;
                }
            }
    qt_status s = qt_batch_config_ex(b, best_l, best_w, best_p);
    if (lanes) *lanes = best_l;
    if (wgs_per_cu) *wgs_per_cu = best_w;
    if (ms) *ms = best;
    return s;
}

qt_status qt_batch_info(const qt_batch* b, int64_t* cmacs, int* lanes, int* wgs_per_cu, int* global_scratch,
                        int* scratch_elems) {
    if (!b) return QT_INVALID_INPUT;
    if (cmacs) *cmacs = b->cmacs_total;
    if (lanes) *lanes = b->lanes;
    if (wgs_per_cu) *wgs_per_cu = b->wgs_per_cu;
    if (global_scratch) *global_scratch = (b->use_global_scratch ? 1 : 0) | (b->stage_leaves ? 2 : 0);
    if (scratch_elems) *scratch_elems = b->scratch_elems;
    return QT_OK;
}

qt_status qt_batch_destroy(qt_batch* b) {
    if (!b) return QT_INVALID_INPUT;
    {
        std::lock_guard<std::recursive_mutex> g(b->ctx->mu);
        (void)hipSetDevice(b->ctx->device);
        (void)hipStreamSynchronize(b->ctx->stream);
        if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
        if (b->graph) (void)hipGraphDestroy(b->graph);
        if (b->ev0) (void)hipEventDestroy(b->ev0);
        if (b->ev1) (void)hipEventDestroy(b->ev1);
        if (b->ev_done) (void)hipEventDestroy(b->ev_done);
        qt::drop_specialised(b);
        for (void* p : {(void*)b->d_progs, (void*)b->d_steps, (void*)b->d_tabs, (void*)b->d_leaf_refs, b->d_pool,
                        (void*)b->d_inst, (void*)b->d_out, (void*)b->d_counter, (void*)b->d_seg, (void*)b->d_members,
                        (void*)b->d_group, b->d_global_scratch, (void*)b->d_chunks, b->d_pool_stage,
                        (void*)b->d_leaf_lds, (void*)b->d_csts, (void*)b->d_tout, (void*)b->d_schunks,
                        (void*)b->d_out_prev, (void*)b->d_param_status, (void*)b->d_angles, b->d_rot_desc})
            qt::batch_free(b->ctx, p);
        --b->ctx->live;
    }
    delete b;
    return QT_OK;
}

}  // extern "C"
