// Pairwise contraction kernels of libqtorch_hip.so (gfx950).
//
// C[c] = sum_s A[a(c,s)] * B[b(c,s)] (Network::ContractIndices, reference
// src/Network.h:876-971), on the bit-level description qt::BitMap.  Three regimes:
//
//   k_direct        : one lane per output element, the summed index in the inner
//                     loop in the reference's order (last pair fastest).  In the
//                     plain complex128 form floating-point contraction is off, so
//                     results are bit-identical to the reference's loop.  Covers
//                     every shape; chosen for small summed extents (memory-bound).
//   k_skinny        : at most 5 free bits on one side of C: one lane per element of
//                     the large side holds all outputs of the small side, summed in
//                     k_direct's order (bit-identical to it); the large operand is
//                     read once instead of once per small-side output.
//   k_reduce_*      : small outputs with a large summed extent (e.g. a scalar from
//                     two big tensors): the summed index is split over workgroups,
//                     wavefront/LDS reductions, then an ordered second pass.
//   k_permute +     : large GEMM-shaped steps: operands are rewritten by the
//   k_gemm_c128       LDS-staged bit-permute kernel into A'[s][m], B'[s][n], then a
//                     complex FP64 GEMM on MFMA (v_mfma_f64_16x16x4f64) writes
//                     C[m + M n], which is already the reference's C layout
//                     (A's free axes low, B's high: src/Network.h:739-769).
#include <algorithm>
#include <cstring>

#include "qt_internal.h"

namespace qt {

namespace {

constexpr int kThreads = 256;

enum { MODE_PLAIN = 0, MODE_4M = 1, MODE_3M = 2, MODE_BF16X3 = 3 };

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

// Accumulator of one complex output.  The 3M form keeps the three real sums.
template <class V, int MODE>
struct Acc {
    using T = typename Real<V>::T;
    T re = 0, im = 0, p3 = 0;
};

__device__ __forceinline__ void mac_plain(Acc<double2, MODE_PLAIN>& acc, double2 a, double2 b) {
#pragma clang fp contract(off)
    // std::complex<double> product then +=, exactly as the reference evaluates it
    const double pr = a.x * b.x - a.y * b.y;
    const double pi = a.x * b.y + a.y * b.x;
    acc.re = acc.re + pr;
    acc.im = acc.im + pi;
}

__device__ __forceinline__ float bf16_round(float x) {
    // round-to-nearest-even to bfloat16, returned widened to float
    uint32_t u = __float_as_uint(x);
    const uint32_t lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;
    u &= 0xffff0000u;
    return __uint_as_float(u);
}

template <class V, int MODE>
__device__ __forceinline__ void mac(Acc<V, MODE>& acc, V a, V b) {
    if constexpr (MODE == MODE_PLAIN) {
        mac_plain(acc, a, b);
    } else if constexpr (MODE == MODE_4M) {
        acc.re = fmaf(a.x, b.x, acc.re);
        acc.re = fmaf(-a.y, b.y, acc.re);
        acc.im = fmaf(a.x, b.y, acc.im);
        acc.im = fmaf(a.y, b.x, acc.im);
    } else if constexpr (MODE == MODE_3M) {
        acc.re = fmaf(a.x, b.x, acc.re);          // P1 = sum ar br
        acc.im = fmaf(a.y, b.y, acc.im);          // P2 = sum ai bi
        acc.p3 = fmaf(a.x + a.y, b.x + b.y, acc.p3);  // P3 = sum (ar+ai)(br+bi)
    } else {
        // bf16x3: x = hi + lo with hi, lo bfloat16; a*b ~ ah*bh + ah*bl + al*bh
        const float arh = bf16_round(a.x), arl = bf16_round(a.x - arh);
        const float aih = bf16_round(a.y), ail = bf16_round(a.y - aih);
        const float brh = bf16_round(b.x), brl = bf16_round(b.x - brh);
        const float bih = bf16_round(b.y), bil = bf16_round(b.y - bih);
        float re = acc.re, im = acc.im;
        re = fmaf(arh, brh, re); re = fmaf(arh, brl, re); re = fmaf(arl, brh, re);
        re = fmaf(-aih, bih, re); re = fmaf(-aih, bil, re); re = fmaf(-ail, bih, re);
        im = fmaf(arh, bih, im); im = fmaf(arh, bil, im); im = fmaf(arl, bih, im);
        im = fmaf(aih, brh, im); im = fmaf(aih, brl, im); im = fmaf(ail, brh, im);
        acc.re = re;
        acc.im = im;
    }
}

template <class V, int MODE>
__device__ __forceinline__ V finish(const Acc<V, MODE>& acc) {
    V r;
    if constexpr (MODE == MODE_3M) {
        r.x = acc.re - acc.im;
        r.y = acc.p3 - acc.re - acc.im;
    } else {
        r.x = acc.re;
        r.y = acc.im;
    }
    return r;
}

// offsets of the summed bits of s in A and B
__device__ __forceinline__ void s_offsets(const BitMap& m, int64_t s, int64_t& ga, int64_t& gb) {
    ga = 0;
    gb = 0;
    for (int j = 0; j < m.nbits_s; ++j)
        if ((s >> j) & 1) {
            ga += int64_t(1) << m.s_a[j];
            gb += int64_t(1) << m.s_b[j];
        }
}

__device__ __forceinline__ void c_offsets(const BitMap& m, int64_t c, int lo, int hi, int64_t& oa,
                                          int64_t& ob) {
    oa = 0;
    ob = 0;
    for (int i = lo; i < hi; ++i)
        if ((c >> i) & 1) {
            const int src = m.c_src[i];
            if (src < 64) oa += int64_t(1) << src;
            else ob += int64_t(1) << (src - 64);
        }
}

template <class V, int MODE>
__global__ __launch_bounds__(kThreads) void k_direct(const V* __restrict__ A, const V* __restrict__ B,
                                                     V* __restrict__ C, BitMap m, int64_t nC) {
    __shared__ int64_t low_a[kThreads], low_b[kThreads];
    __shared__ int64_t ga[kThreads], gb[kThreads];
    const int tid = threadIdx.x;
    const int lo = m.nbits_c < 8 ? m.nbits_c : 8;
    {
        int64_t oa, ob;
        c_offsets(m, tid, 0, lo, oa, ob);
        low_a[tid] = oa;
        low_b[tid] = ob;
    }
    const int64_t nS = int64_t(1) << m.nbits_s;
    const bool one_chunk = nS <= kThreads;
    if (one_chunk) {
        if (tid < nS) {
            int64_t a, b;
            s_offsets(m, tid, a, b);
            ga[tid] = a;
            gb[tid] = b;
        }
    }
    __syncthreads();
    const int64_t stride = int64_t(gridDim.x) * kThreads;
    for (int64_t base = int64_t(blockIdx.x) * kThreads; base < nC; base += stride) {
        int64_t ha, hb;
        c_offsets(m, base, lo, m.nbits_c, ha, hb);  // uniform across the workgroup
        const int64_t c = base + tid;
        const int64_t oa = ha + low_a[tid], ob = hb + low_b[tid];
        Acc<V, MODE> acc;
        if (one_chunk) {
            if (c < nC)
                for (int j = 0; j < int(nS); ++j) mac<V, MODE>(acc, A[oa + ga[j]], B[ob + gb[j]]);
        } else {
            for (int64_t s0 = 0; s0 < nS; s0 += kThreads) {
                __syncthreads();
                {
                    int64_t a, b;
                    s_offsets(m, s0 + tid, a, b);
                    ga[tid] = a;
                    gb[tid] = b;
                }
                __syncthreads();
                if (c < nC)
                    for (int j = 0; j < kThreads; ++j) mac<V, MODE>(acc, A[oa + ga[j]], B[ob + gb[j]]);
            }
        }
        if (c < nC) C[c] = finish<V, MODE>(acc);
    }
}

// Steps with at most 5 free bits on one side of C (a gate, or a thin tensor, against a
// large tensor; C = [A's free bits low | B's free bits high]).  One lane per element
// of the large side's part of C holds all 2^NB outputs of the small side, so the large
// operand is read once (k_direct reads it once per small-side output); the small
// operand's elements for a chunk of summed indices are staged in LDS and read there
// by every lane at once.  Each output is summed in k_direct's order with the same
// arithmetic: bit-identical to k_direct.
// SMALL_B: B is the small side, output o of lane t is C[t + (o << lo_bits)];
// otherwise A is (lo_bits == NB), output o of lane t is C[o + (t << NB)].
template <class V, int MODE, int NB, bool SMALL_B>
__global__ __launch_bounds__(kThreads) void k_skinny(const V* __restrict__ A, const V* __restrict__ B,
                                                     V* __restrict__ C, BitMap m, int lo_bits, int64_t n_big) {
    constexpr int NO = 1 << NB;
    constexpr int SC = NB >= 5 ? 32 : 64;  // summed indices per staged chunk (<= 16 KB complex128)
    constexpr int LU = sizeof(V) == 16 && NB >= 5 ? 4 : 8;  // loads in flight (registers: 175 at most)
    __shared__ int64_t g_big[SC];          // the large operand's offset of each summed index
    __shared__ V ys[SC][NO];               // the small operand's elements
    __shared__ int64_t small_off[NO];
    const int tid = threadIdx.x;
    if (tid < NO) {
        int64_t oa, ob;
        if constexpr (SMALL_B) {
            c_offsets(m, int64_t(tid) << lo_bits, lo_bits, m.nbits_c, oa, ob);
            small_off[tid] = ob;
        } else {
            c_offsets(m, tid, 0, NB, oa, ob);
            small_off[tid] = oa;
        }
    }
    const int64_t nS = int64_t(1) << m.nbits_s;
    const int64_t stride = int64_t(gridDim.x) * kThreads;
    for (int64_t base = int64_t(blockIdx.x) * kThreads; base < n_big; base += stride) {
        const int64_t t = base + tid;
        int64_t big_off;
        {
            int64_t oa, ob;
            if constexpr (SMALL_B) {
                c_offsets(m, t, 0, lo_bits, oa, ob);
                big_off = oa;
            } else {
                c_offsets(m, t << NB, NB, m.nbits_c, oa, ob);
                big_off = ob;
            }
        }
        Acc<V, MODE> acc[NO];
        for (int64_t s0 = 0; s0 < nS; s0 += SC) {
            const int n_s = int(nS - s0 < SC ? nS - s0 : SC);
            __syncthreads();  // small_off is written; the previous chunk's reads are done
            for (int e = tid; e < n_s * NO; e += kThreads) {
                const int j = e / NO, o = e % NO;
                int64_t a, b;
                s_offsets(m, s0 + j, a, b);
                if (o == 0) g_big[j] = SMALL_B ? a : b;
                ys[j][o] = SMALL_B ? B[small_off[o] + b] : A[small_off[o] + a];
            }
            __syncthreads();
            if (t < n_big)
                // LU loads of the large operand in flight before their products (one at a
                // time left the kernel waiting on each load: 14-20 % slower on 28-bit
                // slices); the products keep j's order
                for (int j0 = 0; j0 < n_s; j0 += LU) {
                    V x[LU];
#pragma unroll
                    for (int u = 0; u < LU; ++u)
                        if (j0 + u < n_s) x[u] = SMALL_B ? A[big_off + g_big[j0 + u]] : B[big_off + g_big[j0 + u]];
#pragma unroll
                    for (int u = 0; u < LU; ++u)
                        if (j0 + u < n_s)
#pragma unroll
                            for (int o = 0; o < NO; ++o) {
                                if constexpr (SMALL_B) mac<V, MODE>(acc[o], x[u], ys[j0 + u][o]);
                                else mac<V, MODE>(acc[o], ys[j0 + u][o], x[u]);
                            }
                }
        }
        if (t < n_big)
#pragma unroll
            for (int o = 0; o < NO; ++o) {
                const int64_t idx = SMALL_B ? t + (int64_t(o) << lo_bits) : int64_t(o) + (t << NB);
                C[idx] = finish<V, MODE>(acc[o]);
            }
    }
}

// k_skinny's steps when the large side's free bits keep their order in C (a gate on
// a large tensor, as in a circuit sweep) and at most 4 bits are summed: lane t's
// element of the large operand X sits at t with zero bits inserted at X's summed
// positions (a few shifts instead of a walk over every bit of C), the small operand's
// elements are uniform (scalar loads from the kernel's arguments' offsets), and each
// lane issues its 2^NS loads of X before any arithmetic.  Same summation order and
// arithmetic as k_skinny: bit-identical to it.
struct ThinArgs {
    int64_t n_big;
    int32_t lo_bits;        // small B: output o of lane t is C[t + (o << lo_bits)]
    int32_t n_ins;          // X's summed positions, ascending
    int32_t ins_pos[4];
    int64_t g_big[16];      // X offset of summed index j (k_direct's digit order)
    int32_t g_small[16];    // small operand offset of summed index j
    int32_t small_off[16];  // small operand offset of its free part o
};

template <class V, int MODE, int NS, int NB, bool SMALL_B>
__global__ __launch_bounds__(kThreads) void k_thin(const V* __restrict__ X, const V* __restrict__ G,
                                                   V* __restrict__ C, ThinArgs r) {
    constexpr int NJ = 1 << NS, NO = 1 << NB;
    const int64_t stride = int64_t(gridDim.x) * kThreads;
    for (int64_t t = int64_t(blockIdx.x) * kThreads + threadIdx.x; t < r.n_big; t += stride) {
        int64_t x = t;
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            const int p = r.ins_pos[i];
            x = ((x >> p) << (p + 1)) | (x & ((int64_t(1) << p) - 1));
        }
        V xs[NJ];
#pragma unroll
        for (int j = 0; j < NJ; ++j) xs[j] = X[x + r.g_big[j]];
        Acc<V, MODE> acc[NO];
#pragma unroll
        for (int j = 0; j < NJ; ++j)
#pragma unroll
            for (int o = 0; o < NO; ++o) {
                const V g = G[r.small_off[o] + r.g_small[j]];
                if constexpr (SMALL_B) mac<V, MODE>(acc[o], xs[j], g);
                else mac<V, MODE>(acc[o], g, xs[j]);
            }
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            const int64_t idx = SMALL_B ? t + (int64_t(o) << r.lo_bits) : int64_t(o) + (t << NB);
            C[idx] = finish<V, MODE>(acc[o]);
        }
    }
}

// Split reduction for outputs of at most kReduceMaxOut elements.  Block b sums the
// summed-index chunk [b*chunk, (b+1)*chunk) for every output and writes
// partials[b * nC + c]; k_reduce_final adds the partials in block order.
constexpr int kReduceMaxOut = 64;
constexpr int64_t kReduceMaxBlocks = int64_t(1) << 15;

template <class V>
__device__ __forceinline__ V shfl_down_v(V v, int d) {
    V r;
    r.x = __shfl_down(v.x, d, 64);
    r.y = __shfl_down(v.y, d, 64);
    return r;
}

// Rows of an operand: its elements that share one value of its free bits.  One pass of
// k_reduce_rows reads GA rows of A and GB rows of B and accumulates the GA x GB outputs
// they form, so an operand element is read once per pass, not once per output of the
// other operand.  The summed index s splits into the lane (bits 0-7), the workgroup
// (the next log2(gridDim.x) bits) and the step (the bits above), so the workgroups that
// run together read neighbouring 256-element runs (with a contiguous chunk of s per
// workgroup the kernel measured 2.7x slower on 2^24 summed indices).  Every output is
// summed in the same order whatever the grouping: each lane over its steps in order,
// then the lanes of a wave by shuffles, then the waves in order.
struct RowSet {
    int64_t oa[16], ob[16];  // free-bit offsets of the pass's rows of A and of B
    int16_t c[16];           // output index of row pair (i, j) at c[i * GB + j]
};

template <class V, int MODE, int GA, int GB>
__global__ __launch_bounds__(kThreads) void k_reduce_rows(const V* __restrict__ A, const V* __restrict__ B,
                                                          V* __restrict__ partials, BitMap m, int64_t nC,
                                                          int step_bits, RowSet r) {
    __shared__ V wsum[GA * GB][kThreads / 64];
    // offsets of the step bits 0-5 and 6-11: plan bytes read inside the step loop
    // would be vector loads whose waits also wait for every operand load in flight
    __shared__ int64_t t_lo[2][64], t_hi[2][64];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int first_step_bit = m.nbits_s - step_bits;  // 8 + log2(gridDim.x)
    int64_t la = 0, lb = 0;
    for (int j = 0; j < 8; ++j)
        if ((tid >> j) & 1) {
            la += int64_t(1) << m.s_a[j];
            lb += int64_t(1) << m.s_b[j];
        }
    if (tid < 128) {
        const int x = tid & 63, first = first_step_bit + (tid < 64 ? 0 : 6);
        int64_t oa = 0, ob = 0;
        for (int k = 0; k < 6; ++k)
            if (((x >> k) & 1) && first + k < m.nbits_s) {
                oa += int64_t(1) << m.s_a[first + k];
                ob += int64_t(1) << m.s_b[first + k];
            }
        if (tid < 64) {
            t_lo[0][x] = oa;
            t_lo[1][x] = ob;
        } else {
            t_hi[0][x] = oa;
            t_hi[1][x] = ob;
        }
    }
    int64_t base_a = la, base_b = lb;  // lane and workgroup bits
    for (int j = 8; j < first_step_bit; ++j)
        if ((blockIdx.x >> (j - 8)) & 1) {
            base_a += int64_t(1) << m.s_a[j];
            base_b += int64_t(1) << m.s_b[j];
        }
    __syncthreads();
    Acc<V, MODE> acc[GA][GB];
    // U steps per iteration, all loads issued before the first multiply (the steps are
    // a power of two of at least 16, so U divides them).  Measured on 2^24 summed
    // indices: 0.19 ms; with the U loads of an iteration on neighbouring runs instead
    // of U steps apart, 0.60 ms
    constexpr int U0 = GA + GB >= 16 ? 1 : 16 / (GA + GB);
    constexpr int U = U0 >= 8 ? 8 : U0 >= 4 ? 4 : U0 >= 2 ? 2 : 1;
    const int steps = 1 << step_bits;
    for (int k0 = 0; k0 < steps; k0 += U) {
        V a[U][GA], b[U][GB];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u;
            const int64_t ha = base_a + t_lo[0][k & 63] + t_hi[0][k >> 6];
            const int64_t hb = base_b + t_lo[1][k & 63] + t_hi[1][k >> 6];
#pragma unroll
            for (int i = 0; i < GA; ++i) a[u][i] = A[r.oa[i] + ha];
#pragma unroll
            for (int j = 0; j < GB; ++j) b[u][j] = B[r.ob[j] + hb];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int i = 0; i < GA; ++i)
#pragma unroll
                for (int j = 0; j < GB; ++j) mac<V, MODE>(acc[i][j], a[u][i], b[u][j]);
    }
#pragma unroll
    for (int q = 0; q < GA * GB; ++q) {
        V v = finish<V, MODE>(acc[q / GB][q % GB]);
        for (int d = 32; d >= 1; d >>= 1) {
            V o = shfl_down_v(v, d);
            v.x += o.x;
            v.y += o.y;
        }
        if (lane == 0) wsum[q][wave] = v;
    }
    __syncthreads();
    if (tid < GA * GB) {
        V t = wsum[tid][0];
        for (int w = 1; w < kThreads / 64; ++w) {
            t.x += wsum[tid][w].x;
            t.y += wsum[tid][w].y;
        }
        partials[int64_t(blockIdx.x) * nC + r.c[tid]] = t;
    }
}

template <class V>
__global__ __launch_bounds__(kThreads) void k_reduce_final(const V* __restrict__ partials, V* __restrict__ C,
                                                           int64_t nC, int64_t nblocks) {
    // one workgroup per output: strided partial sums, then a fixed-shape tree (the
    // order does not depend on timing, so the result is reproducible)
    __shared__ V part[kThreads];
    const int64_t c = blockIdx.x;
    V t;
    t.x = 0;
    t.y = 0;
    for (int64_t b = threadIdx.x; b < nblocks; b += kThreads) {
        t.x += partials[b * nC + c].x;
        t.y += partials[b * nC + c].y;
    }
    part[threadIdx.x] = t;
    __syncthreads();
    for (int d = kThreads / 2; d >= 1; d >>= 1) {
        if (int(threadIdx.x) < d) {
            part[threadIdx.x].x += part[threadIdx.x + d].x;
            part[threadIdx.x].y += part[threadIdx.x + d].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) C[c] = part[0];
}

// LDS-staged bit permutation (see PermPlan).  Reads walk the input-low bits
// contiguously, writes walk the output-low bits contiguously.
// Tile element t sits at slot t with its low 6 bits xor-ed with its row t / 64: a read
// that walks the tile 64 elements at a time (an input-low bit moved to an output-high
// position) hits a different bank with every lane, and the tile stays 2^tbits slots.
// P = 1 (complex64): each lane moves two elements that are neighbours on the read side
// (input bit 0) as one 16-byte load, and two that are neighbours on the write side
// (output bit 0) as one 16-byte store; 8-byte accesses measured well below the rate of
// 16-byte ones.
__device__ __forceinline__ int tile_slot(int t) { return t ^ ((t >> 6) & 63); }

template <class V, int P>
__global__ __launch_bounds__(kThreads) void k_permute(const V* __restrict__ in, V* __restrict__ out,
                                                      PermPlan p) {
    extern __shared__ unsigned char smem_raw[];
    V* tile = reinterpret_cast<V*>(smem_raw);
    const int64_t rb = int64_t(blockIdx.x) + int64_t(blockIdx.y) * gridDim.x;
    int64_t base_in = 0, base_out = 0;
    for (int i = 0; i < p.n_rest; ++i)
        if ((rb >> i) & 1) {
            base_in += int64_t(1) << p.rest_in[i];
            base_out += int64_t(1) << p.rest_out[i];
        }
    const int tbits = p.tl + p.n_extra_in;
    const int n = 1 << tbits;
    const int tid = threadIdx.x;
    // element e = (tid << P) + (q << (8 + P)): the offsets split into the thread's part
    // (tile bits P to P + 7, computed once) and the step's part (bits 8 + P and up,
    // uniform); with P = 1 bit 0 selects the pair's second element
    auto in_bit = [&](int b) -> int64_t { return b < p.tl ? int64_t(1) << b : int64_t(1) << p.extra_in_pos[b - p.tl]; };
    auto out_bit = [&](int b) -> int64_t { return b < p.tl ? int64_t(1) << b : int64_t(1) << p.extra_out_pos[b - p.tl]; };
    auto tile_bit = [&](int b) -> int { return b < p.tl ? p.lowout_tile[b] : p.extra_out_tile[b - p.tl]; };
    constexpr int hi = 8 + P;
    const int lo_bits = tbits < hi ? tbits : hi;
    int64_t in_t = base_in, out_t = base_out;
    int slot_t = 0;
    for (int b = P; b < lo_bits; ++b)
        if ((tid >> (b - P)) & 1) {
            in_t += in_bit(b);
            out_t += out_bit(b);
            slot_t |= 1 << tile_bit(b);
        }
    constexpr int per = kThreads << P;
    const int steps = n > per ? n / per : 1;
    // the step bits' offsets, read from the plan before the first data load: a plan
    // byte read inside the step loops is a vector load whose wait also waits for every
    // element load in flight
    constexpr int kStepBits = 4;  // tiles of at most 2^12 elements
    int64_t step_in[kStepBits], step_out[kStepBits];
    int step_tile[kStepBits];
#pragma unroll
    for (int k = 0; k < kStepBits; ++k) {
        const bool on = hi + k < tbits;
        step_in[k] = on ? in_bit(hi + k) : 0;
        step_out[k] = on ? out_bit(hi + k) : 0;
        step_tile[k] = on ? 1 << tile_bit(hi + k) : 0;
    }
    const int pair_bit = P == 1 ? 1 << tile_bit(0) : 0;
    if ((tid << P) < n) {
        for (int q = 0; q < steps; ++q) {
            int64_t off = in_t;
#pragma unroll
            for (int k = 0; k < kStepBits; ++k)
                if ((q >> k) & 1) off += step_in[k];
            const int t = (tid << P) + (q << hi);
            if constexpr (P == 1) {
                const float4 v = *reinterpret_cast<const float4*>(in + off);
                tile[tile_slot(t)] = make_float2(v.x, v.y);
                tile[tile_slot(t + 1)] = make_float2(v.z, v.w);
            } else {
                tile[tile_slot(t)] = in[off];
            }
        }
    }
    __syncthreads();
    if ((tid << P) < n) {
        for (int q = 0; q < steps; ++q) {
            int64_t off = out_t;
            int t = slot_t;
#pragma unroll
            for (int k = 0; k < kStepBits; ++k)
                if ((q >> k) & 1) {
                    off += step_out[k];
                    t |= step_tile[k];
                }
            if constexpr (P == 1) {
                const V a = tile[tile_slot(t)], b2 = tile[tile_slot(t | pair_bit)];
                *reinterpret_cast<float4*>(out + off) = make_float4(a.x, a.y, b2.x, b2.y);
            } else {
                out[off] = tile[tile_slot(t)];
            }
        }
    }
}

// Complex FP64 GEMM on MFMA: C[m + M n] = sum_s A'[s][m] * B'[s][n].
// 64x64 output tile per workgroup, four waves of 32x32; each 16x16 block of a wave
// takes four real v_mfma_f64_16x16x4f64 per k-step of 4 (re*re, -im*im, re*im, im*re).
// The MFMA row operand carries B' (rows = n) and the column operand A' (cols = m), so
// the lanes of one accumulator row hold 16 consecutive m and store contiguously.
constexpr int GM = 64, GN = 64, GK = 16;
typedef double d4_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kThreads) void k_gemm_c128(const double2* __restrict__ Ap,
                                                        const double2* __restrict__ Bp,
                                                        double2* __restrict__ C, int64_t M,
                                                        int64_t N, int64_t K) {
    __shared__ double as_re[GK][GM + 1], as_im[GK][GM + 1];
    __shared__ double bs_re[GK][GN + 1], bs_im[GK][GN + 1];
    // split K: workgroups with blockIdx.z = z take summed rows [z K, (z + 1) K) and write
    // their own M x N partial product at C + z M N
    Ap += int64_t(blockIdx.z) * K * M;
    Bp += int64_t(blockIdx.z) * K * N;
    C += int64_t(blockIdx.z) * M * N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t tiles_m = M / GM;
    const int64_t tile = int64_t(blockIdx.x) + int64_t(blockIdx.y) * gridDim.x;
    if (tile >= tiles_m * (N / GN)) return;
    const int64_t m0 = (tile % tiles_m) * GM, n0 = (tile / tiles_m) * GN;
    d4_t acc_re[2][2], acc_im[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            acc_re[i][j] = d4_t{0, 0, 0, 0};
            acc_im[i][j] = d4_t{0, 0, 0, 0};
        }
    const int li = lane & 15, lk = lane >> 4;
    // each thread stages kPer elements of A' and of B' per k-tile; the next k-tile is
    // loaded into registers while the current one is multiplied from LDS.  A last
    // k-tile beyond K (K = 8) is filled with zeros, which add nothing.
    constexpr int kPer = GK * GM / kThreads;
    static_assert(GK * GM == GK * GN && kPer * kThreads == GK * GM, "tile staging");
    double2 ra[kPer], rb[kPer];
    auto fetch = [&](int64_t k0) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kThreads, kk = e / GM, rr = e % GM;
            const bool in_k = k0 + kk < K;
            ra[q] = in_k ? Ap[(k0 + kk) * M + m0 + rr] : double2{0.0, 0.0};
            rb[q] = in_k ? Bp[(k0 + kk) * N + n0 + rr] : double2{0.0, 0.0};
        }
    };
    fetch(0);
    for (int64_t k0 = 0; k0 < K; k0 += GK) {
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = tid + q * kThreads, kk = e / GM, rr = e % GM;
            as_re[kk][rr] = ra[q].x;
            as_im[kk][rr] = ra[q].y;
            bs_re[kk][rr] = rb[q].x;
            bs_im[kk][rr] = rb[q].y;
        }
        __syncthreads();
        if (k0 + GK < K) fetch(k0 + GK);
#pragma unroll
        for (int kk = 0; kk < GK; kk += 4) {
            const int kl = kk + lk;
            double are[2], aim[2];
#pragma unroll
            for (int bm = 0; bm < 2; ++bm) {
                const int mm = wm * 32 + bm * 16 + li;
                are[bm] = as_re[kl][mm];
                aim[bm] = as_im[kl][mm];
            }
            double bre[2], bim[2];
#pragma unroll
            for (int bn = 0; bn < 2; ++bn) {
                const int nn = wn * 32 + bn * 16 + li;
                bre[bn] = bs_re[kl][nn];
                bim[bn] = bs_im[kl][nn];
            }
            // the four products of a block go to two accumulators; issuing them in
            // four rounds over the blocks keeps dependent MFMAs apart
#pragma unroll
            for (int bn = 0; bn < 2; ++bn)
#pragma unroll
                for (int bm = 0; bm < 2; ++bm)
                    acc_re[bn][bm] = __builtin_amdgcn_mfma_f64_16x16x4f64(bre[bn], are[bm], acc_re[bn][bm], 0, 0, 0);
#pragma unroll
            for (int bn = 0; bn < 2; ++bn)
#pragma unroll
                for (int bm = 0; bm < 2; ++bm)
                    acc_im[bn][bm] = __builtin_amdgcn_mfma_f64_16x16x4f64(bim[bn], are[bm], acc_im[bn][bm], 0, 0, 0);
#pragma unroll
            for (int bn = 0; bn < 2; ++bn)
#pragma unroll
                for (int bm = 0; bm < 2; ++bm)
                    acc_re[bn][bm] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bim[bn], aim[bm], acc_re[bn][bm], 0, 0, 0);
#pragma unroll
            for (int bn = 0; bn < 2; ++bn)
#pragma unroll
                for (int bm = 0; bm < 2; ++bm)
                    acc_im[bn][bm] = __builtin_amdgcn_mfma_f64_16x16x4f64(bre[bn], aim[bm], acc_im[bn][bm], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
        for (int bm = 0; bm < 2; ++bm)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // f64 16x16x4 accumulator: item r of lane l is row (l / 16) + 4 r, column l % 16
                const int64_t n = n0 + wn * 32 + bn * 16 + lk + 4 * r;
                const int64_t mm = m0 + wm * 32 + bm * 16 + li;
                C[mm + M * n] = double2{acc_re[bn][bm][r], acc_im[bn][bm][r]};
            }
}

// Complex64 GEMM on MFMA, C[m + M n] = sum_s A'[s][m] * B'[s][n], in the three
// complex64 forms.  64x64 output tile per workgroup, four waves of 32x32, k-tiles
// of 32.  4M and 3M use v_mfma_f32_16x16x4f32 (four / three real products per
// complex product); bf16x3 splits every FP32 operand into bf16 hi + lo when staging
// it in LDS and takes hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_bf16, all with
// FP32 accumulation.  f32 16x16 accumulator layout: item r of lane l is row
// 4 (l / 16) + r, column l % 16; rows are n (the B' operand), columns m (A').
constexpr int HK = 32;
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t bf16_bits(float x) {  // round to nearest even
    return uint16_t(__float_as_uint(bf16_round(x)) >> 16);
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_gemm_c64(const float2* __restrict__ Ap, const float2* __restrict__ Bp,
                                                       float2* __restrict__ C, int64_t M, int64_t N, int64_t K) {
    // f32 forms: [k][row] planes of re / im; bf16x3: [row][k] planes of re_hi, re_lo,
    // im_hi, im_lo (k contiguous: one 16-byte load is a lane's 8 operands)
    constexpr int kPlaneF = HK * (GM + 1);
    constexpr int kRowB = HK + 8;
    constexpr int kPlaneB = GM * kRowB;
    __shared__ __align__(16) unsigned char smem[MODE == MODE_BF16X3 ? 2 * 4 * kPlaneB * 2 : 2 * 2 * kPlaneF * 4];
    // split K: workgroups with blockIdx.z = z take summed rows [z K, (z + 1) K) and write
    // their own M x N partial product at C + z M N
    Ap += int64_t(blockIdx.z) * K * M;
    Bp += int64_t(blockIdx.z) * K * N;
    C += int64_t(blockIdx.z) * M * N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t tiles_m = M / GM;
    const int64_t tile = int64_t(blockIdx.x) + int64_t(blockIdx.y) * gridDim.x;
    if (tile >= tiles_m * (N / GN)) return;
    const int64_t m0 = (tile % tiles_m) * GM, n0 = (tile / tiles_m) * GN;
    f4_t acc0[2][2], acc1[2][2], acc2[2][2];  // 4M / bf16x3: re, im; 3M: P1, P2, P3
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc0[i][j] = acc1[i][j] = acc2[i][j] = f4_t{0, 0, 0, 0};
    for (int64_t k0 = 0; k0 < K; k0 += HK) {
        if constexpr (MODE == MODE_BF16X3) {
            uint16_t* sa = reinterpret_cast<uint16_t*>(smem);  // 4 planes of A', then 4 of B'
            uint16_t* sb = sa + 4 * kPlaneB;
            for (int e = tid; e < HK * GM; e += kThreads) {
                const int kk = e / GM, rr = e % GM;
                const bool in_k = k0 + kk < K;  // a last k-tile beyond K is zero-filled
                const float2 va = in_k ? Ap[(k0 + kk) * M + m0 + rr] : float2{0.0f, 0.0f};
                const float2 vb = in_k ? Bp[(k0 + kk) * N + n0 + rr] : float2{0.0f, 0.0f};
                const int o = rr * kRowB + kk;
                const float arh = bf16_round(va.x), aih = bf16_round(va.y);
                const float brh = bf16_round(vb.x), bih = bf16_round(vb.y);
                sa[o] = bf16_bits(arh);
                sa[kPlaneB + o] = bf16_bits(va.x - arh);
                sa[2 * kPlaneB + o] = bf16_bits(aih);
                sa[3 * kPlaneB + o] = bf16_bits(va.y - aih);
                sb[o] = bf16_bits(brh);
                sb[kPlaneB + o] = bf16_bits(vb.x - brh);
                sb[2 * kPlaneB + o] = bf16_bits(bih);
                sb[3 * kPlaneB + o] = bf16_bits(vb.y - bih);
            }
            __syncthreads();
            auto ld = [&](const uint16_t* plane, int row) {
                return *reinterpret_cast<const bf8_t*>(plane + row * kRowB + 8 * lk);
            };
            auto neg = [](bf8_t v) {
                uint4 u = *reinterpret_cast<uint4*>(&v);
                u.x ^= 0x80008000u;
                u.y ^= 0x80008000u;
                u.z ^= 0x80008000u;
                u.w ^= 0x80008000u;
                return *reinterpret_cast<bf8_t*>(&u);
            };
            bf8_t a[2][4];
#pragma unroll
            for (int bm = 0; bm < 2; ++bm)
#pragma unroll
                for (int q = 0; q < 4; ++q) a[bm][q] = ld(sa + q * kPlaneB, wm * 32 + bm * 16 + li);
            bf8_t b[2][6];  // per bn: re_hi, re_lo, im_hi, im_lo, -im_hi, -im_lo
#pragma unroll
            for (int bn = 0; bn < 2; ++bn) {
#pragma unroll
                for (int q = 0; q < 4; ++q) b[bn][q] = ld(sb + q * kPlaneB, wn * 32 + bn * 16 + li);
                b[bn][4] = neg(b[bn][2]);
                b[bn][5] = neg(b[bn][3]);
            }
            // re += br ar - bi ai; im += bi ar + br ai, each product hi*hi + hi*lo + lo*hi;
            // every round touches each of the eight accumulators once, so no MFMA waits
            // for the one before it.  (B operand index, A operand index) per round:
            constexpr int kRe[6][2] = {{0, 0}, {0, 1}, {1, 0}, {4, 2}, {4, 3}, {5, 2}};
            constexpr int kIm[6][2] = {{2, 0}, {2, 1}, {3, 0}, {0, 2}, {0, 3}, {1, 2}};
#pragma unroll
            for (int t = 0; t < 6; ++t)
#pragma unroll
                for (int bn = 0; bn < 2; ++bn)
#pragma unroll
                    for (int bm = 0; bm < 2; ++bm) {
                        acc0[bn][bm] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[bn][kRe[t][0]], a[bm][kRe[t][1]],
                                                                               acc0[bn][bm], 0, 0, 0);
                        acc1[bn][bm] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[bn][kIm[t][0]], a[bm][kIm[t][1]],
                                                                               acc1[bn][bm], 0, 0, 0);
                    }
        } else {
            float* sa = reinterpret_cast<float*>(smem);  // re plane, im plane of A', then of B'
            float* sb = sa + 2 * kPlaneF;
            for (int e = tid; e < HK * GM; e += kThreads) {
                const int kk = e / GM, rr = e % GM;
                const bool in_k = k0 + kk < K;  // a last k-tile beyond K is zero-filled
                const float2 va = in_k ? Ap[(k0 + kk) * M + m0 + rr] : float2{0.0f, 0.0f};
                const float2 vb = in_k ? Bp[(k0 + kk) * N + n0 + rr] : float2{0.0f, 0.0f};
                const int o = kk * (GM + 1) + rr;
                sa[o] = va.x;
                sa[kPlaneF + o] = va.y;
                sb[o] = vb.x;
                sb[kPlaneF + o] = vb.y;
            }
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < HK; kk += 4) {
                const int o = (kk + lk) * (GM + 1);
                float are[2], aim[2];
#pragma unroll
                for (int bm = 0; bm < 2; ++bm) {
                    are[bm] = sa[o + wm * 32 + bm * 16 + li];
                    aim[bm] = sa[kPlaneF + o + wm * 32 + bm * 16 + li];
                }
                float bre[2], bim[2];
#pragma unroll
                for (int bn = 0; bn < 2; ++bn) {
                    bre[bn] = sb[o + wn * 32 + bn * 16 + li];
                    bim[bn] = sb[kPlaneF + o + wn * 32 + bn * 16 + li];
                }
                // rounds over the four blocks (dependent MFMAs four apart)
                if constexpr (MODE == MODE_3M) {
#pragma unroll
                    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
                        for (int bm = 0; bm < 2; ++bm) {
                            acc0[bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(bre[bn], are[bm], acc0[bn][bm], 0, 0, 0);
                            acc1[bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(bim[bn], aim[bm], acc1[bn][bm], 0, 0, 0);
                            acc2[bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(bre[bn] + bim[bn], are[bm] + aim[bm],
                                                                               acc2[bn][bm], 0, 0, 0);
                        }
                } else {
#pragma unroll
                    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
                        for (int bm = 0; bm < 2; ++bm) {
                            acc0[bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(bre[bn], are[bm], acc0[bn][bm], 0, 0, 0);
                            acc1[bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(bim[bn], are[bm], acc1[bn][bm], 0, 0, 0);
                        }
#pragma unroll
                    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
                        for (int bm = 0; bm < 2; ++bm) {
                            acc0[bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(-bim[bn], aim[bm], acc0[bn][bm], 0, 0, 0);
                            acc1[bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(bre[bn], aim[bm], acc1[bn][bm], 0, 0, 0);
                        }
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int bn = 0; bn < 2; ++bn)
#pragma unroll
        for (int bm = 0; bm < 2; ++bm)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t n = n0 + wn * 32 + bn * 16 + 4 * lk + r;
                const int64_t mm = m0 + wm * 32 + bm * 16 + li;
                float2 v;
                if constexpr (MODE == MODE_3M) {
                    v.x = acc0[bn][bm][r] - acc1[bn][bm][r];
                    v.y = acc2[bn][bm][r] - acc0[bn][bm][r] - acc1[bn][bm][r];
                } else {
                    v.x = acc0[bn][bm][r];
                    v.y = acc1[bn][bm][r];
                }
                C[mm + M * n] = v;
            }
}

__global__ void k_fill_zero(uint4* p, int64_t n16) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n16; i += stride)
        p[i] = uint4{0, 0, 0, 0};
}

inline unsigned grid_for(int64_t items, int64_t per_block, unsigned cap) {
    int64_t g = (items + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return unsigned(g);
}

template <class V, int MODE>
qt_status run_direct(Ctx* c, const V* A, const V* B, V* C, const BitMap& m) {
    const int64_t nC = int64_t(1) << m.nbits_c;
    const unsigned grid = grid_for(nC, kThreads, 16384);
    hipLaunchKernelGGL((k_direct<V, MODE>), dim3(grid), dim3(kThreads), 0, c->stream, A, B, C, m, nC);
    return check_hip(c, hipGetLastError(), "k_direct");
}

template <class V, int MODE, int NB>
qt_status launch_skinny(Ctx* c, const V* A, const V* B, V* C, const BitMap& m, int mbits, bool small_b) {
    const int64_t n_big = int64_t(1) << (m.nbits_c - NB);
    const unsigned grid = grid_for(n_big, kThreads, 16384);
    if (small_b)
        hipLaunchKernelGGL((k_skinny<V, MODE, NB, true>), dim3(grid), dim3(kThreads), 0, c->stream, A, B, C, m,
                           mbits, n_big);
    else
        hipLaunchKernelGGL((k_skinny<V, MODE, NB, false>), dim3(grid), dim3(kThreads), 0, c->stream, A, B, C, m,
                           mbits, n_big);
    return check_hip(c, hipGetLastError(), "k_skinny");
}

template <class V, int MODE, int NS, int NB>
qt_status launch_thin_nb(Ctx* c, const V* X, const V* G, V* C, const ThinArgs& r, bool small_b) {
    // one element per lane, no grid-stride reuse: on MI355X streaming kernels reach
    // their highest HBM rates with a workgroup per 256 elements (a capped grid looping
    // over the tensor measured 10-30 % slower, two adjacent elements per lane 5-8 %)
    const unsigned grid = grid_for(r.n_big, kThreads, 1u << 30);
    if (small_b)
        hipLaunchKernelGGL((k_thin<V, MODE, NS, NB, true>), dim3(grid), dim3(kThreads), 0, c->stream, X, G, C, r);
    else
        hipLaunchKernelGGL((k_thin<V, MODE, NS, NB, false>), dim3(grid), dim3(kThreads), 0, c->stream, X, G, C, r);
    return check_hip(c, hipGetLastError(), "k_thin");
}

template <class V, int MODE, int NS>
qt_status launch_thin_ns(Ctx* c, const V* X, const V* G, V* C, const ThinArgs& r, int nb, bool small_b) {
    switch (nb) {
        case 0: return launch_thin_nb<V, MODE, NS, 0>(c, X, G, C, r, small_b);
        case 1: return launch_thin_nb<V, MODE, NS, 1>(c, X, G, C, r, small_b);
        case 2: return launch_thin_nb<V, MODE, NS, 2>(c, X, G, C, r, small_b);
        case 3: return launch_thin_nb<V, MODE, NS, 3>(c, X, G, C, r, small_b);
        default: return launch_thin_nb<V, MODE, NS, 4>(c, X, G, C, r, small_b);
    }
}

// k_thin when it applies (1-4 summed bits, at most 4 bits on the small side, the
// large side's free bits in their own order in C); false otherwise.
template <class V, int MODE>
bool try_thin(Ctx* c, const V* A, const V* B, V* C, const BitMap& m, int mbits, bool small_b, qt_status* st) {
    const int nb = small_b ? m.nbits_c - mbits : mbits;
    const int ns = m.nbits_s;
    if (ns < 1 || ns > 4 || nb > 4) return false;
    // the large side's C bits, in C order, must be its free bits in increasing order
    const int lo = small_b ? 0 : nb, hi = small_b ? mbits : m.nbits_c;
    int prev = -1;
    for (int i = lo; i < hi; ++i) {
        const int src = small_b ? m.c_src[i] : m.c_src[i] - 64;
        if (src <= prev) return false;
        prev = src;
    }
    ThinArgs r;
    std::memset(&r, 0, sizeof r);
    r.n_big = int64_t(1) << (hi - lo);
    r.lo_bits = mbits;
    int pos[4];
    for (int d = 0; d < ns; ++d) pos[d] = small_b ? m.s_a[d] : m.s_b[d];
    std::sort(pos, pos + ns);
    r.n_ins = ns;
    for (int d = 0; d < ns; ++d) r.ins_pos[d] = pos[d];
    for (int j = 0; j < (1 << ns); ++j) {
        int64_t gb = 0, gs = 0;
        for (int d = 0; d < ns; ++d)
            if ((j >> d) & 1) {
                gb += int64_t(1) << (small_b ? m.s_a[d] : m.s_b[d]);
                gs += int64_t(1) << (small_b ? m.s_b[d] : m.s_a[d]);
            }
        r.g_big[j] = gb;
        r.g_small[j] = int32_t(gs);
    }
    for (int o = 0; o < (1 << nb); ++o) {
        int64_t off = 0;
        for (int i = 0; i < nb; ++i)
            if ((o >> i) & 1) {
                const int src = small_b ? m.c_src[mbits + i] - 64 : m.c_src[i];
                off += int64_t(1) << src;
            }
        r.small_off[o] = int32_t(off);
    }
    const V* X = small_b ? A : B;
    const V* G = small_b ? B : A;
    switch (ns) {
        case 1: *st = launch_thin_ns<V, MODE, 1>(c, X, G, C, r, nb, small_b); break;
        case 2: *st = launch_thin_ns<V, MODE, 2>(c, X, G, C, r, nb, small_b); break;
        case 3: *st = launch_thin_ns<V, MODE, 3>(c, X, G, C, r, nb, small_b); break;
        default: *st = launch_thin_ns<V, MODE, 4>(c, X, G, C, r, nb, small_b); break;
    }
    return true;
}

// C = [A part (mbits) | B part]; the side with at most 5 bits is the small one
template <class V, int MODE>
qt_status run_skinny(Ctx* c, const V* A, const V* B, V* C, const BitMap& m, int mbits) {
    const int nbits_n = m.nbits_c - mbits;
    const bool small_b = nbits_n <= 5;
    qt_status st = QT_OK;
    if (try_thin<V, MODE>(c, A, B, C, m, mbits, small_b, &st)) return st;
    switch (small_b ? nbits_n : mbits) {
        case 0: return launch_skinny<V, MODE, 0>(c, A, B, C, m, mbits, small_b);
        case 1: return launch_skinny<V, MODE, 1>(c, A, B, C, m, mbits, small_b);
        case 2: return launch_skinny<V, MODE, 2>(c, A, B, C, m, mbits, small_b);
        case 3: return launch_skinny<V, MODE, 3>(c, A, B, C, m, mbits, small_b);
        case 4: return launch_skinny<V, MODE, 4>(c, A, B, C, m, mbits, small_b);
        default: return launch_skinny<V, MODE, 5>(c, A, B, C, m, mbits, small_b);
    }
}

template <class V, int MODE, int GA, int GB>
qt_status launch_rows(Ctx* c, dim3 grid, const V* A, const V* B, V* part, const BitMap& m, int64_t nC,
                      int step_bits, const RowSet& r) {
    hipLaunchKernelGGL((k_reduce_rows<V, MODE, GA, GB>), grid, dim3(kThreads), 0, c->stream, A, B, part, m, nC,
                       step_bits, r);
    return check_hip(c, hipGetLastError(), "k_reduce_rows");
}

template <class V, int MODE>
qt_status launch_reduce_rows(Ctx* c, int ga, int gb, dim3 grid, const V* A, const V* B, V* part, const BitMap& m,
                             int64_t nC, int step_bits, const RowSet& r) {
    switch (ga * 32 + gb) {
        case 1 * 32 + 1: return launch_rows<V, MODE, 1, 1>(c, grid, A, B, part, m, nC, step_bits, r);
        case 1 * 32 + 2: return launch_rows<V, MODE, 1, 2>(c, grid, A, B, part, m, nC, step_bits, r);
        case 1 * 32 + 4: return launch_rows<V, MODE, 1, 4>(c, grid, A, B, part, m, nC, step_bits, r);
        case 1 * 32 + 8: return launch_rows<V, MODE, 1, 8>(c, grid, A, B, part, m, nC, step_bits, r);
        case 1 * 32 + 16: return launch_rows<V, MODE, 1, 16>(c, grid, A, B, part, m, nC, step_bits, r);
        case 2 * 32 + 1: return launch_rows<V, MODE, 2, 1>(c, grid, A, B, part, m, nC, step_bits, r);
        case 2 * 32 + 2: return launch_rows<V, MODE, 2, 2>(c, grid, A, B, part, m, nC, step_bits, r);
        case 2 * 32 + 4: return launch_rows<V, MODE, 2, 4>(c, grid, A, B, part, m, nC, step_bits, r);
        case 2 * 32 + 8: return launch_rows<V, MODE, 2, 8>(c, grid, A, B, part, m, nC, step_bits, r);
        case 4 * 32 + 1: return launch_rows<V, MODE, 4, 1>(c, grid, A, B, part, m, nC, step_bits, r);
        case 4 * 32 + 2: return launch_rows<V, MODE, 4, 2>(c, grid, A, B, part, m, nC, step_bits, r);
        case 4 * 32 + 4: return launch_rows<V, MODE, 4, 4>(c, grid, A, B, part, m, nC, step_bits, r);
        case 8 * 32 + 1: return launch_rows<V, MODE, 8, 1>(c, grid, A, B, part, m, nC, step_bits, r);
        case 8 * 32 + 2: return launch_rows<V, MODE, 8, 2>(c, grid, A, B, part, m, nC, step_bits, r);
        case 16 * 32 + 1: return launch_rows<V, MODE, 16, 1>(c, grid, A, B, part, m, nC, step_bits, r);
        default: return set_error(c, QT_UNSUPPORTED, "split reduction: unsupported row grouping");
    }
}

template <class V, int MODE>
qt_status run_reduce(Ctx* c, const V* A, const V* B, V* C, const BitMap& m, const Workspace* ws) {
    const int64_t nC = int64_t(1) << m.nbits_c;
    const int64_t nS = int64_t(1) << m.nbits_s;
    // each workgroup takes 16 steps of 256 summed indices, more only beyond 2^15
    // workgroups, and at most 2^12 steps (k_reduce_rows' offset tables cover 12 step
    // bits; nS >= 2^13 is a power of two, so the steps divide it)
    int64_t chunk = kThreads * 16;
    while (nS / chunk > kReduceMaxBlocks && chunk < (int64_t(1) << 20)) chunk <<= 1;
    int step_bits = 0;
    while ((int64_t(kThreads) << step_bits) < chunk) ++step_bits;
    const int64_t nblocks = (nS + chunk - 1) / chunk;
    const size_t need = size_t(nblocks * nC) * sizeof(V);
    V* part = nullptr;
    if (ws) {
        if (need > ws->bytes) return set_error(c, QT_INVALID_INPUT, "split reduction: partial sums exceed the workspace");
        part = reinterpret_cast<V*>(ws->base);
    } else if (c->partials_bytes < need) {
        if (c->partials) {
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(c->partials);
            c->partials = nullptr;
            c->partials_bytes = 0;
        }
        qt_status s = check_hip(c, hipMalloc(&c->partials, need), "partials");
        if (s != QT_OK) return s;
        c->partials_bytes = need;
    }
    if (!part) part = static_cast<V*>(c->partials);
    // rows: A's free bits enumerate na rows, B's nb rows; passes of ga x gb row pairs
    // (at most 16 outputs per pass, so the accumulators stay in registers)
    int64_t row_a[64], row_b[64];
    int16_t c_of[64];
    int na = 0, nb = 0;
    {
        int a_bits[6], b_bits[6], ka = 0, kb = 0;
        for (int i = 0; i < m.nbits_c; ++i) {
            if (m.c_src[i] < 64) a_bits[ka++] = i;
            else b_bits[kb++] = i;
        }
        na = 1 << ka;
        nb = 1 << kb;
        for (int x = 0; x < na; ++x) {
            row_a[x] = 0;
            for (int k = 0; k < ka; ++k)
                if ((x >> k) & 1) row_a[x] += int64_t(1) << m.c_src[a_bits[k]];
        }
        for (int y = 0; y < nb; ++y) {
            row_b[y] = 0;
            for (int k = 0; k < kb; ++k)
                if ((y >> k) & 1) row_b[y] += int64_t(1) << (m.c_src[b_bits[k]] - 64);
        }
        for (int64_t cc = 0; cc < nC; ++cc) {
            int x = 0, y = 0;
            for (int k = 0; k < ka; ++k) x |= int((cc >> a_bits[k]) & 1) << k;
            for (int k = 0; k < kb; ++k) y |= int((cc >> b_bits[k]) & 1) << k;
            c_of[x * nb + y] = int16_t(cc);
        }
    }
    int ga = na < 16 ? na : 16, gb = nb < 16 ? nb : 16;
    while (ga * gb > 16) {
        if (ga >= gb) ga >>= 1;
        else gb >>= 1;
    }
    qt_status s = QT_OK;
    for (int x0 = 0; x0 < na && s == QT_OK; x0 += ga)
        for (int y0 = 0; y0 < nb && s == QT_OK; y0 += gb) {
            RowSet r;
            std::memset(&r, 0, sizeof r);
            for (int i = 0; i < ga; ++i) r.oa[i] = row_a[x0 + i];
            for (int j = 0; j < gb; ++j) r.ob[j] = row_b[y0 + j];
            for (int i = 0; i < ga; ++i)
                for (int j = 0; j < gb; ++j) r.c[i * gb + j] = c_of[(x0 + i) * nb + y0 + j];
            s = launch_reduce_rows<V, MODE>(c, ga, gb, dim3(unsigned(nblocks)), A, B, part, m, nC, step_bits, r);
        }
    if (s != QT_OK) return s;

    hipLaunchKernelGGL((k_reduce_final<V>), dim3(unsigned(nC)), dim3(kThreads), 0, c->stream, part, C, nC, nblocks);
    return check_hip(c, hipGetLastError(), "k_reduce_final");
}

template <class V>
qt_status run_permute(Ctx* c, const V* in, V* out, const PermPlan& p) {
    const int64_t blocks = int64_t(1) << p.n_rest;
    const unsigned gx = unsigned(blocks < 65536 ? blocks : 65536);
    const unsigned gy = unsigned((blocks + gx - 1) / gx);
    const size_t lds = (size_t(1) << (p.tl + p.n_extra_in)) * sizeof(V);
    // element pairs need input and output bit 0 inside the tile (tl >= 1) and 16-byte
    // aligned buffers
    const bool pairs = p.tl >= 1 && (reinterpret_cast<uintptr_t>(in) & 15) == 0 &&
                       (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    if constexpr (sizeof(V) == 8) {
        if (pairs) {
            hipLaunchKernelGGL((k_permute<V, 1>), dim3(gx, gy), dim3(kThreads), lds, c->stream, in, out, p);
            return check_hip(c, hipGetLastError(), "k_permute");
        }
    }
    hipLaunchKernelGGL((k_permute<V, 0>), dim3(gx, gy), dim3(kThreads), lds, c->stream, in, out, p);
    return check_hip(c, hipGetLastError(), "k_permute");
}

bool is_identity(const int* src, int n) {
    for (int i = 0; i < n; ++i)
        if (src[i] != i) return false;
    return true;
}

}  // namespace

PermPlan make_perm_plan(int nbits, const int* src, int tile_bits_max) {
    PermPlan p;
    std::memset(&p, 0, sizeof p);
    p.nbits = nbits;
    // tile of 2^tbits elements, tbits = tl + (output-low bits that come from outside the
    // input-low ones): the largest tl whose tile has at most 2^tile_bits_max elements
    // and leaves at least 2^10 tiles (tbits never falls with tl).  Measured on 2^22- to
    // 2^28-element operands: the largest tiles (2^12 elements) move the most bytes per
    // second (2^28 complex64 elements, low and high 14 bits exchanged: 1.14 ms with
    // 2^12-element tiles, 2.12 ms with 2^10), and tiles under 2^8 leave most of a
    // workgroup's lanes idle (a permutation that keeps its low 6 bits in place ran at
    // 0.3 TB/s with tiles of 2^6)
    auto tbits_for = [&](int t) {
        int n = t;
        for (int j = 0; j < t; ++j) n += src[j] >= t ? 1 : 0;
        return n;
    };
    int cap = nbits - 10 < tile_bits_max ? nbits - 10 : tile_bits_max;
    if (cap < 6) cap = 6;
    int tl = 1;
    while (tl < nbits && tbits_for(tl + 1) <= cap) ++tl;
    if (tl > nbits) tl = nbits;
    p.tl = tl;
    int inv[64];
    for (int j = 0; j < nbits; ++j) inv[src[j]] = j;
    int tile_of_input[64];
    for (int i = 0; i < 64; ++i) tile_of_input[i] = -1;
    for (int i = 0; i < tl; ++i) tile_of_input[i] = i;
    for (int j = 0; j < tl; ++j) {
        const int s = src[j];
        if (s >= tl) {
            p.extra_in_pos[p.n_extra_in] = uint8_t(s);
            p.extra_in_out[p.n_extra_in] = uint8_t(j);
            tile_of_input[s] = tl + p.n_extra_in;
            ++p.n_extra_in;
        }
    }
    for (int i = 0; i < tl; ++i) p.lowin_out[i] = uint8_t(inv[i]);
    for (int j = 0; j < tl; ++j) p.lowout_tile[j] = uint8_t(tile_of_input[src[j]]);
    for (int i = 0; i < tl; ++i)
        if (inv[i] >= tl) {
            p.extra_out_pos[p.n_extra_out] = uint8_t(inv[i]);
            p.extra_out_tile[p.n_extra_out] = uint8_t(i);
            ++p.n_extra_out;
        }
    for (int i = 0; i < nbits; ++i)
        if (tile_of_input[i] < 0) {
            p.rest_in[p.n_rest] = uint8_t(i);
            p.rest_out[p.n_rest] = uint8_t(inv[i]);
            ++p.n_rest;
        }
    return p;
}

qt_status permute(Ctx* c, const void* in, void* out, int dtype, const PermPlan& p) {
    if (dtype == QT_DTYPE_C128)
        return run_permute<double2>(c, static_cast<const double2*>(in), static_cast<double2*>(out), p);
    return run_permute<float2>(c, static_cast<const float2*>(in), static_cast<float2*>(out), p);
}

namespace {

// Large complex64 GEMMs in the 4M / 3M forms: 128x128 output tile per workgroup,
// four waves of 64x64 (4x4 blocks of v_mfma_f32_16x16x4f32), k-tiles of 16; the next
// k-tile is loaded into registers while the current one is multiplied from LDS.
// LDS rows are 144 floats apart, so the four k-rows a wave reads at once sit in
// different bank quarters.
constexpr int BT = 128, BK = 16, BROW = BT + 16;
constexpr int64_t kSplitRows = 256;

// C = the sum of S partial products part[z] (z in order), one complex128 per lane
__global__ __launch_bounds__(kThreads) void k_sum_splits_c128(const double2* __restrict__ part,
                                                              double2* __restrict__ C, int64_t n, int S) {
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n) return;
    double2 a = part[i];
    for (int z = 1; z < S; ++z) {
        const double2 b = part[int64_t(z) * n + i];
        a.x += b.x;
        a.y += b.y;
    }
    C[i] = a;
}

// C = the sum of S partial products part[z] (z in order), two complex64 per lane
__global__ __launch_bounds__(kThreads) void k_sum_splits(const float4* __restrict__ part, float4* __restrict__ C,
                                                         int64_t n4, int S) {
    const int64_t i = int64_t(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= n4) return;
    float4 a = part[i];
    for (int z = 1; z < S; ++z) {
        const float4 b = part[int64_t(z) * n4 + i];
        a.x += b.x;
        a.y += b.y;
        a.z += b.z;
        a.w += b.w;
    }
    C[i] = a;
}

template <int MODE>
__global__ __launch_bounds__(kThreads) void k_gemm_c64_big(const float2* __restrict__ Ap,
                                                           const float2* __restrict__ Bp, float2* __restrict__ C,
                                                           int64_t M, int64_t N, int64_t K) {
    __shared__ float sa[2][BK * BROW], sb[2][BK * BROW];  // [re|im][k][row]
    // split K: workgroups with blockIdx.z = z take summed rows [z K, (z + 1) K) and write
    // their own M x N partial product at C + z M N
    Ap += int64_t(blockIdx.z) * K * M;
    Bp += int64_t(blockIdx.z) * K * N;
    C += int64_t(blockIdx.z) * M * N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int li = lane & 15, lk = lane >> 4;
    const int64_t tiles_m = M / BT;
    const int64_t tile = int64_t(blockIdx.x) + int64_t(blockIdx.y) * gridDim.x;
    if (tile >= tiles_m * (N / BT)) return;
    const int64_t m0 = (tile % tiles_m) * BT, n0 = (tile / tiles_m) * BT;
    constexpr int NA = MODE == MODE_3M ? 3 : 2;
    f4_t acc[NA][4][4];
#pragma unroll
    for (int q = 0; q < NA; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[q][i][j] = f4_t{0, 0, 0, 0};
    // each thread moves 8 elements of A' and 8 of B' per k-tile: element e = tid + 256 q
    float2 ra[8], rb[8];
    auto fetch = [&](int64_t k0) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + kThreads * q, kk = e / BT, rr = e % BT;
            ra[q] = Ap[(k0 + kk) * M + m0 + rr];
            rb[q] = Bp[(k0 + kk) * N + n0 + rr];
        }
    };
    fetch(0);
    for (int64_t k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int e = tid + kThreads * q, o = (e / BT) * BROW + e % BT;
            sa[0][o] = ra[q].x;
            sa[1][o] = ra[q].y;
            sb[0][o] = rb[q].x;
            sb[1][o] = rb[q].y;
        }
        __syncthreads();
        if (k0 + BK < K) fetch(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int o = (kk + lk) * BROW;
            float ar[4], ai[4], br[4], bi[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                ar[i] = sa[0][o + wm * 64 + i * 16 + li];
                ai[i] = sa[1][o + wm * 64 + i * 16 + li];
                br[i] = sb[0][o + wn * 64 + i * 16 + li];
                bi[i] = sb[1][o + wn * 64 + i * 16 + li];
            }
            // one round of MFMAs over all 16 blocks per product, so that no MFMA waits
            // for the one just before it
            if constexpr (MODE == MODE_3M) {
                float sr[4], si[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    sr[i] = br[i] + bi[i];
                    si[i] = ar[i] + ai[i];
                }
#pragma unroll
                for (int bn = 0; bn < 4; ++bn)
#pragma unroll
                    for (int bm = 0; bm < 4; ++bm)
                        acc[0][bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(br[bn], ar[bm], acc[0][bn][bm], 0, 0, 0);
#pragma unroll
                for (int bn = 0; bn < 4; ++bn)
#pragma unroll
                    for (int bm = 0; bm < 4; ++bm)
                        acc[1][bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(bi[bn], ai[bm], acc[1][bn][bm], 0, 0, 0);
#pragma unroll
                for (int bn = 0; bn < 4; ++bn)
#pragma unroll
                    for (int bm = 0; bm < 4; ++bm)
                        acc[2][bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(sr[bn], si[bm], acc[2][bn][bm], 0, 0, 0);
            } else {
#pragma unroll
                for (int bn = 0; bn < 4; ++bn)
#pragma unroll
                    for (int bm = 0; bm < 4; ++bm)
                        acc[0][bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(br[bn], ar[bm], acc[0][bn][bm], 0, 0, 0);
#pragma unroll
                for (int bn = 0; bn < 4; ++bn)
#pragma unroll
                    for (int bm = 0; bm < 4; ++bm)
                        acc[1][bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(bi[bn], ar[bm], acc[1][bn][bm], 0, 0, 0);
#pragma unroll
                for (int bn = 0; bn < 4; ++bn)
#pragma unroll
                    for (int bm = 0; bm < 4; ++bm)
                        acc[0][bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(-bi[bn], ai[bm], acc[0][bn][bm], 0, 0, 0);
#pragma unroll
                for (int bn = 0; bn < 4; ++bn)
#pragma unroll
                    for (int bm = 0; bm < 4; ++bm)
                        acc[1][bn][bm] = __builtin_amdgcn_mfma_f32_16x16x4f32(br[bn], ai[bm], acc[1][bn][bm], 0, 0, 0);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int bn = 0; bn < 4; ++bn)
#pragma unroll
        for (int bm = 0; bm < 4; ++bm)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t n = n0 + wn * 64 + bn * 16 + 4 * lk + r;
                const int64_t mm = m0 + wm * 64 + bm * 16 + li;
                float2 v;
                if constexpr (MODE == MODE_3M) {
                    v.x = acc[0][bn][bm][r] - acc[1][bn][bm][r];
                    v.y = acc[2][bn][bm][r] - acc[0][bn][bm][r] - acc[1][bn][bm][r];
                } else {
                    v.x = acc[0][bn][bm][r];
                    v.y = acc[1][bn][bm][r];
                }
                C[mm + M * n] = v;
            }
}

// Largest |re|, |im| of a complex64 tensor in two launches and no atomics: each
// workgroup of k_absmax_c64 writes the maximum of its grid-stride share to parts[];
// every workgroup of k_rescale_c64 reduces the (at most kAbsParts) partial maxima
// itself, then scales its share by the power of two that brings the maximum into
// [0.5, 1).
constexpr int kAbsParts = 1024;

__device__ __forceinline__ float block_max(float m, float* lds) {
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_down(m, d, 64));
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
    __syncthreads();
    m = lds[0];
    for (int w = 1; w < kThreads / 64; ++w) m = fmaxf(m, lds[w]);
    return m;
}

__global__ __launch_bounds__(kThreads) void k_absmax_c64(const float2* __restrict__ x, int64_t n,
                                                         float* __restrict__ parts) {
    __shared__ float lds[kThreads / 64];
    float m = 0.0f;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        m = fmaxf(m, fmaxf(fabsf(x[i].x), fabsf(x[i].y)));
    m = block_max(m, lds);
    if (threadIdx.x == 0) parts[blockIdx.x] = m;
}

__global__ __launch_bounds__(kThreads) void k_rescale_c64(float2* __restrict__ x, int64_t n,
                                                          const float* __restrict__ parts, int n_parts,
                                                          const int32_t* __restrict__ ea,
                                                          const int32_t* __restrict__ eb, int32_t* __restrict__ ec) {
    __shared__ float lds[kThreads / 64];
    float m = 0.0f;
    for (int j = threadIdx.x; j < n_parts; j += kThreads) m = fmaxf(m, parts[j]);
    m = block_max(m, lds);
    int e = 0;
    if (m > 0.0f && isfinite(m)) (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
    const float s = ldexpf(1.0f, -e);                 // exact power of two
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
        x[i] = float2{x[i].x * s, x[i].y * s};
    if (blockIdx.x == 0 && threadIdx.x == 0) *ec = (ea ? *ea : 0) + (eb ? *eb : 0) + e;
}

// tensors of up to kRescaleSmall elements: the maximum, the scale and the exponent in
// one single-workgroup launch
constexpr int64_t kRescaleSmall = 1 << 14;
__global__ __launch_bounds__(kThreads) void k_rescale_small_c64(float2* __restrict__ x, int64_t n,
                                                                const int32_t* __restrict__ ea,
                                                                const int32_t* __restrict__ eb,
                                                                int32_t* __restrict__ ec) {
    __shared__ float part[kThreads / 64];
    float m = 0.0f;
    for (int64_t i = threadIdx.x; i < n; i += kThreads) m = fmaxf(m, fmaxf(fabsf(x[i].x), fabsf(x[i].y)));
    for (int d = 32; d >= 1; d >>= 1) m = fmaxf(m, __shfl_down(m, d, 64));
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    m = part[0];
    for (int w = 1; w < kThreads / 64; ++w) m = fmaxf(m, part[w]);
    int e = 0;
    if (m > 0.0f && isfinite(m)) (void)frexpf(m, &e);
    const float s = ldexpf(1.0f, -e);
    for (int64_t i = threadIdx.x; i < n; i += kThreads) x[i] = float2{x[i].x * s, x[i].y * s};
    if (threadIdx.x == 0) *ec = (ea ? *ea : 0) + (eb ? *eb : 0) + e;
}


template <class V>
__global__ void k_accumulate(const V* __restrict__ t, const int32_t* __restrict__ e, double2* __restrict__ acc) {
    const int x = e ? *e : 0;
    const double2 a = *acc;
    *acc = double2{a.x + ldexp(double(t->x), x), a.y + ldexp(double(t->y), x)};
}

}  // namespace

qt_status accumulate_scalar(Ctx* c, const Tensor* t, Tensor* acc) {
    const int32_t* e = t->exp_slot >= 0 ? c->d_exp + t->exp_slot : nullptr;
    if (t->dtype == QT_DTYPE_C128)
        hipLaunchKernelGGL(k_accumulate<double2>, dim3(1), dim3(1), 0, c->stream,
                           static_cast<const double2*>(t->data), e, static_cast<double2*>(acc->data));
    else
        hipLaunchKernelGGL(k_accumulate<float2>, dim3(1), dim3(1), 0, c->stream, static_cast<const float2*>(t->data),
                           e, static_cast<double2*>(acc->data));
    return check_hip(c, hipGetLastError(), "k_accumulate");
}

namespace {

// The device exponent table (first use) and a slot of it for C.
qt_status acquire_exponent(Ctx* c, Tensor* C) {
    if (!c->d_exp) {
        c->exp_cap = 1 << 20;
        qt_status s = check_hip(c, hipMalloc(reinterpret_cast<void**>(&c->d_exp), size_t(c->exp_cap) * 4), "exponents");
        if (s == QT_OK)
            s = check_hip(c, hipMalloc(reinterpret_cast<void**>(&c->d_absmax), kAbsParts * sizeof(float)), "absmax");
        if (s != QT_OK) return s;
        for (int32_t i = c->exp_cap - 1; i >= 0; --i) c->exp_free.push_back(i);
    }
    if (C->exp_slot < 0) {
        if (c->exp_free.empty()) return set_error(c, QT_OOM, "rescale: exponent slots exhausted");
        C->exp_slot = c->exp_free.back();
        c->exp_free.pop_back();
    }
    return QT_OK;
}

__global__ void k_carry_exponent(const int32_t* __restrict__ ea, const int32_t* __restrict__ eb,
                                 int32_t* __restrict__ ec) {
    *ec = (ea ? *ea : 0) + (eb ? *eb : 0);
}

}  // namespace

qt_status carry_exponent(Ctx* c, const Tensor* A, const Tensor* B, Tensor* C) {
    if (A->exp_slot < 0 && B->exp_slot < 0) {
        release_exponent(c, C);  // C holds plain values
        return QT_OK;
    }
    qt_status s = acquire_exponent(c, C);
    if (s != QT_OK) return s;
    const int32_t* ea = A->exp_slot >= 0 ? c->d_exp + A->exp_slot : nullptr;
    const int32_t* eb = B->exp_slot >= 0 ? c->d_exp + B->exp_slot : nullptr;
    hipLaunchKernelGGL(k_carry_exponent, dim3(1), dim3(1), 0, c->stream, ea, eb, c->d_exp + C->exp_slot);
    return check_hip(c, hipGetLastError(), "k_carry_exponent");
}

qt_status rescale_c64(Ctx* c, const Tensor* A, const Tensor* B, Tensor* C) {
    qt_status st = acquire_exponent(c, C);
    if (st != QT_OK) return st;
    const int64_t n = C->numel;
    float2* x = static_cast<float2*>(C->data);
    const int32_t* ea = A->exp_slot >= 0 ? c->d_exp + A->exp_slot : nullptr;
    const int32_t* eb = B->exp_slot >= 0 ? c->d_exp + B->exp_slot : nullptr;
    if (n <= kRescaleSmall) {
        hipLaunchKernelGGL(k_rescale_small_c64, dim3(1), dim3(kThreads), 0, c->stream, x, n, ea, eb,
                           c->d_exp + C->exp_slot);
        return check_hip(c, hipGetLastError(), "k_rescale_small_c64");
    }
    const unsigned parts = grid_for(n, kThreads, kAbsParts);
    hipLaunchKernelGGL(k_absmax_c64, dim3(parts), dim3(kThreads), 0, c->stream, x, n, c->d_absmax);
    hipLaunchKernelGGL(k_rescale_c64, dim3(grid_for(n, kThreads, 4096)), dim3(kThreads), 0, c->stream, x, n,
                       c->d_absmax, int(parts), ea, eb, c->d_exp + C->exp_slot);
    return check_hip(c, hipGetLastError(), "k_rescale_c64");
}

qt_status read_exponent(Ctx* c, const Tensor* t, int32_t* e) {
    *e = 0;
    if (t->exp_slot < 0) return QT_OK;
    qt_status s = check_hip(c, hipStreamSynchronize(c->stream), "sync");
    if (s != QT_OK) return s;
    return check_hip(c, hipMemcpy(e, c->d_exp + t->exp_slot, 4, hipMemcpyDeviceToHost), "exponent");
}

void release_exponent(Ctx* c, Tensor* t) {
    if (t->exp_slot >= 0) c->exp_free.push_back(t->exp_slot);
    t->exp_slot = -1;
}

qt_status launch_fill_zero(Ctx* c, void* p, size_t bytes) {
    const int64_t n16 = int64_t((bytes + 15) / 16);  // pool blocks are >= 256 B, 16 B aligned
    hipLaunchKernelGGL(k_fill_zero, dim3(grid_for(n16, 256, 8192)), dim3(256), 0, c->stream,
                       static_cast<uint4*>(p), n16);
    return check_hip(c, hipGetLastError(), "k_fill_zero");
}

namespace {

// Large steps: permute into GEMM layout (skipped when already there), MFMA GEMM.
template <class V, int MODE>
qt_status run_gemm(Ctx* c, const Tensor* A, const Tensor* B, Tensor* C, const BitMap& m, int mbits,
                   const Workspace* ws) {
    // operand permutes: tiles of up to 2^12 elements (64 KB of LDS in complex128)
    constexpr int kPermTileBits = 12;
    const int nbits_n = m.nbits_c - mbits;
    const int sb = m.nbits_s;
    int src_a[64], src_b[64];
    for (int j = 0; j < mbits; ++j) src_a[j] = m.c_src[j];
    for (int j = 0; j < sb; ++j) src_a[mbits + j] = m.s_a[j];
    for (int j = 0; j < nbits_n; ++j) src_b[j] = m.c_src[mbits + j] - 64;
    for (int j = 0; j < sb; ++j) src_b[nbits_n + j] = m.s_b[j];
    const V* ap = static_cast<const V*>(A->data);
    const V* bp = static_cast<const V*>(B->data);
    void* abuf = nullptr;
    void* bbuf = nullptr;
    const bool perm_a = !is_identity(src_a, mbits + sb), perm_b = !is_identity(src_b, nbits_n + sb);
    // a workspace holds A's permute at its start and B's after it (256-byte aligned)
    const size_t ws_b = perm_a ? (A->bytes + 255) / 256 * 256 : 0;
    if (ws && ws_b + (perm_b ? B->bytes : 0) > ws->bytes)
        return set_error(c, QT_INVALID_INPUT, "gemm: operand permutes exceed the workspace");
    qt_status s = QT_OK;
    if (perm_a) {
        hipError_t e = hipSuccess;
        abuf = ws ? static_cast<void*>(ws->base) : c->pool.alloc(A->bytes, &e);
        if (!abuf) return check_hip(c, e, "gemm scratch A");
        s = permute(c, A->data, abuf, A->dtype, make_perm_plan(mbits + sb, src_a, kPermTileBits));
        ap = static_cast<const V*>(abuf);
    }
    if (s == QT_OK && perm_b) {
        hipError_t e = hipSuccess;
        bbuf = ws ? static_cast<void*>(ws->base + ws_b) : c->pool.alloc(B->bytes, &e);
        if (!bbuf) {
            if (!ws) c->pool.release(abuf);
            return check_hip(c, e, "gemm scratch B");
        }
        s = permute(c, B->data, bbuf, B->dtype, make_perm_plan(nbits_n + sb, src_b, kPermTileBits));
        bp = static_cast<const V*>(bbuf);
    }
    if (s == QT_OK) {
        const int64_t M = int64_t(1) << mbits, N = int64_t(1) << nbits_n, K = int64_t(1) << sb;
        const int64_t tiles = (M / GM) * (N / GN);
        const unsigned gx = unsigned(tiles < 65536 ? tiles : 65536);
        const unsigned gy = unsigned((tiles + gx - 1) / gx);
        if constexpr (MODE == MODE_PLAIN) {
            // as for complex64 below: fewer than 512 tiles (64 x 64 here) split K
            int S = 1;
            if (!ws && K % GK == 0)
                while (tiles * S < 512 && K / (2 * S) >= kSplitRows && S < 64) S *= 2;
            double2* part = nullptr;
            if (S > 1) {
                hipError_t e = hipSuccess;
                part = static_cast<double2*>(c->pool.alloc(size_t(S) * size_t(M * N) * sizeof(double2), &e));
                if (!part) {
                    (void)hipGetLastError();  // no room: one product, unsplit
                    S = 1;
                }
            }
            double2* out = S > 1 ? part : static_cast<double2*>(C->data);
            hipLaunchKernelGGL(k_gemm_c128, dim3(gx, gy, unsigned(S)), dim3(kThreads), 0, c->stream, ap, bp, out, M,
                               N, K / S);
            s = check_hip(c, hipGetLastError(), "k_gemm_c128");
            if (S > 1) {
                const int64_t n = M * N;
                if (s == QT_OK) {
                    hipLaunchKernelGGL(k_sum_splits_c128, dim3(unsigned((n + kThreads - 1) / kThreads)), dim3(kThreads),
                                       0, c->stream, part, static_cast<double2*>(C->data), n, S);
                    s = check_hip(c, hipGetLastError(), "k_sum_splits_c128");
                }
                c->pool.release(part);
            }
        } else {
            bool done = false;
            if constexpr (MODE == MODE_4M || MODE == MODE_3M) {
                if (M >= BT && N >= BT && K % BK == 0) {
                    const int64_t big = (M / BT) * (N / BT);
                    const unsigned bx = unsigned(big < 65536 ? big : 65536);
                    const unsigned by = unsigned((big + bx - 1) / bx);
                    // fewer output tiles than two per CU: split K (at least kSplitRows
                    // summed rows per split) into S partial products, then add them in
                    // order; not with a caller's workspace, which holds no room for them
                    int S = 1;
                    if (!ws)
                        while (big * S < 512 && K / (2 * S) >= kSplitRows && S < 64) S *= 2;
                    float2* part = nullptr;
                    if (S > 1) {
                        hipError_t e = hipSuccess;
                        part = static_cast<float2*>(c->pool.alloc(size_t(S) * size_t(M * N) * sizeof(float2), &e));
                        if (!part) {
                            (void)hipGetLastError();  // no room: one product, unsplit
                            S = 1;
                        }
                    }
                    float2* out = S > 1 ? part : static_cast<float2*>(C->data);
                    hipLaunchKernelGGL((k_gemm_c64_big<MODE>), dim3(bx, by, unsigned(S)), dim3(kThreads), 0, c->stream,
                                       ap, bp, out, M, N, K / S);
                    s = check_hip(c, hipGetLastError(), "k_gemm_c64_big");
                    if (S > 1) {
                        const int64_t n4 = M * N / 2;
                        if (s == QT_OK) {
                            hipLaunchKernelGGL(k_sum_splits, dim3(unsigned((n4 + kThreads - 1) / kThreads)),
                                               dim3(kThreads), 0, c->stream, reinterpret_cast<const float4*>(part),
                                               reinterpret_cast<float4*>(C->data), n4, S);
                            s = check_hip(c, hipGetLastError(), "k_sum_splits");
                        }
                        c->pool.release(part);
                    }
                    done = true;
                }
            }
            if (!done) {
                // fewer than 512 tiles (64 x 64 here): split K as above
                int S = 1;
                if (!ws && K % HK == 0)
                    while (tiles * S < 512 && K / (2 * S) >= kSplitRows && S < 64) S *= 2;
                float2* part = nullptr;
                if (S > 1) {
                    hipError_t e = hipSuccess;
                    part = static_cast<float2*>(c->pool.alloc(size_t(S) * size_t(M * N) * sizeof(float2), &e));
                    if (!part) {
                        (void)hipGetLastError();  // no room: one product, unsplit
                        S = 1;
                    }
                }
                float2* out = S > 1 ? part : static_cast<float2*>(C->data);
                hipLaunchKernelGGL((k_gemm_c64<MODE>), dim3(gx, gy, unsigned(S)), dim3(kThreads), 0, c->stream, ap, bp,
                                   out, M, N, K / S);
                s = check_hip(c, hipGetLastError(), "k_gemm_c64");
                if (S > 1) {
                    const int64_t n4 = M * N / 2;
                    if (s == QT_OK) {
                        hipLaunchKernelGGL(k_sum_splits, dim3(unsigned((n4 + kThreads - 1) / kThreads)), dim3(kThreads),
                                           0, c->stream, reinterpret_cast<const float4*>(part),
                                           reinterpret_cast<float4*>(C->data), n4, S);
                        s = check_hip(c, hipGetLastError(), "k_sum_splits");
                    }
                    c->pool.release(part);
                }
            }
        }
    }
    // stream-ordered reuse: later work on this stream may take these blocks
    if (!ws) {
        c->pool.release(abuf);
        c->pool.release(bbuf);
    }
    return s;
}

// a split C (A's free bits low, B's high) with at most 5 bits on one side and at
// least 10 on the other
inline bool skinny_shape(const BitMap& m, bool split, int mbits) {
    const int nbits_n = m.nbits_c - mbits;
    return split && std::min(mbits, nbits_n) <= 5 && std::max(mbits, nbits_n) >= 10;
}

template <class V, int MODE>
qt_status dispatch_small(Ctx* c, const Tensor* A, const Tensor* B, Tensor* C, const BitMap& m, bool split,
                         int mbits, const Workspace* ws) {
    const V* a = static_cast<const V*>(A->data);
    const V* b = static_cast<const V*>(B->data);
    V* cc = static_cast<V*>(C->data);
    const int64_t nC = int64_t(1) << m.nbits_c;
    const int64_t nS = int64_t(1) << m.nbits_s;
    if (nC <= kReduceMaxOut && nS >= 8192) return run_reduce<V, MODE>(c, a, b, cc, m, ws);
    if (skinny_shape(m, split, mbits)) return run_skinny<V, MODE>(c, a, b, cc, m, mbits);
    return run_direct<V, MODE>(c, a, b, cc, m);
}

}  // namespace

qt_status contract_pair(Ctx* c, const Tensor* A, const Tensor* B, Tensor* C, const BitMap& m,
                        int precision, const Workspace* ws) {
    // GEMM shape: C's low bits all from A, high bits all from B
    int mbits = 0;
    while (mbits < m.nbits_c && m.c_src[mbits] < 64) ++mbits;
    bool split = true;
    for (int i = mbits; i < m.nbits_c; ++i)
        if (m.c_src[i] < 64) split = false;
    const int nbits_n = m.nbits_c - mbits;
    const bool gemm = split && mbits >= 6 && nbits_n >= 6;
    // short summed extents (8, or 16 in complex64: below a k-tile) take the GEMM, with
    // the last k-tile zero-filled, when the output is large: k_direct would re-read
    // the operands once per output element of the other side (measured: from 2^16
    // outputs in complex64, from 2^24 in complex128)
    if (A->dtype == QT_DTYPE_C128) {
        const bool short_k = m.nbits_s >= 3 && m.nbits_c >= 24;
        if (gemm && (m.nbits_s >= 4 || short_k)) return run_gemm<double2, MODE_PLAIN>(c, A, B, C, m, mbits, ws);
        return dispatch_small<double2, MODE_PLAIN>(c, A, B, C, m, split, mbits, ws);
    }
    const bool short_k = m.nbits_s >= 3 && m.nbits_c >= 16;
    const bool gemm64 = gemm && (m.nbits_s >= 5 || short_k);  // k-tiles of 32
    switch (precision) {
        case QT_C64_4M:
            return gemm64 ? run_gemm<float2, MODE_4M>(c, A, B, C, m, mbits, ws)
                          : dispatch_small<float2, MODE_4M>(c, A, B, C, m, split, mbits, ws);
        case QT_C64_3M:
            return gemm64 ? run_gemm<float2, MODE_3M>(c, A, B, C, m, mbits, ws)
                          : dispatch_small<float2, MODE_3M>(c, A, B, C, m, split, mbits, ws);
        default:
            return gemm64 ? run_gemm<float2, MODE_BF16X3>(c, A, B, C, m, mbits, ws)
                          : dispatch_small<float2, MODE_BF16X3>(c, A, B, C, m, split, mbits, ws);
    }
}

}  // namespace qt
