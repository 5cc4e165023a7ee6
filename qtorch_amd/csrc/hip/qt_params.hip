// Rotation-gate parameter blocks written on the device from angles
// (qt_batch_rotation_params).
//
// An optimiser's outer loop changes only the angles of Rx / Ry / Rz gates between
// evaluations.  The reference rebuilds every light-cone network from text for each
// evaluation: the angle is printed at the default ostream precision (6 significant
// digits; src/maxcut.h:211,217), parsed back with std::stof (src/Network.h:342), and
// the gate's superoperator is built from that float in FP64 (src/Node.h:197-898).
// Here one thread per (angle set, gate) does the same arithmetic on the device and
// writes the gate's block into the batch's leaf pool, so new angles cost a few
// microseconds instead of a host rebuild and a pool upload.
//
// The text round trip is exact: angle_text_float() gives the float that strtof
// returns for the %g text of x, bit for bit, without producing text (exact products
// and remainders with FMA; argued inline).  The trigonometric values come from the
// device's cos / sin in FP64, so a block element may differ from the host's glibc
// value in the last bit or two.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "qt_batch_impl.h"

namespace qt {

namespace {

// 10^k for 0 <= k <= 22 by squaring: every power of ten up to 10^22 is a double, so
// each product that reaches the result is exact (10^32, formed last for k >= 16, is
// never used).
__host__ __device__ inline double pow10_exact(int k) {
    double p = 1.0, b = 10.0;
    for (; k > 0; k >>= 1, b *= b)
        if (k & 1) p *= b;
    return p;
}

// Sign of (hi + lo) - m, where hi + lo is an exact double-double (|lo| <= ulp(hi)/2).
// hi - m is exact when m/2 <= hi <= 2m (Sterbenz); otherwise |hi - m| dwarfs |lo| and
// its rounding keeps the sign.  Rounding of d + lo never changes a sign.
__host__ __device__ inline int sign_dd_minus(double hi, double lo, double m) {
    const double r = (hi - m) + lo;
    return (r > 0) - (r < 0);
}

__host__ __device__ inline float next_float(float f, bool up) {  // f > 0, finite
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    u = up ? u + 1 : u - 1;
    float g;
    __builtin_memcpy(&g, &u, 4);
    return g;
}

}  // namespace

// The float std::stof returns for the text that `std::ostream << x` writes at the
// default precision, i.e. for x rounded to 6 significant decimal digits (%g: ties
// to even on the exact binary value), then rounded to the nearest float (ties to
// even).  Exact for x == 0 and 1e-16 <= |x| < 1e15; *ok = 0 (and NaN) otherwise.
__host__ __device__ float angle_text_float(double x, int* ok) {
    *ok = 1;
    if (x == 0.0) return __builtin_signbit(x) ? -0.0f : 0.0f;  // "-0" parses to -0.0f
    const double a = fabs(x);
    if (!(a >= 1e-16 && a < 1e15)) {
        *ok = 0;
        return __builtin_nanf("");
    }
    // N = a * 10^(5 - e) rounded to an integer (half to even), e = floor(log10(a)):
    // 10^5 <= N <= 10^6.  With a in [2^(b-1), 2^b), floor((b - 1) log10 2) is e or one
    // less; the exact comparisons below correct it.
    int b2 = 0;
    (void)frexp(a, &b2);
    int e = int(floor(double(b2 - 1) * 0.30102999566398120));
    double n_dig = 0;
    for (int iter = 0; iter < 4; ++iter) {
        const int m = 5 - e;  // in [-10, 22]
        if (m >= 0) {
            const double s = pow10_exact(m);
            const double hi = a * s, lo = fma(a, s, -hi);  // a s = hi + lo exactly
            if (sign_dd_minus(hi, lo, 1e5) < 0) {
                --e;
                continue;
            }
            if (sign_dd_minus(hi, lo, 1e6) >= 0) {
                ++e;
                continue;
            }
            // rint(hi) is the answer unless hi sits on a half: then lo decides (lo == 0
            // is a tie, which rint already broke to even)
            const double n0 = rint(hi), r = hi - n0;  // r exact: hi, n0 in [1e5, 1e6]
            n_dig = n0;
            if (r == 0.5 && lo > 0) n_dig = n0 + 1;
            if (r == -0.5 && lo < 0) n_dig = n0 - 1;
        } else {
            const int j = -m;  // a >= 10^6; 10^(6 + j) <= 10^16 is exact
            const double s = pow10_exact(j);
            if (a < 1e5 * s) {
                --e;
                continue;
            }
            if (a >= pow10_exact(6 + j)) {
                ++e;
                continue;
            }
            // a / s rounded, then compared exactly with the neighbouring half-integers
            // times s (M s = hi + lo exactly; a - hi is exact, a being that close)
            double n0 = rint(a / s);
            const double mu = n0 + 0.5, ml = n0 - 0.5;
            const double hu = mu * s, lu = fma(mu, s, -hu);
            const double hl = ml * s, ll = fma(ml, s, -hl);
            const int cu = sign_dd_minus(hu, lu, a);  // sign of mu s - a
            const int cl = sign_dd_minus(hl, ll, a);
            if (cu < 0 || (cu == 0 && fmod(n0, 2.0) != 0)) n0 += 1;
            else if (cl > 0 || (cl == 0 && fmod(n0, 2.0) != 0)) n0 -= 1;
            n_dig = n0;
        }
        break;
    }
    if (n_dig == 0) {  // log10 was off by more than the loop corrects (not expected)
        *ok = 0;
        return __builtin_nanf("");
    }
    // the decimal value D = n_dig * 10^(e - 5), correctly rounded to float
    const int k = e - 5;
    float f;
    if (k >= 0) {
        f = float(n_dig * pow10_exact(k));  // n_dig 10^k <= 10^15: an exact double
    } else {
        const double p = pow10_exact(-k);  // -k <= 22
        const double q = n_dig / p;
        f = float(q);
        // float(q) is right unless q is exactly halfway between two floats: the exact
        // quotient may lie on either side of that midpoint (or on it)
        if (double(f) != q) {
            const float g = next_float(f, q > double(f));
            const double mid = 0.5 * (double(f) + double(g));  // exact
            if (mid == q) {
                const double hi = mid * p, lo = fma(mid, p, -hi);  // mid p exactly
                const int c = sign_dd_minus(hi, lo, n_dig);        // sign of mid - D (times p)
                if (c < 0) f = fmaxf(f, g);                         // D above the midpoint
                else if (c > 0) f = fminf(f, g);
                // c == 0: a true tie, which float(q) broke to even
            }
        }
    }
    return x < 0 ? -f : f;
}

namespace {

constexpr int kMaxRot = 32;

// A call's gates (kernel argument: no device copy of the description).
struct RotArgs {
    int32_t n_rot, dl, n_angles, n_sets;
    int64_t base, stride;
    int32_t kind[kMaxRot];
    int32_t angle[kMaxRot];
    double scale[kMaxRot];
    int64_t offset[kMaxRot];
};

// The elements a launch writes per angle set, (gate << 4) | element: every element
// the first time a layout is written, afterwards those whose value depends on the
// angle (the others already hold their table value).
constexpr int kMaxWrites = kMaxRot * 16;
struct RotWrites {
    int32_t n;
    uint16_t w[kMaxWrites];
};

// What a launch reads besides the angles, kept in device memory (uploaded when it
// changes): as kernel arguments its 1.9 KB made every launch start about 5 us late.
struct RotDesc {
    RotArgs r;
    RotWrites wr;
};

// Values a gate block takes from its angle t (the parsed float, widened): the
// expressions of the host's gate tables.
struct RotScalars {
    double c2, s2, h, ct, st;  // superoperators: cos^2(t/2), sin^2(t/2), sin t / 2, cos t, sin t
};

// Only what the kind's table reads: Rz superoperators cos t and sin t; Rx and Ry
// cos^2(t/2), sin^2(t/2) and sin t / 2 (as sin(t/2) cos(t/2): one sincos); amplitude
// gates cos(t/2) and sin(t/2).
__host__ __device__ inline RotScalars rot_scalars(int dl, int kind, double t) {
    RotScalars r{0, 0, 0, 0, 0};
    if (dl == 2 && kind == QT_ROT_RZ) {
        sincos(t, &r.st, &r.ct);
        return r;
    }
    double ch, sh;
    sincos(t / 2.0, &sh, &ch);
    if (dl == 2) {
        r.c2 = ch * ch;
        r.s2 = sh * sh;
        r.h = sh * ch;  // sin t / 2 (within an ulp or so of the host's sin(t) / 2.0)
        r.st = 2.0 * r.h;
    } else {
        r.c2 = ch;
        r.s2 = sh;
    }
    return r;
}

// Symbols of the superoperator tables (src/Node.h:561-898), one nibble per element
// e = in + 4 out: the element's value in terms of the angle's scalars.
enum : int { S_Z0 = 0, S_ONE, S_C2, S_S2, S_NS2, S_H, S_NH, S_IH, S_NIH, S_EM, S_EP };

// Row `in` lists the symbols for out = 0..3; packed so element e sits at bits 4e.
constexpr uint64_t pack_rows(const int (&t)[4][4]) {
    uint64_t v = 0;
    for (int in = 0; in < 4; ++in)
        for (int out = 0; out < 4; ++out) v |= uint64_t(t[in][out]) << (4 * (in + 4 * out));
    return v;
}
constexpr int kTabRx[4][4] = {{S_C2, S_IH, S_NIH, S_S2}, {S_IH, S_C2, S_S2, S_NIH},
                              {S_NIH, S_S2, S_C2, S_IH}, {S_S2, S_NIH, S_IH, S_C2}};
constexpr int kTabRy[4][4] = {{S_C2, S_H, S_H, S_S2}, {S_NH, S_C2, S_NS2, S_H},
                              {S_NH, S_NS2, S_C2, S_H}, {S_S2, S_NH, S_NH, S_C2}};
constexpr int kTabRz[4][4] = {{S_ONE, S_Z0, S_Z0, S_Z0}, {S_Z0, S_EM, S_Z0, S_Z0},
                              {S_Z0, S_Z0, S_EP, S_Z0}, {S_Z0, S_Z0, S_Z0, S_ONE}};
constexpr uint64_t kPackRx = pack_rows(kTabRx), kPackRy = pack_rows(kTabRy), kPackRz = pack_rows(kTabRz);

// Element e of a gate block.  Superoperators (dl = 2): the tables above, with the
// host's expressions (NS2 = -s2, NH = -(sin t)/2, NIH = -i (sin t)/2 are exact
// negations).  Amplitude gates (dl = 1): element in + 2 out = U[out][in] with
// Rx = [[c, -is], [-is, c]], Ry = [[c, -s], [s, c]], Rz = diag(c - is, c + is).
__host__ __device__ inline double2 rot_element(int kind, int dl, const RotScalars& v, int e) {
    if (dl == 1) {
        const double c = v.c2, s = v.s2;
        switch (kind) {
            case QT_ROT_RX: return (e == 0 || e == 3) ? double2{c, 0} : double2{0, -s};
            case QT_ROT_RY: return (e == 0 || e == 3) ? double2{c, 0} : (e == 1 ? double2{-s, 0} : double2{s, 0});
            default: return e == 0 ? double2{c, -s} : (e == 3 ? double2{c, s} : double2{0, 0});
        }
    }
    const uint64_t pack = kind == QT_ROT_RX ? kPackRx : kind == QT_ROT_RY ? kPackRy : kPackRz;
    switch (int((pack >> (4 * e)) & 15)) {
        case S_ONE: return double2{1.0, 0};
        case S_C2: return double2{v.c2, 0};
        case S_S2: return double2{v.s2, 0};
        case S_NS2: return double2{-v.s2, 0};
        case S_H: return double2{v.h, 0};
        case S_NH: return double2{-v.h, 0};
        case S_IH: return double2{0, v.h};
        case S_NIH: return double2{0, -v.h};
        case S_EM: return double2{v.ct, -v.st};
        case S_EP: return double2{v.ct, v.st};
        default: return double2{0, 0};
    }
}

// 128 (set, gate) pairs per workgroup of 128 threads: 2,048 workgroups for 131,072
// sets of two gates, one wavefront per gate kind.
constexpr int kPairsPerBlock = 128, kRotThreads = 128;

// Element e of a superoperator block without branches (lanes of one wave hold gates of
// different kinds): the symbol picks the value through selects.
__device__ inline double2 rot_element_select(uint64_t pack, const RotScalars& v, int e) {
    const int sym = int((pack >> (4 * e)) & 15);
    double re = 0, im = 0;
    re = sym == S_ONE ? 1.0 : re;
    re = sym == S_C2 ? v.c2 : re;
    re = sym == S_S2 ? v.s2 : re;
    re = sym == S_NS2 ? -v.s2 : re;
    re = sym == S_H ? v.h : re;
    re = sym == S_NH ? -v.h : re;
    re = (sym == S_EM || sym == S_EP) ? v.ct : re;
    im = sym == S_IH ? v.h : im;
    im = sym == S_NIH ? -v.h : im;
    im = sym == S_EM ? -v.st : im;
    im = sym == S_EP ? v.st : im;
    return double2{re, im};
}

// One thread per (set, gate) computes the angle's scalars into LDS; then the
// workgroup writes the listed elements of its sets' blocks, consecutive lanes on
// consecutive list entries (the blocks of a set's gates, and of consecutive sets, are
// usually adjacent in the pool), each lane selecting its element's value without
// branches.  An angle outside angle_text_float's range sets *status and writes NaN.
template <class V>
__global__ __launch_bounds__(kRotThreads) void k_rotation_params(V* __restrict__ pool,
                                                                    const double* __restrict__ angles,
                                                                    const RotDesc* __restrict__ desc,
                                                                    int* __restrict__ status) {
    const RotArgs& r = desc->r;
    const RotWrites& wr = desc->wr;
    __shared__ RotScalars sc[kPairsPerBlock];
    // the write list decoded once per workgroup: (gate << 4 | element, element offset
    // in a set's region, symbol table or kind); the loop below then reads LDS only
    extern __shared__ int64_t rot_lds[];  // wr.n entries of each (rot_lds_bytes)
    int64_t* w_off = rot_lds;
    uint64_t* w_pack = reinterpret_cast<uint64_t*>(rot_lds + wr.n);
    int32_t* w_rot_e = reinterpret_cast<int32_t*>(rot_lds + 2 * wr.n);
    const int sets_per_block = kPairsPerBlock / r.n_rot;
    const int first_set = int(blockIdx.x) * sets_per_block;
    const int sets_here = r.n_sets - first_set < sets_per_block ? r.n_sets - first_set : sets_per_block;
    const int tid = threadIdx.x;
    for (int k = tid; k < wr.n; k += kRotThreads) {
        const int w = wr.w[k], g = w >> 4, e = w & 15, kind = r.kind[g];
        w_rot_e[k] = w;
        w_off[k] = r.offset[g] + e;
        w_pack[k] = r.dl == 1 ? uint64_t(kind) : kind == QT_ROT_RX ? kPackRx : kind == QT_ROT_RY ? kPackRy : kPackRz;
    }
    // gate-major within the workgroup, so that the lanes of a wavefront usually share
    // a gate kind (and its branch of rot_scalars)
    const int rot = tid / sets_per_block, la = tid % sets_per_block;
    if (rot < r.n_rot && la < sets_here) {
        int ok = 1;
        const double x = r.scale[rot] * angles[int64_t(first_set + la) * r.n_angles + r.angle[rot]];
        const float f = angle_text_float(x, &ok);
        RotScalars v = rot_scalars(r.dl, r.kind[rot], double(f));
        if (!ok) {
            atomicOr(status, 1);
            const double nan = __builtin_nan("");
            v = RotScalars{nan, nan, nan, nan, nan};
        }
        sc[la * r.n_rot + rot] = v;
    }
    __syncthreads();
    // entry idx = set la, list entry k (idx = la n + k), stepped without division
    const int n = wr.n;
    int wl = tid / n, wk = tid - (tid / n) * n;
    const int step_l = kRotThreads / n, step_k = kRotThreads - step_l * n;
    V* const region = pool + r.base + int64_t(first_set) * r.stride;
    for (; wl < sets_here; wl += step_l, wk += step_k) {
        if (wk >= n) {
            wk -= n;
            ++wl;
            if (wl >= sets_here) break;
        }
        const int w = w_rot_e[wk];
        const RotScalars& v = sc[wl * r.n_rot + (w >> 4)];
        const double2 val = r.dl == 2 ? rot_element_select(w_pack[wk], v, w & 15) : rot_element(int(w_pack[wk]), 1, v, w & 15);
        V out;
        out.x = decltype(out.x)(val.x);
        out.y = decltype(out.y)(val.y);
        region[int64_t(wl) * r.stride + w_off[wk]] = out;
    }
}

// Whether the blocks a call writes keep every parameter literal of the generated
// kernel: a literal inside a written block must sit at an element that holds the same
// value for every angle (a zero or a one of the table), with that value.
bool literals_survive(const qt_batch* b, const RotArgs& r) {
    const int nel = 1 << (2 * r.dl);
    const bool c128 = b->precision == QT_PLAIN_C128;
    // angle-independent parts of each gate's elements: equal at two unrelated angles
    struct Fixed {
        bool re, im;
        double2 v;
    };
    std::vector<std::vector<Fixed>> fixed(size_t(r.n_rot));
    for (int g = 0; g < r.n_rot; ++g) {
        const RotScalars v1 = rot_scalars(r.dl, r.kind[g], 0.3), v2 = rot_scalars(r.dl, r.kind[g], 1.1);
        for (int e = 0; e < nel; ++e) {
            double2 a = rot_element(r.kind[g], r.dl, v1, e), c = rot_element(r.kind[g], r.dl, v2, e);
            if (!c128) {
                a = double2{double(float(a.x)), double(float(a.y))};
                c = double2{double(float(c.x)), double(float(c.y))};
            }
            fixed[size_t(g)].push_back(Fixed{a.x == c.x, a.y == c.y, a});
        }
    }
    // gate and element of each offset inside a set's region (blocks do not overlap)
    std::vector<std::pair<int64_t, int>> starts;
    for (int g = 0; g < r.n_rot; ++g) starts.push_back({r.offset[g], g});
    std::sort(starts.begin(), starts.end());
    const int64_t span = int64_t(r.n_sets) * r.stride;
    // one element of one set's region: is a literal there kept?
    auto ok_at = [&](int64_t o, const ParamLit& l) {
        auto it = std::upper_bound(starts.begin(), starts.end(), std::make_pair(o, INT32_MAX));
        if (it == starts.begin()) return true;
        --it;
        const int64_t el = o - it->first;
        if (el >= nel) return true;  // between blocks: not written
        const Fixed& f = fixed[size_t(it->second)][size_t(el)];
        return (!l.has_re || (f.re && f.v.x == l.re)) && (!l.has_im || (f.im && f.v.y == l.im));
    };
    // usual layout: every instance's block starts a set's region, so each program's
    // literals land on the same offsets in every instance
    bool aligned = true;
    std::vector<char> used(b->spec_param_lits.size(), 0);
    for (int i = 0; i < b->n_inst && aligned; ++i) {
        const int64_t d = b->h_pbase[size_t(i)] - r.base;
        aligned = d >= 0 && d % r.stride == 0 && d / r.stride < r.n_sets;
        const size_t p = size_t(b->h_prog[size_t(i)]);
        if (p < used.size()) used[p] = 1;
    }
    if (aligned) {
        for (size_t p = 0; p < used.size(); ++p)
            if (used[p])
                for (const ParamLit& l : b->spec_param_lits[p]) {
                    if (l.off < 0 || l.off >= r.stride) {
                        aligned = false;  // reaches into another set's region: check per instance
                        break;
                    }
                    if (!ok_at(l.off, l)) return false;
                }
        if (aligned) return true;
    }
    for (int i = 0; i < b->n_inst; ++i) {
        const size_t p = size_t(b->h_prog[size_t(i)]);
        if (p >= b->spec_param_lits.size()) continue;
        const int64_t pbase = b->h_pbase[size_t(i)];
        for (const ParamLit& l : b->spec_param_lits[p]) {
            const int64_t q = pbase + l.off - r.base;
            if (q < 0 || q >= span) continue;
            if (!ok_at(q % r.stride, l)) return false;
        }
    }
    return true;
}

}  // namespace

qt_status refresh_pool_mirror(qt_batch* b) {
    if (!b->mirror_stale) return QT_OK;
    Ctx* c = b->ctx;
    qt_status s = check_hip(c, hipStreamSynchronize(c->stream), "sync");
    if (s != QT_OK) return s;
    const int64_t n = b->pool_elems;
    std::vector<double> m(size_t(2 * n), 0.0);
    if (b->precision == QT_PLAIN_C128) {
        s = check_hip(c, hipMemcpy(m.data(), b->d_pool, size_t(n) * 16, hipMemcpyDeviceToHost), "pool read");
    } else {
        std::vector<float> f(size_t(2 * n));
        s = check_hip(c, hipMemcpy(f.data(), b->d_pool, size_t(n) * 8, hipMemcpyDeviceToHost), "pool read");
        for (size_t i = 0; i < f.size(); ++i) m[i] = double(f[i]);
    }
    if (s != QT_OK) return s;
    b->h_pool_mirror.swap(m);
    b->mirror_stale = false;
    return QT_OK;
}

}  // namespace qt

using qt::check_hip;
using qt::set_error;

extern "C" {

qt_status qt_angle_text_float(double x, float* out) {
    if (!out) return QT_INVALID_INPUT;
    int ok = 0;
    *out = qt::angle_text_float(x, &ok);
    return ok ? QT_OK : QT_INVALID_INPUT;
}

qt_status qt_batch_rotation_params(qt_batch* b, int dim_log2, int n_rot, const qt_rotation* rots, int n_sets,
                                   int n_angles, const double* angles, int angles_on_host, int64_t base,
                                   int64_t set_stride) {
    if (!b) return QT_INVALID_INPUT;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if ((dim_log2 != 1 && dim_log2 != 2) || n_rot < 1 || n_rot > qt::kMaxRot || !rots || n_sets < 1 ||
        n_angles < 1 || !angles || base < 0 || set_stride < 1)
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_rotation_params: bad arguments");
    if (int64_t(n_sets) * n_rot >= (int64_t(1) << 31))
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_rotation_params: more than 2^31 gate blocks");
    const int nel = 1 << (2 * dim_log2);
    qt::RotArgs r;
    std::memset(&r, 0, sizeof r);
    r.n_rot = n_rot;
    r.dl = dim_log2;
    r.n_angles = n_angles;
    r.n_sets = n_sets;
    r.base = base;
    r.stride = set_stride;
    std::vector<std::pair<int64_t, int64_t>> blocks;
    for (int i = 0; i < n_rot; ++i) {
        const qt_rotation& q = rots[i];
        if (q.kind < QT_ROT_RX || q.kind > QT_ROT_RZ || q.angle < 0 || q.angle >= n_angles || q.offset < 0 ||
            !std::isfinite(q.scale))
            return set_error(ctx, QT_INVALID_INPUT, "qt_batch_rotation_params: bad gate " + std::to_string(i));
        r.kind[i] = q.kind;
        r.angle[i] = q.angle;
        r.scale[i] = q.scale;
        r.offset[i] = q.offset;
        blocks.push_back({q.offset, q.offset + nel});
    }
    std::sort(blocks.begin(), blocks.end());
    for (size_t i = 1; i < blocks.size(); ++i)
        if (blocks[i].first < blocks[i - 1].second)
            return set_error(ctx, QT_INVALID_INPUT, "qt_batch_rotation_params: gate blocks overlap");
    if (blocks.back().second > set_stride)
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_rotation_params: set_stride below a set's blocks");
    const int64_t lo = base + blocks.front().first;
    const int64_t hi = base + (int64_t(n_sets) - 1) * set_stride + blocks.back().second;
    if (hi > b->pool_elems || hi < lo)
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_rotation_params: blocks outside the leaf pool");
    // the generated kernel: its constants, conjugate symmetry (rotation superoperators
    // keep it; amplitude gates have none) and literals
    if (b->specialised) {
        bool keep = lo >= int64_t(b->spec_consts.size() / 2) && !(b->spec_hermitian && dim_log2 != 2);
        if (keep) {
            const char* raw = reinterpret_cast<const char*>(&r);
            const bool same = b->rot_checked_gen == b->layout_gen && b->rot_checked_args.size() == sizeof r &&
                              std::memcmp(b->rot_checked_args.data(), raw, sizeof r) == 0;
            if (!same) {
                keep = qt::literals_survive(b, r);
                if (keep) {
                    b->rot_checked_gen = b->layout_gen;
                    b->rot_checked_args.assign(raw, raw + sizeof r);
                }
            }
        }
        if (!keep) {
            qt::drop_specialised(b);
            qt_status s = qt::after_specialised_drop(b);
            if (s != QT_OK) return s;
        } else {
            b->mirror_stale = true;
        }
    }
    // every element the first time this layout is written into the current pool,
    // afterwards only the angle-dependent ones
    qt::RotWrites wr;
    std::memset(&wr, 0, sizeof wr);
    {
        const char* raw = reinterpret_cast<const char*>(&r);
        const bool full_done = b->rot_full_valid && b->rot_full_args.size() == sizeof r &&
                               std::memcmp(b->rot_full_args.data(), raw, sizeof r) == 0;
        for (int g2 = 0; g2 < n_rot; ++g2) {
            const qt::RotScalars v1 = qt::rot_scalars(dim_log2, r.kind[g2], 0.3);
            const qt::RotScalars v2 = qt::rot_scalars(dim_log2, r.kind[g2], 1.1);
            for (int e = 0; e < nel; ++e) {
                const double2 x1 = qt::rot_element(r.kind[g2], dim_log2, v1, e);
                const double2 x2 = qt::rot_element(r.kind[g2], dim_log2, v2, e);
                if (!full_done || x1.x != x2.x || x1.y != x2.y) wr.w[wr.n++] = uint16_t((g2 << 4) | e);
            }
        }
    }
    qt_status s = QT_OK;
    if (!b->d_param_status) {
        s = check_hip(ctx, qt::batch_malloc(ctx, reinterpret_cast<void**>(&b->d_param_status), 256), "status");
        if (s == QT_OK) s = qt::launch_fill_zero(ctx, b->d_param_status, 256);
        if (s != QT_OK) return s;
    }
    const double* d_ang = angles;
    if (angles_on_host) {
        const size_t n = size_t(n_sets) * size_t(n_angles);
        if (n > b->cap_angles) {
            qt::batch_free(ctx, b->d_angles);
            b->d_angles = nullptr;
            b->cap_angles = 0;
            s = check_hip(ctx, qt::batch_malloc(ctx, reinterpret_cast<void**>(&b->d_angles), n * 8), "angles");
            if (s != QT_OK) return s;
            b->cap_angles = n;
        }
        // the previous call's kernel may still read the buffer: the copy is ordered
        // after it on the stream, and the host array is taken before returning
        s = check_hip(ctx, hipMemcpyAsync(b->d_angles, angles, n * 8, hipMemcpyHostToDevice, ctx->stream), "angles");
        if (s == QT_OK) s = check_hip(ctx, hipStreamSynchronize(ctx->stream), "angles");
        if (s != QT_OK) return s;
        d_ang = b->d_angles;
    }
    if (wr.n == 0) return QT_OK;  // nothing depends on the angles (not expected for rotations)
    const int sets_per_block = qt::kPairsPerBlock / n_rot;
    const size_t lds = size_t(wr.n) * 20;  // k_rotation_params' decoded write list
    const unsigned grid = unsigned((n_sets + sets_per_block - 1) / sets_per_block);
    // the description in device memory, uploaded when it changes (the first calls of a
    // layout; a synchronous copy, so the host copy may change afterwards)
    {
        qt::RotDesc desc;
        std::memset(&desc, 0, sizeof desc);
        desc.r = r;
        desc.wr = wr;
        const char* raw = reinterpret_cast<const char*>(&desc);
        if (!b->d_rot_desc || b->rot_desc_host.size() != sizeof desc ||
            std::memcmp(b->rot_desc_host.data(), raw, sizeof desc) != 0) {
            if (!b->d_rot_desc)
                s = check_hip(ctx, qt::batch_malloc(ctx, &b->d_rot_desc, sizeof desc), "rotation description");
            if (s == QT_OK) s = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
            if (s == QT_OK)
                s = check_hip(ctx, hipMemcpy(b->d_rot_desc, &desc, sizeof desc, hipMemcpyHostToDevice),
                              "rotation description");
            if (s != QT_OK) return s;
            b->rot_desc_host.assign(raw, raw + sizeof desc);
        }
    }
    const auto* dd = static_cast<const qt::RotDesc*>(b->d_rot_desc);
    if (b->precision == QT_PLAIN_C128)
        hipLaunchKernelGGL(qt::k_rotation_params<double2>, dim3(grid), dim3(qt::kRotThreads), lds, ctx->stream,
                           static_cast<double2*>(b->d_pool), d_ang, dd, b->d_param_status);
    else
        hipLaunchKernelGGL(qt::k_rotation_params<float2>, dim3(grid), dim3(qt::kRotThreads), lds, ctx->stream,
                           static_cast<float2*>(b->d_pool), d_ang, dd, b->d_param_status);
    b->in_use = true;  // the kernel writes the pool
    s = check_hip(ctx, hipGetLastError(), "k_rotation_params");
    if (s == QT_OK) {
        b->params_unchecked = true;
        const char* raw = reinterpret_cast<const char*>(&r);
        b->rot_full_args.assign(raw, raw + sizeof r);
        b->rot_full_valid = true;
    }
    return s;
}

}  // extern "C"
