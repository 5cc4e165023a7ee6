// Specialised executors: the programs of a batch compiled to straight-line code.
//
// The table executors (qt_batch.hip) read every step's index maps from packed offset
// tables and gather operands from LDS, so a complex MAC costs table loads, address
// arithmetic and two LDS gathers besides its four FMAs.  When a batch's programs are
// fixed and its instances differ only in their parameter blocks (the edge light
// cones of one graph run for many angle sets), qt_batch_compile writes one HIP
// kernel in which every program is a branch of straight-line code:
//   * each output element of each step is an explicit chain of FMAs over named
//     values, in the order and with the operand roles of the table executors' Acc
//     (so results agree bit for bit, up to the sign of zero);
//   * constant leaves become literals and terms with an exact-zero constant factor
//     are dropped;
//   * parameter leaves are loaded once per instance; intermediates are local values
//     that the compiler keeps in registers (no LDS, no tables);
//   * one lane runs one instance, a wavefront up to 64 consecutive instances of one
//     program (the branch is wave-uniform).
// The kernel is compiled at run time for the device's architecture (hipRTC) and
// loaded as a module; the constant leaves are read from the batch's pool when it is
// compiled (writes over them afterwards drop the compiled kernel).
#include <hip/hiprtc.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <tuple>
#include <unordered_map>
#include <sstream>
#include <string>
#include <vector>

#include "qt_batch_impl.h"

namespace qt {
namespace {

// Straight-line code grows with the programs' work: beyond this many FMAs in total
// (the 45 light cones of a p = 1 QAOA graph need about 90 thousand, compiled in about
// ten seconds) the compile time and code size stop paying off.
constexpr int64_t kMaxSpecialisedFmas = 250000;
// Blocks of 64 groups per tile of the group-sums-only kernel: 64 x 64 angle sets'
// parameter blocks (2 MiB of complex128 p = 1 blocks) stay in an XCD's 4 MiB L2 while
// all 16 slices run over them.
constexpr int kSumsTile = 64;

struct ProgView {
    const ParamLits* param_lits = nullptr;
    const StepDesc* steps = nullptr;
    int n_steps = 0;
    const uint32_t* tabs = nullptr;  // steps[i].tab indexes this array
    const int64_t* refs = nullptr;
    int n_leaves = 0;
    int64_t const_base = 0;
    int dl = 2;  // index bits per axis
};

// One part (real or imaginary) of an operand element: a literal or a variable.
struct Part {
    bool lit = true;
    double v = 0;          // literal value (already rounded for complex64)
    std::string name;      // variable
    bool neg = false;      // the variable's negation (conjugate mirror elements)
};
struct Val {
    Part re, im;
    int step = -1, elem = -1;  // intermediate element it reads (liveness), else -1
};

class Writer {
public:
    explicit Writer(bool c128) : c128_(c128) {}
    std::string lit(double x) const {
        char buf[64];
        if (c128_) std::snprintf(buf, sizeof buf, "%a", x);
        else std::snprintf(buf, sizeof buf, "%af", double(float(x)));
        return std::signbit(x) ? "(" + std::string(buf) + ")" : std::string(buf);
    }
    std::string str(const Part& p) const { return p.lit ? lit(p.v) : (p.neg ? "(-" + p.name + ")" : p.name); }
    std::string neg(const Part& p) const { return p.lit ? lit(-p.v) : (p.neg ? p.name : "(-" + p.name + ")"); }
    const char* fma() const { return c128_ ? "__builtin_fma" : "__builtin_fmaf"; }
    const char* real() const { return c128_ ? "double" : "float"; }

private:
    bool c128_;
};

bool zero(const Part& p) { return p.lit && p.v == 0.0; }

// The four FMAs of Acc<V, 4M / C128>::mac(a, b) in their order: re += a.x b.x;
// re += -a.y b.y; im += a.x b.y; im += a.y b.x.  A term with an exact-zero literal
// factor adds exactly zero and is dropped.
struct Term {
    Val a, b;
    bool f[4];
};

struct Program {
    std::string text;
    int64_t fmas = 0;  // FMAs the generated code executes per instance
};

// Source of one program: a device function prog<id>(const real* P, double* out).
// The conjugate mirror of a flat index with two bits per axis (d = 2 ket + bra): ket
// and bra bits swapped on every axis.
inline int64_t mirror_index(int64_t x) {
    const int64_t lo = 0x5555555555555555LL;
    return ((x & lo) << 1) | ((x >> 1) & lo);
}

// With hermitian set, every leaf satisfies T(mirror(x)) = conj T(x) (superoperators
// U (x) U* of gates, Hermitian measurement vectors), and so does every intermediate:
// of each mirror pair only the lower index is computed (the other is its conjugate)
// and self-mirror elements are real (their imaginary chains are not computed).
bool emit_program(int id, const ProgView& pv, bool c128, bool hermitian, const double* pool, int64_t pool_elems,
                  Program* out, std::string* err) {
    Writer w(c128);
    const int64_t kParam = QT_LEAF_PARAM;
    const std::string tag = "program " + std::to_string(id) + ": ";
    // which step's element each scratch element holds (the table executors' reuse)
    std::map<int32_t, std::pair<int, int>> scratch;
    // per step, per output element: its terms; parts that stay exactly zero
    std::vector<std::vector<std::vector<Term>>> terms(size_t(pv.n_steps));
    std::vector<std::vector<std::pair<bool, bool>>> has(size_t(pv.n_steps));  // (re non-empty, im non-empty)
    std::vector<std::vector<int>> mirror(size_t(pv.n_steps));  // computed element whose conjugate it is, or -1
    if (hermitian && pv.dl != 2) {
        *err = tag + "conjugate symmetry needs two index bits per axis";
        return false;
    }
    std::unordered_map<int64_t, ParamLit> plit;
    if (pv.param_lits)
        for (const ParamLit& l : *pv.param_lits) plit[l.off] = l;
    auto operand = [&](int32_t code, int64_t idx, Val* v) -> bool {
        if (code >= 0) {
            auto it = scratch.find(code + int32_t(idx));
            if (it == scratch.end()) {
                *err = tag + "read of an unwritten scratch element";
                return false;
            }
            const int st = it->second.first;
            int e = it->second.second;
            bool conj = false;
            if (mirror[size_t(st)][size_t(e)] >= 0) {
                e = mirror[size_t(st)][size_t(e)];
                conj = true;
            }
            const auto h = has[size_t(st)][size_t(e)];
            const std::string base = "t" + std::to_string(st) + "_" + std::to_string(e);
            v->re = h.first ? Part{false, 0, base + "r", false} : Part{};
            v->im = h.second ? Part{false, 0, base + "i", conj} : Part{};
            v->step = st;
            v->elem = e;
            return true;
        }
        const int li = -1 - code;
        if (li < 0 || li >= pv.n_leaves) {
            *err = tag + "bad leaf";
            return false;
        }
        const int64_t r = pv.refs[li];
        if (r & kParam) {
            const int64_t e = (r & ~kParam) + idx;
            v->re = Part{false, 0, "p" + std::to_string(e) + "r"};
            v->im = Part{false, 0, "p" + std::to_string(e) + "i"};
            auto lit = plit.find(e);
            if (lit != plit.end()) {  // parts that are the same in every instance
                if (lit->second.has_re) v->re = Part{true, lit->second.re, "", false};
                if (lit->second.has_im) v->im = Part{true, lit->second.im, "", false};
            }
            return true;
        }
        const int64_t e = pv.const_base + r + idx;
        if (e < 0 || e >= pool_elems) {
            *err = tag + "constant leaf outside the pool";
            return false;
        }
        v->re = Part{true, c128 ? pool[2 * e] : double(float(pool[2 * e])), ""};
        v->im = Part{true, c128 ? pool[2 * e + 1] : double(float(pool[2 * e + 1])), ""};
        return true;
    };
    if (hermitian) {
        // every constant leaf must be conjugate-symmetric (to FP64 rounding)
        std::vector<int> leaf_bits(size_t(pv.n_leaves), -1);
        for (int si = 0; si < pv.n_steps; ++si) {
            const StepDesc& st = pv.steps[si];
            if (st.a < 0 && -1 - st.a < pv.n_leaves) leaf_bits[size_t(-1 - st.a)] = st.pad & 0xff;
            if (st.b < 0 && -1 - st.b < pv.n_leaves) leaf_bits[size_t(-1 - st.b)] = (st.pad >> 8) & 0xff;
        }
        for (int li = 0; li < pv.n_leaves; ++li) {
            const int64_t r = pv.refs[li];
            if ((r & kParam) || leaf_bits[size_t(li)] < 0) continue;
            const int64_t n = int64_t(1) << leaf_bits[size_t(li)], base = pv.const_base + r;
            if (base < 0 || base + n > pool_elems) {
                *err = tag + "constant leaf outside the pool";
                return false;
            }
            double mx = 0;
            for (int64_t e = 0; e < n; ++e)
                mx = std::max({mx, std::fabs(pool[2 * (base + e)]), std::fabs(pool[2 * (base + e) + 1])});
            for (int64_t e = 0; e < n; ++e) {
                const int64_t m = mirror_index(e);
                if (std::fabs(pool[2 * (base + m)] - pool[2 * (base + e)]) > 1e-10 * mx ||
                    std::fabs(pool[2 * (base + m) + 1] + pool[2 * (base + e) + 1]) > 1e-10 * mx) {
                    *err = tag + "a constant leaf is not conjugate-symmetric";
                    return false;
                }
            }
        }
    }
    // pass 1: the terms of every output element, with zeros propagated
    int64_t n_terms = 0;
    for (int si = 0; si < pv.n_steps; ++si) {
        const StepDesc& st = pv.steps[si];
        const int64_t nC = int64_t(1) << st.nc_bits, K = int64_t(1) << st.k_bits;
        const uint32_t* fc = pv.tabs + st.tab;
        const uint32_t* gs = fc + nC;
        const bool final_step = (st.flags & 1) != 0;
        if (final_step && (nC != 1 || si + 1 != pv.n_steps)) {
            *err = tag + "a specialised program must end in a scalar";
            return false;
        }
        auto& ts = terms[size_t(si)];
        auto& hs = has[size_t(si)];
        ts.resize(size_t(nC));
        hs.resize(size_t(nC));
        mirror[size_t(si)].assign(size_t(nC), -1);
        for (int64_t c = 0; c < nC; ++c) {
            const int64_t mc = hermitian ? mirror_index(c) : c;
            if (mc < c) {  // the conjugate of an element computed before
                mirror[size_t(si)][size_t(c)] = int(mc);
                hs[size_t(c)] = hs[size_t(mc)];
                continue;
            }
            const bool real_only = hermitian && mc == c;
            const uint32_t oa = fc[c] & 0xffffu, ob = fc[c] >> 16;
            bool hr = false, hi = false;
            for (int64_t s = 0; s < K; ++s) {
                Term t;
                if (!operand(st.a, int64_t(oa + (gs[s] & 0xffffu)), &t.a) ||
                    !operand(st.b, int64_t(ob + (gs[s] >> 16)), &t.b))
                    return false;
                t.f[0] = !zero(t.a.re) && !zero(t.b.re);
                t.f[1] = !zero(t.a.im) && !zero(t.b.im);
                t.f[2] = !real_only && !zero(t.a.re) && !zero(t.b.im);
                t.f[3] = !real_only && !zero(t.a.im) && !zero(t.b.re);
                if (!(t.f[0] || t.f[1] || t.f[2] || t.f[3])) continue;
                hr = hr || t.f[0] || t.f[1];
                hi = hi || t.f[2] || t.f[3];
                ts[size_t(c)].push_back(t);
                if (++n_terms > kMaxSpecialisedFmas) {  // bounded memory for programs that will not fit
                    *err = tag + "too large to specialise; the table executor runs it";
                    return false;
                }
            }
            hs[size_t(c)] = {hr, hi};
        }
        if (!final_step)
            for (int64_t c = 0; c < nC; ++c) scratch[st.c + int32_t(c)] = {si, int(c)};
    }
    // pass 2: liveness from the final scalar backwards
    std::vector<std::vector<char>> live(size_t(pv.n_steps));
    for (int si = 0; si < pv.n_steps; ++si) live[size_t(si)].assign(terms[size_t(si)].size(), 0);
    live.back()[0] = 1;
    for (int si = pv.n_steps - 1; si >= 0; --si)
        for (size_t c = 0; c < terms[size_t(si)].size(); ++c) {
            if (!live[size_t(si)][c]) continue;
            for (const Term& t : terms[size_t(si)][c])
                for (const Val* v : {&t.a, &t.b})
                    if (v->step >= 0) live[size_t(v->step)][size_t(v->elem)] = 1;
        }
    // pass 3: code.  A chain of FMAs identical to one already emitted (same operands
    // in the same order) is the same value: its variable is reused.
    std::ostringstream body;
    std::set<std::string> params;
    std::unordered_map<std::string, std::string> alias, chain_var;
    auto plain = [&](const Part& p) -> std::string {
        if (p.name[0] == 'p') params.insert(p.name.substr(0, p.name.size() - 1));
        auto it = alias.find(p.name);
        return it == alias.end() ? p.name : it->second;
    };
    auto name = [&](const Part& p) -> std::string {
        if (p.lit) return w.lit(p.v);
        return p.neg ? "(-" + plain(p) + ")" : plain(p);
    };
    auto neg = [&](const Part& p) -> std::string {
        if (p.lit) return w.lit(-p.v);
        return p.neg ? plain(p) : "(-" + plain(p) + ")";
    };
    auto emit = [&](const std::string& var, const std::vector<std::pair<std::string, std::string>>& ops) {
        std::string key;
        for (const auto& o : ops) key += o.first + "," + o.second + ";";
        auto it = chain_var.find(key);
        if (it != chain_var.end()) {
            alias[var] = it->second;
            return;
        }
        chain_var.emplace(key, var);
        body << "  " << w.real() << " " << var << " = 0;\n";
        for (const auto& o : ops)
            body << "  " << var << " = " << w.fma() << "(" << o.first << ", " << o.second << ", " << var << ");\n";
        out->fmas += int64_t(ops.size());
    };
    for (int si = 0; si < pv.n_steps; ++si)
        for (size_t c = 0; c < terms[size_t(si)].size(); ++c) {
            if (!live[size_t(si)][c] || mirror[size_t(si)][c] >= 0) continue;
            const std::string base = "t" + std::to_string(si) + "_" + std::to_string(c);
            const std::string vr = base + "r", vi = base + "i";
            std::vector<std::pair<std::string, std::string>> ore, oim;
            for (const Term& t : terms[size_t(si)][c]) {
                if (t.f[0]) ore.push_back({name(t.a.re), name(t.b.re)});
                if (t.f[1]) ore.push_back({neg(t.a.im), name(t.b.im)});
                if (t.f[2]) oim.push_back({name(t.a.re), name(t.b.im)});
                if (t.f[3]) oim.push_back({name(t.a.im), name(t.b.re)});
            }
            if (!ore.empty()) emit(vr, ore);
            if (!oim.empty()) emit(vi, oim);
            if (si + 1 == pv.n_steps) {
                const std::string zero = c128 ? "0.0" : "0.0f";
                body << "  out[0] = double(" << (ore.empty() ? zero : name(Part{false, 0, vr})) << ");\n";
                body << "  out[1] = double(" << (oim.empty() ? zero : name(Part{false, 0, vi})) << ");\n";
            }
        }
    std::ostringstream os;
    os << "static __device__ __attribute__((always_inline)) void prog" << id << "(const " << w.real()
       << "* __restrict__ P, double* __restrict__ out) {\n";
    for (const std::string& p : params) {
        const std::string e = p.substr(1);
        os << "  const " << w.real() << " " << p << "r = P[2 * " << e << "], " << p << "i = P[2 * " << e << " + 1];\n";
    }
    os << body.str() << "}\n\n";
    out->text = os.str();
    return true;
}

std::string generate(const std::vector<ProgView>& progs, bool c128, bool hermitian, const double* pool,
                     int64_t pool_elems, std::vector<int64_t>* fmas, std::string* err,
                     const std::vector<int32_t>* member_prog = nullptr) {
    std::ostringstream os;
    os << "// specialised executor: " << progs.size() << " programs, " << (c128 ? "complex128" : "complex64 4M")
       << (hermitian ? ", conjugate-symmetric leaves" : "") << "\n\n";
    if (fmas) fmas->clear();
    int64_t total = 0;
    for (size_t i = 0; i < progs.size(); ++i) {
        Program p;
        if (!emit_program(int(i), progs[i], c128, hermitian, pool, pool_elems, &p, err)) return std::string();
        total += p.fmas;
        if (total > kMaxSpecialisedFmas) {
            *err = "programs too large to specialise (more than " + std::to_string(kMaxSpecialisedFmas) +
                   " FMAs of straight-line code); the table executor runs them";
            return std::string();
        }
        os << p.text;
        if (fmas) fmas->push_back(p.fmas);
    }
    const char* real = c128 ? "double" : "float";
    // Instance records: int32 program, int32 pad, int64 parameter base (16 bytes)
    os << "extern \"C\" __global__ __launch_bounds__(64) void qt_batch_specialised(const " << real
       << "* __restrict__ pool, const long long* __restrict__ inst, const int* __restrict__ chunks, int n_chunks,\n"
       << "    double* __restrict__ out, int grid_m, int grid_sb, const double* __restrict__ prev,\n"
       << "    double* __restrict__ sums, int n_groups, int m, double c0, double c1, int n_sum_blocks,\n"
       << "    long long pb0, long long pb_stride, int pb_mod) {\n"
       // the first n_sum_blocks workgroups add up the previous launch's results per
       // group (member j of group g = instance j n_groups + g), in the order of the
       // library's strided group-sum kernel: 16 slices j = slice, slice + 16, ..., then
       // the slices in order
       << "  if (int(blockIdx.x) < n_sum_blocks) {  // 4 groups x 16 slices per workgroup\n"
       << "    __shared__ double part[16][4];\n"
       << "    const int gl = int(threadIdx.x) & 3, slice = int(threadIdx.x) >> 2;\n"
       << "    const int g = int(blockIdx.x) * 4 + gl;\n"
       << "    double p = 0.0;\n"
       << "    if (g < n_groups)\n"
       << "      for (int j = slice; j < m; j += 16) p += __builtin_fma(c1, prev[2 * ((long long)j * n_groups + g)], c0);\n"
       << "    part[slice][gl] = p;\n"
       << "    __syncthreads();\n"
       << "    if (slice == 0 && g < n_groups) {\n"
       << "      double t = part[0][gl];\n"
       << "      for (int k = 1; k < 16; ++k) t += part[k][gl];\n"
       << "      sums[g] = t;\n"
       << "    }\n"
       << "    return;\n"
       << "  }\n"
       << "  const unsigned bid = blockIdx.x - unsigned(n_sum_blocks), nb = gridDim.x - unsigned(n_sum_blocks);\n"
       // Workgroups go round-robin to the 8 XCDs.  Chunks in an edge-major grid (chunk
       // j * grid_sb + sb: program j, block sb of 64 instances sharing parameter blocks
       // with block sb of every other program): XCD x owns the blocks
       // [x per, (x + 1) per) and runs them in tiles of T blocks, all programs of a tile
       // before the next, so a tile's parameter blocks stay in that XCD's L2.  Otherwise
       // XCD x takes the contiguous chunk range [x per, (x + 1) per), so the waves
       // resident on a CU run few distinct programs.
       << "  int ci;\n"
       << "  if (grid_sb > 0) {\n"
       << "    const int per = grid_sb / 8, T = per < 16 ? per : 16;\n"
       << "    const int x = int(bid % 8u), k = int(bid / 8u);\n"
       << "    const int tile = k / (T * grid_m), r = k % (T * grid_m);\n"
       << "    ci = (r / T) * grid_sb + x * per + tile * T + r % T;\n"
       << "  } else {\n"
       << "    const int per = int(nb) / 8;\n"
       << "    ci = int(bid % 8u) * per + int(bid / 8u);\n"
       << "  }\n"
       << "  if (ci >= n_chunks) return;\n"
       << "  const int first = chunks[2 * ci], count = chunks[2 * ci + 1];\n"
       << "  if (int(threadIdx.x) >= count) return;\n"
       << "  const int i = first + int(threadIdx.x);\n"
       << "  const int prog = __builtin_amdgcn_readfirstlane(reinterpret_cast<const int*>(inst)[4 * first]);\n"
       // parameter base: pb0 + pb_stride (i mod pb_mod) when every instance has that
       // form (angle set i mod n_groups of an edge-major batch), else the record's
       << "  const long long pbase = pb_mod > 0 ? pb0 + pb_stride * (long long)(i % pb_mod) : inst[2 * i + 1];\n"
       << "  const " << real << "* P = pool + 2 * pbase;\n"
       << "  double* o = out + 2 * i;\n"
       << "  switch (prog) {\n";
    for (size_t i = 0; i < progs.size(); ++i) os << "    case " << i << ": prog" << i << "(P, o); break;\n";
    os << "    default: break;\n  }\n}\n";
    if (member_prog && !member_prog->empty()) {
        // Group sums only.  The first n_sum_blocks workgroups add up the previous
        // launch's 16 partial sums of 64 groups each, in slice order; then workgroup
        // (slice s, block of 64 groups) runs members j = s, s + 16, ... of its groups
        // and adds c0 + c1 Re(value) in that order: the strided group sums' order, so
        // the energies equal the table executor's bit for bit.  Workgroups go
        // round-robin to the 8 XCDs: XCD x runs slices 2x and 2x + 1 (few programs per
        // XCD for the instruction cache) of every block of groups.
        const int m = int(member_prog->size());
        os << "\nextern \"C\" __global__ __launch_bounds__(64) void qt_batch_specialised_sums(const " << real
           << "* __restrict__ pool, int n_groups, double* __restrict__ part, const double* __restrict__ prev,\n"
           << "    double* __restrict__ sums, int n_sum_blocks, double c0, double c1, long long pb0, long long pb_stride) {\n"
           << "  if (int(blockIdx.x) < n_sum_blocks) {\n"
           << "    const int g = int(blockIdx.x) * 64 + int(threadIdx.x);\n"
           << "    if (g < n_groups) {\n"
           << "      double t = prev[g];\n"
           << "      for (int k = 1; k < 16; ++k) t += prev[(long long)k * n_groups + g];\n"
           << "      sums[g] = t;\n"
           << "    }\n"
           << "    return;\n"
           << "  }\n"
           // XCD x (workgroups go round-robin to the 8 XCDs) owns the blocks of 64 groups
           // [x per, (x + 1) per) and runs them in tiles of T blocks, every slice of a
           // tile before the next tile: a tile's parameter blocks (T x 64 sets) stay in
           // that XCD's L2 for all 16 slices, and the waves in flight run few slices'
           // code at a time
           << "  const int bid = int(blockIdx.x) - n_sum_blocks;\n"
           << "  const int n_sb = (n_groups + 63) / 64, per = (n_sb + 7) / 8, T = per < " << kSumsTile << " ? per : "
           << kSumsTile << ";\n"
           << "  const int x = bid % 8, k = bid / 8;\n"
           << "  const int t = k / (16 * T), r = k % (16 * T);\n"
           << "  const int s = r / T, local = t * T + r % T;\n"
           << "  if (local >= per) return;\n"
           << "  const int g = (x * per + local) * 64 + int(threadIdx.x);\n"
           << "  if (g >= n_groups) return;\n"
           << "  const " << real << "* P = pool + 2 * (pb0 + pb_stride * (long long)g);\n"
           << "  double acc = 0.0, o[2];\n"
           << "  switch (s) {\n";
        for (int sl = 0; sl < 16; ++sl) {
            os << "    case " << sl << ":";
            for (int j = sl; j < m; j += 16)
                os << " prog" << (*member_prog)[size_t(j)] << "(P, o); acc += __builtin_fma(c1, o[0], c0);";
            os << " break;\n";
        }
        os << "    default: break;\n  }\n"
           << "  part[(long long)s * n_groups + g] = acc;\n}\n";
    }
    return os.str();
}

std::vector<ProgView> batch_views(const qt_batch* b) {
    std::vector<ProgView> v;
    for (size_t p = 0; p < b->h_hdr.size(); ++p) {
        const ProgHdr& h = b->h_hdr[p];
        ProgView pv;
        if (p < b->spec_param_lits.size()) pv.param_lits = &b->spec_param_lits[p];
        pv.steps = b->h_steps.data() + h.step_begin;
        pv.n_steps = h.n_steps;
        pv.tabs = b->h_tabs.data();
        pv.refs = b->h_refs.data() + h.leaf_begin;
        pv.n_leaves = h.n_leaves;
        pv.const_base = h.const_base;
        pv.dl = h.pad;  // the program's index bits per axis
        v.push_back(pv);
    }
    return v;
}

// Compiles generated source for `arch` (e.g. "gfx950" or a device's full gcnArchName)
// with hipRTC; packed FP32 instructions stay off, as for the library's own kernels.
bool rtc_compile(const std::string& src, const std::string& arch, std::vector<char>* code, std::string* err) {
    hiprtcProgram prog = nullptr;
    if (hiprtcCreateProgram(&prog, src.c_str(), "qt_batch_specialised.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        *err = "hiprtcCreateProgram failed";
        return false;
    }
    const std::string a = "--offload-arch=" + arch;
    const char* opts[] = {a.c_str(), "-O3", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult rc = hiprtcCompileProgram(prog, 6, opts);
    if (rc != HIPRTC_SUCCESS) {
        size_t n = 0;
        (void)hiprtcGetProgramLogSize(prog, &n);
        std::string log(n + 1, '\0');
        (void)hiprtcGetProgramLog(prog, &log[0]);
        (void)hiprtcDestroyProgram(&prog);
        *err = std::string(hiprtcGetErrorString(rc)) + ": " + log.substr(0, 2000);
        return false;
    }
    size_t size = 0;
    (void)hiprtcGetCodeSize(prog, &size);
    code->assign(size, 0);
    (void)hiprtcGetCode(prog, code->data());
    (void)hiprtcDestroyProgram(&prog);
    return true;
}

// Parameter leaves of a program: (offset in the instance's block, element bits).
std::vector<std::pair<int64_t, int>> param_leaves(const qt_batch* b, const ProgHdr& h) {
    std::vector<std::pair<int64_t, int>> out;
    for (int si = 0; si < h.n_steps; ++si) {
        const StepDesc& st = b->h_steps[size_t(h.step_begin + si)];
        for (int side = 0; side < 2; ++side) {
            const int32_t code = side == 0 ? st.a : st.b;
            if (code >= 0) continue;
            const int64_t r = b->h_refs[size_t(h.leaf_begin + (-1 - code))];
            if (r & QT_LEAF_PARAM)
                out.push_back({r & ~QT_LEAF_PARAM, side == 0 ? (st.pad & 0xff) : ((st.pad >> 8) & 0xff)});
        }
    }
    return out;
}

// Per program, the parameter elements with one value across all current instances.
std::vector<ParamLits> collect_param_lits(const qt_batch* b, const std::vector<double>& pool) {
    std::vector<ParamLits> lits(b->h_hdr.size());
    std::vector<std::vector<int64_t>> bases(b->h_hdr.size());
    for (int i = 0; i < b->n_inst; ++i) bases[size_t(b->h_prog[size_t(i)])].push_back(b->h_pbase[size_t(i)]);
    const int64_t n_pool = int64_t(pool.size() / 2);
    for (size_t p = 0; p < b->h_hdr.size(); ++p) {
        auto& bs = bases[p];
        std::sort(bs.begin(), bs.end());
        bs.erase(std::unique(bs.begin(), bs.end()), bs.end());
        if (bs.empty()) continue;
        std::set<int64_t> done;
        for (const auto& [r, bits] : param_leaves(b, b->h_hdr[p]))
            for (int64_t e = r; e < r + (int64_t(1) << bits); ++e) {
                if (!done.insert(e).second) continue;
                bool same_re = true, same_im = true;
                for (int64_t base : bs) {
                    if (base + e >= n_pool) {
                        same_re = same_im = false;
                        break;
                    }
                    same_re = same_re && pool[size_t(2 * (base + e))] == pool[size_t(2 * (bs[0] + e))];
                    same_im = same_im && pool[size_t(2 * (base + e) + 1)] == pool[size_t(2 * (bs[0] + e) + 1)];
                    if (!same_re && !same_im) break;
                }
                if (same_re || same_im)
                    lits[p].push_back(ParamLit{e, same_re, same_im, pool[size_t(2 * (bs[0] + e))],
                                               pool[size_t(2 * (bs[0] + e) + 1)]});
            }
    }
    return lits;
}

// Every parameter leaf of every instance satisfies T(mirror(x)) = conj T(x) (to
// rounding): the leaves' element counts come from the steps that read them.
bool params_conjugate_symmetric(const qt_batch* b, const std::vector<double>& pool) {
    std::set<std::pair<int64_t, int>> checked;  // (pool offset, element bits)
    for (int i = 0; i < b->n_inst; ++i) {
        const ProgHdr& h = b->h_hdr[size_t(b->h_prog[size_t(i)])];
        const int64_t pbase = b->h_pbase[size_t(i)];
        for (int si = 0; si < h.n_steps; ++si) {
            const StepDesc& st = b->h_steps[size_t(h.step_begin + si)];
            for (int side = 0; side < 2; ++side) {
                const int32_t code = side == 0 ? st.a : st.b;
                if (code >= 0) continue;
                const int64_t r = b->h_refs[size_t(h.leaf_begin + (-1 - code))];
                if (!(r & QT_LEAF_PARAM)) continue;
                const int bits = side == 0 ? (st.pad & 0xff) : ((st.pad >> 8) & 0xff);
                const int64_t base = pbase + (r & ~QT_LEAF_PARAM), n = int64_t(1) << bits;
                if (!checked.insert({base, bits}).second) continue;
                if (base + n > int64_t(pool.size() / 2)) return false;
                double mx = 0;
                for (int64_t e = 0; e < n; ++e)
                    mx = std::max({mx, std::fabs(pool[size_t(2 * (base + e))]), std::fabs(pool[size_t(2 * (base + e) + 1)])});
                for (int64_t e = 0; e < n; ++e) {
                    const int64_t m = mirror_index(e);
                    if (std::fabs(pool[size_t(2 * (base + m))] - pool[size_t(2 * (base + e))]) > 1e-6 * mx ||
                        std::fabs(pool[size_t(2 * (base + m) + 1)] + pool[size_t(2 * (base + e) + 1)]) > 1e-6 * mx)
                        return false;
                }
            }
        }
    }
    return true;
}

}  // namespace

// The generated kernel's assumptions about parameter blocks still hold for the
// current instances and the pool's host mirror: its parameter literals, and conjugate
// symmetry when it was generated with it.
bool specialised_params_hold(const qt_batch* b) {
    const std::vector<double>& pool = b->h_pool_mirror;
    const int64_t n_pool = int64_t(pool.size() / 2);
    for (int i = 0; i < b->n_inst; ++i) {
        const size_t p = size_t(b->h_prog[size_t(i)]);
        const int64_t base = b->h_pbase[size_t(i)];
        if (p >= b->spec_param_lits.size()) return false;
        for (const ParamLit& l : b->spec_param_lits[p])
            if (base + l.off >= n_pool || (l.has_re && pool[size_t(2 * (base + l.off))] != l.re) ||
                (l.has_im && pool[size_t(2 * (base + l.off) + 1)] != l.im))
                return false;
    }
    return !b->spec_hermitian || params_conjugate_symmetric(b, pool);
}

// Chunks of up to 64 consecutive instances of one program (one lane each).
qt_status specialised_chunks(qt_batch* b) {
    std::vector<int2> ch;
    for (int i = 0; i < b->n_inst;) {
        int j = i + 1;
        while (j < b->n_inst && j - i < 64 && b->h_prog[size_t(j)] == b->h_prog[size_t(i)]) ++j;
        ch.push_back(int2{i, j - i});
        i = j;
    }
    b->n_schunks = int(ch.size());
    // parameter bases of the form pb0 + stride (i mod n_groups): computed, not read
    b->spec_pb_mod = 0;
    if (b->n_groups > 0 && b->n_inst > 0 && b->n_inst % b->n_groups == 0) {
        const int64_t pb0 = b->h_pbase[0], stride = b->n_groups > 1 ? b->h_pbase[1] - b->h_pbase[0] : 0;
        bool affine = true;
        for (int i = 0; i < b->n_inst && affine; ++i) affine = b->h_pbase[size_t(i)] == pb0 + stride * (i % b->n_groups);
        if (affine) {
            b->spec_pb0 = pb0;
            b->spec_pb_stride = stride;
            b->spec_pb_mod = b->n_groups;
        }
    }
    // an edge-major grid of full chunks (member j of group g = instance j n_groups + g)
    // whose parameter blocks exceed an XCD's L2 (4 MiB): run in the tiled order
    b->spec_grid_m = b->spec_grid_sb = 0;
    int64_t param_end = 0;
    for (int64_t e : b->prog_param_end) param_end = std::max(param_end, e);
    const int64_t param_bytes = int64_t(b->n_groups) * param_end * (b->precision == QT_PLAIN_C128 ? 16 : 8);
    if (b->group_stride_m > 0 && b->n_groups % 64 == 0 && param_bytes > (int64_t(4) << 20)) {
        const int m = b->group_stride_m, n_sb = b->n_groups / 64, per = n_sb / 8;
        bool grid = n_sb % 8 == 0 && per > 0 && (per < 16 || per % 16 == 0) && int(ch.size()) == m * n_sb;
        for (size_t k = 0; k < ch.size() && grid; ++k) grid = ch[k].x == int(64 * k) && ch[k].y == 64;
        if (grid) {
            b->spec_grid_m = m;
            b->spec_grid_sb = n_sb;
        }
    }
    check_sums_fit(b);
    if (ch.empty()) return QT_OK;
    if (ch.size() > b->cap_schunks) {
        batch_free(b->ctx, b->d_schunks);
        b->d_schunks = nullptr;
        b->cap_schunks = 0;
        qt_status s = check_hip(b->ctx, batch_malloc(b->ctx, reinterpret_cast<void**>(&b->d_schunks), ch.size() * sizeof(int2)),
                                "specialised chunks");
        if (s != QT_OK) return s;
        b->cap_schunks = ch.size();
    }
    return check_hip(b->ctx, hipMemcpy(b->d_schunks, ch.data(), ch.size() * sizeof(int2), hipMemcpyHostToDevice),
                     "specialised chunks");
}

void drop_specialised(qt_batch* b) {
    if (b->sums_pending) (void)flush_sums(b);  // the table executor's launches sum their own results
    if (b->module) (void)hipModuleUnload(b->module);
    b->module = nullptr;
    b->kernel = nullptr;
    b->kernel_sums = nullptr;
    b->spec_member_prog.clear();
    b->spec_sums_ok = false;
    b->specialised = false;
    b->spec_consts.clear();
    b->spec_param_lits.clear();
    std::vector<double>().swap(b->h_pool_mirror);
    b->mirror_stale = false;
    ++b->layout_gen;
}

bool specialised_consts_change(const qt_batch* b, int64_t offset, const double* values, int64_t n) {
    const int64_t n_const = int64_t(b->spec_consts.size() / 2);
    const bool c128 = b->precision == QT_PLAIN_C128;
    for (int64_t e = std::max<int64_t>(offset, 0); e < std::min(offset + n, n_const); ++e)
        for (int t = 0; t < 2; ++t) {
            const double v = values[2 * (e - offset) + t];
            if ((c128 ? v : double(float(v))) != b->spec_consts[size_t(2 * e + t)]) return true;
        }
    return false;
}

void check_sums_fit(qt_batch* b) {
    b->spec_sums_ok = false;
    const int m = int(b->spec_member_prog.size());
    if (!b->kernel_sums || m == 0 || m != b->group_stride_m || b->spec_pb_mod != b->n_groups || b->n_groups == 0 ||
        int64_t(b->n_inst) != int64_t(m) * b->n_groups)
        return;
    for (int j = 0; j < m; ++j)
        for (int g = 0; g < b->n_groups; ++g)
            if (b->h_prog[size_t(j) * size_t(b->n_groups) + size_t(g)] != b->spec_member_prog[size_t(j)]) return;
    b->spec_sums_ok = true;
}

qt_status launch_part_sums(qt_batch* b) {
    b->in_use = true;  // queued work reads the batch's arrays
    Ctx* c = b->ctx;
    const void* pool = b->d_pool;
    int n_groups = b->n_groups;
    void* part = nullptr;
    const void* prev = b->d_out;
    double* sums = b->d_group;
    int n_sum_blocks = (b->n_groups + 63) / 64;
    double c0 = b->c0, c1 = b->c1;
    long long pb0 = b->spec_pb0, pb_stride = b->spec_pb_stride;
    void* params[] = {&pool, &n_groups, &part, &prev, &sums, &n_sum_blocks, &c0, &c1, &pb0, &pb_stride};
    return check_hip(c, hipModuleLaunchKernel(b->kernel_sums, unsigned(n_sum_blocks), 1, 1, 64, 1, 1, 0, c->stream,
                                              params, nullptr),
                     "specialised sums");
}

qt_status launch_specialised_sums(qt_batch* b, bool capture) {
    b->in_use = true;  // queued work reads the batch's arrays
    Ctx* c = b->ctx;
    if (b->sums_pending && !b->pending_parts) {  // a previous launch's per-instance values
        qt_status s = flush_sums(b);
        if (s != QT_OK) return s;
    }
    const size_t need = size_t(16) * size_t(b->n_groups);
    const int n_sb = (b->n_groups + 63) / 64;
    // partial sums of this launch in d_out, the previous one's in d_out_prev
    if (!capture) {
        std::swap(b->d_out, b->d_out_prev);
        std::swap(b->cap_out, b->cap_out_prev);
    }
    for (auto [buf, cap] : {std::make_pair(&b->d_out, &b->cap_out), std::make_pair(&b->d_out_prev, &b->cap_out_prev)}) {
        if (*cap >= need) continue;
        if (buf == &b->d_out_prev && (b->sums_pending || capture)) continue;  // holds pending partial sums
        if (capture) return set_error(c, QT_INVALID_INPUT, "capture: partial-sum buffer not reserved");
        batch_free(c, *buf);
        *buf = nullptr;
        *cap = 0;
        qt_status s = check_hip(c, batch_malloc(c, reinterpret_cast<void**>(buf), need * sizeof(double)),
                                "partial sums");
        if (s != QT_OK) return s;
        *cap = need;
    }
    const double* prev = nullptr;
    int n_sum_blocks = 0;
    if (!capture && b->sums_pending) {
        prev = b->d_out_prev;
        n_sum_blocks = (n_sb + 7) / 8 * 8;  // keeps the XCD of each slice workgroup
    }
    const void* pool = b->d_pool;
    int n_groups = b->n_groups;
    void* part = b->d_out;
    double* sums = b->d_group;
    double c0 = b->c0, c1 = b->c1;
    long long pb0 = b->spec_pb0, pb_stride = b->spec_pb_stride;
    void* params[] = {&pool, &n_groups, &part, &prev, &sums, &n_sum_blocks, &c0, &c1, &pb0, &pb_stride};
    // compute workgroups: 8 XCDs x n_tiles tiles x 16 slices x T blocks (the kernel's order)
    const int per = (n_sb + 7) / 8, T = std::min(per, kSumsTile), n_tiles = (per + T - 1) / T;
    const unsigned grid = unsigned(n_sum_blocks + 8 * 16 * T * n_tiles);
    qt_status s = check_hip(c, hipModuleLaunchKernel(b->kernel_sums, grid, 1, 1, 64, 1, 1, 0, c->stream, params,
                                                     nullptr),
                            "specialised sums");
    if (s != QT_OK) return s;
    b->values_valid = false;
    b->sums_pending = true;
    b->pending_parts = true;
    if (capture) {  // no deferral inside a graph: the sums follow at once
        s = launch_part_sums(b);
        b->sums_pending = false;
        b->pending_parts = false;
    }
    return s;
}

qt_status reserve_parts(qt_batch* b) {
    const size_t need = size_t(16) * size_t(b->n_groups);
    if (b->cap_out >= need) return QT_OK;
    qt_status s = flush_sums(b);  // whatever d_out holds is summed first
    if (s != QT_OK) return s;
    batch_free(b->ctx, b->d_out);
    b->d_out = nullptr;
    b->cap_out = 0;
    s = check_hip(b->ctx, batch_malloc(b->ctx, reinterpret_cast<void**>(&b->d_out), need * sizeof(double)),
                  "partial sums");
    if (s == QT_OK) b->cap_out = need;
    return s;
}

qt_status launch_specialised(qt_batch* b, bool defer_sums) {
    b->in_use = true;  // queued work reads the batch's arrays
    Ctx* c = b->ctx;
    if (b->sums_pending && b->pending_parts) {  // partial sums of a sums-only launch
        qt_status s = flush_sums(b);
        if (s != QT_OK) return s;
    }
    b->values_valid = true;
    if (b->needs_tout) return set_error(c, QT_INVALID_INPUT, "batch: specialised programs end in scalars");
    if (b->n_schunks == 0) return QT_OK;
    // deferred group sums: this launch writes the other buffer and sums the previous
    // launch's results (when they are still unsummed) in its first workgroups
    const double* prev = nullptr;
    int n_sum_blocks = 0;
    if (defer_sums) {
        std::swap(b->d_out, b->d_out_prev);
        std::swap(b->cap_out, b->cap_out_prev);  // capacities follow their buffers
        if (b->sums_pending) {
            prev = b->d_out_prev;
            n_sum_blocks = (b->n_groups + 3) / 4;
            n_sum_blocks = (n_sum_blocks + 7) / 8 * 8;  // keeps the XCD of each chunk workgroup
        }
        b->sums_pending = true;
    }
    double* sums = b->d_group;
    int n_groups = b->n_groups, m = b->group_stride_m;
    double c0 = b->c0, c1 = b->c1;
    const void* pool = b->d_pool;
    const void* inst = b->d_inst;
    const void* chunks = b->d_schunks;
    int n_chunks = b->n_schunks;
    void* out = b->d_out;
    int grid_m = b->spec_grid_m, grid_sb = b->spec_grid_sb;
    long long pb0 = b->spec_pb0, pb_stride = b->spec_pb_stride;
    int pb_mod = b->spec_pb_mod;
    void* params[] = {&pool, &inst, &chunks, &n_chunks, &out, &grid_m, &grid_sb, &prev, &sums, &n_groups,
                      &m, &c0, &c1, &n_sum_blocks, &pb0, &pb_stride, &pb_mod};
    // a multiple of the XCD count (exact for a grid), after the sum workgroups
    const unsigned grid = unsigned(n_sum_blocks + (n_chunks + 7) / 8 * 8);
    return check_hip(c, hipModuleLaunchKernel(b->kernel, grid, 1, 1, 64, 1, 1, 0, c->stream, params, nullptr),
                     "specialised kernel");
}

}  // namespace qt

using qt::check_hip;
using qt::set_error;

extern "C" {

qt_status qt_program_source(int n, qt_program* const* progs, const int64_t* const_base, const double* pool,
                            int64_t pool_elems, int precision, int flags, char* buf, size_t cap, size_t* len) {
    if ((flags & ~QT_SPECIALISE_HERMITIAN) != 0) return QT_INVALID_INPUT;
    if (n < 1 || !progs || !const_base || (pool_elems > 0 && !pool) || pool_elems < 0 || !len ||
        (precision != QT_C64_4M && precision != QT_PLAIN_C128))
        return QT_INVALID_INPUT;
    std::vector<qt::ProgView> views;
    for (int i = 0; i < n; ++i) {
        if (!progs[i]) return QT_INVALID_INPUT;
        qt::ProgView pv;
        pv.steps = progs[i]->steps.data();
        pv.n_steps = int(progs[i]->steps.size());
        pv.tabs = progs[i]->tables.data();
        pv.refs = progs[i]->leaf_refs.data();
        pv.n_leaves = int(progs[i]->leaf_refs.size());
        pv.const_base = const_base[i];
        pv.dl = progs[i]->dim_log2;
        views.push_back(pv);
    }
    std::string err;
    const std::string src = qt::generate(views, precision == QT_PLAIN_C128, (flags & QT_SPECIALISE_HERMITIAN) != 0,
                                         pool, pool_elems, nullptr, &err);
    if (src.empty()) return QT_INVALID_INPUT;
    *len = src.size();
    if (buf && cap > 0) {
        const size_t m = std::min(cap - 1, src.size());
        std::memcpy(buf, src.data(), m);
        buf[m] = '\0';
    }
    return QT_OK;
}

qt_status qt_program_compile(const char* source, const char* arch, void* code, size_t cap, size_t* len) {
    if (!source || !arch || !len) return QT_INVALID_INPUT;
    std::vector<char> co;
    std::string err;
    if (!qt::rtc_compile(source, arch, &co, &err)) return QT_DEVICE;
    *len = co.size();
    if (code && cap >= co.size()) std::memcpy(code, co.data(), co.size());
    return QT_OK;
}

qt_status qt_batch_compile(qt_batch* b, int flags, float* compile_ms) {
    if (!b || (flags & ~QT_SPECIALISE_HERMITIAN) != 0) return QT_INVALID_INPUT;
    const bool hermitian = (flags & QT_SPECIALISE_HERMITIAN) != 0;
    auto* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    if (b->precision != QT_C64_4M && b->precision != QT_PLAIN_C128)
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_compile: only the complex64 4M and complex128 forms");
    if (b->n_programs == 0 || b->pool_elems == 0)
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_compile: set programs and leaves first");
    for (int32_t oe : b->prog_out_elems)
        if (oe != 1) return set_error(ctx, QT_INVALID_INPUT, "qt_batch_compile: programs must end in scalars");
    const auto t0 = std::chrono::steady_clock::now();
    // constant leaves as FP64 (complex64 pools convert exactly); with conjugate
    // symmetry the whole pool, to check the instances' parameter blocks too
    // the whole pool: constants, and the parameter blocks (their common elements
    // become literals; conjugate symmetry is checked on them)
    const int64_t n_const = b->pool_elems;
    std::vector<double> pool(size_t(2 * std::max<int64_t>(n_const, 1)), 0.0);
    qt_status s = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (s == QT_OK && n_const > 0) {
        if (b->precision == QT_PLAIN_C128) {
            s = check_hip(ctx, hipMemcpy(pool.data(), b->d_pool, size_t(n_const) * 16, hipMemcpyDeviceToHost),
                          "pool read");
        } else {
            std::vector<float> f(size_t(2 * n_const));
            s = check_hip(ctx, hipMemcpy(f.data(), b->d_pool, size_t(n_const) * 8, hipMemcpyDeviceToHost), "pool read");
            for (size_t i = 0; i < f.size(); ++i) pool[i] = double(f[i]);
        }
    }
    if (s != QT_OK) return s;
    if (hermitian && !qt::params_conjugate_symmetric(b, pool))
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_compile: a parameter block is not conjugate-symmetric");
    std::vector<qt::ParamLits> lits = qt::collect_param_lits(b, pool);
    std::swap(lits, b->spec_param_lits);  // read by batch_views
    std::string err;
    std::vector<int64_t> fmas;
    // an edge-major batch whose members each run one program in every group also gets
    // the group-sums-only kernel
    std::vector<int32_t> member_prog;
    if (b->group_stride_m > 0 && b->n_groups > 0 && int64_t(b->n_inst) == int64_t(b->group_stride_m) * b->n_groups) {
        member_prog.resize(size_t(b->group_stride_m));
        bool ok = true;
        for (int j = 0; j < b->group_stride_m && ok; ++j) {
            member_prog[size_t(j)] = b->h_prog[size_t(j) * size_t(b->n_groups)];
            for (int g = 1; g < b->n_groups && ok; ++g)
                ok = b->h_prog[size_t(j) * size_t(b->n_groups) + size_t(g)] == member_prog[size_t(j)];
        }
        if (!ok) member_prog.clear();
    }
    const std::string src = qt::generate(qt::batch_views(b), b->precision == QT_PLAIN_C128, hermitian, pool.data(),
                                         n_const, &fmas, &err, &member_prog);
    if (src.empty()) {
        std::swap(lits, b->spec_param_lits);  // keep the state of the kernel that runs
        return set_error(ctx, QT_INVALID_INPUT, "qt_batch_compile: " + err);
    }
    hipDeviceProp_t prop;
    s = check_hip(ctx, hipGetDeviceProperties(&prop, ctx->device), "device properties");
    if (s != QT_OK) return s;
    std::vector<char> code;
    if (!qt::rtc_compile(src, prop.gcnArchName, &code, &err)) {
        std::swap(lits, b->spec_param_lits);
        return set_error(ctx, QT_DEVICE, "qt_batch_compile: " + err);
    }
    hipModule_t mod = nullptr;
    hipFunction_t fn = nullptr;
    s = check_hip(ctx, hipModuleLoadData(&mod, code.data()), "module load");
    if (s == QT_OK) s = check_hip(ctx, hipModuleGetFunction(&fn, mod, "qt_batch_specialised"), "module function");
    hipFunction_t fn_sums = nullptr;
    if (s == QT_OK && !member_prog.empty())
        s = check_hip(ctx, hipModuleGetFunction(&fn_sums, mod, "qt_batch_specialised_sums"), "module function");
    if (s != QT_OK) {
        if (mod) (void)hipModuleUnload(mod);
        std::swap(lits, b->spec_param_lits);
        return s;
    }
    std::vector<qt::ParamLits> new_lits = std::move(b->spec_param_lits);
    qt::drop_specialised(b);
    b->spec_param_lits = std::move(new_lits);
    b->h_pool_mirror = pool;
    b->mirror_stale = false;
    ++b->layout_gen;
    b->module = mod;
    b->kernel = fn;
    b->kernel_sums = fn_sums;
    b->spec_member_prog = member_prog;
    b->spec_consts.assign(pool.begin(), pool.begin() + 2 * std::min(n_const, b->max_leaf_ref_end));
    b->spec_hermitian = hermitian;
    b->spec_fmas = fmas;
    s = qt::specialised_chunks(b);
    if (s != QT_OK) {
        qt::drop_specialised(b);
        return s;
    }
    b->specialised = true;
    // a captured graph holds the table executor
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
    if (compile_ms)
        *compile_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return QT_OK;
}

qt_status qt_batch_specialised_info(const qt_batch* b, int* active, int64_t* fmas_per_launch) {
    if (!b) return QT_INVALID_INPUT;
    if (active) *active = b->specialised ? 1 : 0;
    if (fmas_per_launch) {
        int64_t n = 0;
        if (b->specialised)
            for (int i = 0; i < b->n_inst; ++i) n += b->spec_fmas[size_t(b->h_prog[size_t(i)])];
        *fmas_per_launch = n;
    }
    return QT_OK;
}

}  // extern "C"
