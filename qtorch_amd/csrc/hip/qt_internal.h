// Internal declarations shared by the HIP sources of libqtorch_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "qtorch_hip.h"

namespace qt {

// Caching device allocator: stream-ordered reuse on the context's single stream,
// so a block freed by one node can back the next node without hipFree/hipMalloc
// (hipFree synchronises the whole device).
class Pool {
public:
    void* alloc(size_t bytes, hipError_t* err);
    void release(void* p);
    void trim();
    int64_t reserved() const { return reserved_; }
    static size_t round(size_t bytes);

private:
    std::map<size_t, std::vector<void*>> free_;
    std::map<void*, size_t> sizes_;
    int64_t reserved_ = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    Pool pool;
    std::string err;
    std::recursive_mutex mu;
    // partial sums of the split reduction kernel
    void* partials = nullptr;
    size_t partials_bytes = 0;
    // pinned host staging for batched uploads: two buffers used in turn, each reused
    // once the copy that last read it has finished (its event)
    void* staging[2] = {nullptr, nullptr};
    size_t staging_bytes[2] = {0, 0};
    hipEvent_t staging_ev[2] = {nullptr, nullptr};
    bool staging_pending[2] = {false, false};
    int staging_next = 0;
    // tensors and batches alive on this context: it is not destroyed under them
    int64_t live = 0;
    // power-of-two exponents of rescaled complex64 tensors (value = data * 2^exp)
    int32_t* d_exp = nullptr;
    std::vector<int32_t> exp_free;
    int32_t exp_cap = 0;
    float* d_absmax = nullptr;  // per-workgroup maxima of one rescale
    // qt_ctx_set_time_limit: work requested after this instant returns QT_TIMEOUT
    bool has_deadline = false;
    std::chrono::steady_clock::time_point deadline{};
};

// QT_TIMEOUT (with the message set) when the context's time limit has passed.
qt_status check_deadline(Ctx* c, const char* what);

struct Tensor {
    void* data = nullptr;
    // tensors created together share one allocation; the last one freed returns it
    void* block = nullptr;
    int* block_refs = nullptr;
    int64_t numel = 0;
    int rank = 0;
    int dim_log2 = 2;
    int dtype = QT_DTYPE_C128;
    size_t bytes = 0;
    int32_t exp_slot = -1;  // slot in Ctx::d_exp once rescaled (-1: exponent 0)
};

inline size_t elem_bytes(int dtype) { return dtype == QT_DTYPE_C128 ? 16 : 8; }

// Bit-level description of one pairwise contraction.  Every D-ary axis is
// dim_log2 bits of the flat index; the engine works on bits throughout.
//   c_src[i] : source of output bit i: A bit p -> p, B bit p -> 64 + p
//   s_a[j], s_b[j] : A and B bit positions of summed bit j
struct BitMap {
    int32_t nbits_c = 0;
    int32_t nbits_s = 0;
    uint8_t c_src[64];
    uint8_t s_a[64];
    uint8_t s_b[64];
};

// Generic bit permutation for the permute kernel: output bit j <- input bit src[j].
struct PermPlan {
    int32_t nbits = 0;
    int32_t tl = 0;          // low bits handled by the tile on each side
    int32_t n_extra_in = 0;  // tile bits beyond the input-low ones (output-low bits with src >= tl)
    int32_t n_extra_out = 0; // tile bits beyond the output-low ones (input-low bits not in output-low)
    int32_t n_rest = 0;      // bits outside the tile, enumerated by blockIdx
    uint8_t extra_in_pos[16];   // input bit positions of the extra tile bits (read side)
    uint8_t extra_in_out[16];   // output bit positions of those same bits
    uint8_t lowin_out[16];      // output bit position of input-low bit i (i < tl)
    uint8_t extra_out_pos[16];  // output bit positions of the extra bits on the write side
    uint8_t extra_out_tile[16]; // tile-index bit of those bits
    uint8_t lowout_tile[16];    // tile-index bit of output-low bit j (j < tl)
    uint8_t rest_in[64];
    uint8_t rest_out[64];
};

qt_status set_error(Ctx* c, qt_status s, const std::string& msg);
qt_status check_hip(Ctx* c, hipError_t e, const char* what);

// Kernel launchers (qt_contract.hip)
qt_status launch_fill_zero(Ctx* c, void* p, size_t bytes);
// Device scratch a contraction may use instead of allocating: the GEMM operand
// permutes and the split reduction's partial sums come from [base, base + bytes).
struct Workspace {
    char* base = nullptr;
    size_t bytes = 0;
};
// ws == nullptr: scratch from the context's caching allocator.  With a workspace,
// a step whose scratch exceeds it returns QT_INVALID_INPUT before launching anything.
qt_status contract_pair(Ctx* c, const Tensor* A, const Tensor* B, Tensor* C, const BitMap& m,
                        int precision, const Workspace* ws = nullptr);
qt_status permute(Ctx* c, const void* in, void* out, int dtype, const PermPlan& p);
// Scales complex64 C by 2^-e so that its largest |re|, |im| lies in [0.5, 1) and
// records exp(C) = exp(A) + exp(B) + e on the device (no host synchronisation).
qt_status rescale_c64(Ctx* c, const Tensor* A, const Tensor* B, Tensor* C);
// Reads a tensor's exponent (0 if never rescaled); synchronises.
qt_status read_exponent(Ctx* c, const Tensor* t, int32_t* e);
void release_exponent(Ctx* c, Tensor* t);
// exp(C) = exp(A) + exp(B) on the device, for a step without rescaling whose
// operands carry exponents from earlier rescaled steps.
qt_status carry_exponent(Ctx* c, const Tensor* A, const Tensor* B, Tensor* C);
// acc (complex128 scalar) += t (scalar, exponent applied), in FP64 on the stream.
qt_status accumulate_scalar(Ctx* c, const Tensor* t, Tensor* acc);
// tile_bits_max: a tile holds at most 2^tile_bits_max elements; its input-low and
// output-low bits (contiguous runs of 2^tl elements) are as many as that allows
PermPlan make_perm_plan(int nbits, const int* src, int tile_bits_max = 12);

}  // namespace qt

struct qt_ctx {
    qt::Ctx impl;
};
struct qt_tensor {
    qt::Tensor impl;
};
