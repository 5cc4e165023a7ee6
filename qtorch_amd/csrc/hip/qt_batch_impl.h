// Internal structures of programs and batches, shared by qt_batch.hip (table
// executors) and qt_codegen.hip (specialised executors).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <tuple>
#include <vector>

#include "qt_internal.h"

namespace qt {

// Parts of parameter elements (offset in the instance's block) that hold one value in
// every instance of a program (a real part that is always 0, say, while the imaginary
// part varies); generated code has them as literals.
struct ParamLit {
    int64_t off;
    bool has_re, has_im;  // which parts are literals
    double re, im;
};
using ParamLits = std::vector<ParamLit>;

struct ProgHdr {
    int32_t step_begin, n_steps;
    int32_t leaf_begin, n_leaves;
    int32_t scratch_elems;
    int32_t flags;
    int64_t const_base;  // pool offset added to the program's constant leaf references
    int32_t cst_begin, n_cst;  // the program's distinct constant leaves (CstEntry list)
    int32_t param_elems;       // extent of an instance's parameter block that it reads
    int32_t pad;               // index bits per axis (host use only)
};

// A distinct constant leaf of a program and its place in a wave's LDS leaf area.
struct CstEntry {
    int64_t src;  // pool offset relative to the program's const_base
    int32_t n;    // elements
    int32_t dst;  // element offset in the wave's constant area
};

struct StepDesc {
    int32_t a, b;       // >= 0: scratch element offset; < 0: leaf -(1 + index)
    int32_t c;          // scratch element offset of the result (final step: unused)
    int32_t tab;        // offset of the step's tables in the table array (uint32 units)
    int32_t nc_bits;    // output elements = 1 << nc_bits
    int32_t k_bits;     // summed elements = 1 << k_bits
    int32_t flags;      // bit 0: final step
    int32_t pad;
};

struct Instance {
    int32_t prog;
    int32_t pad;
    int64_t param_base;
};

}  // namespace qt

struct qt_program {
    std::vector<qt::StepDesc> steps;  // offsets relative to this program's table block
    std::vector<uint32_t> tables;
    std::vector<int64_t> leaf_refs;
    std::vector<int32_t> leaf_ranks;
    int dim_log2 = 2;
    int64_t cmacs = 0;
    int32_t scratch_elems = 0;
    int32_t max_tensor_elems = 0;
    int32_t out_elems = 1;  // elements of the last step's result
};

struct qt_batch {
    qt::Ctx* ctx = nullptr;
    int precision = QT_C64_4M;
    int n_programs = 0;
    int32_t scratch_elems = 0;  // max over programs
    int64_t cmacs_total = 0;    // per-instance sum, for instances set
    std::vector<int64_t> prog_cmacs;
    // device arrays
    qt::ProgHdr* d_progs = nullptr;
    qt::StepDesc* d_steps = nullptr;
    uint32_t* d_tabs = nullptr;
    int64_t* d_leaf_refs = nullptr;
    void* d_pool = nullptr;  // leaf pool in the batch precision
    int64_t pool_elems = 0;
    int64_t max_leaf_ref_end = 0;  // absolute part
    int64_t max_param_ref_end = 0; // relative part
    qt::Instance* d_inst = nullptr;
    int n_inst = 0;
    double* d_out = nullptr;       // 2 * n_inst
    size_t cap_inst = 0, cap_out = 0, cap_seg = 0, cap_members = 0, cap_group = 0, cap_chunks = 0;
    int32_t* d_counter = nullptr;
    // group sums
    int32_t* d_seg = nullptr;
    int32_t* d_members = nullptr;
    double* d_group = nullptr;
    int n_groups = 0;
    int group_stride_m = 0;  // > 0: group g = instances j * n_groups + g, j < m (k_group_sum_strided)
    double c0 = 0.0, c1 = 1.0;
    void* d_global_scratch = nullptr;
    size_t global_scratch_bytes = 0;
    // launch configuration
    int lanes = 64;  // 256: a workgroup per instance; 64: a wavefront; < 64: lanes per instance in a wave
    int wgs_per_cu = 4;
    int scratch_place = QT_SCRATCH_AUTO;  // with QT_SCRATCH_STAGE_LEAVES or-ed in
    bool stage_leaves = false;
    bool shape_ok = false;  // device arrays fit the executor shape
    int32_t slot_stride = 0;      // elements between the scratch areas of two executors
    int32_t pstage = 0, cstage = 0;  // LDS leaf areas: parameter block per slot, constant leaves per wave
    std::vector<int32_t> h_prog;  // program of each instance (host copy, for chunking)
    std::vector<int32_t> prog_out_elems;  // elements of each program's result
    void** d_tout = nullptr;      // per instance: output tensor of the last step (qt_batch_set_outputs)
    size_t cap_tout = 0;
    bool needs_tout = false;      // an instance's program ends in a tensor
    bool tensor_out = false;      // output tensors are set for the current instances
    int2* d_chunks = nullptr;
    int n_chunks = 0;
    int n_cus = 0;
    bool use_global_scratch = false;
    bool rescale = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_done = nullptr;
    bool ev_pending = false;  // launches since ev_done was last recorded
    bool in_use = false;      // work that reads or writes the batch's device arrays may be queued
    // host copies of the device program arrays (programs can be appended)
    std::vector<qt::ProgHdr> h_hdr;
    std::vector<qt::StepDesc> h_steps;
    std::vector<uint32_t> h_tabs;
    std::vector<int64_t> h_refs;
    std::vector<int32_t> h_leaf_lds;     // per leaf: offset in the constant area, or (1 << 31) | parameter offset
    std::vector<qt::CstEntry> h_csts;
    int32_t* d_leaf_lds = nullptr;
    qt::CstEntry* d_csts = nullptr;
    size_t cap_leaf_lds = 0, cap_csts = 0;
    std::vector<int64_t> prog_const_end, prog_param_end;  // pool extents each program reads
    size_t cap_progs = 0, cap_steps = 0, cap_tabs = 0, cap_refs = 0;
    int64_t pool_cap = 0;           // elements allocated for the pool
    void* d_pool_stage = nullptr;   // FP64 staging of pool writes for complex64 batches
    int64_t pool_stage_cap = 0;
    // specialised executor (qt_batch_compile): generated kernel and its chunks of up to
    // 64 consecutive instances of one program (one lane each)
    hipModule_t module = nullptr;
    hipFunction_t kernel = nullptr;
    bool specialised = false;
    std::vector<double> spec_consts;  // constant leaves the kernel was generated with (FP64)
    std::vector<int64_t> spec_fmas;   // FMAs per instance of each program in the generated code
    bool spec_hermitian = false;      // generated with conjugate-symmetric leaves
    std::vector<qt::ParamLits> spec_param_lits;  // per program: parameter elements generated as literals
    std::vector<double> h_pool_mirror;  // host copy of the pool while generated code runs (complex64 values rounded)
    std::vector<int64_t> h_pbase;     // parameter base of each instance (host copy)
    int2* d_schunks = nullptr;
    size_t cap_schunks = 0;
    int n_schunks = 0;
    int spec_grid_m = 0, spec_grid_sb = 0;  // > 0: chunks form an edge-major grid (programs x 64-instance blocks)
    int64_t spec_pb0 = 0, spec_pb_stride = 0;  // parameter base of instance i = pb0 + stride (i mod pb_mod)
    int spec_pb_mod = 0;                       // 0: read from the instance records
    // Deferred group sums (specialised kernel, strided groups, direct launches): each
    // launch's kernel also sums the previous launch's results, kept in the other result
    // buffer; the last launch's sums run when results are read, polled or flushed.
    double* d_out_prev = nullptr;
    size_t cap_out_prev = 0;
    bool sums_pending = false;  // d_out's group sums have not been launched yet
    // Group sums only (qt_batch_set_sums_only): for an edge-major batch whose members
    // j < m each run one program in every group, a second generated kernel has each
    // wavefront run 64 groups through the members j = s, s + 16, ... of one slice s,
    // adding c0 + c1 Re(value) in that order (the strided group sums' order), and
    // writes 16 partial sums per group; the sums over the slices run in the next
    // launch's first workgroups (or at a flush).  No per-instance value is written.
    hipFunction_t kernel_sums = nullptr;
    std::vector<int32_t> spec_member_prog;  // the program of member j the kernel was generated with
    bool sums_only = false;                 // requested by the caller
    bool spec_sums_ok = false;              // the instances fit the generated sums kernel
    bool pending_parts = false;             // the pending sums are over partial sums (d_out)
    bool values_valid = true;               // d_out holds the last launch's per-instance values
    // rotation blocks written on the device (qt_batch_rotation_params)
    bool mirror_stale = false;        // h_pool_mirror misses device-written blocks
    int* d_param_status = nullptr;    // set by the kernel for an angle out of range
    bool params_unchecked = false;    // a rotation launch since the status was last read
    double* d_angles = nullptr;       // host-given angles, copied
    size_t cap_angles = 0;
    uint64_t layout_gen = 1;          // changes with the instances or the generated kernel
    uint64_t rot_checked_gen = 0;     // layout_gen and arguments the literal check last passed for
    std::vector<char> rot_checked_args;
    bool rot_full_valid = false;      // every element of the rot_full_args layout holds its table value
    std::vector<char> rot_full_args;
    void* d_rot_desc = nullptr;       // the last launch's gate description (device copy)
    std::vector<char> rot_desc_host;  // and its host copy
};

namespace qt {
// Device arrays of a batch come from its context's caching allocator: no hipMalloc /
// hipFree (which synchronises the whole device) once the cache is warm.  A block is
// returned after the context's stream has drained, so work in flight never sees it
// reused.
inline hipError_t batch_malloc(Ctx* c, void** p, size_t bytes) {
    hipError_t e = hipSuccess;
    *p = c->pool.alloc(bytes, &e);
    return *p ? hipSuccess : (e == hipSuccess ? hipErrorOutOfMemory : e);
}
inline void batch_free(Ctx* c, void* p) {
    if (!p) return;
    (void)hipStreamSynchronize(c->stream);
    c->pool.release(p);
}
// specialised executors (qt_codegen.hip)
qt_status specialised_chunks(qt_batch* b);
void drop_specialised(qt_batch* b);
qt_status launch_specialised(qt_batch* b, bool defer_sums);
// group sums only: partial sums per slice, the sums over slices deferred (or, when
// captured, launched right after)
qt_status launch_specialised_sums(qt_batch* b, bool capture);
// the sums of the pending partial sums (a sums-only launch's), into d_group
qt_status launch_part_sums(qt_batch* b);
// whether the current instances fit the generated sums kernel (sets spec_sums_ok)
void check_sums_fit(qt_batch* b);
// sizes d_out for the partial sums before a capture (no allocation inside one)
qt_status reserve_parts(qt_batch* b);
// group sums of d_out on the context's stream (qt_batch.hip)
qt_status launch_group_sums(qt_batch* b);
// launches the group sums still pending for the last launch
qt_status flush_sums(qt_batch* b);
// a pool write of n values at offset would change constants the kernel embeds
bool specialised_consts_change(const qt_batch* b, int64_t offset, const double* values, int64_t n);
// the generated kernel's assumptions on parameter blocks hold for the current
// instances and the pool's host mirror
bool specialised_params_hold(const qt_batch* b);
// waits for the batch's queued work before its device arrays are replaced (qt_batch.hip);
// a batch that has queued nothing since the last wait returns at once
void quiesce(qt_batch* b);
// re-reads h_pool_mirror from the device after device-side parameter writes (qt_params.hip)
qt_status refresh_pool_mirror(qt_batch* b);
// a dropped specialised executor: rebuilds the captured graph's launch state (qt_batch.hip)
qt_status after_specialised_drop(qt_batch* b);
}  // namespace qt
