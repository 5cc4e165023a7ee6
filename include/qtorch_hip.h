/*
 * qtorch_hip.h — C ABI of libqtorch_hip.so, the MI355X (gfx950) contraction engine
 * behind qTorch's pairwise contraction.
 *
 * Plain C types only (no C++ or torch types cross this boundary); every entry
 * point returns a qt_status and never throws.  Tensors are dense, device
 * resident, interleaved complex (re, im) with the reference's flat layout:
 * index = sum_i d_i * D^i over the node's wire order, D = 4 for qTorch's
 * Liouville wires (reference src/Node.h:112-113, 168-186) and D = 2 for
 * amplitude networks (an extension of this engine).
 *
 * Threading: a context owns one HIP stream and is internally serialised; every
 * call on a context is stream ordered.  Use one context per host thread for
 * concurrency.  Calls that return host data (downloads, batch results)
 * synchronise the context's stream.
 */
#ifndef QTORCH_HIP_H
#define QTORCH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qt_ctx qt_ctx;
typedef struct qt_tensor qt_tensor;
typedef struct qt_program qt_program;

typedef enum {
    QT_OK = 0,
    QT_INVALID_INPUT = 1, /* maps to qtorch::InvalidFunctionInput (src/Exceptions.h:44) */
    QT_OOM = 2,
    QT_DEVICE = 3,        /* HIP runtime or kernel failure */
    QT_TIMEOUT = 4,       /* maps to qtorch::ContractionFailure (src/Exceptions.h:36) */
    QT_UNSUPPORTED = 5
} qt_status;

/* Element types of device tensors. */
typedef enum {
    QT_DTYPE_C128 = 0, /* std::complex<double>, the reference's Node storage (src/Node.h:160) */
    QT_DTYPE_C64 = 1   /* std::complex<float> */
} qt_dtype;

/* Arithmetic forms of a contraction. */
typedef enum {
    QT_PLAIN_C128 = 0, /* complex128 operands and accumulation: the drop-in form */
    QT_C64_4M = 1,     /* complex64, 4 real products per complex MAC, FP32 accumulation */
    QT_C64_3M = 2,     /* complex64, 3 real products per complex MAC (Gauss), FP32 accumulation */
    QT_C64_BF16X3 = 3  /* complex64 operands split hi+lo into bf16; hi*hi + hi*lo + lo*hi on
                          MFMA with FP32 accumulation */
} qt_precision;

/* Or-ed into a complex64 precision of qt_contract_pair: after the step, C is scaled
 * by a power of two so that its largest |re|, |im| lies in [0.5, 1), and the
 * exponent is kept on the device beside C (value = data * 2^exponent, exponents of A
 * and B included), so that long chains of small amplitudes cannot underflow FP32.
 * A step without the flag whose operands carry exponents gives C their sum. */
#define QT_RESCALE (1 << 8)

/* Bit 30 of a free-axis code marks an axis of A (else of B). */
#define QT_FROM_A (1 << 30)

/* ---- context ----------------------------------------------------------------- */

/* Creates a context on HIP device `device` (stream, caching allocator). */
qt_status qt_ctx_create(int device, qt_ctx** out);
qt_status qt_ctx_destroy(qt_ctx* ctx);
/* Time limit for work on the context, counted from this call: once `seconds` have
 * passed, qt_contract_pair and qt_batch_launch return QT_TIMEOUT without queueing
 * anything (checked when each call is made, i.e. between pairwise steps; work
 * already queued completes).  seconds <= 0 removes the limit.  It replaces the
 * reference's totTimer / maxTime checks (src/Network.h:50-51,899-910,
 * src/ContractionTools.h:271,332), which stop a contraction once maxTime seconds
 * have passed and end in ContractionFailure (src/ContractionTools.h:380-386). */
qt_status qt_ctx_set_time_limit(qt_ctx* ctx, double seconds);
/* Blocks until every call issued on the context has finished. */
qt_status qt_sync(qt_ctx* ctx);
/* Message of the last error on the context ("" if none).  Owned by the context. */
const char* qt_last_error(const qt_ctx* ctx);
/* Device bytes held by the context's allocator (in use + cached). */
int64_t qt_mem_reserved(const qt_ctx* ctx);
/* Returns the allocator's cached (unused) blocks to the device after the context's
 * stream has drained; tensors in use keep their memory. */
qt_status qt_ctx_trim(qt_ctx* ctx);

/* ---- tensors (replace Node::mVals, src/Node.h:112-113,137,151,160) ---------------- */

/* Allocates a zero-filled tensor of D^rank elements, D = 1 << dim_log2 (Node ctor,
 * src/Node.h:112-113). */
qt_status qt_tensor_alloc(qt_ctx* ctx, int rank, int dim_log2, int dtype, qt_tensor** out);
/* Same without the zero fill: for contraction outputs, which are written in full. */
qt_status qt_tensor_alloc_empty(qt_ctx* ctx, int rank, int dim_log2, int dtype, qt_tensor** out);
/* Host -> device copy of interleaved (re, im) values; nbytes must equal the tensor size. */
qt_status qt_tensor_upload(qt_ctx* ctx, qt_tensor* t, const void* host, int64_t nbytes);
/* Device -> host copy (synchronises); replaces reading Node::GetTensorVals (src/Node.h:151). */
qt_status qt_tensor_download(qt_ctx* ctx, const qt_tensor* t, void* host, int64_t nbytes);
/* Creates n tensors at once from host values (leaf gate/state/measure tensors of a
 * network, src/Node.h:197-898): one device allocation and one asynchronous copy
 * from pinned staging instead of one of each per tensor.  host[i] holds
 * D^ranks[i] interleaved values and may be reused when the call returns. */
qt_status qt_tensor_create_many(qt_ctx* ctx, int n, const int* ranks, int dim_log2, int dtype,
                                const void* const* host, qt_tensor** out);
/* Returns the storage to the context's cache (Node::ClearNodeData, src/Node.h:137). */
qt_status qt_tensor_free(qt_ctx* ctx, qt_tensor* t);
int64_t qt_tensor_numel(const qt_tensor* t);
int qt_tensor_rank(const qt_tensor* t);

/* ---- the pairwise contraction (Network::ContractIndices, src/Network.h:876-971) ----
 *
 * C[c] = sum_s A[a(c,s)] * B[b(c,s)] over the k shared axes:
 *   pair_a[i], pair_b[i]  : toSumOn[i] (axis of A, axis of B), i < k      (src/Network.h:793-800)
 *   free_src[i], i < rank_c: toNotSumOn[i]: axis | QT_FROM_A for an axis of A, plain axis of B;
 *                           digit i of C's flat index                      (src/Network.h:902-907)
 * C must be a distinct, already allocated tensor of rank_c axes with the same
 * dtype and dim as A and B.  `precision` must be QT_PLAIN_C128 for C128 tensors
 * and one of the C64 forms (optionally | QT_RESCALE) for C64 tensors.  Large
 * GEMM-shaped steps run on MFMA: FP64 for complex128, FP32 for 4M / 3M, bf16 for
 * bf16x3.  On return *cmacs (if non-NULL) holds
 * D^(rank_c + k), the reference's complex multiply-add counter (src/Network.h:884-885).
 * The call is asynchronous on the context's stream. */
qt_status qt_contract_pair(qt_ctx* ctx, const qt_tensor* A, int rank_a, const qt_tensor* B,
                           int rank_b, int k, const int32_t* pair_a, const int32_t* pair_b,
                           int rank_c, const int32_t* free_src, qt_tensor* C, int precision,
                           int64_t* cmacs);

/* The same contraction without allocating: no hipMalloc, hipFree or other allocating
 * call.  It has no counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The caller passes A, B and an allocated C as for qt_contract_pair; the
 * step's scratch (GEMM operand permutes: A's and B's sizes, 256-byte aligned; split
 * reduction partial sums: at most 32 MiB) comes from a fixed 64 MiB device array per
 * GPU, held from the first qt_ctx_create on that GPU until the process ends.  A step
 * whose scratch exceeds 64 MiB returns QT_INVALID_INPUT before launching anything, as
 * does QT_RESCALE (its exponent table is allocated on first use: use
 * qt_contract_pair).  Every call on a GPU shares that GPU's array and nothing in it
 * needs to start at zero; calls that overlap on one GPU (from several contexts or
 * threads) race on it.  A GEMM-shaped step with fewer than 512 output tiles and at
 * least 512 summed indices does not split its summed range here (qt_contract_pair
 * does), so its result can differ from qt_contract_pair's in the last bits. */
qt_status qt_contract_pair_noalloc(qt_ctx* ctx, const qt_tensor* A, int rank_a, const qt_tensor* B,
                                   int rank_b, int k, const int32_t* pair_a, const int32_t* pair_b,
                                   int rank_c, const int32_t* free_src, qt_tensor* C, int precision,
                                   int64_t* cmacs);

/* Reads the rank-0 result of a finished contraction (the mFinalVal read of
 * src/Network.h:961-969 and src/LineGraph.h:388), with a rescaled tensor's exponent
 * applied in FP64; synchronises. */
qt_status qt_tensor_read_scalar(qt_ctx* ctx, const qt_tensor* t, double* re, double* im);
/* acc += t for rank-0 tensors, in FP64 on the stream (acc complex128; t's exponent
 * applied): sums of sliced partial amplitudes stay on the device. */
qt_status qt_tensor_accumulate(qt_ctx* ctx, const qt_tensor* t, qt_tensor* acc);
/* Power-of-two exponent of a rescaled complex64 tensor (0 if it never was):
 * its values are the downloaded data times 2^exponent.  Synchronises. */
qt_status qt_tensor_exponent(qt_ctx* ctx, const qt_tensor* t, int32_t* exponent);

/* ---- programs and batches (many networks per launch) ---------------------------------
 * A program is a network's whole sequence of ContractNodes steps (the calls a
 * qTorch ordering makes: LGContract src/LineGraph.h:357-376, ParallelContract
 * src/ContractionTools.h:219-388, ContractGivenSequence :392-411) compiled for the
 * device.  Operand ids: leaves 0..n_leaves-1, then the result of step i is id
 * n_leaves + i; every intermediate is consumed exactly once and the last step
 * yields the network's scalar (or, for a subtree of a larger network, a tensor
 * that a batch writes to per-instance output tensors, qt_batch_set_outputs).
 *   step_desc : 6 ints per step: a, b, rank_a, rank_b, k, rank_c
 *   step_axes : per step pair_a[k], pair_b[k], free_src[rank_c] (the same conventions
 *               as qt_contract_pair), concatenated
 *   leaf_refs : element offset of each leaf in the batch's leaf pool; with bit 62 set
 *               the offset is relative to the instance's parameter base
 * Tensors of a program are limited to 2^16 elements (D = 4: rank 8; D = 2: rank 16).
 * Compilation runs on the host only; ctx may be NULL (errors are then only returned). */
#define QT_LEAF_PARAM (1LL << 62)
qt_status qt_program_create(qt_ctx* ctx, int dim_log2, int n_leaves, const int32_t* leaf_ranks,
                            const int64_t* leaf_refs, int n_steps, const int32_t* step_desc,
                            const int32_t* step_axes, qt_program** out);
/* complex MACs (src/Network.h:884-885 counter), scratch elements and step count */
qt_status qt_program_stats(const qt_program* p, int64_t* cmacs, int64_t* scratch_elems, int* n_steps);
qt_status qt_program_destroy(qt_program* p);

typedef struct qt_batch qt_batch;
/* A batch holds programs on the device in one arithmetic form (qt_precision);
 * n_programs may be 0 (programs added later).  Constant leaf references of these
 * programs are absolute pool offsets. */
qt_status qt_batch_create(qt_ctx* ctx, int n_programs, qt_program* const* progs, int precision,
                          qt_batch** out);
/* Appends programs (program ids first_id, first_id + 1, ...); the constant leaf
 * references of program i are offsets from const_base[i] in the pool.  Waits for
 * the batch's work in flight. */
qt_status qt_batch_add_programs(qt_batch* b, int n, qt_program* const* progs, const int64_t* const_base,
                                int32_t* first_id);
/* Leaf pool: interleaved complex FP64 values (converted on the device for C64 forms). */
qt_status qt_batch_set_leaves(qt_batch* b, const double* pool, int64_t n_elems);
/* Writes n_elems values at pool offset `offset`, growing the pool (contents kept) as
 * needed; waits for the batch's work in flight. */
qt_status qt_batch_write_leaves(qt_batch* b, int64_t offset, const double* values, int64_t n_elems);
/* Instances: program index and parameter base of each; optional FP64 group sums
 * sum_{i in g} (c0 + c1 * Re value_i) (the per-graph energy 0.5 * (1 - <ZZ>) of
 * src/maxcut.cpp:196 with c0 = 0.5, c1 = -0.5); group -1 puts an instance in no group.
 * Waits for the batch's work in flight. */
qt_status qt_batch_set_instances(qt_batch* b, int n, const int32_t* prog, const int64_t* param_base,
                                 int n_groups, const int32_t* group, double c0, double c1);
/* Output tensors of the current instances (one per instance, in instance order, set
 * again after every qt_batch_set_instances): each instance's last step writes its
 * whole result (tensor or scalar) there, in the batch's dtype, in the step's axis
 * order (its free_src), instead of the FP64 value array of qt_batch_results.  Needed
 * when a program ends in a tensor; such batches run with 64 or 256 lanes per
 * network.  The tensors are written on the context's stream by qt_batch_launch and
 * can feed qt_contract_pair calls queued after it.  No counterpart in the
 * reference: it lets the small subtrees of a large network run as one launch. */
qt_status qt_batch_set_outputs(qt_batch* b, int n, qt_tensor* const* outs);
/* Executor shape: lanes per network (4, 8, 16 or 32 = that many lanes of a
 * wavefront per network, several networks of one program side by side; 64 = one
 * wavefront; 256 = one workgroup) and workgroups per compute unit (workgroups are
 * 64 threads for fewer than 64 lanes, else 256).  Intermediates live in LDS when
 * they fit, else in a device scratch area. */
qt_status qt_batch_config(qt_batch* b, int lanes, int wgs_per_cu);
enum { QT_SCRATCH_AUTO = 0, QT_SCRATCH_LDS = 1, QT_SCRATCH_GLOBAL = 2, QT_SCRATCH_STAGE_LEAVES = 4 };
/* The same with an explicit place for the intermediates (QT_INVALID_INPUT, and the
 * previous shape kept, when LDS is asked for and a workgroup's share exceeds 64 KiB);
 * QT_SCRATCH_STAGE_LEAVES or-ed in makes lane-grouped waves copy each step's leaf
 * operands to LDS first (else they are read from the pool through the caches). */
qt_status qt_batch_config_ex(qt_batch* b, int lanes, int wgs_per_cu, int scratch);
/* Times every executor shape, scratch place and leaf staging choice on the current
 * instances and keeps the fastest. */
qt_status qt_batch_autotune(qt_batch* b, int* lanes, int* wgs_per_cu, float* ms);
/* Asynchronous run of every instance (replays the captured graph if there is one). */
qt_status qt_batch_launch(qt_batch* b);
/* *done = 1 once the last launch has finished (0 while it runs); never blocks. */
/* Group sums only (no reference counterpart): launches of a compiled batch whose
 * groups are strided (qt_batch_set_instances with n_groups and member j of group g =
 * instance j n_groups + g), each member running one program in every group, compute
 * the group sums without writing per-instance values: a generated kernel runs, per
 * wavefront, 64 groups through the members j = s, s + 16, ... of one slice s and
 * adds c0 + c1 Re(value) in that order, the order of the ordinary group sums, so the
 * sums are bit-identical to theirs.  qt_batch_results then accepts values == NULL
 * only.  *active (may be NULL) = 1 when launches currently take this path (compiled,
 * with instances that fit); otherwise launches run as before.  on = 0 restores
 * per-instance values. */
/* Device bytes a batch holds for its programs and leaf pool (device arrays of the
 * program store and the pool, allocated capacity): what stays resident between
 * launches as programs are added. */
int64_t qt_batch_resident_bytes(const qt_batch* b);
qt_status qt_batch_set_sums_only(qt_batch* b, int on, int* active);
qt_status qt_batch_poll(qt_batch* b, int* done);

/* ---- rotation gates whose angles change between launches ------------------------- */

/* Gate kinds of qt_batch_rotation_params (Rz also stands for PHASE: same tables). */
typedef enum { QT_ROT_RX = 0, QT_ROT_RY = 1, QT_ROT_RZ = 2 } qt_rotation_kind;

/* One rotation gate of an angle set's parameter region. */
typedef struct qt_rotation {
    int32_t kind;   /* qt_rotation_kind */
    int32_t angle;  /* index of the gate's angle in a set's row of angles */
    double scale;   /* the gate's angle is scale * row[angle] (-1 for Rz(-gamma), 2 for Rx(2 beta)) */
    int64_t offset; /* element offset of the gate's block in the set's parameter region */
} qt_rotation;

/* Writes the blocks of rotation gates whose angles come from an array, on the device:
 * for every set a < n_sets and gate r < n_rot, the block of gate r with angle
 * scale * angles[a * n_angles + angle] at pool element base + a * set_stride +
 * rots[r].offset (16 elements: the superoperator over (in, out), dim_log2 = 2; or 4: the
 * amplitude gate, dim_log2 = 1).  The values are those of a network built from the
 * gate's text: the angle printed at the default ostream precision and parsed with
 * std::stof (src/maxcut.h:211,217; src/Network.h:342), then the gate table in FP64
 * (src/Node.h:197-898) -- the round trip is exact, the cos / sin values are the
 * device's (within a few units in the last place of the host's).  It replaces
 * rebuilding every light-cone network from text for each evaluation of new angles.
 * angles: n_sets x n_angles doubles in device memory (angles_on_host = 0: read in
 * stream order on the context's stream; writes to them from other streams must have
 * completed) or in host memory (angles_on_host = 1: copied first).  Asynchronous.
 * Blocks must lie inside the pool and not overlap (set_stride at least the extent
 * of a set's blocks); at most 32 gates per set.  An angle that is not 0 and whose
 * magnitude lies outside [1e-16, 1e15), or that is not finite, gives NaN values and
 * makes the next qt_batch_results return QT_INVALID_INPUT.  A specialised executor
 * whose literals the new blocks could change is dropped (the table executor runs). */
qt_status qt_batch_rotation_params(qt_batch* b, int dim_log2, int n_rot, const qt_rotation* rots, int n_sets,
                                   int n_angles, const double* angles, int angles_on_host, int64_t base,
                                   int64_t set_stride);
/* The float that the text round trip above gives x (the host instantiation of the
 * routine the device runs): QT_INVALID_INPUT outside its range. */
qt_status qt_angle_text_float(double x, float* out);
/* Captures the launch sequence (counter reset, persistent kernel, group sums) into a
 * HIP graph; later launches replay it. */
qt_status qt_batch_capture(qt_batch* b);
/* Drops the captured graph: launches are direct again. */
qt_status qt_batch_uncapture(qt_batch* b);
/* Synchronises and copies per-instance values (2 doubles each) and group sums. */
qt_status qt_batch_results(qt_batch* b, double* values, double* group_sums);
/* A specialised batch with group sums of the edge-major layout (member j of group g =
 * instance j * n_groups + g) launched directly leaves each launch's sums to its next
 * launch's kernel (the results alternate between two buffers); this launches the last
 * launch's sums now, asynchronously.  qt_batch_results, qt_batch_poll, qt_batch_time
 * and qt_batch_capture do it themselves; a caller that waits on the stream alone
 * calls it first.  No counterpart in the reference. */
qt_status qt_batch_flush(qt_batch* b);
/* Mean time of `reps` launches, by HIP events on the context's stream. */
qt_status qt_batch_time(qt_batch* b, int reps, float* ms_per_launch);
/* The same for the contraction kernel alone, the dominant kernel's time per launch:
 * for the table executors without the group sums (their work-counter reset, a
 * one-thread kernel, is included); for a specialised kernel with deferred group sums
 * as it runs in a stream of direct launches (each summing the previous launch's
 * results in its first workgroups). */
qt_status qt_batch_time_kernel(qt_batch* b, int reps, float* ms_per_launch);
/* complex MACs of one launch, executor shape, flags (1: intermediates in device
 * memory, 2: leaves staged in LDS) and scratch elements per executor */
qt_status qt_batch_info(const qt_batch* b, int64_t* cmacs, int* lanes, int* wgs_per_cu,
                        int* global_scratch, int* scratch_elems);
qt_status qt_batch_destroy(qt_batch* b);

/* ---- specialised executors -------------------------------------------------------------
 * No counterpart in the reference (its ContractIndices walks index vectors per MAC,
 * src/Network.h:876-971): for a batch whose programs stay fixed while instances
 * change only their parameter blocks, qt_batch_compile generates one kernel in which
 * every program is straight-line code (each output element an explicit FMA chain in
 * the order of the table executors, constant leaves as literals with exact-zero
 * terms dropped, intermediates in registers, one lane per instance), compiles it for
 * the device at run time (hipRTC) and makes qt_batch_launch run it.  Only the
 * complex64 4M and complex128 forms, only programs that end in scalars, and at most
 * 250,000 FMAs of generated code in all (QT_INVALID_INPUT otherwise; the table
 * executor keeps running).  The constants are read from the pool now: a later pool
 * write that changes them, new programs, or an explicit qt_batch_config /
 * qt_batch_config_ex drop the generated kernel (the table executor runs again).
 * flags: 0, or QT_SPECIALISE_HERMITIAN for networks in qTorch's density-matrix picture
 * (two index bits per axis, d = 2 ket + bra) whose every leaf satisfies
 * T(x') = conj T(x), x' being x with ket and bra swapped on every axis (gate
 * superoperators U (x) U*, Hermitian measurement vectors): every intermediate then has
 * the same symmetry, so of each pair x, x' only one element is computed and
 * self-mirror elements are real (about half the FMAs).  The constant leaves and the
 * current instances' parameter blocks are checked (QT_INVALID_INPUT if one is not
 * symmetric); parameter blocks written later must keep the symmetry.  Results then
 * differ from the table executors' by rounding only.  *compile_ms (if non-NULL):
 * generation, compilation and loading time. */
enum { QT_SPECIALISE_HERMITIAN = 1 };
qt_status qt_batch_compile(qt_batch* b, int flags, float* compile_ms);
/* *active = 1 while the generated kernel is the one qt_batch_launch runs;
 * *fmas_per_launch: FMAs its code executes for the current instances (each a real
 * multiply-add of 2 FLOPs; terms with exact-zero constant factors and values that
 * never reach the result are not executed) */
qt_status qt_batch_specialised_info(const qt_batch* b, int* active, int64_t* fmas_per_launch);
/* Host only: the HIP source qt_batch_compile would generate for these programs, with
 * program i's constant leaf references offset by const_base[i] into `pool`
 * (interleaved FP64, pool_elems elements).  Writes at most cap bytes (NUL-terminated)
 * to buf (may be NULL) and the full length to *len. */
qt_status qt_program_source(int n, qt_program* const* progs, const int64_t* const_base, const double* pool,
                            int64_t pool_elems, int precision, int flags, char* buf, size_t cap, size_t* len);
/* Host only: compiles such source for an architecture ("gfx950", or a device's full
 * gcnArchName) with the options qt_batch_compile uses (-O3, packed FP32 instructions
 * off); *len receives the code object's size, which is copied to code when cap is
 * large enough (QT_DEVICE if compilation fails). */
qt_status qt_program_compile(const char* source, const char* arch, void* code, size_t cap, size_t* len);

#ifdef __cplusplus
}
#endif

#endif /* QTORCH_HIP_H */
