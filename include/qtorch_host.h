/*
 * qtorch_host.h — C ABI of libqtorch_host.so, the qTorch host surface (QASM parsing
 * into Network/Node/Wire, the orderings, ReduceCircuit, LineGraph, the maxcut
 * light cones) over libqtorch_hip.so.  It is the binding a non-C++ caller (ctypes,
 * cgo, JNI) uses; C++ callers can use qtorch_amd/csrc/host/qtorch.hpp directly.
 *
 * Every function returns a qth status code and never throws; qth_last_error()
 * gives the message of the calling thread's last failure.  Networks created
 * with ctx == NULL are plan-only: the ordering logic runs and every step is
 * recorded, but no values are computed (the value outputs are NaN).
 */
#ifndef QTORCH_HOST_H
#define QTORCH_HOST_H

#include <stdint.h>

#include "qtorch_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: one per reference exception type (src/Exceptions.h:24-59) */
enum {
    QTH_OK = 0,
    QTH_INVALID_FUNCTION_INPUT = 1, /* InvalidFunctionInput */
    QTH_INVALID_FILE = 2,           /* InvalidFile */
    QTH_INVALID_FILE_FORMAT = 3,    /* InvalidFileFormat (and std::stoi/stof failures) */
    QTH_CONTRACTION_FAILURE = 4,    /* ContractionFailure */
    QTH_INVALID_USER_SEQUENCE = 5,  /* InvalidUserContractionSequence */
    QTH_INVALID_METHOD = 6,         /* InvalidContractionMethod */
    QTH_QBB_FAILURE = 7,            /* QbbFailure */
    QTH_DEVICE = 8,                 /* engine/device error */
    QTH_OTHER = 9
};

typedef struct qth_network qth_network;

const char* qth_last_error(void);

/* The reference's process-wide totTimer / maxTime (src/Network.h:50-51, src/Timer.h):
 * once the timer is started and maxTime seconds have passed, contractions stop
 * before their next pairwise step and return QTH_CONTRACTION_FAILURE
 * (src/Network.h:899-910, src/ContractionTools.h:271,332,380-386).  maxTime
 * starts at 60; the timer is only started by qth_timer_start and qth_preprocess. */
int qth_set_max_time(double seconds);
double qth_get_max_time(void);
int qth_timer_start(void);
int qth_timer_reset(void);
double qth_timer_elapsed(void);
/* preProcess (src/preprocess.h:28-52): up to 100 seeded stochastic contractions
 * (attempt i with seed + i), each timed from a fresh timer start with maxTime =
 * time_threshold; *found = 1 and the first in-time contraction's mCreatedFrom
 * record (2 ints per step, *n steps; up to cap written) when one finishes within
 * time_threshold seconds, else *found = 0.  ctx NULL: plan-only contractions.
 * maxTime keeps time_threshold afterwards, as in the reference. */
int qth_preprocess(qt_ctx* ctx, const char* qasm, const char* measure, double time_threshold, uint64_t seed,
                   int32_t* pairs, int cap, int* n, int* found);

/* Network(qasm, measure) (src/Network.h:148-151, 246-304) from in-memory text;
 * measure may be NULL (every qubit traced); base_dir resolves def1/def2 paths. */
int qth_network_create(qt_ctx* ctx, const char* qasm, const char* measure, const char* base_dir,
                       qth_network** out);
/* Same, with the wire dimension: dim_log2 = 2 is qTorch's Liouville picture, 1 the
 * amplitude picture (gates as unitaries on kets, the measurement string is the
 * output bitstring x and the value is <x|U|0...0>). */
int qth_network_create_ex(qt_ctx* ctx, const char* qasm, const char* measure, const char* base_dir,
                          int dim_log2, qth_network** out);
/* Same, with the output wires of open_qubits[0..n_open) left open (no measurement
 * or trace node; the reference has no counterpart, where every qubit ends in one, and
 * gets a probability vector from one measured network per output string,
 * src/Network.h:199-240): the contraction ends in a tensor over those wires instead
 * of a scalar (qth_network_result_tensor).  Contract it with QTH_ORDER_GREEDY
 * (qth_network_order) or a given sequence. */
int qth_network_create_open(qt_ctx* ctx, const char* qasm, const char* measure, const char* base_dir, int dim_log2,
                            const int32_t* open_qubits, int n_open, qth_network** out);
/* The tensor a finished network with open qubits ends in: D^rank interleaved values
 * (up to cap_elems written) and the qubit of each axis (axis_qubits, rank entries;
 * digit i of the flat index belongs to axis_qubits[i]).  For D = 4 the digit of a
 * qubit is 2 ket + bra of the output density matrix, so P(x) is the element whose
 * digits are 3 x_q.  A network without open qubits gives its scalar (rank 0). */
int qth_network_result_tensor(qth_network* net, double* values, int64_t cap_elems, int32_t* axis_qubits, int* rank);
int qth_network_destroy(qth_network* net);
/* Network::ReduceCircuit (src/Network.h:1015-1118) */
int qth_network_reduce(qth_network* net);
/* Network::ContractNodes on node ids (src/Network.h:715-864); *result_id = -1 when
 * the pair was rejected (nullptr in the reference) or the network finished. */
int qth_contract_nodes(qth_network* net, int a_id, int b_id, int threshold, int* result_id);
/* ContractionTools::Contract(Stochastic) with a seed (src/ContractionTools.h:219-388) */
int qth_contract_stochastic(qth_network* net, uint64_t seed, double* re, double* im);
/* ContractionTools::ContractGivenSequence (src/ContractionTools.h:392-411); pairs = 2n ids */
int qth_contract_given(qth_network* net, const int32_t* pairs, int n, double* re, double* im);
/* ContractUserDefinedSequenceOfWires (src/ContractionTools.h:154-214) */
int qth_contract_user(qth_network* net, const int32_t* pairs, int n, double* re, double* im);
/* LineGraph (src/LineGraph.h:92-171) of the network's current uncontracted nodes:
 * the lg.cnf text, the min-fill order (1-based wire ids) and LGContract in a given order. */
int qth_linegraph_cnf(qth_network* net, char* buf, int64_t cap, int64_t* len);
int qth_linegraph_minfill(qth_network* net, int32_t* order, int cap, int* n);
int qth_linegraph_contract(qth_network* net, const int32_t* order, int n, double* re, double* im);
/* Contracts the network in one of the QTH_ORDER_* orders (below); on a plan-only
 * network this only records the plan. */
int qth_network_order(qth_network* net, int order, uint64_t seed);
/* mCreatedFrom record (src/preprocess.h:41-46): 2 ints per contraction */
int qth_record(qth_network* net, int32_t* pairs, int cap, int* n);
/* Node::mNumFloatOps counter (src/Network.h:884-885): complex MACs */
int64_t qth_flops(qth_network* net);
/* mFinalVal (src/Network.h:61) of a contracted network */
int qth_network_value(qth_network* net, double* re, double* im);
int qth_counts(qth_network* net, int* all_nodes, int* uncontracted, int* original, int* done);
/* Recorded plan steps: 8 ints each (a_id, b_id, c_id, rank_a, rank_b, k, rank_c, 0) */
int qth_plan(qth_network* net, int32_t* steps, int cap, int* n);
/* Compiles the recorded plan into a device program (no device needed) and reports
 * its complex-MAC count, the scratch elements one executor needs (steps scheduled
 * to keep the fewest live) and the step count. */
int qth_program_stats(qth_network* net, int64_t* cmacs, int64_t* scratch_elems, int* n_steps);

/* QAOA MaxCut (src/maxcut.h:31-233): per-edge light-cone texts of F_p for a .dgf
 * graph.  ref_map != 0 keeps the reference's light-cone qubit map. */
int qth_qaoa_num_edges(const char* dgf, int* n_edges, int* n_vertices);
int qth_qaoa_edge_text(const char* dgf, int p, int ref_map, int e, const double* betas_gammas,
                       char* qasm, int64_t qasm_cap, char* measure, int64_t measure_cap);
/* F_p (src/maxcut.cpp:162-204) on the device, one stochastically ordered network per
 * edge (seed + e); zz receives <Z Z> per edge (may be NULL). */
int qth_qaoa_energy_plain(qt_ctx* ctx, const char* dgf, int p, int ref_map, const double* betas_gammas,
                          uint64_t seed, double* zz, double* energy);

/* ---- batched evaluation (device programs, include/qtorch_hip.h qt_batch_*) -------------
 * order: 0 = given sequences (ContractGivenSequence), 1 = greedy cost order,
 *        2 = the reference's stochastic order (seeded), 3 = line-graph min-fill order,
 *        4 = circuit sweep (no reference counterpart): local merges, then one growing
 *            tensor absorbing the rest along a line through the qubits (Fiedler order
 *            of their interaction graph); for a sliced amplitude network it also
 *            splits diagonal two-qubit gates (CZ, CPHASE, CRk) into two rank-3
 *            tensors joined by a bond. */
enum { QTH_ORDER_GIVEN = 0, QTH_ORDER_GREEDY = 1, QTH_ORDER_STOCHASTIC = 2, QTH_ORDER_LINEGRAPH = 3,
       QTH_ORDER_SWEEP = 4 };

/* F_p (src/maxcut.cpp:162-204) for many angle sets per launch: one program per edge
 * light cone; instances = angle sets x edges; FP64 energy per angle set. */
typedef struct qth_qaoa_batch qth_qaoa_batch;
int qth_qaoa_batch_create(qt_ctx* ctx, const char* dgf, int p, int ref_map, int precision, int order,
                          uint64_t seed, qth_qaoa_batch** out);
/* bg: n_sets x 2p doubles, each set [betas..., gammas...] */
int qth_qaoa_batch_set_angles(qth_qaoa_batch* b, int n_sets, const double* bg);
/* New angles for the same number of sets, without rebuilding anything on the host:
 * the Rz(-gamma) / Rx(2 beta) blocks are written on the device from bg
 * (qt_batch_rotation_params), with the values the gates' text would give.  bg:
 * n_sets x 2p doubles as above, in host memory (on_host = 1) or device memory
 * (on_host = 0; read in order on the context's stream).  Asynchronous.  It replaces
 * the reference's rebuild of every edge network from text for each evaluation
 * (src/maxcut.cpp:162-204, src/maxcut.h:194-233). */
int qth_qaoa_batch_update_angles(qth_qaoa_batch* b, int n_sets, const double* bg, int on_host);
/* energies: n_sets; zz: n_sets x n_edges values of <Z Z> (either may be NULL); synchronises */
int qth_qaoa_batch_results(qth_qaoa_batch* b, double* energies, double* zz);
qt_batch* qth_qaoa_batch_device(qth_qaoa_batch* b);
int64_t qth_qaoa_batch_cmacs(qth_qaoa_batch* b); /* complex MACs of one energy evaluation */
int qth_qaoa_batch_destroy(qth_qaoa_batch* b);
/* Host only: the specialised-executor source (qt_batch_compile) of the edge programs
 * first_edge .. first_edge + n_edges - 1, as program 0, 1, ... (precision
 * QT_C64_4M or QT_PLAIN_C128; flags as qt_batch_compile's); *len receives the full
 * length. */
int qth_qaoa_source(const char* dgf, int p, int ref_map, int order, uint64_t seed, int precision, int flags,
                    int first_edge, int n_edges, char* buf, size_t cap, size_t* len);
/* The 32 p complex parameter elements (64 p doubles, interleaved) of one angle set
 * [betas..., gammas...]: Rz(-gamma_i) at element 32 i, Rx(2 beta_i) at 32 i + 16. */
int qth_qaoa_param_block(int p, const double* betas_gammas, double* out);

/* The QAOA outer loop (src/maxcut.cpp:145-222): maximises F_p over 2p angles with
 * n_random seeded random starts evaluated in one batch, then a batched Nelder-Mead
 * simplex (the reference uses nlopt COBYLA).  angles receives 2p doubles. */
int qth_qaoa_optimize(qt_ctx* ctx, const char* dgf, int p, int precision, int n_random, int max_evals,
                      uint64_t seed, double* angles, double* energy, int* evaluations);
/* maxcutGetFinalString (src/maxcut.cpp:29-140): greedy bit-by-bit cut from full-circuit
 * projector probabilities; bits receives one int per vertex. */
int qth_maxcut_final_string(qt_ctx* ctx, const char* dgf, int p, const double* betas_gammas, uint64_t seed,
                            int* bits, int* cut_edges);
/* Same, also writing the conditional probability P(bit i = 0 | the bits chosen
 * before it) that decided each bit (cond_probs: one double per vertex; may be NULL). */
int qth_maxcut_final_string_ex(qt_ctx* ctx, const char* dgf, int p, const double* betas_gammas, uint64_t seed,
                               int* bits, int* cut_edges, double* cond_probs);

/* Any list of networks (QASM + measurement text) in one batch.  given / given_len:
 * concatenated node-id pairs per network for QTH_ORDER_GIVEN (else NULL). */
typedef struct qth_netbatch qth_netbatch;
int qth_netbatch_create(qt_ctx* ctx, int n, const char* const* qasm, const char* const* measure,
                        const char* base_dir, int dim_log2, int order, uint64_t seed, const int32_t* given,
                        const int32_t* given_len, int precision, qth_netbatch** out);
qt_batch* qth_netbatch_device(qth_netbatch* b);
int qth_netbatch_results(qth_netbatch* b, double* values); /* 2 doubles per network; synchronises */
int qth_netbatch_cmacs(qth_netbatch* b, int64_t* cmacs);   /* complex MACs per network */
int qth_netbatch_destroy(qth_netbatch* b);

/* ---- serving: continuous batching of QAOA-energy and amplitude requests ----------
 * (no counterpart in the reference, which evaluates one network at a time)
 * Requests are submitted from any thread and finish asynchronously.  Each request's
 * network is planned once per structure and then runs as a program instance in the
 * next launch of an idle stream (or, when its tensors exceed a program's limit, on
 * the large-network path, complex128 by default: sliced, small subtrees fused, MFMA
 * GEMMs),
 * admitted while the planned device bytes of the
 * requests in flight fit hbm_budget.  Results do not depend on what else shares a
 * launch. */
typedef struct qth_server qth_server;
enum { QTH_SERVE_QAOA_ENERGY = 0, QTH_SERVE_AMPLITUDE = 1 };
enum { QTH_ROUTE_SMALL = 0, QTH_ROUTE_MEDIUM = 1, QTH_ROUTE_EAGER = 2 };
int qth_server_create(int device, int slots, int precision, int64_t hbm_budget, int max_requests_per_launch,
                      qth_server** out);
/* The same with the large-network path's arithmetic form (complex128 in
 * qth_server_create; a complex64 form halves its bytes and runs its GEMMs on FP32
 * MFMA, with that form's error). */
int qth_server_create_ex(int device, int slots, int precision, int eager_precision, int64_t hbm_budget,
                         int max_requests_per_launch, qth_server** out);
int qth_server_destroy(qth_server* s); /* finishes the submitted requests first */
/* QAOA energy F_p of a .dgf graph for one angle set [betas..., gammas...] (2p doubles). */
int qth_server_submit_qaoa(qth_server* s, const char* dgf, int p, int ref_map, const double* betas_gammas,
                           int64_t* id);
/* <bits|C|0...0> of a QASM circuit (amplitude picture), bits = one '0'/'1' per qubit. */
int qth_server_submit_amplitude(qth_server* s, const char* qasm, const char* bits, int64_t* id);
/* Non-blocking: *done = 1 and the result once request `id` has finished (it is then
 * forgotten).  status: the request's QTH_* code; value: energy (re) or amplitude. */
int qth_server_poll(qth_server* s, int64_t id, int* done, int* status, double* re, double* im, double* latency_ms,
                    int* route);
/* Blocking variant; returns the request's QTH_* code. */
int qth_server_wait(qth_server* s, int64_t id, double* re, double* im, double* latency_ms, int* route);
int qth_server_drain(qth_server* s);
/* counts: submitted, completed, rejected, launches, QAOA edge networks evaluated;
 * ms: p50, p99, mean latency, time with two or more launches in flight, time with
 * any launch in flight; peak: largest planned device bytes admitted at once. */
int qth_server_stats(qth_server* s, int64_t* counts, double* ms, int64_t* peak);
/* Device bytes of the programs and constants resident in the slots' batches (counted
 * against hbm_budget with the planned bytes of the requests in flight), and how many
 * times a slot dropped its programs because they exceeded a quarter of the budget per
 * slot (structures are added again on their next use). */
int qth_server_memory(qth_server* s, int64_t* resident_bytes, int64_t* slot_resets);
/* Route and planned device bytes of a request (planning only; no device needed).
 * kind QTH_SERVE_QAOA_ENERGY: text = .dgf, p, ref_map; QTH_SERVE_AMPLITUDE: text =
 * QASM, bits = the output bitstring. */
int qth_serve_plan(int kind, const char* text, int p, int ref_map, const char* bits, int precision,
                   int* route, int64_t* planned_bytes, int64_t* cmacs);

/* ---- one large network, sliced --------------------------------------------------
 * (no counterpart in the reference, whose ContractionTools contract one unsliced network)
 * Planned once in `order` (QTH_ORDER_*), then wires are sliced until every tensor
 * has at most max_bits index bits; the D^s slices run step by step (MFMA GEMMs for
 * large steps) in the given form (a complex64 form may be or-ed with QT_RESCALE for
 * per-step power-of-two rescaling) and their scalars are summed in FP64 on the device.
 * dim_log2 = 1: amplitude picture (measure = output bitstring); 2: qTorch's picture. */
typedef struct qth_sliced qth_sliced;
int qth_sliced_create(const char* qasm, const char* measure, int dim_log2, int order, uint64_t seed, int max_bits,
                      qth_sliced** out);
/* plan: sliced wires, peak bits, slices, complex MACs per slice, of those in
 * GEMM-shaped steps, peak bytes of a slice (complex128), fused subtrees, steps in
 * them (small steps whose subtrees run as device programs, one batch launch per
 * chunk of slices) */
int qth_sliced_plan(qth_sliced* s, int64_t* values8);
/* slices_per_launch = 0: every step of a slice runs through qt_contract_pair (no
 * fused subtrees); < 0 (default): fused, slices per launch bounded by memory;
 * > 0: fused, at most that many slices per launch. */
int qth_sliced_set_fusion(qth_sliced* s, int64_t slices_per_launch);
/* Later runs contract only slices [first, first + count) (count < 0: to the last
 * slice) and return their sum: a bounded sample of a network whose full slice count
 * is out of reach.  No reference counterpart (the reference does not slice). */
int qth_sliced_set_range(qth_sliced* s, int64_t first, int64_t count);
/* Later runs contract exactly slices[0..n) (each in [0, slices)) and return their sum. */
int qth_sliced_set_slices(qth_sliced* s, const int64_t* slices, int64_t n);
/* Up to max_count slices, in increasing order, in which no leaf's sub-tensor is all
 * zero (the others contribute exactly zero: diagonal gates tie the wires around
 * them); written to out[0..*n_out).  Host only. */
int qth_sliced_nonzero_slices(qth_sliced* s, int64_t max_count, int64_t* out, int64_t* n_out);
/* The same with the search budget (partial assignments of the sliced wires visited;
 * qth_sliced_nonzero_slices uses 2^24 and returns QTH_OTHER, with what it found
 * written, when the budget runs out first): *complete = 1 when the list is full
 * or every slice was visited, 0 when the budget ran out (the list may then miss
 * nonzero slices). */
int qth_sliced_nonzero_slices_ex(qth_sliced* s, int64_t max_count, int64_t budget, int64_t* out, int64_t* n_out,
                                 int* complete);
int qth_sliced_run(qth_sliced* s, qt_ctx* ctx, int precision, double* re, double* im, double* ms);
int qth_sliced_destroy(qth_sliced* s);

#ifdef __cplusplus
}
#endif

#endif /* QTORCH_HOST_H */
