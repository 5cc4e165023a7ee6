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

/* Network(qasm, measure) (src/Network.h:148-151, 246-304) from in-memory text;
 * measure may be NULL (every qubit traced); base_dir resolves def1/def2 paths. */
int qth_network_create(qt_ctx* ctx, const char* qasm, const char* measure, const char* base_dir,
                       qth_network** out);
/* Same, with the wire dimension: dim_log2 = 2 is qTorch's Liouville picture, 1 the
 * amplitude picture (gates as unitaries on kets, the measurement string is the
 * output bitstring x and the value is <x|U|0...0>). */
int qth_network_create_ex(qt_ctx* ctx, const char* qasm, const char* measure, const char* base_dir,
                          int dim_log2, qth_network** out);
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
int qth_counts(qth_network* net, int* all_nodes, int* uncontracted, int* original, int* done);
/* Recorded plan steps: 8 ints each (a_id, b_id, c_id, rank_a, rank_b, k, rank_c, 0) */
int qth_plan(qth_network* net, int32_t* steps, int cap, int* n);

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
 *        2 = the reference's stochastic order (seeded), 3 = line-graph min-fill order. */
enum { QTH_ORDER_GIVEN = 0, QTH_ORDER_GREEDY = 1, QTH_ORDER_STOCHASTIC = 2, QTH_ORDER_LINEGRAPH = 3 };

/* F_p (src/maxcut.cpp:162-204) for many angle sets per launch: one program per edge
 * light cone; instances = angle sets x edges; FP64 energy per angle set. */
typedef struct qth_qaoa_batch qth_qaoa_batch;
int qth_qaoa_batch_create(qt_ctx* ctx, const char* dgf, int p, int ref_map, int precision, int order,
                          uint64_t seed, qth_qaoa_batch** out);
/* bg: n_sets x 2p doubles, each set [betas..., gammas...] */
int qth_qaoa_batch_set_angles(qth_qaoa_batch* b, int n_sets, const double* bg);
/* energies: n_sets; zz: n_sets x n_edges values of <Z Z> (either may be NULL); synchronises */
int qth_qaoa_batch_results(qth_qaoa_batch* b, double* energies, double* zz);
qt_batch* qth_qaoa_batch_device(qth_qaoa_batch* b);
int64_t qth_qaoa_batch_cmacs(qth_qaoa_batch* b); /* complex MACs of one energy evaluation */
int qth_qaoa_batch_destroy(qth_qaoa_batch* b);

/* The QAOA outer loop (src/maxcut.cpp:145-222): maximises F_p over 2p angles with
 * n_random seeded random starts evaluated in one batch, then a batched Nelder-Mead
 * simplex (the reference uses nlopt COBYLA).  angles receives 2p doubles. */
int qth_qaoa_optimize(qt_ctx* ctx, const char* dgf, int p, int precision, int n_random, int max_evals,
                      uint64_t seed, double* angles, double* energy, int* evaluations);
/* maxcutGetFinalString (src/maxcut.cpp:29-140): greedy bit-by-bit cut from full-circuit
 * projector probabilities; bits receives one int per vertex. */
int qth_maxcut_final_string(qt_ctx* ctx, const char* dgf, int p, const double* betas_gammas, uint64_t seed,
                            int* bits, int* cut_edges);

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

#ifdef __cplusplus
}
#endif

#endif /* QTORCH_HOST_H */
