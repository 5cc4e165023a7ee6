/*
 * ORACLE — test infrastructure only.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this; the product never does.
 *
 * Plain-C restatement of qTorch's pairwise contraction arithmetic,
 * Network::ContractIndices (reference src/Network.h:876-971), loop nest
 * src/Network.h:899-933:
 *
 *   for every output index c (flat, little-endian base-D digits):
 *       digit i of c -> free entry i (toNotSumOn[i]); entry.first ? A-axis : B-axis   (:902-907)
 *       for every summed index s < D^k:
 *           digit i of s -> pair toSumOn[k-1-i]  (the LAST pair varies fastest)      (:912-916)
 *           C[c] += A[a(c,s)] * B[b(c,s)]     (std::complex<double> multiply-add) (:919-931)
 *
 * Tensors are interleaved (re, im) doubles, flat index sum_i d_i * D^i over the
 * node's wire order (src/Node.h:168-186).  The reference has D = 4 (Liouville
 * wires); D = 2 (amplitude networks) is this build's generalisation and uses the
 * identical loop with base-2 digits.
 *
 * The complex product is evaluated as (ar*br - ai*bi, ar*bi + ai*br) with
 * floating-point contraction disabled (-ffp-contract=off), which is what the
 * reference's unoptimised x86-64 build computes for finite operands, so for the
 * same summation order the result is bit-identical to the reference.
 *
 * Parity pinned by: tests/test_oracle.py (bit-exact replay against oracle/_ref,
 * the reference compiled from its sources; numpy.einsum on random tensors; the
 * reference's known answers).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Returns 0 on success, -1 on invalid arguments (InvalidFunctionInput,
 * src/Network.h:938-940). */
int oracle_contract_indices(const double* A, int rA, const double* B, int rB,
                            int nfree, const int* free_from_a, const int* free_axis,
                            int k, const int* pair_a, const int* pair_b,
                            int dim_log2, double* C) {
    if (!A || !B || !C || rA < 0 || rB < 0 || k < 0 || nfree < 0) return -1;
    if (dim_log2 != 1 && dim_log2 != 2) return -1;
    if (nfree + k > 62 / dim_log2) return -1;
    const uint64_t D = 1ull << dim_log2;
    const uint64_t mask = D - 1;
    int idxA[64], idxB[64];
    memset(idxA, 0, sizeof idxA);
    memset(idxB, 0, sizeof idxB);
    const uint64_t nC = 1ull << (dim_log2 * nfree);
    const uint64_t nS = 1ull << (dim_log2 * k);
    for (uint64_t c = 0; c < nC; ++c) {
        for (int i = 0; i < nfree; ++i) {
            int d = (int)((c >> (dim_log2 * i)) & mask);
            if (free_from_a[i]) idxA[free_axis[i]] = d; else idxB[free_axis[i]] = d;
        }
        double accr = C[2 * c], acci = C[2 * c + 1];
        for (uint64_t s = 0; s < nS; ++s) {
            for (int i = 0; i < k; ++i) {
                int d = (int)((s >> (dim_log2 * i)) & mask);
                idxA[pair_a[k - 1 - i]] = d;
                idxB[pair_b[k - 1 - i]] = d;
            }
            uint64_t a = 0, b = 0;
            for (int i = 0; i < rA; ++i) a += (uint64_t)idxA[i] << (dim_log2 * i);
            for (int i = 0; i < rB; ++i) b += (uint64_t)idxB[i] << (dim_log2 * i);
            const double ar = A[2 * a], ai = A[2 * a + 1];
            const double br = B[2 * b], bi = B[2 * b + 1];
            const double pr = ar * br - ai * bi;
            const double pi = ar * bi + ai * br;
            accr += pr;
            acci += pi;
        }
        C[2 * c] = accr;
        C[2 * c + 1] = acci;
    }
    return 0;
}

/* std::stof as used for rotation angles (src/Network.h:342,383,402,422):
 * decimal text -> nearest float, then widened to double by the node ctor. */
double oracle_stof(const char* text) { return (double)strtof(text, NULL); }

/* std::stod as used for CPHASE angles (src/Network.h:615). */
double oracle_stod(const char* text) { return strtod(text, NULL); }
