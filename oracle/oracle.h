/*
 * oracle.h — CPU restatement of the local SpGEMM path (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP path and nothing else: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product library (libtrident_hip.so) never links or calls it.
 *
 * Conventions shared with the HIP path (the contract this build defines
 * where the reference delegates to Kokkos-Kernels / cuSPARSE, whose exact
 * order and structural-zero policy are not reachable offline):
 *  - int32 CSR (or CSC) with 0-based indices, fp64 values;
 *  - output columns sorted ascending within every row;
 *  - structural zeros are kept: an output entry exists iff at least one
 *    product a_ik*b_kj lands on it, whatever its numeric value;
 *  - canonical accumulation order for C = A*B: for row i, walk A's entries
 *    in stored order p = ptrA[i]..ptrA[i+1]-1 (k = idxA[p]), and B_k's
 *    entries in stored order; every product updates c_ij = fma(a_ik, b_kj,
 *    c_ij) starting from c_ij = +0.0;
 *  - C = A + B: entries in both operands are a + b.
 */
#ifndef TRIDENT_ORACLE_H
#define TRIDENT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bitmask generation: bit (i % word_bits) of word i / word_bits is set iff
 * segment i is non-empty (ptr[i] != ptr[i+1]).  word_bits in {8, 32, 64};
 * `words` receives ceil(n / word_bits) words of that width. */
void or_bitmask_gen(const int32_t *ptr, int32_t n, int32_t word_bits, void *words);

/* out[w] = a[w] & b[w] for 64-bit words. */
void or_bitmask_and(const uint64_t *a, const uint64_t *b, int32_t nwords, uint64_t *out);

/* Compression: keep the entries of segment i iff bit i of `mask` (64-bit
 * words) is set; the dimensions are unchanged and dropped segments become
 * empty.  Writes optr (nseg+1), oidx/oval (returned count).  Any of
 * oidx/oval may be NULL to count only. */
int64_t or_compress(int32_t nseg, const int32_t *ptr, const int32_t *idx, const double *val,
                    const uint64_t *mask, int32_t *optr, int32_t *oidx, double *oval);

/* Per-row nnz of C = A*B (structure = boolean product).  Returns nnz(C). */
int64_t or_spgemm_rownnz(int32_t m, int32_t n, const int32_t *ap, const int32_t *ai,
                         const int32_t *bp, const int32_t *bi, int64_t *rownnz);

/* Numeric C = A*B in the canonical order; cp (m+1, int64) from the counts. */
void or_spgemm_numeric(int32_t m, int32_t n, const int32_t *ap, const int32_t *ai, const double *av,
                       const int32_t *bp, const int32_t *bi, const double *bv,
                       const int64_t *cp, int32_t *ci, double *cv);

/* Number of multiplications sum_k nnz(A[:,k]) * nnz(B[k,:]). */
int64_t or_spgemm_flops(int32_t m, const int32_t *ap, const int32_t *ai, const int32_t *bp);

/* C = A + B for sorted operands. */
int64_t or_spadd_rownnz(int32_t m, const int32_t *ap, const int32_t *ai,
                        const int32_t *bp, const int32_t *bi, int64_t *rownnz);
void or_spadd_numeric(int32_t m, const int32_t *ap, const int32_t *ai, const double *av,
                      const int32_t *bp, const int32_t *bi, const double *bv,
                      const int64_t *cp, int32_t *ci, double *cv);

/* CSR -> CSC (equivalently the CSR of the transpose); rows ascending within
 * every column. tp has ncols+1 entries. */
void or_transpose(int32_t nrows, int32_t ncols, const int32_t *ap, const int32_t *ai, const double *av,
                  int32_t *tp, int32_t *ti, double *tv);

/* 1 iff every row has strictly increasing column indices. */
int32_t or_rows_sorted(int32_t m, const int32_t *ap, const int32_t *ai);

/* Number of OpenMP threads the oracle runs with (for reporting). */
int32_t or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
