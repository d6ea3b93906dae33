/*
 * trident_gen.h — seeded synthetic operand generators (host, CPU).
 *
 * These build the input matrices the local SpGEMM path is exercised on:
 * Erdos-Renyi, R-MAT (Graph500 style, duplicates merged), the 27-point 3D
 * Laplacian and a smoothed-aggregation prolongator.  They are test and
 * benchmark infrastructure, not part of the drop-in boundary
 * (trident_hip.h); the reference reads its operands from files through
 * dmmio (src/run_spgemm.cu:119-141) and generates AMG operators with
 * amg_matrices/multigrid_solver_matrix_generator.cpp.
 *
 * Every generator is a pure function of its arguments: random draws come
 * from a counter-based SplitMix64 stream keyed by (seed, element index), so
 * the result does not depend on the number of OpenMP threads.
 *
 * Output matrices are int32 CSR with sorted, unique column indices in every
 * row and fp64 values, allocated with malloc; release with tr_host_csr_free.
 */
#ifndef TRIDENT_GEN_H
#define TRIDENT_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tr_host_csr {
    int32_t  nrows;
    int32_t  ncols;
    int64_t  nnz;
    int32_t *ptr;   /* nrows+1 */
    int32_t *idx;   /* nnz, sorted within each row */
    double  *val;   /* nnz */
} tr_host_csr_t;

/* 0 on success, -1 on bad arguments, -2 on allocation failure. */

/* n x n, `per_row` uniform random columns per row (duplicates merged),
 * values U[-1,1). */
int tr_gen_er(int32_t n, int32_t per_row, uint64_t seed, tr_host_csr_t *out);

/* R-MAT: 2^scale vertices, edge_factor * 2^scale edge draws with quadrant
 * probabilities (a, b, c, 1-a-b-c); duplicate (row,col) pairs merged;
 * values U[-1,1) assigned after merging. */
int tr_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                uint64_t seed, tr_host_csr_t *out);

/* 27-point Laplacian on an n^3 grid: diagonal `diag`, every in-bounds
 * neighbour `off`.  Row index x + n*(y + n*z). */
int tr_gen_laplace27(int32_t n, double diag, double off, tr_host_csr_t *out);

/* Smoothed-aggregation prolongator for the n^3 27-point Laplacian with
 * diagonal `diag` and off-diagonal `off`: aggregates of agg^3 grid points,
 * tentative P0 = aggregate indicator (unnormalised, value 1.0), damped-Jacobi
 * smoothing P = (I - omega D^-1 A) P0.  Shape n^3 x ceil(n/agg)^3. */
int tr_gen_sa_prolongator(int32_t n, int32_t agg, double diag, double off,
                          double omega, tr_host_csr_t *out);

/* Host transpose (CSR of A^T), rows of the result sorted. */
int tr_host_transpose(const tr_host_csr_t *a, tr_host_csr_t *out);

void tr_host_csr_free(tr_host_csr_t *m);

#ifdef __cplusplus
}
#endif
#endif
