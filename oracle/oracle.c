/*
 * oracle.c — CPU restatement of the local SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the benchmark's
 * CPU-baseline leg use this as the checker; the HIP product path never
 * links it.  Conventions are stated in oracle.h.
 *
 * Reference behaviour restated here (file:line under the reference tree):
 *  - bitmask generation: include/spacomm.cuh:29-117 (gen_bitmask: bit
 *    i % MASK_SIZE of word i / MASK_SIZE set iff ptr[i] != ptr[i+1]);
 *  - mask intersection: include/spacomm.cuh:126-148 (intersect_bitmasks);
 *  - compression: include/spacomm.cuh:360-402 (KeepFlagByIndex: entry kept
 *    iff the mask bit of its segment is set), :556-623 (select_ptrs:
 *    adjacent difference, zero the counts of dropped segments, inclusive
 *    rescan), :962-1091 (Compress keeps nrows/ncols and majordim);
 *  - C = A*B: include/KokkosWrap.hpp:367-391 (LocalMatrix::spgemm, the
 *    KokkosSparse::spgemm call at :375; the arithmetic order is the
 *    canonical order defined in oracle.h, since Kokkos-Kernels 4.7.01 is not
 *    vendored);
 *  - C = A + C: include/KokkosWrap.hpp:394-439 (LocalMatrix::spadd) and
 *    include/cusparse_helpers.cuh:799-905 (csrgeam2 with alpha=beta=1);
 *  - CSC -> CSR: include/KokkosWrap.hpp:441-531 (csc_to_csr via
 *    cusparseCsr2cscEx2).
 *
 * Parity pinning (which checks fix each function) is listed in the header
 * of tests/test_oracle.py.
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <omp.h>

int32_t or_num_threads(void) { return (int32_t)omp_get_max_threads(); }

void or_bitmask_gen(const int32_t *ptr, int32_t n, int32_t word_bits, void *words) {
    int32_t nw = (n + word_bits - 1) / word_bits;
    for (int32_t w = 0; w < nw; ++w) {
        uint64_t acc = 0;
        for (int32_t b = 0; b < word_bits; ++b) {
            int32_t i = w * word_bits + b;
            if (i < n && ptr[i] != ptr[i + 1]) acc |= (uint64_t)1 << b;
        }
        if (word_bits == 8) ((uint8_t *)words)[w] = (uint8_t)acc;
        else if (word_bits == 32) ((uint32_t *)words)[w] = (uint32_t)acc;
        else ((uint64_t *)words)[w] = acc;
    }
}

void or_bitmask_and(const uint64_t *a, const uint64_t *b, int32_t nwords, uint64_t *out) {
    for (int32_t w = 0; w < nwords; ++w) out[w] = a[w] & b[w];
}

int64_t or_compress(int32_t nseg, const int32_t *ptr, const int32_t *idx, const double *val,
                    const uint64_t *mask, int32_t *optr, int32_t *oidx, double *oval) {
    int64_t o = 0;
    if (optr) optr[0] = 0;
    for (int32_t s = 0; s < nseg; ++s) {
        int keep = (int)((mask[s >> 6] >> (s & 63)) & 1u);
        if (keep) {
            for (int32_t p = ptr[s]; p < ptr[s + 1]; ++p, ++o) {
                if (oidx) oidx[o] = idx[p];
                if (oval) oval[o] = val[p];
            }
        }
        if (optr) optr[s + 1] = (int32_t)o;
    }
    return o;
}

int64_t or_spgemm_flops(int32_t m, const int32_t *ap, const int32_t *ai, const int32_t *bp) {
    int64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 1024)
    for (int32_t i = 0; i < m; ++i)
        for (int32_t p = ap[i]; p < ap[i + 1]; ++p) total += bp[ai[p] + 1] - bp[ai[p]];
    return total;
}

int64_t or_spgemm_rownnz(int32_t m, int32_t n, const int32_t *ap, const int32_t *ai,
                         const int32_t *bp, const int32_t *bi, int64_t *rownnz) {
    int64_t total = 0;
#pragma omp parallel reduction(+ : total)
    {
        int32_t *mark = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        for (int32_t j = 0; j < n; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 64)
        for (int32_t i = 0; i < m; ++i) {
            int64_t cnt = 0;
            for (int32_t p = ap[i]; p < ap[i + 1]; ++p) {
                int32_t k = ai[p];
                for (int32_t q = bp[k]; q < bp[k + 1]; ++q) {
                    int32_t j = bi[q];
                    if (mark[j] != i) { mark[j] = i; ++cnt; }
                }
            }
            rownnz[i] = cnt;
            total += cnt;
        }
        free(mark);
    }
    return total;
}

static int cmp_i32(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

void or_spgemm_numeric(int32_t m, int32_t n, const int32_t *ap, const int32_t *ai, const double *av,
                       const int32_t *bp, const int32_t *bi, const double *bv,
                       const int64_t *cp, int32_t *ci, double *cv) {
#pragma omp parallel
    {
        int32_t *mark = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        double *acc = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
        for (int32_t j = 0; j < n; ++j) mark[j] = -1;
#pragma omp for schedule(dynamic, 64)
        for (int32_t i = 0; i < m; ++i) {
            int64_t o = cp[i];
            for (int32_t p = ap[i]; p < ap[i + 1]; ++p) {
                const int32_t k = ai[p];
                const double a = av[p];
                for (int32_t q = bp[k]; q < bp[k + 1]; ++q) {
                    const int32_t j = bi[q];
                    if (mark[j] != i) {
                        mark[j] = i;
                        acc[j] = 0.0;
                        ci[o++] = j;
                    }
                    acc[j] = fma(a, bv[q], acc[j]);
                }
            }
            qsort(ci + cp[i], (size_t)(o - cp[i]), sizeof(int32_t), cmp_i32);
            for (int64_t t = cp[i]; t < o; ++t) cv[t] = acc[ci[t]];
        }
        free(mark);
        free(acc);
    }
}

int64_t or_spadd_rownnz(int32_t m, const int32_t *ap, const int32_t *ai,
                        const int32_t *bp, const int32_t *bi, int64_t *rownnz) {
    int64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 1024)
    for (int32_t i = 0; i < m; ++i) {
        int32_t p = ap[i], pe = ap[i + 1], q = bp[i], qe = bp[i + 1];
        int64_t c = 0;
        while (p < pe && q < qe) {
            if (ai[p] < bi[q]) ++p;
            else if (ai[p] > bi[q]) ++q;
            else { ++p; ++q; }
            ++c;
        }
        c += (pe - p) + (qe - q);
        rownnz[i] = c;
        total += c;
    }
    return total;
}

void or_spadd_numeric(int32_t m, const int32_t *ap, const int32_t *ai, const double *av,
                      const int32_t *bp, const int32_t *bi, const double *bv,
                      const int64_t *cp, int32_t *ci, double *cv) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int32_t i = 0; i < m; ++i) {
        int32_t p = ap[i], pe = ap[i + 1], q = bp[i], qe = bp[i + 1];
        int64_t o = cp[i];
        while (p < pe || q < qe) {
            if (q >= qe || (p < pe && ai[p] < bi[q])) { ci[o] = ai[p]; cv[o] = av[p]; ++p; }
            else if (p >= pe || ai[p] > bi[q]) { ci[o] = bi[q]; cv[o] = bv[q]; ++q; }
            else { ci[o] = ai[p]; cv[o] = av[p] + bv[q]; ++p; ++q; }
            ++o;
        }
    }
}

void or_transpose(int32_t nrows, int32_t ncols, const int32_t *ap, const int32_t *ai, const double *av,
                  int32_t *tp, int32_t *ti, double *tv) {
    int64_t *cur = (int64_t *)calloc((size_t)ncols + 1, sizeof(int64_t));
    for (int32_t i = 0; i < nrows; ++i)
        for (int32_t p = ap[i]; p < ap[i + 1]; ++p) cur[ai[p] + 1]++;
    for (int32_t j = 0; j < ncols; ++j) cur[j + 1] += cur[j];
    for (int32_t j = 0; j <= ncols; ++j) tp[j] = (int32_t)cur[j];
    for (int32_t i = 0; i < nrows; ++i)
        for (int32_t p = ap[i]; p < ap[i + 1]; ++p) {
            int64_t q = cur[ai[p]]++;
            ti[q] = i;
            tv[q] = av[p];
        }
    free(cur);
}

int32_t or_rows_sorted(int32_t m, const int32_t *ap, const int32_t *ai) {
    int32_t ok = 1;
#pragma omp parallel for reduction(&& : ok) schedule(static)
    for (int32_t i = 0; i < m; ++i)
        for (int32_t p = ap[i] + 1; p < ap[i + 1]; ++p)
            if (ai[p - 1] >= ai[p]) ok = 0;
    return ok;
}
