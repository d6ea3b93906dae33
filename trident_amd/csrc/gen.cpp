// Seeded synthetic operand generators (host C++ with OpenMP).
//
// Test and benchmark infrastructure for the local SpGEMM path; see
// include/trident_gen.h for the contract.  The reference obtains operands
// from files (dmmio, src/run_spgemm.cu:119-141) and from
// amg_matrices/multigrid_solver_matrix_generator.cpp (27-point stencil at
// :140-181, which uses 26/-4/-2/-1 weights; the generator here takes the
// weights as arguments so the 26/-1 operator of the AMG workload is built).
#include "../../include/trident_gen.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Counter-based stream: draw t of element e under seed s.
struct Stream {
    uint64_t x;
    Stream(uint64_t seed, uint64_t elem) : x(mix64(seed * kGolden + 0x632BE59BD9B4E019ull) ^ (elem * kGolden)) {}
    uint64_t next() { x += kGolden; return mix64(x); }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
};

// Value stream for the merged matrix: position p of the final CSR.
inline double value_at(uint64_t seed, int64_t p) {
    uint64_t z = mix64(mix64(seed ^ 0xA5A5A5A55A5A5A5Aull) + (uint64_t)p * kGolden);
    return (double)(z >> 11) * 0x1.0p-52 - 1.0;  // [-1,1)
}

bool alloc_csr(tr_host_csr_t *out, int32_t nrows, int32_t ncols, int64_t nnz) {
    out->nrows = nrows;
    out->ncols = ncols;
    out->nnz = nnz;
    out->ptr = (int32_t *)malloc(sizeof(int32_t) * ((size_t)nrows + 1));
    out->idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(nnz, 1));
    out->val = (double *)malloc(sizeof(double) * (size_t)std::max<int64_t>(nnz, 1));
    if (!out->ptr || !out->idx || !out->val) {
        free(out->ptr); free(out->idx); free(out->val);
        out->ptr = nullptr; out->idx = nullptr; out->val = nullptr;
        return false;
    }
    return true;
}

// Build a CSR pattern from coordinate pairs: bucket by row, sort and merge
// duplicate columns per row.  Values are filled by the caller.
int coo_to_csr_pattern(int32_t nrows, int32_t ncols, const std::vector<int32_t> &rows,
                       const std::vector<int32_t> &cols, tr_host_csr_t *out) {
    const int64_t ne = (int64_t)rows.size();
    std::vector<int64_t> start((size_t)nrows + 1, 0);
    {
        std::vector<std::atomic<int64_t>> cnt((size_t)nrows);
        for (auto &c : cnt) c.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < ne; ++e) cnt[rows[e]].fetch_add(1, std::memory_order_relaxed);
        for (int32_t r = 0; r < nrows; ++r) start[r + 1] = start[r] + cnt[r].load();
    }
    std::vector<int32_t> bucket((size_t)std::max<int64_t>(ne, 1));
    {
        std::vector<std::atomic<int64_t>> cur((size_t)nrows);
        for (int32_t r = 0; r < nrows; ++r) cur[r].store(start[r], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
        for (int64_t e = 0; e < ne; ++e) {
            int64_t p = cur[rows[e]].fetch_add(1, std::memory_order_relaxed);
            bucket[p] = cols[e];
        }
    }
    std::vector<int32_t> uniq((size_t)nrows, 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t r = 0; r < nrows; ++r) {
        int32_t *b = bucket.data() + start[r];
        int64_t len = start[r + 1] - start[r];
        std::sort(b, b + len);
        int64_t u = 0;
        for (int64_t t = 0; t < len; ++t)
            if (t == 0 || b[t] != b[t - 1]) b[u++] = b[t];
        uniq[r] = (int32_t)u;
    }
    int64_t nnz = 0;
    for (int32_t r = 0; r < nrows; ++r) nnz += uniq[r];
    if (nnz > INT32_MAX) return -1;
    if (!alloc_csr(out, nrows, ncols, nnz)) return -2;
    out->ptr[0] = 0;
    for (int32_t r = 0; r < nrows; ++r) out->ptr[r + 1] = out->ptr[r] + uniq[r];
#pragma omp parallel for schedule(dynamic, 256)
    for (int32_t r = 0; r < nrows; ++r)
        std::memcpy(out->idx + out->ptr[r], bucket.data() + start[r], sizeof(int32_t) * uniq[r]);
    return 0;
}

void fill_values(tr_host_csr_t *m, uint64_t seed) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < m->nnz; ++p) m->val[p] = value_at(seed, p);
}

}  // namespace

extern "C" {

void tr_host_csr_free(tr_host_csr_t *m) {
    if (!m) return;
    free(m->ptr); free(m->idx); free(m->val);
    m->ptr = nullptr; m->idx = nullptr; m->val = nullptr;
    m->nnz = 0;
}

int tr_gen_er(int32_t n, int32_t per_row, uint64_t seed, tr_host_csr_t *out) {
    if (!out || n <= 0 || per_row < 0) return -1;
    std::vector<int32_t> rows((size_t)n * per_row), cols((size_t)n * per_row);
#pragma omp parallel for schedule(static)
    for (int32_t r = 0; r < n; ++r) {
        Stream s(seed, (uint64_t)r);
        for (int32_t t = 0; t < per_row; ++t) {
            rows[(size_t)r * per_row + t] = r;
            cols[(size_t)r * per_row + t] = (int32_t)(s.next() % (uint64_t)n);
        }
    }
    int rc = coo_to_csr_pattern(n, n, rows, cols, out);
    if (rc) return rc;
    fill_values(out, seed);
    return 0;
}

int tr_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                uint64_t seed, tr_host_csr_t *out) {
    if (!out || scale < 1 || scale > 30 || edge_factor < 0) return -1;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return -1;
    const int32_t n = (int32_t)1 << scale;
    const int64_t ne = (int64_t)edge_factor * n;
    std::vector<int32_t> rows((size_t)ne), cols((size_t)ne);
    const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; ++e) {
        Stream s(seed, (uint64_t)e);
        uint32_t r = 0, cc = 0;
        for (int32_t l = 0; l < scale; ++l) {
            double u = s.uniform();
            uint32_t rb = (u >= ab) ? 1u : 0u;                       // quadrants c, d
            uint32_t cb = ((u >= a && u < ab) || u >= abc) ? 1u : 0u;  // quadrants b, d
            r = (r << 1) | rb;
            cc = (cc << 1) | cb;
        }
        rows[e] = (int32_t)r;
        cols[e] = (int32_t)cc;
    }
    int rc = coo_to_csr_pattern(n, n, rows, cols, out);
    if (rc) return rc;
    fill_values(out, seed);
    return 0;
}

int tr_gen_laplace27(int32_t n, double diag, double off, tr_host_csr_t *out) {
    if (!out || n <= 0 || (int64_t)n * n * n > INT32_MAX) return -1;
    const int32_t N = n * n * n;
    std::vector<int32_t> cnt((size_t)N);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < N; ++i) {
        int x = i % n, y = (i / n) % n, z = i / (n * n);
        int cx = 1 + (x > 0) + (x < n - 1), cy = 1 + (y > 0) + (y < n - 1), cz = 1 + (z > 0) + (z < n - 1);
        cnt[i] = cx * cy * cz;
    }
    int64_t nnz = 0;
    for (int32_t i = 0; i < N; ++i) nnz += cnt[i];
    if (nnz > INT32_MAX) return -1;
    if (!alloc_csr(out, N, N, nnz)) return -2;
    out->ptr[0] = 0;
    for (int32_t i = 0; i < N; ++i) out->ptr[i + 1] = out->ptr[i] + cnt[i];
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < N; ++i) {
        int x = i % n, y = (i / n) % n, z = i / (n * n);
        int32_t p = out->ptr[i];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int xx = x + dx, yy = y + dy, zz = z + dz;
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= n || yy >= n || zz >= n) continue;
                    out->idx[p] = xx + n * (yy + n * zz);
                    out->val[p] = (dx == 0 && dy == 0 && dz == 0) ? diag : off;
                    ++p;
                }
    }
    return 0;
}

int tr_gen_sa_prolongator(int32_t n, int32_t agg, double diag, double off, double omega,
                          tr_host_csr_t *out) {
    if (!out || n <= 0 || agg <= 0 || (int64_t)n * n * n > INT32_MAX || diag == 0.0) return -1;
    const int32_t N = n * n * n;
    const int32_t nc1 = (n + agg - 1) / agg;
    const int32_t NC = nc1 * nc1 * nc1;
    auto agg_of = [&](int x, int y, int z) { return (x / agg) + nc1 * ((y / agg) + nc1 * (z / agg)); };
    // Row i of P: for every aggregate c touched by the 27-neighbourhood of i,
    // P[i,c] = [agg(i)==c] - (omega/diag) * sum_{j in N(i), agg(j)=c} A_ij.
    // At most 8 distinct aggregates can be touched (2 per axis).
    std::vector<int32_t> cnt((size_t)N);
    auto row_entries = [&](int32_t i, int32_t *cols, double *vals) -> int {
        int x = i % n, y = (i / n) % n, z = i / (n * n);
        int k = 0;
        int32_t ck[8];
        double sk[8];
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int xx = x + dx, yy = y + dy, zz = z + dz;
                    if (xx < 0 || yy < 0 || zz < 0 || xx >= n || yy >= n || zz >= n) continue;
                    double aij = (dx == 0 && dy == 0 && dz == 0) ? diag : off;
                    int32_t c = agg_of(xx, yy, zz);
                    int t = 0;
                    while (t < k && ck[t] != c) ++t;
                    if (t == k) { ck[k] = c; sk[k] = 0.0; ++k; }
                    sk[t] += aij;
                }
        const int32_t own = agg_of(x, y, z);
        const double scale = omega / diag;
        // sort by column (insertion sort on <= 8 entries)
        for (int s = 1; s < k; ++s)
            for (int t = s; t > 0 && ck[t - 1] > ck[t]; --t) {
                std::swap(ck[t - 1], ck[t]);
                std::swap(sk[t - 1], sk[t]);
            }
        if (cols) {
            for (int t = 0; t < k; ++t) {
                cols[t] = ck[t];
                vals[t] = (ck[t] == own ? 1.0 : 0.0) - scale * sk[t];
            }
        }
        return k;
    };
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < N; ++i) cnt[i] = row_entries(i, nullptr, nullptr);
    int64_t nnz = 0;
    for (int32_t i = 0; i < N; ++i) nnz += cnt[i];
    if (!alloc_csr(out, N, NC, nnz)) return -2;
    out->ptr[0] = 0;
    for (int32_t i = 0; i < N; ++i) out->ptr[i + 1] = out->ptr[i] + cnt[i];
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < N; ++i) row_entries(i, out->idx + out->ptr[i], out->val + out->ptr[i]);
    return 0;
}

int tr_host_transpose(const tr_host_csr_t *a, tr_host_csr_t *out) {
    if (!a || !out) return -1;
    if (!alloc_csr(out, a->ncols, a->nrows, a->nnz)) return -2;
    std::vector<int64_t> cnt((size_t)a->ncols + 1, 0);
    for (int64_t p = 0; p < a->nnz; ++p) cnt[a->idx[p] + 1]++;
    for (int32_t j = 0; j < a->ncols; ++j) cnt[j + 1] += cnt[j];
    for (int32_t j = 0; j <= a->ncols; ++j) out->ptr[j] = (int32_t)cnt[j];
    // rows visited in increasing order -> each output row sorted
    for (int32_t i = 0; i < a->nrows; ++i)
        for (int32_t p = a->ptr[i]; p < a->ptr[i + 1]; ++p) {
            int64_t q = cnt[a->idx[p]]++;
            out->idx[q] = i;
            out->val[q] = a->val[p];
        }
    return 0;
}

}  // extern "C"
