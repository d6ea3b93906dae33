// Drives the KokkosWrap::LocalMatrix shim (include/KokkosWrap.hpp) the way
// the reference's test/test_kokkos.cu:159-162 does — sp_mma twice into one
// accumulator — plus spgemm, spadd and a CSC operand, and checks every
// result against a dense host computation in the canonical order
// (c = fma(a, b, c) over A's stored entries).  Exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "KokkosWrap.hpp"

using LM = KokkosWrap::LocalMatrix<int32_t, int32_t, double>;

struct HostCsr {
    int32_t n, m;
    std::vector<int32_t> ptr, idx;
    std::vector<double> val;
};

static HostCsr random_csr(int32_t n, int32_t m, int per_row, unsigned seed) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    HostCsr h{n, m, {0}, {}, {}};
    for (int32_t r = 0; r < n; ++r) {
        std::vector<char> used(m, 0);
        int k = (int)(g() % (per_row + 1));
        for (int t = 0; t < k; ++t) used[g() % m] = 1;
        for (int32_t c = 0; c < m; ++c)
            if (used[c]) {
                h.idx.push_back(c);
                h.val.push_back(u(g));
            }
        h.ptr.push_back((int32_t)h.idx.size());
    }
    return h;
}

// dense canonical product, presence tracked separately (structural zeros)
static void dense_product(const HostCsr &A, const HostCsr &B, std::vector<double> &D, std::vector<char> &P) {
    D.assign((size_t)A.n * B.m, 0.0);
    P.assign((size_t)A.n * B.m, 0);
    for (int32_t i = 0; i < A.n; ++i)
        for (int32_t p = A.ptr[i]; p < A.ptr[i + 1]; ++p) {
            int32_t k = A.idx[p];
            for (int32_t q = B.ptr[k]; q < B.ptr[k + 1]; ++q) {
                size_t o = (size_t)i * B.m + B.idx[q];
                D[o] = std::fma(A.val[p], B.val[q], D[o]);
                P[o] = 1;
            }
        }
}

static tr_csx_t upload(const HostCsr &h, int32_t majordim = TR_ROWS) {
    tr_csx_t d{};
    int32_t nr = majordim == TR_ROWS ? h.n : h.m, nc = majordim == TR_ROWS ? h.m : h.n;
    KokkosWrap::detail::check(tr_csx_upload(KokkosWrap::context(), majordim, nr, nc, (int64_t)h.idx.size(),
                                            h.ptr.data(), h.idx.data(), h.val.data(), &d),
                              "upload");
    return d;
}

static mmio::CSX<int32_t, double> as_csx(const tr_csx_t &d) {
    return mmio::CSX<int32_t, double>{d.majordim == TR_ROWS ? MajorDim::ROWS : MajorDim::COLS,
                                      d.nrows, d.ncols, (int32_t)d.nnz, d.ptr_vec, d.idx_vec, d.val,
                                      false, nullptr, 0};
}

static int compare(LM &C, const std::vector<double> &D, const std::vector<char> &P, int32_t n, int32_t m,
                   const char *what) {
    mmio::CSX<int32_t, double> *c = C.get_csx();
    std::vector<int32_t> ptr(n + 1), idx(c->nnz);
    std::vector<double> val(c->nnz);
    tr_csx_t t = C.storage;
    KokkosWrap::detail::check(tr_csx_download(KokkosWrap::context(), &t, ptr.data(), idx.data(), val.data()),
                              "download");
    std::free(c);
    int64_t expect = 0;
    for (char x : P) expect += x;
    if ((int64_t)idx.size() != expect) {
        std::printf("%s: nnz %zu != %lld\n", what, idx.size(), (long long)expect);
        return 1;
    }
    for (int32_t i = 0; i < n; ++i)
        for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) {
            if (p > ptr[i] && idx[p] <= idx[p - 1]) {
                std::printf("%s: row %d unsorted\n", what, i);
                return 1;
            }
            size_t o = (size_t)i * m + idx[p];
            if (!P[o] || std::memcmp(&D[o], &val[p], sizeof(double)) != 0) {
                std::printf("%s: mismatch at (%d,%d): %.17g vs %.17g\n", what, i, idx[p], val[p], D[o]);
                return 1;
            }
        }
    std::printf("%s: ok (%lld entries)\n", what, (long long)expect);
    return 0;
}

int main() {
    const int32_t n = 300, k = 200, m = 250;
    HostCsr A1 = random_csr(n, k, 12, 1), B1 = random_csr(k, m, 9, 2);
    HostCsr A2 = random_csr(n, k, 12, 3), B2 = random_csr(k, m, 9, 4);
    tr_csx_t dA1 = upload(A1), dB1 = upload(B1), dA2 = upload(A2), dB2 = upload(B2);
    auto cA1 = as_csx(dA1), cB1 = as_csx(dB1), cA2 = as_csx(dA2), cB2 = as_csx(dB2);
    LM a1(&cA1), b1(&cB1), a2(&cA2), b2(&cB2);
    int fails = 0;

    // spgemm
    std::vector<double> D1, D2;
    std::vector<char> P1, P2;
    dense_product(A1, B1, D1, P1);
    LM C1 = LM::spgemm(a1, b1);
    fails += compare(C1, D1, P1, n, m, "spgemm");

    // sp_mma twice into one accumulator (test_kokkos.cu:159-162)
    dense_product(A2, B2, D2, P2);
    LM acc;
    LM::sp_mma(a1, b1, acc);
    LM::sp_mma(a2, b2, acc);
    std::vector<double> S(D1.size());
    std::vector<char> PS(P1.size());
    for (size_t o = 0; o < S.size(); ++o) {
        PS[o] = P1[o] | P2[o];
        // the accumulator is C_prod + C_local: (A2*B2) + (A1*B1)
        S[o] = (P1[o] && P2[o]) ? D2[o] + D1[o] : (P1[o] ? D1[o] : D2[o]);
    }
    fails += compare(acc, S, PS, n, m, "sp_mma x2");

    // spadd: acc2 = C1 + C1 (uninitialised accumulator first copies)
    LM acc2;
    LM::spadd(C1, acc2);
    LM::spadd(C1, acc2);
    std::vector<double> T(D1.size());
    for (size_t o = 0; o < T.size(); ++o) T[o] = D1[o] + D1[o];
    fails += compare(acc2, T, P1, n, m, "spadd");

    // CSC operand converted on the device
    std::vector<int32_t> cptr(k + 1, 0), cidx(A1.idx.size());
    std::vector<double> cval(A1.idx.size());
    for (int32_t c : A1.idx) cptr[c + 1]++;
    for (int32_t c = 0; c < k; ++c) cptr[c + 1] += cptr[c];
    std::vector<int32_t> cur(cptr.begin(), cptr.end() - 1);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t p = A1.ptr[i]; p < A1.ptr[i + 1]; ++p) {
            int32_t q = cur[A1.idx[p]]++;
            cidx[q] = i;
            cval[q] = A1.val[p];
        }
    HostCsr Acsc{k, n, cptr, cidx, cval};
    tr_csx_t dAc = upload(Acsc, TR_COLS);
    auto cAc = as_csx(dAc);
    LM ac(&cAc);
    LM C3 = LM::spgemm(ac, b1);
    fails += compare(C3, D1, P1, n, m, "csc operand");

    C1.freeBuffers();
    acc.freeBuffers();
    acc2.freeBuffers();
    C3.freeBuffers();
    ac.freeBuffers();
    tr_ctx_t ctx = KokkosWrap::context();
    tr_csx_free(ctx, &dA1);
    tr_csx_free(ctx, &dB1);
    tr_csx_free(ctx, &dA2);
    tr_csx_free(ctx, &dB2);
    tr_csx_free(ctx, &dAc);
    std::printf(fails ? "FAILED\n" : "ALL OK\n");
    return fails ? 1 : 0;
}
