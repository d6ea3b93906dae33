// A C++ caller that follows the reference's asynchronous stage loop
// (src/hns_spgemm.cu:587-588, 640-650, 688-690, 771-1007) for every rank of a
// P x P grid whose tiles all live on one GPU, through this repository's
// drop-in headers:
//   - TileHolder receives each stage's A and B tiles (load_tile_contig in
//     place of recv_tile_contig, copy_device_local_csx for a rank's own tile);
//   - the sending rank compresses its tile with SpaCommHandler::Compress and
//     the receiver's index, as SpaComm does before a send;
//   - the A tile arrives as CSC and becomes CSR through
//     CusparseCSX(&handle, csx, conversion_buffs);
//   - LocalMatrix A_p(A_remote->mat), B_p(B_node->mat), then either
//     sp_mma_hybrid(&handle, A_p, B_p, &C_local, &C_accum, did_one_spgemm)
//     (mode 0), or the accumulation thread (ACCUM_THREAD, mode 1 with the
//     double-buffered partial C, mode 2 with a LocalMatrix accumulator);
//   - colAtoGet / rowBtoGet advance in Cannon order.
// The matrix is the R-MAT generator's (C = A * A); every rank's partial C is
// written to a binary file that tests/test_gpu_stage_loop.py checks against
// the oracle's stage products summed in the same order.
//
// usage: test_stage_loop <scale> <edge_factor> <seed> <P> <mode> <out_file>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "AccumCSX.hpp"
#include "AccumThread.hpp"
#include "KokkosWrap.hpp"
#include "SpaCommHandler.hpp"
#include "TileHolder.hpp"
#include "trident_gen.h"

using IT = int32_t;
using VT = double;
using namespace KokkosWrap;
using LM = LocalMatrix<IT, IT, VT>;

struct HostTile {
    IT nrows, ncols;
    std::vector<IT> ptr, idx;
    std::vector<VT> val;
};

// tile (bi, bj) of an n x n CSR, as CSR (csc = false) or CSC (csc = true)
static HostTile cut(const tr_host_csr_t &A, IT m, int bi, int bj, bool csc) {
    HostTile t{m, m, {}, {}, {}};
    std::vector<std::vector<std::pair<IT, VT>>> seg(m);
    for (IT r = 0; r < m; ++r) {
        const IT g = bi * m + r;
        for (int64_t p = A.ptr[g]; p < A.ptr[g + 1]; ++p) {
            const IT c = A.idx[p] - bj * m;
            if (c < 0 || c >= m) continue;
            if (csc) seg[c].push_back({r, A.val[p]});
            else seg[r].push_back({c, A.val[p]});
        }
    }
    t.ptr.assign(m + 1, 0);
    for (IT s = 0; s < m; ++s) {
        for (auto &e : seg[s]) {  // rows of a column arrive in increasing order
            t.idx.push_back(e.first);
            t.val.push_back(e.second);
        }
        t.ptr[s + 1] = (IT)t.idx.size();
    }
    return t;
}

static void check(int rc, const char *what) { KokkosWrap::detail::check(rc, what); }

int main(int argc, char **argv) {
    if (argc != 7) {
        std::fprintf(stderr, "usage: %s scale edge_factor seed P mode out_file\n", argv[0]);
        return 2;
    }
    const int scale = std::atoi(argv[1]), ef = std::atoi(argv[2]), P = std::atoi(argv[4]), mode = std::atoi(argv[5]);
    const uint64_t seed = std::strtoull(argv[3], nullptr, 10);
    tr_host_csr_t H;
    if (tr_gen_rmat(scale, ef, 0.57, 0.19, 0.19, seed, &H) != 0) return 2;
    const IT n = H.nrows, m = n / P;
    if (m * P != n) return 2;
    tr_ctx_t ctx = KokkosWrap::context();

    // every rank's local A (CSC) and B (CSR) tile, on the device
    std::vector<tr_csx_t> dA(P * P), dB(P * P);
    std::vector<mmio::CSX<IT, VT>> cA(P * P), cB(P * P);
    IT A_max_nnz = 1, B_max_nnz = 1;
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < P; ++j) {
            HostTile a = cut(H, m, i, j, true), b = cut(H, m, i, j, false);
            check(tr_csx_upload(ctx, TR_COLS, m, m, (int64_t)a.idx.size(), a.ptr.data(), a.idx.data(),
                                a.val.data(), &dA[i * P + j]),
                  "upload A");
            check(tr_csx_upload(ctx, TR_ROWS, m, m, (int64_t)b.idx.size(), b.ptr.data(), b.idx.data(),
                                b.val.data(), &dB[i * P + j]),
                  "upload B");
            for (int t = 0; t < 2; ++t) {
                const tr_csx_t &d = t ? dB[i * P + j] : dA[i * P + j];
                (t ? cB : cA)[i * P + j] = mmio::CSX<IT, VT>{t ? MajorDim::ROWS : MajorDim::COLS, m, m,
                                                             (IT)d.nnz, d.ptr_vec, d.idx_vec, d.val, false,
                                                             nullptr, 0};
            }
            A_max_nnz = std::max(A_max_nnz, (IT)a.idx.size());
            B_max_nnz = std::max(B_max_nnz, (IT)b.idx.size());
        }
    tr_host_csr_free(&H);

    // each rank's SpaComm filters (spcomm_2D's result on this grid)
    std::vector<SpaCommHandler<IT, VT> *> spacomm(P * P);
    for (int r = 0; r < P; ++r)
        for (int c = 0; c < P; ++c) {
            LocalGrid<IT, VT> grid;
            grid.row_size = grid.col_size = P;
            grid.row_rank = r;
            grid.col_rank = c;
            for (int t = 0; t < P * P; ++t) {
                grid.A_tiles.push_back(&cA[t]);
                grid.B_tiles.push_back(&cB[t]);
            }
            spacomm[r * P + c] = new SpaCommHandler<IT, VT>(&cA[r * P + c], &cB[r * P + c], &grid);
        }

    FILE *out = std::fopen(argv[6], "wb");
    if (!out) return 2;
    KokkosWrap::Handle handle;
    double loop_ms = 0.0;  // the ranks' stage loops (multiplies, adds, tile movement), summed
    for (int r = 0; r < P; ++r)
        for (int c = 0; c < P; ++c) {
            // buffers of the rank (hns_spgemm.cu:640-650)
            const size_t A_buf_size = mmio::CSX_buf_size<IT, VT>(m, m, (IT)(A_max_nnz * 1.5), MajorDim::COLS);
            const size_t B_buf_size = mmio::CSX_buf_size<IT, VT>(m, m, (IT)(B_max_nnz * 1.5), MajorDim::ROWS);
            TileHolder<IT, VT> A_holder(A_buf_size, m, (IT)(A_max_nnz * 1.5));
            TileHolder<IT, VT> B_holder(B_buf_size, m, (IT)(B_max_nnz * 1.5));
            void *stream = tr_ctx_stream(ctx);
            auto *conversion_buffs = new CsxBuffers<IT, VT>(A_max_nnz * 1.5, m + 1, m, &stream, 1, true);
            auto *C_accum_buffs = new CsxBuffers<IT, VT>(A_max_nnz * 1.5, m + 1, m, &stream, 1, true);
            auto *C_local_buffs = new CsxBuffers<IT, VT>(A_max_nnz * 1.5, m + 1, m, &stream, 1, true);
            CusparseCSX<IT, VT> *C_accum = new CusparseCSX<IT, VT>(C_accum_buffs);
            CusparseCSX<IT, VT> *C_local = new CusparseCSX<IT, VT>(C_local_buffs);
            LM C_lm;  // the LocalMatrix accumulator of mode 2

            AccumThreadHandle<IT, VT> *accum_handle = nullptr;
            std::thread accum_thread;
            if (mode == 1) accum_handle = new AccumThreadHandle<IT, VT>(&handle, C_local, C_accum, P, 0);
            if (mode == 2) accum_handle = new AccumThreadHandle<IT, VT>(&C_lm, P, 0);
            if (accum_handle) accum_thread = std::thread([&] { accum_thread_loop(*accum_handle); });

            check(tr_ctx_synchronize(ctx), "synchronize");
            const auto t_loop = std::chrono::steady_clock::now();
            bool did_one_spgemm = false;
            int colAtoGet = (r + c) % P;  // stagger left (hns_spgemm.cu:587)
            int rowBtoGet = (c + r) % P;  // stagger down (:588)
            SpaCommBuffers<IT, VT> sendA, sendB;
            for (int iter = 0; iter < P; ++iter) {
                // the owners compress their tiles for this receiver and send
                // them; a rank's own tile is copied (hns_spgemm.cu:847-866)
                IT A_tile_nnz, B_tile_nnz;
                const int a_src = r * P + colAtoGet, b_src = rowBtoGet * P + c;
                mmio::CSX<IT, VT> *sa = spacomm[a_src]->Compress(&cA[a_src], c, &sendA);
                mmio::CSX<IT, VT> *sb = spacomm[b_src]->Compress(&cB[b_src], r, &sendB);
                if (colAtoGet != c) A_tile_nnz = A_holder.load_tile_contig(sa->buf, sa->nnz, sa->buf_size);
                else A_tile_nnz = A_holder.copy_device_local_csx(sa, stream);
                if (rowBtoGet != r) B_tile_nnz = B_holder.load_tile_contig(sb->buf, sb->nnz, sb->buf_size);
                else B_tile_nnz = B_holder.copy_device_local_csx(sb, stream);
                std::free(sa);
                std::free(sb);

                // CSC -> CSR of A into the conversion buffers (:892-897)
                mmio::CSX<IT, VT> *A_hdr = A_holder.form_mmiocsx(m, m, A_tile_nnz, MajorDim::COLS);
                CusparseCSX<IT, VT> *A_remote = new CusparseCSX<IT, VT>(&handle, A_hdr, conversion_buffs);
                CusparseCSX<IT, VT> *B_node =
                    new CusparseCSX<IT, VT>(B_holder.form_mmiocsx(m, m, B_tile_nnz, MajorDim::ROWS));

                if (A_remote->nnz() > 0 && B_node->nnz() > 0) {  // :953-981
                    LM A_p(A_remote->mat);
                    LM B_p(B_node->mat);
                    if (accum_handle) {
                        LM C_prod = LM::spgemm(A_p, B_p);
                        if (iter > 0) accum_handle->wait_until_idle();
                        accum_handle->start_accum(std::move(C_prod));
                    } else {
                        sp_mma_hybrid(&handle, A_p, B_p, &C_local, &C_accum, did_one_spgemm);
                    }
                } else if (accum_handle) {
                    if (iter > 0) accum_handle->wait_until_idle();
                    accum_handle->signal_empty();
                }
                // the headers formed over the holders were malloc'd
                std::free(A_hdr);
                std::free(B_node->mat);
                delete A_remote;
                delete B_node;
                colAtoGet = (colAtoGet + 1) % P;  // shift left (:1000)
                rowBtoGet = (rowBtoGet + 1) % P;  // shift down (:1001)
            }
            if (accum_handle) accum_thread.join();
            check(tr_ctx_synchronize(ctx), "synchronize");
            loop_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_loop).count();

            // the rank's partial C
            tr_csx_t C{};
            if (mode == 1) C = accum_handle->C_local_csx->view();
            else if (mode == 2) C = C_lm.initialized ? C_lm.storage : tr_csx_t{};
            else C = C_local->view();
            int64_t nnz = C.ptr_vec ? C.nnz : 0;
            std::vector<IT> ptr(m + 1, 0), idx(nnz);
            std::vector<VT> val(nnz);
            if (C.ptr_vec) check(tr_csx_download(ctx, &C, ptr.data(), idx.data(), val.data()), "download");
            const int32_t hdr[4] = {r, c, m, (int32_t)nnz};
            std::fwrite(hdr, sizeof(hdr), 1, out);
            std::fwrite(ptr.data(), sizeof(IT), ptr.size(), out);
            std::fwrite(idx.data(), sizeof(IT), idx.size(), out);
            std::fwrite(val.data(), sizeof(VT), val.size(), out);
            std::printf("rank (%d,%d): nnz(C) = %lld%s\n", r, c, (long long)nnz,
                        mode == 0 && !did_one_spgemm ? " (no product)" : "");

            delete accum_handle;  // releases the accumulation context (mode 2's C with it)
            C_lm = LM();
            C_accum_buffs->explicitFree();
            C_local_buffs->explicitFree();
            delete C_accum;
            delete C_local;
            delete C_accum_buffs;
            delete C_local_buffs;
            delete conversion_buffs;
        }
    std::fclose(out);
    for (auto *s : spacomm) delete s;
    for (auto &d : dA) tr_csx_free(ctx, &d);
    for (auto &d : dB) tr_csx_free(ctx, &d);
    std::printf("stage loops: %.3f ms (mode %d)\n", loop_ms, mode);
    std::printf("ALL OK\n");
    return 0;
}
