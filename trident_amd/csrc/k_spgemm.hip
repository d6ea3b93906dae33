// Local SpGEMM C = A*B on gfx950 (int32 CSR, fp64 values).
//
// Replaces LocalMatrix::spgemm / KokkosSparse::spgemm
// (include/KokkosWrap.hpp:367-391, call at :375).  Pipeline:
//
//  1. row analysis   one wave per row: flops_i = sum_{k in A_i} nnz(B_k);
//                    rows are binned by flops (empty / light classes by hash
//                    table size / heavy).
//  2. symbolic       light rows (flops <= the tuned cut): one wave per row, LDS
//                    open-addressing hash of column keys;
//                    heavy rows: one 1024-thread block per row, a bitmap of
//                    2^20 columns in LDS (128 KiB of the 160 KiB), popcounted
//                    per 128-column fine block; the block packs fine blocks
//                    into column windows of at most H+127 entries and
//                    appends them to a device work list.
//  3. scan           row counts -> 64-bit row offsets; C is cut into int32
//                    tiles (panels) when it exceeds INT32_MAX entries.
//  4. numeric        light rows: one wave per row, LDS hash of (key, fp64),
//                    A's entries walked in stored order with the lanes over
//                    B_k (no two lanes share a column, so values need no
//                    atomics and every output sees its products in order);
//                    light rows are compacted and bitonic-sorted in LDS;
//                    heavy rows: one workgroup per item (a run of windows of
//                    one row) sharing one LDS table per window, products of
//                    the window flattened over the workgroup in A-order,
//                    column conflicts inside a step applied in thread (= A)
//                    order, sorted ranks from a presence bitmap.
//
// Every output entry therefore receives c = fma(a_ik, b_kj, c) in the order
// of A's stored entries, starting from +0.0: the results are bit-identical to
// the CPU oracle and do not depend on scheduling.
#include "dev_common.h"
#include "tr_internal.h"
#include "spgemm_types.h"

#include <algorithm>
#include <vector>


namespace tr {


namespace {

using namespace dev;

__device__ __forceinline__ int32_t lower_bound_dev(const TR_GLOBAL int32_t *__restrict__ a, int32_t lo, int32_t hi,
                                                   int32_t key) {
    while (lo < hi) {
        int32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// first q in (lo, e] with a[q] >= key (q == e if none), galloping from lo,
// where a[lo] < key is known; vq receives a[q] (INT32_MAX when q == e), the
// column the next window of a heavy row starts from
__device__ __forceinline__ int32_t gallop_from(const TR_GLOBAL int32_t *__restrict__ a, int32_t lo, int32_t e,
                                               int32_t key, int32_t &vq) {
    int32_t step = 1, hi;
    vq = INT32_MAX;
    while (true) {
        hi = lo + step;
        if (hi >= e) {
            hi = e;
            vq = INT32_MAX;
            break;
        }
        const int32_t v = a[hi];
        if (v >= key) {
            vq = v;
            break;
        }
        lo = hi;
        step <<= 1;
    }
    while (hi - lo > 1) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        const int32_t v = a[mid];
        if (v < key) {
            lo = mid;
        } else {
            hi = mid;
            vq = v;
        }
    }
    return hi;
}

// Visit the column indices of B row segment [s, e) with the wave's lanes,
// four loads in flight per lane before the first is consumed.
template <typename F>
__device__ __forceinline__ void stream_cols(const TR_GLOBAL int32_t *__restrict__ bi, int32_t s, int32_t e, F &&f) {
    int32_t q = s + lane_id();
    for (; q + 3 * kWave < e; q += 4 * kWave) {
        const int32_t j0 = bi[q], j1 = bi[q + kWave], j2 = bi[q + 2 * kWave], j3 = bi[q + 3 * kWave];
        f(j0);
        f(j1);
        f(j2);
        f(j3);
    }
    for (; q < e; q += kWave) f(bi[q]);
}

// Same visit with 16-byte loads for the aligned body of the segment (four
// columns per lane per load, four loads in flight), scalar head and tail.
// Alignment is taken from the address: caller tiles may place indices at any
// 4-byte offset.  kEager issues the head, the tail and the first two groups
// per lane at once (more registers: for kernels whose occupancy LDS bounds).
template <bool kEager, typename F>
__device__ __forceinline__ void stream_cols_v4(const TR_GLOBAL int32_t *__restrict__ bi, int32_t s, int32_t e,
                                               F &&f) {
    const int l = lane_id();
    const int32_t mis = (int32_t)(((uintptr_t)(bi + s) >> 2) & 3);  // dwords past a 16-byte boundary
    const int32_t sa = min(e, s + ((4 - mis) & 3));                  // first 16-byte aligned index
    const int32_t nv = (e - sa) >> 2;                                // whole int4 groups
    typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
    const TR_GLOBAL i32x4 *__restrict__ b4 = (const TR_GLOBAL i32x4 *)(bi + sa);
    const int32_t ta = sa + (nv << 2);
    if constexpr (kEager) {
    // head, tail and the first two groups per lane are loaded together: a
    // segment of up to 2 * 64 groups costs one round trip
    const bool hh = s + l < sa, ht = ta + l < e, h0 = l < nv, h1 = l + kWave < nv;
    int32_t jh = 0, jt = 0;
    i32x4 g0 = {0, 0, 0, 0}, g1 = {0, 0, 0, 0};
    if (hh) jh = bi[s + l];
    if (ht) jt = bi[ta + l];
    if (h0) g0 = b4[l];
    if (h1) g1 = b4[l + kWave];
    if (hh) f(jh);
    if (ht) f(jt);
    if (h0) f(g0.x), f(g0.y), f(g0.z), f(g0.w);
    if (h1) f(g1.x), f(g1.y), f(g1.z), f(g1.w);
    }
    int32_t v = kEager ? l + 2 * kWave : l;
    if (!kEager && s + l < sa) f(bi[s + l]);
    for (; v + 3 * kWave < nv; v += 4 * kWave) {
        const i32x4 x0 = b4[v], x1 = b4[v + kWave], x2 = b4[v + 2 * kWave], x3 = b4[v + 3 * kWave];
        f(x0.x), f(x0.y), f(x0.z), f(x0.w);
        f(x1.x), f(x1.y), f(x1.z), f(x1.w);
        f(x2.x), f(x2.y), f(x2.z), f(x2.w);
        f(x3.x), f(x3.y), f(x3.z), f(x3.w);
    }
    for (; v < nv; v += kWave) {
        const i32x4 x = b4[v];
        f(x.x), f(x.y), f(x.z), f(x.w);
    }
    if (!kEager && ta + l < e) f(bi[ta + l]);
}

// Streaming of B rows with their loads one round trip ahead, for kernels
// whose LDS leaves registers to spare (the 2^20-bit heavy symbolic runs one
// 16-wave workgroup per CU): each row's first round trip (its unaligned head
// and tail and its first two 16-byte groups per lane, as stream_cols_v4<true>)
// is issued while the previous row is applied, and a long row's next 1024
// columns while the current ones are.
typedef int32_t i32x4_t __attribute__((ext_vector_type(4)));

struct RowHead {
    int32_t s, e, sa, nv, jh, jt;
    i32x4_t g0, g1;
};

// Loads below are issued unconditionally so that the compiler can keep the
// next round trip's loads in flight while it waits for the current ones (a
// branch around a load makes it wait for every load): a lane with nothing to
// read loads the aligned 16 bytes around an element known to exist (the
// row's first, or bi[0], which a heavy row's operand always has), which
// stays on a mapped page, and its value is dropped.
__device__ __forceinline__ const TR_GLOBAL i32x4_t *safe_block(const TR_GLOBAL int32_t *__restrict__ bi, int32_t s,
                                                             int32_t e) {
    return (const TR_GLOBAL i32x4_t *)((uintptr_t)(bi + (e > s ? s : 0)) & ~(uintptr_t)15);
}

__device__ __forceinline__ RowHead row_head_issue(const TR_GLOBAL int32_t *__restrict__ bi, int32_t s, int32_t e) {
    RowHead h;
    const int l = lane_id();
    const int32_t mis = (int32_t)(((uintptr_t)(bi + s) >> 2) & 3);
    h.s = s;
    h.e = e;
    h.sa = min(e, s + ((4 - mis) & 3));
    h.nv = (e - h.sa) >> 2;
    const int32_t ta = h.sa + (h.nv << 2);
    const TR_GLOBAL i32x4_t *__restrict__ b4 = (const TR_GLOBAL i32x4_t *)(bi + h.sa);
    const TR_GLOBAL i32x4_t *__restrict__ sb = safe_block(bi, s, e);
    const TR_GLOBAL int32_t *__restrict__ s1 = (const TR_GLOBAL int32_t *)sb;
    h.jh = *(s + l < h.sa ? bi + s + l : s1);
    h.jt = *(ta + l < e ? bi + ta + l : s1);
    h.g0 = *(l < h.nv ? b4 + l : sb);
    h.g1 = *(l + kWave < h.nv ? b4 + l + kWave : sb);
    return h;
}

template <typename F>
__device__ __forceinline__ void row_stream_apply(const TR_GLOBAL int32_t *__restrict__ bi, const RowHead &h, F &f) {
    const int l = lane_id();
    const int32_t ta = h.sa + (h.nv << 2);
    if (h.s + l < h.sa) f(h.jh);
    if (ta + l < h.e) f(h.jt);
    if (l < h.nv) f(h.g0.x), f(h.g0.y), f(h.g0.z), f(h.g0.w);
    if (l + kWave < h.nv) f(h.g1.x), f(h.g1.y), f(h.g1.z), f(h.g1.w);
    if (h.nv <= 2 * kWave) return;  // wave-uniform
    const TR_GLOBAL i32x4_t *__restrict__ b4 = (const TR_GLOBAL i32x4_t *)(bi + h.sa);
    const TR_GLOBAL i32x4_t *__restrict__ sb = safe_block(bi, h.s, h.e);
    i32x4_t c[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t v = 2 * kWave + k * kWave + l;
        c[k] = *(v < h.nv ? b4 + v : sb);
    }
    for (int32_t base = 2 * kWave; base < h.nv; base += 4 * kWave) {  // wave-uniform
        i32x4_t n[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t v = base + 4 * kWave + k * kWave + l;
            n[k] = *(v < h.nv ? b4 + v : sb);
        }
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (base + k * kWave + l < h.nv) f(c[k].x), f(c[k].y), f(c[k].z), f(c[k].w);
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = n[k];
    }
}

// the nk rows [bs, be) held by lanes 0..nk-1, each streamed over the wave
template <typename F>
__device__ __forceinline__ void stream_rows_ahead(const TR_GLOBAL int32_t *__restrict__ bi, int32_t bs, int32_t be,
                                                  int nk, F &&f) {
    if (nk <= 0) return;
    RowHead cur = row_head_issue(bi, __shfl(bs, 0, kWave), __shfl(be, 0, kWave));
    for (int t = 0; t < nk; ++t) {
        RowHead nxt = cur;
        if (t + 1 < nk) nxt = row_head_issue(bi, __shfl(bs, t + 1, kWave), __shfl(be, t + 1, kWave));
        row_stream_apply(bi, cur, f);
        cur = nxt;
    }
}

// Same with the values (all entries of one B row have distinct columns, so
// the order in which one lane applies them does not matter).
template <typename F>
__device__ __forceinline__ void stream_entries(const TR_GLOBAL int32_t *__restrict__ bi,
                                               const TR_GLOBAL double *__restrict__ bv, int32_t s, int32_t e, F &&f) {
    int32_t q = s + lane_id();
    for (; q + kWave < e; q += 2 * kWave) {
        const int32_t j0 = bi[q], j1 = bi[q + kWave];
        const double b0 = bv[q], b1 = bv[q + kWave];
        f(j0, b0);
        f(j1, b1);
    }
    for (; q < e; q += kWave) f(bi[q], bv[q]);
}

// task owning global row g (largest t with row_base <= g)
__device__ __forceinline__ int task_of(const TaskDev *__restrict__ tasks, int nt, int32_t g) {
    int lo = 0, hi = nt - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tasks[mid].row_base <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ const Panel &find_panel(const Panel *__restrict__ panels, int np, int32_t row) {
    int lo = 0, hi = np - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (panels[mid].row_begin <= row) lo = mid;
        else hi = mid - 1;
    }
    return panels[lo];
}

// ---------------------------------------------------------------- analysis
// 256 global rows per block: a thread per row for short rows, the whole wave
// for rows with more than 32 entries; class counters are aggregated in LDS
// so there is one global atomic per class per block.
constexpr int kRowsPerBlock = 256;

__global__ void k_zero_counters(Counters *__restrict__ cnt) {
    if (threadIdx.x < sizeof(Counters) / 4) ((int32_t *)cnt)[threadIdx.x] = 0;
}

__global__ void k_zero_i64(int64_t *__restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

__global__ __launch_bounds__(kRowsPerBlock) void k_row_flops(const TaskDev *__restrict__ tasks, int nt,
                                                             int32_t rows_total, int64_t *__restrict__ flops,
                                                             int32_t *__restrict__ lists,
                                                             Counters *__restrict__ cnt, int32_t light_max) {
    __shared__ int32_t s_cnt[kNumClasses], s_base[kNumClasses];
    __shared__ unsigned long long s_f, s_hf;
    const int tid = threadIdx.x, l = lane_id();
    const int32_t g = blockIdx.x * kRowsPerBlock + tid;
    if (tid < kNumClasses) s_cnt[tid] = 0;
    if (tid == 0) s_f = s_hf = 0ull;
    __syncthreads();
    int32_t a0 = 0, a1 = 0, ncols = 0;
    const int32_t *ai = nullptr, *bp = nullptr;
    if (g < rows_total) {
        const TaskDev &T = tasks[task_of(tasks, nt, g)];
        const int32_t r = g - T.row_base;
        a0 = T.ap[r];
        a1 = T.ap[r + 1];
        ai = T.ai;
        bp = T.bp;
        ncols = T.n;
    }
    int64_t f = 0;
    if (a1 - a0 <= 32) {
        for (int32_t p = a0; p < a1; ++p) {
            const int32_t k = ai[p];
            f += bp[k + 1] - bp[k];
        }
    }
    uint64_t longm = ballot(a1 - a0 > 32);
    while (longm) {
        const int src = __ffsll((unsigned long long)longm) - 1;
        longm &= longm - 1;
        const int32_t s0 = __shfl(a0, src, kWave), s1 = __shfl(a1, src, kWave);
        const int32_t *sai = (const int32_t *)__shfl((long long)ai, src, kWave);
        const int32_t *sbp = (const int32_t *)__shfl((long long)bp, src, kWave);
        int64_t acc = 0;
        for (int32_t p = s0 + l; p < s1; p += kWave) {
            const int32_t k = sai[p];
            acc += sbp[k + 1] - sbp[k];
        }
        acc = wave_sum(acc);
        if (l == src) f = acc;
    }
    int c = -1, lpos = 0;
    if (g < rows_total && f > 0) {
        if (f <= light_max) {
            const int64_t bound = f < ncols ? f : ncols;
            int64_t T = 64;
            c = 0;
            while (T < 2 * bound) {
                T <<= 1;
                ++c;
            }
        } else {
            c = f <= kMidMaxF ? kMidClass : kHeavyClass;
        }
        lpos = atomicAdd(&s_cnt[c], 1);
    }
    const unsigned long long wf = (unsigned long long)wave_sum(f);
    const unsigned long long whf = (unsigned long long)wave_sum(c >= kHeavyClass ? f : (int64_t)0);
    if (l == 0) {
        atomicAdd(&s_f, wf);
        atomicAdd(&s_hf, whf);
    }
    __syncthreads();
    if (tid < kNumClasses && s_cnt[tid] > 0) s_base[tid] = atomicAdd(&cnt->cls[tid], s_cnt[tid]);
    if (tid == 0) {
        if (s_f) atomicAdd(&cnt->flops, s_f);
        if (s_hf) atomicAdd(&cnt->heavy_flops, s_hf);
    }
    __syncthreads();
    if (c >= 0) lists[(int64_t)c * rows_total + s_base[c] + lpos] = g;
    if (g < rows_total) flops[g] = f;
}

// ---------------------------------------------------------------- symbolic, light rows
// One wave per row, rows taken from the class list (grid-stride, count read
// on the device so the launch can be replayed from a graph).
template <int T>
__global__ __launch_bounds__(256) void k_sym_light(const TaskDev *__restrict__ tasks, int nt,
                                                   const int32_t *__restrict__ rows, const Counters *__restrict__ cnt,
                                                   int cls, int64_t *__restrict__ rownnz) {
    __shared__ int32_t tabs[4][T];
    constexpr uint32_t kLog = __builtin_ctz(T);
    // (marking the row wave-uniform here moves its values to scalar
    // registers, 101 SGPRs: 7 waves per SIMD instead of 8, and slower)
    const int w = threadIdx.x / kWave, l = lane_id();
    int32_t *tab = tabs[w];
    const int32_t nrows = cnt->cls[cls];
    for (int32_t li = blockIdx.x * 4 + w; li < nrows; li += gridDim.x * 4) {
        const int32_t g = rows[li];
        const TaskDev &TK = tasks[task_of(tasks, nt, g)];
        const int32_t r = g - TK.row_base;
        const TR_GLOBAL int32_t *__restrict__ ai = as_global(TK.ai);
        const TR_GLOBAL int32_t *__restrict__ bp = as_global(TK.bp);
        const TR_GLOBAL int32_t *__restrict__ bi = as_global(TK.bi);
        for (int t = l; t < T; t += kWave) tab[t] = kEmpty;
        wave_sync();
        const int32_t a0 = as_global(TK.ap)[r], a1 = as_global(TK.ap)[r + 1];
        int cntv = 0;
        auto insert = [&](int32_t j) {
            uint32_t h = hash_slot(j, kLog);
            for (int probe = 0; probe < T; ++probe) {
                const int32_t old = atomicCAS(&tab[h], kEmpty, j);
                if (old == kEmpty) {
                    ++cntv;
                    break;
                }
                if (old == j) break;
                h = (h + 1) & (T - 1);
            }
        };
        for (int32_t pb = a0; pb < a1; pb += kWave) {
            int32_t bs = 0, be = 0;
            if (pb + l < a1) {
                const int32_t k = ai[pb + l];
                bs = bp[k];
                be = bp[k + 1];
            }
            const int nk = min(kWave, a1 - pb);
            const int32_t len = be - bs;
            const int32_t inc = wave_incl_scan(len);
            const int32_t total = __shfl(inc, kWave - 1, kWave);
            if (total >= 32 * nk) {
                // long B rows: lanes stream one B row at a time
                for (int t = 0; t < nk; ++t) {
                    const int32_t s = __shfl(bs, t, kWave), e = __shfl(be, t, kWave);
                    stream_cols(bi, s, e, insert);
                }
            } else {
                // short B rows: the chunk's products flattened over the lanes
                // (counting distinct columns does not depend on the order);
                // product x belongs to the last lane whose prefix is <= x
                const int32_t excl = inc - len;
                for (int32_t x0 = 0; x0 < total; x0 += kWave) {
                    const int32_t x = x0 + l;
                    int t = 0;
#pragma unroll
                    for (int step = kWave / 2; step > 0; step >>= 1)
                        if (__shfl(excl, t + step, kWave) <= x) t += step;
                    const int32_t q = __shfl(bs, t, kWave) + (x - __shfl(excl, t, kWave));
                    if (x < total) insert(bi[q]);
                }
            }
        }
        cntv = wave_sum(cntv);
        if (l == 0) rownnz[g] = cntv;
        wave_sync();
    }
}

// LDS word of presence-bitmap word wd: with kSwz, its bank bits (the low
// five) XORed with the next two groups of five, a bijection on each aligned
// group of 32 words (clearing whole groups clears every word in use)
template <bool kSwz>
__device__ __forceinline__ int32_t bm_word(int32_t wd) {
    return kSwz ? wd ^ (((wd >> 5) ^ (wd >> 10)) & 31) : wd;
}

// ---------------------------------------------------------------- symbolic, heavy rows
// One NTH-thread block per row (grid-stride over the heavy list): a bitmap
// of 2^PB columns in LDS per pass, popcounts per 128-column fine block, and
// the row's column windows (<= win_h + 127 entries, <= kWinMaxWidth columns)
// appended to the device work lists.
template <int PB, int NTH>
// Narrow bitmaps (2^18 bits on 512 threads, 2^19 on 1024) leave LDS for 32
// waves per CU; the register budget is held to that too (64 VGPRs and a
// few spilled bytes per lane): Galerkin 128^3 12.5 -> 12.0 ms, R-MAT s19
// (2^19 columns) 17.8 -> 15.0 ms.
__global__ __launch_bounds__(NTH, (PB <= 18 || (PB == 19 && NTH == 1024)) ? 8 : 1) void k_sym_heavy(
    const TaskDev *__restrict__ tasks, int nt, const int32_t *__restrict__ rows, Counters *__restrict__ cnt, int cls,
    int64_t *__restrict__ rownnz, int emit, int32_t win_h, int32_t item_windows, Window *__restrict__ wins,
    int32_t win_cap, Item *__restrict__ items, int32_t item_cap, int32_t *__restrict__ lists, int32_t rows_total) {
    constexpr int kWords = 1 << (PB - 5);          // bitmap words of one pass
    constexpr int kFine = 1 << (PB - kFineShift);  // fine blocks of one pass
    constexpr int kPerThread = kFine / NTH;        // consecutive fine blocks per thread: 2..8
    constexpr int kWavesB = NTH / kWave;
    static_assert(kPerThread >= 1 && kPerThread <= 8 && (kPerThread & (kPerThread - 1)) == 0, "fine blocks per thread");
    // B rows are streamed with eager loads where the bitmap's LDS already
    // limits the kernel to four waves per SIMD (2^20 bits), so the extra
    // registers cost no occupancy
    constexpr bool kEagerB = PB == 20 || (PB == 19 && NTH == 512);
    // R-MAT column bits are mostly 0, so a quarter of the marks would share a
    // bank; the 2^20-bit kernel XORs the word's bank bits with higher ones
    // (R-MAT s20: 34.8 -> 33.2 ms).  In the 64-VGPR kernels the extra work
    // costs more than it saves (2^19 bits: 16.0 -> 18.4 ms).
    constexpr bool kSwizzle = PB == 20;
    static_assert(kWords % 32 == 0, "the bitmap holds whole 32-word swizzle groups");
    __shared__ __align__(16) uint32_t bm[kWords];
    __shared__ int32_t wsum[kWavesB];
    __shared__ int32_t sh_wbase;

    const int tid = threadIdx.x, w = tid / kWave, l = lane_id();
    int32_t *fex_lds = (int32_t *)bm;            // [0, kFine) after marking
    int32_t *first_lds = (int32_t *)bm + kFine;  // [kFine, 2 kFine)
    const int32_t nheavy = cnt->cls[cls];
    for (int32_t hi = blockIdx.x; hi < nheavy; hi += gridDim.x) {
        const int32_t g = rows[hi];
        const TaskDev &TK = tasks[task_of(tasks, nt, g)];
        const int32_t r = g - TK.row_base;
        const TR_GLOBAL int32_t *__restrict__ ai = as_global(TK.ai);
        const TR_GLOBAL int32_t *__restrict__ bp = as_global(TK.bp);
        const TR_GLOBAL int32_t *__restrict__ bi = as_global(TK.bi);
        const int32_t ncols = TK.n;
        const int32_t a0 = as_global(TK.ap)[r], a1 = as_global(TK.ap)[r + 1];
        int32_t row_total = 0;
        for (int64_t base = 0; base < ncols; base += (1 << PB)) {
            const int32_t width = (int32_t)min<int64_t>(1 << PB, (int64_t)ncols - base);
            const int32_t nwords = (width + 31) >> 5;
            const int32_t nfine = (width + (1 << kFineShift) - 1) >> kFineShift;
            // only the pass's words are marked and counted (a narrow matrix
            // leaves the rest of the bitmap unused)
            for (int t = tid; t < (kSwizzle ? ((nwords + 31) & ~31) : ((nwords + 3) & ~3)); t += NTH) bm[t] = 0u;
            __syncthreads();
            // mark: entry pb + w + 16 t goes to wave w
            for (int32_t pb = a0; pb < a1; pb += kWavesB * kWave) {
                int32_t bs = 0, be = 0;
                const int32_t p = pb + w + kWavesB * l;
                if (p < a1) {
                    const int32_t k = ai[p];
                    bs = bp[k];
                    be = bp[k + 1];
                }
                const int32_t rem = a1 - (pb + w);
                const int nk = rem <= 0 ? 0 : min(kWave, (rem + kWavesB - 1) / kWavesB);
                auto mark = [&](int32_t j) {
                    const int64_t jj = (int64_t)j - base;
                    if (jj >= 0 && jj < width) atomicOr(&bm[bm_word<kSwizzle>((int32_t)(jj >> 5))], 1u << (jj & 31));
                };
                if constexpr (kEagerB) {
                    stream_rows_ahead(bi, bs, be, nk, mark);
                } else {
                    for (int t = 0; t < nk; ++t) {
                        const int32_t s = __shfl(bs, t, kWave), e = __shfl(be, t, kWave);
                        stream_cols_v4<false>(bi, s, e, mark);
                    }
                }
            }
            __syncthreads();
            // fine-block counts, kPerThread consecutive fine blocks per thread.
            // A fine block's four words are one 16-byte read (the swizzle
            // keeps them in their aligned 16 bytes), and read k of lane l
            // takes block (k + l / (8 / kPerThread)) mod kPerThread: eight
            // consecutive lanes then read eight different 16-byte bank groups
            // (in order, lanes kPerThread blocks apart would share one or two:
            // a 32- or 64-way conflict on every read)
            int32_t c[kPerThread];
            int32_t tsum = 0;
#pragma unroll
            for (int u = 0; u < kPerThread; ++u) c[u] = 0;
#pragma unroll
            for (int k = 0; k < kPerThread; ++k) {
                const int u = (k + l / (8 / kPerThread)) & (kPerThread - 1);
                const int f = tid * kPerThread + u;
                int32_t v = 0;
                if (f < nfine) {
                    const uint4 x = *reinterpret_cast<const uint4 *>(&bm[bm_word<kSwizzle>(f * 4) & ~3]);
                    v = __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);
                }
#pragma unroll
                for (int uu = 0; uu < kPerThread; ++uu) c[uu] = uu == u ? v : c[uu];
                tsum += v;
            }
            const int32_t inc = wave_incl_scan(tsum);
            if (l == kWave - 1) wsum[w] = inc;
            __syncthreads();
            int32_t off = 0, pass_total = 0;
            for (int v = 0; v < kWavesB; ++v) {
                if (v < w) off += wsum[v];
                pass_total += wsum[v];
            }
            int32_t fex[kPerThread];
            {
                int32_t run = off + inc - tsum;
#pragma unroll
                for (int u = 0; u < kPerThread; ++u) {
                    fex[u] = run;
                    run += c[u];
                }
            }
            if (pass_total == 0 || !emit) {
                row_total += pass_total;
                __syncthreads();
                continue;
            }
            // A row done in one pass whose product has few entries (many
            // multiplications per entry: the coarse Galerkin product has about
            // 55) goes to the light numeric, whose small table and lane-mask
            // ordering suit it, instead of to windows: appended to the light
            // class whose table holds twice its entries.
            if (ncols <= (1 << PB) && pass_total <= kCompactMax) {
                if (tid == 0) {
                    const int c = pass_total <= 32 ? 0 : pass_total <= 64 ? 1 : 2;
                    const int32_t pos = atomicAdd(&cnt->cls[c], 1);
                    lists[(int64_t)c * rows_total + pos] = g;
                }
                row_total += pass_total;
                __syncthreads();
                continue;
            }
            // 16-byte writes: a thread's offsets are consecutive words
            if constexpr (kPerThread % 4 == 0) {
#pragma unroll
                for (int u = 0; u < kPerThread; u += 4)
                    *reinterpret_cast<int4 *>(&fex_lds[tid * kPerThread + u]) =
                        make_int4(fex[u], fex[u + 1], fex[u + 2], fex[u + 3]);
            } else {
#pragma unroll
                for (int u = 0; u < kPerThread; ++u) fex_lds[tid * kPerThread + u] = fex[u];
            }
            __syncthreads();
            // window starts: a count quantum or a width quantum begins.  A
            // sparse pass (few entries per width quantum) is cut by a smaller
            // count quantum only; its windows may be as wide as the pass and
            // are ranked without a bitmap by the numeric kernel.
            const bool sparse_pass = (int64_t)pass_total * (1 << kWidthShift) <= (int64_t)kSparseAvg * nfine;
            const int32_t q = sparse_pass ? kSparseH : win_h;
            int flags = 0, nflag = 0;
#pragma unroll
            for (int u = 0; u < kPerThread; ++u) {
                const int f = tid * kPerThread + u;
                bool first = false;
                // the previous block's offset is this thread's, except for its first block
                const int32_t prev = u == 0 ? (f > 0 ? fex_lds[f - 1] : 0) : fex[u - 1];
                if (f < nfine)
                    first = (f == 0) || (fex[u] / q != prev / q) ||
                            (!sparse_pass && ((f >> kWidthShift) != ((f - 1) >> kWidthShift)));
                if (first) {
                    flags |= 1 << u;
                    ++nflag;
                }
            }
            const int32_t finc = wave_incl_scan(nflag);
            __syncthreads();
            if (l == kWave - 1) wsum[w] = finc;
            __syncthreads();
            int32_t foff = 0, nwin = 0;
            for (int v = 0; v < kWavesB; ++v) {
                if (v < w) foff += wsum[v];
                nwin += wsum[v];
            }
            int32_t rk = foff + finc - nflag;
#pragma unroll
            for (int u = 0; u < kPerThread; ++u)
                if (flags & (1 << u)) first_lds[rk++] = tid * kPerThread + u;
            // rows of more than kLongRow entries get a quarter of the
            // windows per item: their items are the longest, and the last
            // of them would otherwise run alone at the end of the numeric
            const int32_t iw = (a1 - a0 > kLongRow) ? max(1, item_windows / 4) : item_windows;
            if (tid == 0) {
                const int32_t nitems = (nwin + iw - 1) / iw;
                const int32_t wb = atomicAdd(&cnt->windows, nwin);
                const int32_t ib = atomicAdd(&cnt->items, nitems);
                if (wb + nwin > win_cap || ib + nitems > item_cap) {
                    atomicOr(&cnt->overflow, 1);
                    sh_wbase = -1;
                } else {
                    sh_wbase = wb;
                    for (int32_t it = 0; it < nitems; ++it) {
                        Item I;
                        I.row = g;
                        I.wbeg = wb + it * iw;
                        I.wend = min(wb + nwin, I.wbeg + iw);
                        I.pad = 0;
                        items[ib + it] = I;
                    }
                }
            }
            __syncthreads();
            const int32_t wbase = sh_wbase;
            if (wbase >= 0) {
                for (int32_t rr = tid; rr < nwin; rr += NTH) {
                    const int32_t f = first_lds[rr];
                    const int32_t fn = (rr + 1 < nwin) ? first_lds[rr + 1] : nfine;
                    Window W;
                    W.c0 = (int32_t)(base + ((int64_t)f << kFineShift));
                    W.c1 = (int32_t)min<int64_t>(base + ((int64_t)fn << kFineShift), base + width);
                    W.off = row_total + fex_lds[f];
                    W.cnt = (fn < nfine ? fex_lds[fn] : pass_total) - fex_lds[f];
                    wins[wbase + rr] = W;
                }
            }
            row_total += pass_total;
            __syncthreads();
        }
        if (tid == 0) rownnz[g] = row_total;
        __syncthreads();
    }
}

// per item: cursor slots it needs (multi-window items of rows longer than
// reg_max entries; shorter rows keep their cursors in registers)
__global__ void k_item_cursor_len(const Item *__restrict__ items, const Counters *__restrict__ cnt,
                                  const TaskDev *__restrict__ tasks, int nt, int32_t cap, int32_t reg_max,
                                  int32_t *__restrict__ len) {
    const int32_t it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= cap) return;
    int32_t v = 0;
    if (it < cnt->items) {
        const Item I = items[it];
        if (I.wend - I.wbeg > 1) {
            const TaskDev &TK = tasks[task_of(tasks, nt, I.row)];
            const int32_t r = I.row - TK.row_base;
            v = as_global(TK.ap)[r + 1] - as_global(TK.ap)[r];
            if (v <= reg_max) v = 0;
        }
    }
    len[it] = v;
}

// ---------------------------------------------------------------- heavy item order
// Heavy items are processed in the column order of their first window, so
// the workgroups resident at one time work in one column region of B and
// share its lines through L2 (the items of one row still run left to right
// within their item).  Counting sort by column bucket: a block-aggregated
// histogram, a scan, and a scatter with block-local ranks.
// bucket = first window's c0 >> shift; the host picks the smallest shift
// that maps the widest operand onto the buckets
constexpr int kOrderBuckets = 4096;

__global__ void k_fill_u64(unsigned long long *__restrict__ p, int32_t n) {
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0ull;
}

__device__ __forceinline__ int item_bucket(const Item *__restrict__ items, const Window *__restrict__ wins,
                                           int32_t it, int shift) {
    const int32_t c0 = wins[items[it].wbeg].c0;
    return min(c0 >> shift, kOrderBuckets - 1);
}

__global__ __launch_bounds__(256) void k_item_hist(const Item *__restrict__ items, const Window *__restrict__ wins,
                                                   const Counters *__restrict__ cnt, int shift,
                                                   unsigned long long *__restrict__ hist) {
    __shared__ int32_t h[kOrderBuckets];
    for (int t = threadIdx.x; t < kOrderBuckets; t += 256) h[t] = 0;
    __syncthreads();
    const int32_t n = cnt->items;
    for (int32_t it = blockIdx.x * 256 + threadIdx.x; it < n; it += gridDim.x * 256)
        atomicAdd(&h[item_bucket(items, wins, it, shift)], 1);
    __syncthreads();
    for (int t = threadIdx.x; t < kOrderBuckets; t += 256)
        if (h[t]) atomicAdd(&hist[t], (unsigned long long)h[t]);
}

// hist[0..kOrderBuckets) -> exclusive offsets, in place (one block)
__global__ __launch_bounds__(1024) void k_item_hist_scan(unsigned long long *__restrict__ hist) {
    __shared__ unsigned long long part[1024 / kWave];
    constexpr int kPer = kOrderBuckets / 1024;
    const int tid = threadIdx.x, w = tid / kWave, l = lane_id();
    unsigned long long v[kPer], sum = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        v[u] = hist[tid * kPer + u];
        sum += v[u];
    }
    const unsigned long long inc = wave_incl_scan(sum);
    if (l == kWave - 1) part[w] = inc;
    __syncthreads();
    unsigned long long off = 0;
    for (int u = 0; u < w; ++u) off += part[u];
    unsigned long long run = off + inc - sum;
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        hist[tid * kPer + u] = run;
        run += v[u];
    }
}

// The same grid walk as k_item_hist: block-local ranks per bucket, one
// global cursor bump per (block, bucket)
__global__ __launch_bounds__(256) void k_item_scatter(const Item *__restrict__ items, const Window *__restrict__ wins,
                                                      const Counters *__restrict__ cnt, int shift,
                                                      unsigned long long *__restrict__ cursor,
                                                      int32_t *__restrict__ order) {
    __shared__ int32_t h[kOrderBuckets];
    __shared__ unsigned long long base[kOrderBuckets];
    const int32_t n = cnt->items;
    for (int32_t it0 = blockIdx.x * 256; it0 < n; it0 += gridDim.x * 256) {
        for (int t = threadIdx.x; t < kOrderBuckets; t += 256) h[t] = 0;
        __syncthreads();
        const int32_t it = it0 + threadIdx.x;
        int b = 0, r = 0;
        if (it < n) {
            b = item_bucket(items, wins, it, shift);
            r = atomicAdd(&h[b], 1);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < kOrderBuckets; t += 256)
            if (h[t]) base[t] = atomicAdd(&cursor[t], (unsigned long long)h[t]);
        __syncthreads();
        if (it < n) order[base[b] + r] = it;
        __syncthreads();
    }
}

// panel p starts at global row row_begin: its first entry's global offset
__global__ void k_panel_bases(Panel *__restrict__ panels, int np, const int64_t *__restrict__ rowptr) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < np) panels[p].nnz_begin = rowptr[panels[p].row_begin];
}

// ---------------------------------------------------------------- numeric, light rows
constexpr int kLightPre = 4;  // B rows prefetched together per batch of a light row

// WPB waves per block, one row per wave (big tables use one-wave blocks so
// that LDS, not the block shape, bounds the waves per CU)
template <int T, int WPB>
// Tables of <= 256 slots leave LDS for 8 waves per SIMD; the registers are
// held to that too (63 VGPRs, 12 spilled bytes per lane; about 1% on stage
// 22:3 and Galerkin 128^3 against 65 VGPRs and 7 waves).
__global__ __launch_bounds__(WPB * kWave, T <= 256 ? 8 : 1) void k_num_light(const TaskDev *__restrict__ tasks, int nt,
                                                           const int32_t *__restrict__ rows,
                                                           Counters *__restrict__ cnt, int cls,
                                                           const int64_t *__restrict__ rowptr,
                                                           const Panel *__restrict__ panels, int npanels) {
    // tables of <= 512 slots also flatten chunks of short B rows over the
    // lanes; per-slot words order the chunk's products: lane masks (<= 256
    // slots) or owner lanes (512 slots, where 64-bit masks would cost
    // occupancy)
    constexpr bool kFlat = T <= 512;
    constexpr bool kMask = T <= 256;
    constexpr int kOwn = kFlat ? T : 1;
    using OwnT = typename std::conditional<kMask, unsigned long long, uint32_t>::type;
    __shared__ int32_t keys_s[WPB][T];
    __shared__ double vals_s[WPB][T];
    __shared__ OwnT own_s[WPB][kOwn];
    constexpr uint32_t kLog = __builtin_ctz(T);
    // the wave index and everything per row are wave-uniform (scalar loads
    // and registers)
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), l = lane_id();
    int32_t *keys = keys_s[w];
    double *vals = vals_s[w];
    OwnT *own = own_s[w];
    if (kFlat)
        for (int t = l; t < kOwn; t += kWave) own[t] = kMask ? 0 : kWave;
    const int32_t nrows = cnt->cls[cls];
    const int32_t stride = gridDim.x * WPB;
    int32_t g_next = 0;
    if (blockIdx.x * WPB + w < nrows) g_next = rows[blockIdx.x * WPB + w];
    for (int32_t li = blockIdx.x * WPB + w; li < nrows; li += stride) {
        // a row is a chain of dependent loads (row, A entries, B row
        // pointers, B entries); the next row's index and this row's output
        // place are loaded early so that their latency overlaps the chain
        const int32_t g = __builtin_amdgcn_readfirstlane(g_next);
        if (li + stride < nrows) g_next = rows[li + stride];
        const TaskDev &TK = tasks[task_of(tasks, nt, g)];
        const int32_t r = g - TK.row_base;
        const TR_GLOBAL int32_t *__restrict__ ai = as_global(TK.ai);
        const TR_GLOBAL double *__restrict__ av = as_global(TK.av);
        const TR_GLOBAL int32_t *__restrict__ bp = as_global(TK.bp);
        const TR_GLOBAL int32_t *__restrict__ bi = as_global(TK.bi);
        const TR_GLOBAL double *__restrict__ bv = as_global(TK.bv);
        const int32_t a0 = as_global(TK.ap)[r], a1 = as_global(TK.ap)[r + 1];
        const int64_t rs = rowptr[g];
        const int64_t re = rowptr[g + 1];
        const Panel P = find_panel(panels, npanels, g);
        for (int t = l; t < T; t += kWave) keys[t] = kEmpty;
        wave_sync();
        for (int32_t pb = a0; pb < a1; pb += kWave) {
            int32_t bs = 0, be = 0;
            double a = 0.0;
            if (pb + l < a1) {
                const int32_t k = ai[pb + l];
                a = av[pb + l];
                bs = bp[k];
                be = bp[k + 1];
            }
            const int nk = min(kWave, a1 - pb);
            if (kFlat) {
                const int32_t len = be - bs;
                const int32_t inc = wave_incl_scan(len);
                const int32_t total = __shfl(inc, kWave - 1, kWave);
                if (total < 24 * nk) {
                    // Short B rows: 64 products of the chunk per round, flattened
                    // in A-order over the lanes (product x belongs to the last
                    // lane whose prefix is <= x).  Products that meet in a slot
                    // are applied in lane (= A) order: each pass, the lowest
                    // pending lane of every slot applies its product.
                    const int32_t excl = inc - len;
                    for (int32_t x0 = 0; x0 < total; x0 += kWave) {
                        const int32_t x = x0 + l;
                        int t = 0;
                        int32_t q;
                        if constexpr (kMask && T >= 2 * kWave) {
                            // The lanes whose B rows meet this round write (lane,
                            // B offset) at the position where their rows start in
                            // it (or 0) and mark it in a start mask; position x
                            // belongs to the highest start at or below it.  The
                            // first 65 lane-mask words, all zero between rounds,
                            // hold both and are cleared again before the round
                            // uses them.  Three round trips instead of the six
                            // dependent shuffles of a search over the lanes.
                            const bool meets = len > 0 && excl < x0 + kWave && excl + len > x0;
                            const int32_t pos = max(excl - x0, 0);
                            if (meets) {
                                own[pos] = ((unsigned long long)(uint32_t)l << 32) | (uint32_t)(bs - excl);
                                atomicOr(&own[kWave], 1ull << pos);
                            }
                            wave_sync();
                            const unsigned long long starts = own[kWave] & (~0ull >> (63 - l));
                            const unsigned long long e = own[63 - __clzll((long long)starts)];
                            t = (int)(e >> 32);
                            q = (int32_t)(uint32_t)e + x;
                            wave_sync();
                            own[l] = 0ull;
                            if (l == 0) own[kWave] = 0ull;
                            wave_sync();
                        } else {
#pragma unroll
                            for (int step = kWave / 2; step > 0; step >>= 1)
                                if (__shfl(excl, t + step, kWave) <= x) t += step;
                            q = __shfl(bs, t, kWave) + (x - __shfl(excl, t, kWave));
                        }
                        const double at = __shfl(a, t, kWave);
                        bool pend = x < total;
                        uint32_t h = 0;
                        double b = 0.0;
                        if (pend) {
                            const int32_t j = bi[q];
                            b = bv[q];
                            h = hash_slot(j, kLog);
                            bool placed = false;
                            for (int probe = 0; probe < T; ++probe) {
                                const int32_t old = atomicCAS(&keys[h], kEmpty, j);
                                if (old == kEmpty) {
                                    vals[h] = 0.0;
                                    placed = true;
                                    break;
                                }
                                if (old == j) {
                                    placed = true;
                                    break;
                                }
                                h = (h + 1) & (T - 1);
                            }
                            if (!placed) {
                                atomicOr(&cnt->overflow, 2);  // table full: symbolic counts were wrong
                                pend = false;
                            }
                        }
                        wave_sync();
                        if constexpr (kMask) {
                            // each slot's lanes as a mask; its lowest lane
                            // applies them all in lane order, fetching the
                            // other lanes' operands across the wave
                            if (pend) atomicOr(&own[h], 1ull << l);
                            wave_sync();
                            unsigned long long rem = 0ull;
                            double v = 0.0;
                            if (pend) {
                                const unsigned long long m = own[h];
                                if (__ffsll(m) - 1 == l) {
                                    rem = m;
                                    v = vals[h];
                                }
                            }
                            while (ballot(rem != 0ull)) {  // every lane takes part in the fetches
                                const int t = rem ? __ffsll(rem) - 1 : l;
                                const double ta = __shfl(at, t, kWave);
                                const double tb = __shfl(b, t, kWave);
                                if (rem) {
                                    v = __fma_rn(ta, tb, v);
                                    rem &= rem - 1ull;
                                }
                            }
                            wave_sync();
                            if (pend && (__ffsll(own[h]) - 1 == l)) {
                                vals[h] = v;
                                own[h] = 0ull;
                            }
                            wave_sync();
                        } else {
                            while (ballot(pend)) {
                                if (pend) atomicMin(&own[h], (uint32_t)l);
                                wave_sync();
                                if (pend && own[h] == (uint32_t)l) {
                                    vals[h] = __fma_rn(at, b, vals[h]);
                                    own[h] = kWave;
                                    pend = false;
                                }
                                wave_sync();
                            }
                        }
                    }
                    continue;
                }
            }
            auto ins = [&](int32_t j, double b, double at) {
                uint32_t h = hash_slot(j, kLog);
                for (int probe = 0; probe < T; ++probe) {
                    const int32_t old = atomicCAS(&keys[h], kEmpty, j);
                    if (old == kEmpty) {
                        vals[h] = __fma_rn(at, b, 0.0);
                        break;
                    }
                    if (old == j) {
                        vals[h] = __fma_rn(at, b, vals[h]);
                        break;
                    }
                    h = (h + 1) & (T - 1);
                }
            };
            // the first 64 entries of the first kLightPre B rows are loaded
            // together (one round trip instead of one per row), then applied
            // row by row in A-order
            int32_t pj[kLightPre];
            double pv[kLightPre];
#pragma unroll
            for (int t = 0; t < kLightPre; ++t) {
                pj[t] = 0;
                pv[t] = 0.0;
                const int32_t s = __shfl(bs, t, kWave), e = __shfl(be, t, kWave);
                if (t < nk && s + l < e) {
                    pj[t] = bi[s + l];
                    pv[t] = bv[s + l];
                }
            }
#pragma unroll
            for (int t = 0; t < kLightPre; ++t) {
                if (t < nk) {
                    const int32_t s = __shfl(bs, t, kWave), e = __shfl(be, t, kWave);
                    const double at = __shfl(a, t, kWave);
                    if (s + l < e) ins(pj[t], pv[t], at);
                    stream_entries(bi, bv, s + kWave, e, [&](int32_t j, double b) { ins(j, b, at); });
                }
            }
            for (int t = kLightPre; t < nk; ++t) {
                const int32_t s = __shfl(bs, t, kWave), e = __shfl(be, t, kWave);
                const double at = __shfl(a, t, kWave);
                stream_entries(bi, bv, s, e, [&](int32_t j, double b) { ins(j, b, at); });
            }
        }
        wave_sync();
        const int n = wave_compact_lds<double>(keys, vals, T);
        wave_sort_lds<double>(keys, vals, n);
        const int expect = (int)(re - rs);
        if (l == 0 && n != expect) atomicOr(&cnt->overflow, 2);
        const int nw = n < expect ? n : expect;  // never write past the row
        TR_GLOBAL int32_t *oi = as_global(P.idx) + (rs - P.nnz_begin);
        TR_GLOBAL double *ov = as_global(P.val) + (rs - P.nnz_begin);
        for (int t = l; t < nw; t += kWave) {
            oi[t] = keys[t];
            ov[t] = vals[t];
        }
        wave_sync();
    }
}

// ---------------------------------------------------------------- numeric, heavy rows (workgroup)
// One NT-thread workgroup per item (a run of windows of one row),
// sharing one T-slot LDS table per window.  For each window the row's A
// entries are taken NT at a time (one per thread): each thread finds the
// part of B_k inside the window (register cursor for rows with <= NT
// entries, the item's cursor array otherwise), a workgroup scan flattens the
// products in A-order, and each step applies NT of them.  Slot keys carry the
// column offset in the high bits and the slot's last pusher of the step in
// the low bits: products of one step that meet in a slot form a list that
// the last pusher applies in thread order (= A-order).  A presence bitmap
// over the window, built after the window's last step, gives every entry its
// sorted rank (prefix popcounts) for the write-out.
//
// Occupancy: the kernel is bound by the latency of its dependent LDS and
// memory chains, so resident waves matter most.  The bitmap shares LDS with
// the step arrays (31 KB per 2048-slot, 256-thread workgroup: five per CU)
// and the register budget is held to 96 VGPRs (five waves per SIMD); the
// 4096-slot table (62 KB, 512 threads) fits two workgroups per CU.
constexpr int kRegSlots = 1;  // register-held A entries per thread in the heavy numeric kernel
constexpr int kSmallNT = 256;  // threads of the 2048-slot heavy numeric
constexpr uint32_t kSlotFree = 0xffffffffu;
constexpr int kTagBits = 10;                        // thread ids < 1023
constexpr uint32_t kNoTag = (1u << kTagBits) - 1u;  // slot key = column offset << kTagBits | tag

template <int T, int NT>
struct HeavyWgLds {
    uint32_t keytag[T];
    double vals[T];
    int32_t wsum[NT / kWave];
    int32_t nocc;
    union {
        struct {                    // product steps of a window
            double pa[NT], pb[NT];  // the step's products (A value, B value) by thread
            // segment starts of three rotating product steps: bit p of
            // smask[b] is set where a segment starts at step position p
            // (position 0 always holds the segment that covers the step's
            // first product), smark[b][p] is that segment's index
            unsigned long long smask[3][NT / kWave];
            int32_t seg_d[NT];  // B position minus flattened product position of segment (A entry) o
            uint16_t link[NT];  // next older product of the same slot in this step
            typename std::conditional<(NT <= 256), uint8_t, uint16_t>::type smark[3][NT];
        };
        struct {  // extraction of a window: presence bitmap and its prefix counts
            unsigned long long bm[kBmWords];
            uint16_t wpre[kBmWords];
        };
    };
};

// marks the segment [excl, excl + len) of the calling thread in the step that
// starts at product `base`: at its first position in the step, and at every
// later position of the step that starts a 64-position word, so that each
// word's first position names the segment covering it
template <int NT, typename M>
__device__ __forceinline__ void mark_segment(unsigned long long *smask, M *smark, int32_t base, int32_t excl,
                                             int32_t len) {
    if (len <= 0) return;
    const int32_t rel = excl - base;
    const int32_t lo = rel > 0 ? rel : 0, hi = rel + len < NT ? rel + len : NT;
    if (lo >= hi) return;  // the segment has no product in this step
    smark[lo] = (M)threadIdx.x;
    atomicOr(&smask[lo >> 6], 1ull << (lo & 63));
    for (int32_t ws = (lo & ~63) + 64; ws < hi; ws += 64) {
        smark[ws] = (M)threadIdx.x;
        atomicOr(&smask[ws >> 6], 1ull);
    }
}

// segment of step position rel: the mark at the highest set bit <= rel,
// always inside rel's own word (its first position is marked)
template <typename M>
__device__ __forceinline__ int owner_marked(const unsigned long long *smask, const M *smark, int rel) {
    const int w = rel >> 6;
    const unsigned long long m = smask[w] & (~0ull >> (63 - (rel & 63)));
    return (int)smark[w * 64 + 63 - __clzll((long long)m)];
}

// orders two products (thread id, A value, B value) by thread id
__device__ __forceinline__ void order_products(uint32_t &u, double &a, double &b, uint32_t &w, double &c,
                                               double &d) {
    if (w < u) {
        const uint32_t tu = u;
        const double ta = a, tb = b;
        u = w;
        a = c;
        b = d;
        w = tu;
        c = ta;
        d = tb;
    }
}

// Every word is written with a select, so the mask stays in registers
// (a conditional store to one element becomes an indexed private array).
template <int NT>
__device__ __forceinline__ void mask_set(uint64_t (&mk)[NT / kWave], uint32_t t) {
    const uint64_t bit = 1ull << (t & 63);
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) mk[w] |= ((int)(t >> 6) == w) ? bit : 0ull;
}

// workgroup exclusive scan of v; returns the exclusive prefix, *total = sum
template <int NT>
__device__ __forceinline__ int32_t wg_excl_scan(int32_t v, int32_t *wsum, int32_t *total) {
    const int w = threadIdx.x / kWave, l = lane_id();
    const int32_t inc = wave_incl_scan(v);
    if (l == kWave - 1) wsum[w] = inc;
    __syncthreads();
    int32_t off = 0, tot = 0;
#pragma unroll
    for (int u = 0; u < NT / kWave; ++u) {
        const int32_t s = wsum[u];
        off += (u < w) ? s : 0;
        tot += s;
    }
    *total = tot;
    return off + inc - v;
}

template <int T, int NT>
__global__ __launch_bounds__(NT, (NT <= 256 ? 5 : 4)) void k_num_heavy_wg(const TaskDev *__restrict__ tasks, int nt,
                                                     const Item *__restrict__ items, Counters *__restrict__ cnt,
                                                     const Window *__restrict__ wins,
                                                     const int32_t *__restrict__ order,
                                                     const int64_t *__restrict__ cursor_base,
                                                     int4 *__restrict__ cursors, const int64_t *__restrict__ rowptr,
                                                     const Panel *__restrict__ panels, int npanels,
                                                     int32_t dense_width) {
    constexpr uint32_t kLog = __builtin_ctz(T);
    __shared__ HeavyWgLds<T, NT> L;
    const int tid = threadIdx.x, l = lane_id();
    for (int t = tid; t < T; t += NT) L.keytag[t] = kSlotFree;
    if (tid == 0) L.nocc = 0;
    __syncthreads();
    const int32_t nitems = cnt->items;
    const int32_t slot = (int32_t)blockIdx.x;
    for (int32_t ord = slot; ord < nitems; ord += gridDim.x) {
        const int32_t itx = order[ord];
        const Item I = items[itx];
        const int32_t g = I.row;
        const TaskDev &TK = tasks[task_of(tasks, nt, g)];
        const int32_t r = g - TK.row_base;
        const TR_GLOBAL int32_t *__restrict__ ai = as_global(TK.ai);
        const TR_GLOBAL double *__restrict__ av = as_global(TK.av);
        const TR_GLOBAL int32_t *__restrict__ bp = as_global(TK.bp);
        const TR_GLOBAL int32_t *__restrict__ bi = as_global(TK.bi);
        const TR_GLOBAL double *__restrict__ bv = as_global(TK.bv);
        const int32_t ncols = TK.n;
        const int32_t a0 = as_global(TK.ap)[r], a1 = as_global(TK.ap)[r + 1];
        // cursor of A entry p: (first B position at or after the next window,
        // its column, end of B_k)
        int4 *cur = cursors + cursor_base[itx];
        const int64_t rs = rowptr[g];
        const Panel &P = find_panel(panels, npanels, g);
        TR_GLOBAL int32_t *oi_row = as_global(P.idx) + (rs - P.nnz_begin);
        TR_GLOBAL double *ov_row = as_global(P.val) + (rs - P.nnz_begin);
        // rows with <= NT * kRegSlots entries: the B ranges of thread tid's A
        // entries (slots tid, tid + NT, ...) and their cursors live in
        // registers; longer rows keep cursors in the item's cursor array
        const bool reg_row = (a1 - a0) <= NT * kRegSlots;
        // (before the item's first window the cursor holds the start of B_k)
        int32_t rk_be[kRegSlots], rk_cur[kRegSlots], rk_nxt[kRegSlots];
#pragma unroll
        for (int u = 0; u < kRegSlots; ++u) {
            rk_be[u] = rk_cur[u] = rk_nxt[u] = 0;
            const int32_t p = a0 + u * NT + tid;
            if (reg_row && p < a1) {
                const int32_t k = ai[p];
                rk_cur[u] = bp[k];
                rk_be[u] = bp[k + 1];
            }
        }
        bool have_cursor = false;
        for (int32_t wi = I.wbeg; wi < I.wend; ++wi) {
            const Window W = wins[wi];
            if (W.cnt == 0) continue;  // block-uniform
            const bool full = (W.c0 == 0 && W.c1 >= ncols);
            const bool last = (wi == I.wend - 1);
            // windows up to kWinMaxWidth columns are ranked by presence
            // bitmaps of kBmWidth columns, one pass each; wider (sparse)
            // windows hold <= kSparseH + 127 entries and are ranked by
            // counting smaller keys
            const bool dense = (W.c1 - W.c0) <= kWinMaxWidth;
            // windows no wider than dense_width (<= T) address their slots by
            // column offset: no probing, no ranking pass.  The slot is
            // (off * hmul) >> (32 - kLog) either way, with hmul = 2^(32 - kLog)
            // for the identity and the golden-ratio multiplier for the hash.
            const uint32_t hmul = (W.c1 - W.c0) <= dense_width ? (1u << (32 - kLog)) : 0x9E3779B1u;
            int nfresh = 0;
            for (int32_t pb = a0, slot = 0; pb < a1; pb += NT, ++slot) {
                if (tid < 2 * (NT / kWave)) (&L.smask[0][0])[tid] = 0ull;
                // this thread's k-slot: A entry p, B segment [s, e) inside the window
                const int32_t p = pb + tid;
                int32_t s = 0, e = 0;
                if (p < a1) {
                    // c: cursor (first entry at or after this window's start),
                    // v: its column (INT32_MAX at the end of B_k); known once the
                    // item's first window has been visited
                    int32_t bs = 0, be = 0, c = 0, v = 0;
                    if (reg_row) {
                        // uniform slot index: selects, not dynamic register indexing
#pragma unroll
                        for (int u = 0; u < kRegSlots; ++u)
                            if (u == slot) {
                                be = rk_be[u];
                                c = rk_cur[u];
                                v = rk_nxt[u];
                            }
                        bs = c;
                    } else if (have_cursor && !full) {
                        const int4 cv = cur[p - a0];
                        c = cv.x;
                        v = cv.y;
                        be = cv.z;
                    } else {
                        const int32_t k = ai[p];
                        bs = bp[k];
                        be = bp[k + 1];
                    }
                    if (full) {
                        s = bs;
                        e = be;
                    } else {
                        if (!have_cursor) {
                            // an item that starts at the row's first column
                            // starts at the front of B_k: no search
                            c = W.c0 == 0 ? bs : lower_bound_dev(bi, bs, be, W.c0);
                            v = c < be ? bi[c] : INT32_MAX;
                        }
                        s = c;
                        // an A entry with nothing in this window costs no load
                        e = v >= W.c1 ? s : gallop_from(bi, s, be, W.c1, v);
                        if (reg_row) {
#pragma unroll
                            for (int u = 0; u < kRegSlots; ++u)
                                if (u == slot) {
                                    rk_cur[u] = e;
                                    rk_nxt[u] = v;
                                }
                        } else if (!last) {
                            cur[p - a0] = make_int4(e, v, be, 0);
                        }
                    }
                }
                int32_t total;
                const int32_t slen = e - s;
                const int32_t excl = wg_excl_scan<NT>(slen, L.wsum, &total);
                L.seg_d[tid] = s - excl;
                // segment starts of the first two steps (their masks were
                // cleared before the scan's barrier)
                mark_segment<NT>(L.smask[0], L.smark[0], 0, excl, slen);
                mark_segment<NT>(L.smask[1], L.smark[1], NT, excl, slen);
                __syncthreads();
                // products x = x0 + tid, flattened in A-order; the B entry of the
                // next step is loaded while the current step is applied
                int32_t jn = 0;
                double bn = 0.0, an = 0.0;
                if (tid < total) {
                    const int o = owner_marked(L.smask[0], L.smark[0], tid);
                    const int32_t q = L.seg_d[o] + tid;
                    an = av[pb + o];
                    jn = bi[q];
                    bn = bv[q];
                }
                // step k reads the marks of step k + 1 (buffer bA), clears
                // buffer bC before its first barrier and marks step k + 2 in it
                // between its barriers
                int bA = 1, bC = 2;
                for (int32_t x0 = 0; x0 < total; x0 += NT) {
                    const int32_t x = x0 + tid;
                    bool act = x < total;
                    const int32_t j = jn;
                    const double b = bn, ao = an;
                    const int32_t xn = x + NT;
                    // claim or join this product's slot first: the CAS round
                    // trip overlaps the next step's owner lookup below
                    uint32_t h = 0, key = 0, me = 0, old = kSlotFree;
                    if (act) {
                        const uint32_t off = (uint32_t)(j - W.c0);
                        key = off << kTagBits;
                        me = key | (uint32_t)tid;
                        h = (off * hmul) >> (32 - kLog);
                        old = atomicCAS(&L.keytag[h], kSlotFree, me);
                    }
                    if (xn < total) {
                        const int o = owner_marked(L.smask[bA], L.smark[bA], tid);
                        const int32_t q = L.seg_d[o] + xn;
                        an = av[pb + o];
                        jn = bi[q];
                        bn = bv[q];
                    }
                    if (tid < NT / kWave) L.smask[bC][tid] = 0ull;
                    // insert: the first product of a slot in this step creates or
                    // claims it; every later one pushes itself onto the slot's
                    // list (the tag field holds the last pusher, link[] the rest)
                    if (act) {
                        const uint32_t off = key >> kTagBits;
                        // probe to a free slot (claimed) or this key's slot; the
                        // bound only guards a table the symbolic counts overfilled
                        for (int probe = 1; old != kSlotFree && (old >> kTagBits) != off && probe < T; ++probe) {
                            h = (h + 1) & (T - 1);
                            old = atomicCAS(&L.keytag[h], kSlotFree, me);
                        }
                        uint32_t prev = kNoTag;
                        if (old == kSlotFree) {
                            L.vals[h] = 0.0;
                            ++nfresh;
                        } else if ((old >> kTagBits) == off) {
                            prev = atomicExch(&L.keytag[h], me) & kNoTag;
                        } else {
                            atomicOr(&cnt->overflow, 8);  // table full: symbolic counts were wrong
                            act = false;
                        }
                        if (act) {
                            L.link[tid] = (uint16_t)prev;
                            L.pa[tid] = ao;
                            L.pb[tid] = b;
                        }
                    }
                    __syncthreads();
                    // the last pusher of each slot applies the slot's products of
                    // this step in thread (= A) order.  Each list member's
                    // operands are read with its link; lists of up to three
                    // are ordered in registers, longer ones walked through a
                    // bitmask of thread ids (reading two or four operands per
                    // round trip there measured 2% and 7% slower)
                    if (act) {
                        const uint32_t kt = L.keytag[h];
                        const uint32_t t1 = L.link[tid];
                        double v = L.vals[h];
                        if ((kt & kNoTag) == (uint32_t)tid) {
                            if (t1 == kNoTag) {
                                v = __fma_rn(ao, b, v);
                            } else {
                                const uint32_t t2 = L.link[t1];
                                const double a1 = L.pa[t1], b1 = L.pb[t1];
                                if (t2 == kNoTag) {
                                    if ((uint32_t)tid < t1) {
                                        v = __fma_rn(ao, b, v);
                                        v = __fma_rn(a1, b1, v);
                                    } else {
                                        v = __fma_rn(a1, b1, v);
                                        v = __fma_rn(ao, b, v);
                                    }
                                } else {
                                    const uint32_t t3 = L.link[t2];
                                    const double a2 = L.pa[t2], b2 = L.pb[t2];
                                    if (t3 == kNoTag) {
                                        uint32_t u0 = (uint32_t)tid, u1 = t1, u2 = t2;
                                        double ea0 = ao, eb0 = b, ea1 = a1, eb1 = b1, ea2 = a2, eb2 = b2;
                                        order_products(u0, ea0, eb0, u1, ea1, eb1);
                                        order_products(u1, ea1, eb1, u2, ea2, eb2);
                                        order_products(u0, ea0, eb0, u1, ea1, eb1);
                                        v = __fma_rn(ea0, eb0, v);
                                        v = __fma_rn(ea1, eb1, v);
                                        v = __fma_rn(ea2, eb2, v);
                                    } else {
                                        uint64_t mk[NT / kWave];
#pragma unroll
                                        for (int w = 0; w < NT / kWave; ++w) mk[w] = 0ull;
                                        mask_set<NT>(mk, (uint32_t)tid);
                                        mask_set<NT>(mk, t1);
                                        mask_set<NT>(mk, t2);
                                        for (uint32_t t = t3; t != kNoTag; t = L.link[t]) mask_set<NT>(mk, t);
#pragma unroll
                                        for (int w = 0; w < NT / kWave; ++w) {
                                            uint64_t mw = mk[w];
                                            while (mw) {
                                                const int t = w * kWave + __ffsll((unsigned long long)mw) - 1;
                                                mw &= mw - 1ull;
                                                v = __fma_rn(L.pa[t], L.pb[t], v);
                                            }
                                        }
                                    }
                                }
                            }
                            L.vals[h] = v;
                            L.keytag[h] = key | kNoTag;
                        }
                    }
                    if (x0 + 2 * NT < total) mark_segment<NT>(L.smask[bC], L.smark[bC], x0 + 2 * NT, excl, slen);
                    __syncthreads();
                    bA = bC;
                    bC = bC == 2 ? 0 : bC + 1;
                }
                __syncthreads();  // seg_* are rewritten by the next batch
            }
            if (!full) have_cursor = true;
            // entries created in this window (each thread counted its fresh inserts)
            const int32_t wf = wave_sum(nfresh);
            if (l == 0 && wf) atomicAdd(&L.nocc, wf);
            TR_GLOBAL int32_t *oi = oi_row + W.off;
            TR_GLOBAL double *ov = ov_row + W.off;
            // Each thread's entries as rank << 20 | column offset (kNone: no
            // entry) and value.  The table's first three quarters [0, kStage)
            // are also the staging area (a window holds at most kStage
            // entries): their values are stashed in registers before a
            // barrier, the last quarter is read and staged after it.
            constexpr int kSt = (T + NT - 1) / NT;
            constexpr int kStage = 3 * T / 4;
            constexpr int kLo = (kStage + NT - 1) / NT;
            constexpr int kSp = (kSparseH + 127 + NT - 1) / NT;  // sparse entries per thread
            constexpr int kStash = kLo > kSp ? kLo : kSp;
            static_assert(kSp <= kSt && kStage % NT == 0, "stash slots are exactly the staging area's");
            constexpr uint32_t kNone = 0xffffffffu;
            uint32_t st_rk[kSt];
            double st_v[kStash];
            const int32_t lim = min(W.cnt, kStage);  // never write past the window
            int32_t *sidx = reinterpret_cast<int32_t *>(L.keytag);
            if (dense) {
                // the thread's slots (h = tid + u NT) by column offset, ranked
                // one kBmWidth-column pass at a time with a presence bitmap
                // (it shares LDS with the step arrays, which are dead now).
                // A direct window's slots are its column offsets, so a ballot
                // over 64 consecutive slots is its presence word, and its one
                // pass (width <= T < kBmWidth) needs no clearing or atomics.
                const int32_t width = W.c1 - W.c0;
                const bool direct = width <= dense_width;
#pragma unroll
                for (int u = 0; u < kSt; ++u) {
                    const int h = tid + u * NT;
                    const uint32_t kt = h < T ? L.keytag[h] : kSlotFree;
                    st_rk[u] = kt == kSlotFree ? kNone : kt >> kTagBits;
                }
                if (direct) {
                    const int wsl = __builtin_amdgcn_readfirstlane(tid / kWave);  // word of slot tid in round 0
#pragma unroll 1
                    for (int u = 0; u < kSt; ++u) {
                        const uint64_t m = ballot(L.keytag[tid + u * NT] != kSlotFree);
                        if (l == 0) L.bm[u * (NT / kWave) + wsl] = m;
                    }
                }
                static_assert(T / kWave <= kBmWords, "a direct window's slots fit one pass");
                int32_t base_rank = 0;
                for (int32_t lo = 0; lo < width; lo += kBmWidth) {  // block-uniform
                    const int nwords = min(kBmWords, (width - lo + 63) >> 6);
                    if (!direct) {
                        for (int t = tid; t < nwords; t += NT) L.bm[t] = 0ull;
                        __syncthreads();
#pragma unroll
                        for (int u = 0; u < kSt; ++u) {
                            const uint32_t rel = st_rk[u] - (uint32_t)lo;  // kNone wraps past kBmWidth
                            if (rel < (uint32_t)kBmWidth) atomicOr(&L.bm[rel >> 6], 1ull << (rel & 63));
                        }
                    }
                    __syncthreads();
                    // prefix popcounts of the presence words
                    constexpr int kPer = (kBmWords + NT - 1) / NT;
                    int32_t wc[kPer];
                    int32_t tsum = 0;
#pragma unroll
                    for (int u = 0; u < kPer; ++u) {
                        const int wd = tid * kPer + u;
                        wc[u] = (wd < nwords) ? __popcll(L.bm[wd]) : 0;
                        tsum += wc[u];
                    }
                    int32_t tot;
                    int32_t run = wg_excl_scan<NT>(tsum, L.wsum, &tot);
#pragma unroll
                    for (int u = 0; u < kPer; ++u) {
                        const int wd = tid * kPer + u;
                        if (wd < kBmWords) L.wpre[wd] = (uint16_t)run;
                        run += wc[u];
                    }
                    __syncthreads();
#pragma unroll
                    for (int u = 0; u < kSt; ++u) {
                        const uint32_t rel = st_rk[u] - (uint32_t)lo;
                        if (rel < (uint32_t)kBmWidth) {
                            const int wd = (int)(rel >> 6);
                            const uint32_t rr = (uint32_t)(base_rank + L.wpre[wd]) +
                                                __popcll(L.bm[wd] & ((1ull << (rel & 63)) - 1ull));
                            st_rk[u] |= rr << 20;
                        }
                    }
                    base_rank += tot;
                    __syncthreads();  // wsum, bm and wpre are reused by the next pass
                }
            }
            if (dense) {
                if (tid == 0 && L.nocc != W.cnt) atomicOr(&cnt->overflow, 4);
#pragma unroll
                for (int u = 0; u < kStash; ++u) {
                    const int h = tid + u * NT;
                    if (u < kLo && h < kStage && st_rk[u] != kNone) st_v[u] = L.vals[h];
                }
                __syncthreads();
#pragma unroll
                for (int u = 0; u < kSt; ++u) {
                    const int h = tid + u * NT;
                    if (h < kStage || h >= T || st_rk[u] == kNone) continue;
                    const int32_t rr = (int32_t)(st_rk[u] >> 20);
                    if (rr < lim) {
                        const double v = L.vals[h];
                        sidx[rr] = W.c0 + (int32_t)(st_rk[u] & 0xfffffu);
                        L.vals[rr] = v;
                    }
                }
            } else {
                // compact the occupied slots (keys into the bitmap's space,
                // slot numbers into the prefix array), then rank each entry by
                // the number of smaller keys
                uint32_t *lkey = reinterpret_cast<uint32_t *>(L.bm);
                uint16_t *lslot = L.wpre;
                int32_t mine = 0;
                for (int h = tid; h < T; h += NT) mine += (L.keytag[h] != kSlotFree) ? 1 : 0;
                int32_t tot;
                int32_t pos = wg_excl_scan<NT>(mine, L.wsum, &tot);
                const bool fits = tot <= kSparseH + 127;  // bound of a sparse window
                if (fits)
                    for (int h = tid; h < T; h += NT) {
                        const uint32_t kt = L.keytag[h];
                        if (kt == kSlotFree) continue;
                        lkey[pos] = kt >> kTagBits;
                        lslot[pos] = (uint16_t)h;
                        ++pos;
                    }
                __syncthreads();
                if (tid == 0 && (L.nocc != W.cnt || !fits)) atomicOr(&cnt->overflow, 4);
#pragma unroll
                for (int u = 0; u < kStash; ++u) {
                    const int e = tid + u * NT;
                    st_rk[u] = kNone;
                    if (u >= kSp || !fits || e >= tot) continue;
                    const uint32_t k = lkey[e];
                    uint32_t rr = 0;
                    for (int q = 0; q < tot; ++q) rr += lkey[q] < k ? 1u : 0u;
                    st_rk[u] = (rr << 20) | k;
                    st_v[u] = L.vals[lslot[e]];
                }
                __syncthreads();
            }
            // stage the stashed entries in rank order, then write the window
            // out with coalesced stores
#pragma unroll
            for (int u = 0; u < kStash; ++u) {
                const int32_t rr = (int32_t)(st_rk[u] >> 20);
                if (st_rk[u] != kNone && rr < lim) {
                    sidx[rr] = W.c0 + (int32_t)(st_rk[u] & 0xfffffu);
                    L.vals[rr] = st_v[u];
                }
            }
            __syncthreads();
            // streaming stores: C is written once and should not evict the B
            // rows that later windows read again
            for (int e = tid; e < W.cnt; e += NT) {
                __builtin_nontemporal_store(sidx[e], &oi[e]);
                __builtin_nontemporal_store(L.vals[e], &ov[e]);
            }
            __syncthreads();
            for (int t = tid; t < T; t += NT) L.keytag[t] = kSlotFree;
            if (tid == 0) L.nocc = 0;
            __syncthreads();
        }
    }
}

}  // namespace

// ======================================================================= host side

void *Workspace::get(int slot, size_t bytes) {
    req[slot] = std::max(req[slot], bytes);
    if (!owned) return scratch(ctx, slot, bytes);
    if (bytes == 0) bytes = 1;
    if (sz[slot] < bytes) {
        if (p[slot]) ctx->pool.release(p[slot]);
        p[slot] = nullptr;
        sz[slot] = 0;
        p[slot] = ctx->pool.alloc(bytes);
        sz[slot] = bytes;
    }
    return p[slot];
}

void Workspace::release() {
    if (!owned || !ctx) return;
    for (size_t s = 0; s < p.size(); ++s)
        if (p[s]) {
            try {
                ctx->pool.release(p[s]);
            } catch (...) {
            }
            p[s] = nullptr;
            sz[s] = 0;
        }
}

namespace {

uint32_t cap_grid(int64_t need, int64_t cap) {
    if (need <= 0) return 0;
    return (uint32_t)std::min<int64_t>(need, cap);
}


template <int T>
void launch_sym_light(tr_ctx *ctx, MulPlan &pl, int c) {
    const int32_t n = pl.h.cls[c];
    TR_LAUNCH(ctx, "sym_light", k_sym_light<T>, cap_grid(div_up((uint64_t)n, 4), 1 << 20), 256, 0, pl.d_tasks,
              pl.nt, pl.lists + (int64_t)c * pl.rows_total, pl.cnt, c, pl.rownnz);
}

template <int T>
void launch_num_light(tr_ctx *ctx, MulPlan &pl, int c, Panel *panels, int np) {
    constexpr int kWpb = T >= 1024 ? 1 : 4;
    const int32_t n = pl.h.cls[c];
    TR_LAUNCH(ctx, "num_light", (k_num_light<T, kWpb>), cap_grid(div_up((uint64_t)n, kWpb), 1 << 20), kWpb * kWave,
              0, pl.d_tasks, pl.nt, pl.lists + (int64_t)c * pl.rows_total, pl.cnt, c, pl.rowptr, panels, np);
}

}  // namespace

void mul_setup(tr_ctx *ctx, Workspace &ws, MulPlan &pl, const std::vector<TaskDev> &tasks) {
    pl.nt = (int)tasks.size();
    pl.h_tasks = tasks;
    pl.tune = ctx->tune;
    int64_t rows = 0;
    for (auto &t : pl.h_tasks) {
        t.row_base = (int32_t)rows;
        rows += t.m;
    }
    TR_REQUIRE(rows < INT32_MAX, "spgemm: total rows of the batch exceed INT32_MAX");
    pl.rows_total = (int32_t)rows;
    const size_t R = (size_t)rows + 1;
    pl.d_tasks = (TaskDev *)ws.get(S_PANELS + 0, sizeof(TaskDev) * (size_t)std::max(pl.nt, 1));
    pl.flops = (int64_t *)ws.get(S_FLOPS, sizeof(int64_t) * R);
    pl.lists = (int32_t *)ws.get(S_BINS, sizeof(int32_t) * (size_t)kNumClasses * R);
    pl.cnt = (Counters *)ws.get(S_BINCNT, sizeof(Counters));
    pl.rownnz = (int64_t *)ws.get(S_ROWNNZ, sizeof(int64_t) * R);
    pl.rowptr = (int64_t *)ws.get(S_ROWPTR64, sizeof(int64_t) * R);
    pl.scan_tmp = ws.owned ? (int64_t *)ws.get(S_SCAN_TMP, scan_tmp_bytes(rows)) : nullptr;
    TR_HIP(hipMemcpyAsync(pl.d_tasks, pl.h_tasks.data(), sizeof(TaskDev) * (size_t)pl.nt, hipMemcpyHostToDevice,
                          ctx->stream));
}

void mul_analyze(tr_ctx *ctx, MulPlan &pl) {
    TR_LAUNCH(ctx, "zero_counters", k_zero_counters, 1, 64, 0, pl.cnt);
    TR_LAUNCH(ctx, "zero_rownnz", k_zero_i64, cap_grid(div_up((uint64_t)pl.rows_total + 1, 256), 4096), 256, 0,
              pl.rownnz, (int64_t)pl.rows_total + 1);
    TR_LAUNCH(ctx, "row_flops", k_row_flops, div_up((uint64_t)pl.rows_total, kRowsPerBlock), kRowsPerBlock, 0,
              pl.d_tasks, pl.nt, pl.rows_total, pl.flops, pl.lists, pl.cnt,
              pl.tune.light_max_flops);
}

void mul_read_counts(tr_ctx *ctx, MulPlan &pl) { read_array(ctx, &pl.h, pl.cnt, sizeof(Counters)); }

template <int PB, int NTH>
void launch_sym_heavy(tr_ctx *ctx, MulPlan &pl, int cls, int per_cu, bool emit) {
    TR_LAUNCH(ctx, "sym_heavy", (k_sym_heavy<PB, NTH>), cap_grid(pl.h.cls[cls], (int64_t)ctx->cu_count * per_cu), NTH,
              0, pl.d_tasks, pl.nt, pl.lists + (int64_t)cls * pl.rows_total, pl.cnt, cls, pl.rownnz, emit ? 1 : 0,
              pl.tune.window_entries, pl.tune.item_windows, pl.wins, pl.win_cap, pl.items, pl.item_cap, pl.lists,
              pl.rows_total);
}

void mul_symbolic(tr_ctx *ctx, Workspace &ws, MulPlan &pl, bool emit, bool capture) {
    if (pl.h.cls[0]) launch_sym_light<64>(ctx, pl, 0);
    if (pl.h.cls[1]) launch_sym_light<128>(ctx, pl, 1);
    if (pl.h.cls[2]) launch_sym_light<256>(ctx, pl, 2);
    if (pl.h.cls[3]) launch_sym_light<512>(ctx, pl, 3);
    if (pl.h.cls[4]) launch_sym_light<1024>(ctx, pl, 4);
    if (pl.h.cls[5]) launch_sym_light<2048>(ctx, pl, 5);
    const int32_t nh = pl.h.cls[kHeavyClass] + pl.h.cls[kMidClass];
    if (nh) {
        int32_t max_n = 0;
        for (auto &t : pl.h_tasks) max_n = std::max(max_n, t.n);
        if (emit && !capture) {
            // windows per pass <= ceil(pass_nnz / H) + width splits (mid rows use the shorter pass)
            const int64_t passes = ((int64_t)max_n + (1 << kMidPassBits) - 1) >> kMidPassBits;
            const int64_t wsplits = (int64_t)max_n / kWinMaxWidth + 1;
            const int64_t q = std::min<int64_t>(pl.tune.window_entries, kSparseH);
            const int64_t wcap = (int64_t)(pl.h.heavy_flops / q) + (int64_t)nh * (2 * passes + wsplits) + 16;
            const int64_t icap = wcap / std::max(1, pl.tune.item_windows / 4) + (int64_t)nh * passes + 16;
            TR_REQUIRE(wcap < INT32_MAX && icap < INT32_MAX, "spgemm: heavy work list exceeds int32 capacity");
            pl.win_cap = (int32_t)wcap;
            pl.item_cap = (int32_t)icap;
            pl.wins = (Window *)ws.get(S_WINDOWS, sizeof(Window) * (size_t)wcap);
            pl.items = (Item *)ws.get(S_ITEMS, sizeof(Item) * (size_t)icap);
        }
        // The bitmap covers 2^PB columns per pass and its LDS bounds the
        // resident workgroups: 2^20 bits (128 KiB) hold one per CU, 2^19 two,
        // 2^18 four.  The narrowest pass that still covers the widest
        // operand in one pass is used (same windows, same results).  The
        // grids hold many more workgroups than fit at once: rows differ
        // widely in cost, and queued workgroups go to whichever CU frees
        // first (Galerkin 128^3: 12.9 -> 12.3 ms against one resident wave
        // of workgroups).
        const int pb = max_n <= (1 << 18) ? 18 : max_n <= (1 << 19) ? 19 : kPassBits;
        if (pl.h.cls[kHeavyClass]) {
            if (pb == 18) launch_sym_heavy<18, kMidThreads>(ctx, pl, kHeavyClass, 64, emit);
            else if (pb == 19) launch_sym_heavy<19, kHeavyThreads>(ctx, pl, kHeavyClass, 32, emit);
            else launch_sym_heavy<kPassBits, kHeavyThreads>(ctx, pl, kHeavyClass, 32, emit);
        }
        if (pl.h.cls[kMidClass]) {
            if (pb == 18) launch_sym_heavy<18, kMidThreads>(ctx, pl, kMidClass, 64, emit);
            else launch_sym_heavy<kMidPassBits, kHeavyThreads>(ctx, pl, kMidClass, 32, emit);
        }
    }
    launch_exclusive_scan_i64(ctx, pl.rownnz, pl.rowptr, pl.rows_total, pl.scan_tmp);
}

int64_t mul_cursor_prep(tr_ctx *ctx, Workspace &ws, MulPlan &pl, bool capture) {
    if (!pl.h.cls[kHeavyClass] && !pl.h.cls[kMidClass]) return 0;
    if (!capture) {
        pl.clen = (int32_t *)ws.get(S_MISC, sizeof(int32_t) * ((size_t)pl.item_cap + 1));
        pl.cbase = (int64_t *)ws.get(S_MISC2, sizeof(int64_t) * ((size_t)pl.item_cap + 1));
    }
    if (ws.owned && !capture) pl.scan_tmp = (int64_t *)ws.get(S_SCAN_TMP, scan_tmp_bytes(std::max<int64_t>(pl.item_cap, pl.rows_total)));
    TR_LAUNCH(ctx, "item_cursor_len", k_item_cursor_len, div_up((uint64_t)pl.item_cap, 256), 256, 0, pl.items, pl.cnt,
              pl.d_tasks, pl.nt, pl.item_cap, (pl.tune.heavy_table == 4096 ? 512 : kSmallNT) * kRegSlots, pl.clen);
    // column order of the items
    if (!capture) {
        pl.order = (int32_t *)ws.get(S_ORDER, sizeof(int32_t) * ((size_t)pl.item_cap + 1));
        pl.order_hist = (unsigned long long *)ws.get(S_ORDER_HIST, sizeof(unsigned long long) * kOrderBuckets);
    }
    const uint32_t og = cap_grid(div_up((uint64_t)pl.item_cap, 256), (int64_t)ctx->cu_count * 8);
    int32_t maxn = 1;
    for (const TaskDev &t : pl.h_tasks) maxn = std::max(maxn, t.n);
    int shift = 0;
    while (((int64_t)kOrderBuckets << shift) < (int64_t)maxn) ++shift;
    TR_LAUNCH(ctx, "item_order", k_fill_u64, div_up((uint64_t)kOrderBuckets, 256), 256, 0, pl.order_hist,
              kOrderBuckets);
    TR_LAUNCH(ctx, "item_order", k_item_hist, og, 256, 0, pl.items, pl.wins, pl.cnt, shift, pl.order_hist);
    TR_LAUNCH(ctx, "item_order", k_item_hist_scan, 1, 1024, 0, pl.order_hist);
    TR_LAUNCH(ctx, "item_order", k_item_scatter, og, 256, 0, pl.items, pl.wins, pl.cnt, shift, pl.order_hist,
              pl.order);
    launch_exclusive_scan_i32_to_i64(ctx, pl.clen, pl.cbase, pl.item_cap, pl.scan_tmp);
    if (capture) return pl.cursor_cap;
    return read_scalar<int64_t>(ctx, pl.cbase + pl.item_cap);
}

void mul_panel_bases(tr_ctx *ctx, MulPlan &pl, Panel *d_panels, int np) {
    TR_LAUNCH(ctx, "panel_bases", k_panel_bases, div_up((uint64_t)np, 64), 64, 0, d_panels, np, pl.rowptr);
}

void mul_numeric(tr_ctx *ctx, MulPlan &pl, Panel *dp, int np) {
    if (pl.h.cls[0]) launch_num_light<64>(ctx, pl, 0, dp, np);
    if (pl.h.cls[1]) launch_num_light<128>(ctx, pl, 1, dp, np);
    if (pl.h.cls[2]) launch_num_light<256>(ctx, pl, 2, dp, np);
    if (pl.h.cls[3]) launch_num_light<512>(ctx, pl, 3, dp, np);
    if (pl.h.cls[4]) launch_num_light<1024>(ctx, pl, 4, dp, np);
    if (pl.h.cls[5]) launch_num_light<2048>(ctx, pl, 5, dp, np);
    if (pl.h.items > 0) {
        // one workgroup per item, grid-stride beyond 2^22 workgroups (the
        // dispatch packet counts work-items in 32 bits)
        const uint32_t grid = cap_grid(pl.h.items, 1 << 22);
        if (pl.tune.heavy_table == 4096)
            TR_LAUNCH(ctx, "num_heavy", (k_num_heavy_wg<4096, 512>), grid, 512, 0, pl.d_tasks, pl.nt, pl.items,
                      pl.cnt, pl.wins, pl.order, pl.cbase, pl.cursors, pl.rowptr, dp, np, pl.tune.dense_width);
        else
            TR_LAUNCH(ctx, "num_heavy", (k_num_heavy_wg<2048, kSmallNT>), grid, kSmallNT, 0, pl.d_tasks, pl.nt, pl.items,
                      pl.cnt, pl.wins, pl.order, pl.cbase, pl.cursors, pl.rowptr, dp, np, pl.tune.dense_width);
    }
}

void mul_check(tr_ctx *ctx, MulPlan &pl) {
    const int32_t ov = read_scalar<int32_t>(ctx, &pl.cnt->overflow);
    if (ov & 1) fail(TR_ERR_HIP, "spgemm: heavy work list capacity exceeded (internal bound violated)");
    if (ov & 14) fail(TR_ERR_HIP, "spgemm: numeric row counts disagree with the symbolic phase (internal error)");
}

namespace {

TaskDev task_of_pair(const tr_csx_t *A, const tr_csx_t *B) {
    TaskDev t{};
    t.ap = A->ptr_vec;
    t.ai = A->idx_vec;
    t.av = A->val;
    t.bp = B->ptr_vec;
    t.bi = B->idx_vec;
    t.bv = B->val;
    t.m = A->nrows;
    t.n = B->ncols;
    return t;
}

}  // namespace

void check_operand_shapes(const tr_csx_t *A, const tr_csx_t *B) {
    check_csx(A, "A");
    check_csx(B, "B");
    TR_REQUIRE(A->majordim == TR_ROWS && B->majordim == TR_ROWS, "spgemm operands must be CSR (TR_ROWS)");
    TR_REQUIRE(A->ncols == B->nrows, "spgemm: A.ncols != B.nrows");
}

void validate_operands(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *B) {
    check_operand_shapes(A, B);
    int32_t *st = (int32_t *)scratch(ctx, S_TMP_CSX, 64);
    TR_HIP(hipMemsetAsync(st, 0, 2 * sizeof(int32_t), ctx->stream));
    launch_validate(ctx, A, st);
    launch_validate(ctx, B, st + 1);
    int32_t hs[2];
    read_array(ctx, hs, st, sizeof(hs));
    if (hs[0] & 3) fail(TR_ERR_INVALID, "spgemm: A is not a valid CSR (ptr or column index out of range)");
    if (hs[1] & 3) fail(TR_ERR_INVALID, "spgemm: B is not a valid CSR (ptr or column index out of range)");
    if (hs[1] & 4) fail(TR_ERR_INVALID, "spgemm: B has unsorted or duplicate columns; call tr_csx_sort_rows");
    if (hs[0] & 4) fail(TR_ERR_INVALID, "spgemm: A has unsorted or duplicate columns; call tr_csx_sort_rows");
}

int64_t spgemm_flops(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *B) {
    validate_operands(ctx, A, B);
    Workspace ws(ctx, false);
    MulPlan pl;
    mul_setup(ctx, ws, pl, {task_of_pair(A, B)});
    mul_analyze(ctx, pl);
    mul_read_counts(ctx, pl);
    return (int64_t)pl.h.flops;
}

void spgemm_symbolic_only(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t *nnz, size_t *bytes) {
    validate_operands(ctx, A, B);
    Workspace ws(ctx, false);
    MulPlan pl;
    mul_setup(ctx, ws, pl, {task_of_pair(A, B)});
    mul_analyze(ctx, pl);
    mul_read_counts(ctx, pl);
    // the same steps as the product up to its numeric phase, so every work
    // list and the cursor array are sized exactly as the product sizes them
    mul_symbolic(ctx, ws, pl, true, false);
    const int64_t ncur = mul_cursor_prep(ctx, ws, pl, false);
    const int64_t total = read_scalar<int64_t>(ctx, pl.rowptr + pl.rows_total);
    if (nnz) *nnz = total;
    if (bytes) {
        // Upper bound of the device memory tr_spgemm takes from a context
        // pool with nothing cached: every scratch slot at the size the product
        // requests (grown to 1.25x, in pool granules), the cursor array, the
        // operand checks and the output tile.  A batch places the product in
        // two blocks (indices, values) and its row pointers in a third.
        std::vector<size_t> req = ws.req;
        if (pl.h.items > 0) req[S_CURSORS] = std::max(req[S_CURSORS], sizeof(int4) * (size_t)std::max<int64_t>(ncur, 1));
        req[S_MISC2] = std::max(req[S_MISC2], validate_tmp_bytes(std::max(A->nnz, B->nnz)));
        req[S_TMP_CSX] = std::max(req[S_TMP_CSX], std::max<size_t>(64, sizeof(Panel)));
        size_t b = 0;
        for (size_t r : req)
            if (r) b += scratch_block_bytes(r);
        const size_t rows1 = (size_t)pl.rows_total + 1;
        b += pool_block_bytes((size_t)total * 12 + rows1 * 4);
        b += pool_block_bytes((size_t)total * 4) + pool_block_bytes((size_t)total * 8) + pool_block_bytes(rows1 * 4) +
             2 * pool_block_bytes(sizeof(Panel) + rows1 * 4);
        *bytes = b;
    }
}

int spgemm_panels(tr_ctx *ctx, const tr_csx_t *A, const tr_csx_t *B, int64_t max_panel_nnz, tr_csx_t *panels,
                  int32_t *row_begin, int32_t max_panels, int32_t *npanels) {
    TR_REQUIRE(panels && row_begin && npanels && max_panels >= 1, "spgemm: null panel outputs");
    TR_REQUIRE(max_panel_nnz >= 1 && max_panel_nnz <= INT32_MAX, "spgemm: max_panel_nnz must be in [1, INT32_MAX]");
    validate_operands(ctx, A, B);
    Workspace ws(ctx, false);
    MulPlan pl;
    mul_setup(ctx, ws, pl, {task_of_pair(A, B)});
    mul_analyze(ctx, pl);
    mul_read_counts(ctx, pl);
    mul_symbolic(ctx, ws, pl, true, false);
    const int64_t ncur = mul_cursor_prep(ctx, ws, pl, false);
    const int32_t m = pl.rows_total;
    const int64_t total = read_scalar<int64_t>(ctx, pl.rowptr + m);
    mul_read_counts(ctx, pl);
    if (pl.h.overflow & 1) fail(TR_ERR_HIP, "spgemm: heavy work list capacity exceeded (internal bound violated)");

    // panel split
    std::vector<int32_t> rb{0};
    std::vector<int64_t> hrp;
    if (total > max_panel_nnz) {
        hrp.resize((size_t)m + 1);
        read_array(ctx, hrp.data(), pl.rowptr, sizeof(int64_t) * ((size_t)m + 1));
        int64_t start = 0;
        for (int32_t r = 0; r < m; ++r) {
            if (hrp[r + 1] - hrp[r] > max_panel_nnz)
                fail(TR_ERR_OVERFLOW, "spgemm: a single row of C exceeds the panel capacity");
            if (hrp[r + 1] - start > max_panel_nnz) {
                rb.push_back(r);
                start = hrp[r];
            }
        }
    }
    rb.push_back(m);
    const int32_t np = (int32_t)rb.size() - 1;
    if (np > max_panels) {
        *npanels = np;
        fail(TR_ERR_OVERFLOW, "spgemm: nnz(C) = " + std::to_string(total) + " needs " + std::to_string(np) +
                                  " panels of at most " + std::to_string(max_panel_nnz) + " entries");
    }
    std::vector<Panel> hp(np);
    for (int32_t p = 0; p < np; ++p) {
        const int64_t nb0 = np == 1 ? 0 : hrp[rb[p]];
        const int64_t nb1 = np == 1 ? total : hrp[rb[p + 1]];
        tr_csx_t *C = &panels[p];
        alloc_tile(ctx, TR_ROWS, rb[p + 1] - rb[p], pl.h_tasks[0].n, nb1 - nb0, C);
        launch_rowptr_i64_to_i32(ctx, pl.rowptr, rb[p], rb[p + 1], C->ptr_vec);
        hp[p].nnz_begin = nb0;
        hp[p].row_begin = rb[p];
        hp[p].pad = 0;
        hp[p].idx = C->idx_vec;
        hp[p].val = C->val;
        row_begin[p] = rb[p];
    }
    row_begin[np] = m;
    *npanels = np;
    Panel *dp = (Panel *)ws.get(S_TMP_CSX, sizeof(Panel) * (size_t)np);
    TR_HIP(hipMemcpyAsync(dp, hp.data(), sizeof(Panel) * (size_t)np, hipMemcpyHostToDevice, ctx->stream));
    if (pl.h.items > 0)
        pl.cursors = (int4 *)ws.get(S_CURSORS, sizeof(int4) * (size_t)std::max<int64_t>(ncur, 1));
    mul_numeric(ctx, pl, dp, np);
    mul_check(ctx, pl);
    return TR_OK;
}

}  // namespace tr
