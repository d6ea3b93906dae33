// Device-side helpers shared by the gfx950 kernels: wave64 primitives,
// LDS ordering and the sorting network used to emit sorted rows.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tr {
namespace dev {

constexpr int kWave = 64;
constexpr int32_t kEmpty = -1;          // empty hash slot (columns are >= 0)
constexpr int32_t kKeyPad = 0x7fffffff;  // sort padding key

// Global-memory (address space 1) view of a device pointer.  Accesses
// through it are global instructions, which count on vmcnt only; a flat
// access (what a pointer read from memory compiles to) also counts on
// lgkmcnt, so every wait for an LDS result would wait for it as well.
#define TR_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ TR_GLOBAL T *as_global(T *p) {
    return (TR_GLOBAL T *)p;
}

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Order this wave's LDS accesses across the call (and wait for outstanding
// LDS operations); no cross-wave synchronisation.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t ballot(bool p) { return (uint64_t)__ballot(p ? 1 : 0); }

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ int rank_below(uint64_t mask) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, kWave);
        v = v > w ? v : w;
    }
    return v;
}

// Inclusive prefix sum over the wave.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int l = lane_id();
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T w = __shfl_up(v, o, kWave);
        if (l >= o) v += w;
    }
    return v;
}

// Multiplicative hash of a column index into a table of 2^log2T slots.
__device__ __forceinline__ uint32_t hash_slot(int32_t key, uint32_t log2T) {
    return ((uint32_t)key * 0x9E3779B1u) >> (32u - log2T);
}

// Largest lane l with excl[l] <= x, given non-decreasing per-lane `excl`
// (excl[0] == 0).  Binary search over lanes through shuffles.
__device__ __forceinline__ int owner_lane(int64_t excl, int64_t x) {
    int lo = 0;
#pragma unroll
    for (int step = 32; step > 0; step >>= 1) {
        int probe = lo + step;
        int64_t v = __shfl(excl, probe & 63, kWave);
        if (probe < kWave && v <= x) lo = probe;
    }
    return lo;
}

// Sort keys[0..n) (with vals) ascending in LDS by one wave.  The table
// region must hold at least N = max(64, next_pow2(n)) entries; entries
// [n, N) are overwritten with padding.
template <typename V>
__device__ void wave_sort_lds(int32_t *keys, V *vals, int n) {
    const int l = lane_id();
    int N = 64;
    while (N < n) N <<= 1;
    for (int t = n + l; t < N; t += kWave) keys[t] = kKeyPad;
    wave_sync();
    if (N == 64) {
        // register bitonic sort, one entry per lane
        int32_t k = keys[l];
        V v = vals[l];
#pragma unroll
        for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
            for (int stride = size >> 1; stride > 0; stride >>= 1) {
                int32_t pk = __shfl_xor(k, stride, kWave);
                V pv = __shfl_xor(v, stride, kWave);
                bool up = ((l & size) == 0);
                bool lower = ((l & stride) == 0);
                bool take_min = (lower == up);
                bool take_partner = take_min ? (pk < k) : (pk > k);
                if (take_partner) {
                    k = pk;
                    v = pv;
                }
            }
        }
        wave_sync();
        keys[l] = k;
        vals[l] = v;
        wave_sync();
        return;
    }
    for (int size = 2; size <= N; size <<= 1) {
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = l; i < (N >> 1); i += kWave) {
                int lo = ((i & ~(stride - 1)) << 1) | (i & (stride - 1));
                int hi = lo + stride;
                bool up = ((lo & size) == 0);
                int32_t ka = keys[lo], kb = keys[hi];
                if ((ka > kb) == up) {
                    V va = vals[lo], vb = vals[hi];
                    keys[lo] = kb;
                    keys[hi] = ka;
                    vals[lo] = vb;
                    vals[hi] = va;
                }
            }
            wave_sync();
        }
    }
}

// Move the occupied slots of a hash table to its front, preserving nothing
// about order.  Returns the number of occupied slots (wave-uniform).
template <typename V>
__device__ int wave_compact_lds(int32_t *keys, V *vals, int T) {
    const int l = lane_id();
    int n = 0;
    for (int base = 0; base < T; base += kWave) {
        int32_t k = keys[base + l];
        V v = vals[base + l];
        bool occ = (k != kEmpty);
        uint64_t m = ballot(occ);
        int r = rank_below(m);
        wave_sync();
        if (occ) {
            keys[n + r] = k;
            vals[n + r] = v;
        }
        wave_sync();
        n += __popcll(m);
    }
    return n;
}

}  // namespace dev
}  // namespace tr
