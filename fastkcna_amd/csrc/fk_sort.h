// Block-level sorting and scanning helpers (LDS resident).
#pragma once
#include "fk_device.h"

// Bitonic sort of n keys (n a power of two, n <= blockDim.x * per-thread
// share) held in LDS, ascending.  Every thread of the block must call it.
template <typename K>
__device__ void block_bitonic_sort(K *s, uint32_t n) {
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const K a = s[i], b = s[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) { s[i] = b; s[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
}

// Bitonic sort of n 64-bit keys in LDS (n a power of two, 2 <= n <= 512)
// by a 256-thread block, ascending, with the same comparisons as
// block_bitonic_sort.  Each wave keeps 128 consecutive keys in registers
// (two per lane) and takes every step whose partner lies in the same wave
// with lane shuffles; only the steps with partner distance 128 or 256 go
// through LDS, so the sort has at most three block barriers per pass
// instead of one per step.  Every thread of the block must call it.
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, uint32_t m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, (int)m, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), (int)m, 64);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline void block_sort512(uint64_t *s, uint32_t n) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t i0 = w * 128 + lane * 2;
    const bool act = i0 < n;
    uint64_t a0 = 0, a1 = 0;
    if (act) { a0 = s[i0]; a1 = s[i0 + 1]; }
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const bool up = (i0 & k) == 0;  // the same for i0 + 1 (k >= 2)
            if (j >= 128) {
                __syncthreads();
                if (act) { s[i0] = a0; s[i0 + 1] = a1; }
                __syncthreads();
                if (act) {
                    const uint64_t b0 = s[i0 ^ j], b1 = s[(i0 + 1) ^ j];
                    const bool lo = (i0 & j) == 0;  // the same for i0 + 1
                    a0 = (lo == up) ? (a0 < b0 ? a0 : b0) : (a0 < b0 ? b0 : a0);
                    a1 = (lo == up) ? (a1 < b1 ? a1 : b1) : (a1 < b1 ? b1 : a1);
                }
            } else if (j == 1) {
                if ((a0 > a1) == up) { const uint64_t t = a0; a0 = a1; a1 = t; }
            } else {
                const uint64_t b0 = shfl_xor_u64(a0, j >> 1), b1 = shfl_xor_u64(a1, j >> 1);
                const bool lo = (i0 & j) == 0;
                a0 = (lo == up) ? (a0 < b0 ? a0 : b0) : (a0 < b0 ? b0 : a0);
                a1 = (lo == up) ? (a1 < b1 ? a1 : b1) : (a1 < b1 ? b1 : a1);
            }
        }
    }
    __syncthreads();
    if (act) { s[i0] = a0; s[i0 + 1] = a1; }
    __syncthreads();
}

// Bitonic sort of n 64-bit keys in LDS (n a power of two, 2 <= n <= 128)
// by one wavefront, ascending: two keys per lane in registers, every step by
// lane shuffles.  The caller orders the LDS accesses around it (every lane's
// earlier stores before, the stores it makes before later reads).
__device__ inline void wave_sort128(uint64_t *s, uint32_t n) {
    const uint32_t i0 = (threadIdx.x & 63) * 2;
    const bool act = i0 < n;
    uint64_t a0 = 0, a1 = 0;
    if (act) { a0 = s[i0]; a1 = s[i0 + 1]; }
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const bool up = (i0 & k) == 0;
            if (j == 1) {
                if ((a0 > a1) == up) { const uint64_t t = a0; a0 = a1; a1 = t; }
            } else {
                const uint64_t b0 = shfl_xor_u64(a0, j >> 1), b1 = shfl_xor_u64(a1, j >> 1);
                const bool lo = (i0 & j) == 0;
                a0 = (lo == up) ? (a0 < b0 ? a0 : b0) : (a0 < b0 ? b0 : a0);
                a1 = (lo == up) ? (a1 < b1 ? a1 : b1) : (a1 < b1 ? b1 : a1);
            }
        }
    }
    if (act) { s[i0] = a0; s[i0 + 1] = a1; }
}

// Same, sorting (key, value) pairs by key.
template <typename K, typename V>
__device__ void block_bitonic_sort_kv(K *s, V *v, uint32_t n) {
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const K a = s[i], b = s[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        s[i] = b; s[ixj] = a;
                        const V t = v[i]; v[i] = v[ixj]; v[ixj] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Wave-level sort of up to 64 keys, one per lane (bitonic through shuffles);
// lanes >= n must hold the maximal key.
__device__ __forceinline__ uint64_t wave_sort64(uint64_t v) {
    const uint32_t l = lane_id();
    for (uint32_t k = 2; k <= 64; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint64_t o = __shfl_xor(v, j, 64);
            const bool up = (l & k) == 0;
            const bool lower = (l & j) == 0;
            const uint64_t mn = v < o ? v : o, mx = v < o ? o : v;
            v = (lower == up) ? mn : mx;
        }
    }
    return v;
}

__device__ __forceinline__ uint32_t next_pow2(uint32_t x) {
    uint32_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

// Number of keys strictly less than x in sorted a[0..n).
template <typename K>
__device__ __forceinline__ uint32_t lower_bound_cnt(const K *a, uint32_t n, K x) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}
