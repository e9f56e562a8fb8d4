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
