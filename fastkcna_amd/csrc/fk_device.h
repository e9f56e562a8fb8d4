// Device-side building blocks shared by every kernel of the engine.
//
// Pool entries are packed 64-bit keys: the float32 squared distance's bit
// pattern in the high word and the node id in the low word.  Squared L2 is
// never negative, so the unsigned order of the key is the (dist, id) order
// that the engine uses everywhere as its canonical tie order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FK_WAVE 64
#define FK_KEY_EMPTY 0xFFFFFFFFFFFFFFFFull

__device__ __forceinline__ uint32_t key_id(uint64_t k) { return (uint32_t)k; }
__device__ __forceinline__ float key_dist(uint64_t k) { return __uint_as_float((uint32_t)(k >> 32)); }
__device__ __forceinline__ uint64_t make_key(float d, uint32_t id) {
    return ((uint64_t)__float_as_uint(d) << 32) | (uint64_t)id;
}

// One 8-float step of the canonical squared-L2: lane k of the reference's
// 256-bit accumulator gets fma(a[k]-b[k], a[k]-b[k], acc[k]).  g++ -O3
// -march=native contracts the sub/mul/add macro of src/metric.cpp:37-42 into
// exactly this fused form.
__device__ __forceinline__ void l2_step8(const float *__restrict__ a, const float *__restrict__ b,
                                         float acc[8]) {
    const float4 a0 = *reinterpret_cast<const float4 *>(a);
    const float4 a1 = *reinterpret_cast<const float4 *>(a + 4);
    const float4 b0 = *reinterpret_cast<const float4 *>(b);
    const float4 b1 = *reinterpret_cast<const float4 *>(b + 4);
    float t;
    t = a0.x - b0.x; acc[0] = __builtin_fmaf(t, t, acc[0]);
    t = a0.y - b0.y; acc[1] = __builtin_fmaf(t, t, acc[1]);
    t = a0.z - b0.z; acc[2] = __builtin_fmaf(t, t, acc[2]);
    t = a0.w - b0.w; acc[3] = __builtin_fmaf(t, t, acc[3]);
    t = a1.x - b1.x; acc[4] = __builtin_fmaf(t, t, acc[4]);
    t = a1.y - b1.y; acc[5] = __builtin_fmaf(t, t, acc[5]);
    t = a1.z - b1.z; acc[6] = __builtin_fmaf(t, t, acc[6]);
    t = a1.w - b1.w; acc[7] = __builtin_fmaf(t, t, acc[7]);
}

// Canonical squared L2 of two zero-padded rows of dpad floats (dpad % 8 == 0,
// 16-byte aligned), in the exact order of float_l2sqr_avx
// (src/metric.cpp:45-78): the dpad % 32 tail blocks first (offsets +16, +8,
// +0 past the last full 32-block), then every 8-block from 0 upward, each
// into 8 lane accumulators; then the lanes summed left to right.  The result
// is bit-identical to the reference's distance and symmetric in (a, b).
__device__ __forceinline__ float l2_canon(const float *__restrict__ a, const float *__restrict__ b,
                                          uint32_t dpad) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const uint32_t dr = dpad & 31u;
    const uint32_t dd = dpad - dr;
    if (dr == 24u) {
        l2_step8(a + dd + 16, b + dd + 16, acc);
        l2_step8(a + dd + 8, b + dd + 8, acc);
        l2_step8(a + dd, b + dd, acc);
    } else if (dr == 16u) {
        l2_step8(a + dd + 8, b + dd + 8, acc);
        l2_step8(a + dd, b + dd, acc);
    } else if (dr == 8u) {
        l2_step8(a + dd, b + dd, acc);
    }
    for (uint32_t i = 0; i < dd; i += 8) l2_step8(a + i, b + i, acc);
    float r = acc[0] + acc[1];
    r += acc[2];
    r += acc[3];
    r += acc[4];
    r += acc[5];
    r += acc[6];
    r += acc[7];
    return r;
}

// Canonical squared L2 with one operand held in registers (DP floats,
// compile-time): identical order and result to l2_canon.
template <int DP>
__device__ __forceinline__ void l2_step8_reg(const float4 &a0, const float4 &a1, const float *__restrict__ b,
                                             float acc[8]) {
    const float4 b0 = *reinterpret_cast<const float4 *>(b);
    const float4 b1 = *reinterpret_cast<const float4 *>(b + 4);
    float t;
    t = a0.x - b0.x; acc[0] = __builtin_fmaf(t, t, acc[0]);
    t = a0.y - b0.y; acc[1] = __builtin_fmaf(t, t, acc[1]);
    t = a0.z - b0.z; acc[2] = __builtin_fmaf(t, t, acc[2]);
    t = a0.w - b0.w; acc[3] = __builtin_fmaf(t, t, acc[3]);
    t = a1.x - b1.x; acc[4] = __builtin_fmaf(t, t, acc[4]);
    t = a1.y - b1.y; acc[5] = __builtin_fmaf(t, t, acc[5]);
    t = a1.z - b1.z; acc[6] = __builtin_fmaf(t, t, acc[6]);
    t = a1.w - b1.w; acc[7] = __builtin_fmaf(t, t, acc[7]);
}

template <int DP>
__device__ __forceinline__ float l2_canon_reg(const float4 (&a)[DP / 4], const float *__restrict__ b) {
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    constexpr int dr = DP & 31;
    constexpr int dd = DP - dr;
    if (dr == 24) {
        l2_step8_reg<DP>(a[(dd + 16) / 4], a[(dd + 16) / 4 + 1], b + dd + 16, acc);
        l2_step8_reg<DP>(a[(dd + 8) / 4], a[(dd + 8) / 4 + 1], b + dd + 8, acc);
        l2_step8_reg<DP>(a[dd / 4], a[dd / 4 + 1], b + dd, acc);
    } else if (dr == 16) {
        l2_step8_reg<DP>(a[(dd + 8) / 4], a[(dd + 8) / 4 + 1], b + dd + 8, acc);
        l2_step8_reg<DP>(a[dd / 4], a[dd / 4 + 1], b + dd, acc);
    } else if (dr == 8) {
        l2_step8_reg<DP>(a[dd / 4], a[dd / 4 + 1], b + dd, acc);
    }
#pragma unroll
    for (int i = 0; i < dd; i += 8) l2_step8_reg<DP>(a[i / 4], a[i / 4 + 1], b + i, acc);
    float r = acc[0] + acc[1];
    r += acc[2];
    r += acc[3];
    r += acc[4];
    r += acc[5];
    r += acc[6];
    r += acc[7];
    return r;
}

// Octet (8-lane) form of the canonical squared L2.
// The 8 lane accumulators of float_l2sqr_avx are independent fma chains over
// the 8-blocks in canonical order (the dpad % 32 tail blocks from the last
// down, then the full blocks ascending); lane k of an octet runs chain k.
// The chain-major copy of the dataset keeps, per row, each chain's
// nb = dpad / 8 elements contiguous in that order:
//     chains[i * dpad + k * nb + j] = row_i[B_j + k],  B_j = j-th block.
// An octet then reads a row as 8 coalesced runs, and the result, summed
// lane 0 to lane 7 left to right, is bit-identical to l2_canon.
__device__ __forceinline__ uint32_t chain_block(uint32_t j, uint32_t dpad) {
    const uint32_t dr = dpad & 31u, dd = dpad - dr, ntail = dr >> 3;
    return j < ntail ? dd + 8u * (ntail - 1u - j) : 8u * (j - ntail);
}

// Ordered sum of the 8 accumulators of this lane's octet, in every lane of
// the octet.  The 8 lanes must be active together.
__device__ __forceinline__ float octet_sum(float acc) {
    // every lane adds its own partial to its lower neighbour's running sum
    // (DPP row_shr:1 within the lane's row of 16): after step j lane L
    // holds the left-to-right sum of the partials of lanes L - j .. L, so
    // after seven steps lane 7 of the octet holds ((p0 + p1) + p2) + ... +
    // p7, the same additions in the same order (what lanes 0..6 and the
    // lanes reading across an octet boundary hold is not used).  Then lane
    // 7's value to the whole octet: mirror the 8 lanes (lane 0 gets it),
    // broadcast lane 0 of each quad, and lanes 4..7 take lanes 0..3's
    // (row_shr:4)
    const uint32_t k = threadIdx.x & 7u;
    float s = acc;
#pragma unroll
    for (uint32_t i = 1; i < 8; ++i)
        s = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0x111, 0xF, 0xF, true)) + acc;
    const int m = __builtin_amdgcn_mov_dpp(__float_as_int(s), 0x141, 0xF, 0xF, true);  // row_half_mirror
    const int q = __builtin_amdgcn_mov_dpp(m, 0x00, 0xF, 0xF, true);                    // quad_perm [0,0,0,0]
    const int h = __builtin_amdgcn_mov_dpp(q, 0x114, 0xF, 0xF, true);                   // row_shr:4
    return __int_as_float(k < 4 ? q : h);
}

// Sum over the 8 lanes of this lane's octet of 32-bit integers (exact),
// as a float (every sum here is below 2^24).
// Integer sums are exact in any order: a butterfly of three DPP steps
// (swap lanes 1 apart and 2 apart within quads, then mirror the 8 lanes)
// leaves the octet's total in all of its lanes.
__device__ __forceinline__ float octet_sum_u32(uint32_t acc) {
    uint32_t r = acc;
    r += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    r += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    r += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)r, 0x141, 0xF, 0xF, false);  // row_half_mirror
    return (float)r;
}

// One lane's chain of one row, in registers (NB = dpad / 8, a multiple of
// 4) or read from memory per distance (NB == 0).
template <int NB>
struct OctChain {
    float4 v[NB / 4];
    __device__ __forceinline__ void load(const float *__restrict__ c, uint32_t) {
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) v[q] = reinterpret_cast<const float4 *>(c)[q];
    }
    __device__ __forceinline__ void store(float *__restrict__ c, uint32_t) const {
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) reinterpret_cast<float4 *>(c)[q] = v[q];
    }
    // distance to another row held in registers
    __device__ __forceinline__ float dist_reg(const OctChain &o, uint32_t) const {
        float acc = 0.f, t;
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) {
            t = v[q].x - o.v[q].x; acc = __builtin_fmaf(t, t, acc);
            t = v[q].y - o.v[q].y; acc = __builtin_fmaf(t, t, acc);
            t = v[q].z - o.v[q].z; acc = __builtin_fmaf(t, t, acc);
            t = v[q].w - o.v[q].w; acc = __builtin_fmaf(t, t, acc);
        }
        return octet_sum(acc);
    }
    // distance to the row whose chain (same lane) starts at o
    __device__ __forceinline__ float dist(const float *__restrict__ o, uint32_t) const {
        float acc = 0.f, t;
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) {
            const float4 b = reinterpret_cast<const float4 *>(o)[q];
            t = v[q].x - b.x; acc = __builtin_fmaf(t, t, acc);
            t = v[q].y - b.y; acc = __builtin_fmaf(t, t, acc);
            t = v[q].z - b.z; acc = __builtin_fmaf(t, t, acc);
            t = v[q].w - b.w; acc = __builtin_fmaf(t, t, acc);
        }
        return octet_sum(acc);
    }
    // Half-precision copy of the chain (round to nearest even), NB / 4
    // 8-byte stores; the chain of lane k starts k * NB halves in (8-byte
    // aligned when NB % 4 == 0).
    __device__ __forceinline__ void store_h(_Float16 *__restrict__ c) const {
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) {
            const _Float16 h[4] = {(_Float16)v[q].x, (_Float16)v[q].y, (_Float16)v[q].z, (_Float16)v[q].w};
            reinterpret_cast<uint2 *>(c)[q] = *reinterpret_cast<const uint2 *>(h);
        }
    }
    // Squared distance to a half-precision chain, the halves widened exactly
    // to FP32: an approximation of the distance to the row the halves were
    // made from (any summation order; used with an error bound only).
    __device__ __forceinline__ float dist_h(const _Float16 *__restrict__ o) const {
        float acc = 0.f, t;
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) {
            const uint2 w = reinterpret_cast<const uint2 *>(o)[q];
            const _Float16 *h = reinterpret_cast<const _Float16 *>(&w);
            t = v[q].x - (float)h[0]; acc = __builtin_fmaf(t, t, acc);
            t = v[q].y - (float)h[1]; acc = __builtin_fmaf(t, t, acc);
            t = v[q].z - (float)h[2]; acc = __builtin_fmaf(t, t, acc);
            t = v[q].w - (float)h[3]; acc = __builtin_fmaf(t, t, acc);
        }
        return octet_sum(acc);
    }
    // Squared norm of the row (FP32, any order)
    __device__ __forceinline__ float norm2() const {
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) {
            acc = __builtin_fmaf(v[q].x, v[q].x, acc);
            acc = __builtin_fmaf(v[q].y, v[q].y, acc);
            acc = __builtin_fmaf(v[q].z, v[q].z, acc);
            acc = __builtin_fmaf(v[q].w, v[q].w, acc);
        }
        return octet_sum(acc);
    }
    // dot products (the exact form: see exact_dist)
    __device__ __forceinline__ float dot_reg(const OctChain &o, uint32_t) const {
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) {
            acc = __builtin_fmaf(v[q].x, o.v[q].x, acc);
            acc = __builtin_fmaf(v[q].y, o.v[q].y, acc);
            acc = __builtin_fmaf(v[q].z, o.v[q].z, acc);
            acc = __builtin_fmaf(v[q].w, o.v[q].w, acc);
        }
        return octet_sum(acc);
    }
    __device__ __forceinline__ float dot(const float *__restrict__ o, uint32_t) const {
        float acc = 0.f;
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) {
            const float4 b = reinterpret_cast<const float4 *>(o)[q];
            acc = __builtin_fmaf(v[q].x, b.x, acc);
            acc = __builtin_fmaf(v[q].y, b.y, acc);
            acc = __builtin_fmaf(v[q].z, b.z, acc);
            acc = __builtin_fmaf(v[q].w, b.w, acc);
        }
        return octet_sum(acc);
    }
};
// Chain sums of two chains given by pointers, any length nb: 16-byte
// loads when nb % 4 == 0 (every chain then starts 16-byte aligned), four of
// them in flight ahead of the dependent fma chain; the fma order is the
// element order either way.
__device__ __forceinline__ float chain_l2_acc(const float *__restrict__ a, const float *__restrict__ b, uint32_t nb) {
    float acc = 0.f, t;
    if ((nb & 3u) == 0) {
        const float4 *a4 = reinterpret_cast<const float4 *>(a);
        const float4 *b4 = reinterpret_cast<const float4 *>(b);
#pragma unroll 4
        for (uint32_t q = 0; q < (nb >> 2); ++q) {
            const float4 x = a4[q], y = b4[q];
            t = x.x - y.x; acc = __builtin_fmaf(t, t, acc);
            t = x.y - y.y; acc = __builtin_fmaf(t, t, acc);
            t = x.z - y.z; acc = __builtin_fmaf(t, t, acc);
            t = x.w - y.w; acc = __builtin_fmaf(t, t, acc);
        }
    } else {
        for (uint32_t j = 0; j < nb; ++j) {
            t = a[j] - b[j];
            acc = __builtin_fmaf(t, t, acc);
        }
    }
    return acc;
}
__device__ __forceinline__ float chain_dot_acc(const float *__restrict__ a, const float *__restrict__ b, uint32_t nb) {
    float acc = 0.f;
    if ((nb & 3u) == 0) {
        const float4 *a4 = reinterpret_cast<const float4 *>(a);
        const float4 *b4 = reinterpret_cast<const float4 *>(b);
#pragma unroll 4
        for (uint32_t q = 0; q < (nb >> 2); ++q) {
            const float4 x = a4[q], y = b4[q];
            acc = __builtin_fmaf(x.x, y.x, acc);
            acc = __builtin_fmaf(x.y, y.y, acc);
            acc = __builtin_fmaf(x.z, y.z, acc);
            acc = __builtin_fmaf(x.w, y.w, acc);
        }
    } else {
        for (uint32_t j = 0; j < nb; ++j) acc = __builtin_fmaf(a[j], b[j], acc);
    }
    return acc;
}

template <>
struct OctChain<0> {
    const float *p;
    __device__ __forceinline__ void load(const float *__restrict__ c, uint32_t) { p = c; }
    __device__ __forceinline__ void store(float *__restrict__, uint32_t) const {}
    __device__ __forceinline__ float dist_reg(const OctChain &o, uint32_t nb) const { return dist(o.p, nb); }
    __device__ __forceinline__ float dist(const float *__restrict__ o, uint32_t nb) const {
        return octet_sum(chain_l2_acc(p, o, nb));
    }
    __device__ __forceinline__ float dot_reg(const OctChain &o, uint32_t nb) const { return dot(o.p, nb); }
    __device__ __forceinline__ float dot(const float *__restrict__ o, uint32_t nb) const {
        return octet_sum(chain_dot_acc(p, o, nb));
    }
};

// Byte form of a lane's chain (exact-form rows whose values are integers in
// [0, 255]): NB bytes, read as 16-byte vectors; dot products with
// v_dot4_u32_u8, exact in 32-bit integers.
template <int NB>
struct OctChainU8 {
    static_assert(NB % 16 == 0, "byte chains are read 16 bytes at a time");
    uint4 v[NB / 16];
    __device__ __forceinline__ void load(const uint8_t *__restrict__ c, uint32_t) {
#pragma unroll
        for (int q = 0; q < NB / 16; ++q) v[q] = reinterpret_cast<const uint4 *>(c)[q];
    }
    __device__ __forceinline__ void store(uint8_t *__restrict__ c, uint32_t) const {
#pragma unroll
        for (int q = 0; q < NB / 16; ++q) reinterpret_cast<uint4 *>(c)[q] = v[q];
    }
    __device__ __forceinline__ static uint32_t dot4x4(const uint4 &a, const uint4 &b, uint32_t acc) {
        acc = __builtin_amdgcn_udot4(a.x, b.x, acc, false);
        acc = __builtin_amdgcn_udot4(a.y, b.y, acc, false);
        acc = __builtin_amdgcn_udot4(a.z, b.z, acc, false);
        return __builtin_amdgcn_udot4(a.w, b.w, acc, false);
    }
    __device__ __forceinline__ float dot_reg(const OctChainU8 &o, uint32_t) const {
        uint32_t acc = 0;
#pragma unroll
        for (int q = 0; q < NB / 16; ++q) acc = dot4x4(v[q], o.v[q], acc);
        return octet_sum_u32(acc);
    }
    __device__ __forceinline__ float dot(const uint8_t *__restrict__ o, uint32_t) const {
        uint32_t acc = 0;
#pragma unroll
        for (int q = 0; q < NB / 16; ++q) acc = dot4x4(v[q], reinterpret_cast<const uint4 *>(o)[q], acc);
        return octet_sum_u32(acc);
    }
    // exact squared distance is only taken in the dot form
    __device__ __forceinline__ float dist(const uint8_t *__restrict__, uint32_t) const { return 0.f; }
    __device__ __forceinline__ float dist_reg(const OctChainU8 &, uint32_t) const { return 0.f; }
};

// Exact form of the squared L2 for rows that the Gram preparation marked
// exact (integer values, |x_i| <= 256, small norms): every product, partial
// sum and the result are integers below 2^24, so |a|^2 + |b|^2 - 2 a.b is
// the canonical distance bit for bit, in any summation order, at one fused
// multiply-add per element instead of a subtract and a multiply-add.
__device__ __forceinline__ float exact_dist(float na, float nb, float dot) { return (na + nb) - 2.f * dot; }

// Octet distance between two chain-major rows given by pointers (this
// lane's chain of each); a is the minuend, as in OctChain::dist.
template <int NB>
__device__ __forceinline__ float octet_dist2(const float *__restrict__ a, const float *__restrict__ b, uint32_t nb) {
    float acc = 0.f, t;
    if constexpr (NB > 0) {
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) {
            const float4 x = reinterpret_cast<const float4 *>(a)[q];
            const float4 y = reinterpret_cast<const float4 *>(b)[q];
            t = x.x - y.x; acc = __builtin_fmaf(t, t, acc);
            t = x.y - y.y; acc = __builtin_fmaf(t, t, acc);
            t = x.z - y.z; acc = __builtin_fmaf(t, t, acc);
            t = x.w - y.w; acc = __builtin_fmaf(t, t, acc);
        }
    } else {
        acc = chain_l2_acc(a, b, nb);
    }
    return octet_sum(acc);
}

template <int NB>
__device__ __forceinline__ float octet_dot2(const float *__restrict__ a, const float *__restrict__ b, uint32_t nb) {
    float acc = 0.f;
    if constexpr (NB > 0) {
#pragma unroll
        for (int q = 0; q < NB / 4; ++q) {
            const float4 x = reinterpret_cast<const float4 *>(a)[q];
            const float4 y = reinterpret_cast<const float4 *>(b)[q];
            acc = __builtin_fmaf(x.x, y.x, acc);
            acc = __builtin_fmaf(x.y, y.y, acc);
            acc = __builtin_fmaf(x.z, y.z, acc);
            acc = __builtin_fmaf(x.w, y.w, acc);
        }
    } else {
        acc = chain_dot_acc(a, b, nb);
    }
    return octet_sum(acc);
}

__device__ __forceinline__ float octet_dot2(const uint8_t *__restrict__ a, const uint8_t *__restrict__ b,
                                            uint32_t nb) {
    uint32_t acc = 0;
    for (uint32_t j = 0; j < nb; j += 16)
        acc = OctChainU8<16>::dot4x4(*reinterpret_cast<const uint4 *>(a + j), *reinterpret_cast<const uint4 *>(b + j),
                                     acc);
    return octet_sum_u32(acc);
}

// Octet distance / dot of two chains given by pointers, for either row type.
template <int NB>
__device__ __forceinline__ float chain_dot2(const float *__restrict__ a, const float *__restrict__ b, uint32_t nb) {
    return octet_dot2<NB>(a, b, nb);
}
template <int NB>
__device__ __forceinline__ float chain_dot2(const uint8_t *__restrict__ a, const uint8_t *__restrict__ b,
                                            uint32_t nb) {
    return octet_dot2(a, b, nb);
}
template <int NB>
__device__ __forceinline__ float chain_dist2(const float *__restrict__ a, const float *__restrict__ b, uint32_t nb) {
    return octet_dist2<NB>(a, b, nb);
}
template <int NB>
__device__ __forceinline__ float chain_dist2(const uint8_t *__restrict__, const uint8_t *__restrict__, uint32_t) {
    return 0.f;  // byte rows only take the dot form
}

// Row kinds of the octet kernels: MODE 0 FP32 rows, canonical distance;
// MODE 1 FP32 exact-form rows, dot form; MODE 2 byte exact-form rows, dot
// form (NB a multiple of 16).
template <int NB, int MODE>
struct RowKind {
    using E = float;
    using Chain = OctChain<NB>;
    static constexpr bool dot = MODE != 0;
};
template <int NB>
struct RowKind<NB, 2> {
    using E = uint8_t;
    using Chain = OctChainU8<NB>;
    static constexpr bool dot = true;
};

// Occlusion rules of the three pruner families; dqt is the kept neighbour's
// distance to the pool owner, djk the kept neighbour to candidate distance,
// pd the candidate's distance to the owner.
//   NSG / HNSW layer: src/kgraph.cpp:1836     djk < p.dist
//   tau-MNG:          src/kgraph.cpp:2859     djk < p.dist - 3*tau
//   alpha-PG:         src/kgraph.cpp:3184-3187
//        cos = (dqt + djk - p.dist) / 2 / sqrt(dqt * djk); djk < p.dist && cos < thr
enum { FK_RULE_NSG = 0, FK_RULE_TAU = 1, FK_RULE_ALPHA = 2 };

template <int RULE>
__device__ __forceinline__ bool occludes(float djk, float pd, float dqt, float tau3, float cos_thr) {
    if (RULE == FK_RULE_NSG) return djk < pd;
    if (RULE == FK_RULE_TAU) return djk < pd - tau3;
    float num = dqt + djk;
    num = num - pd;
    num = num / 2.0f;
    const float den = __builtin_sqrtf(dqt * djk);
    const float c = num / den;
    return (djk < pd) && (c < cos_thr);
}

// Wave-wide helpers (wave64).
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint32_t lane_id() { return threadIdx.x & (FK_WAVE - 1); }

// Block-wide sum (blockDim.x a multiple of 64, at most 1024); the result is
// valid in thread 0.  Every thread of the block must call it.
template <typename T>
__device__ __forceinline__ T block_sum(T v) {
    __shared__ T s_part[16];
    v = wave_sum(v);
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) s_part[w] = v;
    __syncthreads();
    T r = 0;
    if (threadIdx.x == 0)
        for (uint32_t i = 0; i < (blockDim.x >> 6); ++i) r += s_part[i];
    __syncthreads();
    return r;
}

// XCD-aware block order.  Workgroups are dispatched to the 8 XCDs round
// robin (block b -> XCD b % 8), each XCD with its own L2; launching
// 8 * per_xcd blocks and giving XCD x the contiguous run
// [x * per_xcd, (x + 1) * per_xcd) of a locality order keeps neighbouring
// work on one L2.  Returns the position in that order.
#define FK_NUM_XCD 8
__device__ __forceinline__ uint32_t xcd_position(uint32_t per_xcd) {
    return (blockIdx.x % FK_NUM_XCD) * per_xcd + blockIdx.x / FK_NUM_XCD;
}

// Counter-based RNG: splitmix64 finaliser over a 64-bit counter.
__device__ __host__ __forceinline__ uint64_t fk_mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __host__ __forceinline__ uint32_t fk_rand32(uint64_t seed, uint64_t stream, uint64_t k) {
    return (uint32_t)(fk_mix64(seed * 0xD1B54A32D192ED03ull ^ fk_mix64(stream * 0x8CB92BA72F3D8DD7ull + k)) >> 32);
}
