// Gram-form building blocks shared by the local join (fk_gram.hip) and the
// Gram-form prune (fk_prune_gram.hip): 16-bit operand conversion, one
// 16 x 16 x 32 MFMA step, a 16 x 16 Gram tile from FP32 rows, and the error
// model bounding the canonical FP32 distance from the Gram-form value.
#pragma once
#include "fk_device.h"
#include "fk_internal.h"

namespace fk {

typedef short fk_s16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 fk_bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 fk_f16x8 __attribute__((ext_vector_type(8)));
typedef float fk_f32x4 __attribute__((ext_vector_type(4)));

// BF16 bits of x, rounded to nearest even (finite inputs).
__device__ __forceinline__ uint32_t bf16_bits(float x) {
    const uint32_t b = __float_as_uint(x);
    return (b + 0x7FFFu + ((b >> 16) & 1u)) >> 16;
}

template <int FMT>
__device__ __forceinline__ float quantize(float x) {
    if constexpr (FMT == FK_GRAM_F16) return (float)(_Float16)x;
    else return __uint_as_float(bf16_bits(x) << 16);
}

// One K-step of a 16 x 16 x 32 tile (the gfx950 16-bit MFMA shape): a, b
// are this lane's 8 consecutive values of its A row and B row (already
// scaled), converted here to FP16 or to BF16 rounded to nearest even.
template <int FMT>
__device__ __forceinline__ fk_f32x4 mfma_step(const float (&a)[8], const float (&b)[8], fk_f32x4 acc) {
    if constexpr (FMT == FK_GRAM_F16) {
        fk_f16x8 ha, hb;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            ha[i] = (_Float16)a[i];
            hb[i] = (_Float16)b[i];
        }
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(ha, hb, acc, 0, 0, 0);
    } else {
        fk_s16x8 sa, sb;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            sa[i] = (short)bf16_bits(a[i]);
            sb[i] = (short)bf16_bits(b[i]);
        }
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(fk_bf16x8, sa),
                                                       __builtin_bit_cast(fk_bf16x8, sb), acc, 0, 0, 0);
    }
}

// 16 x 16 Gram tile.  Lane l supplies row (l & 15) of the A block (arow) and
// of the B block (brow); it reads the values at k = 32 s + 8 (l >> 4) .. +7
// for every step s, zero past dpad (a multiple of 8).  On return lane l
// holds G[4 (l >> 4) + e][l & 15], e = 0..3 (the MFMA output layout), in
// the scaled units (scale^2 G).  Every lane of the wave must call it.
template <int FMT>
__device__ __forceinline__ fk_f32x4 gram_tile(const float *__restrict__ arow, const float *__restrict__ brow,
                                              uint32_t dpad, float scale) {
    fk_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
    const uint32_t kq = (lane_id() >> 4) * 8u;
    const uint32_t steps = (dpad + 31u) >> 5;
    for (uint32_t s = 0; s < steps; ++s) {
        const uint32_t k = s * 32u + kq;
        float a[8], b[8];
        if (k < dpad) {
            const float4 a0 = *reinterpret_cast<const float4 *>(arow + k);
            const float4 a1 = *reinterpret_cast<const float4 *>(arow + k + 4);
            const float4 b0 = *reinterpret_cast<const float4 *>(brow + k);
            const float4 b1 = *reinterpret_cast<const float4 *>(brow + k + 4);
            a[0] = a0.x; a[1] = a0.y; a[2] = a0.z; a[3] = a0.w;
            a[4] = a1.x; a[5] = a1.y; a[6] = a1.z; a[7] = a1.w;
            b[0] = b0.x; b[1] = b0.y; b[2] = b0.z; b[3] = b0.w;
            b[4] = b1.x; b[5] = b1.y; b[6] = b1.z; b[7] = b1.w;
        } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) a[i] = b[i] = 0.f;
        }
        if constexpr (FMT == FK_GRAM_F16) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                a[i] *= scale;
                b[i] *= scale;
            }
        }
        acc = mfma_step<FMT>(a, b, acc);
    }
    return acc;
}

// One K-step of a 16 x 16 x 32 tile from operands already in the 16-bit
// format (the values mfma_step converts to), as raw bits.
template <int FMT>
__device__ __forceinline__ fk_f32x4 mfma_step16(fk_s16x8 a, fk_s16x8 b, fk_f32x4 acc) {
    if constexpr (FMT == FK_GRAM_F16)
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(fk_f16x8, a), __builtin_bit_cast(fk_f16x8, b),
                                                      acc, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(fk_bf16x8, a),
                                                       __builtin_bit_cast(fk_bf16x8, b), acc, 0, 0, 0);
}

// Error model of one prepared dataset (see the file comment).
struct GramBound {
    int exact;        // d~ is the canonical distance
    float inv_s2;     // 1 / scale^2 (exact power of two)
    float u;          // relative rounding of a row, with slack
    float e_abs;      // absolute rounding of a row (FP16 underflow), unscaled
    float f_rel;      // FP32 evaluation of d~, relative to |x~|^2 + |y~|^2
    float gam;        // canonical FP32 distance, relative to a^2
    float ymax;       // largest row norm |x|, rounded up
};

// Lower bound on the canonical distance given d~ (0 when nothing is known).
__device__ __forceinline__ float gram_lower(float dt, float2 nx, float2 ny, const GramBound &g) {
    const float f = g.f_rel * (nx.x + ny.x);
    const float e = g.u * (nx.y + ny.y) + g.e_abs;
    const float lo = __builtin_sqrtf(fmaxf(dt - f, 0.f)) - e;
    return lo > 0.f ? lo * lo * (1.f - g.gam) : 0.f;
}
__device__ __forceinline__ float gram_upper(float dt, float2 nx, float2 ny, const GramBound &g) {
    const float f = g.f_rel * (nx.x + ny.x);
    const float e = g.u * (nx.y + ny.y) + g.e_abs;
    const float hi = __builtin_sqrtf(dt + f) + e;
    return hi * hi * (1.f + g.gam);
}


// The error model of the prepared dataset (its Gram format and scale).
GramBound gram_bound(const Dataset &ds);

}  // namespace fk
