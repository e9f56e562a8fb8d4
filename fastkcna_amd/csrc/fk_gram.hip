// Local join in Gram form on the MFMA units, with an FP32 recheck.
//
// The local join (src/kgraph.cpp:970-987) needs, per node, the distance of
// every (new, new) and (new, old) pair of its sampled list.  Here a node's
// pairs are 16 x 16 tiles of the Gram matrix G = X Y^T, computed by
// v_mfma_f32_16x16x32 on 16-bit operands (BF16, or FP16 for unit-normalised
// rows) with FP32 accumulation, and
//     d~(x, y) = |x~|^2 + |y~|^2 - 2 G(x, y)
// with FP32 norms of the rounded rows.  Pools only ever hold exact FP32
// distances (the canonical order of src/metric.cpp:45-78):
//   * exact form -- integer-valued rows with |x_i| <= 256 and small norms
//     (the integer mixtures): the 16-bit operands are the values themselves
//     and every sum is an integer below 2^24, so d~ IS the canonical
//     distance, bit for bit, and no recheck is needed;
//   * bounded forms -- every pair whose canonical distance could be at or
//     below either endpoint's pool threshold (the join's filter) is
//     recomputed in FP32 from the staged rows; the others can neither enter
//     a pool nor mark a join, so the pools come out identical to the plain
//     FP32 join's.  The bound is rigorous:
//       |x~ - x| <= u |x| per row (u = 2^-8 BF16, 2^-11 FP16, round to
//       nearest), so the exact distance a of x, y and a~ of x~, y~ satisfy
//       |a - a~| <= e = u (|x| + |y|) (+ an FP16 underflow term);
//       the FP32 evaluation of d~ is within f = f_rel (|x~|^2 + |y~|^2) of
//       a~^2; and the canonical FP32 distance c is within a factor
//       (1 +- gam) of a^2.  A pair is skipped only when
//       (sqrt(max(d~ - f, 0)) - e)^2 (1 - gam) > max(thr_x, thr_y),
//     which implies c > both thresholds.
// Exact-form rows with every value in [0, 255] are joined from their byte
// copy instead (k_join_gram8: v_mfma_i32_16x16x64_i8 on the values shifted by
// -128, integer norms), with the same pairs, distances and counts.
#include <math.h>

#include <algorithm>

#include "fk_build.h"
#include "fk_device.h"
#include "fk_internal.h"
#include "fk_gram.h"
#include "fk_join.h"

namespace fk {

// ---------------------------------------------------------------------------
// Dataset statistics and per-row norms.
// ---------------------------------------------------------------------------
// out[0] AND of "row is integer-valued", out[1] max |x_i| (float bits),
// out[2] OR of "row has a negative value", out[3] max |x|^2, out[4] min
// |x|^2 (float bits, rounded outward).
__global__ __launch_bounds__(256) void k_gram_stats(const float *__restrict__ rows, uint32_t stride, uint32_t d,
                                                    uint32_t n, uint32_t *__restrict__ out) {
    uint32_t all_int = 1, neg = 0;
    float mabs = 0.f, nmax = 0.f, nmin = INFINITY;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float *r = rows + (size_t)i * stride;
        double s = 0.0;
        for (uint32_t k = 0; k < d; ++k) {
            const float v = r[k];
            all_int &= (v == rintf(v)) ? 1u : 0u;
            neg |= v < 0.f ? 1u : 0u;
            mabs = fmaxf(mabs, fabsf(v));
            s += (double)v * (double)v;
        }
        nmax = fmaxf(nmax, (float)(s * (1.0 + 1e-7)));
        nmin = fminf(nmin, (float)(s * (1.0 - 1e-7)));
    }
    atomicAnd(&out[0], all_int);
    atomicMax(&out[1], __float_as_uint(mabs));
    atomicOr(&out[2], neg);
    atomicMax(&out[3], __float_as_uint(nmax));
    atomicMin(&out[4], __float_as_uint(nmin));
}

// nrm[i] = (|x~|^2 in FP32 (unscaled), |x| rounded up).
template <int FMT>
__global__ __launch_bounds__(256) void k_gram_norms(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                    uint32_t n, float scale, float inv_s2, float2 *__restrict__ nrm) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = rows + (size_t)i * stride;
    float q2 = 0.f;
    double s = 0.0;
    for (uint32_t k = 0; k < dpad; ++k) {
        const float v = r[k];
        const float q = quantize<FMT>(FMT == FK_GRAM_F16 ? v * scale : v);
        q2 = q2 + q * q;
        s += (double)v * (double)v;
    }
    nrm[i] = make_float2(q2 * inv_s2, (float)(sqrt(s) * (1.0 + 1e-6)));
}

// norms8[i] = sum over the dpad bytes b of row i of (b - 128)^2.
__global__ void k_norms8(const uint8_t *__restrict__ c8, uint32_t n, uint32_t dpad, int32_t *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint8_t *r = c8 + (size_t)i * dpad;
    int32_t s = 0;
    for (uint32_t k = 0; k < dpad; ++k) {
        const int32_t v = (int32_t)r[k] - 128;
        s += v * v;
    }
    out[i] = s;
}

// rows16 = the 16-bit operands of every row, dpad wide: FP16 of the
// prescaled value, or BF16 rounded to nearest even (the values mfma_step
// makes from the FP32 row).  One wave per row.
template <int FMT>
__global__ __launch_bounds__(256) void k_rows16(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                uint32_t n, float scale, uint16_t *__restrict__ out) {
    const uint32_t lane = lane_id();
    for (uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6); i < n; i += gridDim.x * 4u) {
        const float *r = rows + (size_t)i * stride;
        uint16_t *o = out + (size_t)i * dpad;
        for (uint32_t k = lane; k < dpad; k += 64) {
            const float v = r[k];
            if constexpr (FMT == FK_GRAM_F16) o[k] = __builtin_bit_cast(uint16_t, (_Float16)(v * scale));
            else o[k] = (uint16_t)bf16_bits(v);
        }
    }
}

// rows8q: each row of width D (96 or 256) as 8-bit integers q with one
// scale s = max |x| / 127 (q = x / s rounded, so |x - s q| <= s / 2 per
// value), then s, the error norm |x - s q| (exact products and differences
// in double, rounded up) and |s q|^2 (s^2 sum q^2 in double, rounded once)
// as FP32, the rest of the row's lines zero: 128 bytes (one cache line) at
// width 96, 384 at width 256.  One wave per row, lane l holds values l + 64 j.
template <int D>
__global__ __launch_bounds__(256) void k_rows8q(const float *__restrict__ rows, uint32_t stride, uint32_t n,
                                                uint4 *__restrict__ out) {
    constexpr uint32_t J = (D + 63) / 64, RB = D == 96 ? 128u : 384u;
    const uint32_t lane = lane_id();
    for (uint32_t i = blockIdx.x * 4u + (threadIdx.x >> 6); i < n; i += gridDim.x * 4u) {
        const float *r = rows + (size_t)i * stride;
        float x[J];
        float m = 0.f;
#pragma unroll
        for (uint32_t j = 0; j < J; ++j) {
            x[j] = lane + 64u * j < (uint32_t)D ? r[lane + 64u * j] : 0.f;
            m = fmaxf(m, fabsf(x[j]));
        }
        for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        const float s = m / 127.f;
        double err2 = 0.0;
        int qq = 0;
        int8_t *ob = reinterpret_cast<int8_t *>(out + (size_t)i * (RB / 16u));
#pragma unroll
        for (uint32_t j = 0; j < J; ++j) {
            int q = 0;
            if (s > 0.f) q = (int)fminf(fmaxf(rintf(x[j] / s), -127.f), 127.f);
            const double dd = (double)x[j] - (double)s * q;
            err2 += dd * dd;
            qq += q * q;
            if (lane + 64u * j < (uint32_t)D) ob[lane + 64u * j] = (int8_t)q;
        }
        for (int o = 32; o >= 1; o >>= 1) {
            err2 += __shfl_xor(err2, o, 64);
            qq += __shfl_xor(qq, o, 64);
        }
        if (lane < (RB - D) / 4u) {
            float v = 0.f;
            if (lane == 0) v = s;
            if (lane == 1) v = (float)(sqrt(err2) * (1.0 + 1e-6));
            if (lane == 2) v = (float)((double)s * (double)s * (double)qq);
            reinterpret_cast<float *>(ob + D)[lane] = v;
        }
    }
}

__global__ void k_to_bytes(const float *__restrict__ src, size_t n, uint8_t *__restrict__ dst) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = (uint8_t)src[i];
}

GramBound gram_bound(const Dataset &ds) {
    GramBound g{};
    g.exact = ds.gram_fmt == FK_GRAM_EXACT;
    g.inv_s2 = 1.f / (ds.gram_scale * ds.gram_scale);
    g.ymax = ds.gram_ymax;
    if (g.exact) return g;
    const double dp = ds.dpad;
    const double u = ds.gram_fmt == FK_GRAM_F16 ? ldexp(1.0, -11) : ldexp(1.0, -8);
    g.u = (float)(u * 1.001);
    // FP16: values below the smallest normal are rounded with absolute
    // error <= 2^-25 (scaled units) per element
    g.e_abs = ds.gram_fmt == FK_GRAM_F16 ? (float)(2.0 * sqrt(dp) * ldexp(1.0, -25) / ds.gram_scale) : 0.f;
    // norms: dpad products and sums; Gram: dpad + 16 accumulations; two
    // final operations
    g.f_rel = (float)((2.0 * dp + 16.0) * ldexp(1.0, -24) * 1.01 + ldexp(1.0, -21));
    // canonical distance: subtraction, square and fma per element, dpad / 8
    // terms per lane and 7 lane sums
    g.gam = (float)((dp + 16.0) * ldexp(1.0, -23));
    return g;
}

void gram_prepare(hipStream_t s, Dataset &ds, uint32_t flags) {
    const uint32_t want = flags & (FK_FLAG_GRAM_BF16 | FK_FLAG_GRAM_F16);
    if (ds.gram_fmt != FK_GRAM_NONE && ds.gram_req == want) return;
    DevBuf<uint32_t> st(8);
    const uint32_t init[5] = {1u, 0u, 0u, 0u, 0x7F800000u};
    FK_HIP(hipMemcpyAsync(st.p, init, sizeof(init), hipMemcpyHostToDevice, s));
    const unsigned grid = std::min<unsigned>(2048, (ds.n_total + 255) / 256);
    hipLaunchKernelGGL(k_gram_stats, dim3(grid), dim3(256), 0, s, ds.rows.p, ds.stride, ds.d, ds.n_total, st.p);
    FK_LAUNCH_CHECK();
    uint32_t h[5];
    FK_HIP(hipMemcpyAsync(h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
    FK_HIP(hipStreamSynchronize(s));
    float mabs, nmax, nmin;
    memcpy(&mabs, &h[1], 4);
    memcpy(&nmax, &h[3], 4);
    memcpy(&nmin, &h[4], 4);
    int fmt;
    if (want & FK_FLAG_GRAM_F16) fmt = FK_GRAM_F16;
    else if (want & FK_FLAG_GRAM_BF16) fmt = FK_GRAM_BF16;
    else if (h[0] && mabs <= 256.f && ((!h[2] && nmax < 8388608.f) || nmax < 4194304.f)) fmt = FK_GRAM_EXACT;
    else if (nmin >= 0.998f && nmax <= 1.002f) fmt = FK_GRAM_F16;  // unit-normalised
    else fmt = FK_GRAM_BF16;
    float scale = 1.f;
    if (fmt == FK_GRAM_F16) {
        FK_CHECK_ARG(mabs > 0.f && std::isfinite(mabs), "FP16 Gram form needs finite, non-zero rows");
        // largest power of two keeping every scaled value <= 2^14
        scale = ldexpf(1.f, 14 - ilogbf(mabs) - 1);
    }
    ds.gram_nrm.alloc(ds.n_total);
    const unsigned g2 = (ds.n_total + 255) / 256;
    const float inv_s2 = 1.f / (scale * scale);
    if (fmt == FK_GRAM_F16)
        hipLaunchKernelGGL(k_gram_norms<FK_GRAM_F16>, dim3(g2), dim3(256), 0, s, ds.rows.p, ds.stride, ds.dpad,
                           ds.n_total, scale, inv_s2, ds.gram_nrm.p);
    else
        hipLaunchKernelGGL(k_gram_norms<FK_GRAM_BF16>, dim3(g2), dim3(256), 0, s, ds.rows.p, ds.stride, ds.dpad,
                           ds.n_total, scale, inv_s2, ds.gram_nrm.p);
    FK_LAUNCH_CHECK();
    // the 16-bit operand copy (the Gram-form prune's tiles; exact forms as
    // BF16, which holds their integers exactly)
    ds.rows16.alloc((size_t)ds.n_total * ds.dpad);
    const unsigned g16 = std::min<unsigned>(65536, (ds.n_total + 3) / 4);
    if (fmt == FK_GRAM_F16)
        hipLaunchKernelGGL(k_rows16<FK_GRAM_F16>, dim3(g16), dim3(256), 0, s, ds.rows.p, ds.stride, ds.dpad,
                           ds.n_total, scale, ds.rows16.p);
    else
        hipLaunchKernelGGL(k_rows16<FK_GRAM_BF16>, dim3(g16), dim3(256), 0, s, ds.rows.p, ds.stride, ds.dpad,
                           ds.n_total, scale, ds.rows16.p);
    FK_LAUNCH_CHECK();
    // widths 96 and 256 in the FP16 form: the 8-bit copy of BridgeView's
    // bound pass
    ds.rows8q.release();
    if (fmt == FK_GRAM_F16 && ds.dpad == 96) {
        ds.rows8q.alloc((size_t)ds.n_total * 8u);
        hipLaunchKernelGGL(k_rows8q<96>, dim3(g16), dim3(256), 0, s, ds.rows.p, ds.stride, ds.n_total, ds.rows8q.p);
        FK_LAUNCH_CHECK();
    } else if (fmt == FK_GRAM_F16 && ds.dpad == 256) {
        ds.rows8q.alloc((size_t)ds.n_total * 24u);
        hipLaunchKernelGGL(k_rows8q<256>, dim3(g16), dim3(256), 0, s, ds.rows.p, ds.stride, ds.n_total,
                           ds.rows8q.p);
        FK_LAUNCH_CHECK();
    }
    // exact form with values in [0, 255]: a chain-major byte copy for the
    // octet kernels (dot products in 32-bit integers, exact)
    ds.chains8.release();
    ds.norms8.release();
    if (fmt == FK_GRAM_EXACT && !h[2] && mabs <= 255.f) {
        ds.chains8.alloc((size_t)ds.n_total * ds.dpad);
        const size_t total = (size_t)ds.n_total * ds.dpad;
        hipLaunchKernelGGL(k_to_bytes, dim3((unsigned)std::min<size_t>((total + 255) / 256, 65536)), dim3(256), 0, s,
                           ds.chains.p, total, ds.chains8.p);
        FK_LAUNCH_CHECK();
        ds.norms8.alloc(ds.n_total);
        hipLaunchKernelGGL(k_norms8, dim3(g2), dim3(256), 0, s, ds.chains8.p, ds.n_total, ds.dpad, ds.norms8.p);
        FK_LAUNCH_CHECK();
    }
    FK_HIP(hipStreamSynchronize(s));
    ds.gram_fmt = fmt;
    ds.gram_scale = scale;
    // nmax bounds every squared norm from above (a double sum rounded up)
    ds.gram_ymax = (float)(std::sqrt((double)nmax) * (1.0 + 1e-6));
    ds.gram_req = want;
}

// ---------------------------------------------------------------------------
// Pass 1 of the local join in Gram form: same outputs as k_join_count (pair
// distances for pass 2, per-slot candidate counts, bucket reservations).
// One workgroup (4 waves) per node; the sampled rows are staged in LDS in
// FP32 as in the plain join (tiles of x rows resident, y rows streamed when
// the list is longer than the staging area); each wave takes 16 x 16 pair
// tiles round robin.
// ---------------------------------------------------------------------------
template <int FMT>
__global__ __launch_bounds__(256) void k_join_gram(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                   const float2 *__restrict__ nrm, float scale, GramBound gb,
                                                   const uint32_t *__restrict__ lists,
                                                   const uint64_t *__restrict__ list_off,
                                                   const uint32_t *__restrict__ n_new,
                                                   const uint64_t *__restrict__ pair_off,
                                                   const float *__restrict__ thr, float *__restrict__ pdist,
                                                   uint32_t *__restrict__ resv, uint32_t *__restrict__ bcnt,
                                                   uint32_t N, const uint32_t *__restrict__ order, uint32_t tcap,
                                                   unsigned long long *__restrict__ n_recheck, uint32_t rows_bytes,
                                                   uint32_t t_lo, uint32_t t_hi) {
    extern __shared__ __align__(16) float s_rows[];
    float2 *s_n = reinterpret_cast<float2 *>(s_rows + rows_bytes / 4);
    float *s_thr = reinterpret_cast<float *>(s_n + tcap);
    uint32_t *s_c = reinterpret_cast<uint32_t *>(s_thr + tcap);
    uint32_t *s_id = s_c + tcap;
    const uint32_t node = node_of_block(order, N);
    if (node >= N) return;
    const uint64_t lo = list_off[node];
    const uint32_t T = (uint32_t)(list_off[node + 1] - lo);
    if (T <= t_lo || T > t_hi) return;  // another launch's lists
    const uint32_t nn = n_new[node];
    const uint32_t *lst = lists + lo;
    float *o = pdist + pair_off[node];
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const uint32_t id = lst[t];
        s_id[t] = id;
        s_thr[t] = thr[id];
        s_n[t] = nrm[id];
        s_c[t] = 0;
    }
    __syncthreads();
    const uint32_t ls = dpad + 4;
    const uint32_t cap = rows_bytes / (ls * 4);
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    uint32_t rechecks = 0;
    // x chunk [x0, x0 + nx) at s_rows, y chunk [y0, y0 + ny) at s_y
    auto tiles = [&](uint32_t x0, uint32_t nx, uint32_t y0, uint32_t ny, const float *s_y) {
        const uint32_t ntx = (nx + 15) >> 4, nty = (ny + 15) >> 4;
        for (uint32_t t = wave; t < ntx * nty; t += 4) {
            const uint32_t tx = t / nty, ty = t - tx * nty;
            // no pair of the tile has y > x
            if (y0 + ty * 16 + std::min(16u, ny - ty * 16) - 1 <= x0 + tx * 16) continue;
            const uint32_t ra = std::min(tx * 16 + (lane & 15), nx - 1);
            const uint32_t rb = std::min(ty * 16 + (lane & 15), ny - 1);
            const fk_f32x4 g = gram_tile<FMT>(s_rows + ra * ls, s_y + rb * ls, dpad, scale);
            // lane (G, j) = (lane >> 4, lane & 15) holds the pairs (x, y) with
            // x = G * 4 + e (e < 4) and y = j of the tile.  Counts: the hits
            // of x by a ballot over its 16 lanes, those of y summed over e
            // and the four groups, then two atomics per tile on distinct slots
            const uint32_t yi = ty * 16 + (lane & 15);
            const uint32_t y = y0 + yi;
            const bool yok = yi < ny;
            const uint32_t yc = yok ? y : y0;
            const float2 ny2 = s_n[yc];
            const float thy = s_thr[yc];
            const uint32_t idy = s_id[yc];
            uint32_t cy = 0, cx = 0;
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                const uint32_t xi = tx * 16 + (lane >> 4) * 4 + e;
                const uint32_t x = x0 + xi;
                bool hx = false;
                if (xi < nx && yok && y > x) {
                    const float2 nx2 = s_n[x];
                    const float thx = s_thr[x];
                    const float dt = (nx2.x + ny2.x) - 2.f * (g[e] * gb.inv_s2);
                    float dd = dt;
                    if (!gb.exact && gram_lower(dt, nx2, ny2, gb) <= fmaxf(thx, thy)) {
                        dd = l2_canon(s_rows + xi * ls, s_y + yi * ls, dpad);
                        ++rechecks;
                    }
                    o[join_pair_index(x, y, T)] = dd;
                    if (s_id[x] != idy) {
                        hx = dd <= thx;
                        cy += dd <= thy ? 1u : 0u;
                    }
                }
                const uint64_t m = wave_ballot(hx);
                if ((lane & 15) == e) cx = (uint32_t)__builtin_popcountll((m >> (lane & 48)) & 0xFFFFull);
            }
            cy += (uint32_t)__shfl_xor((int)cy, 16, 64);
            cy += (uint32_t)__shfl_xor((int)cy, 32, 64);
            if ((lane & 15) < 4 && cx) atomicAdd(&s_c[x0 + tx * 16 + (lane >> 4) * 4 + (lane & 15)], cx);
            if (lane < 16 && cy) atomicAdd(&s_c[y], cy);
        }
    };
    if (T <= cap) {
        join_stage_rows(rows, stride, dpad, s_id, 0, T, s_rows);
        __syncthreads();
        tiles(0, nn, 0, T, s_rows);
    } else {
        const uint32_t bx = nn < cap / 2 ? nn : cap / 2, by = cap - bx;
        float *s_y = s_rows + bx * ls;
        for (uint32_t x0 = 0; x0 < nn; x0 += bx) {
            const uint32_t nx = (nn - x0) < bx ? (nn - x0) : bx;
            __syncthreads();  // the previous tiles are done with the x rows
            join_stage_rows(rows, stride, dpad, s_id, x0, nx, s_rows);
            for (uint32_t y0 = x0 + 1; y0 < T; y0 += by) {
                const uint32_t ny = (T - y0) < by ? (T - y0) : by;
                if (y0 != x0 + 1) __syncthreads();  // the previous tile is done with the y rows
                join_stage_rows(rows, stride, dpad, s_id, y0, ny, s_y);
                __syncthreads();
                tiles(x0, nx, y0, ny, s_y);
            }
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const uint32_t c = s_c[t];
        resv[lo + t] = c ? atomicAdd(&bcnt[s_id[t]], c) : 0u;
    }
    if (!gb.exact) {
        const unsigned long long r = block_sum((unsigned long long)rechecks);
        if (threadIdx.x == 0 && r) tally_add(n_recheck, r);
    }
}

size_t join_gram_lds(uint32_t tcap, uint32_t rows_bytes) { return rows_bytes + (size_t)tcap * 20; }

// ---------------------------------------------------------------------------
// The same pass from byte rows (exact form, values in [0, 255]): the staged
// rows are the chain-major byte copy (dpad bytes, a fixed permutation of the
// dimensions, which a dot product does not see), shifted to signed bytes
// b - 128 (the bit pattern b ^ 0x80) for v_mfma_i32_16x16x32_i8.  With
// n'(x) = sum (x_k - 128)^2 over the dpad bytes,
//     |x - y|^2 = n'(x) + n'(y) - 2 (x - 128) . (y - 128)
// exactly in 32-bit integers, and the value (an integer below 2^24 in the
// exact form) converts to the canonical FP32 distance without rounding.
// A quarter of the staging bytes of the FP32 rows, half the MFMA steps.
// ---------------------------------------------------------------------------
typedef int fk_i32x4 __attribute__((ext_vector_type(4)));

// Row stride of staged byte rows: 16 rows of a tile read 8 bytes each at one
// offset, and dpad + 8 bytes between rows puts them in distinct banks.
__host__ __device__ inline uint32_t join8_row_bytes(uint32_t dpad) { return dpad + 8u; }

__device__ __forceinline__ void join_stage_bytes(const uint8_t *__restrict__ c8, uint32_t dpad, const uint32_t *s_id,
                                                 uint32_t first, uint32_t count, uint8_t *s_dst) {
    const uint32_t ls = join8_row_bytes(dpad), v8 = dpad >> 3;
    for (uint32_t i = threadIdx.x; i < count * v8; i += blockDim.x) {
        const uint32_t r = i / v8, c = i - r * v8;
        *reinterpret_cast<uint2 *>(s_dst + r * ls + c * 8) =
            *reinterpret_cast<const uint2 *>(c8 + (size_t)s_id[first + r] * dpad + c * 8);
    }
}

// 16 x 16 tile of shifted dot products (the gfx950 8-bit MFMA shape,
// 16 x 16 x 64).  Lane l supplies row (l & 15) of the A block (arow) and of
// the B block (brow), 16 bytes at k = 64 s + 16 (l >> 4) for every step s,
// zero past dpad (a multiple of 8, so a lane's bytes are all in, all out or
// the first 8 in); A and B take the same bytes of their rows, so every
// product pairs equal dimensions.  Rows are 8-byte aligned (read as two
// 8-byte halves).  On return lane l holds G[4 (l >> 4) + e][l & 15],
// e = 0..3.  Every lane of the wave must call it.
__device__ __forceinline__ fk_i32x4 gram_tile8(const uint8_t *__restrict__ arow, const uint8_t *__restrict__ brow,
                                               uint32_t dpad) {
    fk_i32x4 acc = {0, 0, 0, 0};
    const uint32_t kq = (lane_id() >> 4) * 16u;
    constexpr uint64_t flip = 0x8080808080808080ull;
    for (uint32_t k0 = 0; k0 < dpad; k0 += 64u) {
        const uint32_t k = k0 + kq;
        uint64_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
        if (k < dpad) {
            a0 = *reinterpret_cast<const uint64_t *>(arow + k) ^ flip;
            b0 = *reinterpret_cast<const uint64_t *>(brow + k) ^ flip;
            if (k + 8u < dpad) {
                a1 = *reinterpret_cast<const uint64_t *>(arow + k + 8) ^ flip;
                b1 = *reinterpret_cast<const uint64_t *>(brow + k + 8) ^ flip;
            }
        }
        const fk_i32x4 a = {(int)(uint32_t)a0, (int)(uint32_t)(a0 >> 32), (int)(uint32_t)a1, (int)(uint32_t)(a1 >> 32)};
        const fk_i32x4 b = {(int)(uint32_t)b0, (int)(uint32_t)(b0 >> 32), (int)(uint32_t)b1, (int)(uint32_t)(b1 >> 32)};
        acc = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, acc, 0, 0, 0);
    }
    return acc;
}

__global__ __launch_bounds__(256) void k_join_gram8(const uint8_t *__restrict__ c8, uint32_t dpad,
                                                    const int32_t *__restrict__ norms8,
                                                    const uint32_t *__restrict__ lists,
                                                    const uint64_t *__restrict__ list_off,
                                                    const uint32_t *__restrict__ n_new,
                                                    const uint64_t *__restrict__ pair_off,
                                                    const float *__restrict__ thr, float *__restrict__ pdist,
                                                    uint32_t *__restrict__ resv, uint32_t *__restrict__ bcnt,
                                                    uint32_t N, const uint32_t *__restrict__ order, uint32_t tcap,
                                                    uint32_t rows_bytes, uint32_t t_lo, uint32_t t_hi) {
    extern __shared__ __align__(16) uint8_t s_rows8[];
    int32_t *s_n = reinterpret_cast<int32_t *>(s_rows8 + rows_bytes);
    float *s_thr = reinterpret_cast<float *>(s_n + tcap);
    uint32_t *s_c = reinterpret_cast<uint32_t *>(s_thr + tcap);
    uint32_t *s_id = s_c + tcap;
    const uint32_t node = node_of_block(order, N);
    if (node >= N) return;
    const uint64_t lo = list_off[node];
    const uint32_t T = (uint32_t)(list_off[node + 1] - lo);
    if (T <= t_lo || T > t_hi) return;  // another launch's lists
    const uint32_t nn = n_new[node];
    const uint32_t *lst = lists + lo;
    float *o = pdist + pair_off[node];
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const uint32_t id = lst[t];
        s_id[t] = id;
        s_thr[t] = thr[id];
        s_n[t] = norms8[id];
        s_c[t] = 0;
    }
    __syncthreads();
    const uint32_t ls = join8_row_bytes(dpad);
    const uint32_t cap = rows_bytes / ls;
    const uint32_t wave = threadIdx.x >> 6, lane = lane_id();
    // x chunk [x0, x0 + nx) at s_rows8, y chunk [y0, y0 + ny) at s_y; the
    // epilogue is k_join_gram's (slot counts by ballot and register sums)
    auto tiles = [&](uint32_t x0, uint32_t nx, uint32_t y0, uint32_t ny, const uint8_t *s_y) {
        const uint32_t ntx = (nx + 15) >> 4, nty = (ny + 15) >> 4;
        for (uint32_t t = wave; t < ntx * nty; t += 4) {
            const uint32_t tx = t / nty, ty = t - tx * nty;
            if (y0 + ty * 16 + std::min(16u, ny - ty * 16) - 1 <= x0 + tx * 16) continue;
            const uint32_t ra = std::min(tx * 16 + (lane & 15), nx - 1);
            const uint32_t rb = std::min(ty * 16 + (lane & 15), ny - 1);
            const fk_i32x4 g = gram_tile8(s_rows8 + ra * ls, s_y + rb * ls, dpad);
            const uint32_t yi = ty * 16 + (lane & 15);
            const uint32_t y = y0 + yi;
            const bool yok = yi < ny;
            const uint32_t yc = yok ? y : y0;
            const int32_t nyv = s_n[yc];
            const float thy = s_thr[yc];
            const uint32_t idy = s_id[yc];
            uint32_t cy = 0, cx = 0;
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                const uint32_t xi = tx * 16 + (lane >> 4) * 4 + e;
                const uint32_t x = x0 + xi;
                bool hx = false;
                if (xi < nx && yok && y > x) {
                    const float thx = s_thr[x];
                    const float dd = (float)(s_n[x] + nyv - 2 * g[e]);
                    o[join_pair_index(x, y, T)] = dd;
                    if (s_id[x] != idy) {
                        hx = dd <= thx;
                        cy += dd <= thy ? 1u : 0u;
                    }
                }
                const uint64_t m = wave_ballot(hx);
                if ((lane & 15) == e) cx = (uint32_t)__builtin_popcountll((m >> (lane & 48)) & 0xFFFFull);
            }
            cy += (uint32_t)__shfl_xor((int)cy, 16, 64);
            cy += (uint32_t)__shfl_xor((int)cy, 32, 64);
            if ((lane & 15) < 4 && cx) atomicAdd(&s_c[x0 + tx * 16 + (lane >> 4) * 4 + (lane & 15)], cx);
            if (lane < 16 && cy) atomicAdd(&s_c[y], cy);
        }
    };
    if (T <= cap) {
        join_stage_bytes(c8, dpad, s_id, 0, T, s_rows8);
        __syncthreads();
        tiles(0, nn, 0, T, s_rows8);
    } else {
        const uint32_t bx = nn < cap / 2 ? nn : cap / 2, by = cap - bx;
        uint8_t *s_y = s_rows8 + bx * ls;
        for (uint32_t x0 = 0; x0 < nn; x0 += bx) {
            const uint32_t nx = (nn - x0) < bx ? (nn - x0) : bx;
            __syncthreads();  // the previous tiles are done with the x rows
            join_stage_bytes(c8, dpad, s_id, x0, nx, s_rows8);
            for (uint32_t y0 = x0 + 1; y0 < T; y0 += by) {
                const uint32_t ny = (T - y0) < by ? (T - y0) : by;
                if (y0 != x0 + 1) __syncthreads();  // the previous tile is done with the y rows
                join_stage_bytes(c8, dpad, s_id, y0, ny, s_y);
                __syncthreads();
                tiles(x0, nx, y0, ny, s_y);
            }
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const uint32_t c = s_c[t];
        resv[lo + t] = c ? atomicAdd(&bcnt[s_id[t]], c) : 0u;
    }
}

size_t join_gram8_lds(uint32_t tcap, uint32_t rows_bytes) { return rows_bytes + (size_t)tcap * 16; }

// Lists of up to FK_JOIN_SMALL_ROWS rows first with a small staging area
// (more workgroups per CU), then the longer ones with the full area.
void launch_join_gram(BuildState &b, uint32_t tcap, unsigned long long *n_recheck) {
    const Dataset &ds = *b.ds;
    const uint32_t *ord = b.has_order ? b.order.p : nullptr;
    if (ds.gram_fmt == FK_GRAM_EXACT && ds.chains8.p && ds.norms8.p) {
        const uint32_t rb8 = join8_row_bytes(ds.dpad);
        const uint32_t small8 = std::min<uint32_t>(FK_JOIN_SMALL_ROWS * rb8, FK_JOIN_LDS_ROWS_BYTES) & ~15u;
        const uint32_t big8 = FK_JOIN_LDS_ROWS_BYTES & ~15u;
        const uint32_t small_t8 = small8 / rb8;
        FK_CHECK_ARG(small_t8 >= 2, "rows too long for the byte-row join");
        auto launch8 = [&](uint32_t tc, uint32_t rb, uint32_t t_lo, uint32_t t_hi) {
            hipLaunchKernelGGL(k_join_gram8, dim3(node_grid(b)), dim3(256), join_gram8_lds(tc, rb), b.s,
                               ds.chains8.p, ds.dpad, ds.norms8.p, b.lists.p, b.list_off.p, b.n_new.p,
                               b.pair_off.p, b.thr.p, b.pdist.p, b.resv.p, b.bcnt.p, b.N, ord, tc, rb, t_lo, t_hi);
            FK_LAUNCH_CHECK();
        };
        launch8(std::min(tcap, small_t8), small8, 0u, small_t8);
        if (tcap > small_t8) launch8(tcap, big8, small_t8, 0xFFFFFFFFu);
        return;
    }
    const GramBound gb = gram_bound(ds);
    const uint32_t row_bytes = (ds.dpad + 4) * 4;
    const uint32_t small_bytes = std::min<uint32_t>(FK_JOIN_SMALL_ROWS * row_bytes, FK_JOIN_LDS_ROWS_BYTES);
    const uint32_t small_t = small_bytes / row_bytes;
    auto launch = [&](uint32_t tc, uint32_t rb, uint32_t t_lo, uint32_t t_hi) {
        if (ds.gram_fmt == FK_GRAM_F16)
            hipLaunchKernelGGL(k_join_gram<FK_GRAM_F16>, dim3(node_grid(b)), dim3(256), join_gram_lds(tc, rb), b.s,
                               ds.rows.p, ds.stride, ds.dpad, ds.gram_nrm.p, ds.gram_scale, gb, b.lists.p,
                               b.list_off.p, b.n_new.p, b.pair_off.p, b.thr.p, b.pdist.p, b.resv.p, b.bcnt.p, b.N,
                               ord, tc, n_recheck, rb, t_lo, t_hi);
        else
            hipLaunchKernelGGL(k_join_gram<FK_GRAM_BF16>, dim3(node_grid(b)), dim3(256), join_gram_lds(tc, rb), b.s,
                               ds.rows.p, ds.stride, ds.dpad, ds.gram_nrm.p, ds.gram_scale, gb, b.lists.p,
                               b.list_off.p, b.n_new.p, b.pair_off.p, b.thr.p, b.pdist.p, b.resv.p, b.bcnt.p, b.N,
                               ord, tc, n_recheck, rb, t_lo, t_hi);
        FK_LAUNCH_CHECK();
    };
    launch(std::min(tcap, small_t), small_bytes, 0u, small_t);
    if (tcap > small_t) launch(tcap, FK_JOIN_LDS_ROWS_BYTES, small_t, 0xFFFFFFFFu);
}

// ---------------------------------------------------------------------------
// Gram-form distances of arbitrary pairs with their bounds (the unit entry
// point): one wavefront per 16 pairs, whose rows form the A and B blocks;
// the pair distances are the tile's diagonal.
// ---------------------------------------------------------------------------
template <int FMT>
__global__ __launch_bounds__(256) void k_gram_pairs(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                    const float2 *__restrict__ nrm, float scale, GramBound gb,
                                                    const uint32_t *__restrict__ a, const uint32_t *__restrict__ b,
                                                    size_t n, float *__restrict__ out, float *__restrict__ lo,
                                                    float *__restrict__ hi) {
    const size_t base = ((size_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16;
    if (base >= n) return;  // whole wave
    const uint32_t lane = lane_id();
    const size_t pr = std::min<size_t>(base + (lane & 15), n - 1);
    const uint32_t ia = a[pr], ib = b[pr];
    const fk_f32x4 g = gram_tile<FMT>(rows + (size_t)ia * stride, rows + (size_t)ib * stride, dpad, scale);
    // lane l holds rows 4 (l >> 4) .. +3 of column l & 15: the diagonal
    // element of column j sits in the lane group j >> 2, entry j & 3
    const uint32_t j = lane & 15;
    if ((j >> 2) != (lane >> 4)) return;
    const size_t p = base + j;
    if (p >= n) return;
    const float2 nx2 = nrm[ia], ny2 = nrm[ib];
    const float dt = (nx2.x + ny2.x) - 2.f * (g[j & 3] * gb.inv_s2);
    out[p] = dt;
    lo[p] = gb.exact ? dt : gram_lower(dt, nx2, ny2, gb);
    hi[p] = gb.exact ? dt : gram_upper(dt, nx2, ny2, gb);
}

void launch_gram_pairs(hipStream_t s, const Dataset &ds, const uint32_t *a, const uint32_t *b, size_t n,
                       float *out, float *lo, float *hi) {
    if (n == 0) return;
    const GramBound gb = gram_bound(ds);
    const unsigned grid = (unsigned)((n + 63) / 64);  // 4 waves x 16 pairs
    if (ds.gram_fmt == FK_GRAM_F16)
        hipLaunchKernelGGL(k_gram_pairs<FK_GRAM_F16>, dim3(grid), dim3(256), 0, s, ds.rows.p, ds.stride, ds.dpad,
                           ds.gram_nrm.p, ds.gram_scale, gb, a, b, n, out, lo, hi);
    else
        hipLaunchKernelGGL(k_gram_pairs<FK_GRAM_BF16>, dim3(grid), dim3(256), 0, s, ds.rows.p, ds.stride, ds.dpad,
                           ds.gram_nrm.p, ds.gram_scale, gb, a, b, n, out, lo, hi);
    FK_LAUNCH_CHECK();
}

}  // namespace fk
