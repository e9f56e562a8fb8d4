// Learned distance screen for the NN-descent local join.
//
// Model, trained during each build on the exact pair distances of the first
// NN-descent iteration:
//     z = P x                          P: 32 x d, a linear projection
//     u = (z_a - z_b)^2 / mu           elementwise; mu = mean training distance
//     h = relu(W1 u + b1)              W1: 64 x 32
//     y = w2 . h + b2                  the predicted distance over mu
// loss = mean over the batch of ((y - t) / t)^2, t = max(dist / mu, 0.01)
// (relative error: the join needs the screen accurate on near pairs, which
// are rare among the first join's random pairs), minimised with Adam
// (Kingma and Ba, 2015) by the kernels below: a forward and backward pass per
// pair, then the weight gradients as reductions over the batch.  The initial
// weights make the network the plain random-projection estimate
// (Var P = 1 / 32, W1 >= 0 with row sums about 1, w2 = 1/2), which is what
// FK_FLAG_SCREEN_NOTRAIN keeps: the method without the trained network.
//
// In the join (iterations 2+), a pair whose predicted distance times
// (1 - margin) exceeds both endpoints' pool thresholds is skipped: no exact
// distance is computed and the pair can enter neither pool.  The margin is
// recalibrated before every screened join on that join's own pairs: for a
// seeded sample of about one node in 64, every pair's exact distance is
// computed, and the margin is the (1 - 1e-3) quantile of (pred - dist) / pred
// over the sampled pairs that a threshold admits, so about one such pair in a
// thousand is expected to be lost.  (Calibrating once on the training pairs
// does not transfer: the first join's pairs are random, later joins' pairs
// are neighbours.)  One skipped pair in 64 is still computed exactly and the
// candidates it would have given are counted (fk_stats.screen_missed).
// The projections are stored in BF16 and the inference weights rounded to
// BF16; the arithmetic is FP32.
#include <math.h>

#include <algorithm>
#include <random>
#include <vector>

#include "fk_api_impl.h"
#include "fk_build.h"
#include "fk_device.h"
#include "fk_internal.h"
#include "fk_join.h"

namespace fk {

#define FK_SCR_Z 32
#define FK_SCR_H 64
// per-pair training record: u[32] | h[64] | dh[64] | dz[32] | dy | pair
#define FK_SCR_ACT 196
#define FK_SCR_BATCH 32768
#define FK_SCR_AUDIT 64

struct ScreenLayout {
    uint32_t pt, w1, b1, w2, b2, n;
};
static ScreenLayout screen_layout(uint32_t dpad) {
    ScreenLayout l;
    l.pt = 0;
    l.w1 = dpad * FK_SCR_Z;
    l.b1 = l.w1 + FK_SCR_H * FK_SCR_Z;
    l.w2 = l.b1 + FK_SCR_H;
    l.b2 = l.w2 + FK_SCR_H;
    l.n = l.b2 + 1;
    return l;
}
// inference weights: W1 | b1 | w2 | b2
#define FK_SCR_WI_B1 (FK_SCR_H * FK_SCR_Z)
#define FK_SCR_WI_W2 (FK_SCR_WI_B1 + FK_SCR_H)
#define FK_SCR_WI_B2 (FK_SCR_WI_W2 + FK_SCR_H)
#define FK_SCR_WI_N (FK_SCR_WI_B2 + 1)

__device__ __forceinline__ uint16_t f2bf(float x) {
    const uint32_t b = __float_as_uint(x);
    return (uint16_t)((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
static float host_bf16_round(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    b = ((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16) << 16;
    float r;
    memcpy(&r, &b, 4);
    return r;
}

// Screen output (distance / mu) for two BF16 projections.
__device__ __forceinline__ float screen_eval(const uint16_t *__restrict__ za, const uint16_t *__restrict__ zb,
                                             const float *__restrict__ wi) {
    float u[FK_SCR_Z];
#pragma unroll
    for (int q = 0; q < FK_SCR_Z / 8; ++q) {
        const uint4 a = reinterpret_cast<const uint4 *>(za)[q];
        const uint4 b = reinterpret_cast<const uint4 *>(zb)[q];
        const uint32_t av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float d0 = __uint_as_float(av[r] << 16) - __uint_as_float(bv[r] << 16);
            const float d1 = __uint_as_float(av[r] & 0xFFFF0000u) - __uint_as_float(bv[r] & 0xFFFF0000u);
            u[q * 8 + r * 2] = d0 * d0;
            u[q * 8 + r * 2 + 1] = d1 * d1;
        }
    }
    float y = wi[FK_SCR_WI_B2];
    for (int i = 0; i < FK_SCR_H; ++i) {
        float hp = wi[FK_SCR_WI_B1 + i];
#pragma unroll
        for (int j = 0; j < FK_SCR_Z; ++j) hp = __builtin_fmaf(wi[i * FK_SCR_Z + j], u[j], hp);
        y = __builtin_fmaf(wi[FK_SCR_WI_W2 + i], fmaxf(hp, 0.f), y);
    }
    return y;
}

// ---------------------------------------------------------------------------
// Training.
// ---------------------------------------------------------------------------
// Pairs of the first iteration's join (distinct ids) of the nodes selected
// by a seeded hash (one in q), with their exact distances.
__global__ __launch_bounds__(256) void k_scr_sample(const uint32_t *__restrict__ lists,
                                                    const uint64_t *__restrict__ list_off,
                                                    const uint32_t *__restrict__ n_new,
                                                    const uint64_t *__restrict__ pair_off,
                                                    const float *__restrict__ pdist, uint32_t N, uint32_t q,
                                                    uint64_t seed, uint32_t cap, uint32_t *__restrict__ oa,
                                                    uint32_t *__restrict__ ob, float *__restrict__ ot,
                                                    uint32_t *__restrict__ counter) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N || fk_mix64(seed ^ ((uint64_t)n * 0xD6E8FEB86659FD93ull)) % q != 0) return;
    const uint64_t lo = list_off[n];
    const uint32_t T = (uint32_t)(list_off[n + 1] - lo), nn = n_new[n];
    const uint32_t *l = lists + lo;
    const float *o = pdist + pair_off[n];
    uint32_t c = 0;
    for (uint32_t x = 0; x < nn; ++x)
        for (uint32_t y = x + 1; y < T; ++y) c += l[x] != l[y];
    if (c == 0) return;
    const uint32_t base = atomicAdd(counter, c);
    uint32_t k = base;
    for (uint32_t x = 0; x < nn; ++x)
        for (uint32_t y = x + 1; y < T; ++y) {
            if (l[x] == l[y]) continue;
            if (k < cap) {
                oa[k] = l[x];
                ob[k] = l[y];
                ot[k] = o[join_pair_index(x, y, T)];
            }
            ++k;
        }
}

// Batch element e is training pair ((start + e) * mul) mod n_all (mul coprime
// to n_all: a permutation).  Forward and backward of one pair per thread;
// the record keeps what the weight gradients need.  loss += sum of squared
// residuals.
__global__ __launch_bounds__(256) void k_scr_fwdbwd(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                    const float *__restrict__ W, const uint32_t *__restrict__ ta,
                                                    const uint32_t *__restrict__ tb, const float *__restrict__ tt,
                                                    uint32_t n_all, uint64_t start, uint64_t mul, uint32_t B,
                                                    float inv_mu, float *__restrict__ act, double *__restrict__ loss) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    double sq = 0.0;
    if (e < B) {
        const uint32_t p = (uint32_t)(((start + e) * mul) % n_all);
        const float *xa = rows + (size_t)ta[p] * stride, *xb = rows + (size_t)tb[p] * stride;
        const float *PT = W, *W1 = W + (size_t)dpad * FK_SCR_Z, *b1 = W1 + FK_SCR_H * FK_SCR_Z, *w2 = b1 + FK_SCR_H,
                    *b2 = w2 + FK_SCR_H;
        float z[FK_SCR_Z];
#pragma unroll
        for (int j = 0; j < FK_SCR_Z; ++j) z[j] = 0.f;
        for (uint32_t k = 0; k < dpad; k += 4) {
            const float4 a = *reinterpret_cast<const float4 *>(xa + k);
            const float4 b = *reinterpret_cast<const float4 *>(xb + k);
            const float dv[4] = {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w};
#pragma unroll
            for (int r = 0; r < 4; ++r)
#pragma unroll
                for (int j = 0; j < FK_SCR_Z; ++j) z[j] = __builtin_fmaf(PT[(k + r) * FK_SCR_Z + j], dv[r], z[j]);
        }
        float *A = act + (size_t)e * FK_SCR_ACT;
        float u[FK_SCR_Z];
#pragma unroll
        for (int j = 0; j < FK_SCR_Z; ++j) {
            u[j] = z[j] * z[j] * inv_mu;
            A[j] = u[j];
        }
        float y = b2[0];
        for (int i = 0; i < FK_SCR_H; ++i) {
            float hp = b1[i];
#pragma unroll
            for (int j = 0; j < FK_SCR_Z; ++j) hp = __builtin_fmaf(W1[i * FK_SCR_Z + j], u[j], hp);
            const float h = fmaxf(hp, 0.f);
            A[32 + i] = h;
            y = __builtin_fmaf(w2[i], h, y);
        }
        const float t = fmaxf(tt[p] * inv_mu, 1e-2f);
        const float r = (y - t) / t;
        sq = (double)r * (double)r;
        const float dy = 2.f * r / (t * (float)B);
        float du[FK_SCR_Z];
#pragma unroll
        for (int j = 0; j < FK_SCR_Z; ++j) du[j] = 0.f;
        for (int i = 0; i < FK_SCR_H; ++i) {
            const float dh = A[32 + i] > 0.f ? dy * w2[i] : 0.f;
            A[96 + i] = dh;
#pragma unroll
            for (int j = 0; j < FK_SCR_Z; ++j) du[j] = __builtin_fmaf(W1[i * FK_SCR_Z + j], dh, du[j]);
        }
#pragma unroll
        for (int j = 0; j < FK_SCR_Z; ++j) A[160 + j] = du[j] * 2.f * z[j] * inv_mu;
        A[192] = dy;
        A[193] = __uint_as_float(p);
    }
    sq = block_sum(sq);
    if (threadIdx.x == 0 && sq != 0.0) atomicAdd(loss, sq);
}

// dP^T[k][j] = sum_e (x_a - x_b)[k] dz_e[j]; grid (dpad / 8, chunks),
// block = 8 dimensions x 32 outputs.
__global__ __launch_bounds__(256) void k_scr_gproj(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                   const uint32_t *__restrict__ ta, const uint32_t *__restrict__ tb,
                                                   const float *__restrict__ act, uint32_t B, uint32_t chunk,
                                                   float *__restrict__ G) {
    const uint32_t k = blockIdx.x * 8 + (threadIdx.x >> 5), j = threadIdx.x & 31u;
    const uint32_t e0 = blockIdx.y * chunk, e1 = std::min(e0 + chunk, B);
    if (k >= dpad) return;
    float acc = 0.f;
    for (uint32_t e = e0; e < e1; ++e) {
        const float *A = act + (size_t)e * FK_SCR_ACT;
        const uint32_t p = __float_as_uint(A[193]);
        const float dv = rows[(size_t)ta[p] * stride + k] - rows[(size_t)tb[p] * stride + k];
        acc = __builtin_fmaf(dv, A[160 + j], acc);
    }
    if (acc != 0.f) atomicAdd(&G[(size_t)k * FK_SCR_Z + j], acc);
}

// W1, b1, w2 and b2 gradients; grid = batch chunks, 256 threads: thread t
// owns W1 row t / 4, columns 8 (t % 4) .. + 7, and one of b1 (t < 64), w2
// (64 <= t < 128) or b2 (t = 128).
__global__ __launch_bounds__(256) void k_scr_gmlp(const float *__restrict__ act, uint32_t B, uint32_t chunk,
                                                  float *__restrict__ Gw1) {
    const uint32_t t = threadIdx.x, i = t >> 2, j0 = (t & 3u) * 8;
    const uint32_t e0 = blockIdx.x * chunk, e1 = std::min(e0 + chunk, B);
    float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float aux = 0.f;
    for (uint32_t e = e0; e < e1; ++e) {
        const float *A = act + (size_t)e * FK_SCR_ACT;
        const float dh = A[96 + i];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc[q] = __builtin_fmaf(dh, A[j0 + q], acc[q]);
        if (t < 64) aux += A[96 + t];
        else if (t < 128) aux = __builtin_fmaf(A[192], A[32 + t - 64], aux);
        else if (t == 128) aux += A[192];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q)
        if (acc[q] != 0.f) atomicAdd(&Gw1[i * FK_SCR_Z + j0 + q], acc[q]);
    if (t <= 128 && aux != 0.f) atomicAdd(&Gw1[FK_SCR_H * FK_SCR_Z + t], aux);  // b1 | w2 | b2 follow W1
}

__global__ void k_scr_adam(float *__restrict__ W, const float *__restrict__ G, float *__restrict__ M,
                           float *__restrict__ V, uint32_t n, float lr, float bc1, float bc2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = G[i];
    const float m = 0.9f * M[i] + 0.1f * g;
    const float v = 0.999f * V[i] + 0.001f * g * g;
    M[i] = m;
    V[i] = v;
    W[i] -= lr * (m / bc1) / (sqrtf(v / bc2) + 1e-8f);
}

// FP32 forward of arbitrary pairs (test entry point).
__global__ __launch_bounds__(256) void k_scr_predict(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                     const float *__restrict__ W, const uint32_t *__restrict__ a,
                                                     const uint32_t *__restrict__ b, uint32_t n, float inv_mu,
                                                     float *__restrict__ out) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float *xa = rows + (size_t)a[e] * stride, *xb = rows + (size_t)b[e] * stride;
    const float *PT = W, *W1 = W + (size_t)dpad * FK_SCR_Z, *b1 = W1 + FK_SCR_H * FK_SCR_Z, *w2 = b1 + FK_SCR_H,
                *b2 = w2 + FK_SCR_H;
    float z[FK_SCR_Z];
#pragma unroll
    for (int j = 0; j < FK_SCR_Z; ++j) z[j] = 0.f;
    for (uint32_t k = 0; k < dpad; ++k) {
        const float dv = xa[k] - xb[k];
#pragma unroll
        for (int j = 0; j < FK_SCR_Z; ++j) z[j] = __builtin_fmaf(PT[k * FK_SCR_Z + j], dv, z[j]);
    }
    float y = b2[0];
    for (int i = 0; i < FK_SCR_H; ++i) {
        float hp = b1[i];
#pragma unroll
        for (int j = 0; j < FK_SCR_Z; ++j) hp = __builtin_fmaf(W1[i * FK_SCR_Z + j], z[j] * z[j] * inv_mu, hp);
        y = __builtin_fmaf(w2[i], fmaxf(hp, 0.f), y);
    }
    out[e] = y;
}

// Z[n] = BF16(P (x_n - center) * sc), one row per thread.
__global__ __launch_bounds__(256) void k_scr_project(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                     uint32_t N, const float *__restrict__ W,
                                                     const float *__restrict__ center, float sc,
                                                     uint16_t *__restrict__ Z) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const float *x = rows + (size_t)n * stride;
    float z[FK_SCR_Z];
#pragma unroll
    for (int j = 0; j < FK_SCR_Z; ++j) z[j] = 0.f;
    for (uint32_t k = 0; k < dpad; k += 4) {
        const float4 v = *reinterpret_cast<const float4 *>(x + k);
        const float4 c = *reinterpret_cast<const float4 *>(center + k);
        const float dv[4] = {v.x - c.x, v.y - c.y, v.z - c.z, v.w - c.w};
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int j = 0; j < FK_SCR_Z; ++j) z[j] = __builtin_fmaf(W[(k + r) * FK_SCR_Z + j], dv[r], z[j]);
    }
    uint16_t *o = Z + (size_t)n * FK_SCR_Z;
#pragma unroll
    for (int j = 0; j < FK_SCR_Z; ++j) o[j] = f2bf(z[j] * sc);
}

// Recalibration sample: node k * q + off for k < nodes; one wavefront per
// node computes the exact distance of each of its join pairs (distinct ids)
// and, for the pairs a threshold admits, appends (pred - dist) / pred with
// the inference path (-1 where pred <= 0).
__global__ __launch_bounds__(64) void k_scr_recal(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                  const uint32_t *__restrict__ lists,
                                                  const uint64_t *__restrict__ list_off,
                                                  const uint32_t *__restrict__ n_new, const float *__restrict__ thr,
                                                  uint32_t N, uint32_t q, uint32_t off,
                                                  const uint16_t *__restrict__ Z, const float *__restrict__ wi,
                                                  float mu, uint32_t cap, float *__restrict__ rho,
                                                  uint32_t *__restrict__ counter) {
    const uint32_t n = blockIdx.x * q + off;
    if (n >= N) return;
    const uint64_t lo = list_off[n];
    const uint32_t T = (uint32_t)(list_off[n + 1] - lo), nn = n_new[n];
    const uint32_t *l = lists + lo;
    for (uint32_t c = threadIdx.x; c < nn * T; c += blockDim.x) {
        const uint32_t x = c / T, y = c - x * T;
        if (y <= x) continue;
        const uint32_t ix = l[x], iy = l[y];
        if (ix == iy) continue;
        const float tm = fmaxf(thr[ix], thr[iy]);
        if (!(tm < INFINITY)) continue;
        const float d = l2_canon(rows + (size_t)ix * stride, rows + (size_t)iy * stride, dpad);
        if (d > tm) continue;
        const float pred = screen_eval(Z + (size_t)ix * FK_SCR_Z, Z + (size_t)iy * FK_SCR_Z, wi) * mu;
        const uint32_t k = atomicAdd(counter, 1u);
        if (k < cap) rho[k] = pred > 0.f ? (pred - d) / pred : -1.f;
    }
}

// Shared training step (used by the build and by fk_screen_grad): forward,
// backward and weight gradients of one batch into G (zeroed here).
static void screen_grad_step(hipStream_t s, const Dataset &ds, const float *W, const uint32_t *ta,
                             const uint32_t *tb, const float *tt, uint32_t n_all, uint64_t start, uint64_t mul,
                             uint32_t B, float inv_mu, float *act, float *G, double *loss) {
    const ScreenLayout L = screen_layout(ds.dpad);
    FK_HIP(hipMemsetAsync(G, 0, (size_t)L.n * 4, s));
    hipLaunchKernelGGL(k_scr_fwdbwd, dim3((B + 255) / 256), dim3(256), 0, s, ds.rows.p, ds.stride, ds.dpad, W, ta,
                       tb, tt, n_all, start, mul, B, inv_mu, act, loss);
    FK_LAUNCH_CHECK();
    const uint32_t chunk = 256, nch = (B + chunk - 1) / chunk;
    hipLaunchKernelGGL(k_scr_gproj, dim3(ds.dpad / 8, nch), dim3(256), 0, s, ds.rows.p, ds.stride, ds.dpad, ta, tb,
                       act, B, chunk, G);
    FK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scr_gmlp, dim3(nch), dim3(256), 0, s, act, B, chunk, G + L.w1);
    FK_LAUNCH_CHECK();
}

static void screen_init_weights(uint32_t dpad, uint32_t seed, std::vector<float> &w) {
    const ScreenLayout L = screen_layout(dpad);
    w.assign(L.n, 0.f);
    std::mt19937_64 g(0xC0FFEEull ^ ((uint64_t)seed << 20));
    std::normal_distribution<float> nd(0.f, 1.f / sqrtf((float)FK_SCR_Z));
    std::uniform_real_distribution<float> ud(0.f, 2.f / FK_SCR_Z);
    for (uint32_t i = L.pt; i < L.w1; ++i) w[i] = nd(g);
    for (uint32_t i = L.w1; i < L.b1; ++i) w[i] = ud(g);
    for (uint32_t i = L.w2; i < L.b2; ++i) w[i] = 0.5f;
}

static uint64_t coprime_multiplier(uint64_t n, uint64_t seed) {
    uint64_t m = (fk_mix64(seed) % n) | 1u;
    auto gcd = [](uint64_t a, uint64_t b) { while (b) { const uint64_t t = a % b; a = b; b = t; } return a; };
    while (gcd(m, n) != 1) m = (m + 2) % n ? (m + 2) % n : 1;
    return m;
}

void screen_train(BuildState &b) {
    const double t0 = now_sec();
    ScreenState &S = b.scr;
    const Dataset &ds = *b.ds;
    const uint32_t N = b.N;
    const ScreenLayout L = screen_layout(ds.dpad);
    // training pairs: about FK screen_pairs of the first join's pairs
    uint64_t pairs = 0;
    FK_HIP(hipMemcpyAsync(&pairs, b.pair_off.p + N, 8, hipMemcpyDeviceToHost, b.s));
    b.sync();
    const uint32_t want = 5000000u;
    const uint32_t q = (uint32_t)std::max<uint64_t>(1, pairs / want);
    const uint32_t cap = (uint32_t)std::min<uint64_t>(pairs, (uint64_t)want * 2);
    DevBuf<uint32_t> ta(cap), tb(cap), cnt(1);
    DevBuf<float> tt(cap);
    FK_HIP(hipMemsetAsync(cnt.p, 0, 4, b.s));
    hipLaunchKernelGGL(k_scr_sample, dim3((N + 255) / 256), dim3(256), 0, b.s, b.lists.p, b.list_off.p, b.n_new.p,
                       b.pair_off.p, b.pdist.p, N, q, (uint64_t)b.P.seed, cap, ta.p, tb.p, tt.p, cnt.p);
    FK_LAUNCH_CHECK();
    uint32_t n_all = 0;
    FK_HIP(hipMemcpyAsync(&n_all, cnt.p, 4, hipMemcpyDeviceToHost, b.s));
    b.sync();
    n_all = std::min(n_all, cap);
    FK_CHECK_ARG(n_all >= 1024, "too few join pairs to train the distance screen");
    // mean training distance
    std::vector<float> host_t(n_all);
    FK_HIP(hipMemcpyAsync(host_t.data(), tt.p, (size_t)n_all * 4, hipMemcpyDeviceToHost, b.s));
    b.sync();
    double sum = 0;
    for (float v : host_t) sum += v;
    S.mu = (float)std::max(sum / n_all, 1e-30);
    const float inv_mu = 1.f / S.mu;
    const uint32_t n_train = n_all;
    std::vector<float> w;
    screen_init_weights(ds.dpad, b.P.seed, w);
    S.w.alloc(L.n);
    S.g.alloc(L.n);
    S.m.alloc(L.n);
    S.v.alloc(L.n);
    FK_HIP(hipMemcpyAsync(S.w.p, w.data(), (size_t)L.n * 4, hipMemcpyHostToDevice, b.s));
    FK_HIP(hipMemsetAsync(S.m.p, 0, (size_t)L.n * 4, b.s));
    FK_HIP(hipMemsetAsync(S.v.p, 0, (size_t)L.n * 4, b.s));
    DevBuf<double> loss(1);
    const uint32_t B = std::min<uint32_t>(FK_SCR_BATCH, n_train);
    DevBuf<float> act((size_t)B * FK_SCR_ACT);
    const uint32_t steps = (b.P.flags & FK_FLAG_SCREEN_NOTRAIN) ? 0u : 60u;
    const uint64_t mul = coprime_multiplier(n_train, b.P.seed);
    double last_loss = 0;
    for (uint32_t st = 0; st < steps; ++st) {
        const bool report = st + 1 == steps;
        if (report) FK_HIP(hipMemsetAsync(loss.p, 0, 8, b.s));
        screen_grad_step(b.s, ds, S.w.p, ta.p, tb.p, tt.p, n_train, (uint64_t)st * B, mul, B, inv_mu, act.p, S.g.p,
                         loss.p);
        const float lr = st < steps / 2 ? 1e-2f : (st < 5 * steps / 6 ? 3e-3f : 1e-3f);
        const float bc1 = 1.f - powf(0.9f, (float)(st + 1)), bc2 = 1.f - powf(0.999f, (float)(st + 1));
        hipLaunchKernelGGL(k_scr_adam, dim3((L.n + 255) / 256), dim3(256), 0, b.s, S.w.p, S.g.p, S.m.p, S.v.p, L.n,
                           lr, bc1, bc2);
        FK_LAUNCH_CHECK();
        if (report) {
            FK_HIP(hipMemcpyAsync(&last_loss, loss.p, 8, hipMemcpyDeviceToHost, b.s));
            b.sync();
            last_loss /= B;
        }
    }
    // inference weights rounded to BF16
    FK_HIP(hipMemcpyAsync(w.data(), S.w.p, (size_t)L.n * 4, hipMemcpyDeviceToHost, b.s));
    b.sync();
    std::vector<float> wi(FK_SCR_WI_N);
    for (uint32_t i = 0; i < FK_SCR_WI_N; ++i) wi[i] = host_bf16_round(w[L.w1 + i]);
    S.wi.alloc(FK_SCR_WI_N);
    FK_HIP(hipMemcpyAsync(S.wi.p, wi.data(), (size_t)FK_SCR_WI_N * 4, hipMemcpyHostToDevice, b.s));
    // projections of every row of the build, centred, scaled by 1/sqrt(mu)
    DevBuf<float> center(ds.dpad);
    FK_HIP(hipMemsetAsync(center.p, 0, center.bytes(), b.s));
    column_mean(b.s, ds, center.p);
    S.z.alloc((size_t)N * FK_SCR_Z);
    hipLaunchKernelGGL(k_scr_project, dim3((N + 255) / 256), dim3(256), 0, b.s, ds.rows.p, ds.stride, ds.dpad, N,
                       S.w.p, center.p, sqrtf(inv_mu), S.z.p);
    FK_LAUNCH_CHECK();
    S.cnt.alloc(4);
    FK_HIP(hipMemsetAsync(S.cnt.p, 0, 32, b.s));
    b.sync();
    S.ready = true;
    b.st->screen_loss = (float)last_loss;
    b.st->screen_train_pairs = n_train;
    b.st->t_screen += now_sec() - t0;
    if (b.verbose)
        fprintf(stderr, "[fk] screen: %u training pairs, mu %g, %u steps, loss %g (%.3f s)\n", n_train, S.mu, steps,
                last_loss, now_sec() - t0);
}

// ---------------------------------------------------------------------------
// Join pass 1 with the screen: as k_join_count, but a pair whose screened
// distance clears both thresholds is not computed (distance +inf: it enters
// no bucket).  Lists longer than the LDS staging area take the unscreened
// tiled path.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_join_screen(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                     const uint32_t *__restrict__ lists,
                                                     const uint64_t *__restrict__ list_off,
                                                     const uint32_t *__restrict__ n_new,
                                                     const uint64_t *__restrict__ pair_off,
                                                     const float *__restrict__ thr, float *__restrict__ pdist,
                                                     uint32_t *__restrict__ resv, uint32_t *__restrict__ bcnt,
                                                     uint32_t N, const uint32_t *__restrict__ order, uint32_t tcap,
                                                     const uint16_t *__restrict__ Z, const float *__restrict__ wi,
                                                     float mu_keep, unsigned long long *__restrict__ counters) {
    extern __shared__ __align__(16) float s_rows[];
    float *s_thr = s_rows + FK_JOIN_LDS_ROWS_BYTES / 4;
    uint32_t *s_c = reinterpret_cast<uint32_t *>(s_thr + tcap);
    uint32_t *s_id = s_c + tcap;
    uint16_t *s_z = reinterpret_cast<uint16_t *>(s_id + tcap);  // 16-byte aligned: tcap % 4 == 0
    const uint32_t node = node_of_block(order, N);
    if (node >= N) return;
    const uint64_t lo = list_off[node];
    const uint32_t T = (uint32_t)(list_off[node + 1] - lo);
    const uint32_t nn = n_new[node];
    const uint32_t *lst = lists + lo;
    float *o = pdist + pair_off[node];
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const uint32_t id = lst[t];
        s_id[t] = id;
        s_thr[t] = thr[id];
        s_c[t] = 0;
    }
    const uint32_t ls = dpad + 4;
    const uint32_t cap = FK_JOIN_LDS_ROWS_BYTES / (ls * 4);
    uint32_t screened = 0, skipped = 0, audited = 0, missed = 0;
    if (T <= cap) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < T * 4; i += blockDim.x)
            reinterpret_cast<uint4 *>(s_z)[i] =
                reinterpret_cast<const uint4 *>(Z + (size_t)s_id[i >> 2] * FK_SCR_Z)[i & 3];
        join_stage_rows(rows, stride, dpad, s_id, 0, T, s_rows);
        __syncthreads();
        const uint32_t cells = nn * T;
        for (uint32_t c = threadIdx.x; c < cells; c += blockDim.x) {
            const uint32_t x = c / T, y = c - x * T;
            if (y <= x) continue;
            const uint32_t ix = s_id[x], iy = s_id[y];
            const float tx = s_thr[x], ty = s_thr[y], tm = fmaxf(tx, ty);
            float dd;
            bool skip = false;
            if (ix != iy && tm < INFINITY) {
                ++screened;
                skip = screen_eval(s_z + x * FK_SCR_Z, s_z + y * FK_SCR_Z, wi) * mu_keep > tm;
            }
            if (skip) {
                ++skipped;
                dd = INFINITY;
                if (fk_mix64(((uint64_t)ix << 32) ^ iy) % FK_SCR_AUDIT == 0) {
                    ++audited;
                    const float e = l2_canon(s_rows + x * ls, s_rows + y * ls, dpad);
                    missed += (e <= tx) + (e <= ty);
                }
            } else {
                dd = l2_canon(s_rows + x * ls, s_rows + y * ls, dpad);
            }
            o[join_pair_index(x, y, T)] = dd;
            if (ix != iy) {
                if (dd <= tx) atomicAdd(&s_c[x], 1u);
                if (dd <= ty) atomicAdd(&s_c[y], 1u);
            }
        }
    } else {
        __syncthreads();
        join_block(rows, stride, dpad, s_id, T, nn, s_rows, [&](uint32_t x, uint32_t y, float dd) {
            o[join_pair_index(x, y, T)] = dd;
            if (s_id[x] != s_id[y]) {
                if (dd <= s_thr[x]) atomicAdd(&s_c[x], 1u);
                if (dd <= s_thr[y]) atomicAdd(&s_c[y], 1u);
            }
        });
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const uint32_t c = s_c[t];
        resv[lo + t] = c ? atomicAdd(&bcnt[s_id[t]], c) : 0u;
    }
    const unsigned long long a0 = block_sum((unsigned long long)screened);
    const unsigned long long a1 = block_sum((unsigned long long)skipped);
    const unsigned long long a2 = block_sum((unsigned long long)audited);
    const unsigned long long a3 = block_sum((unsigned long long)missed);
    if (threadIdx.x == 0) {
        if (a0) atomicAdd(&counters[0], a0);
        if (a1) atomicAdd(&counters[1], a1);
        if (a2) atomicAdd(&counters[2], a2);
        if (a3) atomicAdd(&counters[3], a3);
    }
}

// Margin for the coming join from its own pairs (see the file comment).
#define FK_SCR_RECAL_EPS 1e-3
void screen_calibrate(BuildState &b) {
    const double t0 = now_sec();
    ScreenState &S = b.scr;
    const uint32_t q = 64, off = (uint32_t)(fk_mix64(b.P.seed + 977u * b.st->nnd_iterations) % q);
    const uint32_t nodes = (b.N + q - 1 - std::min(off, b.N)) / q;
    const uint32_t cap = 1u << 22;
    S.rho.ensure(cap);
    S.rcnt.ensure(1);
    FK_HIP(hipMemsetAsync(S.rcnt.p, 0, 4, b.s));
    if (nodes)
        hipLaunchKernelGGL(k_scr_recal, dim3(nodes), dim3(64), 0, b.s, b.ds->rows.p, b.ds->stride, b.ds->dpad,
                           b.lists.p, b.list_off.p, b.n_new.p, b.thr.p, b.N, q, off, S.z.p, S.wi.p, S.mu, cap,
                           S.rho.p, S.rcnt.p);
    FK_LAUNCH_CHECK();
    uint32_t n = 0;
    FK_HIP(hipMemcpyAsync(&n, S.rcnt.p, 4, hipMemcpyDeviceToHost, b.s));
    b.sync();
    n = std::min(n, cap);
    float m = 0.99f;  // too few admitted pairs to calibrate: skip almost nothing
    if (n >= 1000) {
        std::vector<float> r(n);
        FK_HIP(hipMemcpyAsync(r.data(), S.rho.p, (size_t)n * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
        const size_t qi = std::min<size_t>(n - 1, (size_t)((double)n * (1.0 - FK_SCR_RECAL_EPS)));
        std::nth_element(r.begin(), r.begin() + qi, r.end());
        m = std::min(std::max(r[qi], 0.f), 0.99f);
    }
    S.margin = m;
    b.st->screen_margin = m;
    b.st->screen_calib_pairs += n;
    b.st->t_screen += now_sec() - t0;
}

size_t join_screen_lds(uint32_t tcap) {
    tcap = (tcap + 3u) & ~3u;
    return FK_JOIN_LDS_ROWS_BYTES + (size_t)tcap * 12 + (size_t)tcap * FK_SCR_Z * 2;
}

void launch_join_screen(BuildState &b, uint32_t tcap) {
    tcap = (tcap + 3u) & ~3u;
    const uint32_t *ord = b.has_order ? b.order.p : nullptr;
    hipLaunchKernelGGL(k_join_screen, dim3(node_grid(b)), dim3(256), join_screen_lds(tcap), b.s, b.ds->rows.p,
                       b.ds->stride, b.ds->dpad, b.lists.p, b.list_off.p, b.n_new.p, b.pair_off.p, b.thr.p, b.pdist.p,
                       b.resv.p, b.bcnt.p, b.N, ord, tcap, b.scr.z.p, b.scr.wi.p, b.scr.mu * (1.f - b.scr.margin),
                       b.scr.cnt.p);
    FK_LAUNCH_CHECK();
}

void screen_collect(BuildState &b) {
    unsigned long long c[4];
    FK_HIP(hipMemcpyAsync(c, b.scr.cnt.p, 32, hipMemcpyDeviceToHost, b.s));
    FK_HIP(hipMemsetAsync(b.scr.cnt.p, 0, 32, b.s));
    b.sync();
    b.st->d_screened += c[0];
    b.st->d_screen_skip += c[1];
    b.st->screen_audit += c[2];
    b.st->screen_missed += c[3];
}

}  // namespace fk

using namespace fk;

extern "C" {

size_t fk_screen_param_count(uint32_t d) { return screen_layout((d + 7u) & ~7u).n; }

fk_status fk_screen_grad(fk_handle *h, const float *weights, const uint32_t *a, const uint32_t *b, const float *dist,
                         uint32_t n, float inv_mu, float *grad, double *loss) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        FK_CHECK_ARG(weights && a && b && dist && grad && loss, "null pointer");
        FK_CHECK_ARG(n >= 1 && n <= 65536, "batch must hold 1 to 65536 pairs");
        for (uint32_t i = 0; i < n; ++i)
            FK_CHECK_ARG(a[i] < h->ds.n_total && b[i] < h->ds.n_total, "row id out of range");
        select_device(h);
        const Dataset &ds = h->ds;
        const ScreenLayout L = screen_layout(ds.dpad);
        DevBuf<float> W(L.n), G(L.n), T(n), act((size_t)n * FK_SCR_ACT);
        DevBuf<uint32_t> A(n), Bv(n);
        DevBuf<double> ls(1);
        FK_HIP(hipMemcpyAsync(W.p, weights, (size_t)L.n * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(A.p, a, (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(Bv.p, b, (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(T.p, dist, (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemsetAsync(ls.p, 0, 8, h->stream));
        screen_grad_step(h->stream, ds, W.p, A.p, Bv.p, T.p, n, 0, 1, n, inv_mu, act.p, G.p, ls.p);
        FK_HIP(hipMemcpyAsync(grad, G.p, (size_t)L.n * 4, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipMemcpyAsync(loss, ls.p, 8, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipStreamSynchronize(h->stream));
        *loss /= n;
    });
}

fk_status fk_screen_predict(fk_handle *h, const float *weights, const uint32_t *a, const uint32_t *b, uint32_t n,
                            float inv_mu, float *out) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        FK_CHECK_ARG(n == 0 || (weights && a && b && out), "null pointer");
        for (uint32_t i = 0; i < n; ++i)
            FK_CHECK_ARG(a[i] < h->ds.n_total && b[i] < h->ds.n_total, "row id out of range");
        if (n == 0) return;
        select_device(h);
        const Dataset &ds = h->ds;
        const ScreenLayout L = screen_layout(ds.dpad);
        DevBuf<float> W(L.n), O(n);
        DevBuf<uint32_t> A(n), Bv(n);
        FK_HIP(hipMemcpyAsync(W.p, weights, (size_t)L.n * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(A.p, a, (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
        FK_HIP(hipMemcpyAsync(Bv.p, b, (size_t)n * 4, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_scr_predict, dim3((n + 255) / 256), dim3(256), 0, h->stream, ds.rows.p, ds.stride,
                           ds.dpad, W.p, A.p, Bv.p, n, inv_mu, O.p);
        FK_LAUNCH_CHECK();
        FK_HIP(hipMemcpyAsync(out, O.p, (size_t)n * 4, hipMemcpyDeviceToHost, h->stream));
        FK_HIP(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
