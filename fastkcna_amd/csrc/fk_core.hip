// Core kernels: pair distances, the local-join distance block, the occlusion
// pruner, prefix scans and the brute-force exact kNN.
#include "fk_device.h"
#include "fk_internal.h"
#include "fk_join.h"
#include "fk_sort.h"

namespace fk {

// ---------------------------------------------------------------------------
// Pair distances: one thread per (a[k], b[k]) pair.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_pair_distances(const float *__restrict__ rows, uint32_t stride,
                                                        uint32_t dpad, const uint32_t *__restrict__ a,
                                                        const uint32_t *__restrict__ b, size_t n,
                                                        float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = l2_canon(rows + (size_t)a[i] * stride, rows + (size_t)b[i] * stride, dpad);
}

void launch_pair_distances(hipStream_t s, const Dataset &ds, const uint32_t *a, const uint32_t *b,
                           size_t n, float *out) {
    if (n == 0) return;
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_pair_distances, dim3(grid), dim3(256), 0, s, ds.rows.p, ds.stride, ds.dpad, a, b,
                       n, out);
    FK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Local-join distance block (the NN-descent join's join_block), one
// workgroup per node.  Pair (x, y), x < n_new, x < y < T, goes to
// out[pair_off + rowstart(x) + y - x - 1], rowstart(x) = x(T-1) - x(x-1)/2.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_join_distances(const float *__restrict__ rows, uint32_t stride,
                                                        uint32_t dpad, const uint32_t *__restrict__ ids,
                                                        const uint64_t *__restrict__ list_off,
                                                        const uint32_t *__restrict__ n_new,
                                                        const uint64_t *__restrict__ pair_off,
                                                        float *__restrict__ out) {
    extern __shared__ __align__(16) float s_rows[];
    uint32_t *s_id = reinterpret_cast<uint32_t *>(s_rows + FK_JOIN_LDS_ROWS_BYTES / 4);
    const uint32_t node = blockIdx.x;
    const uint64_t lo = list_off[node];
    const uint32_t T = (uint32_t)(list_off[node + 1] - lo);
    const uint32_t nn = n_new[node];
    float *o = out + pair_off[node];
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) s_id[t] = ids[lo + t];
    __syncthreads();
    join_block(rows, stride, dpad, s_id, T, nn, s_rows,
               [&](uint32_t x, uint32_t y, float dd) { o[join_pair_index(x, y, T)] = dd; });
}

void launch_join_distances(hipStream_t s, const Dataset &ds, const uint32_t *ids, const uint64_t *list_off,
                           const uint32_t *n_new, const uint64_t *pair_off, uint32_t n_nodes, uint32_t tmax,
                           float *out) {
    if (n_nodes == 0) return;
    FK_CHECK_ARG(2ull * (ds.dpad + 4) * 4 <= FK_JOIN_LDS_ROWS_BYTES, "rows too wide for the join block");
    FK_CHECK_ARG(tmax <= 16384, "sample list longer than 16384");
    hipLaunchKernelGGL(k_join_distances, dim3(n_nodes), dim3(256), FK_JOIN_LDS_ROWS_BYTES + (size_t)tmax * 4, s,
                       ds.rows.p, ds.stride, ds.dpad, ids, list_off, n_new, pair_off, out);
    FK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Occlusion prune, one wavefront per pool.
//
// Sequential-accept semantics of *_sync_prune (src/kgraph.cpp:1809-1851):
// pool[start] is always kept (start skips a leading self entry); every later
// candidate at position < min(len, maxc) is examined in order until nsg_R are
// kept; it is dropped if an already kept entry has its id or occludes it.
// Each candidate checks the kept list in kept order and stops at the first
// occluder, as the reference does, so the decisions and the distance count
// are the reference's exactly.
// ---------------------------------------------------------------------------
#define FK_PRUNE_WPB 4
#define FK_PRUNE_RMAX 256

// One wavefront per pool; candidates go in batches of 8, one per octet
// (8 lanes), each octet holding its candidate's chains in registers and
// computing each occlusion distance as one octet distance.  Within a batch
// the accepts are resolved in candidate order with a ballot: the first
// octet that no kept neighbour occludes is accepted, the later ones are
// checked against it, and so on, which is the reference's sequential loop.
//
// Latency: the keys of batch b + 2 and the rows of batch b + 1 are loaded
// while batch b is decided, and the chains of the first PruneMC kept
// neighbours (which decide most occlusions) are kept in LDS, so a batch
// rarely waits on a dependent global load.
// Rows of the first kept neighbours cached in LDS per wave: a byte budget
// and a row cap per rule and row kind, and the number of them (FAST) an
// octet checks its own candidate against before the survivors go on one at
// a time.  Under the NSG and tau rules most candidates fall to one of the
// first few kept neighbours, and a longer check of one octet only holds the
// other seven; the alpha rule at small angles occludes rarely, so every
// candidate is checked against most of the kept list and each cached row
// saves a dependent global read.  Byte rows are small enough to cache 32,
// read by the survivor steps.  (Measured on the 1M x 128 NSG build and the
// 256-set alpha-PG batch.)
#ifndef FK_PRUNE_MC_BYTES
#define FK_PRUNE_MC_BYTES 4096
#endif
#ifndef FK_PRUNE_MC_CAP
#define FK_PRUNE_MC_CAP 8
#endif
#ifndef FK_PRUNE_MC_BYTES_ALPHA
#define FK_PRUNE_MC_BYTES_ALPHA 8192
#endif
#ifndef FK_PRUNE_MC_CAP_ALPHA
#define FK_PRUNE_MC_CAP_ALPHA 16
#endif
#ifndef FK_PRUNE_MC_BYTES8
#define FK_PRUNE_MC_BYTES8 4096
#endif
#ifndef FK_PRUNE_MC_CAP8
#define FK_PRUNE_MC_CAP8 32
#endif
#ifndef FK_PRUNE_FAST
#define FK_PRUNE_FAST 64
#endif
#ifndef FK_PRUNE_FAST8
#define FK_PRUNE_FAST8 8
#endif
#ifndef FK_PRUNE_DB
#define FK_PRUNE_DB 1  // cached kept chains double-buffered in registers
#endif
template <int RULE, int NB, int MODE>
struct PruneMC {
    static constexpr bool wide = RULE == FK_RULE_ALPHA && MODE != 2;
    static constexpr uint32_t row = (uint32_t)NB * 8u * (MODE == 2 ? 1u : 4u);
    static constexpr uint32_t budget = MODE == 2 ? FK_PRUNE_MC_BYTES8 : wide ? FK_PRUNE_MC_BYTES_ALPHA : FK_PRUNE_MC_BYTES;
    static constexpr uint32_t cap = MODE == 2 ? FK_PRUNE_MC_CAP8 : wide ? FK_PRUNE_MC_CAP_ALPHA : FK_PRUNE_MC_CAP;
    static constexpr uint32_t raw = NB == 0 ? 0u : budget / row;
    static constexpr uint32_t value = raw < cap ? raw : cap;
};

// Variant 2 caches the kept rows as halves: up to 64 rows in 8 KiB per wave
// (FK_PRUNE_MC16_BYTES; 16 KiB halves the waves per SIMD and was slower on
// every measured shape).
#ifndef FK_PRUNE_MC16_BYTES
#define FK_PRUNE_MC16_BYTES 8192
#endif
template <int NB>
struct PruneMC16 {
    static constexpr uint32_t raw = NB == 0 ? 0u : (uint32_t)FK_PRUNE_MC16_BYTES / ((uint32_t)NB * 16u);
    static constexpr uint32_t value = raw < 64u ? raw : 64u;
};
template <int RULE, int NB, int MODE, int FV>
struct PruneCache {
    static constexpr uint32_t value = FV == 2 ? PruneMC16<NB>::value : PruneMC<RULE, NB, MODE>::value;
};

// LDS per wave: the cached kept chains (mc x dpad), one staging row
// (dpad), kept ids, distances and norms (R each).
//
// MODE (RowKind): 0 canonical FP32; 1 exact-form rows (fk_gram.hip), every
// occlusion distance taken as |a|^2 + |b|^2 - 2 a.b (exact_dist),
// bit-identical to the canonical distance at half the arithmetic; 2 the same
// from byte rows (a quarter of the bytes, v_dot4_u32_u8).
__host__ __device__ inline uint32_t prune_wave_bytes(uint32_t mc, uint32_t dpad, uint32_t R, uint32_t elem) {
    return ((mc + 1u) * dpad * elem + R * 12u + 15u) & ~15u;
}

// FV (variant): 0 -- each octet checks its own candidate against the
// first cached kept neighbours, and candidates that survive those go on
// one at a time with the wave's eight octets splitting the remaining kept
// neighbours (best when most candidates fall to the first few kept
// neighbours); 1 -- each octet checks its own candidate against every kept
// neighbour, the cached ones from LDS and the rest from global memory (best
// when occlusion is rare and most candidates are checked against the whole
// kept list); 2 -- like 0, with up to 64 kept neighbours cached as
// half-precision rows: a check is decided by a rigorous bound on the
// distance to the cached halves and computed exactly from the FP32 row in
// global memory only when the bound cannot decide (for rows that no longer
// fit the cache, when the HBM re-reads of uncached kept rows set the time).
// All make the same decisions and count the same distances.
template <int RULE, int NB, int MODE, int FV>
struct PruneWave {
    using E = typename RowKind<NB, MODE>::E;
    using Chain = typename RowKind<NB, MODE>::Chain;
    static constexpr bool DOT = RowKind<NB, MODE>::dot;
    uint32_t oracle_n, R;
    float tau3, cos_thr;
    const E *lane_chains;      // chain k of row 0
    E *lane_cache;             // chain k of cached kept neighbour 0
    _Float16 *lane_cache16;    // variant 2: the same as halves
    E *lane_stage;             // chain k of the staging row
    uint32_t *so;
    float *sd, *sn;
    const float2 *nrm;  // DOT: per-row norms
    uint32_t dpad, nb, g, k, lane, limit;
    uint32_t nout;
    bool stop;
    unsigned long long cc, exam, uncomputed;

    __device__ uint64_t key_at(const uint64_t *c, uint32_t idx) const {
        return idx < limit ? c[idx] : FK_KEY_EMPTY;
    }
    __device__ const E *row_of(uint64_t key) const {
        const uint32_t id = key_id(key);
        return lane_chains + (size_t)((key != FK_KEY_EMPTY && id < oracle_n) ? id : 0u) * dpad;
    }
    __device__ float norm_of(uint64_t key) const {
        if constexpr (!DOT) return 0.f;
        const uint32_t id = key_id(key);
        return nrm[(key != FK_KEY_EMPTY && id < oracle_n) ? id : 0u].x;
    }
    // Distance of this octet's candidate to kept neighbour t: its cached row
    // from LDS, else its row from global memory (two branches, so each load
    // keeps its address space instead of becoming a flat load).
    __device__ float member_dist(const Chain &crow, float cn, uint32_t t, uint32_t tid) const {
        auto dist_to = [&](const E *o) __attribute__((always_inline)) -> float {
            if constexpr (DOT) return exact_dist(cn, sn[t], crow.dot(o, nb));
            else return crow.dist(o, nb);
        };
        if (FV != 2 && t < PruneCache<RULE, NB, MODE, FV>::value) return dist_to(lane_cache + (size_t)t * dpad);
        return dist_to(lane_chains + (size_t)tid * dpad);
    }
    // Variant 2: whether the canonical distance djk of a candidate at
    // distance pd to the kept neighbour whose half row is at distance df
    // (dist_h) occludes it: 1 / 0 when the bounds decide, -1 otherwise.
    // With x the candidate, y the kept row and y~ its halves, ||y - y~|| <=
    // rho (kept per row); df is ||x - y~||^2 in FP32 and djk ||x - y||^2 in
    // the canonical FP32 order, each within a relative g of its exact value
    // (non-negative terms); so sqrt(djk) lies within rho of sqrt(df), with g
    // and a final 2^-17 margin for the FP32 evaluation of the bounds.
    __device__ int filt_decide(float df, float rho, float pd) const {
        if (!(df < 3.0e38f)) return -1;  // a half row that overflowed
        const float gm = (float)(dpad + 32u) * 5.9604645e-8f;
        const float sq = __builtin_sqrtf(df);
        const float slo = sq * (1.f - gm) - rho, shi = sq * (1.f + gm) + rho;
        const float lo = (slo > 0.f ? slo * slo : 0.f) * (1.f - gm) * (1.f - 7.62939453125e-06f);
        const float hi = shi * shi * (1.f + gm) * (1.f + 7.62939453125e-06f);
        const float th = RULE == FK_RULE_TAU ? pd - tau3 : pd;  // the rule's comparison value
        if (lo >= th) return 0;
        if (RULE != FK_RULE_ALPHA && hi < th) return 1;
        return -1;
    }
    // Variant 2: rho of a row held in this octet's chain registers
    __device__ float half_rho(const Chain &r) const {
        const float gm = (float)(dpad + 32u) * 5.9604645e-8f;
        return (4.8828125e-4f * __builtin_sqrtf(r.norm2()) * (1.f + gm) + 2.9802322e-8f * __builtin_sqrtf((float)dpad)) *
               (1.f + 7.62939453125e-06f);
    }
    // Decides the batch at pos whose key (this octet's), chain and norm
    // (DOT) are given.
    __device__ void batch(uint32_t pos, uint64_t key, const Chain &crow, float cn) {
        const uint32_t nbat = (limit - pos) < 8u ? (limit - pos) : 8u;
        const bool valid = g < nbat;
        const uint32_t pid = key_id(key);
        const float pd = key_dist(key);
        const bool skip = !valid || pid >= oracle_n;
        // tau rule with pd - 3 tau <= 0: djk < pd - 3 tau never holds for a
        // squared distance djk >= 0, so no kept neighbour can occlude the
        // candidate and only one with its id removes it; the distances the
        // sequential loop evaluates are counted but not computed
        const bool nodist = RULE == FK_RULE_TAU && !(pd - tau3 > 0.f);
        bool occl = false;
        uint32_t mycc = 0, myfilt = 0;
        __builtin_amdgcn_wave_barrier();
        const uint32_t n0 = nout;
        constexpr uint32_t MC = PruneCache<RULE, NB, MODE, FV>::value;
        // each octet checks its own candidate against the first nA kept
        // neighbours; the survivors go on one at a time, the wave's octets
        // taking eight kept neighbours per step
        constexpr uint32_t FAST = MODE == 2 ? FK_PRUNE_FAST8 : FK_PRUNE_FAST;
        constexpr uint32_t MF = FV == 1 ? 0xFFFFFFFFu : FV == 2 ? MC : (MC < FAST ? MC : FAST);
        const uint32_t nA = n0 < MF ? n0 : MF;
        const uint32_t nC = nA < MC ? nA : MC;  // the part of the fast checks with cached rows
        if (!skip && nodist) {
            uint32_t t = 0;
            while (t < n0 && so[t] != pid) ++t;
            mycc += t;
            occl = t < n0;
        } else if (!skip) {
            // the cached kept neighbours decide most candidates; octet-
            // uniform: every lane of an octet sees the same pid, tid, djk.
            // The next cached chain is read from LDS while one is checked
            // (two register chains, unrolled by two so neither is copied;
            // chains of up to 16 registers, where the second one costs no
            // occupancy)
            if constexpr (FV == 2) {
                // the cached kept neighbours as halves: bound first, the
                // FP32 row only when the bound cannot decide
                for (uint32_t t = 0; t < nC; ++t) {
                    const uint32_t tid = so[t];
                    if (tid == pid) { occl = true; break; }
                    ++mycc;
                    const int dec = filt_decide(crow.dist_h(lane_cache16 + (size_t)t * dpad), sn[t], pd);
                    bool o;
                    if (dec >= 0) {
                        o = dec != 0;
                        ++myfilt;
                    } else {
                        o = occludes<RULE>(crow.dist(lane_chains + (size_t)tid * dpad, nb), pd, sd[t], tau3, cos_thr);
                    }
                    if (o) { occl = true; break; }
                }
            } else if constexpr (FK_PRUNE_DB && MC > 0 && sizeof(Chain) <= 64) {
                auto check = [&](uint32_t t, const Chain &o) __attribute__((always_inline)) -> bool {
                    if (so[t] == pid) return true;
                    ++mycc;
                    const float djk = DOT ? exact_dist(cn, sn[t], crow.dot_reg(o, nb)) : crow.dist_reg(o, nb);
                    return occludes<RULE>(djk, pd, sd[t], tau3, cos_thr);
                };
                Chain ca, cb;
                if (nC > 0) ca.load(lane_cache, nb);
                for (uint32_t t = 0; t < nC; t += 2) {
                    if (t + 1 < nC) cb.load(lane_cache + (size_t)(t + 1) * dpad, nb);
                    if (check(t, ca)) { occl = true; break; }
                    if (t + 1 >= nC) break;
                    if (t + 2 < nC) ca.load(lane_cache + (size_t)(t + 2) * dpad, nb);
                    if (check(t + 1, cb)) { occl = true; break; }
                }
            } else {
                for (uint32_t t = 0; t < nC; ++t) {
                    const uint32_t tid = so[t];
                    if (tid == pid) { occl = true; break; }
                    ++mycc;
                    const float djk = member_dist(crow, cn, t, tid);
                    if (occludes<RULE>(djk, pd, sd[t], tau3, cos_thr)) { occl = true; break; }
                }
            }
            if (FV && !occl) {
                // variant 1: the kept neighbours past the cached ones, rows
                // from global memory
                for (uint32_t t = nC; t < nA; ++t) {
                    const uint32_t tid = so[t];
                    if (tid == pid) { occl = true; break; }
                    ++mycc;
                    const float djk = member_dist(crow, cn, t, tid);
                    if (occludes<RULE>(djk, pd, sd[t], tau3, cos_thr)) { occl = true; break; }
                }
            }
        }
        if (n0 > nA) {
            // the rest, for the survivors one at a time, eight kept
            // neighbours per step (one per octet) so their rows load
            // together; the first occluder in kept order ends the check,
            // and the distances counted are those the sequential loop makes
            uint64_t surv = wave_ballot(!skip && !nodist && !occl && k == 0);
            while (surv) {
                const uint32_t sl = (uint32_t)__builtin_ctzll(surv);
                surv &= surv - 1;
                const uint32_t sg = sl >> 3;
                const uint32_t spid = __shfl(pid, (int)sl, 64);
                const float spd = __shfl(pd, (int)sl, 64);
                const float spn = DOT ? __shfl(cn, (int)sl, 64) : 0.f;
                const E *cp;
                if constexpr (NB > 0) {
                    if (g == sg) crow.store(lane_stage, nb);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    cp = lane_stage;
                } else {
                    cp = lane_chains + (size_t)spid * dpad;
                }
                uint32_t first = 0xFFFFFFFFu;
                bool dup = false;
                for (uint32_t t0 = nA; t0 < n0; t0 += 8) {
                    const uint32_t t = t0 + g;
                    bool hit = false, isdup = false;
                    if (t < n0) {  // octet-uniform
                        const uint32_t tid = so[t];
                        if (tid == spid) {
                            hit = isdup = true;
                        } else {
                            // a cached row from LDS, else from global memory (two
                            // branches, so each load keeps its address space)
                            auto dist_to = [&](const E *kr) __attribute__((always_inline)) -> float {
                                return DOT ? exact_dist(spn, sn[t], chain_dot2<NB>(cp, kr, nb))
                                           : chain_dist2<NB>(cp, kr, nb);
                            };
                            float djk;
                            if (MF < MC && t < MC) djk = dist_to(lane_cache + (size_t)t * dpad);
                            else djk = dist_to(lane_chains + (size_t)tid * dpad);
                            hit = occludes<RULE>(djk, spd, sd[t], tau3, cos_thr);
                        }
                    }
                    const uint64_t mh = wave_ballot(hit && k == 0);
                    if (mh) {
                        const uint32_t fo = (uint32_t)__builtin_ctzll(mh) >> 3;
                        first = t0 + fo;
                        dup = __shfl(isdup ? 1 : 0, (int)(fo * 8), 64) != 0;
                        break;
                    }
                }
                const uint32_t add = first == 0xFFFFFFFFu ? n0 - nA : first - nA + (dup ? 0u : 1u);
                if (g == sg) {
                    mycc += add;
                    occl = first != 0xFFFFFFFFu;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();  // the staging row is free again
            }
        }
        uint32_t done = 0;
        uint32_t stop_g = 7;
        while (true) {
            const uint64_t m = wave_ballot(!skip && !occl && g >= done);
            if (m == 0) break;
            const uint32_t al = (uint32_t)__builtin_ctzll(m) >> 3;
            const uint32_t aid = __shfl(pid, (int)(al * 8), 64);
            const float ad = __shfl(pd, (int)(al * 8), 64);
            const float an = DOT ? __shfl(cn, (int)(al * 8), 64) : 0.f;
            const uint32_t slot = nout;
            float arho = 0.f;
            if constexpr (FV == 2) arho = __shfl(half_rho(crow), (int)(al * 8), 64);
            if (lane == 0) {
                so[slot] = aid;
                sd[slot] = ad;
                if (DOT) sn[slot] = an;
                if (FV == 2) sn[slot] = arho;
            }
            if (slot < PruneCache<RULE, NB, MODE, FV>::value && g == al) {
                if constexpr (FV == 2) crow.store_h(lane_cache16 + (size_t)slot * dpad);
                else crow.store(lane_cache + (size_t)slot * dpad, nb);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            ++nout;
            done = al + 1;
            if (nout >= R) { stop = true; stop_g = al; break; }
            if (!skip && !occl && g > al) {
                if (aid == pid) {
                    occl = true;
                } else {
                    ++mycc;
                    if (!nodist) {
                        const float djk = member_dist(crow, cn, slot, aid);
                        if (occludes<RULE>(djk, pd, ad, tau3, cos_thr)) occl = true;
                    }
                }
            }
        }
        if (k == 0 && g <= stop_g) {
            cc += mycc;
            uncomputed += (nodist && !skip) ? mycc : myfilt;
            exam += skip ? 0u : 1u;
        }
    }
};

// Waves per SIMD the prune is compiled for (0: the compiler's choice).
#ifndef FK_PRUNE_WPE
#define FK_PRUNE_WPE 0
#endif
#if FK_PRUNE_WPE > 0
#define FK_PRUNE_ATTR __attribute__((amdgpu_waves_per_eu(FK_PRUNE_WPE)))
#else
#define FK_PRUNE_ATTR
#endif

template <int RULE, int NB, int MODE, int FV>
__global__ __launch_bounds__(256) FK_PRUNE_ATTR void k_prune(const void *__restrict__ chains_v, uint32_t dpad,
                                                             PruneLaunch a, uint32_t wave_bytes) {
    using PW = PruneWave<RULE, NB, MODE, FV>;
    using E = typename PW::E;
    using Chain = typename PW::Chain;
    constexpr bool DOT = PW::DOT;
    const E *chains = reinterpret_cast<const E *>(chains_v);
    extern __shared__ __align__(16) uint8_t smem[];
    constexpr uint32_t MC = PruneCache<RULE, NB, MODE, FV>::value;
    const uint32_t lane = lane_id();
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t vb = a.xcd_order ? xcd_position((gridDim.x + FK_NUM_XCD - 1) / FK_NUM_XCD) : blockIdx.x;
    const uint32_t task = vb * FK_PRUNE_WPB + w;
    if (task >= a.n_tasks) return;
    const uint32_t slot = a.task_list ? a.task_list[a.task_base + task] : a.task_base + task;
    const uint64_t *c = a.cand + (a.cand_off ? a.cand_off[slot] : (uint64_t)slot * a.cand_stride);
    const uint32_t len = a.cand_len[slot];
    const uint32_t q = a.qid ? a.qid[slot] : slot;
    uint32_t *oid = a.out_ids + (uint64_t)slot * a.out_stride;
    float *od = a.out_d ? a.out_d + (uint64_t)slot * a.out_stride : nullptr;
    uint32_t *olen = a.out_len + slot;
    uint8_t *wb = smem + (size_t)w * wave_bytes;

    PW pw;
    pw.oracle_n = a.oracle_n;
    pw.R = a.R;
    pw.tau3 = a.tau3;
    pw.cos_thr = a.cos_thr;
    pw.g = lane >> 3;
    pw.k = lane & 7u;
    pw.lane = lane;
    pw.dpad = dpad;
    pw.nb = NB ? (uint32_t)NB : (dpad >> 3);
    pw.lane_chains = chains + pw.k * pw.nb;
    // LDS of the wave: MC cached rows (halves for variant 2), the staging
    // row, then kept ids, distances and norms (or variant 2's row bounds)
    const size_t cache_bytes = (size_t)MC * dpad * (FV == 2 ? 2 : sizeof(E));
    pw.lane_cache = reinterpret_cast<E *>(wb) + pw.k * pw.nb;
    pw.lane_cache16 = reinterpret_cast<_Float16 *>(wb) + pw.k * pw.nb;
    pw.lane_stage = reinterpret_cast<E *>(wb + cache_bytes) + pw.k * pw.nb;
    pw.so = reinterpret_cast<uint32_t *>(wb + cache_bytes + (size_t)dpad * sizeof(E));
    pw.sd = reinterpret_cast<float *>(pw.so + a.R);
    pw.sn = pw.sd + a.R;
    pw.nrm = a.nrm;
    pw.cc = 0;
    pw.exam = 0;
    pw.uncomputed = 0;

    uint32_t start = 0;
    if (len > 0 && key_id(c[0]) == q) start = 1;
    if (start >= len || a.R == 0) {
        if (lane == 0) *olen = 0;
        return;
    }
    pw.limit = len < a.maxc ? len : a.maxc;
    {
        // the first candidate is always kept
        const uint64_t k0 = c[start];
        if (lane == 0) {
            pw.so[0] = key_id(k0);
            pw.sd[0] = key_dist(k0);
            if (DOT) pw.sn[0] = a.nrm[key_id(k0)].x;  // kept whatever its id, like its row
        }
        if (MC > 0 && pw.g == 0) {
            Chain t0;
            t0.load(pw.lane_chains + (size_t)key_id(k0) * dpad, pw.nb);
            if constexpr (FV == 2) {
                t0.store_h(pw.lane_cache16);
                const float rho = pw.half_rho(t0);
                if (lane == 0) pw.sn[0] = rho;
            } else {
                t0.store(pw.lane_cache, pw.nb);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    pw.nout = 1;
    pw.stop = pw.nout >= a.R;
    const uint32_t p0 = start + 1;
    // software pipeline, unrolled by two so the prefetched chain is never
    // copied (a copy would wait for its load)
    // invariant at the top: ka, kb = this octet's keys of batches 2j, 2j + 1;
    // ra = chain of batch 2j (loaded or in flight)
    uint64_t ka = pw.key_at(c, p0 + pw.g), kb = pw.key_at(c, p0 + 8 + pw.g);
    Chain ra, rb;
    ra.load(pw.row_of(ka), pw.nb);
    float na = pw.norm_of(ka), nbn = 0.f;
    for (uint32_t pos = p0; !pw.stop && pos < pw.limit; pos += 16) {
        rb.load(pw.row_of(kb), pw.nb);
        nbn = pw.norm_of(kb);
        const uint64_t kc = pw.key_at(c, pos + 16 + pw.g);
        pw.batch(pos, ka, ra, na);
        if (pw.stop || pos + 8 >= pw.limit) break;
        ra.load(pw.row_of(kc), pw.nb);
        na = pw.norm_of(kc);
        const uint64_t kd = pw.key_at(c, pos + 24 + pw.g);
        pw.batch(pos + 8, kb, rb, nbn);
        ka = kc;
        kb = kd;
    }
    __builtin_amdgcn_wave_barrier();
    for (uint32_t t = lane; t < pw.nout; t += 64) {
        oid[t] = pw.so[t];
        if (od) od[t] = pw.sd[t];
    }
    if (lane == 0) *olen = pw.nout;
    if (a.cc) {
        const unsigned long long s = wave_sum(pw.cc);
        if (lane == 0 && s) tally_add(a.cc, s);
    }
    if (a.examined) {
        const unsigned long long s = wave_sum(pw.exam);
        if (lane == 0 && s) tally_add(a.examined, s);
    }
    if ((RULE == FK_RULE_TAU || FV == 2) && a.uncomputed) {
        const unsigned long long s = wave_sum(pw.uncomputed);
        if (lane == 0 && s) tally_add(a.uncomputed, s);
    }
}

template <int RULE, int NB, int FV>
static void launch_prune_nbv(hipStream_t s, const Dataset &ds, const PruneLaunch &a, unsigned grid) {
    if constexpr (NB == 16 || NB == 32) {
        if (a.nrm && a.chains8) {
            const uint32_t wb = prune_wave_bytes(PruneMC<RULE, NB, 2>::value, ds.dpad, a.R, 1);
            hipLaunchKernelGGL((k_prune<RULE, NB, 2, FV>), dim3(grid), dim3(64 * FK_PRUNE_WPB),
                               (size_t)wb * FK_PRUNE_WPB, s, (const void *)a.chains8, ds.dpad, a, wb);
            return;
        }
    }
    const uint32_t wave_bytes = prune_wave_bytes(PruneMC<RULE, NB, 0>::value, ds.dpad, a.R, 4);  // modes 0 and 1
    if constexpr (NB == 12 || NB == 16 || NB == 32) {
        if (a.nrm) {
            hipLaunchKernelGGL((k_prune<RULE, NB, 1, FV>), dim3(grid), dim3(64 * FK_PRUNE_WPB),
                               (size_t)wave_bytes * FK_PRUNE_WPB, s, (const void *)ds.chains.p, ds.dpad, a,
                               wave_bytes);
            return;
        }
    }
    hipLaunchKernelGGL((k_prune<RULE, NB, 0, FV>), dim3(grid), dim3(64 * FK_PRUNE_WPB),
                       (size_t)wave_bytes * FK_PRUNE_WPB, s, (const void *)ds.chains.p, ds.dpad, a, wave_bytes);
}

template <int RULE, int NB>
static void launch_prune_nb(hipStream_t s, const Dataset &ds, const PruneLaunch &a, unsigned grid) {
    if constexpr (NB == 12 || NB == 16 || NB == 32) {
        // variant 2: FP32 rows (the exact forms' rows are cheaper already)
        if (a.variant == 2 && !a.nrm) {
            constexpr uint32_t MC = PruneCache<RULE, NB, 0, 2>::value;
            const uint32_t wb = (uint32_t)(((size_t)MC * ds.dpad * 2 + (size_t)ds.dpad * 4 + (size_t)a.R * 12 + 15) & ~15ull);
            hipLaunchKernelGGL((k_prune<RULE, NB, 0, 2>), dim3(grid), dim3(64 * FK_PRUNE_WPB), (size_t)wb * FK_PRUNE_WPB,
                               s, (const void *)ds.chains.p, ds.dpad, a, wb);
            return;
        }
    }
    if (a.variant == 1) launch_prune_nbv<RULE, NB, 1>(s, ds, a, grid);
    else launch_prune_nbv<RULE, NB, 0>(s, ds, a, grid);  // variant 2 where it does not apply runs as 0
}

template <int RULE>
static void launch_prune_rule(hipStream_t s, const Dataset &ds, const PruneLaunch &a, unsigned grid) {
    switch (ds.dpad) {
        case 32: launch_prune_nb<RULE, 4>(s, ds, a, grid); break;
        case 64: launch_prune_nb<RULE, 8>(s, ds, a, grid); break;
        case 96: launch_prune_nb<RULE, 12>(s, ds, a, grid); break;
        case 128: launch_prune_nb<RULE, 16>(s, ds, a, grid); break;
        case 256: launch_prune_nb<RULE, 32>(s, ds, a, grid); break;
        default: launch_prune_nb<RULE, 0>(s, ds, a, grid); break;
    }
}

// Chain-major copy of the rows (layout of OctChain).
__global__ void k_chain_layout(const float *__restrict__ rows, uint32_t stride, uint64_t n, uint32_t dpad,
                               float *__restrict__ chains) {
    const uint32_t nb = dpad >> 3;
    const uint64_t total = n * dpad;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t i = e / dpad;
        const uint32_t x = (uint32_t)(e - i * dpad);  // position in the chain-major row
        const uint32_t kk = x / nb, j = x - kk * nb;
        chains[e] = rows[i * stride + chain_block(j, dpad) + kk];
    }
}

void build_chains(hipStream_t s, Dataset &ds) {
    ds.chains.alloc((size_t)ds.n_total * ds.dpad);
    const uint64_t total = (uint64_t)ds.n_total * ds.dpad;
    const unsigned grid = (unsigned)std::min<uint64_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(k_chain_layout, dim3(grid), dim3(256), 0, s, ds.rows.p, ds.stride, (uint64_t)ds.n_total,
                       ds.dpad, ds.chains.p);
    FK_LAUNCH_CHECK();
}

// Order-independent digest of a prune's output lists (ids, distance bits
// and positions of every task's list, and its length), summed over tasks:
// two launches that wrote the same lists give the same digest.
__device__ __forceinline__ uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

__global__ __launch_bounds__(256) void k_prune_digest(PruneLaunch a, unsigned long long *__restrict__ out) {
    const uint32_t task = blockIdx.x * 4 + (threadIdx.x >> 6);
    uint64_t h = 0;
    if (task < a.n_tasks) {
        const uint32_t slot = a.task_list ? a.task_list[a.task_base + task] : a.task_base + task;
        const uint32_t len = a.out_len[slot];
        const uint32_t *ids = a.out_ids + (uint64_t)slot * a.out_stride;
        const float *d = a.out_d ? a.out_d + (uint64_t)slot * a.out_stride : nullptr;
        for (uint32_t k = lane_id(); k < len; k += 64) {
            const uint64_t v = ((uint64_t)ids[k] << 32) | (d ? __float_as_uint(d[k]) : 0u);
            h += fmix64(v ^ ((uint64_t)slot * 0x9E3779B97F4A7C15ull) ^ ((uint64_t)k << 20));
        }
        if (lane_id() == 0) h += fmix64(((uint64_t)slot << 32) | len);
    }
    h = wave_sum(h);
    if (lane_id() == 0 && h) atomicAdd(out, (unsigned long long)h);
}

void launch_prune_digest(hipStream_t s, const PruneLaunch &a, unsigned long long *out) {
    if (a.n_tasks == 0) return;
    hipLaunchKernelGGL(k_prune_digest, dim3((a.n_tasks + 3) / 4), dim3(256), 0, s, a, out);
    FK_LAUNCH_CHECK();
}

bool prune_variant_applies(const Dataset &ds, const PruneLaunch &a, int v) {
    switch (v) {
        case 0:
        case 1: return true;
        case 2: return (ds.dpad == 96 || ds.dpad == 128 || ds.dpad == 256) && !a.nrm;
        case 3: return prune_gram_applies(ds, a);
        default: return false;
    }
}

void launch_prune(hipStream_t s, const Dataset &ds, const PruneLaunch &a) {
    if (a.n_tasks == 0) return;
    FK_CHECK_ARG(a.R <= FK_PRUNE_RMAX, "nsg_R above the supported maximum (256)");
    FK_CHECK_ARG(a.variant >= 0 && a.variant < FK_PRUNE_VARIANTS, "unknown prune variant");
    unsigned grid = (a.n_tasks + FK_PRUNE_WPB - 1) / FK_PRUNE_WPB;
    if (a.xcd_order) grid = (grid + FK_NUM_XCD - 1) / FK_NUM_XCD * FK_NUM_XCD;
    if (a.variant == 3 && launch_prune_gram(s, ds, a)) return;  // else it runs as variant 0
    switch (a.rule) {
        case FK_RULE_NSG: launch_prune_rule<FK_RULE_NSG>(s, ds, a, grid); break;
        case FK_RULE_TAU: launch_prune_rule<FK_RULE_TAU>(s, ds, a, grid); break;
        default: launch_prune_rule<FK_RULE_ALPHA>(s, ds, a, grid); break;
    }
    FK_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Exclusive prefix sums (three phases: block sums, scan of block sums, fix-up).
// ---------------------------------------------------------------------------
#define FK_SCAN_BLOCK 256
#define FK_SCAN_PER_THREAD 8
#define FK_SCAN_TILE (FK_SCAN_BLOCK * FK_SCAN_PER_THREAD)

template <typename T>
__device__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *s_tmp, uint64_t *total) {
    // Hillis-Steele over the block in LDS.
    s_tmp[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t o = 1; o < blockDim.x; o <<= 1) {
        const uint64_t t = threadIdx.x >= o ? s_tmp[threadIdx.x - o] : 0;
        __syncthreads();
        s_tmp[threadIdx.x] += t;
        __syncthreads();
    }
    const uint64_t incl = s_tmp[threadIdx.x];
    *total = s_tmp[blockDim.x - 1];
    __syncthreads();
    return incl - v;
}

template <typename TI>
__global__ __launch_bounds__(FK_SCAN_BLOCK) void k_scan_partials(const TI *__restrict__ in, size_t n,
                                                                 uint64_t *__restrict__ part) {
    __shared__ uint64_t s_tmp[FK_SCAN_BLOCK];
    const size_t base = (size_t)blockIdx.x * FK_SCAN_TILE + (size_t)threadIdx.x * FK_SCAN_PER_THREAD;
    uint64_t v = 0;
    for (int k = 0; k < FK_SCAN_PER_THREAD; ++k)
        if (base + k < n) v += (uint64_t)in[base + k];
    uint64_t tot;
    block_exclusive_scan<uint64_t>(v, s_tmp, &tot);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ __launch_bounds__(FK_SCAN_BLOCK) void k_scan_single(uint64_t *__restrict__ part, size_t n,
                                                               uint64_t *__restrict__ total) {
    __shared__ uint64_t s_tmp[FK_SCAN_BLOCK];
    uint64_t carry = 0;
    for (size_t base = 0; base < n; base += FK_SCAN_BLOCK) {
        const size_t i = base + threadIdx.x;
        const uint64_t v = i < n ? part[i] : 0;
        uint64_t tot;
        const uint64_t ex = block_exclusive_scan<uint64_t>(v, s_tmp, &tot);
        if (i < n) part[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

template <typename TI>
__global__ __launch_bounds__(FK_SCAN_BLOCK) void k_scan_apply(const TI *__restrict__ in, size_t n,
                                                              const uint64_t *__restrict__ part,
                                                              uint64_t *__restrict__ out) {
    __shared__ uint64_t s_tmp[FK_SCAN_BLOCK];
    const size_t base = (size_t)blockIdx.x * FK_SCAN_TILE + (size_t)threadIdx.x * FK_SCAN_PER_THREAD;
    TI vals[FK_SCAN_PER_THREAD];
    uint64_t v = 0;
    for (int k = 0; k < FK_SCAN_PER_THREAD; ++k) {
        vals[k] = base + k < n ? in[base + k] : (TI)0;
        v += (uint64_t)vals[k];
    }
    uint64_t tot;
    uint64_t run = block_exclusive_scan<uint64_t>(v, s_tmp, &tot) + part[blockIdx.x];
    for (int k = 0; k < FK_SCAN_PER_THREAD; ++k) {
        if (base + k < n) out[base + k] = run;
        run += (uint64_t)vals[k];
    }
}

template <typename TI>
static void exclusive_scan_impl(hipStream_t s, const TI *in, uint64_t *out, size_t n, uint64_t *total,
                                DevBuf<uint64_t> &part) {
    if (n == 0) {
        if (total) FK_HIP(hipMemsetAsync(total, 0, sizeof(uint64_t), s));
        return;
    }
    const size_t nb = (n + FK_SCAN_TILE - 1) / FK_SCAN_TILE;
    if (part.n < nb) {
        FK_HIP(hipStreamSynchronize(s));  // the old scratch may still be in use
        part.alloc(nb);
    }
    hipLaunchKernelGGL(k_scan_partials<TI>, dim3((unsigned)nb), dim3(FK_SCAN_BLOCK), 0, s, in, n, part.p);
    FK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_single, dim3(1), dim3(FK_SCAN_BLOCK), 0, s, part.p, nb, total);
    FK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_apply<TI>, dim3((unsigned)nb), dim3(FK_SCAN_BLOCK), 0, s, in, n, part.p, out);
    FK_LAUNCH_CHECK();
}

void exclusive_scan_u32_to_u64(hipStream_t s, const uint32_t *in, uint64_t *out, size_t n, uint64_t *total,
                               DevBuf<uint64_t> &tmp) {
    exclusive_scan_impl<uint32_t>(s, in, out, n, total, tmp);
}
void exclusive_scan_u64(hipStream_t s, const uint64_t *in, uint64_t *out, size_t n, uint64_t *total,
                        DevBuf<uint64_t> &tmp) {
    exclusive_scan_impl<uint64_t>(s, in, out, n, total, tmp);
}

// ---------------------------------------------------------------------------
// Brute-force exact kNN.
// Phase 1: keys[q][i] = (dist(q, row i), i) for a batch of queries; a block
// stages 64 data rows in LDS and sweeps the batch's queries 4 at a time.
// Phase 2: one block per query finds the K-th smallest 64-bit key by an
// 8-pass byte radix select, gathers the K keys <= it and sorts them.
// ---------------------------------------------------------------------------
#define FK_BF_ROWS 64

__global__ __launch_bounds__(256) void k_bf_keys(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                 uint32_t n_used, const float *__restrict__ qv,
                                                 uint32_t q_stride, uint32_t n_q,
                                                 const uint32_t *__restrict__ self_ids,
                                                 uint64_t *__restrict__ keys) {
    extern __shared__ __align__(16) float s_rows[];
    const uint32_t ls = dpad + 4;
    const uint32_t r0 = blockIdx.x * FK_BF_ROWS;
    const uint32_t nr = (n_used - r0) < FK_BF_ROWS ? (n_used - r0) : FK_BF_ROWS;
    const uint32_t v4 = dpad >> 2;
    for (uint32_t i = threadIdx.x; i < nr * v4; i += blockDim.x) {
        const uint32_t r = i / v4, c = i - r * v4;
        *reinterpret_cast<float4 *>(s_rows + r * ls + c * 4) =
            *reinterpret_cast<const float4 *>(rows + (size_t)(r0 + r) * stride + c * 4);
    }
    __syncthreads();
    const uint32_t r = threadIdx.x & 63u;
    for (uint32_t q = threadIdx.x >> 6; q < n_q; q += 4) {
        if (r < nr) {
            const uint32_t id = r0 + r;
            uint64_t k;
            if (self_ids && self_ids[q] == id) k = FK_KEY_EMPTY;
            else k = make_key(l2_canon(qv + (size_t)q * q_stride, s_rows + r * ls, dpad), id);
            keys[(size_t)q * n_used + id] = k;
        }
    }
}

// Wide rows (the 64-row LDS tile of k_bf_keys does not fit): one octet per
// row, rows read once per group of FK_BF_QG queries from the chain-major
// copy, the group's queries staged in LDS in chain-major form.
#define FK_BF_QG 8
#define FK_BF_WROWS 256

__global__ __launch_bounds__(256) void k_bf_keys_wide(const float *__restrict__ chains, uint32_t dpad,
                                                      uint32_t n_used, const float *__restrict__ qv,
                                                      uint32_t q_stride, uint32_t n_q,
                                                      const uint32_t *__restrict__ self_ids,
                                                      uint64_t *__restrict__ keys) {
    extern __shared__ __align__(16) float s_q[];  // FK_BF_QG x dpad, chain-major
    const uint32_t oc = threadIdx.x >> 3, k = threadIdx.x & 7u, nb = dpad >> 3;
    const uint32_t r0 = blockIdx.x * FK_BF_WROWS;
    for (uint32_t q0 = 0; q0 < n_q; q0 += FK_BF_QG) {
        const uint32_t nq = (n_q - q0) < FK_BF_QG ? (n_q - q0) : FK_BF_QG;
        __syncthreads();  // the previous group is done with s_q
        for (uint32_t e = threadIdx.x; e < nq * dpad; e += blockDim.x) {
            const uint32_t qq = e / dpad, x = e - qq * dpad;
            const uint32_t kk = x / nb, j = x - kk * nb;
            s_q[e] = qv[(size_t)(q0 + qq) * q_stride + chain_block(j, dpad) + kk];
        }
        __syncthreads();
        for (uint32_t rr = oc; rr < FK_BF_WROWS; rr += (blockDim.x >> 3)) {
            const uint32_t id = r0 + rr;
            if (id >= n_used) break;  // octet-uniform
            const float *rc = chains + (size_t)id * dpad + k * nb;
            float acc[FK_BF_QG];
#pragma unroll
            for (uint32_t qq = 0; qq < FK_BF_QG; ++qq) acc[qq] = 0.f;
            for (uint32_t j = 0; j < nb; ++j) {
                const float r = rc[j];
#pragma unroll
                for (uint32_t qq = 0; qq < FK_BF_QG; ++qq) {
                    const float t = s_q[qq * dpad + k * nb + j] - r;
                    acc[qq] = __builtin_fmaf(t, t, acc[qq]);
                }
            }
#pragma unroll
            for (uint32_t qq = 0; qq < FK_BF_QG; ++qq) {
                const float dd = octet_sum(acc[qq]);
                if (qq < nq && k == 0) {
                    const uint32_t q = q0 + qq;
                    keys[(size_t)q * n_used + id] =
                        (self_ids && self_ids[q] == id) ? FK_KEY_EMPTY : make_key(dd, id);
                }
            }
        }
    }
}

#define FK_SEL_THREADS 1024
#define FK_SEL_MAXK 2048

// The K smallest keys of kq[0, n) (all of them when n < K, then padded with
// FK_KEY_EMPTY) in ascending order, handed to emit(i, key) for i < K: a
// radix select on the key bytes (one 256-bin histogram per byte, most
// significant first) finds the K-th key, the keys up to it are gathered and
// sorted.  The whole workgroup calls it.
template <class Emit>
__device__ __forceinline__ void topk_block(const uint64_t *__restrict__ kq, uint32_t n, uint32_t K, Emit emit) {
    __shared__ uint32_t hist[256];
    __shared__ uint64_t s_sel[FK_SEL_MAXK];
    __shared__ uint32_t s_cnt;
    __shared__ uint64_t s_prefix, s_mask;
    __shared__ uint32_t s_need;
    if (threadIdx.x == 0) { s_prefix = 0; s_mask = 0; s_need = K; s_cnt = 0; }
    __syncthreads();
    for (int byte = 7; byte >= 0; --byte) {
        for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) hist[i] = 0;
        __syncthreads();
        const uint64_t prefix = s_prefix, mask = s_mask;
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
            const uint64_t k = kq[i];
            if ((k & mask) == prefix) atomicAdd(&hist[(uint32_t)(k >> (8 * byte)) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t need = s_need, b = 0;
            for (; b < 256; ++b) {
                if (hist[b] >= need) break;
                need -= hist[b];
            }
            if (b == 256) b = 255;  // fewer than K keys: keep everything
            s_prefix = prefix | ((uint64_t)b << (8 * byte));
            s_mask = mask | (0xFFull << (8 * byte));
            s_need = need;
        }
        __syncthreads();
    }
    const uint64_t kth = s_prefix;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint64_t k = kq[i];
        if (k <= kth && k != FK_KEY_EMPTY) {
            const uint32_t p = atomicAdd(&s_cnt, 1u);
            if (p < FK_SEL_MAXK) s_sel[p] = k;
        }
    }
    __syncthreads();
    const uint32_t cnt = s_cnt < FK_SEL_MAXK ? s_cnt : FK_SEL_MAXK;
    const uint32_t np = next_pow2(cnt < 2 ? 2 : cnt);
    for (uint32_t i = cnt + threadIdx.x; i < np; i += blockDim.x) s_sel[i] = FK_KEY_EMPTY;
    __syncthreads();
    block_bitonic_sort(s_sel, np);
    for (uint32_t i = threadIdx.x; i < K; i += blockDim.x) emit(i, i < cnt ? s_sel[i] : FK_KEY_EMPTY);
}

// One workgroup per query: the K nearest of its n keys as ids (and distances).
__global__ __launch_bounds__(FK_SEL_THREADS) void k_topk_select(const uint64_t *__restrict__ keys, uint32_t n,
                                                                uint32_t K, uint32_t *__restrict__ out_ids,
                                                                float *__restrict__ out_d) {
    topk_block(keys + (size_t)blockIdx.x * n, n, K, [&](uint32_t i, uint64_t k) {
        out_ids[(size_t)blockIdx.x * K + i] = key_id(k);
        if (out_d) out_d[(size_t)blockIdx.x * K + i] = key_dist(k);
    });
}

// First stage for long key rows: workgroup (c, q) keeps the K smallest keys
// of chunk c (len keys) of query q's row in part[(q * C + c) * K ...]; the
// K smallest of the union of the chunks' K smallest are the row's K
// smallest (the keys are distinct), which k_topk_select then takes from the
// C * K kept keys.
__global__ __launch_bounds__(FK_SEL_THREADS) void k_topk_chunks(const uint64_t *__restrict__ keys, uint32_t n,
                                                                uint32_t len, uint32_t K,
                                                                uint64_t *__restrict__ part) {
    const uint32_t c = blockIdx.x, q = blockIdx.y, C = gridDim.x;
    const uint32_t lo = c * len;
    const uint32_t m = lo < n ? (n - lo < len ? n - lo : len) : 0u;
    uint64_t *o = part + ((size_t)q * C + c) * K;
    topk_block(keys + (size_t)q * n + lo, m, K, [&](uint32_t i, uint64_t k) { o[i] = k; });
}

__global__ void k_add_u32(uint32_t *__restrict__ v, size_t n, uint32_t add) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] += add;
}

void bruteforce_knn(hipStream_t s, const Dataset &ds, const float *d_queries, uint32_t q_stride, uint32_t n_q,
                    uint32_t n_used, uint32_t K, const uint32_t *d_self, uint32_t *d_out_ids, float *d_out_dists,
                    uint32_t base) {
    FK_CHECK_ARG(K >= 1 && K <= FK_SEL_MAXK, "K must be in [1, 2048]");
    FK_CHECK_ARG(n_used >= 1, "empty dataset");
    const size_t lds = (size_t)FK_BF_ROWS * (ds.dpad + 4) * sizeof(float);
    const bool wide = lds > 64 * 1024;
    const size_t lds_wide = (size_t)FK_BF_QG * ds.dpad * sizeof(float);
    FK_CHECK_ARG(!wide || lds_wide <= 64 * 1024, "dimension too large for the brute-force kernel (d <= 2048)");
    // Batch queries so the key matrix stays within ~2 GiB.
    size_t qb = ((size_t)2 << 30) / ((size_t)n_used * 8);
    if (qb < 1) qb = 1;
    if (qb > n_q) qb = n_q;
    DevBuf<uint64_t> keys(qb * (size_t)n_used);
    DevBuf<uint64_t> part;  // the chunks' kept keys
    for (uint32_t q0 = 0; q0 < n_q; q0 += (uint32_t)qb) {
        const uint32_t nq = (uint32_t)((n_q - q0) < qb ? (n_q - q0) : qb);
        if (wide) {
            const unsigned gx = (n_used + FK_BF_WROWS - 1) / FK_BF_WROWS;
            hipLaunchKernelGGL(k_bf_keys_wide, dim3(gx), dim3(256), lds_wide, s, ds.chains.p + (size_t)base * ds.dpad,
                               ds.dpad, n_used,
                               d_queries + (size_t)q0 * q_stride, q_stride, nq, d_self ? d_self + q0 : nullptr,
                               keys.p);
        } else {
            const unsigned gx = (n_used + FK_BF_ROWS - 1) / FK_BF_ROWS;
            hipLaunchKernelGGL(k_bf_keys, dim3(gx), dim3(256), lds, s, ds.rows.p + (size_t)base * ds.stride, ds.stride,
                               ds.dpad, n_used,
                               d_queries + (size_t)q0 * q_stride, q_stride, nq, d_self ? d_self + q0 : nullptr,
                               keys.p);
        }
        FK_LAUNCH_CHECK();
        // rows much longer than K: chunks first, so that about a thousand
        // workgroups share the select instead of one per query
        const uint32_t C = std::min<uint32_t>((1024u + nq - 1) / nq, n_used / (16u * K));
        if (C >= 2) {
            const uint32_t len = (n_used + C - 1) / C;
            part.ensure((size_t)nq * C * K);
            hipLaunchKernelGGL(k_topk_chunks, dim3(C, nq), dim3(FK_SEL_THREADS), 0, s, keys.p, n_used, len, K,
                               part.p);
            FK_LAUNCH_CHECK();
            hipLaunchKernelGGL(k_topk_select, dim3(nq), dim3(FK_SEL_THREADS), 0, s, part.p, C * K, K,
                               d_out_ids + (size_t)q0 * K, d_out_dists ? d_out_dists + (size_t)q0 * K : nullptr);
        } else {
            hipLaunchKernelGGL(k_topk_select, dim3(nq), dim3(FK_SEL_THREADS), 0, s, keys.p, n_used, K,
                               d_out_ids + (size_t)q0 * K, d_out_dists ? d_out_dists + (size_t)q0 * K : nullptr);
        }
        FK_LAUNCH_CHECK();
    }
    if (base) {
        const size_t n = (size_t)n_q * K;
        hipLaunchKernelGGL(k_add_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_out_ids, n, base);
        FK_LAUNCH_CHECK();
    }
    FK_HIP(hipStreamSynchronize(s));
}

}  // namespace fk
