// The FastKCNA refinement loops on the GPU and the fk_build / fk_nndescent
// entry points.
//
// Per loop (buildNSG / buildtauMG / buildAlphaPG / buildHNSWLayer,
// src/kgraph.cpp:2314-2337, 3143-3154, 3482-3493):
//   Link      prune every pool into an out-list (full rule), add reverse
//             edges, re-prune over-full lists (nsg_Link :2087-2143)
//   ComputeEp centroid + beam search (fk_search.hip)
//   tree_grow connectivity (fk_search.hip; not for HNSW layers)
//   UpdateNhoods  isnew bits of the out-lists (:3537-3584)
// and between loops BridgeView (fk_nnd.hip).
//
// The pruner always applies the full rule (*_sync_prune).  The reference
// switches to the incremental *_sync_prune2 for nodes that were not
// re-pruned in the previous loop; the paper states that the two agree
// (result-preserving Opt2), and the CPU oracle diffs them.
#include <math.h>
#include <stdlib.h>

#include <algorithm>

#include "fk_api_impl.h"
#include "fk_build.h"
#include "fk_device.h"
#include "fk_sort.h"

namespace fk {

// Parameter shrinking of the KGraphConstructor constructor
// (src/kgraph.cpp:1288-1334) and MAXL.
BuildParams shrink_params(const fk_params &p, uint32_t N) {
    BuildParams P;
    P.K = p.K; P.L = p.L; P.S = p.S; P.R = p.R; P.iterations = p.iterations; P.controls = p.controls;
    P.seed = p.seed; P.delta = p.delta; P.recall = p.recall; P.nsg_R = p.nsg_R; P.search_L = p.search_L;
    P.search_K = p.search_K; P.step = p.step; P.massq_S = p.massq_S; P.loop_i = p.loop_i; P.tau = p.tau;
    P.angle = p.angle; P.flags = p.flags;
    if (N <= P.controls) P.controls = N - 1;
    if (N <= P.L) P.L = N - 1;
    if (N <= P.S) P.S = N - 1;
    if (N <= P.search_L) P.search_L = N - 1;
    if (N <= P.search_K) P.search_K = N - 1;
    if (N <= P.K) {
        P.K = N - 1;
        P.loop_i = 0;
        P.S = N - 1;
        P.iterations = 0;
    }
    P.MAXL = std::max(std::max(P.L, P.K), std::max(P.search_L, P.search_K));
    return P;
}

// ---------------------------------------------------------------------------
// Reverse edges (InterInsert, src/kgraph.cpp:2051-2084): every edge n -> t
// offers n to t's list unless n is already there.  In-edges are gathered by
// count / scan / fill as packed (d(n, t), n) keys.
// ---------------------------------------------------------------------------
__global__ void k_indeg(const uint32_t *__restrict__ nsg_id, const uint32_t *__restrict__ nsg_len, uint32_t N,
                        uint32_t ncap, uint32_t *__restrict__ indeg) {
    const uint32_t n = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (n >= N) return;
    const uint32_t deg = nsg_len[n];
    for (uint32_t k = lane_id(); k < deg; k += 64) atomicAdd(&indeg[nsg_id[(uint64_t)n * ncap + k]], 1u);
}

__global__ void k_in_fill(const uint32_t *__restrict__ nsg_id, const float *__restrict__ nsg_d,
                          const uint32_t *__restrict__ nsg_len, uint32_t N, uint32_t ncap,
                          const uint64_t *__restrict__ in_off, uint32_t *__restrict__ cur, uint64_t *__restrict__ in_key) {
    const uint32_t n = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (n >= N) return;
    const uint32_t deg = nsg_len[n];
    for (uint32_t k = lane_id(); k < deg; k += 64) {
        const uint32_t t = nsg_id[(uint64_t)n * ncap + k];
        const uint64_t slot = in_off[t] + atomicAdd(&cur[t], 1u);
        in_key[slot] = make_key(nsg_d[(uint64_t)n * ncap + k], n);
    }
}

#define FK_INTER_MAX 4096

// Out-list A of node des plus the in-edges B not already in A (one
// workgroup).  If the union fits nsg_R it is A followed by B in (dist, id)
// order; otherwise the union, sorted, is staged for the reversed prune.
__device__ void inter_merge_node(uint32_t des, uint32_t ncap, uint32_t R, uint32_t *__restrict__ nsg_id,
                                 float *__restrict__ nsg_d, uint32_t *__restrict__ nsg_len,
                                 const uint32_t *__restrict__ indeg, const uint64_t *__restrict__ in_off,
                                 const uint64_t *__restrict__ in_key, uint64_t *__restrict__ rp_key,
                                 uint64_t *__restrict__ rp_off, uint32_t *__restrict__ rp_len,
                                 uint32_t *__restrict__ task_list, uint32_t *__restrict__ n_tasks,
                                 uint32_t *__restrict__ err) {
    __shared__ uint64_t s_b[FK_INTER_MAX];
    __shared__ uint32_t s_a[FK_NSG_SLACK + 256];
    __shared__ uint32_t s_nb;
    const uint32_t nA = nsg_len[des];
    const uint32_t nB = indeg[des];
    if (nB == 0) {
        if (threadIdx.x == 0) rp_len[des] = 0;
        return;
    }
    if (nA + nB > FK_INTER_MAX) {
        if (threadIdx.x == 0) { atomicMax(err, nA + nB); rp_len[des] = 0; }
        return;
    }
    uint32_t *ai = nsg_id + (uint64_t)des * ncap;
    float *ad = nsg_d + (uint64_t)des * ncap;
    for (uint32_t k = threadIdx.x; k < nA; k += blockDim.x) s_a[k] = ai[k];
    if (threadIdx.x == 0) s_nb = 0;
    __syncthreads();
    const uint64_t *bk = in_key + in_off[des];
    // B minus A (reference: dedupe against the current list, :2063-2076)
    for (uint32_t k = threadIdx.x; k < nB; k += blockDim.x) {
        const uint64_t key = bk[k];
        const uint32_t src = key_id(key);
        bool dup = false;
        for (uint32_t j = 0; j < nA; ++j)
            if (s_a[j] == src) { dup = true; break; }
        if (!dup) s_b[atomicAdd(&s_nb, 1u)] = key;
    }
    __syncthreads();
    const uint32_t nb = s_nb;
    if (nb == 0) {
        if (threadIdx.x == 0) rp_len[des] = 0;
        return;
    }
    if (nA + nb <= R) {
        const uint32_t np2 = next_pow2(nb < 2 ? 2 : nb);
        for (uint32_t k = nb + threadIdx.x; k < np2; k += blockDim.x) s_b[k] = FK_KEY_EMPTY;
        __syncthreads();
        block_bitonic_sort(s_b, np2);
        for (uint32_t k = threadIdx.x; k < nb; k += blockDim.x) {
            ai[nA + k] = key_id(s_b[k]);
            ad[nA + k] = key_dist(s_b[k]);
        }
        if (threadIdx.x == 0) { nsg_len[des] = nA + nb; rp_len[des] = 0; }
        return;
    }
    // over-full: stage A U B sorted for the reversed prune
    const uint32_t tot = nA + nb;
    for (uint32_t k = threadIdx.x; k < nA; k += blockDim.x) s_b[nb + k] = make_key(ad[k], s_a[k]);
    const uint32_t np2 = next_pow2(tot);
    for (uint32_t k = tot + threadIdx.x; k < np2; k += blockDim.x) s_b[k] = FK_KEY_EMPTY;
    __syncthreads();
    block_bitonic_sort(s_b, np2);
    const uint64_t off = in_off[des] + (uint64_t)des * R;
    for (uint32_t k = threadIdx.x; k < tot; k += blockDim.x) rp_key[off + k] = s_b[k];
    if (threadIdx.x == 0) {
        rp_off[des] = off;
        rp_len[des] = tot;
    }
}

// The nodes k_inter_small leaves (unions above FK_INTER_SMALL entries),
// one workgroup at a time each.
__global__ __launch_bounds__(256) void k_inter_merge(const uint32_t *__restrict__ list,
                                                     const uint32_t *__restrict__ n_list, uint32_t ncap, uint32_t R,
                                                     uint32_t *__restrict__ nsg_id, float *__restrict__ nsg_d,
                                                     uint32_t *__restrict__ nsg_len,
                                                     const uint32_t *__restrict__ indeg,
                                                     const uint64_t *__restrict__ in_off,
                                                     const uint64_t *__restrict__ in_key, uint64_t *__restrict__ rp_key,
                                                     uint64_t *__restrict__ rp_off, uint32_t *__restrict__ rp_len,
                                                     uint32_t *__restrict__ task_list, uint32_t *__restrict__ n_tasks,
                                                     uint32_t *__restrict__ err) {
    const uint32_t n = *n_list;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        inter_merge_node(list[i], ncap, R, nsg_id, nsg_d, nsg_len, indeg, in_off, in_key, rp_key, rp_off, rp_len,
                         task_list, n_tasks, err);
        __syncthreads();
    }
}

// The same merge for the common node, one wavefront per node, when the
// union has at most FK_INTER_SMALL entries: B minus A compacted in order,
// ranked by counting (keys are distinct: one in-edge per source), written
// straight to its slots.  Larger nodes are listed for k_inter_merge.
#define FK_INTER_SMALL 256
#define FK_INTER_WPB 4
__global__ __launch_bounds__(64 * FK_INTER_WPB) void k_inter_small(
    uint32_t N, uint32_t ncap, uint32_t R, uint32_t *__restrict__ nsg_id, float *__restrict__ nsg_d,
    uint32_t *__restrict__ nsg_len, const uint32_t *__restrict__ indeg, const uint64_t *__restrict__ in_off,
    const uint64_t *__restrict__ in_key, uint64_t *__restrict__ rp_key, uint64_t *__restrict__ rp_off,
    uint32_t *__restrict__ rp_len, uint32_t *__restrict__ task_list, uint32_t *__restrict__ n_tasks,
    uint32_t *__restrict__ big_list, uint32_t *__restrict__ n_big) {
    __shared__ uint64_t s_b[FK_INTER_WPB][FK_INTER_SMALL];
    __shared__ uint32_t s_a[FK_INTER_WPB][FK_INTER_SMALL];
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t des = blockIdx.x * FK_INTER_WPB + w;
    if (des >= N) return;
    const uint32_t nA = nsg_len[des];
    const uint32_t nB = indeg[des];
    if (nB == 0) {
        if (lane == 0) rp_len[des] = 0;
        return;
    }
    if (nA + nB > FK_INTER_SMALL) {
        if (lane == 0) big_list[atomicAdd(n_big, 1u)] = des;
        return;
    }
    uint32_t *sa = s_a[w];
    uint64_t *sb = s_b[w];
    uint32_t *ai = nsg_id + (uint64_t)des * ncap;
    float *ad = nsg_d + (uint64_t)des * ncap;
    for (uint32_t k = lane; k < nA; k += 64) sa[k] = ai[k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // B minus A (reference: dedupe against the current list, :2063-2076)
    const uint64_t *bk = in_key + in_off[des];
    uint32_t nb = 0;
    for (uint32_t k0 = 0; k0 < nB; k0 += 64) {
        const uint32_t k = k0 + lane;
        uint64_t key = FK_KEY_EMPTY;
        bool keep = false;
        if (k < nB) {
            key = bk[k];
            const uint32_t src = key_id(key);
            keep = true;
            for (uint32_t j = 0; j < nA; ++j)
                if (sa[j] == src) { keep = false; break; }
        }
        const uint64_t m = __ballot(keep);
        if (keep) sb[nb + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = key;
        nb += (uint32_t)__builtin_popcountll(m);
    }
    if (nb == 0) {
        if (lane == 0) rp_len[des] = 0;
        return;
    }
    const bool fits = nA + nb <= R;
    // over-full: the union A U B (A's keys after B's) is ranked instead
    const uint32_t tot = fits ? nb : nA + nb;
    if (!fits)
        for (uint32_t k = lane; k < nA; k += 64) sb[nb + k] = make_key(ad[k], sa[k]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const uint64_t off = in_off[des] + (uint64_t)des * R;
    for (uint32_t k = lane; k < tot; k += 64) {
        const uint64_t key = sb[k];
        uint32_t r = 0;
        for (uint32_t j = 0; j < tot; ++j) r += sb[j] < key ? 1u : 0u;
        if (fits) {
            ai[nA + r] = key_id(key);
            ad[nA + r] = key_dist(key);
        } else {
            rp_key[off + r] = key;
        }
    }
    if (lane == 0) {
        if (fits) {
            nsg_len[des] = nA + nb;
            rp_len[des] = 0;
        } else {
            rp_off[des] = off;
            rp_len[des] = tot;
        }
    }
}

// Reversed-prune tasks: the nodes whose union was staged (rp_len > 0),
// listed in the locality order (node order when there is none) by a flag
// scan, so the prune's XCD-contiguous runs cover neighbouring nodes -- an
// atomic append per node held every wave behind one counter when most
// out-lists overflow.
__global__ void k_task_flags(const uint32_t *__restrict__ rp_len, const uint32_t *__restrict__ order, uint32_t N,
                             uint32_t *__restrict__ flag) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    flag[i] = rp_len[order ? order[i] : i] > 0 ? 1u : 0u;
}
__global__ void k_task_scatter(const uint32_t *__restrict__ flag, const uint64_t *__restrict__ pos,
                               const uint32_t *__restrict__ order, uint32_t N, uint32_t *__restrict__ tasks) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N || !flag[i]) return;
    tasks[pos[i]] = order ? order[i] : i;
}

// UpdateNhoods: isnew = (first loop) or (id not in the previous out-list);
// then the previous out-list becomes the current one.
__global__ void k_update_nhoods(const uint32_t *__restrict__ nsg_id, const uint32_t *__restrict__ nsg_len,
                                uint8_t *__restrict__ nsg_new, uint32_t *__restrict__ old_id,
                                uint32_t *__restrict__ old_len, uint32_t N, uint32_t ncap, uint32_t first) {
    const uint32_t n = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (n >= N) return;
    const uint32_t deg = nsg_len[n];
    const uint32_t od = first ? 0u : old_len[n];
    const uint32_t *ids = nsg_id + (uint64_t)n * ncap;
    uint32_t *oi = old_id + (uint64_t)n * ncap;
    for (uint32_t k = lane_id(); k < deg; k += 64) {
        uint8_t isnew = 1;
        if (!first) {
            const uint32_t id = ids[k];
            for (uint32_t j = 0; j < od; ++j)
                if (oi[j] == id) { isnew = 0; break; }
        }
        nsg_new[(uint64_t)n * ncap + k] = isnew;
    }
    __builtin_amdgcn_wave_barrier();
    for (uint32_t k = lane_id(); k < deg; k += 64) oi[k] = ids[k];
    if (lane_id() == 0) old_len[n] = deg;
}

// CalculateNSG (src/kgraph.cpp:2273-2293): degree statistics; one atomic
// per block.
__global__ __launch_bounds__(1024) void k_degree_stats(const uint32_t *__restrict__ nsg_len, uint32_t N,
                                                       unsigned long long *__restrict__ acc) {
    __shared__ uint32_t s_mx, s_mn;
    if (threadIdx.x == 0) { s_mx = 0; s_mn = 0xFFFFFFFFu; }
    __syncthreads();
    unsigned long long sum = 0;
    uint32_t mx = 0, mn = 0xFFFFFFFFu;
    for (uint32_t n = blockIdx.x * blockDim.x + threadIdx.x; n < N; n += gridDim.x * blockDim.x) {
        const uint32_t d = nsg_len[n];
        sum += d;
        mx = d > mx ? d : mx;
        mn = d < mn ? d : mn;
    }
    atomicMax(&s_mx, mx);
    atomicMin(&s_mn, mn);
    sum = block_sum(sum);
    if (threadIdx.x == 0) {
        atomicAdd(acc, sum);
        atomicMax(acc + 1, (unsigned long long)s_mx);
        atomicMin(acc + 2, (unsigned long long)s_mn);
    }
}

__global__ void k_pack_csr(const uint32_t *__restrict__ nsg_id, const uint32_t *__restrict__ nsg_len, uint32_t N,
                           uint32_t ncap, const uint64_t *__restrict__ off, uint32_t *__restrict__ out) {
    const uint32_t n = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (n >= N) return;
    const uint32_t deg = nsg_len[n];
    for (uint32_t k = lane_id(); k < deg; k += 64) out[off[n] + k] = nsg_id[(uint64_t)n * ncap + k];
}

// The prune of one link step with the kernel variant for its shape: forced
// by a flag, or taken from the handle, or -- on the shape's first use --
// measured.  Measurement runs every variant twice on a sample (the first
// sixteenth of the tasks, at most 131,072 of them, a contiguous run of the
// locality order), takes each one's faster run (none is judged on caches
// another one warmed), requires the sample's out-lists (order-independent
// digest) and distance counts to agree, keeps the fastest variant on the
// handle and runs it on the remaining tasks: the launch's outputs and
// tallies are those of a single launch (the sample's from its last run,
// variant 0's).  Small launches are measured whole.  Variants that do not
// apply to the shape run as variant 0 and are not timed.  Returns
// 1 + variant.
static uint32_t tuned_prune(BuildState &b, PruneLaunch &a, bool reversed) {
    const uint32_t flags = b.P.flags;
    if (flags & (FK_FLAG_PRUNE_V0 | FK_FLAG_PRUNE_V1 | FK_FLAG_PRUNE_V2 | FK_FLAG_PRUNE_V3)) {
        a.variant = (flags & FK_FLAG_PRUNE_V1) ? 1 : (flags & FK_FLAG_PRUNE_V2) ? 2 : (flags & FK_FLAG_PRUNE_V3) ? 3 : 0;
        launch_prune(b.s, *b.ds, a);
        return 1u + (uint32_t)a.variant;
    }
    const uint32_t kind = (a.chains8 ? 2u : (a.nrm ? 1u : 0u)) | (a.gram ? 4u : 0u);
    const std::array<uint32_t, 6> key = {(uint32_t)a.rule, b.ds->dpad, kind, a.R, a.maxc, reversed ? 1u : 0u};
    auto it = b.h->prune_variant.find(key);
    if (it != b.h->prune_variant.end()) {
        a.variant = it->second;
        launch_prune(b.s, *b.ds, a);
        return 1u + (uint32_t)a.variant;
    }
    const double t0 = now_sec();
    const uint32_t n = a.n_tasks;
    const uint32_t m = n >= 65536 ? std::min<uint32_t>(n / 16, 131072u) : n;  // sample
    PruneLaunch smp = a;
    smp.n_tasks = m;
    hipEvent_t e0, e1;
    FK_HIP(hipEventCreate(&e0));
    FK_HIP(hipEventCreate(&e1));
    float best_ms[FK_PRUNE_VARIANTS];
    for (int v = 0; v < FK_PRUNE_VARIANTS; ++v) best_ms[v] = INFINITY;
    unsigned long long ref[2] = {0, 0};
    bool have_ref = false;
    DevBuf<unsigned long long> dg(1);
    try {
        // every variant twice (variants 3, 2, 1, 0, then again), each timed
        // by its faster run, so none is judged on caches another one warmed;
        // the last run is variant 0's
        for (int r = 2 * FK_PRUNE_VARIANTS - 1; r >= 0; --r) {
            const int v = r % FK_PRUNE_VARIANTS;
            if (!prune_variant_applies(*b.ds, smp, v)) continue;  // it would run as variant 0
            b.tally.clear(b.s, FK_T_PRUNE_CC, 2);
            b.tally.clear(b.s, FK_T_PRUNE_SKIP);
            smp.variant = v;
            FK_HIP(hipEventRecord(e0, b.s));
            launch_prune(b.s, *b.ds, smp);
            FK_HIP(hipEventRecord(e1, b.s));
            FK_HIP(hipMemsetAsync(dg.p, 0, 8, b.s));
            launch_prune_digest(b.s, smp, dg.p);
            unsigned long long got[2] = {0, 0};
            FK_HIP(hipMemcpyAsync(&got[0], dg.p, 8, hipMemcpyDeviceToHost, b.s));
            got[1] = b.tally.read1(b.s, FK_T_PRUNE_CC);  // synchronises the stream
            float ms = 0.f;
            FK_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (!have_ref) {
                ref[0] = got[0];
                ref[1] = got[1];
                have_ref = true;
            } else if (got[0] != ref[0] || got[1] != ref[1]) {
                throw Error(FK_E_INTERNAL, "prune variants disagree (variant " + std::to_string(v) + ")");
            }
            best_ms[v] = std::min(best_ms[v], ms);
        }
    } catch (...) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        throw;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    int best = 0;
    for (int v = 1; v < FK_PRUNE_VARIANTS; ++v)
        if (best_ms[v] < best_ms[best]) best = v;
    b.h->prune_variant[key] = best;
    a.variant = best;
    b.st->t_tune += now_sec() - t0;
    if (m < n) {  // the rest of the tasks; the tallies keep the sample's counts (the same for every variant)
        PruneLaunch rest = a;
        rest.task_base = a.task_base + m;
        rest.n_tasks = n - m;
        launch_prune(b.s, *b.ds, rest);
    }
    if (b.verbose)
        fprintf(stderr, "[fk] prune variant %d chosen for rule %d, dpad %u, rows %u, R %u%s (sample of %u: %.3f / %.3f / %.3f / %.3f ms)\n",
                best, a.rule, b.ds->dpad, kind, a.R, reversed ? ", reversed" : "", m, best_ms[0], best_ms[1], best_ms[2],
                best_ms[3]);
    return 1u + (uint32_t)best;
}

static void link(BuildState &b, fk_pg_type type, uint32_t li) {
    const double t0 = now_sec();
    const uint32_t N = b.N;
    const BuildParams &P = b.P;
    const unsigned gw = (N + 3) / 4;
    // 1. prune every pool
    b.tally.clear(b.s, FK_T_PRUNE_CC, 2);
    b.tally.clear(b.s, FK_T_PRUNE_SKIP);
    PruneLaunch a;
    a.nrm = exact_norms(b);
    a.chains8 = a.nrm ? b.ds->chains8.p : nullptr;
    a.cand = b.pool.p;
    a.cand_len = b.cnt.p;
    a.cand_stride = b.cap;
    a.n_tasks = N;
    a.oracle_n = N;
    a.maxc = P.search_K;
    a.R = P.nsg_R;
    a.tau3 = 3.0f * P.tau;
    a.cos_thr = alpha_cos_threshold(P.angle);
    a.rule = rule_of(type);
    a.out_ids = b.nsg_id.p;
    a.out_d = b.nsg_d.p;
    a.out_stride = b.ncap;
    a.out_len = b.nsg_len.p;
    a.cc = b.tally.at(FK_T_PRUNE_CC);
    a.examined = b.tally.at(FK_T_PRUNE_EXAM);
    a.uncomputed = b.tally.at(FK_T_PRUNE_SKIP);
    a.gram = (P.flags & FK_FLAG_GRAM) != 0;
    if (b.has_order) {
        a.task_list = b.order.p;
        a.xcd_order = true;
    }
    hipEvent_t pe = b.prof_begin();
    b.st->prune_variant = tuned_prune(b, a, false);
    unsigned long long cc2[2] = {0, 0};
    b.tally.read(b.s, FK_T_PRUNE_CC, 2, cc2);
    const unsigned long long cc = cc2[0];
    // every examined candidate's row and key read once, kept ids and
    // distances written
    b.prof_end(pe, FK_PROF_PRUNE, (double)cc2[1] * (4.0 * b.ds->d + 8.0) + 8.0 * (double)N * P.nsg_R);
    b.prof_flush();
    b.width = P.nsg_R;
    b.set_width.assign(b.n_sets, P.nsg_R);
    b.st->d_prune += cc;
    b.st->d_prune_uncomputed += b.tally.read1(b.s, FK_T_PRUNE_SKIP);
    const double t1 = now_sec();
    if (li < FK_MAX_LOOPS) {
        b.st->prune_cc[li] = (float)((double)cc / ((double)N * (double)(N - 1) / 2.0));
        b.st->prune_time[li] = t1 - t0;
    }
    b.st->t_prune += t1 - t0;
    // 2. reverse edges
    FK_HIP(hipMemsetAsync(b.bcnt.p, 0, (size_t)N * 4, b.s));
    pe = b.prof_begin();
    hipLaunchKernelGGL(k_indeg, dim3(gw), dim3(256), 0, b.s, b.nsg_id.p, b.nsg_len.p, N, b.ncap, b.bcnt.p);
    FK_LAUNCH_CHECK();
    exclusive_scan_u32_to_u64(b.s, b.bcnt.p, b.boff.p, N, b.boff.p + N, b.scan_tmp);
    uint64_t n_in = 0;
    FK_HIP(hipMemcpyAsync(&n_in, b.boff.p + N, 8, hipMemcpyDeviceToHost, b.s));
    b.sync();
    b.bkey.ensure(n_in);
    FK_HIP(hipMemsetAsync(b.pair_cnt.p, 0, (size_t)N * 4, b.s));
    hipLaunchKernelGGL(k_in_fill, dim3(gw), dim3(256), 0, b.s, b.nsg_id.p, b.nsg_d.p, b.nsg_len.p, N, b.ncap,
                       b.boff.p, b.pair_cnt.p, b.bkey.p);
    FK_LAUNCH_CHECK();
    DevBuf<uint64_t> rp_key(n_in + (uint64_t)N * P.nsg_R), rp_off(N);
    DevBuf<uint32_t> rp_len(N), tasks(N), big(N), counters(3);
    FK_HIP(hipMemsetAsync(counters.p, 0, 12, b.s));
    hipLaunchKernelGGL(k_inter_small, dim3((N + FK_INTER_WPB - 1) / FK_INTER_WPB), dim3(64 * FK_INTER_WPB), 0, b.s, N,
                       b.ncap, P.nsg_R, b.nsg_id.p, b.nsg_d.p, b.nsg_len.p, b.bcnt.p, b.boff.p, b.bkey.p, rp_key.p,
                       rp_off.p, rp_len.p, tasks.p, counters.p, big.p, counters.p + 2);
    FK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_inter_merge, dim3(256), dim3(256), 0, b.s, big.p, counters.p + 2, b.ncap, P.nsg_R,
                       b.nsg_id.p, b.nsg_d.p, b.nsg_len.p, b.bcnt.p, b.boff.p, b.bkey.p, rp_key.p, rp_off.p, rp_len.p,
                       tasks.p, counters.p, counters.p + 1);
    FK_LAUNCH_CHECK();
    DevBuf<uint64_t> tpos((size_t)N + 1);
    const uint32_t *ord = b.has_order ? b.order.p : nullptr;
    hipLaunchKernelGGL(k_task_flags, dim3((N + 255) / 256), dim3(256), 0, b.s, rp_len.p, ord, N, big.p);
    FK_LAUNCH_CHECK();
    exclusive_scan_u32_to_u64(b.s, big.p, tpos.p, N, tpos.p + N, b.scan_tmp);
    hipLaunchKernelGGL(k_task_scatter, dim3((N + 255) / 256), dim3(256), 0, b.s, big.p, tpos.p, ord, N, tasks.p);
    FK_LAUNCH_CHECK();
    uint32_t hc[2] = {0, 0};
    uint64_t n_tasks = 0;
    FK_HIP(hipMemcpyAsync(hc, counters.p, 8, hipMemcpyDeviceToHost, b.s));
    FK_HIP(hipMemcpyAsync(&n_tasks, tpos.p + N, 8, hipMemcpyDeviceToHost, b.s));
    b.sync();
    hc[0] = (uint32_t)n_tasks;
    // out-lists read twice (count, fill) and rewritten, in-edges written and
    // read once as 8-byte keys
    b.prof_end(pe, FK_PROF_INTER, 3.0 * 8.0 * (double)n_in + 16.0 * (double)n_in);
    b.prof_flush();
    if (hc[1]) throw Error(FK_E_INTERNAL, "in-degree above the reverse-edge merge capacity (" +
                                              std::to_string(hc[1]) + " > 4096)");
    // 3. reversed prune of over-full lists (no search_K cap, no size guard)
    if (hc[0]) {
        b.tally.clear(b.s, FK_T_PRUNE_CC, 2);
        b.tally.clear(b.s, FK_T_PRUNE_SKIP);
        PruneLaunch r = a;
        r.cand = rp_key.p;
        r.cand_off = rp_off.p;
        r.cand_len = rp_len.p;
        r.task_list = tasks.p;
        r.n_tasks = hc[0];
        r.oracle_n = 0xFFFFFFFFu;
        r.maxc = 0xFFFFFFFFu;
        r.examined = b.tally.at(FK_T_PRUNE_EXAM);
        hipEvent_t pr = b.prof_begin();
        b.st->prune_rev_variant = tuned_prune(b, r, true);
        unsigned long long rcc[2] = {0, 0};
        b.tally.read(b.s, FK_T_PRUNE_CC, 2, rcc);
        const unsigned long long rc = rcc[0];
        b.prof_end(pr, FK_PROF_PRUNE_REV, (double)rcc[1] * (4.0 * b.ds->d + 8.0) + 8.0 * (double)hc[0] * P.nsg_R);
        b.prof_flush();
        b.st->d_prune_rev += rc;
        b.st->d_prune_uncomputed += b.tally.read1(b.s, FK_T_PRUNE_SKIP);
    }
    const double t2 = now_sec();
    if (li < FK_MAX_LOOPS) b.st->inter_time[li] = t2 - t1;
    b.st->t_inter += t2 - t1;
    if (b.verbose)
        fprintf(stderr, "[fk] link: prune dists %llu, in-edges %llu, re-pruned %u (%.3f s)\n", cc,
                (unsigned long long)n_in, hc[0], t2 - t0);
}

static void update_nhoods(BuildState &b) {
    const unsigned gw = (b.N + 3) / 4;
    hipLaunchKernelGGL(k_update_nhoods, dim3(gw), dim3(256), 0, b.s, b.nsg_id.p, b.nsg_len.p, b.nsg_new.p, b.old_id.p,
                       b.old_len.p, b.N, b.ncap, b.loopcount == 0 ? 1u : 0u);
    FK_LAUNCH_CHECK();
}

static void calculate_nsg(BuildState &b) {
    unsigned long long init[3] = {0, 0, ~0ull};
    FK_HIP(hipMemcpyAsync(b.acc.p, init, 24, hipMemcpyHostToDevice, b.s));
    const unsigned g = std::min<unsigned>(1024u, (b.N + 1023) / 1024);
    hipLaunchKernelGGL(k_degree_stats, dim3(g), dim3(1024), 0, b.s, b.nsg_len.p, b.N, b.acc.p);
    FK_LAUNCH_CHECK();
    unsigned long long r[3];
    FK_HIP(hipMemcpyAsync(r, b.acc.p, 24, hipMemcpyDeviceToHost, b.s));
    b.sync();
    // avg is an unsigned in the reference: the mean truncated to an integer
    b.st->aod = (uint32_t)(r[0] / b.N);
    b.st->max_degree = (uint32_t)r[1];
    b.st->min_degree = (uint32_t)r[2];
    if (b.verbose)
        fprintf(stderr, "[fk] Degree Statistics: Max = %u, Min = %u, Avg = %u\n", b.st->max_degree,
                b.st->min_degree, b.st->aod);
}

void build_loops(BuildState &b, fk_pg_type type) {
    const uint32_t N = b.N;
    b.ncap = b.P.nsg_R + FK_NSG_SLACK;
    b.nsg_id.alloc((size_t)N * b.ncap);
    b.nsg_d.alloc((size_t)N * b.ncap);
    b.nsg_new.alloc((size_t)N * b.ncap);
    b.old_id.alloc((size_t)N * b.ncap);
    b.nsg_len.alloc(N);
    b.old_len.alloc(N);
    FK_HIP(hipMemsetAsync(b.old_len.p, 0, (size_t)N * 4, b.s));
    b.loopcount = 0;
    locality_order(b);
    const uint32_t iters = b.P.loop_i + 1;
    const double s0 = b.st->t_total;  // build start offset is tracked by the caller
    (void)s0;
    for (uint32_t iter = 1; iter <= iters; ++iter) {
        const uint32_t li = iter - 1;
        const double tb = now_sec();
        link(b, type, li);
        compute_ep(b);
        if (type != FK_PG_HNSW_LAYER) tree_grow(b);
        update_nhoods(b);
        b.sync();
        if (li < FK_MAX_LOOPS) {
            b.st->buildnsg_time[li] = now_sec() - tb;
            b.st->con_time[li] = now_sec() - b.t0;
        }
        b.loopcount++;
        b.st->n_loops = iter;
        if (iter == iters) {
            calculate_nsg(b);
            break;
        }
        const double ts = now_sec();
        bridgeview(b);
        locality_order(b);
        if (li < FK_MAX_LOOPS) b.st->search_time[li] = now_sec() - ts;
        const float rec = nnd_evaluate(b, b.P.MAXL);
        if (b.st->n_recall < FK_MAX_LOOPS + 64) b.st->recall[b.st->n_recall++] = rec;
        if (b.st->n_kgraph < FK_MAX_LOOPS + 64) b.st->kgraph_time[b.st->n_kgraph++] = now_sec() - b.t0;
        if (b.verbose) fprintf(stderr, "[fk] recall after BridgeView: %g\n", rec);
    }
}

}  // namespace fk

using namespace fk;

static void check_params(const fk_params *p) {
    FK_CHECK_ARG(p != nullptr, "null params");
    FK_CHECK_ARG(p->K >= 1 && p->L >= 1 && p->S >= 1, "K, L and S must be positive");
    FK_CHECK_ARG(p->nsg_R >= 1 && p->nsg_R <= 256, "nsg_R must be in [1, 256]");
    FK_CHECK_ARG(p->R >= 1, "R = 0 (uncapped reverse lists) is not supported");
    FK_CHECK_ARG((uint64_t)p->L + 2ull * p->R <= FK_LIST_MAX, "L + 2R must not exceed 2048");
    FK_CHECK_ARG(p->massq_S >= 1, "massq_S must be positive");
    FK_CHECK_ARG(!((p->flags & FK_FLAG_GRAM_BF16) && (p->flags & FK_FLAG_GRAM_F16)), "conflicting Gram formats");
    FK_CHECK_ARG(!((p->flags & FK_FLAG_GRAM) && (p->flags & FK_FLAG_SCREEN)),
                 "FK_FLAG_SCREEN and FK_FLAG_GRAM cannot be combined");
}

// The graph as CSR in host memory: offsets by a device scan of the
// out-degrees, ids packed on the device, both downloaded in pinned chunks.
static void graph_out(BuildState &b, fk_graph *out) {
    const uint32_t N = b.N;
    DevBuf<uint64_t> doff((size_t)N + 1);
    exclusive_scan_u32_to_u64(b.s, b.nsg_len.p, doff.p, N, doff.p + N, b.scan_tmp);
    uint64_t ne = 0;
    FK_HIP(hipMemcpyAsync(&ne, doff.p + N, 8, hipMemcpyDeviceToHost, b.s));
    b.sync();
    DevBuf<uint32_t> dids(std::max<uint64_t>(ne, 1));
    hipLaunchKernelGGL(k_pack_csr, dim3((N + 3) / 4), dim3(256), 0, b.s, b.nsg_id.p, b.nsg_len.p, N, b.ncap, doff.p,
                       dids.p);
    FK_LAUNCH_CHECK();
    struct Host {
        void *p;
        ~Host() { free(p); }
    } off{host_result_alloc(((size_t)N + 1) * sizeof(uint64_t))}, ids{host_result_alloc(ne * sizeof(uint32_t))};
    if (!off.p || !ids.p) throw std::bad_alloc();
    download_bytes(b.h, off.p, doff.p, ((size_t)N + 1) * sizeof(uint64_t));
    download_bytes(b.h, ids.p, dids.p, ne * sizeof(uint32_t));
    out->n = N;
    out->ep = b.ep;
    out->width = b.width;
    out->offsets = static_cast<uint64_t *>(off.p);
    out->ids = static_cast<uint32_t *>(ids.p);
    off.p = ids.p = nullptr;
}

extern "C" {

fk_status fk_build(fk_handle *h, fk_pg_type type, uint32_t n_used, const fk_params *p, fk_graph *out) {
    if (!h || !out) return FK_E_INVALID;
    memset(out, 0, sizeof(*out));
    return guard(h, [&] {
        require_dataset(h);
        check_params(p);
        FK_CHECK_ARG(type == FK_PG_NSG || type == FK_PG_HNSW_LAYER || type == FK_PG_TAUMNG || type == FK_PG_ALPHAPG,
                     "unknown graph type");
        FK_CHECK_ARG(n_used >= 2 && n_used <= h->ds.n_total, "n_used must be in [2, n_total]");
        select_device(h);
        const double t0 = now_sec();
        BuildState b;
        b.h = h;
        b.s = h->stream;
        b.ds = &h->ds;
        b.N = n_used;
        b.P = shrink_params(*p, n_used);
        b.cap = b.P.MAXL;
        b.st = &out->stats;
        b.verbose = (p->flags & FK_FLAG_VERBOSE) != 0;
        b.init_sets({0u, n_used});
        b.prof = (p->flags & FK_FLAG_PROFILE) != 0;
        b.t0 = t0;
        nnd_controls(b);
        nnd_run(b);
        // buildFast*: evaluate() at MAXL, then the constructor's time
        const float rec = nnd_evaluate(b, b.P.MAXL);
        if (b.st->n_recall < FK_MAX_LOOPS + 64) b.st->recall[b.st->n_recall++] = rec;
        if (b.st->n_kgraph < FK_MAX_LOOPS + 64) b.st->kgraph_time[b.st->n_kgraph++] = now_sec() - t0;
        build_loops(b, type);
        graph_out(b, out);
        out->stats.t_total = now_sec() - t0;
    });
}

fk_status fk_build_batch(fk_handle *h, fk_pg_type type, uint32_t n_sets, const uint64_t *set_off,
                         const fk_params *p, fk_graph *out) {
    if (!h || !out || !set_off || n_sets == 0) return FK_E_INVALID;
    for (uint32_t k = 0; k < n_sets; ++k) memset(&out[k], 0, sizeof(out[k]));
    return guard(h, [&] {
        require_dataset(h);
        check_params(p);
        FK_CHECK_ARG(type == FK_PG_NSG || type == FK_PG_HNSW_LAYER || type == FK_PG_TAUMNG || type == FK_PG_ALPHAPG,
                     "unknown graph type");
        FK_CHECK_ARG(!(p->flags & FK_FLAG_SCREEN), "the learned screen is not supported in batch builds");
        FK_CHECK_ARG(set_off[0] == 0 && set_off[n_sets] <= h->ds.n_total, "sets must lie in the resident rows");
        // every set keeps the unshrunk parameters (src/kgraph.cpp:1288-1334
        // shrinks them only for sets no larger than K, L, S, controls,
        // search_L or search_K)
        const uint64_t need = std::max({(uint64_t)p->K, (uint64_t)p->L, (uint64_t)p->S, (uint64_t)p->controls,
                                        (uint64_t)p->search_L, (uint64_t)p->search_K}) + 1;
        std::vector<uint32_t> lo(n_sets + 1);
        for (uint32_t k = 0; k < n_sets; ++k) {
            FK_CHECK_ARG(set_off[k + 1] >= set_off[k] + need,
                         "every set needs more rows than max(K, L, S, controls, search_L, search_K)");
            lo[k] = (uint32_t)set_off[k];
        }
        lo[n_sets] = (uint32_t)set_off[n_sets];
        select_device(h);
        const double t0 = now_sec();
        fk_stats st;
        memset(&st, 0, sizeof(st));
        BuildState b;
        b.h = h;
        b.s = h->stream;
        b.ds = &h->ds;
        b.N = lo[n_sets];
        b.P = shrink_params(*p, b.N);
        b.cap = b.P.MAXL;
        b.st = &st;
        b.verbose = (p->flags & FK_FLAG_VERBOSE) != 0;
        b.prof = (p->flags & FK_FLAG_PROFILE) != 0;
        b.t0 = t0;
        b.init_sets(lo);
        nnd_controls(b);
        nnd_run(b);
        const float rec = nnd_evaluate(b, b.P.MAXL);
        if (st.n_recall < FK_MAX_LOOPS + 64) st.recall[st.n_recall++] = rec;
        if (st.n_kgraph < FK_MAX_LOOPS + 64) st.kgraph_time[st.n_kgraph++] = now_sec() - t0;
        build_loops(b, type);
        // one CSR for all sets, then per set with ids relative to the set
        std::vector<uint32_t> len(b.N);
        FK_HIP(hipMemcpyAsync(len.data(), b.nsg_len.p, (size_t)b.N * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
        std::vector<uint64_t> off((size_t)b.N + 1, 0);
        for (uint32_t i = 0; i < b.N; ++i) off[i + 1] = off[i] + len[i];
        DevBuf<uint64_t> doff(b.N + 1);
        DevBuf<uint32_t> dids(off[b.N]);
        FK_HIP(hipMemcpyAsync(doff.p, off.data(), off.size() * 8, hipMemcpyHostToDevice, b.s));
        hipLaunchKernelGGL(k_pack_csr, dim3((b.N + 3) / 4), dim3(256), 0, b.s, b.nsg_id.p, b.nsg_len.p, b.N, b.ncap,
                           doff.p, dids.p);
        FK_LAUNCH_CHECK();
        std::vector<uint32_t> ids(std::max<uint64_t>(off[b.N], 1));
        FK_HIP(hipMemcpyAsync(ids.data(), dids.p, off[b.N] * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
        st.t_total = now_sec() - t0;
        std::vector<fk_graph> res(n_sets);
        struct Free {
            std::vector<fk_graph> &r;
            bool keep = false;
            ~Free() {
                if (!keep)
                    for (auto &g : r) { free(g.offsets); free(g.ids); }
            }
        } guard_res{res};
        for (uint32_t k = 0; k < n_sets; ++k) {
            fk_graph &g = res[k];
            memset(&g, 0, sizeof(g));
            const uint32_t a = lo[k], n = lo[k + 1] - a;
            const uint64_t e0 = off[a], e1 = off[lo[k + 1]];
            g.n = n;
            g.ep = b.set_ep[k] - a;
            g.width = b.set_width[k];
            g.offsets = (uint64_t *)malloc(((size_t)n + 1) * sizeof(uint64_t));
            g.ids = (uint32_t *)malloc(std::max<uint64_t>(e1 - e0, 1) * 4);
            if (!g.offsets || !g.ids) throw std::bad_alloc();
            for (uint32_t i = 0; i <= n; ++i) g.offsets[i] = off[a + i] - e0;
            for (uint64_t e = e0; e < e1; ++e) g.ids[e - e0] = ids[e] - a;
            g.stats = st;
        }
        guard_res.keep = true;
        for (uint32_t k = 0; k < n_sets; ++k) out[k] = res[k];
    });
}

}  // extern "C"

namespace fk {

// A one-set build state of s->n nodes for a step entry point: parameters
// shrunk as a build of n nodes shrinks them, counters and scratch allocated.
static void step_begin(BuildState &b, fk_handle *h, const fk_params *p, const fk_state *s, fk_stats *st) {
    FK_CHECK_ARG(s != nullptr, "null state");
    FK_CHECK_ARG(s->n >= 2 && s->n <= h->ds.n_total, "state size must be in [2, resident rows]");
    b.h = h;
    b.s = h->stream;
    b.ds = &h->ds;
    b.N = s->n;
    b.P = shrink_params(*p, s->n);
    b.cap = b.P.MAXL;
    b.st = st;
    b.t0 = now_sec();
    b.init_sets({0u, s->n});
    b.tally.ensure(FK_T_SLOTS);
    b.acc.alloc(8);
    b.scal.alloc(16);
    b.loopcount = s->loopcount;
    if (b.P.flags & FK_FLAG_GRAM) gram_prepare(b.s, *b.ds, b.P.flags);
}

// Host rows of `cols` entries at pitch `hpitch` (elements) <-> device rows
// at pitch `dpitch`.
template <typename T>
static void put_rows(BuildState &b, T *dev, size_t dpitch, const T *host, size_t hpitch, size_t cols) {
    FK_HIP(hipMemcpy2DAsync(dev, dpitch * sizeof(T), host, hpitch * sizeof(T), cols * sizeof(T), b.N,
                            hipMemcpyHostToDevice, b.s));
}
template <typename T>
static void get_rows(BuildState &b, T *host, size_t hpitch, const T *dev, size_t dpitch, size_t cols) {
    FK_HIP(hipMemcpy2DAsync(host, hpitch * sizeof(T), dev, dpitch * sizeof(T), cols * sizeof(T), b.N,
                            hipMemcpyDeviceToHost, b.s));
}

static void check_lens(const uint32_t *len, uint32_t n, uint32_t cap, const char *what) {
    FK_CHECK_ARG(len != nullptr, "null length array");
    for (uint32_t i = 0; i < n; ++i) FK_CHECK_ARG(len[i] <= cap, what);
}

// Pool ids must be resident rows below n (the pruners and BridgeView read
// their rows; the reference's pools hold only node ids).
static void check_pool_ids(const fk_state *s) {
    for (uint32_t i = 0; i < s->n; ++i)
        for (uint32_t k = 0; k < s->pool_len[i]; ++k)
            FK_CHECK_ARG((uint32_t)s->pool[(size_t)i * s->pool_cap + k] < s->n, "pool id out of range");
}

static void check_out_ids(const fk_state *s, const uint32_t *ids, const uint32_t *len) {
    for (uint32_t i = 0; i < s->n; ++i)
        for (uint32_t k = 0; k < len[i]; ++k)
            FK_CHECK_ARG(ids[(size_t)i * s->out_cap + k] < s->n, "out-list id out of range");
}

}  // namespace fk

extern "C" {

fk_status fk_step_link(fk_handle *h, fk_pg_type type, const fk_params *p, fk_state *s, fk_stats *st) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        check_params(p);
        FK_CHECK_ARG(type == FK_PG_NSG || type == FK_PG_HNSW_LAYER || type == FK_PG_TAUMNG || type == FK_PG_ALPHAPG,
                     "unknown graph type");
        FK_CHECK_ARG(s && s->pool && s->pool_len && s->out_id && s->out_dist && s->out_len, "null state field");
        FK_CHECK_ARG(s->out_cap >= p->nsg_R, "out_cap < nsg_R");
        FK_CHECK_ARG(s->pool_cap >= 1 && s->pool_cap <= FK_POOL_MAX, "pool_cap must be in [1, 1024]");
        check_lens(s->pool_len, s->n, s->pool_cap, "pool_len > pool_cap");
        check_pool_ids(s);
        select_device(h);
        fk_stats local;
        memset(&local, 0, sizeof(local));
        BuildState b;
        step_begin(b, h, p, s, st ? st : &local);
        if (st) memset(st, 0, sizeof(*st));
        const uint32_t N = b.N;
        b.cap = s->pool_cap;  // the prune reads each pool's first pool_len entries
        b.pool.alloc((size_t)N * b.cap);
        b.cnt.alloc(N);
        b.bcnt.alloc(N);
        b.boff.alloc((size_t)N + 1);
        b.pair_cnt.alloc(N);
        b.ncap = b.P.nsg_R + FK_NSG_SLACK;
        b.nsg_id.alloc((size_t)N * b.ncap);
        b.nsg_d.alloc((size_t)N * b.ncap);
        b.nsg_new.alloc((size_t)N * b.ncap);
        b.nsg_len.alloc(N);
        put_rows(b, b.pool.p, b.cap, s->pool, s->pool_cap, s->pool_cap);
        FK_HIP(hipMemcpyAsync(b.cnt.p, s->pool_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        link(b, type, 0);
        // every list fits nsg_R after the link (kept + reverse edges, or the reversed prune)
        get_rows(b, s->out_id, s->out_cap, b.nsg_id.p, b.ncap, b.P.nsg_R);
        get_rows(b, s->out_dist, s->out_cap, b.nsg_d.p, b.ncap, b.P.nsg_R);
        FK_HIP(hipMemcpyAsync(s->out_len, b.nsg_len.p, (size_t)N * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
        for (uint32_t i = 0; i < N; ++i)  // slots past each list hold zeros, not device bytes
            for (uint32_t k = s->out_len[i]; k < s->out_cap; ++k) {
                s->out_id[(size_t)i * s->out_cap + k] = 0;
                s->out_dist[(size_t)i * s->out_cap + k] = 0.f;
            }
    });
}

fk_status fk_step_update_nhoods(fk_handle *h, fk_state *s) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        FK_CHECK_ARG(s && s->out_id && s->out_isnew && s->out_len && s->prev_id && s->prev_len, "null state field");
        FK_CHECK_ARG(s->out_cap >= 1, "out_cap must be positive");
        check_lens(s->out_len, s->n, s->out_cap, "out_len > out_cap");
        check_lens(s->prev_len, s->n, s->out_cap, "prev_len > out_cap");
        select_device(h);
        fk_params p;
        memset(&p, 0, sizeof(p));
        p.K = p.L = p.S = p.nsg_R = p.search_L = p.search_K = 1;
        fk_stats local;
        memset(&local, 0, sizeof(local));
        BuildState b;
        step_begin(b, h, &p, s, &local);
        const uint32_t N = b.N;
        b.ncap = s->out_cap;
        b.nsg_id.alloc((size_t)N * b.ncap);
        b.nsg_new.alloc((size_t)N * b.ncap);
        b.old_id.alloc((size_t)N * b.ncap);
        b.nsg_len.alloc(N);
        b.old_len.alloc(N);
        put_rows(b, b.nsg_id.p, b.ncap, s->out_id, s->out_cap, s->out_cap);
        put_rows(b, b.nsg_new.p, b.ncap, s->out_isnew, s->out_cap, s->out_cap);  // slots past out_len kept
        put_rows(b, b.old_id.p, b.ncap, s->prev_id, s->out_cap, s->out_cap);
        FK_HIP(hipMemcpyAsync(b.nsg_len.p, s->out_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        FK_HIP(hipMemcpyAsync(b.old_len.p, s->prev_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        update_nhoods(b);
        get_rows(b, s->out_isnew, s->out_cap, b.nsg_new.p, b.ncap, s->out_cap);
        get_rows(b, s->prev_id, s->out_cap, b.old_id.p, b.ncap, s->out_cap);
        FK_HIP(hipMemcpyAsync(s->prev_len, b.old_len.p, (size_t)N * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
    });
}

fk_status fk_step_bridgeview(fk_handle *h, const fk_params *p, fk_state *s, fk_stats *st) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        check_params(p);
        FK_CHECK_ARG(s && s->pool && s->pool_len && s->pool_flags && s->out_id && s->out_isnew && s->out_len,
                     "null state field");
        FK_CHECK_ARG(s->loopcount >= 1, "BridgeView runs after the first loop (loopcount >= 1)");
        const BuildParams P = shrink_params(*p, s->n);
        FK_CHECK_ARG(s->pool_cap == P.MAXL, "pool_cap must equal max(L, K, search_L, search_K)");
        FK_CHECK_ARG(P.MAXL <= FK_POOL_MAX, "max(L, K, search_L, search_K) above 1024 is not supported");
        FK_CHECK_ARG(s->out_cap >= 1, "out_cap must be positive");
        check_lens(s->pool_len, s->n, s->pool_cap, "pool_len > pool_cap");
        check_lens(s->out_len, s->n, s->out_cap, "out_len > out_cap");
        check_pool_ids(s);
        check_out_ids(s, s->out_id, s->out_len);
        select_device(h);
        fk_stats local;
        memset(&local, 0, sizeof(local));
        BuildState b;
        if (st) memset(st, 0, sizeof(*st));
        step_begin(b, h, p, s, st ? st : &local);
        const uint32_t N = b.N;
        b.pool.alloc((size_t)N * b.cap);
        b.pflg.alloc((size_t)N * b.cap);
        b.cnt.alloc(N);
        b.ncap = s->out_cap;
        b.nsg_id.alloc((size_t)N * b.ncap);
        b.nsg_new.alloc((size_t)N * b.ncap);
        b.nsg_len.alloc(N);
        put_rows(b, b.pool.p, b.cap, s->pool, s->pool_cap, s->pool_cap);
        put_rows(b, b.pflg.p, b.cap, s->pool_flags, s->pool_cap, s->pool_cap);
        FK_HIP(hipMemcpyAsync(b.cnt.p, s->pool_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        put_rows(b, b.nsg_id.p, b.ncap, s->out_id, s->out_cap, s->out_cap);
        put_rows(b, b.nsg_new.p, b.ncap, s->out_isnew, s->out_cap, s->out_cap);
        FK_HIP(hipMemcpyAsync(b.nsg_len.p, s->out_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        bridgeview(b);
        // the pass records its rounds and search_cc at index loopcount - 1;
        // the step reports them at index 0
        const uint32_t li = b.loopcount - 1;
        if (li > 0 && li < FK_MAX_LOOPS) {
            b.st->bv_rounds[0] = b.st->bv_rounds[li];
            b.st->search_cc[0] = b.st->search_cc[li];
            b.st->bv_rounds[li] = 0;
            b.st->search_cc[li] = 0.f;
        }
        get_rows(b, s->pool, s->pool_cap, b.pool.p, b.cap, s->pool_cap);
        get_rows(b, s->pool_flags, s->pool_cap, b.pflg.p, b.cap, s->pool_cap);
        FK_HIP(hipMemcpyAsync(s->pool_len, b.cnt.p, (size_t)N * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
    });
}

fk_status fk_step_connectivity(fk_handle *h, const fk_params *p, fk_state *s, uint32_t ep, fk_stats *st) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        check_params(p);
        FK_CHECK_ARG(s && s->pool && s->pool_len && s->out_id && s->out_dist && s->out_len, "null state field");
        FK_CHECK_ARG(s->pool_cap >= 1, "pool_cap must be positive");
        FK_CHECK_ARG(s->out_cap > p->nsg_R, "out_cap must exceed nsg_R (room for attached edges)");
        FK_CHECK_ARG(ep < s->n, "entry point out of range");
        check_lens(s->pool_len, s->n, s->pool_cap, "pool_len > pool_cap");
        check_lens(s->out_len, s->n, s->out_cap, "out_len > out_cap");
        check_pool_ids(s);
        check_out_ids(s, s->out_id, s->out_len);
        select_device(h);
        fk_stats local;
        memset(&local, 0, sizeof(local));
        BuildState b;
        if (st) memset(st, 0, sizeof(*st));
        step_begin(b, h, p, s, st ? st : &local);
        const uint32_t N = b.N;
        b.cap = s->pool_cap;
        b.pool.alloc((size_t)N * b.cap);
        b.cnt.alloc(N);
        b.ncap = s->out_cap;
        b.nsg_id.alloc((size_t)N * b.ncap);
        b.nsg_d.alloc((size_t)N * b.ncap);
        b.nsg_len.alloc(N);
        put_rows(b, b.pool.p, b.cap, s->pool, s->pool_cap, s->pool_cap);
        FK_HIP(hipMemcpyAsync(b.cnt.p, s->pool_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        put_rows(b, b.nsg_id.p, b.ncap, s->out_id, s->out_cap, s->out_cap);
        put_rows(b, b.nsg_d.p, b.ncap, s->out_dist, s->out_cap, s->out_cap);
        FK_HIP(hipMemcpyAsync(b.nsg_len.p, s->out_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        b.ep = ep;
        b.set_ep[0] = ep;
        tree_grow(b);
        get_rows(b, s->out_id, s->out_cap, b.nsg_id.p, b.ncap, s->out_cap);
        get_rows(b, s->out_dist, s->out_cap, b.nsg_d.p, b.ncap, s->out_cap);
        FK_HIP(hipMemcpyAsync(s->out_len, b.nsg_len.p, (size_t)N * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
        for (uint32_t i = 0; i < N; ++i)  // slots past each list hold zeros, not device bytes
            for (uint32_t k = s->out_len[i]; k < s->out_cap; ++k) {
                s->out_id[(size_t)i * s->out_cap + k] = 0;
                s->out_dist[(size_t)i * s->out_cap + k] = 0.f;
            }
    });
}

fk_status fk_step_stop_stats(fk_handle *h, const fk_params *p, const fk_state *s, uint32_t n_ctrl,
                             const uint32_t *ctrl_ids, const uint32_t *gt, uint32_t gt_stride, float *delta,
                             float *recall) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        check_params(p);
        FK_CHECK_ARG(s && s->pool && s->pool_len && s->pool_flags && delta && recall, "null argument");
        FK_CHECK_ARG(n_ctrl == 0 || (ctrl_ids && gt), "null controls");
        FK_CHECK_ARG(s->pool_cap >= 1, "pool_cap must be positive");
        check_lens(s->pool_len, s->n, s->pool_cap, "pool_len > pool_cap");
        const BuildParams P = shrink_params(*p, s->n);
        FK_CHECK_ARG(n_ctrl == 0 || gt_stride >= P.K, "gt_stride < K");
        for (uint32_t c = 0; c < n_ctrl; ++c) FK_CHECK_ARG(ctrl_ids[c] < s->n, "control id out of range");
        select_device(h);
        fk_stats local;
        memset(&local, 0, sizeof(local));
        BuildState b;
        step_begin(b, h, p, s, &local);
        const uint32_t N = b.N;
        b.cap = s->pool_cap;
        b.pool.alloc((size_t)N * b.cap);
        b.pflg.alloc((size_t)N * b.cap);
        b.cnt.alloc(N);
        put_rows(b, b.pool.p, b.cap, s->pool, s->pool_cap, s->pool_cap);
        put_rows(b, b.pflg.p, b.cap, s->pool_flags, s->pool_cap, s->pool_cap);
        FK_HIP(hipMemcpyAsync(b.cnt.p, s->pool_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        DevBuf<uint32_t> dc(n_ctrl), dg((size_t)n_ctrl * gt_stride);
        if (n_ctrl) {
            FK_HIP(hipMemcpyAsync(dc.p, ctrl_ids, (size_t)n_ctrl * 4, hipMemcpyHostToDevice, b.s));
            FK_HIP(hipMemcpyAsync(dg.p, gt, (size_t)n_ctrl * gt_stride * 4, hipMemcpyHostToDevice, b.s));
        }
        nnd_stop_stats(b, b.P.K, n_ctrl, dc.p, dg.p, gt_stride, delta, recall);
    });
}

fk_status fk_step_entry_point(fk_handle *h, const fk_params *p, const fk_state *s, uint32_t init_node,
                              uint32_t *ep) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        check_params(p);
        FK_CHECK_ARG(s && s->pool && s->pool_len && s->out_id && s->out_len && ep, "null argument");
        FK_CHECK_ARG(s->pool_cap >= 1 && s->out_cap >= 1, "capacities must be positive");
        FK_CHECK_ARG(init_node < s->n, "init_node out of range");
        check_lens(s->pool_len, s->n, s->pool_cap, "pool_len > pool_cap");
        check_lens(s->out_len, s->n, s->out_cap, "out_len > out_cap");
        check_pool_ids(s);
        check_out_ids(s, s->out_id, s->out_len);
        select_device(h);
        fk_stats local;
        memset(&local, 0, sizeof(local));
        BuildState b;
        step_begin(b, h, p, s, &local);
        const uint32_t N = b.N;
        FK_CHECK_ARG(b.P.search_L <= 1024, "search_L above 1024 is not supported");
        b.cap = s->pool_cap;
        b.pool.alloc((size_t)N * b.cap);
        b.cnt.alloc(N);
        b.ncap = s->out_cap;
        b.nsg_id.alloc((size_t)N * b.ncap);
        b.nsg_len.alloc(N);
        put_rows(b, b.pool.p, b.cap, s->pool, s->pool_cap, s->pool_cap);
        FK_HIP(hipMemcpyAsync(b.cnt.p, s->pool_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        put_rows(b, b.nsg_id.p, b.ncap, s->out_id, s->out_cap, s->out_cap);
        FK_HIP(hipMemcpyAsync(b.nsg_len.p, s->out_len, (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        compute_ep(b, &init_node);
        *ep = b.ep;
    });
}

fk_status fk_nndescent(fk_handle *h, uint32_t n_used, const fk_params *p, uint64_t *pools, uint32_t *pool_len,
                       uint32_t pool_stride, fk_stats *stats) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        check_params(p);
        FK_CHECK_ARG(n_used >= 2 && n_used <= h->ds.n_total, "n_used must be in [2, n_total]");
        FK_CHECK_ARG(pools && pool_len, "null output");
        select_device(h);
        const double t0 = now_sec();
        fk_stats local;
        memset(&local, 0, sizeof(local));
        BuildState b;
        b.h = h;
        b.s = h->stream;
        b.ds = &h->ds;
        b.N = n_used;
        b.P = shrink_params(*p, n_used);
        FK_CHECK_ARG(pool_stride >= b.P.L, "pool_stride < L");
        b.cap = b.P.MAXL;
        b.st = stats ? stats : &local;
        memset(b.st, 0, sizeof(*b.st));
        b.verbose = (p->flags & FK_FLAG_VERBOSE) != 0;
        b.init_sets({0u, n_used});
        b.prof = (p->flags & FK_FLAG_PROFILE) != 0;
        b.t0 = t0;
        nnd_controls(b);
        nnd_run(b);
        FK_HIP(hipMemcpyAsync(pool_len, b.cnt.p, (size_t)b.N * 4, hipMemcpyDeviceToHost, b.s));
        FK_HIP(hipMemcpy2DAsync(pools, (size_t)pool_stride * 8, b.pool.p, (size_t)b.cap * 8,
                                (size_t)std::min(b.cap, pool_stride) * 8, b.N, hipMemcpyDeviceToHost, b.s));
        b.sync();
        b.st->t_total = now_sec() - t0;
    });
}

}  // extern "C"
