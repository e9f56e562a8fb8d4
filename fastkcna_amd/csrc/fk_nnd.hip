// NN-descent on the GPU: random init, the local join as a distance block per
// node with threshold-filtered candidate buckets, a per-node merge of the
// buckets into the sorted pools, the new/old/reverse resampling step, the
// stop rule over sampled ground truth, and the BridgeView (P2K) rounds.
//
// Reference: KGraphConstructor (src/kgraph.cpp:924-1459), BridgeView
// (src/kgraph.cpp:2478-2663).
//
// Concurrency model.  The reference inserts every join pair into two pools
// under per-node spinlocks while other threads keep joining.  Here a join
// never touches a pool: each pair that passes the target's threshold (the
// pool's L-th distance at iteration start, the reference's `radius` filter)
// becomes a bucket entry for the target, and one workgroup per target then
// merges its bucket into its pool.  The merged pool is the L best distinct
// ids of (pool U candidates) in (dist, id) order, which is what the
// reference's sorted insert converges to up to the order of equal distances,
// and it is deterministic.
#include <math.h>

#include <algorithm>
#include <chrono>

#include "fk_build.h"
#include "fk_device.h"
#include "fk_gram.h"
#include "fk_join.h"
#include "fk_sort.h"

namespace fk {

double now_sec() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void BuildState::sync() { FK_HIP(hipStreamSynchronize(s)); }

void BuildState::init_sets(const std::vector<uint32_t> &lo) {
    FK_CHECK_ARG(lo.size() >= 2 && lo.front() == 0 && lo.back() == N, "sets must cover the nodes");
    n_sets = (uint32_t)lo.size() - 1;
    set_lo = lo;
    set_rng.assign(n_sets, ((uint64_t)P.seed << 32) ^ 0x5DEECE66Dull);
    set_ep.assign(n_sets, 0);
    set_width.assign(n_sets, 0);
    set_active.assign(n_sets, 1);
    set_acc.alloc(2 * (size_t)n_sets);
    if (n_sets > 1) {
        std::vector<uint32_t> of(N);
        for (uint32_t k = 0; k < n_sets; ++k)
            for (uint32_t v = lo[k]; v < lo[k + 1]; ++v) of[v] = k;
        d_set_lo.alloc(n_sets + 1);
        set_of.alloc(N);
        d_set_active.alloc(n_sets);
        FK_HIP(hipMemcpyAsync(d_set_lo.p, lo.data(), lo.size() * 4, hipMemcpyHostToDevice, s));
        FK_HIP(hipMemcpyAsync(set_of.p, of.data(), (size_t)N * 4, hipMemcpyHostToDevice, s));
        FK_HIP(hipMemcpyAsync(d_set_active.p, set_active.data(), n_sets, hipMemcpyHostToDevice, s));
        sync();
    }
}

hipEvent_t BuildState::ev_get() {
    if (!ev_free.empty()) {
        hipEvent_t e = ev_free.back();
        ev_free.pop_back();
        return e;
    }
    hipEvent_t e;
    FK_HIP(hipEventCreate(&e));
    return e;
}

hipEvent_t BuildState::prof_begin() {
    if (!prof) return nullptr;
    hipEvent_t e = ev_get();
    FK_HIP(hipEventRecord(e, s));
    return e;
}

void BuildState::prof_end(hipEvent_t start, int fam, double bytes) {
    if (!prof || !start) return;
    hipEvent_t e = ev_get();
    FK_HIP(hipEventRecord(e, s));
    ev_pending.push_back({start, e, fam, bytes});
}

void BuildState::prof_set_last_bytes(double bytes) {
    if (!ev_pending.empty()) ev_pending.back().bytes = bytes;
}

void BuildState::prof_flush() {
    for (auto &r : ev_pending) {
        FK_HIP(hipEventSynchronize(r.b));
        float ms = 0.f;
        FK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        if (st && r.fam >= 0 && r.fam < FK_PROF_N) {
            st->prof_ms[r.fam] += ms;
            st->prof_bytes[r.fam] += r.bytes;
            st->prof_launches[r.fam] += 1;
        }
        ev_free.push_back(r.a);
        ev_free.push_back(r.b);
    }
    ev_pending.clear();
}

BuildState::~BuildState() {
    if (s) (void)hipStreamSynchronize(s);
    for (auto &r : ev_pending) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : ev_free) (void)hipEventDestroy(e);
}

__device__ __forceinline__ bool sorted_contains(const uint64_t *a, uint32_t n, uint64_t x) {
    const uint32_t i = lower_bound_cnt(a, n, x);
    return i < n && a[i] == x;
}

// ---------------------------------------------------------------------------
// Random init (src/kgraph.cpp:1001-1044).  GenRandom (:82-110): `size`
// distinct values in [0, N) -- draw in [0, N - size), sort, make strictly
// increasing, rotate by a random offset.  One 64-thread block per node.
// ---------------------------------------------------------------------------
__device__ void gen_random_block(uint32_t *a, uint32_t size, uint32_t N, uint64_t seed, uint64_t stream) {
    if (size >= N) {
        for (uint32_t i = threadIdx.x; i < size; i += blockDim.x) a[i] = i % N;
        __syncthreads();
        return;
    }
    const uint32_t np = next_pow2(size);
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x)
        a[i] = i < size ? fk_rand32(seed, stream, i) % (N - size) : 0xFFFFFFFFu;
    __syncthreads();
    block_bitonic_sort(a, np);
    if (threadIdx.x == 0) {
        for (uint32_t i = 1; i < size; ++i)
            if (a[i] <= a[i - 1]) a[i] = a[i - 1] + 1;
    }
    __syncthreads();
    const uint64_t off = fk_rand32(seed, stream, size) % N;
    for (uint32_t i = threadIdx.x; i < size; i += blockDim.x) a[i] = (uint32_t)(((uint64_t)a[i] + off) % N);
    __syncthreads();
}

// With sets, node n of set s draws from the set's own ids, with the streams
// of its id within the set (a build of the set's rows alone draws the same).
__global__ __launch_bounds__(64) void k_nnd_init(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                 uint32_t N, uint32_t S, uint32_t cap, uint32_t Lp, uint64_t seed,
                                                 const uint32_t *__restrict__ set_lo,
                                                 const uint32_t *__restrict__ set_of,
                                                 uint64_t *__restrict__ pool, uint8_t *__restrict__ pflg,
                                                 uint32_t *__restrict__ cnt, uint32_t *__restrict__ Marr,
                                                 float *__restrict__ thr, uint32_t *__restrict__ lists) {
    __shared__ uint32_t s_a[512];
    __shared__ uint64_t s_k[256];
    const uint32_t n = blockIdx.x;
    uint32_t lo = 0, Ns = N;
    if (set_of) {
        const uint32_t k = set_of[n];
        lo = set_lo[k];
        Ns = set_lo[k + 1] - lo;
    }
    const uint32_t ln = n - lo;
    // nn_new: 2S random ids (may include n itself, as in the reference)
    gen_random_block(s_a, 2 * S, Ns, seed, 2ull * ln);
    for (uint32_t i = threadIdx.x; i < 2 * S; i += blockDim.x) lists[(uint64_t)n * 2 * S + i] = lo + s_a[i];
    __syncthreads();
    // pool: S random ids != n from S + 1 draws
    gen_random_block(s_a, S + 1, Ns, seed, 2ull * ln + 1);
    if (threadIdx.x == 0) {
        uint32_t i = 0;
        for (uint32_t l = 0; l < S; ++l) {
            if (s_a[i] == ln) ++i;
            s_a[256 + l] = lo + s_a[i++];
        }
    }
    __syncthreads();
    const uint32_t np = next_pow2(S < 2 ? 2 : S);
    for (uint32_t l = threadIdx.x; l < np; l += blockDim.x) {
        if (l < S) {
            const uint32_t id = s_a[256 + l];
            s_k[l] = make_key(l2_canon(rows + (size_t)id * stride, rows + (size_t)n * stride, dpad), id);
        } else {
            s_k[l] = FK_KEY_EMPTY;
        }
    }
    __syncthreads();
    block_bitonic_sort(s_k, np);
    uint64_t *pr = pool + (uint64_t)n * cap;
    uint8_t *pf = pflg + (uint64_t)n * cap;
    for (uint32_t l = threadIdx.x; l < S; l += blockDim.x) {
        pr[l] = s_k[l];
        pf[l] = FK_FLG_NEW | FK_FLG_ISNEW;
    }
    if (threadIdx.x == 0) {
        cnt[n] = S;
        Marr[n] = S;
        thr[n] = S >= Lp ? key_dist(s_k[Lp - 1]) : INFINITY;
    }
}

// ---------------------------------------------------------------------------
// Pair counts per node: C(nn, 2) + nn * no.
// ---------------------------------------------------------------------------
__global__ void k_pair_counts(const uint64_t *__restrict__ list_off, const uint32_t *__restrict__ n_new, uint32_t N,
                              uint32_t *__restrict__ out) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    const uint32_t T = (uint32_t)(list_off[n + 1] - list_off[n]);
    const uint32_t nn = n_new[n];
    out[n] = nn * (nn ? nn - 1 : 0) / 2 + nn * (T - nn);
}

// ---------------------------------------------------------------------------
// Local join, pass 1: pair distances (kept for pass 2), per-target candidate
// counts and one bucket reservation per (join, target).
// src/kgraph.cpp:970-987 enumerates new x new (i < j) and new x old; a pair
// joins j into pool(i) and i into pool(j).  Pairs of equal ids add nothing.
// ---------------------------------------------------------------------------
// Dynamic LDS of the two passes, sized by the longest list (L + 2R): pass 1
// keeps per-list-slot threshold, count and id and stages the sampled rows
// when they fit; pass 2 keeps ids, thresholds and bucket bases.
size_t join_count_lds(uint32_t tcap, uint32_t rows_bytes) { return rows_bytes + (size_t)tcap * 12; }
size_t join_fill_lds(uint32_t tcap) { return (size_t)tcap * 16; }
#define FK_FILL_GROUP 4  // list slots a fill wave takes at once

__global__ __launch_bounds__(256) void k_join_count(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                                                    const uint32_t *__restrict__ lists,
                                                    const uint64_t *__restrict__ list_off,
                                                    const uint32_t *__restrict__ n_new,
                                                    const uint64_t *__restrict__ pair_off,
                                                    const float *__restrict__ thr, float *__restrict__ pdist,
                                                    uint32_t *__restrict__ resv, uint32_t *__restrict__ bcnt,
                                                    uint32_t N, const uint32_t *__restrict__ order, uint32_t tcap,
                                                    uint32_t rows_bytes, uint32_t t_lo, uint32_t t_hi) {
    extern __shared__ __align__(16) float s_rows[];
    float *s_thr = s_rows + rows_bytes / 4;
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
        s_c[t] = 0;
    }
    __syncthreads();
    join_block(rows, stride, dpad, s_id, T, nn, s_rows, [&](uint32_t x, uint32_t y, float dd) {
        o[join_pair_index(x, y, T)] = dd;
        if (s_id[x] != s_id[y]) {
            if (dd <= s_thr[x]) atomicAdd(&s_c[x], 1u);
            if (dd <= s_thr[y]) atomicAdd(&s_c[y], 1u);
        }
    }, rows_bytes);
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const uint32_t c = s_c[t];
        resv[lo + t] = c ? atomicAdd(&bcnt[s_id[t]], c) : 0u;
    }
}

// Pass 2: write the candidates into their reserved bucket slots.  One
// wavefront per list slot t: t's partners (every other slot when t is new,
// the new slots when t is old: the pairs pass 1 computed) are taken in list
// order 64 at a time, and those with another id and a distance within t's
// threshold fill t's reserved run of slots in that order -- the pairs and
// conditions of pass 1, so the counts agree, and every store instruction
// writes one contiguous run instead of 64 scattered slots.
__global__ __launch_bounds__(256) void k_join_fill(const uint32_t *__restrict__ lists,
                                                   const uint64_t *__restrict__ list_off,
                                                   const uint32_t *__restrict__ n_new,
                                                   const uint64_t *__restrict__ pair_off,
                                                   const float *__restrict__ thr, const float *__restrict__ pdist,
                                                   const uint32_t *__restrict__ resv,
                                                   const uint64_t *__restrict__ boff, uint64_t *__restrict__ bkey,
                                                   uint32_t *__restrict__ bsrc, uint32_t N,
                                                   const uint32_t *__restrict__ order, uint32_t tcap) {
    extern __shared__ __align__(16) uint64_t s_base[];
    float *s_thr = reinterpret_cast<float *>(s_base + tcap);
    uint32_t *s_id = reinterpret_cast<uint32_t *>(s_thr + tcap);
    const uint32_t node = node_of_block(order, N);
    if (node >= N) return;
    const uint64_t lo = list_off[node];
    const uint32_t T = (uint32_t)(list_off[node + 1] - lo);
    const uint32_t nn = n_new[node];
    const uint32_t *lst = lists + lo;
    const float *o = pdist + pair_off[node];
    // every slot's id, threshold and first reserved bucket slot, read at once
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const uint32_t id = lst[t];
        s_id[t] = id;
        s_thr[t] = thr[id];
        s_base[t] = boff[id] + resv[lo + t];
    }
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint32_t nw = blockDim.x >> 6;
    if (T <= 64) {
        // one chunk of partners per slot: FK_FILL_GROUP slots per wave at a
        // time, all their distance loads in flight before any store
        for (uint32_t t0 = threadIdx.x >> 6; t0 < T; t0 += nw * FK_FILL_GROUP) {
            float dd[FK_FILL_GROUP];
            bool ok[FK_FILL_GROUP];
            const uint32_t y = lane;
#pragma unroll
            for (uint32_t g = 0; g < FK_FILL_GROUP; ++g) {
                const uint32_t t = t0 + g * nw;
                ok[g] = false;
                dd[g] = 0.f;
                if (t < T && y < (t < nn ? T : nn) && y != t && s_id[y] != s_id[t]) {
                    ok[g] = true;
                    dd[g] = o[t < y ? join_pair_index(t, y, T) : join_pair_index(y, t, T)];
                }
            }
#pragma unroll
            for (uint32_t g = 0; g < FK_FILL_GROUP; ++g) {
                const uint32_t t = t0 + g * nw;
                if (t >= T) break;  // uniform
                const bool keep = ok[g] && dd[g] <= s_thr[t];
                const uint64_t m = __ballot(keep);
                if (keep) {
                    const uint64_t at = s_base[t] + __builtin_amdgcn_mbcnt_hi(
                                                        (uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                    bkey[at] = make_key(dd[g], s_id[y]);
                    bsrc[at] = node;
                }
            }
        }
        return;
    }
    for (uint32_t t = threadIdx.x >> 6; t < T; t += nw) {
        const uint32_t it = s_id[t];
        const float th = s_thr[t];
        uint64_t sl = s_base[t];
        const uint32_t ylim = t < nn ? T : nn;
        for (uint32_t y0 = 0; y0 < ylim; y0 += 64) {
            const uint32_t y = y0 + lane;
            bool ok = false;
            float dd = 0.f;
            uint32_t iy = 0;
            if (y < ylim && y != t) {
                iy = s_id[y];
                if (iy != it) {
                    dd = o[t < y ? join_pair_index(t, y, T) : join_pair_index(y, t, T)];
                    ok = dd <= th;
                }
            }
            const uint64_t m = __ballot(ok);
            if (ok) {
                const uint64_t at = sl + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                bkey[at] = make_key(dd, iy);
                bsrc[at] = node;
            }
            sl += (uint32_t)__builtin_popcountll(m);
        }
    }
}

// ---------------------------------------------------------------------------
// Bucket merge, one workgroup per target node.  The merged pool is the best
// Lp distinct ids of pool ∪ candidates.  Every candidate id that ends up in
// it without having been in the iteration-start pool records the smallest
// source join among all candidates with its key, and if it ranks before the
// iteration-start K-th entry it marks that join `found`
// (src/kgraph.cpp:1111-1114: an insert at position < K).  Both are
// independent of the order of the bucket, so the merge is deterministic.
// ---------------------------------------------------------------------------

// The bucket is taken in chunks.  A chunk's candidates are de-duplicated by
// id in an LDS hash set before sorting (equal ids carry equal keys: the
// distance to the target is symmetric and computed in one canonical order),
// keeping the smallest source join per id; candidates that cannot enter the
// pool are dropped first, and candidates already in the pool only lower the
// recorded source of an entry inserted by an earlier chunk.  A key that
// cannot enter the pool at its chunk never can later (the pool only
// improves), so each inserted entry's recorded source is the minimum over
// the whole bucket.  Only the distinct new candidates are sorted, then
// merged into the pool in place.  LDS: pool keys, flags and sources,
// FK_NM_CH candidate keys and a 2 x FK_NM_CH-slot id set (18.8 KB at
// MAXL 500: eight workgroups per CU).
#ifndef FK_NM_CH
#define FK_NM_CH 512
#endif
#ifndef FK_NM_RANK
#define FK_NM_RANK 256  // up to this many new keys (at most 256) are ranked by counting, more sorted
#endif
#define FK_NM_HASH (2 * FK_NM_CH)
#define FK_NM_HBITS (FK_NM_CH == 256 ? 9 : FK_NM_CH == 512 ? 10 : 11)
#define FK_NM_NOSRC 0xFFFFFFFFu

__device__ __forceinline__ uint32_t nm_hash(uint32_t id) { return (id * 2654435761u) >> (32 - FK_NM_HBITS); }
static_assert(FK_NM_HASH == (1u << FK_NM_HBITS), "nm_hash yields FK_NM_HBITS bits");

size_t nnd_merge_lds(uint32_t cap) {
    return (size_t)cap * 8 + (size_t)FK_NM_CH * 8 + (size_t)FK_NM_HASH * 8 + (size_t)cap * 4 + (size_t)cap;
}

__global__ __launch_bounds__(256) void k_nnd_merge(uint32_t cap, uint32_t Lp, uint32_t K, uint64_t *__restrict__ pool,
                                                   uint8_t *__restrict__ pflg, uint32_t *__restrict__ cnt,
                                                   float *__restrict__ thr, const uint32_t *__restrict__ bcnt,
                                                   const uint64_t *__restrict__ boff,
                                                   const uint64_t *__restrict__ bkey,
                                                   const uint32_t *__restrict__ bsrc, uint8_t *__restrict__ found) {
    extern __shared__ __align__(16) uint8_t smem[];
    uint64_t *s_p = reinterpret_cast<uint64_t *>(smem);
    uint64_t *s_u = s_p + cap;
    uint32_t *s_tid = reinterpret_cast<uint32_t *>(s_u + FK_NM_CH);
    uint32_t *s_tsrc = s_tid + FK_NM_HASH;
    uint32_t *s_ps = s_tsrc + FK_NM_HASH;  // source of entries inserted by this merge
    uint8_t *s_pf = reinterpret_cast<uint8_t *>(s_ps + cap);
    __shared__ uint32_t s_np, s_nu;
    const uint32_t i = blockIdx.x;
    const uint32_t nb = bcnt[i];
    if (nb == 0) return;
    uint64_t *pr = pool + (uint64_t)i * cap;
    uint8_t *pf = pflg + (uint64_t)i * cap;
    const uint32_t np0 = cnt[i];
    const uint64_t base = boff[i];
    // one chunk of bucket entries (key and source) per thread slot, read
    // unconditionally in bounds (slots past the chunk repeat its last entry)
    constexpr uint32_t CR = FK_NM_CH / 256;
    uint64_t ck[CR];
    uint32_t cs[CR];
    auto load_chunk = [&](uint32_t c0, uint32_t m) __attribute__((always_inline)) {
#pragma unroll
        for (uint32_t r = 0; r < CR; ++r) {
            const uint32_t k = r * 256 + threadIdx.x;
            const uint64_t at = base + c0 + (k < m ? k : m - 1);
            ck[r] = bkey[at];
            cs[r] = bsrc[at];
        }
    };
    // the pool and the first chunk: every load in flight before any store
    {
        constexpr uint32_t PR = FK_POOL_MAX / 256;
        uint64_t pk[PR];
        uint32_t pfl[PR];
#pragma unroll
        for (uint32_t r = 0; r < PR; ++r) {
            const uint32_t j = r * 256 + threadIdx.x;
            if (r * 256 < np0) {
                const uint32_t jj = j < np0 ? j : np0 - 1;
                pk[r] = pr[jj];
                pfl[r] = pf[jj];
            }
        }
        load_chunk(0, nb < FK_NM_CH ? nb : FK_NM_CH);
#pragma unroll
        for (uint32_t r = 0; r < PR; ++r) {
            const uint32_t j = r * 256 + threadIdx.x;
            if (j < np0) {
                s_p[j] = pk[r];
                s_pf[j] = (uint8_t)pfl[r];
                s_ps[j] = FK_NM_NOSRC;
            }
        }
    }
    if (threadIdx.x == 0) s_np = np0;
    __syncthreads();
    const uint64_t kth = np0 >= K ? s_p[K - 1] : FK_KEY_EMPTY;
    uint32_t first_change = 0xFFFFFFFFu;  // smallest pool position rewritten (thread 0's copy)
    for (uint32_t c0 = 0; c0 < nb; c0 += FK_NM_CH) {
        const uint32_t m = (nb - c0) < FK_NM_CH ? (nb - c0) : FK_NM_CH;
        if (c0 > 0) load_chunk(c0, m);
        for (uint32_t j = threadIdx.x; j < FK_NM_HASH; j += blockDim.x) {
            s_tid[j] = 0xFFFFFFFFu;
            s_tsrc[j] = FK_NM_NOSRC;
        }
        if (threadIdx.x == 0) s_nu = 0;
        __syncthreads();
        const uint32_t np = s_np;
        const uint64_t lth = np >= Lp ? s_p[Lp - 1] : FK_KEY_EMPTY;
#pragma unroll
        for (uint32_t r = 0; r < CR; ++r) {
            const uint32_t k = r * 256 + threadIdx.x;
            const uint64_t key = ck[r];
            if (k >= m || key >= lth) continue;
            const uint32_t src = cs[r];
            const uint32_t at = lower_bound_cnt(s_p, np, key);
            if (at < np && s_p[at] == key) {
                if (s_ps[at] != FK_NM_NOSRC) atomicMin(&s_ps[at], src);
                continue;
            }
            const uint32_t id = key_id(key);
            uint32_t h = nm_hash(id);
            while (true) {
                const uint32_t old = atomicCAS(&s_tid[h], 0xFFFFFFFFu, id);
                if (old == 0xFFFFFFFFu) {
                    s_u[atomicAdd(&s_nu, 1u)] = key;
                    break;
                }
                if (old == id) break;
                h = (h + 1) & (FK_NM_HASH - 1);
            }
            atomicMin(&s_tsrc[h], src);
        }
        __syncthreads();
        const uint32_t nu = s_nu;
        if (nu > 0) {
            if (nu <= FK_NM_RANK) {
                // few keys: each one's rank is the number of smaller keys (the
                // ids are distinct, so are the keys), counted against the
                // whole set by broadcast reads; one scatter sorts them
                uint64_t rk = FK_KEY_EMPTY;
                uint32_t rr = 0;
                if (threadIdx.x < nu) {
                    rk = s_u[threadIdx.x];
                    for (uint32_t j = 0; j < nu; ++j) rr += s_u[j] < rk ? 1u : 0u;
                }
                __syncthreads();
                if (threadIdx.x < nu) s_u[rr] = rk;
                __syncthreads();
            } else {
                const uint32_t m2 = next_pow2(nu);
                for (uint32_t k = nu + threadIdx.x; k < m2; k += blockDim.x) s_u[k] = FK_KEY_EMPTY;
                __syncthreads();
                if (m2 <= 512) block_sort512(s_u, m2);
                else block_bitonic_sort(s_u, m2);
            }
            // each candidate's merged position against the pool before it
            // moves, and its smallest source (at most FK_NM_CH / 256 per thread)
            uint32_t upos[FK_NM_CH / 256], usrc[FK_NM_CH / 256];
#pragma unroll
            for (uint32_t r = 0; r < FK_NM_CH / 256; ++r) {
                const uint32_t k = r * 256 + threadIdx.x;
                upos[r] = 0xFFFFFFFFu;
                usrc[r] = FK_NM_NOSRC;
                if (k < nu) {
                    const uint64_t key = s_u[k];
                    upos[r] = k + lower_bound_cnt(s_p, np, key);
                    if (upos[r] < Lp) {
                        const uint32_t id = key_id(key);
                        uint32_t h = nm_hash(id);
                        while (s_tid[h] != id) h = (h + 1) & (FK_NM_HASH - 1);
                        usrc[r] = s_tsrc[h];
                    }
                }
            }
            // pool entries move up to j + (#candidates below), last chunk
            // first, each chunk read before it is written; then the
            // candidates fill the remaining slots
            for (uint32_t ch = (np + 255) / 256; ch-- > 0;) {
                const uint32_t j = ch * 256 + threadIdx.x;
                uint64_t key = 0;
                uint8_t f = 0;
                uint32_t ps = FK_NM_NOSRC;
                uint32_t pos = 0xFFFFFFFFu;
                if (j < np) {
                    key = s_p[j];
                    f = s_pf[j];
                    ps = s_ps[j];
                    pos = j + lower_bound_cnt(s_u, nu, key);
                }
                __syncthreads();
                if (pos < Lp) { s_p[pos] = key; s_pf[pos] = f; s_ps[pos] = ps; }
                __syncthreads();
            }
#pragma unroll
            for (uint32_t r = 0; r < FK_NM_CH / 256; ++r)
                if (upos[r] < Lp) {
                    s_p[upos[r]] = s_u[r * 256 + threadIdx.x];
                    s_pf[upos[r]] = (uint8_t)(FK_FLG_NEW | FK_FLG_ISNEW);
                    s_ps[upos[r]] = usrc[r];
                }
            if (threadIdx.x == 0) {
                const uint32_t u0 = upos[0];  // the smallest candidate's position
                first_change = u0 < first_change ? u0 : first_change;
                s_np = (np + nu) < Lp ? (np + nu) : Lp;
            }
        }
        __syncthreads();
    }
    // entries below the first inserted position are unchanged
    __shared__ uint32_t s_first;
    if (threadIdx.x == 0) s_first = first_change;
    __syncthreads();
    const uint32_t np = s_np;
    const uint32_t st = s_first < np ? s_first : np;
    for (uint32_t j = st + threadIdx.x; j < np; j += blockDim.x) {
        pr[j] = s_p[j];
        pf[j] = s_pf[j];
        if (s_ps[j] != FK_NM_NOSRC && s_p[j] < kth) found[s_ps[j]] = 1;
    }
    if (threadIdx.x == 0) {
        cnt[i] = np;
        thr[i] = np >= Lp ? key_dist(s_p[Lp - 1]) : INFINITY;
    }
}

// ---------------------------------------------------------------------------
// Stop-rule statistics (src/kgraph.cpp:1371-1394).
// delta = mean over nodes of (#flagged among the first K entries) / K;
// recall = mean over controls of |pool[0..min(L,K)) ∩ gt[0..K)| / K.
// ---------------------------------------------------------------------------
// One wavefront per node, grid-stride; one atomic per block.
// With sets, acc[set] per set: each wave takes a contiguous run of nodes
// (a set is a contiguous id range, so a run spans one or two sets) and adds
// its running sums to the set's counters when the set changes and at the
// end -- a few atomics per wave instead of one per node, which held every
// wave of a many-set batch behind the same few addresses.
// This lane's share of the count of FK_FLG_NEW among the first lim flags of
// a pool row (the wave's lanes together count each flag once).  Rows of
// cap flags with cap % 4 == 0 start on 4-byte boundaries and are read four
// flags per load, the last word keeping only its leading bytes.
__device__ __forceinline__ uint32_t lane_new_flags(const uint8_t *pf, uint32_t lim, uint32_t cap) {
    uint32_t cn = 0;
    if ((cap & 3u) == 0) {
        static_assert(FK_FLG_NEW == 1u, "the word count takes bit 0 of every flag byte");
        const uint32_t *pw = reinterpret_cast<const uint32_t *>(pf);
        const uint32_t nw = (lim + 3u) >> 2;
        for (uint32_t w = lane_id(); w < nw; w += 64) {
            uint32_t v = pw[w] & 0x01010101u;
            const uint32_t rem = lim - w * 4u;
            if (rem < 4u) v &= (1u << (8u * rem)) - 1u;
            cn += (uint32_t)__builtin_popcount(v);
        }
    } else {
        for (uint32_t l = lane_id(); l < lim; l += 64) cn += (pf[l] & FK_FLG_NEW) ? 1u : 0u;
    }
    return cn;
}

__global__ __launch_bounds__(1024) void k_delta(const uint8_t *__restrict__ pflg, const uint32_t *__restrict__ cnt,
                                                uint32_t N, uint32_t cap, uint32_t K,
                                                const uint32_t *__restrict__ set_of,
                                                unsigned long long *__restrict__ acc) {
    const uint32_t wpb = blockDim.x >> 6;
    if (set_of) {  // uniform
        const uint32_t nw = gridDim.x * wpb, chunk = (N + nw - 1) / nw;
        const uint32_t lo = (blockIdx.x * wpb + (threadIdx.x >> 6)) * chunk;
        const uint32_t hi = lo + chunk < N ? lo + chunk : N;
        uint32_t cur = 0xFFFFFFFFu;
        unsigned long long run = 0;
        for (uint32_t n = lo; n < hi; ++n) {
            const uint32_t lim = cnt[n] < K ? cnt[n] : K;
            const unsigned long long cn = wave_sum((unsigned long long)lane_new_flags(pflg + (uint64_t)n * cap, lim, cap));
            const uint32_t k = set_of[n];
            if (k != cur) {
                if (lane_id() == 0 && run) atomicAdd(&acc[cur], run);
                cur = k;
                run = 0;
            }
            run += cn;
        }
        if (lane_id() == 0 && run) atomicAdd(&acc[cur], run);
        return;
    }
    unsigned long long c = 0;
    for (uint32_t n = blockIdx.x * wpb + (threadIdx.x >> 6); n < N; n += gridDim.x * wpb) {
        const uint32_t lim = cnt[n] < K ? cnt[n] : K;
        c += lane_new_flags(pflg + (uint64_t)n * cap, lim, cap);
    }
    c = block_sum(c);
    if (threadIdx.x == 0 && c) atomicAdd(acc, c);
}

__global__ __launch_bounds__(256) void k_recall(const uint64_t *__restrict__ pool, const uint32_t *__restrict__ cnt,
                                                uint32_t cap, const uint32_t *__restrict__ ctrl_ids,
                                                const uint32_t *__restrict__ gt, uint32_t gt_stride, uint32_t K,
                                                uint32_t *__restrict__ found_out) {
    __shared__ uint32_t s_gt[2 * FK_POOL_MAX];
    __shared__ uint32_t s_found;
    const uint32_t c = blockIdx.x;
    const uint32_t np2 = next_pow2(K < 2 ? 2 : K);
    for (uint32_t k = threadIdx.x; k < np2; k += blockDim.x) s_gt[k] = k < K ? gt[(uint64_t)c * gt_stride + k] : 0xFFFFFFFFu;
    if (threadIdx.x == 0) s_found = 0;
    __syncthreads();
    block_bitonic_sort(s_gt, np2);
    const uint32_t id = ctrl_ids[c];
    const uint32_t lim = cnt[id] < K ? cnt[id] : K;
    const uint64_t *pr = pool + (uint64_t)id * cap;
    uint32_t f = 0;
    for (uint32_t l = threadIdx.x; l < lim; l += blockDim.x) {
        const uint32_t x = key_id(pr[l]);
        const uint32_t pos = lower_bound_cnt(s_gt, K, x);
        if (pos < K && s_gt[pos] == x) ++f;
    }
    if (f) atomicAdd(&s_found, f);
    __syncthreads();
    if (threadIdx.x == 0) found_out[c] = s_found;
}

// ---------------------------------------------------------------------------
// update() (src/kgraph.cpp:1118-1202).
// ---------------------------------------------------------------------------
// M growth and radiusM.
// A node of a set that has stopped gets M = 0: it samples nothing and joins
// nothing from then on.
__global__ void k_upd_m(const uint64_t *__restrict__ pool, const uint8_t *__restrict__ pflg,
                        const uint32_t *__restrict__ cnt, const uint8_t *__restrict__ found, uint32_t N, uint32_t cap,
                        uint32_t S, const uint32_t *__restrict__ set_of, const uint8_t *__restrict__ set_active,
                        uint32_t *__restrict__ Marr, float *__restrict__ radiusM) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    if (set_of && !set_active[set_of[n]]) {
        Marr[n] = 0;
        radiusM[n] = INFINITY;
        return;
    }
    uint32_t M = Marr[n];
    const uint32_t L = cnt[n];
    if (found[n]) {
        const uint32_t maxl = (M + S) < L ? (M + S) : L;
        const uint8_t *pf = pflg + (uint64_t)n * cap;
        uint32_t c = 0, l = 0;
        while (l < maxl && c < S) {
            if (pf[l] & FK_FLG_NEW) ++c;
            ++l;
        }
        M = l;
    }
    if (M > L) M = L;
    if (M == 0) M = L < 1 ? 0 : 1;
    Marr[n] = M;
    radiusM[n] = M ? key_dist(pool[(uint64_t)n * cap + M - 1]) : INFINITY;
}

// Forward new/old counts and reverse-push counts.
__global__ void k_upd_count(const uint64_t *__restrict__ pool, const uint8_t *__restrict__ pflg, uint32_t N,
                            uint32_t cap, const uint32_t *__restrict__ Marr, const float *__restrict__ radiusM,
                            uint32_t *__restrict__ fcnt_new, uint32_t *__restrict__ fcnt_old,
                            uint32_t *__restrict__ rcnt_new, uint32_t *__restrict__ rcnt_old) {
    const uint32_t n = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (n >= N) return;
    const uint32_t M = Marr[n];
    const uint64_t *pr = pool + (uint64_t)n * cap;
    const uint8_t *pf = pflg + (uint64_t)n * cap;
    uint32_t fn = 0;
    for (uint32_t l = lane_id(); l < M; l += 64) {
        const uint64_t e = pr[l];
        const uint32_t o = key_id(e);
        const bool f = (pf[l] & FK_FLG_NEW) != 0;
        fn += f ? 1u : 0u;
        if (key_dist(e) > radiusM[o]) atomicAdd(f ? &rcnt_new[o] : &rcnt_old[o], 1u);
    }
    fn = wave_sum(fn);
    if (lane_id() == 0) { fcnt_new[n] = fn; fcnt_old[n] = M - fn; }
}

__global__ void k_list_sizes(const uint32_t *__restrict__ fcnt_new, const uint32_t *__restrict__ fcnt_old,
                             const uint32_t *__restrict__ rcnt_new, const uint32_t *__restrict__ rcnt_old, uint32_t N,
                             uint32_t R, uint32_t *__restrict__ n_new, uint32_t *__restrict__ lsize,
                             uint32_t *__restrict__ lmax) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t t = 0;
    if (n < N) {
        const uint32_t rn = (R && rcnt_new[n] > R) ? R : rcnt_new[n];
        const uint32_t ro = (R && rcnt_old[n] > R) ? R : rcnt_old[n];
        n_new[n] = fcnt_new[n] + rn;
        t = fcnt_new[n] + rn + fcnt_old[n] + ro;
        lsize[n] = t;
    }
    for (int o = 32; o > 0; o >>= 1) t = max(t, (uint32_t)__shfl_xor((int)t, o, 64));
    if (lane_id() == 0 && t) atomicMax(lmax, t);
}

// Forward lists, reverse bucket fill, flag clearing.
__global__ void k_upd_fill(const uint64_t *__restrict__ pool, uint8_t *__restrict__ pflg, uint32_t N, uint32_t cap,
                           const uint32_t *__restrict__ Marr, const float *__restrict__ radiusM,
                           const uint64_t *__restrict__ list_off, const uint32_t *__restrict__ n_new,
                           const uint32_t *__restrict__ fcnt_new, uint32_t *__restrict__ lists,
                           const uint64_t *__restrict__ roff_new, const uint64_t *__restrict__ roff_old,
                           uint32_t *__restrict__ rcur_new, uint32_t *__restrict__ rcur_old,
                           uint32_t *__restrict__ rbuf_new, uint32_t *__restrict__ rbuf_old) {
    const uint32_t n = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (n >= N) return;
    const uint32_t lane = lane_id();
    const uint32_t M = Marr[n];
    const uint64_t *pr = pool + (uint64_t)n * cap;
    uint8_t *pf = pflg + (uint64_t)n * cap;
    uint32_t *ln = lists + list_off[n];
    uint32_t *lo = ln + n_new[n];
    uint32_t bn = 0, bo = 0;
    for (uint32_t l0 = 0; l0 < M; l0 += 64) {
        const uint32_t l = l0 + lane;
        const bool v = l < M;
        uint64_t e = 0;
        uint8_t fl = 0;
        if (v) { e = pr[l]; fl = pf[l]; }
        const bool f = v && (fl & FK_FLG_NEW);
        const uint64_t mn = wave_ballot(f), mo = wave_ballot(v && !f);
        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        if (f) ln[bn + __popcll(mn & below)] = key_id(e);
        if (v && !f) lo[bo + __popcll(mo & below)] = key_id(e);
        bn += __popcll(mn);
        bo += __popcll(mo);
        if (v) {
            const uint32_t o = key_id(e);
            if (key_dist(e) > radiusM[o]) {
                if (f) rbuf_new[roff_new[o] + atomicAdd(&rcur_new[o], 1u)] = n;
                else rbuf_old[roff_old[o] + atomicAdd(&rcur_old[o], 1u)] = n;
            }
            if (f) pf[l] = fl & (uint8_t)~FK_FLG_NEW;
        }
    }
    (void)fcnt_new;
}

// Reverse-list cap: when more than R nodes pushed, keep a uniform random R
// subset (the reference shuffles with rand() and truncates, :1189-1200); the
// subset is the R smallest counter-based hashes, so it is reproducible.
#define FK_SEL_CH 2048
#define FK_REV_NPB 4
__global__ __launch_bounds__(256) void k_rev_select(uint32_t N, uint32_t R, uint64_t seed, uint32_t it,
                                                    const uint32_t *__restrict__ rcnt_new,
                                                    const uint32_t *__restrict__ rcnt_old,
                                                    const uint64_t *__restrict__ roff_new,
                                                    const uint64_t *__restrict__ roff_old,
                                                    const uint32_t *__restrict__ rbuf_new,
                                                    const uint32_t *__restrict__ rbuf_old,
                                                    const uint64_t *__restrict__ list_off,
                                                    const uint32_t *__restrict__ n_new,
                                                    const uint32_t *__restrict__ fcnt_new,
                                                    const uint32_t *__restrict__ fcnt_old,
                                                    const uint32_t *__restrict__ set_lo,
                                                    const uint32_t *__restrict__ set_of,
                                                    uint32_t *__restrict__ lists) {
    __shared__ uint64_t s_k[FK_SEL_CH];
    // FK_REV_NPB consecutive nodes per workgroup: most lists are only copied,
    // and a workgroup per node would make the launch dispatch-bound
    const uint32_t o_end = (blockIdx.x + 1) * FK_REV_NPB < N ? (blockIdx.x + 1) * FK_REV_NPB : N;
    for (uint32_t o = blockIdx.x * FK_REV_NPB; o < o_end; ++o) {
        // the hashes take ids within the node's set (a build of the set alone
        // selects the same entries)
        const uint32_t lo = set_of ? set_lo[set_of[o]] : 0u;
        for (int which = 0; which < 2; ++which) {
            const uint32_t c = which == 0 ? rcnt_new[o] : rcnt_old[o];
            const uint32_t *src = which == 0 ? rbuf_new + roff_new[o] : rbuf_old + roff_old[o];
            uint32_t *dst = lists + list_off[o] + (which == 0 ? fcnt_new[o] : n_new[o] + fcnt_old[o]);
            if (c == 0) continue;
            if (R == 0 || c <= R) {
                for (uint32_t k = threadIdx.x; k < c; k += blockDim.x) dst[k] = src[k];
                continue;
            }
            // running best-R selection over chunks
            uint32_t have = 0;
            const uint64_t stream = ((uint64_t)it << 33) | ((uint64_t)which << 32) | (o - lo);
            for (uint32_t c0 = 0; c0 < c; c0 += (FK_SEL_CH - R)) {
                const uint32_t m = (c - c0) < (FK_SEL_CH - R) ? (c - c0) : (FK_SEL_CH - R);
                const uint32_t tot = have + m;
                const uint32_t np2 = next_pow2(tot < 2 ? 2 : tot);
                for (uint32_t k = threadIdx.x; k < np2; k += blockDim.x) {
                    if (k >= have) {
                        if (k < tot) {
                            const uint32_t s = src[c0 + (k - have)];
                            s_k[k] = ((uint64_t)fk_rand32(seed, stream, s - lo) << 32) | s;
                        } else {
                            s_k[k] = FK_KEY_EMPTY;
                        }
                    }
                }
                __syncthreads();
                block_bitonic_sort(s_k, np2);
                have = tot < R ? tot : R;
                __syncthreads();
            }
            for (uint32_t k = threadIdx.x; k < R; k += blockDim.x) dst[k] = (uint32_t)s_k[k];
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(1024) void k_sum_u32(const uint32_t *__restrict__ v, uint32_t n,
                                                  unsigned long long *__restrict__ acc) {
    unsigned long long x = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x += v[i];
    x = block_sum(x);
    if (threadIdx.x == 0 && x) atomicAdd(acc, x);
}

// Out-degree sum of the marked nodes; clears the marks.
__global__ __launch_bounds__(1024) void k_marked_degree(uint32_t *__restrict__ mark,
                                                        const uint32_t *__restrict__ nsg_len, uint32_t n,
                                                        unsigned long long *__restrict__ acc) {
    unsigned long long x = 0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (mark[i]) {
            x += nsg_len[i];
            mark[i] = 0;
        }
    x = block_sum(x);
    if (threadIdx.x == 0 && x) atomicAdd(acc, x);
}

static unsigned reduce_grid(uint32_t n_waves) {
    const unsigned g = (n_waves + 15) / 16;
    return g < 2048 ? (g ? g : 1) : 2048;
}

// ---------------------------------------------------------------------------
// Host phases.
// ---------------------------------------------------------------------------
static void zero_u32(BuildState &b, DevBuf<uint32_t> &x, size_t n) { FK_HIP(hipMemsetAsync(x.p, 0, n * 4, b.s)); }

void nnd_alloc(BuildState &b) {
    const size_t N = b.N;
    b.pool.alloc(N * b.cap);
    b.pflg.alloc(N * b.cap);
    b.cnt.alloc(N);
    b.thr.alloc(N);
    b.M.alloc(N);
    b.radiusM.alloc(N);
    b.found.alloc(N);
    b.list_off.alloc(N + 1);
    b.n_new.alloc(N);
    b.pair_off.alloc(N + 1);
    b.pair_cnt.alloc(N);
    b.bcnt.alloc(N);
    b.boff.alloc(N + 1);
    b.fcnt_new.alloc(N);
    b.fcnt_old.alloc(N);
    b.rcnt_new.alloc(N);
    b.rcnt_old.alloc(N);
    b.roff_new.alloc(N + 1);
    b.roff_old.alloc(N + 1);
    b.acc.alloc(8);
    b.tally.ensure(FK_T_SLOTS);
    b.scal.alloc(16);
}

// q[c] = the row of node ids[c] (dpad floats): the control rows as queries.
__global__ void k_gather_rows(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                              const uint32_t *__restrict__ ids, uint32_t n, float *__restrict__ q) {
    const size_t total = (size_t)n * dpad;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const size_t c = e / dpad;
        q[e] = rows[(size_t)ids[c] * stride + (e - c * dpad)];
    }
}

// GenerateControl (src/kgraph.cpp:363-387): `controls` distinct random nodes
// and their exact MAXL nearest neighbours (self excluded, ties by id); per
// set, within the set, with the set's random stream.  Every set's controls
// are drawn first, uploaded and their rows gathered in one launch (a batch of
// many sets would otherwise pay a copy per control row); then one brute-force
// search per set over the set's rows.
void nnd_controls(BuildState &b) {
    const double t0 = now_sec();
    const uint32_t nc = (b.P.flags & FK_FLAG_NO_CONTROLS) ? 0 : b.P.controls;
    b.n_ctrl = nc * b.n_sets;
    b.ctrl_lo.assign(b.n_sets + 1, 0);
    for (uint32_t k = 0; k <= b.n_sets; ++k) b.ctrl_lo[k] = k * nc;
    if (b.n_ctrl == 0) return;
    b.ctrl_ids.alloc(b.n_ctrl);
    b.ctrl_gt.alloc((size_t)b.n_ctrl * b.P.MAXL);
    std::vector<uint32_t> ids(b.n_ctrl), local(b.n_ctrl);
    for (uint32_t k = 0; k < b.n_sets; ++k) {
        const uint32_t lo = b.set_lo[k], Ns = b.set_n(k);
        std::vector<uint32_t> idx(Ns);
        for (uint32_t i = 0; i < Ns; ++i) idx[i] = i;
        // partial Fisher-Yates with the set's host stream
        for (uint32_t i = 0; i < nc; ++i) {
            const uint32_t j = i + (uint32_t)(b.next_rand(k) % (Ns - i));
            std::swap(idx[i], idx[j]);
        }
        for (uint32_t c = 0; c < nc; ++c) {
            local[b.ctrl_lo[k] + c] = idx[c];
            ids[b.ctrl_lo[k] + c] = lo + idx[c];
        }
    }
    DevBuf<float> q((size_t)b.n_ctrl * b.ds->dpad);
    DevBuf<uint32_t> dself(b.n_ctrl);
    FK_HIP(hipMemcpyAsync(b.ctrl_ids.p, ids.data(), (size_t)b.n_ctrl * 4, hipMemcpyHostToDevice, b.s));
    FK_HIP(hipMemcpyAsync(dself.p, local.data(), (size_t)b.n_ctrl * 4, hipMemcpyHostToDevice, b.s));
    {
        const size_t total = (size_t)b.n_ctrl * b.ds->dpad;
        hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)std::min<size_t>((total + 255) / 256, 65536)), dim3(256), 0,
                           b.s, b.ds->rows.p, b.ds->stride, b.ds->dpad, b.ctrl_ids.p, b.n_ctrl, q.p);
        FK_LAUNCH_CHECK();
    }
    for (uint32_t k = 0; k < b.n_sets; ++k) {
        const uint32_t lo = b.set_lo[k], Ns = b.set_n(k);
        const uint32_t K = b.P.MAXL < Ns - 1 ? b.P.MAXL : Ns - 1;
        uint32_t *gt_out = b.ctrl_gt.p + (size_t)b.ctrl_lo[k] * b.P.MAXL;
        if (K < b.P.MAXL) FK_HIP(hipMemsetAsync(gt_out, 0xFF, (size_t)nc * b.P.MAXL * 4, b.s));
        DevBuf<uint32_t> gt((size_t)nc * K);
        hipEvent_t pe = b.prof_begin();
        bruteforce_knn(b.s, *b.ds, q.p + (size_t)b.ctrl_lo[k] * b.ds->dpad, b.ds->dpad, nc, Ns, K,
                       dself.p + b.ctrl_lo[k], gt.p, nullptr, lo);
        // algorithmic bytes: every row read once, the K ids written per control
        b.prof_end(pe, FK_PROF_CONTROLS, (double)Ns * 4.0 * b.ds->d + (double)nc * K * 4.0);
        FK_HIP(hipMemcpy2DAsync(gt_out, (size_t)b.P.MAXL * 4, gt.p, (size_t)K * 4, (size_t)K * 4, nc,
                                hipMemcpyDeviceToDevice, b.s));
        b.st->d_ctrl += (uint64_t)nc * (Ns - 1);
    }
    b.sync();
    b.st->t_ctrl += now_sec() - t0;
}

float nnd_evaluate(BuildState &b, uint32_t K, std::vector<float> *per_set) {
    if (per_set) per_set->assign(b.n_sets, 0.f);
    if (b.n_ctrl == 0) return 0.f;
    FK_CHECK_ARG(K <= FK_POOL_MAX * 2, "K too large for recall evaluation");
    DevBuf<uint32_t> found(b.n_ctrl);
    hipLaunchKernelGGL(k_recall, dim3(b.n_ctrl), dim3(256), 0, b.s, b.pool.p, b.cnt.p, b.cap, b.ctrl_ids.p,
                       b.ctrl_gt.p, b.P.MAXL, K, found.p);
    FK_LAUNCH_CHECK();
    std::vector<uint32_t> f(b.n_ctrl);
    FK_HIP(hipMemcpyAsync(f.data(), found.p, (size_t)b.n_ctrl * 4, hipMemcpyDeviceToHost, b.s));
    b.sync();
    // mean over controls of found/K (EvaluateRecall, src/kgraph.cpp:113-169),
    // per set and over all controls
    unsigned long long all = 0;
    for (uint32_t k = 0; k < b.n_sets; ++k) {
        unsigned long long fs = 0;
        for (uint32_t c = b.ctrl_lo[k]; c < b.ctrl_lo[k + 1]; ++c) fs += f[c];
        all += fs;
        const uint32_t ns = b.ctrl_lo[k + 1] - b.ctrl_lo[k];
        if (per_set && ns) (*per_set)[k] = (float)((double)fs / K / ns);
    }
    return (float)((double)all / K / b.n_ctrl);
}

// The stop rule's statistics of the current pools (one set): delta at K over
// all nodes, recall at K over n_ctrl controls d_ctrl with exact neighbours
// d_gt (gt_stride ids per control).  Same kernels and arithmetic as the
// NN-descent loop.
void nnd_stop_stats(BuildState &b, uint32_t K, uint32_t n_ctrl, const uint32_t *d_ctrl, const uint32_t *d_gt,
                    uint32_t gt_stride, float *delta, float *recall) {
    FK_CHECK_ARG(K >= 1 && K <= FK_POOL_MAX * 2, "K out of range for the stop statistics");
    FK_HIP(hipMemsetAsync(b.set_acc.p, 0, 8, b.s));
    hipLaunchKernelGGL(k_delta, dim3(reduce_grid(b.N)), dim3(1024), 0, b.s, b.pflg.p, b.cnt.p, b.N, b.cap, K,
                       nullptr, b.set_acc.p);
    FK_LAUNCH_CHECK();
    unsigned long long fl = 0;
    FK_HIP(hipMemcpyAsync(&fl, b.set_acc.p, 8, hipMemcpyDeviceToHost, b.s));
    std::vector<uint32_t> f(n_ctrl);
    if (n_ctrl) {
        DevBuf<uint32_t> found(n_ctrl);
        hipLaunchKernelGGL(k_recall, dim3(n_ctrl), dim3(256), 0, b.s, b.pool.p, b.cnt.p, b.cap, d_ctrl, d_gt,
                           gt_stride, K, found.p);
        FK_LAUNCH_CHECK();
        FK_HIP(hipMemcpyAsync(f.data(), found.p, (size_t)n_ctrl * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
    } else {
        b.sync();
    }
    *delta = (float)((double)fl / K / b.N);
    unsigned long long all = 0;
    for (uint32_t c = 0; c < n_ctrl; ++c) all += f[c];
    *recall = n_ctrl ? (float)((double)all / K / n_ctrl) : 0.f;
}

// Exclusive scan of n counts into out[0..n], out[n] = total; returns total.
static uint64_t scan_to_host(BuildState &b, const uint32_t *in, uint64_t *out, size_t n) {
    exclusive_scan_u32_to_u64(b.s, in, out, n, out + n, b.scan_tmp);
    uint64_t tot = 0;
    FK_HIP(hipMemcpyAsync(&tot, out + n, 8, hipMemcpyDeviceToHost, b.s));
    b.sync();
    return tot;
}

// Gathers v[order[p]] into out[p].
__global__ void k_gather_by_order(const uint32_t *__restrict__ v, const uint32_t *__restrict__ order, uint32_t n,
                                  uint32_t *__restrict__ out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) out[p] = v[order[p]];
}

// Scatters off[p] to boff[order[p]].
__global__ void k_scatter_by_order(const uint64_t *__restrict__ off, const uint32_t *__restrict__ order, uint32_t n,
                                   uint64_t *__restrict__ boff) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) boff[order[p]] = off[p];
}

// Bucket offsets of the join's candidates (boff[N] = total).  With a
// locality order the buckets are laid out in that order, so the buckets one
// node's list writes into (its neighbours') lie close together in memory
// instead of across the whole candidate buffer (address translation and L2
// reuse); every kernel still finds bucket t at boff[t].
static uint64_t bucket_offsets(BuildState &b) {
    const uint32_t N = b.N;
    if (!b.has_order) return scan_to_host(b, b.bcnt.p, b.boff.p, N);
    b.ord_cnt.ensure(N);
    b.ord_off.ensure((size_t)N + 1);
    const unsigned g = (N + 255) / 256;
    hipLaunchKernelGGL(k_gather_by_order, dim3(g), dim3(256), 0, b.s, b.bcnt.p, b.order.p, N, b.ord_cnt.p);
    FK_LAUNCH_CHECK();
    const uint64_t tot = scan_to_host(b, b.ord_cnt.p, b.ord_off.p, N);
    hipLaunchKernelGGL(k_scatter_by_order, dim3(g), dim3(256), 0, b.s, b.ord_off.p, b.order.p, N, b.boff.p);
    FK_LAUNCH_CHECK();
    FK_HIP(hipMemcpyAsync(b.boff.p + N, b.ord_off.p + N, 8, hipMemcpyDeviceToDevice, b.s));
    return tot;
}

static void nnd_join(BuildState &b) {
    const uint32_t N = b.N;
    const unsigned g256 = (N + 255) / 256;
    hipLaunchKernelGGL(k_pair_counts, dim3(g256), dim3(256), 0, b.s, b.list_off.p, b.n_new.p, N, b.pair_cnt.p);
    FK_LAUNCH_CHECK();
    const uint64_t pairs = scan_to_host(b, b.pair_cnt.p, b.pair_off.p, N);
    b.st->d_join += pairs;
    b.pdist.ensure(pairs);
    zero_u32(b, b.bcnt, N);
    FK_HIP(hipMemsetAsync(b.found.p, 0, N, b.s));
    uint64_t nlist = 0;
    FK_HIP(hipMemcpyAsync(&nlist, b.list_off.p + N, 8, hipMemcpyDeviceToHost, b.s));
    b.sync();
    b.resv.ensure(nlist);
    const uint32_t tcap = std::max(b.tmax, 1u);  // longest sample list of this iteration
    FK_CHECK_ARG(tcap <= FK_LIST_MAX, "sample list longer than the join supports");
    FK_CHECK_ARG(2ull * (b.ds->dpad + 4) * 4 <= FK_JOIN_LDS_ROWS_BYTES, "rows too wide for the join block");
    if (b.verbose) {
        // list-size profile: which blocks stage their rows in LDS
        std::vector<uint64_t> lo((size_t)N + 1);
        std::vector<uint32_t> nnv(N);
        FK_HIP(hipMemcpyAsync(lo.data(), b.list_off.p, lo.size() * 8, hipMemcpyDeviceToHost, b.s));
        FK_HIP(hipMemcpyAsync(nnv.data(), b.n_new.p, (size_t)N * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
        const uint32_t fit = FK_JOIN_LDS_ROWS_BYTES / ((b.ds->dpad + 4) * 4);
        uint64_t tmax = 0, nfit = 0, pfit = 0, ptot = 0, nsum = 0;
        for (uint32_t i = 0; i < N; ++i) {
            const uint64_t T = lo[i + 1] - lo[i], nn = nnv[i];
            const uint64_t pr = nn * (nn ? nn - 1 : 0) / 2 + nn * (T - nn);
            tmax = std::max(tmax, T);
            nsum += nn;
            ptot += pr;
            if (T <= fit) { ++nfit; pfit += pr; }
        }
        fprintf(stderr, "[fk] join lists: mean T %.1f, mean new %.1f, max T %llu; %.1f%% of nodes (%.1f%% of pairs) "
                        "stage rows in LDS (T <= %u)\n", (double)lo[N] / N, (double)nsum / N,
                (unsigned long long)tmax, 100.0 * nfit / N, 100.0 * pfit / std::max<uint64_t>(ptot, 1), fit);
    }
    const uint32_t *ord = b.has_order ? b.order.p : nullptr;
    const double d4 = 4.0 * b.ds->d;
    const bool gram = join_in_gram_form(b.P.flags, *b.ds);
    if (gram) b.tally.clear(b.s, FK_T_RECHECK);
    const bool screen = b.scr.ready && join_screen_lds(tcap) <= FK_JOIN_SCREEN_LDS_MAX;
    if (screen) screen_calibrate(b);
    hipEvent_t pe = b.prof_begin();
    if (gram) {
        launch_join_gram(b, tcap, b.tally.at(FK_T_RECHECK));
    } else if (screen) {
        launch_join_screen(b, tcap);
    } else {
        // lists of up to FK_JOIN_SMALL_ROWS rows first, staged in a small LDS
        // area (more workgroups per CU), then the longer ones with the full one
        const uint32_t row_bytes = (b.ds->dpad + 4) * 4;
        const uint32_t small_bytes = std::min<uint32_t>(FK_JOIN_SMALL_ROWS * row_bytes, FK_JOIN_LDS_ROWS_BYTES);
        const uint32_t small_t = small_bytes / row_bytes;
        const uint32_t tsm = std::min(tcap, small_t);
        hipLaunchKernelGGL(k_join_count, dim3(node_grid(b)), dim3(256), join_count_lds(tsm, small_bytes), b.s,
                           b.ds->rows.p, b.ds->stride, b.ds->dpad, b.lists.p, b.list_off.p, b.n_new.p, b.pair_off.p,
                           b.thr.p, b.pdist.p, b.resv.p, b.bcnt.p, N, ord, tsm, small_bytes, 0u, small_t);
        FK_LAUNCH_CHECK();
        if (tcap > small_t) {
            hipLaunchKernelGGL(k_join_count, dim3(node_grid(b)), dim3(256),
                               join_count_lds(tcap, FK_JOIN_LDS_ROWS_BYTES), b.s, b.ds->rows.p, b.ds->stride,
                               b.ds->dpad, b.lists.p, b.list_off.p, b.n_new.p, b.pair_off.p, b.thr.p, b.pdist.p,
                               b.resv.p, b.bcnt.p, N, ord, tcap, (uint32_t)FK_JOIN_LDS_ROWS_BYTES, small_t,
                               0xFFFFFFFFu);
            FK_LAUNCH_CHECK();
        }
    }
    // sampled rows and ids gathered once per join, one float per pair written
    b.prof_end(pe, FK_PROF_JOIN, (double)nlist * (d4 + 4.0) + 4.0 * (double)pairs);
    const uint64_t cands = bucket_offsets(b);
    if (screen) {
        screen_collect(b);
        b.st->screen_cands += cands;
    }
    if (gram) {
        const uint64_t rc = b.tally.read1(b.s, FK_T_RECHECK);
        b.st->d_gram += pairs;
        b.st->d_recheck += rc;
    }
    b.bkey.ensure(cands);
    b.bsrc.ensure(cands);
    double pool_entries = 0;
    if (b.prof) {
        FK_HIP(hipMemsetAsync(b.acc.p + 6, 0, 8, b.s));
        hipLaunchKernelGGL(k_sum_u32, dim3(reduce_grid(N / 64 + 1)), dim3(1024), 0, b.s, b.cnt.p, N, b.acc.p + 6);
        unsigned long long pe_cnt = 0;
        FK_HIP(hipMemcpyAsync(&pe_cnt, b.acc.p + 6, 8, hipMemcpyDeviceToHost, b.s));
        b.sync();
        pool_entries = (double)pe_cnt;
    }
    pe = b.prof_begin();
    hipLaunchKernelGGL(k_join_fill, dim3(node_grid(b)), dim3(256), join_fill_lds(tcap), b.s, b.lists.p, b.list_off.p,
                       b.n_new.p, b.pair_off.p, b.thr.p, b.pdist.p, b.resv.p, b.boff.p, b.bkey.p, b.bsrc.p, N, ord,
                       tcap);
    FK_LAUNCH_CHECK();
    // pair distances and list ids read, (key, source) bucket entries written
    b.prof_end(pe, FK_PROF_JOIN_FILL, 4.0 * (double)pairs + 4.0 * (double)nlist + 12.0 * (double)cands);
    pe = b.prof_begin();
    hipLaunchKernelGGL(k_nnd_merge, dim3(N), dim3(256), nnd_merge_lds(b.cap), b.s, b.cap, b.P.L, b.P.K, b.pool.p,
                       b.pflg.p, b.cnt.p,
                       b.thr.p, b.bcnt.p, b.boff.p, b.bkey.p, b.bsrc.p, b.found.p);
    FK_LAUNCH_CHECK();
    // bucket entries read once; each pool read and rewritten (8-byte key +
    // 1-byte flags per entry), before and after sizes bounded by L
    b.prof_end(pe, FK_PROF_MERGE, 12.0 * (double)cands + 18.0 * pool_entries);
    b.prof_flush();
    if (b.verbose) fprintf(stderr, "[fk] join: pairs %llu candidates %llu\n", (unsigned long long)pairs,
                           (unsigned long long)cands);
}

static void nnd_update(BuildState &b, uint32_t it) {
    const uint32_t N = b.N;
    const unsigned g256 = (N + 255) / 256;
    const unsigned gw = (N + 3) / 4;  // 4 waves per 256-thread block
    hipEvent_t pe = b.prof_begin();
    hipLaunchKernelGGL(k_upd_m, dim3(g256), dim3(256), 0, b.s, b.pool.p, b.pflg.p, b.cnt.p, b.found.p, N, b.cap,
                       b.P.S, b.dev_set_of(), b.n_sets > 1 ? b.d_set_active.p : nullptr, b.M.p, b.radiusM.p);
    FK_LAUNCH_CHECK();
    zero_u32(b, b.rcnt_new, N);
    zero_u32(b, b.rcnt_old, N);
    hipLaunchKernelGGL(k_upd_count, dim3(gw), dim3(256), 0, b.s, b.pool.p, b.pflg.p, N, b.cap, b.M.p, b.radiusM.p,
                       b.fcnt_new.p, b.fcnt_old.p, b.rcnt_new.p, b.rcnt_old.p);
    FK_LAUNCH_CHECK();
    const uint64_t tn = scan_to_host(b, b.rcnt_new.p, b.roff_new.p, N);
    const uint64_t to = scan_to_host(b, b.rcnt_old.p, b.roff_old.p, N);
    b.rbuf_new.ensure(tn);
    b.rbuf_old.ensure(to);
    uint32_t *lmax = reinterpret_cast<uint32_t *>(b.scal.p + 1);
    FK_HIP(hipMemsetAsync(lmax, 0, 4, b.s));
    hipLaunchKernelGGL(k_list_sizes, dim3(g256), dim3(256), 0, b.s, b.fcnt_new.p, b.fcnt_old.p, b.rcnt_new.p,
                       b.rcnt_old.p, N, b.P.R, b.n_new.p, b.pair_cnt.p, lmax);
    FK_LAUNCH_CHECK();
    uint32_t tmax = 0;
    FK_HIP(hipMemcpyAsync(&tmax, lmax, 4, hipMemcpyDeviceToHost, b.s));
    const uint64_t nlist = scan_to_host(b, b.pair_cnt.p, b.list_off.p, N);  // synchronises
    b.tmax = tmax;
    b.lists.ensure(nlist);
    // reverse cursors reuse the bucket and pair counters
    zero_u32(b, b.bcnt, N);
    zero_u32(b, b.pair_cnt, N);
    hipLaunchKernelGGL(k_upd_fill, dim3(gw), dim3(256), 0, b.s, b.pool.p, b.pflg.p, N, b.cap, b.M.p, b.radiusM.p,
                       b.list_off.p, b.n_new.p, b.fcnt_new.p, b.lists.p, b.roff_new.p, b.roff_old.p, b.bcnt.p,
                       b.pair_cnt.p, b.rbuf_new.p, b.rbuf_old.p);
    FK_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_rev_select, dim3((N + FK_REV_NPB - 1) / FK_REV_NPB), dim3(256), 0, b.s, N, b.P.R,
                       (uint64_t)b.P.seed, it, b.rcnt_new.p,
                       b.rcnt_old.p, b.roff_new.p, b.roff_old.p, b.rbuf_new.p, b.rbuf_old.p, b.list_off.p,
                       b.n_new.p, b.fcnt_new.p, b.fcnt_old.p, b.dev_set_lo(), b.dev_set_of(), b.lists.p);
    FK_LAUNCH_CHECK();
    // the first M entries of every pool read twice (keys + flags), lists and
    // reverse buckets written once
    b.prof_end(pe, FK_PROF_UPDATE, 18.0 * (double)nlist + 8.0 * (double)(tn + to) + 4.0 * (double)nlist);
    b.sync();
    b.prof_flush();
}

void nnd_run(BuildState &b) {
    const double t0 = now_sec();
    const uint32_t N = b.N;
    const BuildParams &P = b.P;
    FK_CHECK_ARG(P.S >= 1 && P.S <= 128, "S must be in [1, 128]");
    FK_CHECK_ARG(P.S <= P.L, "S must not exceed L");
    FK_CHECK_ARG(P.MAXL <= FK_POOL_MAX, "max(L, K, search_L, search_K) above 1024 is not supported");
    nnd_alloc(b);
    if (P.flags & FK_FLAG_GRAM) {
        gram_prepare(b.s, *b.ds, P.flags);
        b.st->gram_mode = (uint32_t)b.ds->gram_fmt;
    }
    // init (iteration 1 joins 2S random ids per node, old lists empty)
    b.lists.ensure((size_t)N * 2 * P.S);
    b.tmax = 2 * P.S;
    {
        std::vector<uint64_t> lo((size_t)N + 1);
        for (size_t i = 0; i <= N; ++i) lo[i] = i * 2 * P.S;
        FK_HIP(hipMemcpyAsync(b.list_off.p, lo.data(), lo.size() * 8, hipMemcpyHostToDevice, b.s));
        std::vector<uint32_t> nn(N, 2 * P.S);
        FK_HIP(hipMemcpyAsync(b.n_new.p, nn.data(), (size_t)N * 4, hipMemcpyHostToDevice, b.s));
        b.sync();
    }
    hipEvent_t pe0 = b.prof_begin();
    hipLaunchKernelGGL(k_nnd_init, dim3(N), dim3(64), 0, b.s, b.ds->rows.p, b.ds->stride, b.ds->dpad, N, P.S, b.cap,
                       P.L, (uint64_t)P.seed, b.dev_set_lo(), b.dev_set_of(), b.pool.p, b.pflg.p, b.cnt.p, b.M.p,
                       b.thr.p, b.lists.p);
    FK_LAUNCH_CHECK();
    b.prof_end(pe0, FK_PROF_INIT, (double)N * ((P.S + 1) * (4.0 * b.ds->d) + P.S * 9.0 + 2 * P.S * 4.0));
    b.st->d_init += (uint64_t)N * P.S;
    const double tk0 = now_sec();
    for (uint32_t it = 0; it < P.iterations; ++it) {
        b.st->nnd_iterations++;
        nnd_join(b);
        // the screen learns from the first join's exact pair distances
        if (it == 0 && (P.flags & FK_FLAG_SCREEN) && P.iterations > 1) screen_train(b);
        // statistics and the stop rule, per set (EvaluateDelta / EvaluateRecall)
        FK_HIP(hipMemsetAsync(b.set_acc.p, 0, 8 * (size_t)b.n_sets, b.s));
        hipLaunchKernelGGL(k_delta, dim3(reduce_grid(N)), dim3(1024), 0, b.s, b.pflg.p, b.cnt.p, N, b.cap, P.K,
                           b.dev_set_of(), b.set_acc.p);
        FK_LAUNCH_CHECK();
        std::vector<unsigned long long> fl(b.n_sets);
        FK_HIP(hipMemcpyAsync(fl.data(), b.set_acc.p, 8 * (size_t)b.n_sets, hipMemcpyDeviceToHost, b.s));
        b.sync();
        std::vector<float> rec_set;
        const float recall = nnd_evaluate(b, P.K, &rec_set);
        unsigned long long fl_all = 0;
        uint64_t n_all = 0;
        uint32_t n_active = 0;
        float delta = 0.f;
        for (uint32_t k = 0; k < b.n_sets; ++k) {
            if (!b.set_active[k]) continue;
            fl_all += fl[k];
            n_all += b.set_n(k);
            const float dk = (float)((double)fl[k] / P.K / b.set_n(k));
            uint32_t stop = 0;
            if (dk <= P.delta) stop = 1;
            else if (b.ctrl_lo[k + 1] > b.ctrl_lo[k] && rec_set[k] >= P.recall) stop = 2;
            if (b.n_sets == 1) delta = dk;
            if (stop || it + 1 >= P.iterations) {
                b.set_active[k] = 0;
                if (b.n_sets == 1) b.st->nnd_stop = stop;
            } else {
                ++n_active;
            }
        }
        if (b.n_sets > 1) delta = (float)((double)fl_all / P.K / (double)n_all);
        b.st->nnd_delta = delta;
        b.st->nnd_recall = recall;
        if (b.st->n_kgraph < FK_MAX_LOOPS + 64) b.st->kgraph_time[b.st->n_kgraph++] = now_sec() - b.t0;
        if (b.st->n_recall < FK_MAX_LOOPS + 64) b.st->recall[b.st->n_recall++] = recall;
        if (b.verbose)
            fprintf(stderr, "[fk] iteration: %u recall: %g delta: %g time: %g (%u of %u sets go on)\n", it + 1,
                    recall, delta, now_sec() - tk0, n_active, b.n_sets);
        if (n_active == 0) break;
        if (b.n_sets > 1)
            FK_HIP(hipMemcpyAsync(b.d_set_active.p, b.set_active.data(), b.n_sets, hipMemcpyHostToDevice, b.s));
        // from the third join on, the pools' heads are close enough for a
        // useful locality order; renewed before every other join (measured
        // on the 1M x 128 NSG build: as fast as every join, 4 ms less work)
        if (it >= 1 && (it - 1) % 2 == 0) locality_order(b);
        nnd_update(b, it);
    }
    b.sync();
    b.st->t_nnd += now_sec() - t0;
}

// ---------------------------------------------------------------------------
// BridgeView (P2K batched search, src/kgraph.cpp:2478-2663), pull form: node
// i expands the flagged entries q of its window [0, min(window, L_i)) and
// joins q's out-list into its own pool, which is what the reference's
// reverse lists rnn_new deliver to i.  A pair (q -> t) is skipped when t == i
// or when both the pool entry q and the out-list entry t are old (:2640-2642).
// ---------------------------------------------------------------------------
__global__ void k_bv_init(uint8_t *__restrict__ pflg, const uint32_t *__restrict__ cnt, uint32_t N, uint32_t cap,
                          uint32_t cur_loop, uint32_t pre_L, uint32_t cur_L) {
    const uint32_t n = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (n >= N) return;
    uint8_t *pf = pflg + (uint64_t)n * cap;
    const uint32_t L = cnt[n];
    for (uint32_t j = lane_id(); j < L; j += 64) {
        const uint8_t f = pf[j];
        uint8_t isnew = 0;
        if (cur_loop != 0 && j >= pre_L && j < cur_L) isnew = (f & FK_FLG_NEW) ? FK_FLG_ISNEW : 0;
        pf[j] = (uint8_t)(FK_FLG_NEW | isnew);
    }
}

// Flagged window entries, capped at S per node and all; with sets, into
// acc[2 set] and acc[2 set + 1] (one atomic pair per node).
__global__ __launch_bounds__(1024) void k_bv_count(const uint8_t *__restrict__ pflg, const uint32_t *__restrict__ cnt,
                                                   uint32_t N, uint32_t cap, uint32_t W, uint32_t S,
                                                   const uint32_t *__restrict__ set_of,
                                                   unsigned long long *__restrict__ acc) {
    const uint32_t wpb = blockDim.x >> 6;
    if (set_of) {  // uniform: contiguous runs per wave, as in k_delta
        const uint32_t nw = gridDim.x * wpb, chunk = (N + nw - 1) / nw;
        const uint32_t lo = (blockIdx.x * wpb + (threadIdx.x >> 6)) * chunk;
        const uint32_t hi = lo + chunk < N ? lo + chunk : N;
        uint32_t cur = 0xFFFFFFFFu;
        unsigned long long rc = 0, ra = 0;
        for (uint32_t n = lo; n < hi; ++n) {
            const uint32_t lim = cnt[n] < W ? cnt[n] : W;
            const uint32_t f = wave_sum(lane_new_flags(pflg + (uint64_t)n * cap, lim, cap));
            const uint32_t k = set_of[n];
            if (k != cur) {
                if (lane_id() == 0 && ra) {
                    atomicAdd(&acc[2 * cur], rc);
                    atomicAdd(&acc[2 * cur + 1], ra);
                }
                cur = k;
                rc = ra = 0;
            }
            rc += f < S ? f : S;
            ra += f;
        }
        if (lane_id() == 0 && ra) {
            atomicAdd(&acc[2 * cur], rc);
            atomicAdd(&acc[2 * cur + 1], ra);
        }
        return;
    }
    unsigned long long capped = 0, all = 0;
    for (uint32_t n = blockIdx.x * wpb + (threadIdx.x >> 6); n < N; n += gridDim.x * wpb) {
        const uint8_t *pf = pflg + (uint64_t)n * cap;
        const uint32_t lim = cnt[n] < W ? cnt[n] : W;
        uint32_t f = lane_new_flags(pf, lim, cap);
        f = wave_sum(f);
        if (lane_id() == 0) {
            capped += f < S ? f : S;
            all += f;
        }
    }
    capped = block_sum(capped);
    all = block_sum(all);
    if (threadIdx.x == 0) {
        if (capped) atomicAdd(acc, capped);
        if (all) atomicAdd(acc + 1, all);
    }
}

// One wavefront per node, FK_BVW_WAVES nodes per workgroup and no block
// barriers: a node's round is a chain of dependent gathers (pool -> the
// selected entries' out-lists -> their rows), so many nodes in flight per
// CU is what hides the latency: the out-list entries of up to 256 pairs are
// read with independent loads, and only the search window of the pool is
// read before the selection (the rest only by nodes that select something).
// Per wave, in LDS:
//   u64 p[cap]   the pool (sorted keys), merged in place
//   u64 c[CH]    candidate keys (distance <= the pool's last at round start)
//   u32 sel[W], sdeg[W + 1], todo[FK_BVW_TODO], hs[bvw_hsize(cap)]
//   u8 pf[cap], seln[W]
// hs is a lossy open-addressing set of ids (hs_seen_or_add): the pool's at
// round start, then every pair target whose distance this round computes.
// A target found in it
// is counted (search_cc counts every pair the reference evaluates) but not
// computed again: its key equals the pool entry's or the earlier
// candidate's (the distance is symmetric and computed in one canonical
// order), which the merge drops, and an id pushed out of the pool this
// round has at least cap better keys ahead of it for the rest of the round.
#define FK_BVW_WAVES 2
#define FK_BVW_CH 128
#ifndef FK_BVW_INFLIGHT
#define FK_BVW_INFLIGHT 2  // rows per octet in flight in the distance step
#endif
#ifndef FK_BVW_INFLIGHT8
#define FK_BVW_INFLIGHT8 4  // the same for byte rows (16 bytes per lane and row)
#endif
#ifndef FK_BVW_WPE
#define FK_BVW_WPE 4       // waves per SIMD the FP32 row kinds are compiled for
#endif
#ifndef FK_BVW_WPE32
#define FK_BVW_WPE32 3     // the same at width 256 (its chains take more registers)
#endif
#define FK_BVW_TODO 320    // one read of 64 FK_BVW_PREFETCH pairs + a remainder below 64
#ifndef FK_BVW_PREFETCH
#define FK_BVW_PREFETCH 4  // chunks of 64 out-list entries read at once
#endif
#define FK_BVW_LOADS 8     // chunks of 64 pool entries read at once

// Slots of the id set: a power of two >= cap.  With the short probe runs a
// full set only loses ids (computed again); measured on the 1M x 128 NSG
// build, this size and two probes beat larger sets and longer runs (LDS per
// wave sets the occupancy).
__host__ __device__ inline uint32_t bvw_hsize(uint32_t cap) {
    uint32_t h = 64;
    while (h < cap) h <<= 1;
    return h;
}

__host__ __device__ inline size_t bvw_wave_bytes(uint32_t cap, uint32_t W) {
    const size_t b = 8 * ((size_t)cap + FK_BVW_CH) +
                     4 * ((size_t)W + (W + 1) + FK_BVW_TODO + bvw_hsize(cap)) + (size_t)cap + W;
    return (b + 15) & ~(size_t)15;
}

#define FK_HS_EMPTY 0xFFFFFFFFu

__device__ __forceinline__ uint32_t hs_slot(uint32_t id, uint32_t mask) { return ((id * 0x9E3779B1u) >> 15) & mask; }

// The id set is lossy: plain LDS reads and stores, at most FK_HS_PROBES
// slots probed from the id's home slot.  True when id is in the set;
// otherwise it is stored in the first empty slot of its probe run, if any.
// An id reported present was stored, so a skip is always right; an id lost
// (no empty slot in its run, or overwritten by a lane of the same
// instruction storing another id in the same slot) is only computed again,
// and its key is dropped at the merge.
#define FK_HS_PROBES 2
__device__ __forceinline__ bool hs_seen_or_add(uint32_t *hs, uint32_t mask, uint32_t id) {
    uint32_t s = hs_slot(id, mask);
    for (uint32_t q = 0; q < FK_HS_PROBES; ++q, s = (s + 1) & mask) {
        const uint32_t v = hs[s];
        if (v == id) return true;
        if (v == FK_HS_EMPTY) {
            hs[s] = id;
            return false;
        }
    }
    return false;
}

// Exclusive prefix sum of v over the wave; the wave total in every lane.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, uint32_t &total) {
    const uint32_t lane = lane_id();
    uint32_t x = v;
    for (uint32_t o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sorts c[0, n) (n a power of two) ascending; one wave.
__device__ void wave_bitonic_sort(uint64_t *c, uint32_t n) {
    const uint32_t lane = lane_id();
    for (uint32_t k = 2; k <= n; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = lane; i < n; i += 64) {
                const uint32_t ixj = i ^ j;
                if (ixj > i) {
                    const uint64_t a = c[i], b2 = c[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b2) == up) { c[i] = b2; c[ixj] = a; }
                }
            }
            wave_lds_sync();
        }
    }
}

// Merges the nc candidate keys of c into the pool p / pf (np entries, at
// most cap kept) in place and returns the new size.  One wave.
// `first` receives the smallest pool position written (cap when none); the
// entries below it are unchanged.
__device__ uint32_t bvw_merge(uint64_t *p, uint8_t *pf, uint32_t np, uint64_t *c, uint32_t nc, uint32_t cap,
                              uint32_t &inserted, uint32_t &first) {
    const uint32_t lane = lane_id();
    const uint32_t n2 = next_pow2(nc < 2 ? 2 : nc);
    for (uint32_t x = nc + lane; x < n2; x += 64) c[x] = FK_KEY_EMPTY;
    wave_lds_sync();
    if (n2 <= 128) {
        wave_sort128(c, n2);
        wave_lds_sync();
    } else {
        wave_bitonic_sort(c, n2);
    }
    // keep the first of equal keys that is not in the pool; compacted in
    // place (chunks in order, every lane reads before any lane writes)
    uint32_t nk = 0;
    for (uint32_t x0 = 0; x0 < nc; x0 += 64) {
        const uint32_t x = x0 + lane;
        uint64_t key = FK_KEY_EMPTY;
        bool keep = false;
        if (x < nc) {
            key = c[x];
            keep = (x == 0 || c[x - 1] != key) && !sorted_contains(p, np, key);
        }
        const uint64_t m = wave_ballot(keep);
        wave_lds_sync();
        if (keep) c[nk + lanes_below(m)] = key;
        nk += (uint32_t)__builtin_popcountll(m);
        wave_lds_sync();
    }
    inserted = nk;
    first = cap;
    if (nk == 0) return np;
    // candidate positions against the pool before it moves
    uint32_t cpos[FK_BVW_CH / 64];
#pragma unroll
    for (uint32_t r = 0; r < FK_BVW_CH / 64; ++r) {
        const uint32_t x = r * 64 + lane;
        cpos[r] = x < nk ? x + lower_bound_cnt(p, np, c[x]) : 0xFFFFFFFFu;
    }
    // the smallest candidate (lane 0's first) lands lowest; nothing below it moves
    first = (uint32_t)__shfl((int)cpos[0], 0, 64);
    if (first > cap) first = cap;
    // pool entries move up to j + (#candidates below), last chunk first,
    // each chunk read before it is written
    for (uint32_t ch = (np + 63) / 64; ch-- > first / 64;) {
        const uint32_t j = ch * 64 + lane;
        uint64_t key = 0;
        uint8_t f = 0;
        uint32_t pos = 0xFFFFFFFFu;
        if (j < np) {
            key = p[j];
            f = pf[j];
            pos = j + lower_bound_cnt(c, nk, key);
        }
        wave_lds_sync();
        if (pos < cap) { p[pos] = key; pf[pos] = f; }
        wave_lds_sync();
    }
#pragma unroll
    for (uint32_t r = 0; r < FK_BVW_CH / 64; ++r) {
        const uint32_t x = r * 64 + lane;
        if (cpos[r] < cap) { p[cpos[r]] = c[x]; pf[cpos[r]] = (uint8_t)(FK_FLG_NEW | FK_FLG_ISNEW); }
    }
    wave_lds_sync();
    return (np + nk) < cap ? (np + nk) : cap;
}

// The same merge for at most 64 candidates with a short chain of dependent
// LDS accesses (what bounds a wave that merges a few candidates): the
// candidates are sorted in registers (one per lane), one search of the pool
// per candidate serves both the de-duplication and its final position, and
// the pool entries that move are all read, with their new positions, before
// any is written (FK_BVW_MOVE chunks of 64 at a time, highest first).
#define FK_BVW_SMALL 64
#define FK_BVW_MOVE 4
__device__ uint32_t bvw_merge_small(uint64_t *p, uint8_t *pf, uint32_t np, uint64_t *c, uint32_t nc, uint32_t cap,
                                    uint32_t &inserted, uint32_t &first) {
    const uint32_t lane = lane_id();
    uint64_t key = lane < nc ? c[lane] : FK_KEY_EMPTY;
    key = wave_sort64(key);
    const uint64_t prev = __shfl_up(key, 1, 64);
    const uint32_t pos = key != FK_KEY_EMPTY ? lower_bound_cnt(p, np, key) : np;
    const bool keep = key != FK_KEY_EMPTY && (lane == 0 || prev != key) && !(pos < np && p[pos] == key);
    const uint64_t m = wave_ballot(keep);
    const uint32_t nk = (uint32_t)__builtin_popcountll(m);
    inserted = nk;
    first = cap;
    if (nk == 0) return np;
    const uint32_t cpos = lanes_below(m) + pos;  // final position of a kept candidate
    first = (uint32_t)__shfl((int)cpos, (int)__builtin_ctzll(m), 64);
    if (first > cap) first = cap;
    wave_lds_sync();  // every lane has read its c[lane]
    if (keep) c[lanes_below(m)] = key;
    wave_lds_sync();
    const uint32_t ch_lo = first / 64, ch_hi = (np + 63) / 64;
    for (uint32_t g1 = ch_hi; g1 > ch_lo;) {
        const uint32_t g0 = (g1 - ch_lo) > FK_BVW_MOVE ? g1 - FK_BVW_MOVE : ch_lo;
        uint64_t kk[FK_BVW_MOVE];
        uint8_t ff[FK_BVW_MOVE];
        uint32_t pp[FK_BVW_MOVE];
#pragma unroll
        for (uint32_t r = 0; r < FK_BVW_MOVE; ++r) {
            const uint32_t j = (g0 + r) * 64 + lane;
            pp[r] = 0xFFFFFFFFu;
            if (g0 + r < g1 && j < np) {
                kk[r] = p[j];
                ff[r] = pf[j];
                pp[r] = j + lower_bound_cnt(c, nk, kk[r]);
            }
        }
        wave_lds_sync();
#pragma unroll
        for (uint32_t r = 0; r < FK_BVW_MOVE; ++r)
            if (pp[r] < cap) { p[pp[r]] = kk[r]; pf[pp[r]] = ff[r]; }
        wave_lds_sync();
        g1 = g0;
    }
    if (keep && cpos < cap) { p[cpos] = key; pf[cpos] = (uint8_t)(FK_FLG_NEW | FK_FLG_ISNEW); }
    wave_lds_sync();
    return (np + nk) < cap ? (np + nk) : cap;
}

// Copies the pool entries [a, b) of one node (keys and flag bytes) from
// global memory to LDS.  Every load of a batch of 64 FK_BVW_LOADS entries is
// issued before any LDS store, unconditionally (lanes past b read entry
// b - 1 again) and into whole registers, so a batch costs one memory round
// trip.  One wave; a < b or nothing happens.
__device__ __forceinline__ void bvw_load_pool(uint64_t *p, uint8_t *pf, const uint64_t *__restrict__ pr,
                                              const uint8_t *__restrict__ pfr, uint32_t a, uint32_t b) {
    const uint32_t lane = lane_id();
    for (uint32_t j0 = a; j0 < b; j0 += 64 * FK_BVW_LOADS) {
        uint64_t kk[FK_BVW_LOADS];
        uint32_t ff[FK_BVW_LOADS];
#pragma unroll
        for (uint32_t r = 0; r < FK_BVW_LOADS; ++r) {
            const uint32_t j = j0 + r * 64 + lane;
            const uint32_t jj = j < b ? j : b - 1;
            kk[r] = pr[jj];
            ff[r] = pfr[jj];
        }
#pragma unroll
        for (uint32_t r = 0; r < FK_BVW_LOADS; ++r) {
            const uint32_t j = j0 + r * 64 + lane;
            if (j < b) { p[j] = kk[r]; pf[j] = (uint8_t)ff[r]; }
        }
    }
}

// Lane k's part of a row's 16-bit Gram operands (values k NB .. k NB + NB -
// 1 of rows16, as NB / 2 half pairs), for the BridgeView bound pass; NB =
// 12, 16 or 32 (16-byte pieces, or 8-byte ones at 12).
template <int NB>
struct HalfSlice {
    static constexpr int P = NB / 2;
    uint32_t v[P];
    __device__ __forceinline__ void load(const uint16_t *p) {
        if constexpr (P % 4 == 0) {
#pragma unroll
            for (int j = 0; j < P / 4; ++j) {
                const uint4 q = reinterpret_cast<const uint4 *>(p)[j];
                v[4 * j] = q.x; v[4 * j + 1] = q.y; v[4 * j + 2] = q.z; v[4 * j + 3] = q.w;
            }
        } else {
#pragma unroll
            for (int j = 0; j < P / 2; ++j) {
                const uint2 q = reinterpret_cast<const uint2 *>(p)[j];
                v[2 * j] = q.x; v[2 * j + 1] = q.y;
            }
        }
    }
    // this lane's part of the dot product of two FP16 rows (products exact,
    // FP32 sums)
    __device__ __forceinline__ float dot(const HalfSlice &o) const {
        typedef _Float16 h2 __attribute__((ext_vector_type(2)));
        float s = 0.f;
#pragma unroll
        for (int j = 0; j < P; ++j)
            s = __builtin_amdgcn_fdot2(__builtin_bit_cast(h2, v[j]), __builtin_bit_cast(h2, o.v[j]), s, false);
        return s;
    }
};

// The sum of an octet's 8 values in every lane of the octet (any order: a
// bound's dot product, not a canonical distance).
__device__ __forceinline__ float octet_sum_any(float s) {
    s += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0xB1, 0xF, 0xF, true));   // quad_perm [1,0,3,2]
    s += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0x4E, 0xF, 0xF, true));   // quad_perm [2,3,0,1]
    s += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(s), 0x141, 0xF, 0xF, true));  // row_half_mirror
    return s;
}

// The sum of an octet's 8 32-bit integers (exact), in lane 0 of the octet
// at least (every sum here is below 2^24 in magnitude), as a float.
__device__ __forceinline__ float octet_sum_i32(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // row_half_mirror
    return (float)v;
}
// Lane 6's 32-bit value, as a float, in lane 0 of each octet (row_shl:6).
__device__ __forceinline__ float octet_from6(uint32_t v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, (int)v, 0x106, 0xF, 0xF, false));
}

#ifndef FK_BVW_INFLIGHT8Q
#define FK_BVW_INFLIGHT8Q 8  // 8-bit rows per octet in flight in the bound pass
#endif
#ifndef FK_BVW_INFLIGHT8Q32
#define FK_BVW_INFLIGHT8Q32 4  // the same at width 256 (three pieces per lane and row)
#endif
// Relative error of |x^ - y^|^2 evaluated in FP32 from the stored scales and
// |x^|^2 (each rounded once) and the exact integer dot product, as a share
// of |x^|^2 + |y^|^2: at most about 10 roundings, here 16
#define FK_Q8_FREL 0.00000095367431640625f  // 2^-20

#ifndef FK_BVW_INFLIGHT16
#define FK_BVW_INFLIGHT16 4  // rows per octet in flight in the bound pass, width 96
#endif
#ifndef FK_BVW_INFLIGHT16W
#define FK_BVW_INFLIGHT16W 4  // the same from width 128 up
#endif
// 1 +- 2^-20, 1 +- 2^-18 (exact in FP32): margins of the bound pass's test
#define FK_BVW_UP20 1.00000095367431640625f
#define FK_BVW_UP18 1.000003814697265625f
#define FK_BVW_DN18 0.999996185302734375f

// MODE (RowKind): 0 canonical FP32; 1 exact-form rows (fk_gram.hip),
// distances as |a|^2 + |b|^2 - 2 a.b; 2 the same from byte rows.
//
// Bound pass (MODE 0, NB 12 / 16 / 32, r16 set: FK_FLAG_GRAM on FP16-form
// rows): when the node's pool is full, every target first gets the 16-bit
// Gram-form value d~ of its distance (dot product of the FP16 operand rows)
// and is dropped without an FP32 distance when the rigorous lower bound
// gram_lower(d~) (the Gram prune's error model) exceeds the pool's last
// distance thr: its canonical distance is then above thr and the FP32 test
// dd <= thr would reject it.  The test is taken without a square root,
// d~ - f > (e + T)^2 with T = sqrt(thr / (1 - gam)) (T and e rounded up by
// 2^-20, each side moved by 2^-18, more than the rounding of these
// operations).  A target's |y~|^2 is summed from the 16-bit row the pass
// reads anyway (at most 35 FP32 roundings of non-negative terms, inside the
// norm share of f_rel), and its |y| in e is bounded by the largest row norm
// (gb.ymax; FP16 forms are chosen for unit-normalised rows), so the pass
// reads one row per target and no norm.  The survivors go through the FP32
// step unchanged, so the candidates and the merged pools are the same.  G:
// the kernel with the pass (the others are compiled without it).
template <int NB, int MODE, int G>
__global__ __launch_bounds__(64 * FK_BVW_WAVES) __attribute__((amdgpu_waves_per_eu(MODE == 2 ? 4 : (NB == 32 ? FK_BVW_WPE32 : FK_BVW_WPE)))) void k_bv_round(
    const void *__restrict__ chains_v, const float2 *__restrict__ nrm, uint32_t dpad, uint32_t N, uint32_t cap,
    uint32_t W, uint32_t S_cap_all, const uint32_t *__restrict__ set_of, const uint32_t *__restrict__ set_scap,
    uint64_t *__restrict__ pool, uint8_t *__restrict__ pflg, uint32_t *__restrict__ cnt,
    const uint32_t *__restrict__ nsg_id, const uint32_t *__restrict__ nsg_len, const uint8_t *__restrict__ nsg_new,
    uint32_t ncap, unsigned long long *__restrict__ cc_acc, unsigned long long *__restrict__ comp_acc,
    const uint32_t *__restrict__ order, uint32_t *__restrict__ mark, uint32_t wave_bytes, uint32_t wpb,
    const uint16_t *__restrict__ r16, const float2 *__restrict__ gnrm, GramBound gb,
    const uint4 *__restrict__ q8, unsigned long long *__restrict__ rej_acc) {
    using E = typename RowKind<NB, MODE>::E;
    using Chain = typename RowKind<NB, MODE>::Chain;
    constexpr bool DOT = RowKind<NB, MODE>::dot;
    constexpr uint32_t INF = MODE == 2 ? FK_BVW_INFLIGHT8 : FK_BVW_INFLIGHT;
    const E *chains = reinterpret_cast<const E *>(chains_v);
    extern __shared__ __align__(16) uint8_t smem[];
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const uint32_t per_xcd = ((N + wpb - 1) / wpb + FK_NUM_XCD - 1) / FK_NUM_XCD;
    const uint32_t pos = (order ? xcd_position(per_xcd) : blockIdx.x) * wpb + w;
    if (pos >= N) return;
    const uint32_t i = order ? order[pos] : pos;
    // a set whose rounds are over expands nothing (S_cap 0)
    const uint32_t S_cap = set_of ? set_scap[set_of[i]] : S_cap_all;
    uint8_t *wb = smem + (size_t)w * wave_bytes;
    uint64_t *p = reinterpret_cast<uint64_t *>(wb);
    uint64_t *c = p + cap;
    uint32_t *sel = reinterpret_cast<uint32_t *>(c + FK_BVW_CH);
    uint32_t *sdeg = sel + W;
    uint32_t *todo = sdeg + W + 1;
    const uint32_t hsize = bvw_hsize(cap), hmask = hsize - 1;
    uint32_t *hs = todo + FK_BVW_TODO;
    uint8_t *pf = reinterpret_cast<uint8_t *>(hs + hsize);
    uint8_t *seln = pf + cap;

    uint64_t *pr = pool + (uint64_t)i * cap;
    uint8_t *pfr = pflg + (uint64_t)i * cap;
    const uint32_t np0 = cnt[i];
    const uint32_t lim = np0 < W ? np0 : W;
    // the window only; the rest of the pool is read when there is something
    // to merge into it
    bvw_load_pool(p, pf, pr, pfr, 0, lim);
    wave_lds_sync();
    // select the first S_cap flagged window entries, clear their flags
    uint32_t ns = 0;
    for (uint32_t j0 = 0; j0 < lim && ns < S_cap; j0 += 64) {
        const uint32_t j = j0 + lane;
        const uint8_t f = j < lim ? pf[j] : (uint8_t)0;
        const bool flagged = (f & FK_FLG_NEW) != 0;
        const uint64_t m = wave_ballot(flagged);
        const uint32_t r = ns + lanes_below(m);
        if (flagged && r < S_cap) {
            const uint32_t q = key_id(p[j]);
            sel[r] = q;
            seln[r] = (f & FK_FLG_ISNEW) ? 1 : 0;
            pf[j] = f & (uint8_t)~FK_FLG_NEW;
            if (mark) mark[q] = 1u;  // profiling: q's out-list is read this round
        }
        const uint32_t add = (uint32_t)__builtin_popcountll(m);
        ns = (ns + add) < S_cap ? (ns + add) : S_cap;
    }
    if (ns == 0) return;  // nothing selected, nothing changed
    wave_lds_sync();
    // global loads first: the first 64 selected entries' out-degrees and the
    // first batch of the rest of the pool; while they are in flight the id
    // set is cleared and takes the window's ids
    const uint32_t deg0 = lane < ns ? nsg_len[sel[lane]] : 0u;
    const uint32_t rest_hi = (np0 - lim) > 64 * FK_BVW_LOADS ? lim + 64 * FK_BVW_LOADS : np0;
    uint64_t rk[FK_BVW_LOADS];
    uint32_t rf[FK_BVW_LOADS];
    if (lim < np0) {
#pragma unroll
        for (uint32_t r = 0; r < FK_BVW_LOADS; ++r) {
            const uint32_t j = lim + r * 64 + lane;
            const uint32_t jj = j < rest_hi ? j : rest_hi - 1;
            rk[r] = pr[jj];
            rf[r] = pfr[jj];
        }
    }
    for (uint32_t j = lane; j < hsize; j += 64) hs[j] = FK_HS_EMPTY;
    wave_lds_sync();
    for (uint32_t j = lane; j < lim; j += 64) hs_seen_or_add(hs, hmask, key_id(p[j]));
    if (lim < np0) {
#pragma unroll
        for (uint32_t r = 0; r < FK_BVW_LOADS; ++r) {
            const uint32_t j = lim + r * 64 + lane;
            if (j < rest_hi) {
                p[j] = rk[r];
                pf[j] = (uint8_t)rf[r];
                hs_seen_or_add(hs, hmask, key_id(rk[r]));
            }
        }
    }
    if (rest_hi < np0) {  // pools longer than the window + 64 FK_BVW_LOADS
        bvw_load_pool(p, pf, pr, pfr, rest_hi, np0);
        wave_lds_sync();
        for (uint32_t j = rest_hi + lane; j < np0; j += 64) hs_seen_or_add(hs, hmask, key_id(p[j]));
    }
    wave_lds_sync();
    // exclusive prefix of the selected out-degrees
    uint32_t run = 0;
    for (uint32_t s0 = 0; s0 < ns; s0 += 64) {
        const uint32_t x = s0 + lane;
        const uint32_t v = s0 == 0 ? deg0 : (x < ns ? nsg_len[sel[x]] : 0u);
        uint32_t tot;
        const uint32_t ex = wave_excl_scan(v, tot);
        wave_lds_sync();
        if (x < ns) sdeg[x] = run + ex;
        run += tot;
    }
    if (lane == 0) sdeg[ns] = run;
    wave_lds_sync();
    const uint32_t total = run;
    const float thr = np0 >= cap ? key_dist(pr[cap - 1]) : INFINITY;
    const uint32_t g = lane >> 3, k = lane & 7u;
    const uint32_t nb = NB ? (uint32_t)NB : (dpad >> 3);
    const E *lane_chains = chains + k * nb;
    Chain mine;
    mine.load(lane_chains + (size_t)i * dpad, nb);
    const float ni = DOT ? nrm[i].x : 0.f;
    constexpr bool G16 = G == 1 && MODE == 0 && (NB == 12 || NB == 16 || NB == 32);
    constexpr bool Q8 = G == 2 && MODE == 0 && (NB == 12 || NB == 32);
    // 8-bit rows: width 96 one 128-byte line (lanes 0-5 values, lane 6 the
    // scalars), width 256 three (lane k pieces k and 8 + k, scalars piece 16)
    constexpr uint32_t QU = NB == 32 ? 24u : 8u;
    constexpr uint32_t I8 = NB == 32 ? FK_BVW_INFLIGHT8Q32 : FK_BVW_INFLIGHT8Q;
    constexpr uint32_t NBH = G16 ? (uint32_t)NB : 4u;
    const bool bound16 = G16 && r16 != nullptr && thr < INFINITY;  // wave-uniform
    const bool bound8 = Q8 && q8 != nullptr && thr < INFINITY;     // wave-uniform
    HalfSlice<NBH> mine16;
    float2 n16 = make_float2(0.f, 0.f);
    float t16 = 0.f;
    uint4 mine8 = make_uint4(0u, 0u, 0u, 0u);
    float s8 = 0.f, e8 = 0.f, nq8 = 0.f;  // the node's scale, error norm, |x^|^2 (octet lane 0)
    if (bound16) {
        mine16.load(r16 + (size_t)i * dpad + k * NBH);
        n16 = gnrm[i];
        t16 = __builtin_sqrtf(thr / (1.f - gb.gam)) * FK_BVW_UP20;
    }
    uint4 mine8b = make_uint4(0u, 0u, 0u, 0u);
    if (bound8) {
        mine8 = q8[(size_t)i * QU + k];
        if constexpr (NB == 32) {
            mine8b = q8[(size_t)i * QU + 8u + k];
            const uint4 sc = q8[(size_t)i * QU + 16u];
            s8 = __uint_as_float(sc.x);
            e8 = __uint_as_float(sc.y);
            nq8 = __uint_as_float(sc.z);
        } else {
            s8 = octet_from6(mine8.x);
            e8 = octet_from6(mine8.y);
            nq8 = octet_from6(mine8.z);
        }
        t16 = __builtin_sqrtf(thr / (1.f - gb.gam)) * FK_BVW_UP20;
    }
    uint32_t np = np0, nc = 0, nt = 0;
    unsigned long long cc = 0, computed = 0, rejected = 0;
    bool merged = false;
    uint32_t first_change = cap;  // smallest pool position rewritten
    // merges the candidates of c into the pool
    auto merge_c = [&]() __attribute__((always_inline)) {
        wave_lds_sync();
        uint32_t ins = 0, fst = cap;
        np = nc <= FK_BVW_SMALL ? bvw_merge_small(p, pf, np, c, nc, cap, ins, fst)
                                : bvw_merge(p, pf, np, c, nc, cap, ins, fst);
        if (ins > 0) {
            merged = true;
            first_change = fst < first_change ? fst : first_change;
        }
        nc = 0;
    };
    // the distances of todo[0, m): one octet per distance, INF rows per
    // octet in flight (slots past m compute a discarded distance to row 0);
    // candidates (distance <= thr) go to c, merged before c can overflow
    auto run_todo = [&](uint32_t m) __attribute__((always_inline)) {
        if (bound8) {
            // the bound pass on the 8-bit rows (one 16-byte piece per lane,
            // a whole 128-byte row per octet); survivors compacted as below
            uint32_t m2 = 0;
            for (uint32_t e0 = 0; e0 < m; e0 += 8 * I8) {
                uint4 hq[I8], hq2[NB == 32 ? I8 : 1], hs[NB == 32 ? I8 : 1];
                uint32_t tt[I8];
#pragma unroll
                for (uint32_t u = 0; u < I8; ++u) {
                    const uint32_t e = e0 + u * 8 + g;
                    tt[u] = e < m ? todo[e] : i;
                    hq[u] = q8[(size_t)tt[u] * QU + k];
                    if constexpr (NB == 32) {
                        hq2[u] = q8[(size_t)tt[u] * QU + 8u + k];
                        hs[u] = q8[(size_t)tt[u] * QU + 16u];
                    }
                }
                wave_lds_sync();
#pragma unroll
                for (uint32_t u = 0; u < I8; ++u) {
                    int part = __builtin_amdgcn_sdot4((int)mine8.x, (int)hq[u].x, 0, false);
                    part = __builtin_amdgcn_sdot4((int)mine8.y, (int)hq[u].y, part, false);
                    part = __builtin_amdgcn_sdot4((int)mine8.z, (int)hq[u].z, part, false);
                    part = __builtin_amdgcn_sdot4((int)mine8.w, (int)hq[u].w, part, false);
                    float sy, ey, nx;
                    if constexpr (NB == 32) {
                        part = __builtin_amdgcn_sdot4((int)mine8b.x, (int)hq2[u].x, part, false);
                        part = __builtin_amdgcn_sdot4((int)mine8b.y, (int)hq2[u].y, part, false);
                        part = __builtin_amdgcn_sdot4((int)mine8b.z, (int)hq2[u].z, part, false);
                        part = __builtin_amdgcn_sdot4((int)mine8b.w, (int)hq2[u].w, part, false);
                        sy = __uint_as_float(hs[u].x);
                        ey = __uint_as_float(hs[u].y);
                        nx = nq8 + __uint_as_float(hs[u].z);
                    } else {
                        part = k < 6u ? part : 0;  // lanes 6 and 7 hold the row's scalars
                        sy = octet_from6(hq[u].x);
                        ey = octet_from6(hq[u].y);
                        nx = nq8 + octet_from6(hq[u].z);
                    }
                    const float dot = octet_sum_i32(part);
                    // |x^ - y^|^2 in FP32 (dot exact), its error within FK_Q8_FREL nx
                    const float dt = nx - 2.f * ((s8 * sy) * dot);
                    const float f = FK_Q8_FREL * nx;
                    const float lo = t16 + (e8 + ey) * FK_BVW_UP20;
                    const bool far = (dt - f) * FK_BVW_DN18 > lo * lo * FK_BVW_UP18;
                    const bool live = (e0 + u * 8 + g) < m && k == 0;
                    const uint64_t mk = wave_ballot(live && !far);
                    if (live && !far) todo[m2 + lanes_below(mk)] = tt[u];
                    m2 += (uint32_t)__builtin_popcountll(mk);
                    rejected += (uint32_t)__builtin_popcountll(wave_ballot(live && far));
                }
                wave_lds_sync();
            }
            m = m2;
        }
        if (bound16) {
            // the bound pass: survivors compacted to todo[0, m2) in order
            // (a chunk's entries are read before any is written back, and
            // writes never pass the chunk's end)
            constexpr uint32_t I16 = NBH >= 16 ? FK_BVW_INFLIGHT16W : FK_BVW_INFLIGHT16;
            uint32_t m2 = 0;
            for (uint32_t e0 = 0; e0 < m; e0 += 8 * I16) {
                HalfSlice<NBH> hh[I16];
                uint32_t tt[I16];
#pragma unroll
                for (uint32_t u = 0; u < I16; ++u) {
                    const uint32_t e = e0 + u * 8 + g;
                    tt[u] = e < m ? todo[e] : i;
                    hh[u].load(r16 + (size_t)tt[u] * dpad + k * NBH);
                }
                wave_lds_sync();
#pragma unroll
                for (uint32_t u = 0; u < I16; ++u) {
                    const float dot = octet_sum_any(mine16.dot(hh[u]));
                    // the target's |y~|^2 from the row just read (no norm
                    // read per target: on random targets that is another
                    // cache line each); |y| <= the largest row norm
                    const float ny = octet_sum_any(hh[u].dot(hh[u])) * gb.inv_s2;
                    const float nx = n16.x + ny;
                    const float dt = nx - 2.f * (dot * gb.inv_s2);
                    const float f = gb.f_rel * nx;
                    const float e = (gb.u * (n16.y + gb.ymax) + gb.e_abs) * FK_BVW_UP20;
                    const float lo = t16 + e;
                    const bool far = (dt - f) * FK_BVW_DN18 > lo * lo * FK_BVW_UP18;
                    const bool live = (e0 + u * 8 + g) < m && k == 0;
                    const uint64_t mk = wave_ballot(live && !far);
                    if (live && !far) todo[m2 + lanes_below(mk)] = tt[u];
                    m2 += (uint32_t)__builtin_popcountll(mk);
                    rejected += (uint32_t)__builtin_popcountll(wave_ballot(live && far));
                }
                wave_lds_sync();
            }
            m = m2;
        }
        for (uint32_t e0 = 0; e0 < m; e0 += 8 * INF) {
            if (nc > FK_BVW_CH - 8 * INF) merge_c();
            Chain rr[INF];
            uint32_t te[INF];
            float tn[INF];
#pragma unroll
            for (uint32_t u = 0; u < INF; ++u) {
                const uint32_t e = e0 + u * 8 + g;
                te[u] = e < m ? todo[e] : 0u;
                rr[u].load(lane_chains + (size_t)te[u] * dpad, nb);
                tn[u] = DOT ? nrm[te[u]].x : 0.f;
            }
#pragma unroll
            for (uint32_t u = 0; u < INF; ++u) {
                const float dd = DOT ? exact_dist(ni, tn[u], mine.dot_reg(rr[u], nb)) : mine.dist_reg(rr[u], nb);
                const bool ok = (e0 + u * 8 + g) < m && k == 0 && dd <= thr;
                const uint64_t mc = wave_ballot(ok);
                if (ok) c[nc + lanes_below(mc)] = make_key(dd, te[u]);
                nc += (uint32_t)__builtin_popcountll(mc);
            }
        }
    };
    for (uint32_t s0 = 0; s0 < total; s0 += 64 * FK_BVW_PREFETCH) {
        // the out-list entries of FK_BVW_PREFETCH chunks of 64 (entry,
        // neighbour) pairs are read at once; pair f belongs to entry s with
        // sdeg[s] <= f < sdeg[s + 1]
        uint32_t tr[FK_BVW_PREFETCH];
        uint8_t nr[FK_BVW_PREFETCH], sr[FK_BVW_PREFETCH];
#pragma unroll
        for (uint32_t r = 0; r < FK_BVW_PREFETCH; ++r) {
            const uint32_t f = s0 + r * 64 + lane;
            tr[r] = i;  // "t == i": unused
            nr[r] = sr[r] = 0;
            if (f < total) {
                uint32_t lo = 0, hi = ns;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (sdeg[mid] <= f) lo = mid; else hi = mid;
                }
                const uint64_t at = (uint64_t)sel[lo] * ncap + (f - sdeg[lo]);
                tr[r] = nsg_id[at];
                nr[r] = nsg_new[at];
                sr[r] = seln[lo];
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < FK_BVW_PREFETCH; ++r) {
            const uint32_t f0 = s0 + r * 64;
            if (f0 >= total) break;
            const uint32_t t = tr[r];
            const bool use = t != i && (sr[r] || nr[r]);
            cc += (uint32_t)__builtin_popcountll(wave_ballot(use));  // every pair the reference evaluates
            // computed only when the target is not in the id set (which
            // then takes it)
            const bool fresh = use && !hs_seen_or_add(hs, hmask, t);
            const uint64_t mf = wave_ballot(fresh);
            const uint32_t nf = (uint32_t)__builtin_popcountll(mf);
            if (fresh) todo[nt + lanes_below(mf)] = t;
            nt += nf;
            computed += nf;
        }
        wave_lds_sync();
        if (s0 + 64 * FK_BVW_PREFETCH >= total) {
            run_todo(nt);
            nt = 0;
        } else if (nt >= 64) {
            // whole iterations now; the remainder (< 8 INF) moves to the front
            const uint32_t m = nt / (8 * INF) * (8 * INF);
            run_todo(m);
            const uint32_t rest = nt - m;
            const uint32_t v = lane < rest ? todo[m + lane] : 0u;
            wave_lds_sync();
            if (lane < rest) todo[lane] = v;
            wave_lds_sync();
            nt = rest;
        }
    }
    if (nc > 0) merge_c();
    // the window's flags (the selection) and everything from the first
    // rewritten position on
    const uint32_t fl = lim < first_change ? lim : first_change;
    for (uint32_t j = lane; j < fl; j += 64) pfr[j] = pf[j];
    if (merged) {
        for (uint32_t j = first_change + lane; j < np; j += 64) { pr[j] = p[j]; pfr[j] = pf[j]; }
        if (lane == 0) cnt[i] = np;
    }
    if (lane == 0 && cc) tally_add(cc_acc, cc);  // the same counts in every lane
    if (lane == 0 && computed) tally_add(comp_acc, computed);
    if (lane == 0 && rejected) tally_add(rej_acc, rejected);
}

static void launch_bv_round(BuildState &b, uint32_t W, uint32_t scap, const uint32_t *set_scap) {
    const uint32_t *ord = b.has_order ? b.order.p : nullptr;
    const uint32_t wave_bytes = (uint32_t)bvw_wave_bytes(b.cap, W);
    uint32_t wpb = (64u * 1024u) / wave_bytes;
    wpb = wpb < FK_BVW_WAVES ? wpb : FK_BVW_WAVES;
    FK_CHECK_ARG(wpb >= 1, "pool too large for the BridgeView kernel");
    const size_t lds = (size_t)wave_bytes * wpb;
    const uint32_t nblk = (b.N + wpb - 1) / wpb;
    const unsigned grid = ord ? (nblk + FK_NUM_XCD - 1) / FK_NUM_XCD * FK_NUM_XCD : nblk;
    uint32_t *mark = b.prof ? b.bv_mark.p : nullptr;
    const float2 *nrm = exact_norms(b);
    const uint8_t *c8 = nrm && b.ds->chains8.p ? b.ds->chains8.p : nullptr;
    // the bound pass: FK_FLAG_GRAM with the FP16 form (MODE 0 kernels of
    // widths 96, 128, 256 take it)
    const bool g16 = (b.P.flags & FK_FLAG_GRAM) && b.ds->gram_fmt == FK_GRAM_F16 && b.ds->rows16.p;
    const uint16_t *r16 = g16 ? b.ds->rows16.p : nullptr;
    const GramBound gb = g16 ? gram_bound(*b.ds) : GramBound{};
    // widths 96 and 256: the pass on the 8-bit rows when they are made
    const uint4 *q8 = g16 && b.ds->rows8q.p ? b.ds->rows8q.p : nullptr;
#define FK_BV_LAUNCH_G(NBV, MODE, G)                                                                             \
    hipLaunchKernelGGL((k_bv_round<NBV, MODE, G>), dim3(grid), dim3(64 * wpb), lds, b.s,                       \
                       MODE == 2 ? (const void *)c8 : (const void *)b.ds->chains.p, nrm,                        \
                       b.ds->dpad, b.N, b.cap, W, scap, b.dev_set_of(), set_scap, b.pool.p, b.pflg.p, b.cnt.p,  \
                       b.nsg_id.p, b.nsg_len.p,                                                                 \
                       b.nsg_new.p, b.ncap, b.tally.at(FK_T_BV_CC), b.tally.at(FK_T_BV_COMP), ord, mark,       \
                       wave_bytes, wpb, r16, b.ds->gram_nrm.p, gb, q8, b.tally.at(FK_T_BV_REJ))
#define FK_BV_LAUNCH(NBV, MODE) FK_BV_LAUNCH_G(NBV, MODE, 0)
#define FK_BV_LAUNCH16(NBV) do { if (r16) FK_BV_LAUNCH_G(NBV, 0, 1); else FK_BV_LAUNCH_G(NBV, 0, 0); } while (0)
    switch (b.ds->dpad) {
        case 32: FK_BV_LAUNCH(4, 0); break;
        case 64: FK_BV_LAUNCH(8, 0); break;
        case 96: if (nrm) FK_BV_LAUNCH(12, 1); else if (q8) FK_BV_LAUNCH_G(12, 0, 2); else FK_BV_LAUNCH16(12); break;
        case 128: if (c8) FK_BV_LAUNCH(16, 2); else if (nrm) FK_BV_LAUNCH(16, 1); else FK_BV_LAUNCH16(16); break;
        case 256:
            if (c8) FK_BV_LAUNCH(32, 2);
            else if (nrm) FK_BV_LAUNCH(32, 1);
            else if (q8) FK_BV_LAUNCH_G(32, 0, 2);
            else FK_BV_LAUNCH16(32);
            break;
        default: FK_BV_LAUNCH(0, 0); break;
    }
#undef FK_BV_LAUNCH16
#undef FK_BV_LAUNCH
#undef FK_BV_LAUNCH_G
}

void bridgeview(BuildState &b) {
    const double t0 = now_sec();
    const uint32_t N = b.N;
    const BuildParams &P = b.P;
    const uint32_t cur_loop = b.loopcount - 1;
    const uint32_t cur_L = std::min(P.search_L + cur_loop * P.step, P.MAXL);
    const uint32_t pre_L = cur_loop ? std::min(P.search_L + (cur_loop - 1) * P.step, P.MAXL) : 0;
    const unsigned gw = (N + 3) / 4;
    hipLaunchKernelGGL(k_bv_init, dim3(gw), dim3(256), 0, b.s, b.pflg.p, b.cnt.p, N, b.cap, cur_loop, pre_L, cur_L);
    FK_LAUNCH_CHECK();
    const uint32_t W = std::min(P.search_L + (b.loopcount - 1) * P.step, P.MAXL);
    const uint32_t S = P.massq_S;
    uint64_t cc_total = 0;
    uint32_t rounds = 0;
    // per set: the rounds run until the set's own last round
    std::vector<uint8_t> done(b.n_sets, 0);
    std::vector<uint32_t> scap_set(b.n_sets);
    DevBuf<uint32_t> d_scap(b.n_sets);
    while (true) {
        FK_HIP(hipMemsetAsync(b.set_acc.p, 0, 16 * (size_t)b.n_sets, b.s));
        hipLaunchKernelGGL(k_bv_count, dim3(reduce_grid(N)), dim3(1024), 0, b.s, b.pflg.p, b.cnt.p, N, b.cap, W, S,
                           b.dev_set_of(), b.set_acc.p);
        FK_LAUNCH_CHECK();
        std::vector<unsigned long long> sc(2 * (size_t)b.n_sets);
        FK_HIP(hipMemcpyAsync(sc.data(), b.set_acc.p, 16 * (size_t)b.n_sets, hipMemcpyDeviceToHost, b.s));
        b.sync();
        // delta = cnt / nd_ / S_ in float (src/kgraph.cpp:2492); below 0.5 the
        // round also expands every remaining flagged entry and is the last.
        unsigned long long cnts[2] = {0, 0};
        bool last = true;
        for (uint32_t k = 0; k < b.n_sets; ++k) {
            if (done[k]) { scap_set[k] = 0; continue; }
            const float delta = (float)sc[2 * k] / (float)b.set_n(k) / (float)S;
            const bool lk = delta < 0.5f;
            scap_set[k] = lk ? 0xFFFFFFFFu : S;
            if (lk) done[k] = 1;
            else last = false;
            cnts[0] += sc[2 * k];
            cnts[1] += sc[2 * k + 1];
        }
        const uint32_t scap = scap_set[0];
        if (b.n_sets > 1)
            FK_HIP(hipMemcpyAsync(d_scap.p, scap_set.data(), (size_t)b.n_sets * 4, hipMemcpyHostToDevice, b.s));
        b.tally.clear(b.s, FK_T_BV_CC, 2);  // and FK_T_BV_COMP
        b.tally.clear(b.s, FK_T_BV_REJ);
        if (b.prof && b.bv_mark.n < N) {
            b.bv_mark.alloc(N);
            FK_HIP(hipMemsetAsync(b.bv_mark.p, 0, (size_t)N * 4, b.s));
        }
        hipEvent_t pe = b.prof_begin();
        launch_bv_round(b, W, scap, b.n_sets > 1 ? d_scap.p : nullptr);
        FK_LAUNCH_CHECK();
        b.prof_end(pe, FK_PROF_BV, 0.0);
        unsigned long long cc2[2] = {0, 0};
        b.tally.read(b.s, FK_T_BV_CC, 2, cc2);
        const unsigned long long cc = cc2[0];
        b.st->d_bv_computed += cc2[1];
        b.st->d_bv_rejected += b.tally.read1(b.s, FK_T_BV_REJ);
        if (b.prof) {
            // minimal traffic of a round: the rows (+ ids) of every expansion
            // (|rnn_q|) and of every out-list entry of an expanded q
            // (|nsg_q|), each read once, plus one read and one write of the
            // packed pools (8 bytes per slot, MAXL + 1 slots)
            FK_HIP(hipMemsetAsync(b.acc.p + 7, 0, 8, b.s));
            hipLaunchKernelGGL(k_marked_degree, dim3(reduce_grid(N / 64 + 1)), dim3(1024), 0, b.s, b.bv_mark.p,
                               b.nsg_len.p, N, b.acc.p + 7);
            FK_LAUNCH_CHECK();
            unsigned long long deg = 0;
            FK_HIP(hipMemcpyAsync(&deg, b.acc.p + 7, 8, hipMemcpyDeviceToHost, b.s));
            b.sync();
            const double expansions = (double)(last ? cnts[1] : cnts[0]);
            b.prof_set_last_bytes((expansions + (double)deg) * (4.0 * b.ds->d + 4.0) +
                                  16.0 * (double)N * (double)(P.MAXL + 1));
        }
        b.sync();
        b.prof_flush();
        cc_total += cc;
        ++rounds;
        if (b.verbose)
            fprintf(stderr, "[fk] bridgeview round %u: expansions %llu (all %llu) dists %llu%s\n", rounds,
                    cnts[0], cnts[1], cc, last ? " (last)" : "");
        if (last) break;
    }
    const uint32_t li = b.loopcount - 1;
    if (li < FK_MAX_LOOPS) {
        b.st->search_cc[li] = (float)((double)cc_total / ((double)N * (double)(N - 1) / 2.0));
        b.st->bv_rounds[li] = rounds;
    }
    b.st->d_bv += cc_total;
    b.st->t_bv += now_sec() - t0;
}

}  // namespace fk
