// Graph search and connectivity:
//   * beam search over the out-lists (get_neighbors(const float*, ...),
//     src/kgraph.cpp:1560-1623), one wavefront per query; used for the entry
//     point (ComputeEp, :1528-1536) and, batched, for query evaluation
//   * the centroid (MatrixOracle::Calcenter, src/kgraph-data.h:413-428)
//   * tree_grow connectivity (src/kgraph.cpp:2145-2270): the reference's DFS
//     only decides which nodes are reached, so it is a BFS here; nodes left
//     unreached are attached in ascending id order exactly as findroot does
//     (first reached pool neighbour with degree < nsg_R, else a random
//     reached node), each followed by a BFS from the attached node.
#include <math.h>

#include <algorithm>
#include <memory>

#include "fk_api_impl.h"
#include "fk_build.h"
#include "fk_device.h"
#include "fk_sort.h"

namespace fk {

// ---------------------------------------------------------------------------
// Centroid (MatrixOracle::Calcenter, src/kgraph-data.h:413-428): float sums
// in row order, then a float division by the row count.  The mean of all
// resident rows is made on the host as the rows are uploaded
// (Dataset::center); the mean of a row range (one set of a batched build) by
// k_colmean_ordered: one lane per column adding the rows one after the other,
// which is that order exactly; the loads of 16 rows are issued before their
// adds.  Range k of a launch is rows [lo[k], lo[k] + n[k]) into
// center[k * dpad].
// ---------------------------------------------------------------------------
#define FK_CEN_UNROLL 16
__global__ __launch_bounds__(64) void k_colmean_ordered(const float *__restrict__ rows, uint32_t stride, uint32_t d,
                                                        uint32_t dpad, const uint32_t *__restrict__ lo,
                                                        const uint32_t *__restrict__ n, float *__restrict__ center) {
    const uint32_t k = blockIdx.y, c = blockIdx.x * 64 + threadIdx.x;
    if (c >= d) return;
    const uint32_t r0 = lo[k], m = n[k];
    const float *col = rows + (size_t)r0 * stride + c;
    float s = 0.f;
    uint32_t r = 0;
    for (; r + FK_CEN_UNROLL <= m; r += FK_CEN_UNROLL) {
        float v[FK_CEN_UNROLL];
#pragma unroll
        for (uint32_t u = 0; u < FK_CEN_UNROLL; ++u) v[u] = col[(size_t)(r + u) * stride];
#pragma unroll
        for (uint32_t u = 0; u < FK_CEN_UNROLL; ++u) s += v[u];
    }
    for (; r < m; ++r) s += col[(size_t)r * stride];
    center[(size_t)k * dpad + c] = s / (float)m;
}

// ---------------------------------------------------------------------------
// Beam search, one 64-lane workgroup per query.
// init_mode 0: start from the first L entries of pool(init_node)
//              (get_neighbors, :1566-1571), random fill up to L
// init_mode 1: start from ep and ep's out-list, random fill up to L
// The visited set is a per-query bitmap of N bits.
// ---------------------------------------------------------------------------
#define FK_BEAM_MAX 1024

struct BeamArgs {
    const float *rows;
    uint32_t stride, dpad, N;
    const float *queries;      // n_q x dpad
    uint32_t n_q;
    uint32_t L, K;
    int init_mode;
    uint32_t init_node;        // mode 0 (per launch)
    const uint64_t *pool;      // mode 0
    const uint32_t *pool_cnt;
    uint32_t pool_cap;
    uint32_t ep;               // mode 1
    // graph: either fixed-stride lists or CSR
    const uint32_t *g_ids;
    const uint32_t *g_len;     // fixed stride mode
    uint32_t g_cap;
    const uint64_t *g_off;     // CSR mode when non-null
    uint32_t *visited;         // n_q x ((N + 31) / 32) words, zeroed (or per-query ranges below)
    uint64_t seed;
    uint32_t q_base;           // index of query 0 of this launch (its fill stream is q_base + q)
    // per-query id ranges (sets): query q searches ids [q_lo[q], q_lo[q] + q_n[q]),
    // starts from pool(q_init[q]), fills with stream 0 of q_seed[q] (what a
    // build of the set alone does), and owns visited words from q_vis[q]
    const uint32_t *q_lo, *q_n, *q_init;
    const uint64_t *q_seed, *q_vis;
    uint32_t *out_ids;         // n_q x K
    float *out_d;
    unsigned long long *n_dist;
};

// fk_search's random fill: query i draws fk_rand32(FK_SEARCH_SEED, i, k)
// for k = 0, 1, ... (independent of how the queries are batched)
#define FK_SEARCH_SEED 0x9E3779B97F4A7C15ull

__device__ __forceinline__ bool test_and_set(uint32_t *bm, uint32_t id) {
    const uint32_t bit = 1u << (id & 31u);
    return (atomicOr(&bm[id >> 5], bit) & bit) == 0;
}

__global__ __launch_bounds__(64) void k_beam(BeamArgs a) {
    __shared__ uint64_t s_key[FK_BEAM_MAX + 64];
    __shared__ uint8_t s_flag[FK_BEAM_MAX + 64];
    __shared__ uint64_t s_o[FK_BEAM_MAX + 64];
    __shared__ uint8_t s_of[FK_BEAM_MAX + 64];
    __shared__ uint32_t s_init[FK_BEAM_MAX];
    __shared__ __align__(16) float s_q[1024];
    __shared__ uint32_t s_n;
    const uint32_t q = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const float *qv = a.queries + (size_t)q * a.dpad;
    for (uint32_t k = lane; k < a.dpad; k += 64) s_q[k] = qv[k];
    const uint32_t lo = a.q_lo ? a.q_lo[q] : 0u, Nq = a.q_lo ? a.q_n[q] : a.N;
    // the visited bitmap indexes ids relative to lo
    uint32_t *bm = a.visited + (a.q_vis ? a.q_vis[q] : (size_t)q * ((a.N + 31) / 32));
    const uint32_t L = a.L;
    // initial ids
    if (lane == 0) {
        uint32_t n = 0;
        if (a.init_mode == 0) {
            const uint32_t init = a.q_init ? a.q_init[q] : a.init_node;
            const uint32_t pc = a.pool_cnt[init];
            const uint64_t *pr = a.pool + (uint64_t)init * a.pool_cap;
            for (uint32_t i = 0; i < pc && n < L; ++i) {
                const uint32_t id = key_id(pr[i]);
                if (id < a.N && test_and_set(bm, id - lo)) s_init[n++] = id;
            }
        } else {
            if (test_and_set(bm, a.ep - lo)) s_init[n++] = a.ep;
            const uint64_t o = a.g_off ? a.g_off[a.ep] : (uint64_t)a.ep * a.g_cap;
            const uint32_t deg = a.g_off ? (uint32_t)(a.g_off[a.ep + 1] - o) : a.g_len[a.ep];
            for (uint32_t m = 0; m < deg && n < L; ++m) {
                const uint32_t id = a.g_ids[o + m];
                if (id < a.N && test_and_set(bm, id - lo)) s_init[n++] = id;
            }
        }
        uint64_t k = 0;
        const uint32_t want = L < Nq ? L : Nq;
        const uint64_t seed = a.q_seed ? a.q_seed[q] : a.seed;
        const uint64_t stream = a.q_seed ? 0u : (uint64_t)a.q_base + q;
        while (n < want) {
            const uint32_t id = lo + fk_rand32(seed, stream, k++) % Nq;
            if (test_and_set(bm, id - lo)) s_init[n++] = id;
        }
        s_n = n;
    }
    __syncthreads();
    uint32_t Lc = s_n;
    unsigned long long nd = 0;
    const uint32_t np2 = next_pow2(Lc < 2 ? 2 : Lc);
    for (uint32_t k = lane; k < np2; k += 64) {
        if (k < Lc) {
            const uint32_t id = s_init[k];
            s_key[k] = make_key(l2_canon(s_q, a.rows + (size_t)id * a.stride, a.dpad), id);
            ++nd;
        } else {
            s_key[k] = FK_KEY_EMPTY;
        }
    }
    __syncthreads();
    block_bitonic_sort(s_key, np2);
    for (uint32_t k = lane; k < Lc; k += 64) s_flag[k] = 1;
    __syncthreads();
    uint32_t k = 0;
    while (k < Lc) {
        uint32_t nk = Lc;
        if (s_flag[k]) {
            __syncthreads();
            if (lane == 0) s_flag[k] = 0;
            const uint32_t n = key_id(s_key[k]);
            const uint64_t o = a.g_off ? a.g_off[n] : (uint64_t)n * a.g_cap;
            const uint32_t deg = a.g_off ? (uint32_t)(a.g_off[n + 1] - o) : a.g_len[n];
            for (uint32_t m0 = 0; m0 < deg; m0 += 64) {
                const uint32_t m = m0 + lane;
                uint64_t cand = FK_KEY_EMPTY;
                const float worst = key_dist(s_key[Lc - 1]);
                if (m < deg) {
                    const uint32_t id = a.g_ids[o + m];
                    if (id < a.N && test_and_set(bm, id - lo)) {
                        ++nd;
                        const float dd = l2_canon(s_q, a.rows + (size_t)id * a.stride, a.dpad);
                        if (dd < worst) cand = make_key(dd, id);  // `dist >= pool[L-1].dist` skips
                    }
                }
                const uint64_t sorted = wave_sort64(cand);
                const uint64_t mv = wave_ballot(sorted != FK_KEY_EMPTY);
                const uint32_t nc = (uint32_t)__popcll(mv);
                if (nc == 0) continue;
                // rank-merge the sorted candidates (lanes 0..nc-1) into the pool
                __syncthreads();
                const uint32_t lim = (Lc + nc) < L ? (Lc + nc) : L;
                if (lane < nc) s_o[FK_BEAM_MAX + lane] = sorted;
                __syncthreads();
                const uint64_t *cs = s_o + FK_BEAM_MAX;
                for (uint32_t j = lane; j < Lc; j += 64) {
                    const uint32_t pos = j + lower_bound_cnt(cs, nc, s_key[j]);
                    if (pos < lim) { s_o[pos] = s_key[j]; s_of[pos] = s_flag[j]; }
                }
                uint32_t minpos = 0xFFFFFFFFu;
                if (lane < nc) {
                    const uint32_t pos = lane + lower_bound_cnt(s_key, Lc, sorted);
                    if (pos < lim) { s_o[pos] = sorted; s_of[pos] = 1; minpos = pos; }
                }
                // smallest insertion position (lane 0 holds the smallest key)
                const uint32_t r0 = __shfl(minpos, 0, 64);
                __syncthreads();
                for (uint32_t j = lane; j < lim; j += 64) { s_key[j] = s_o[j]; s_flag[j] = s_of[j]; }
                __syncthreads();
                Lc = lim;
                if (r0 < nk) nk = r0;
            }
        }
        if (nk <= k) k = nk; else ++k;
        __syncthreads();
    }
    const uint32_t kk = a.K;
    for (uint32_t j = lane; j < kk; j += 64) {
        const uint64_t key = j < Lc ? s_key[j] : FK_KEY_EMPTY;
        a.out_ids[(size_t)q * kk + j] = key_id(key);
        if (a.out_d) a.out_d[(size_t)q * kk + j] = key_dist(key);
    }
    if (a.n_dist) {
        nd = wave_sum(nd);
        if (lane == 0) tally_add(a.n_dist, nd);
    }
}

static void launch_beam(hipStream_t s, const BeamArgs &a) {
    FK_CHECK_ARG(a.L >= 1 && a.L <= FK_BEAM_MAX, "search_L must be in [1, 1024]");
    FK_CHECK_ARG(a.dpad <= 1024, "dimension above 1024 is not supported by the beam search");
    FK_CHECK_ARG(a.K <= a.L, "K must not exceed search_L");
    hipLaunchKernelGGL(k_beam, dim3(a.n_q), dim3(64), 0, s, a);
    FK_LAUNCH_CHECK();
}

// Column means of the resident rows [r0, r0 + n) into center[0, d) (the
// caller zeroes the padding); synchronises the stream.
void column_mean(hipStream_t s, const Dataset &ds, float *center, uint32_t r0, uint32_t n) {
    if (n == 0) n = ds.n_total - r0;
    if (r0 == 0 && n == ds.n_total) {  // made at upload
        FK_HIP(hipMemcpyAsync(center, ds.center.p, (size_t)ds.dpad * 4, hipMemcpyDeviceToDevice, s));
        return;
    }
    column_means(s, ds, center, &r0, &n, 1);
}

// The ranges' centroids are made once per uploaded dataset (a batched
// build asks for them in every graph loop, and repeated builds of the same
// sets ask again): ranges not yet in ds.range_center are computed in one
// launch and kept on the host, then every requested one is copied to
// center[k * dpad].
void column_means(hipStream_t s, const Dataset &ds, float *center, const uint32_t *r0, const uint32_t *n,
                  uint32_t count) {
    if (count == 0) return;
    std::vector<uint32_t> miss_lo, miss_n;
    for (uint32_t k = 0; k < count; ++k)
        if (!ds.range_center.count({r0[k], n[k]})) {
            miss_lo.push_back(r0[k]);
            miss_n.push_back(n[k]);
        }
    const uint32_t m = (uint32_t)miss_lo.size();
    if (m) {
        DevBuf<uint32_t> rng(2 * (size_t)m);
        DevBuf<float> out((size_t)m * ds.dpad);
        FK_HIP(hipMemcpyAsync(rng.p, miss_lo.data(), (size_t)m * 4, hipMemcpyHostToDevice, s));
        FK_HIP(hipMemcpyAsync(rng.p + m, miss_n.data(), (size_t)m * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_colmean_ordered, dim3((ds.d + 63) / 64, m), dim3(64), 0, s, ds.rows.p, ds.stride, ds.d,
                           ds.dpad, rng.p, rng.p + m, out.p);
        FK_LAUNCH_CHECK();
        std::vector<float> host((size_t)m * ds.dpad);
        FK_HIP(hipMemcpyAsync(host.data(), out.p, host.size() * 4, hipMemcpyDeviceToHost, s));
        FK_HIP(hipStreamSynchronize(s));  // rng and out are released on return
        for (uint32_t k = 0; k < m; ++k)
            ds.range_center[{miss_lo[k], miss_n[k]}] =
                std::vector<float>(host.begin() + (size_t)k * ds.dpad, host.begin() + (size_t)k * ds.dpad + ds.d);
    }
    std::vector<float> all((size_t)count * ds.dpad, 0.f);
    for (uint32_t k = 0; k < count; ++k) {
        const std::vector<float> &c = ds.range_center.at({r0[k], n[k]});
        std::copy(c.begin(), c.end(), all.begin() + (size_t)k * ds.dpad);
    }
    FK_HIP(hipMemcpyAsync(center, all.data(), all.size() * 4, hipMemcpyHostToDevice, s));
    FK_HIP(hipStreamSynchronize(s));  // all is released on return
}

// ComputeEp (src/kgraph.cpp:1528-1536), per set: the centroid of the set's
// rows (all resident rows for a single build: Calcenter averages the whole
// matrix, not the oracle's prefix), then a beam search from a random node's
// pool (init_nodes[k] instead, when given), one query per set.
void compute_ep(BuildState &b, const uint32_t *init_nodes) {
    const double t0 = now_sec();
    const Dataset &ds = *b.ds;
    const uint32_t ns = b.n_sets;
    DevBuf<float> center((size_t)ns * ds.dpad);
    FK_HIP(hipMemsetAsync(center.p, 0, center.bytes(), b.s));
    hipEvent_t pe = b.prof_begin();
    std::vector<uint32_t> q_lo(ns), q_n(ns), q_init(ns);
    std::vector<uint64_t> q_seed(ns), q_vis(ns);
    uint64_t words = 0;
    double center_rows = 0;
    std::vector<uint32_t> set_n(ns);
    for (uint32_t k = 0; k < ns; ++k) set_n[k] = b.set_n(k);
    if (ns == 1) column_mean(b.s, ds, center.p, 0, ds.n_total);
    else column_means(b.s, ds, center.p, b.set_lo.data(), set_n.data(), ns);
    for (uint32_t k = 0; k < ns; ++k) {
        const uint32_t lo = b.set_lo[k], n = b.set_n(k);
        center_rows += ns == 1 ? ds.n_total : n;
        q_lo[k] = lo;
        q_n[k] = n;
        q_init[k] = lo + (uint32_t)(b.next_rand(k) % n);
        if (init_nodes) q_init[k] = init_nodes[k];
        q_seed[k] = b.next_rand(k);
        q_vis[k] = words;
        words += (n + 31) / 32;
    }
    DevBuf<uint32_t> vis(words), oid(ns), dq(3 * (size_t)ns);
    DevBuf<uint64_t> dq64(2 * (size_t)ns);
    FK_HIP(hipMemsetAsync(vis.p, 0, vis.bytes(), b.s));
    FK_HIP(hipMemcpyAsync(dq.p, q_lo.data(), (size_t)ns * 4, hipMemcpyHostToDevice, b.s));
    FK_HIP(hipMemcpyAsync(dq.p + ns, q_n.data(), (size_t)ns * 4, hipMemcpyHostToDevice, b.s));
    FK_HIP(hipMemcpyAsync(dq.p + 2 * ns, q_init.data(), (size_t)ns * 4, hipMemcpyHostToDevice, b.s));
    FK_HIP(hipMemcpyAsync(dq64.p, q_seed.data(), (size_t)ns * 8, hipMemcpyHostToDevice, b.s));
    FK_HIP(hipMemcpyAsync(dq64.p + ns, q_vis.data(), (size_t)ns * 8, hipMemcpyHostToDevice, b.s));
    BeamArgs a{};
    a.rows = ds.rows.p;
    a.stride = ds.stride;
    a.dpad = ds.dpad;
    a.N = b.N;
    a.queries = center.p;
    a.n_q = ns;
    a.L = b.P.search_L;
    a.K = 1;
    a.init_mode = 0;
    a.pool = b.pool.p;
    a.pool_cnt = b.cnt.p;
    a.pool_cap = b.cap;
    a.g_ids = b.nsg_id.p;
    a.g_len = b.nsg_len.p;
    a.g_cap = b.ncap;
    a.visited = vis.p;
    a.q_lo = dq.p;
    a.q_n = dq.p + ns;
    a.q_init = dq.p + 2 * ns;
    a.q_seed = dq64.p;
    a.q_vis = dq64.p + ns;
    a.out_ids = oid.p;
    a.n_dist = b.tally.at(FK_T_SEARCH);
    b.tally.clear(b.s, FK_T_SEARCH);
    launch_beam(b.s, a);
    FK_HIP(hipMemcpyAsync(b.set_ep.data(), oid.p, (size_t)ns * 4, hipMemcpyDeviceToHost, b.s));
    const unsigned long long nd = b.tally.read1(b.s, FK_T_SEARCH);
    b.ep = b.set_ep[0];
    // the centroids read every row once; the search one row per distance
    b.prof_end(pe, FK_PROF_EP, center_rows * 4.0 * ds.d + (double)nd * 4.0 * ds.d);
    b.prof_flush();
    b.st->d_ep += nd;
    b.st->t_ep += now_sec() - t0;
    if (b.verbose) fprintf(stderr, "[fk] ep: %u\n", b.ep);
}

// ---------------------------------------------------------------------------
// Connectivity (tree_grow, src/kgraph.cpp:2145-2270).
//
// The reference sweeps from ep, then repeatedly takes the smallest unreached
// id u, links it from a root (the first reached pool entry of u with degree
// < nsg_R, else a random reached node with degree < nsg_R * factor) and
// sweeps again from u.  A sweep never leaves through a reached node (every
// out-edge of a reached node was followed when it was reached), so u's sweep
// reaches exactly the unreached nodes v whose smallest unreached reacher is
// u.  Hence one min-label propagation over the unreached subgraph, label(v) =
// min { w unreached : w reaches v }, replaces all sweeps: the attached ids are
// the v with label(v) == v, taken in ascending order, and "v was reached
// before u is attached" is "v reached from ep, or label(v) < u".  Only the
// root choices (a sequential walk over degrees) run on the host.
// ---------------------------------------------------------------------------
#define FK_REACHED 0xFFFFFFFFu

__global__ void k_mark_reached(const uint32_t *__restrict__ ids, uint32_t n, uint32_t *__restrict__ label) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) label[ids[i]] = FK_REACHED;
}

__global__ void k_bfs_expand(const uint32_t *__restrict__ nsg_id, const uint32_t *__restrict__ nsg_len,
                             uint32_t ncap, const uint32_t *__restrict__ front, uint32_t nf,
                             uint32_t *__restrict__ label, uint32_t *__restrict__ next, uint32_t *__restrict__ nnext) {
    const uint32_t w = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (w >= nf) return;
    const uint32_t f = front[w];
    const uint32_t deg = nsg_len[f];
    // one queue reservation per wave and chunk of 64 out-edges
    for (uint32_t k0 = 0; k0 < deg; k0 += 64) {
        const uint32_t k = k0 + lane_id();
        uint32_t t = 0;
        bool fresh = false;
        if (k < deg) {
            t = nsg_id[(uint64_t)f * ncap + k];
            fresh = atomicExch(&label[t], FK_REACHED) != FK_REACHED;
        }
        const uint64_t m = __ballot(fresh);
        if (m == 0) continue;
        const uint32_t lead = (uint32_t)__builtin_ctzll(m);
        uint32_t base = 0;
        if (lane_id() == lead) base = atomicAdd(nnext, (uint32_t)__builtin_popcountll(m));
        base = (uint32_t)__shfl((int)base, (int)lead, 64);
        if (fresh) next[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = t;
    }
}

__global__ void k_label_init(uint32_t *__restrict__ label, uint32_t N) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < N) label[v] = v;
}

// The nodes the search from the entry points did not reach, in any order
// (one atomic per wave).
__global__ void k_list_unreached(const uint32_t *__restrict__ label, uint32_t N, uint32_t *__restrict__ list,
                                 uint32_t *__restrict__ n_list) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    const bool un = v < N && label[v] != FK_REACHED;
    const uint64_t m = __ballot(un);
    if (m == 0) return;
    const uint32_t lead = (uint32_t)__builtin_ctzll(m);
    uint32_t base = 0;
    if (lane_id() == lead) base = atomicAdd(n_list, (uint32_t)__builtin_popcountll(m));
    base = (uint32_t)__shfl((int)base, (int)lead, 64);
    if (un) list[base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u))] = v;
}

// One wavefront per unreached node: push the node's label along its
// out-edges to unreached targets (atomicMin); *changed is set when any
// label dropped.  The fixed point (each label the smallest id that reaches
// the node in the unreached subgraph) does not depend on the list's order.
__global__ void k_label_prop(const uint32_t *__restrict__ nsg_id, const uint32_t *__restrict__ nsg_len,
                             const uint32_t *__restrict__ list, uint32_t n_list, uint32_t ncap,
                             uint32_t *__restrict__ label, uint32_t *__restrict__ changed) {
    const uint32_t w = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (w >= n_list) return;
    const uint32_t u = list[w];
    const uint32_t lu = label[u];
    if (lu == FK_REACHED) return;
    const uint32_t deg = nsg_len[u];
    bool ch = false;
    for (uint32_t k = lane_id(); k < deg; k += 64) {
        const uint32_t t = nsg_id[(uint64_t)u * ncap + k];
        const uint32_t lt = label[t];
        if (lt != FK_REACHED && lt > lu) {
            if (atomicMin(&label[t], lu) > lu) ch = true;
        }
    }
    if (wave_ballot(ch) && lane_id() == 0) *changed = 1u;
}

// Shortcut: a node labelled l takes l's label when smaller (every node
// that reaches l reaches the node through l), which halves the remaining
// path lengths per round; the fixed point is unchanged.
__global__ void k_label_jump(const uint32_t *__restrict__ list, uint32_t n_list, uint32_t *__restrict__ label,
                             uint32_t *__restrict__ changed) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    bool ch = false;
    if (w < n_list) {
        const uint32_t u = list[w];
        const uint32_t l = label[u];
        if (l != u) {
            const uint32_t ll = label[l];
            if (ll < l && atomicMin(&label[u], ll) > ll) ch = true;
        }
    }
    if (wave_ballot(ch) && lane_id() == 0) *changed = 1u;
}

// Pool ids (first cnt entries) of a list of nodes, for the root choices.
__global__ void k_gather_pool_ids(const uint64_t *__restrict__ pool, const uint32_t *__restrict__ cnt, uint32_t cap,
                                  const uint32_t *__restrict__ nodes, uint32_t n, uint32_t *__restrict__ out) {
    const uint32_t w = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (w >= n) return;
    const uint32_t v = nodes[w];
    const uint32_t c = cnt[v];
    uint32_t *o = out + (uint64_t)w * (cap + 1);
    if (lane_id() == 0) o[0] = c;
    for (uint32_t k = lane_id(); k < c; k += 64) o[1 + k] = key_id(pool[(uint64_t)v * cap + k]);
}

__global__ void k_set_edges(const float *__restrict__ rows, uint32_t stride, uint32_t dpad,
                            const uint32_t *__restrict__ src, const uint32_t *__restrict__ slot,
                            const uint32_t *__restrict__ dst, uint32_t n, uint32_t ncap, uint32_t *__restrict__ nsg_id,
                            float *__restrict__ nsg_d, uint32_t *__restrict__ nsg_len) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t p = (uint64_t)src[i] * ncap + slot[i];
    nsg_id[p] = dst[i];
    nsg_d[p] = l2_canon(rows + (size_t)src[i] * stride, rows + (size_t)dst[i] * stride, dpad);
    atomicMax(&nsg_len[src[i]], slot[i] + 1);
}

// The random fallback of findroot (src/kgraph.cpp:2174-2236): try t = 1, 2,
// ... draws a node of the set and takes it when it is reached and its
// out-degree is below nsg_R * (t / n + 1) (and below the out-list capacity).
// When no reached node of the set is below a band's bound (every out-list
// full, as a truncating prune leaves them) every draw of that band fails,
// which costs n draws per attach; such a band is skipped at once and the
// set's stream advanced past its draws (the stream is a counter), so the
// choice is the one the draw-by-draw loop makes.  low[f] counts the set's
// reached nodes below band f's bound: nodes leave a band when their degree
// reaches its bound, an attached component joins the counts once its root
// has been given its edge (later roots are larger ids, so it stays reached).
struct BandCounts {
    std::vector<uint32_t> thr, low;
    uint32_t set = 0xFFFFFFFFu;
    const uint32_t *lab = nullptr;
    const std::vector<uint32_t> *roots = nullptr;  // attached ids, ascending
    std::unique_ptr<uint32_t[]> kidx;              // root id -> its index in roots (written at roots only)
    std::vector<uint32_t> moff, mem;               // the set's unreached nodes grouped by root
    uint32_t k_lo = 0;                             // index of the set's first root
    void reset(uint32_t set_, const uint32_t *deg, uint32_t lo, uint32_t n, uint32_t R, uint32_t cap, uint32_t u,
               uint32_t N) {
        set = set_;
        thr.clear();
        for (uint32_t f = 1;; ++f) {
            const uint32_t t = (uint32_t)std::min<uint64_t>((uint64_t)R * f, cap);
            thr.push_back(t);
            if (t == cap) break;
        }
        low.assign(thr.size(), 0);
        const std::vector<uint32_t> &rs = *roots;
        if (!kidx) {
            kidx.reset(new uint32_t[N]);
            for (uint32_t k = 0; k < (uint32_t)rs.size(); ++k) kidx[rs[k]] = k;
        }
        k_lo = (uint32_t)(std::lower_bound(rs.begin(), rs.end(), lo) - rs.begin());
        const uint32_t k_hi = (uint32_t)(std::lower_bound(rs.begin(), rs.end(), lo + n) - rs.begin());
        moff.assign((size_t)(k_hi - k_lo) + 1, 0);
        // an unreached node's label is its component's smallest id, a root
        for (uint32_t v = lo; v < lo + n; ++v) {
            if (lab[v] == FK_REACHED || lab[v] < u) {
                for (size_t f = 0; f < thr.size(); ++f) low[f] += deg[v] < thr[f] ? 1u : 0u;
            } else {
                ++moff[kidx[lab[v]] - k_lo + 1];
            }
        }
        for (size_t k = 1; k < moff.size(); ++k) moff[k] += moff[k - 1];
        mem.resize(moff.back());
        std::vector<uint32_t> fill(moff.begin(), moff.end() - 1);
        for (uint32_t v = lo; v < lo + n; ++v)
            if (lab[v] != FK_REACHED && lab[v] >= u) mem[fill[kidx[lab[v]] - k_lo]++] = v;
    }
    void add_edge(uint32_t *deg, uint32_t v, uint32_t set_of_v) {  // deg[v] grows by one
        if (set_of_v == set)
            for (size_t f = 0; f < thr.size(); ++f) low[f] -= deg[v] + 1 == thr[f] ? 1u : 0u;
        deg[v]++;
    }
    void reach(const uint32_t *deg, uint32_t u, uint32_t set_of_u) {  // component u is now reached
        if (set_of_u != set) return;
        const uint32_t k = kidx[u] - k_lo;
        for (uint32_t e = moff[k]; e < moff[k + 1]; ++e)
            for (size_t f = 0; f < thr.size(); ++f) low[f] += deg[mem[e]] < thr[f] ? 1u : 0u;
    }
};

void tree_grow(BuildState &b) {
    const double t0 = now_sec();
    const uint32_t N = b.N;
    hipEvent_t pe = b.prof_begin();
    DevBuf<uint32_t> label(N), f0(N), f1(N), cnt(1);
    // 1. nodes reachable from ep are marked FK_REACHED (frontier BFS)
    hipLaunchKernelGGL(k_label_init, dim3((N + 255) / 256), dim3(256), 0, b.s, label.p, N);
    FK_LAUNCH_CHECK();
    // every set's entry point is a source (edges never leave a set)
    {
        std::vector<uint32_t> marks(b.n_sets, FK_REACHED);
        DevBuf<uint32_t> eps(b.n_sets);
        FK_HIP(hipMemcpyAsync(eps.p, b.set_ep.data(), (size_t)b.n_sets * 4, hipMemcpyHostToDevice, b.s));
        FK_HIP(hipMemcpyAsync(f0.p, b.set_ep.data(), (size_t)b.n_sets * 4, hipMemcpyHostToDevice, b.s));
        hipLaunchKernelGGL(k_mark_reached, dim3((b.n_sets + 255) / 256), dim3(256), 0, b.s, eps.p, b.n_sets, label.p);
        FK_LAUNCH_CHECK();
        b.sync();
    }
    uint32_t nf = b.n_sets, levels = 0;
    uint32_t *cur = f0.p, *nxt = f1.p;
    while (nf) {
        FK_HIP(hipMemsetAsync(cnt.p, 0, 4, b.s));
        hipLaunchKernelGGL(k_bfs_expand, dim3((nf + 3) / 4), dim3(256), 0, b.s, b.nsg_id.p, b.nsg_len.p, b.ncap, cur,
                           nf, label.p, nxt, cnt.p);
        FK_LAUNCH_CHECK();
        FK_HIP(hipMemcpyAsync(&nf, cnt.p, 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
        std::swap(cur, nxt);
        ++levels;
    }
    const double t_bfs = now_sec();
    // 2. min-label propagation over the unreached subgraph, whose nodes are
    // listed once (f0 is free after the search)
    FK_HIP(hipMemsetAsync(cnt.p, 0, 4, b.s));
    hipLaunchKernelGGL(k_list_unreached, dim3((N + 255) / 256), dim3(256), 0, b.s, label.p, N, f0.p, cnt.p);
    FK_LAUNCH_CHECK();
    uint32_t n_un = 0;
    FK_HIP(hipMemcpyAsync(&n_un, cnt.p, 4, hipMemcpyDeviceToHost, b.s));
    b.sync();
    uint32_t rounds = 0;
    while (n_un) {
        FK_HIP(hipMemsetAsync(cnt.p, 0, 4, b.s));
        hipLaunchKernelGGL(k_label_prop, dim3((n_un + 3) / 4), dim3(256), 0, b.s, b.nsg_id.p, b.nsg_len.p, f0.p,
                           n_un, b.ncap, label.p, cnt.p);
        FK_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_label_jump, dim3((n_un + 255) / 256), dim3(256), 0, b.s, f0.p, n_un, label.p, cnt.p);
        FK_LAUNCH_CHECK();
        uint32_t ch = 0;
        FK_HIP(hipMemcpyAsync(&ch, cnt.p, 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
        ++rounds;
        if (!ch) break;
    }
    // labels and out-degrees into the handle's pinned buffers (grow-only)
    b.h->host_lab.ensure(N);
    b.h->host_deg.ensure(N);
    FK_HIP(hipMemcpyAsync(b.h->host_lab.p, label.p, (size_t)N * 4, hipMemcpyDeviceToHost, b.s));
    FK_HIP(hipMemcpyAsync(b.h->host_deg.p, b.nsg_len.p, (size_t)N * 4, hipMemcpyDeviceToHost, b.s));
    b.sync();
    uint32_t *lab = b.h->host_lab.p, *deg = b.h->host_deg.p;
    // the search reads every node's out-list and its targets' labels once;
    // each propagation round those of the unreached nodes
    b.prof_end(pe, FK_PROF_CONN, ((double)N + (double)rounds * n_un) * (8.0 + 8.0 * b.P.nsg_R));
    b.prof_flush();
    const double t_lab = now_sec();
    std::vector<uint32_t> roots_of;  // attached ids, ascending
    for (uint32_t v = 0; v < N; ++v)
        if (lab[v] == v) roots_of.push_back(v);
    BandCounts bands;
    bands.lab = lab;
    bands.roots = &roots_of;
    const uint32_t na = (uint32_t)roots_of.size();
    std::vector<uint32_t> e_src, e_slot, e_dst;
    double t_loop = 0;
    uint64_t n_draws = 0;  // random fallback draws
    if (na) {
        // 3. pool rows of the attached ids, one gather
        const uint32_t rs = b.cap + 1;
        DevBuf<uint32_t> dnodes(na), drows((size_t)na * rs);
        PinnedBuf<uint32_t> &rows = b.h->host_rows;
        rows.ensure((size_t)na * rs);
        FK_HIP(hipMemcpyAsync(dnodes.p, roots_of.data(), (size_t)na * 4, hipMemcpyHostToDevice, b.s));
        hipLaunchKernelGGL(k_gather_pool_ids, dim3((na + 3) / 4), dim3(256), 0, b.s, b.pool.p, b.cnt.p, b.cap, dnodes.p,
                           na, drows.p);
        FK_LAUNCH_CHECK();
        FK_HIP(hipMemcpyAsync(rows.p, drows.p, (size_t)na * rs * 4, hipMemcpyDeviceToHost, b.s));
        b.sync();
        const double t_rows = now_sec();
        // 4. root choices in ascending order of the attached ids.  The band
        // counts of a set are made once it has had 4096 roots (or a long
        // fallback): then a root's pool row is scanned only while some
        // reached node of the set is below nsg_R, since no entry can take
        // the edge otherwise (a truncating prune leaves every out-list full)
        uint32_t set_run = 0xFFFFFFFFu, in_set = 0;
        for (uint32_t k = 0; k < na; ++k) {
            const uint32_t u = roots_of[k];
            auto reached = [&](uint32_t v) { return lab[v] == FK_REACHED || lab[v] < u; };
            const uint32_t *prow = rows.p + (size_t)k * rs;
            uint32_t root = N;
            const uint32_t set = b.set_of_node(u), lo = b.set_lo[set], ns = b.set_n(set);
            in_set = set == set_run ? in_set + 1 : 1;
            set_run = set;
            if (in_set == 4096 && bands.set != set) bands.reset(set, deg, lo, ns, b.P.nsg_R, b.ncap, u, N);
            if (bands.set != set || bands.low[0] > 0)
                for (uint32_t i = 0; i < prow[0]; ++i) {
                    const uint32_t p = prow[1 + i];
                    if (p < N && reached(p) && deg[p] < b.P.nsg_R) { root = p; break; }
                }
            if (root == N) {
                // a random reached node of u's set, with the set's stream; the
                // band counts (a pass over the set) are made once a fallback
                // runs long: most end in a few draws
                uint64_t tries = 0;
                while (true) {
                    if (tries == 4096 && bands.set != set) bands.reset(set, deg, lo, ns, b.P.nsg_R, b.ncap, u, N);
                    const uint64_t f = (tries + 1) / ns;  // band of the next draw (factor f + 1)
                    const bool counted = bands.set == set;
                    if (counted && f >= bands.thr.size() - 1 && bands.low.back() == 0)
                        throw Error(FK_E_INTERNAL, "tree_grow: no node can take an edge");
                    if (counted && f < bands.thr.size() - 1 && bands.low[f] == 0) {
                        // no draw of this band can succeed: skip to the next band
                        const uint64_t skip = (f + 1) * ns - 1 - tries;
                        b.set_rng[set] += skip * 0x9E3779B97F4A7C15ull;  // next_rand's counter step
                        tries += skip;
                        continue;
                    }
                    ++tries;
                    ++n_draws;
                    const uint64_t factor = tries / ns + 1;
                    const uint32_t rid = lo + (uint32_t)(b.next_rand(set) % ns);
                    if (reached(rid) && deg[rid] < b.ncap && deg[rid] < b.P.nsg_R * factor) { root = rid; break; }
                    if (tries > 64ull * ns) throw Error(FK_E_INTERNAL, "tree_grow: no node can take an edge");
                }
            }
            e_src.push_back(root);
            e_slot.push_back(deg[root]);
            e_dst.push_back(u);
            bands.add_edge(deg, root, set);
            bands.reach(deg, u, set);
        }
        t_loop = now_sec() - t_rows;
        DevBuf<uint32_t> s(na), sl(na), d(na);
        FK_HIP(hipMemcpyAsync(s.p, e_src.data(), (size_t)na * 4, hipMemcpyHostToDevice, b.s));
        FK_HIP(hipMemcpyAsync(sl.p, e_slot.data(), (size_t)na * 4, hipMemcpyHostToDevice, b.s));
        FK_HIP(hipMemcpyAsync(d.p, e_dst.data(), (size_t)na * 4, hipMemcpyHostToDevice, b.s));
        hipLaunchKernelGGL(k_set_edges, dim3((na + 255) / 256), dim3(256), 0, b.s, b.ds->rows.p, b.ds->stride,
                           b.ds->dpad, s.p, sl.p, d.p, na, b.ncap, b.nsg_id.p, b.nsg_d.p, b.nsg_len.p);
        FK_LAUNCH_CHECK();
        b.sync();
        b.st->d_conn += na;
    }
    b.st->n_attached += na;
    uint32_t mx = b.width;
    for (uint32_t i = 0; i < N; ++i) mx = std::max(mx, deg[i]);
    b.width = mx;
    for (uint32_t k = 0; k < b.n_sets; ++k)
        for (uint32_t i = b.set_lo[k]; i < b.set_lo[k + 1]; ++i) b.set_width[k] = std::max(b.set_width[k], deg[i]);
    const double dt = now_sec() - t0;
    const uint32_t li = b.loopcount;
    if (li < FK_MAX_LOOPS) b.st->tree_grow_time[li] = dt;
    b.st->t_conn += dt;
    if (b.verbose)
        fprintf(stderr,
                "[fk] tree_grow: attached %u (%.3f s; %u BFS levels %.3f s, %u label rounds %.3f s, root choices %.3f s, "
                "%llu random draws)\n",
                na, dt, levels, t_bfs - t0, rounds, t_lab - t_bfs, t_loop, (unsigned long long)n_draws);
}

}  // namespace fk

using namespace fk;

extern "C" fk_status fk_search(fk_handle *h, const uint64_t *offsets, const uint32_t *ids, uint32_t n, uint32_t ep,
                               const float *queries, uint32_t n_q, uint32_t search_L, uint32_t K, uint32_t *out_ids,
                               uint64_t *n_dist) {
    if (!h) return FK_E_INVALID;
    return guard(h, [&] {
        require_dataset(h);
        FK_CHECK_ARG(offsets && queries && out_ids, "null pointer");
        FK_CHECK_ARG(n >= 1 && n <= h->ds.n_total && ep < n, "graph size / entry point out of range");
        FK_CHECK_ARG(search_L >= K && K >= 1, "need 1 <= K <= search_L");
        const uint64_t ne = offsets[n];
        for (uint64_t i = 0; i < ne; ++i) FK_CHECK_ARG(ids[i] < n, "graph id out of range");
        if (n_dist) *n_dist = 0;
        if (n_q == 0) return;
        select_device(h);
        const Dataset &ds = h->ds;
        DevBuf<uint64_t> doff((size_t)n + 1);
        DevBuf<uint32_t> dids(ne);
        FK_HIP(hipMemcpyAsync(doff.p, offsets, ((size_t)n + 1) * 8, hipMemcpyHostToDevice, h->stream));
        if (ne) FK_HIP(hipMemcpyAsync(dids.p, ids, ne * 4, hipMemcpyHostToDevice, h->stream));
        const uint32_t words = (n + 31) / 32;
        uint32_t qb = (uint32_t)std::min<uint64_t>(n_q, std::max<uint64_t>(1, (512ull << 20) / ((uint64_t)words * 4)));
        DevBuf<uint32_t> vis((size_t)qb * words);
        DevBuf<float> dq((size_t)qb * ds.dpad);
        DevBuf<uint32_t> dout((size_t)qb * K);
        Tally cnt;
        cnt.ensure(1);
        cnt.clear(h->stream, 0);
        for (uint32_t q0 = 0; q0 < n_q; q0 += qb) {
            const uint32_t nq = std::min(qb, n_q - q0);
            FK_HIP(hipMemsetAsync(vis.p, 0, (size_t)nq * words * 4, h->stream));
            upload_rows(h, dq.p, ds.dpad, queries + (size_t)q0 * ds.d, (size_t)ds.d * 4, nq, ds.d);
            BeamArgs a{};
            a.rows = ds.rows.p;
            a.stride = ds.stride;
            a.dpad = ds.dpad;
            a.N = n;
            a.queries = dq.p;
            a.n_q = nq;
            a.L = search_L;
            a.K = K;
            a.init_mode = 1;
            a.ep = ep;
            a.g_ids = dids.p;
            a.g_off = doff.p;
            a.visited = vis.p;
            a.seed = FK_SEARCH_SEED;
            a.q_base = q0;
            a.out_ids = dout.p;
            a.n_dist = cnt.at(0);
            launch_beam(h->stream, a);
            FK_HIP(hipMemcpyAsync(out_ids + (size_t)q0 * K, dout.p, (size_t)nq * K * 4, hipMemcpyDeviceToHost,
                                  h->stream));
            FK_HIP(hipStreamSynchronize(h->stream));
        }
        const unsigned long long c = cnt.read1(h->stream, 0);
        if (n_dist) *n_dist = c;
    });
}
