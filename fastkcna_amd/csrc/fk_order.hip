// Locality order of the nodes of a build.
//
// The per-node kernels (pruning, BridgeView rounds, the NN-descent local
// join) gather the rows of a node's pool members: about 500 random 512-byte
// rows per node at d = 128, with no reuse between consecutive node ids.
// Nodes whose pools overlap are placed next to each other here; a kernel that
// walks this order with the XCD-contiguous block mapping (xcd_position) then
// has each XCD work through one run of neighbouring nodes, whose candidate
// rows its L2 already holds.
//
// The order is a sort of the nodes by a two-level key, both levels labels
// of min-label propagation over each pool's first FK_ORDER_FANOUT entries:
//   coarse = smallest id reachable within FK_ORDER_HOPS_COARSE hops,
//   fine   = smallest id reachable within FK_ORDER_HOPS_FINE hops,
// so nodes of one tight group share the coarse label and, inside a group,
// nodes of one sub-neighbourhood share the fine label.  The order only
// changes which blocks run together: every kernel that uses it produces the
// same results for any order.
//
// A batched build's sort is set-major (edges never leave a set, so neither do
// labels).  Its order is then interleaved by set: the XCD-contiguous mapping
// gives XCD x the x-th eighth of every set in turn, so the eight XCDs work on
// one set at a time, as they do in a build of that set alone, and the rows
// they share stay in the last-level cache.
#include <hipcub/hipcub.hpp>

#include "fk_build.h"

namespace fk {

#define FK_ORDER_FANOUT 8
#define FK_ORDER_HOPS_FINE 2
#define FK_ORDER_HOPS_COARSE 6
#ifndef FK_ORDER_INTERLEAVE
#define FK_ORDER_INTERLEAVE 1
#endif

// One Jacobi step of min-label propagation along pool edges.
__global__ void k_label_hop(const uint64_t *__restrict__ pool, const uint32_t *__restrict__ cnt, uint32_t cap,
                            uint32_t N, uint32_t fanout, const uint32_t *__restrict__ lin,
                            uint32_t *__restrict__ lout) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    const uint32_t c = cnt[v];
    const uint32_t m = c < fanout ? c : fanout;
    const uint64_t *p = pool + (uint64_t)v * cap;
    uint32_t l = lin ? lin[v] : v;
    for (uint32_t j = 0; j < m; ++j) {
        const uint32_t u = key_id(p[j]);
        if (u < N) {
            const uint32_t lu = lin ? lin[u] : u;
            l = lu < l ? lu : l;
        }
    }
    lout[v] = l;
}

__global__ void k_order_keys(const uint32_t *__restrict__ coarse, const uint32_t *__restrict__ fine, uint32_t N,
                             uint32_t bits, uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= N) return;
    keys[v] = ((uint64_t)coarse[v] << bits) | (uint64_t)fine[v];
    vals[v] = v;
}

// out[q] = in[src[k] + q - dst[k]] for the slice k holding q (dst ascending).
__global__ void k_slice_gather(const uint32_t *__restrict__ in, uint32_t N, const uint32_t *__restrict__ dst,
                               const uint32_t *__restrict__ src, uint32_t n_slices, uint32_t *__restrict__ out) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= N) return;
    uint32_t lo = 0, hi = n_slices;  // last slice with dst <= q
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (dst[mid] <= q) lo = mid;
        else hi = mid;
    }
    out[q] = in[src[lo] + q - dst[lo]];
}

static void interleave_sets(BuildState &b, uint32_t *tmp) {
    const uint32_t S = b.n_sets, N = b.N;
    std::vector<uint32_t> h((size_t)2 * FK_NUM_XCD * S);
    uint32_t *dst = h.data(), *src = dst + (size_t)FK_NUM_XCD * S;
    uint32_t at = 0, k = 0;
    for (uint32_t x = 0; x < FK_NUM_XCD; ++x)
        for (uint32_t s = 0; s < S; ++s, ++k) {
            const uint64_t n = b.set_n(s);
            const uint32_t a = b.set_lo[s] + (uint32_t)(n * x / FK_NUM_XCD);
            const uint32_t e = b.set_lo[s] + (uint32_t)(n * (x + 1) / FK_NUM_XCD);
            dst[k] = at;
            src[k] = a;
            at += e - a;
        }
    DevBuf<uint32_t> d(h.size());
    FK_HIP(hipMemcpyAsync(d.p, h.data(), h.size() * 4, hipMemcpyHostToDevice, b.s));
    hipLaunchKernelGGL(k_slice_gather, dim3((N + 255) / 256), dim3(256), 0, b.s, b.order.p, N, d.p,
                       d.p + (size_t)FK_NUM_XCD * S, FK_NUM_XCD * S, tmp);
    FK_LAUNCH_CHECK();
    FK_HIP(hipMemcpyAsync(b.order.p, tmp, (size_t)N * 4, hipMemcpyDeviceToDevice, b.s));
    b.sync();  // the slice table is released on return
}

void locality_order(BuildState &b) {
    const uint32_t N = b.N;
    if (N == 0) return;
    const unsigned g = (N + 255) / 256;
    hipEvent_t pe = b.prof_begin();
    b.order.ensure(N);
    b.ord_lab.ensure(3ull * N);
    uint32_t *la = b.ord_lab.p, *lb = la + N, *fine = lb + N;
    const uint32_t *cur = nullptr;
    for (uint32_t h = 0; h < FK_ORDER_HOPS_COARSE; ++h) {
        uint32_t *dst = (h & 1) ? lb : la;
        hipLaunchKernelGGL(k_label_hop, dim3(g), dim3(256), 0, b.s, b.pool.p, b.cnt.p, b.cap, N,
                           (uint32_t)FK_ORDER_FANOUT, cur, dst);
        FK_LAUNCH_CHECK();
        cur = dst;
        if (h + 1 == FK_ORDER_HOPS_FINE)
            FK_HIP(hipMemcpyAsync(fine, cur, (size_t)N * 4, hipMemcpyDeviceToDevice, b.s));
    }
    uint32_t bits = 1;
    while (bits < 32 && (1ull << bits) < N) ++bits;
    b.ord_key.ensure(2ull * N);
    b.ord_val.ensure(N);
    hipLaunchKernelGGL(k_order_keys, dim3(g), dim3(256), 0, b.s, cur, fine, N, bits, b.ord_key.p, b.ord_val.p);
    FK_LAUNCH_CHECK();
    size_t tmp = 0;
    FK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, b.ord_key.p, b.ord_key.p + N, b.ord_val.p, b.order.p,
                                              (int)N, 0, (int)(2 * bits), b.s));
    b.ord_tmp.ensure(tmp);
    FK_HIP(hipcub::DeviceRadixSort::SortPairs(b.ord_tmp.p, tmp, b.ord_key.p, b.ord_key.p + N, b.ord_val.p,
                                              b.order.p, (int)N, 0, (int)(2 * bits), b.s));
    if (FK_ORDER_INTERLEAVE && b.n_sets > 1) interleave_sets(b, b.ord_val.p);
    // pool heads read once per hop (plus the labels they point at), keys
    // written and sorted (radix passes of 8 bits over key + value)
    const double passes = (2.0 * bits + 7) / 8;
    b.prof_end(pe, FK_PROF_ORDER, (double)N * (FK_ORDER_HOPS_COARSE * (8.0 * FK_ORDER_FANOUT + 4.0 * (FK_ORDER_FANOUT + 2)) +
                                               12.0 + passes * 24.0));
    b.has_order = true;
}

}  // namespace fk
