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
#include <hipcub/hipcub.hpp>

#include "fk_build.h"

namespace fk {

#define FK_ORDER_FANOUT 8
#define FK_ORDER_HOPS_FINE 2
#define FK_ORDER_HOPS_COARSE 6

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
    // pool heads read once per hop (plus the labels they point at), keys
    // written and sorted (radix passes of 8 bits over key + value)
    const double passes = (2.0 * bits + 7) / 8;
    b.prof_end(pe, FK_PROF_ORDER, (double)N * (FK_ORDER_HOPS_COARSE * (8.0 * FK_ORDER_FANOUT + 4.0 * (FK_ORDER_FANOUT + 2)) +
                                               12.0 + passes * 24.0));
    b.has_order = true;
}

}  // namespace fk
