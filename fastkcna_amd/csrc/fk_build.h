// Device state of one graph build and the phase entry points shared by the
// NN-descent, loop and search translation units.
//
// Layout in HBM (structure of arrays, N = nodes of the build):
//   pool      u64[N][cap]   packed (dist bits << 32 | id), each node's first
//                           cnt[i] entries sorted ascending; cap = MAXL
//   pflg      u8 [N][cap]   bit0 = flag (not yet expanded), bit1 = isnew
//   cnt       u32[N]        valid pool entries (the reference's Nhood::L)
//   nsg_id    u32[N][ncap]  pruned out-lists, ncap = nsg_R + slack for the
//   nsg_d     f32[N][ncap]  connectivity edges; nsg_len u32[N]
//   nsg_new   u8 [N][ncap]  isnew bit of each out-list entry
//   old_id    u32[N][ncap]  previous loop's out-list (UpdateNhoods)
#pragma once
#include "fk_device.h"
#include "fk_internal.h"

#include <algorithm>

namespace fk {

#define FK_FLG_NEW 0x1u
#define FK_FLG_ISNEW 0x2u
#define FK_POOL_MAX 1024   // largest supported MAXL
#define FK_LIST_MAX 2048   // largest supported NN-descent sample list
#define FK_NSG_SLACK 64    // out-list room for connectivity edges
// tally counters of a build (BuildState::tally)
#define FK_T_PRUNE_CC 0    // prune: distances
#define FK_T_PRUNE_EXAM 1  // prune: examined candidates
#define FK_T_BV_CC 2       // BridgeView round: distances the reference evaluates
#define FK_T_BV_COMP 3     // BridgeView round: distances actually computed (slot after FK_T_BV_CC)
#define FK_T_SEARCH 4      // entry-point search: distances
#define FK_T_RECHECK 5     // Gram join: FP32 rechecks
#define FK_T_PRUNE_SKIP 6  // prune: distances counted but not computed (tau rule, pd <= 3 tau)
#define FK_T_BV_REJ 7      // BridgeView round: targets rejected by the 16-bit bound (no FP32 distance)
#define FK_T_SLOTS 8

struct BuildParams {
    // effective (shrunk) parameters, src/kgraph.cpp:1288-1334
    uint32_t K, L, S, R, iterations, controls, seed;
    float delta, recall;
    uint32_t nsg_R, search_L, search_K, step, massq_S, loop_i;
    float tau, angle;
    uint32_t flags;
    uint32_t MAXL;
};

BuildParams shrink_params(const fk_params &p, uint32_t N);

// Learned distance screen of one build (fk_screen.hip).
struct ScreenState {
    bool ready = false;
    DevBuf<float> w, g, m, v;             // FP32 weights, gradient, Adam moments
    DevBuf<float> wi;                     // inference weights (BF16-rounded MLP)
    DevBuf<uint16_t> z;                   // BF16 projections, N x 32
    DevBuf<unsigned long long> cnt;       // screened, skipped, audited, missed
    DevBuf<float> rho;                    // recalibration residuals
    DevBuf<uint32_t> rcnt;
    float mu = 1.f, margin = 0.f;
};

struct BuildState {
    fk_handle *h = nullptr;
    hipStream_t s = nullptr;
    Dataset *ds = nullptr;
    uint32_t N = 0;        // oracle.size()
    uint32_t cap = 0;      // pool capacity (MAXL)
    uint32_t ncap = 0;     // out-list capacity
    BuildParams P{};
    fk_stats *st = nullptr;
    bool verbose = false;
    double t0 = 0;         // build start (seconds, steady clock)

    DevBuf<uint64_t> pool;
    DevBuf<uint8_t> pflg;
    DevBuf<uint32_t> cnt;
    DevBuf<float> thr;       // join filter threshold per node
    DevBuf<uint32_t> M;
    DevBuf<float> radiusM;
    DevBuf<uint8_t> found;

    // sampling lists: [fwd new | rev new | fwd old | rev old] per node
    DevBuf<uint64_t> list_off;  // N + 1
    DevBuf<uint32_t> n_new;     // N
    DevBuf<uint32_t> lists;
    DevBuf<uint64_t> pair_off;  // N + 1
    DevBuf<uint32_t> pair_cnt;  // scratch N
    DevBuf<float> pdist;
    DevBuf<uint32_t> resv;
    DevBuf<uint32_t> bcnt;      // N
    DevBuf<uint64_t> boff;      // N + 1
    DevBuf<uint32_t> ord_cnt;   // N: bucket sizes in locality order
    DevBuf<uint64_t> ord_off;   // N + 1: their exclusive scan
    DevBuf<uint64_t> bkey;
    DevBuf<uint32_t> bsrc;
    // reverse-sample buckets
    DevBuf<uint32_t> fcnt_new, fcnt_old, rcnt_new, rcnt_old;  // N each
    DevBuf<uint64_t> roff_new, roff_old;                       // N + 1
    DevBuf<uint32_t> rbuf_new, rbuf_old;
    uint32_t tmax = 0;          // longest sample list of the current iteration
    ScreenState scr;

    // sampled ground truth (GenerateControl)
    uint32_t n_ctrl = 0;
    DevBuf<uint32_t> ctrl_ids;   // n_ctrl
    DevBuf<uint32_t> ctrl_gt;    // n_ctrl x MAXL ids
    DevBuf<unsigned long long> acc;  // reduction scratch
    // per-wave counters of the big launches (slots FK_T_*)
    Tally tally;

    // pruned graph
    DevBuf<uint32_t> nsg_id, nsg_len, old_id, old_len;
    DevBuf<float> nsg_d;
    DevBuf<uint8_t> nsg_new;
    uint32_t ep = 0;
    uint32_t width = 0;
    uint32_t loopcount = 0;
    std::vector<float> center;  // Calcenter over all resident rows (cached)

    // Independent sets built together (fk_build_batch): set s is the nodes
    // [set_lo[s], set_lo[s + 1]); every per-set decision (random draws, stop
    // rule, BridgeView rounds, entry point, connectivity) is made per set,
    // so each set's graph is the one a build of its rows alone gives.
    // fk_build is the case n_sets == 1 (set_of empty).
    uint32_t n_sets = 1;
    std::vector<uint32_t> set_lo;       // host, n_sets + 1
    DevBuf<uint32_t> d_set_lo;          // device copy (n_sets > 1)
    DevBuf<uint32_t> set_of;            // node -> set (n_sets > 1)
    std::vector<uint64_t> set_rng;      // per set: the reference's global rand() stream
    std::vector<uint32_t> set_ep;       // per set: entry point (node id)
    std::vector<uint32_t> set_width;    // per set: width (max out-degree seen)
    std::vector<uint8_t> set_active;    // per set: NN-descent still iterating
    DevBuf<uint8_t> d_set_active;
    std::vector<uint32_t> ctrl_lo;      // controls of set s: [ctrl_lo[s], ctrl_lo[s + 1])
    DevBuf<unsigned long long> set_acc; // per-set reduction scratch (2 per set)
    uint32_t set_n(uint32_t s) const { return set_lo[s + 1] - set_lo[s]; }
    uint32_t set_of_node(uint32_t v) const {  // host lookup
        return (uint32_t)(std::upper_bound(set_lo.begin(), set_lo.end(), v) - set_lo.begin()) - 1;
    }
    const uint32_t *dev_set_lo() const { return n_sets > 1 ? d_set_lo.p : nullptr; }
    const uint32_t *dev_set_of() const { return n_sets > 1 ? set_of.p : nullptr; }

    // locality order (fk_order.hip): node at each position of the order
    DevBuf<uint32_t> order;     // N
    bool has_order = false;
    DevBuf<uint32_t> ord_lab, ord_val;
    DevBuf<uint64_t> ord_key;
    DevBuf<uint8_t> ord_tmp;

    // scratch
    DevBuf<uint64_t> scal;      // 16 u64
    DevBuf<uint64_t> scan_tmp;  // prefix-scan partials

    // kernel-family timing (FK_FLAG_PROFILE)
    bool prof = false;
    struct ProfRec {
        hipEvent_t a, b;
        int fam;
        double bytes;
    };
    std::vector<hipEvent_t> ev_free;
    std::vector<ProfRec> ev_pending;
    hipEvent_t ev_get();
    // Returns a start event recorded on the build stream (nullptr when off).
    hipEvent_t prof_begin();
    void prof_end(hipEvent_t start, int fam, double bytes);
    void prof_set_last_bytes(double bytes);  // bytes of the last prof_end, known later
    DevBuf<uint32_t> bv_mark;  // profiling: nodes whose out-list a round reads
    void prof_flush();
    ~BuildState();

    uint64_t next_rand(uint32_t set = 0) {
        uint64_t &r = set_rng[set];
        r += 0x9E3779B97F4A7C15ull;
        uint64_t x = r;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        return x ^ (x >> 31);
    }
    // Sets up the set bookkeeping (n_sets >= 1; lo has n_sets + 1 entries
    // and covers [0, N)).
    void init_sets(const std::vector<uint32_t> &lo);
    void sync();
};

// Phases (fk_nnd.hip)
void nnd_alloc(BuildState &b);
void nnd_controls(BuildState &b);
void nnd_run(BuildState &b);            // init + iterations (constructor)
// recall over controls at K (all controls; per set into *per_set if given)
float nnd_evaluate(BuildState &b, uint32_t K, std::vector<float> *per_set = nullptr);
// stop-rule statistics of the current pools over given controls (one set)
void nnd_stop_stats(BuildState &b, uint32_t K, uint32_t n_ctrl, const uint32_t *d_ctrl, const uint32_t *d_gt,
                    uint32_t gt_stride, float *delta, float *recall);

// Pass 1 of the local join in Gram form (fk_gram.hip); rechecked pairs are
// added to *n_recheck (device counter)
void launch_join_gram(BuildState &b, uint32_t tcap, unsigned long long *n_recheck);

// Learned distance screen (fk_screen.hip): training on the first join's
// pairs, the screened join pass 1, and the per-iteration counters.
void screen_train(BuildState &b);
void screen_calibrate(BuildState &b);
void launch_join_screen(BuildState &b, uint32_t tcap);
size_t join_screen_lds(uint32_t tcap);
#define FK_JOIN_SCREEN_LDS_MAX (64 * 1024)  // longer lists join unscreened
void screen_collect(BuildState &b);

// Pool merge and BridgeView (fk_nnd.hip)
void bridgeview(BuildState &b);

// Locality order of the nodes from the current pools (fk_order.hip)
void locality_order(BuildState &b);

// Node of this block for one-block-per-node kernels: with an order, XCD x
// walks the contiguous run x of the order; without one, block = node.
// Launch with node_grid(b) blocks; blocks past N return.
__device__ __forceinline__ uint32_t node_of_block(const uint32_t *order, uint32_t N) {
    if (!order) return blockIdx.x;
    const uint32_t pos = xcd_position((N + FK_NUM_XCD - 1) / FK_NUM_XCD);
    return pos < N ? order[pos] : 0xFFFFFFFFu;
}
// Per-row norms when the build may take distances in the exact form
// (FK_FLAG_GRAM on exact-form rows, fk_gram.hip), else nullptr.
inline const float2 *exact_norms(const BuildState &b) {
    return ((b.P.flags & FK_FLAG_GRAM) && b.ds->gram_fmt == FK_GRAM_EXACT) ? b.ds->gram_nrm.p : nullptr;
}
inline unsigned node_grid(const BuildState &b) {
    return b.has_order ? (b.N + FK_NUM_XCD - 1) / FK_NUM_XCD * FK_NUM_XCD : b.N;
}

// Loops (fk_build.hip)
void build_loops(BuildState &b, fk_pg_type type);

// Search / graph utilities (fk_search.hip)
void compute_ep(BuildState &b, const uint32_t *init_nodes = nullptr);
void column_mean(hipStream_t s, const Dataset &ds, float *center, uint32_t r0 = 0, uint32_t n = 0);
// means of `count` row ranges [r0[k], r0[k] + n[k]) into center[k * dpad] (host arrays)
void column_means(hipStream_t s, const Dataset &ds, float *center, const uint32_t *r0, const uint32_t *n,
                  uint32_t count);
void tree_grow(BuildState &b);

double now_sec();

}  // namespace fk
