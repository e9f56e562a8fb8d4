/*
 * fk_api.h -- C ABI of the MI355X (gfx950) graph-construction engine.
 *
 * Plain C, no C++ or torch types: pointers, sizes and a status code.  One
 * handle owns one HIP stream and one resident dataset on one GPU.  Calls on a
 * handle are serialised and a handle is not thread-safe; separate handles are
 * independent.  Every entry point returns an fk_status and never throws; the
 * text of the last failure is kept per handle (fk_last_error).
 *
 * Reference interfaces replaced (paths relative to the FastKCNA sources):
 *   fk_build            KGraphImpl::buildFastNSG / buildFastTauMNG /
 *                       buildFastAlphaPG / buildFastHNSW
 *                       (src/kgraph.h:230-233, src/kgraph.cpp:3720-3966)
 *   fk_dataset_upload   MatrixOracle<float, metric::l2sqr> construction over a
 *                       Matrix<float> (src/kgraph-data.h:349-372); the oracle's
 *                       size()/set_size() prefix is fk_build's n_used
 *   fk_pair_distances   IndexOracle::operator()(i, j)
 *                       (src/kgraph.h:72, src/kgraph-data.h:392-395 ->
 *                       src/metric.cpp:45-78)
 *   fk_join_distances   the local join's distance calls, KGraphConstructor::join
 *                       + Nhood::join (src/kgraph.cpp:970-987, 1097-1117)
 *   fk_prune            nsg_sync_prune / taumg_sync_prune / alphapg_sync_prune
 *                       and their *_reversed_prune variants
 *                       (src/kgraph.cpp:1809-1851, 2012-2048, 2833-2876,
 *                       3038-3083, 3157-3203, 3374-3422)
 *   fk_nndescent        the KGraphConstructor constructor's NN-descent loop
 *                       (src/kgraph.cpp:1281-1459)
 *   fk_knn_bruteforce   LinearSearch / GenerateControl (src/kgraph.cpp:309-387)
 *   fk_search           get_neighbors(const float*, ...) beam search
 *                       (src/kgraph.cpp:1560-1623), batched over queries
 */
#ifndef FK_API_H
#define FK_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FK_OK = 0,
    FK_E_INVALID = 1,   /* bad argument; maps to kgraph::invalid_argument   */
    FK_E_NOMEM = 2,     /* device or host allocation failed                */
    FK_E_DEVICE = 3,    /* HIP runtime / launch failure                    */
    FK_E_IO = 4,        /* file error; maps to kgraph::io_error            */
    FK_E_INTERNAL = 5,  /* invariant broken inside the library             */
    FK_E_NODEVICE = 6   /* no usable gfx950 device                         */
} fk_status;

/* Proximity-graph family; values equal the reference's PG_type codes
 * (src/kgraph.cpp:3741, 3816, 3874, 3931). */
typedef enum {
    FK_PG_NSG = 1,
    FK_PG_HNSW_LAYER = 2,
    FK_PG_TAUMNG = 3,
    FK_PG_ALPHAPG = 4
} fk_pg_type;

/* Flags (fk_params.flags). */
#define FK_FLAG_VERBOSE 0x1u     /* per-iteration log lines on stderr            */
#define FK_FLAG_NO_CONTROLS 0x2u /* skip the sampled ground truth (recall = 0)   */
#define FK_FLAG_PROFILE 0x4u     /* time kernel families with HIP events         */
/* Gram-form distances on the MFMA units (16-bit operands, FP32 accumulation
 * and norms); the graph equals the plain FP32 build's.  The operand format is
 * picked from the data (fk_stats.gram_mode) unless forced.  Local join: in
 * Gram form on exact-form rows (FK_GRAM_EXACT), and on the bounded forms only
 * when FK_FLAG_GRAM_BF16 / _F16 force one (every pair that could enter a
 * pool recomputed in exact FP32 -- on float rows a third to a half of the
 * pairs, which makes the plain join the faster one there).  Prunes: prune
 * variant 3 (below) becomes available.  On exact-form rows the prunes and
 * BridgeView also take their distances as |a|^2 + |b|^2 - 2 a.b, which is
 * the FP32 distance bit for bit there (fk_prune honours the flag too).  On
 * FP16-form rows of width 96, 128 or 256, BridgeView first bounds each
 * target's distance from a low-precision copy (8-bit integers with a scale
 * per row at widths 96 and 256, the 16-bit operands at 128) and computes the FP32
 * distance only when the bound cannot put it beyond the pool
 * (fk_stats.d_bv_rejected).  Memory: a 16-bit copy of the rows, plus 128
 * (width 96) or 384 (width 256) bytes per row in the FP16 form. */
#define FK_FLAG_GRAM 0x8u
#define FK_FLAG_GRAM_BF16 0x10u  /* force the bounded BF16 form                  */
#define FK_FLAG_GRAM_F16 0x20u   /* force the bounded FP16 form                  */

/* Learned distance screen in the local join (NN-descent iterations 2+): a
 * d -> 32 projection and a 32-64-1 MLP, trained during the build on the first
 * iteration's exact pair distances, skip the exact distance of pairs it
 * predicts cannot enter either pool (recalibrated before each join on a
 * sample of its own pairs to lose about 1e-3 of the pairs a threshold would
 * admit; audited in fk_stats).  NOTRAIN keeps the
 * initial weights: the plain random-projection estimate, calibrated the same
 * way.  Not combinable with FK_FLAG_GRAM. */
#define FK_FLAG_SCREEN 0x40u
#define FK_FLAG_SCREEN_NOTRAIN 0x80u

/* Prune kernel variant.  Builds pick one per shape (rule, width, row kind,
 * nsg_R, search_K, sync or reversed) on its first use in a handle: each
 * variant prunes the same sample of that launch's pools, timed, the sample's
 * out-lists and distance counts are checked to be identical, the faster
 * variant prunes the rest and is used from then on (fk_stats.prune_variant).  Variant 0 checks each candidate against the
 * first cached kept neighbours and splits the survivors' remaining checks
 * across the wavefront; variant 1 checks each candidate against every kept
 * neighbour itself; variant 2 (FP32 rows of width 96, 128 or 256) caches up
 * to 64 kept neighbours (8 KiB) as half-precision rows and decides a check from a
 * rigorous bound on the distance to them, computing it from the FP32 rows
 * only when the bound cannot decide (those checks are counted in
 * d_prune_uncomputed); variant 3 (with FK_FLAG_GRAM, nsg_R <= 64, width <=
 * 256) computes every (candidate, kept neighbour) distance of a batch of 16
 * candidates as Gram-form MFMA tiles over the dataset's 16-bit operand copy:
 * exact on exact-form rows, otherwise a rigorous bracket decides and
 * undecided pairs are computed in FP32.  These flags force one (no timing; a
 * variant that does not apply runs as 0); fk_prune and fk_prune_noalloc use
 * variant 0 unless FK_FLAG_PRUNE_V1, V2 or V3 is set. */
#define FK_FLAG_PRUNE_V0 0x100u
#define FK_FLAG_PRUNE_V1 0x200u
#define FK_FLAG_PRUNE_V2 0x400u
#define FK_FLAG_PRUNE_V3 0x800u

/* Gram operand formats (fk_stats.gram_mode, fk_gram_pair_distances). */
enum {
    FK_GRAM_NONE = 0,
    FK_GRAM_EXACT = 1, /* integer rows, |x_i| <= 256, small norms: BF16 operands are
                          exact and the Gram distance is the FP32 distance */
    FK_GRAM_BF16 = 2,  /* BF16 operands, FP32 recheck inside a rigorous bound   */
    FK_GRAM_F16 = 3    /* FP16 operands (power-of-two prescaled), same recheck   */
};

/* Kernel families timed under FK_FLAG_PROFILE (fk_stats.prof_*). */
enum {
    FK_PROF_CONTROLS = 0, /* brute-force keys + top-k select (GenerateControl) */
    FK_PROF_INIT = 1,     /* random init                                        */
    FK_PROF_JOIN = 2,     /* local-join distance block + bucket reservation     */
    FK_PROF_JOIN_FILL = 3,/* bucket fill                                        */
    FK_PROF_MERGE = 4,    /* per-node bucket merge into the pools               */
    FK_PROF_UPDATE = 5,   /* resampling (new / old / reverse lists)             */
    FK_PROF_PRUNE = 6,    /* occlusion prune of the pools                       */
    FK_PROF_PRUNE_REV = 7,/* reversed prune of over-full out-lists              */
    FK_PROF_INTER = 8,    /* reverse-edge gather and merge                      */
    FK_PROF_BV = 9,       /* BridgeView rounds                                  */
    FK_PROF_EP = 10,      /* entry-point beam search + centroid                 */
    FK_PROF_CONN = 11,    /* BFS connectivity                                   */
    FK_PROF_ORDER = 12,   /* locality order (label hops + key sort)             */
    FK_PROF_N = 16
};

/* KGraph::IndexParams fields used on the build path (src/kgraph.h:116-170),
 * one for one, plus flags. */
typedef struct {
    uint32_t K, L, S, R, iterations, controls, seed;
    float delta, recall;
    uint32_t nsg_R, search_L, search_K, step, massq_S, loop_i;
    float tau, angle;
    uint32_t flags;
} fk_params;

#define FK_MAX_LOOPS 16

/* Per-build statistics.  Distance-evaluation counters follow the
 * reference's counting points: d_join = n_comps (src/kgraph.cpp:1108),
 * d_prune = prune_cc * N(N-1)/2 (:2115), d_bv = search_cc * N(N-1)/2 (:2503);
 * the others are not counted by the reference.  Times are seconds. */
typedef struct {
    uint64_t d_init, d_ctrl, d_join, d_prune, d_prune_rev, d_inter, d_ep, d_bv, d_conn;
    double t_total, t_ctrl, t_nnd, t_prune, t_inter, t_ep, t_conn, t_bv, t_upload;
    uint32_t nnd_iterations;      /* NN-descent joins run                     */
    uint32_t nnd_stop;            /* 0 iterations, 1 delta, 2 recall           */
    float nnd_recall, nnd_delta;  /* last NN-descent stop-rule values         */
    uint32_t n_loops;             /* prune loops run (loop_i + 1)             */
    uint32_t n_kgraph, n_recall;  /* entries used in kgraph_time / recall     */
    double kgraph_time[FK_MAX_LOOPS + 64];
    float recall[FK_MAX_LOOPS + 64];
    double prune_time[FK_MAX_LOOPS], inter_time[FK_MAX_LOOPS], tree_grow_time[FK_MAX_LOOPS];
    double buildnsg_time[FK_MAX_LOOPS], search_time[FK_MAX_LOOPS], con_time[FK_MAX_LOOPS];
    float prune_cc[FK_MAX_LOOPS], search_cc[FK_MAX_LOOPS];
    uint32_t aod;                 /* average out-degree, integer as reference */
    uint32_t max_degree, min_degree;
    uint32_t n_attached;          /* edges added by the connectivity pass     */
    uint32_t bv_rounds[FK_MAX_LOOPS];
    /* FK_FLAG_PROFILE: per kernel family, summed kernel milliseconds (HIP
     * events on the build's stream), algorithmic HBM bytes and launches. */
    double prof_ms[FK_PROF_N];
    double prof_bytes[FK_PROF_N];
    uint32_t prof_launches[FK_PROF_N];
    /* FK_FLAG_GRAM: operand format, join pairs computed on the MFMA units,
     * and of those the pairs recomputed in FP32 (bounded forms only). */
    uint32_t gram_mode;
    uint32_t reserved0;
    uint64_t d_gram, d_recheck;
    /* FK_FLAG_SCREEN: join pairs the screen evaluated, pairs it skipped, the
     * skipped pairs audited with the exact distance, and the pool candidates
     * those audited pairs would have given (each counted per endpoint). */
    uint64_t d_screened, d_screen_skip, screen_audit, screen_missed;
    float screen_margin, screen_loss;  /* calibrated relative margin, final training loss */
    uint32_t screen_train_pairs, reserved1;
    double t_screen;                   /* training, projection and calibration seconds */
    uint64_t screen_calib_pairs;       /* admitted pairs in the per-join calibration samples */
    uint64_t screen_cands;             /* pool candidates of the screened joins; one skipped
                                          pair in 64 is audited, so about 64 * screen_missed
                                          admitted candidates were lost to the screen */
    /* BridgeView pairs whose distance was computed: d_bv counts every pair
     * the reference evaluates, of which pairs whose target is already in the
     * pool or was computed earlier in the same round for the same node are
     * counted but not computed again (the result is the same). */
    uint64_t d_bv_computed;
    /* Prune and reversed-prune distances counted in d_prune / d_prune_rev
     * but not computed in FP32: under the tau-MNG rule a candidate whose
     * distance to the owner is at most 3 tau cannot be occluded (djk < dist -
     * 3 tau never holds for a squared distance), so only a kept neighbour
     * with its id removes it and the distances the reference evaluates for it
     * decide nothing; prune variant 2 decides most checks from a bound on the
     * distance to a half-precision copy of the kept row. */
    uint64_t d_prune_uncomputed;
    /* Prune variants used by the last main and reversed prune (1 + variant;
     * 0: none ran), and the seconds spent timing variants on samples in this
     * build (a shape's first use in a handle). */
    uint32_t prune_variant, prune_rev_variant;
    double t_tune;
    /* FK_FLAG_GRAM on FP16-form rows of width 96, 128 or 256: BridgeView
     * targets of d_bv_computed that a rigorous bound on a low-precision copy
     * of the rows (8-bit integers with one scale per row at width 96, the
     * 16-bit operands at 128 and 256) showed to be beyond the pool's last
     * distance, so no FP32 distance was computed for them (they cannot
     * enter the pool). */
    uint64_t d_bv_rejected;
} fk_stats;

/* Output graph in CSR form, allocated by the library (fk_graph_free). */
typedef struct {
    uint32_t n, ep, width;
    uint64_t *offsets; /* n + 1 entries */
    uint32_t *ids;     /* offsets[n] entries, each node's list in reference order */
    fk_stats stats;
} fk_graph;

typedef struct fk_handle fk_handle;

/* Open device `device` (HIP ordinal).  Fails with FK_E_NODEVICE when it is
 * not present. */
fk_status fk_open(int device, fk_handle **h);
void fk_close(fk_handle *h);
const char *fk_last_error(const fk_handle *h);
/* Library version string (KGraph::version, src/kgraph.h:266). */
const char *fk_version(void);

/* Copy n_total rows of d float32 values (host memory, row_stride_bytes apart)
 * to the device.  Replaces any previously resident dataset. */
fk_status fk_dataset_upload(fk_handle *h, const float *rows, size_t row_stride_bytes,
                            uint32_t n_total, uint32_t d);

/* Full FastKCNA build over the first n_used resident rows (the oracle's
 * size(); Calcenter still averages all n_total rows, src/kgraph-data.h:413-428).
 * Runs sampled ground truth, NN-descent, loop_i+1 prune loops with reverse
 * edges, entry point and connectivity, and loop_i BridgeView passes.  For
 * FK_PG_HNSW_LAYER the connectivity pass is skipped (src/kgraph.cpp:2326-2337).
 * The graph is returned in *out (free with fk_graph_free). */
fk_status fk_build(fk_handle *h, fk_pg_type type, uint32_t n_used, const fk_params *p,
                   fk_graph *out);
void fk_graph_free(fk_graph *g);

/* Batched build of n_sets independent datasets in one pass (the reference
 * builds one dataset per KGraph::buildFast* call, src/kgraph.cpp:3720-3966;
 * this entry point has no counterpart and is for callers with many datasets).
 * Set s is the resident rows [set_off[s], set_off[s + 1]) (set_off[0] = 0,
 * ascending, set_off[n_sets] <= n_total); every kernel launch covers all
 * sets, and every per-set decision (random draws, controls and stop rule,
 * BridgeView rounds, entry point, connectivity) is made per set, so out[s]
 * is the graph fk_build gives for set s's rows uploaded alone, with ids
 * relative to the set.  Every set must have more rows than max(K, L, S,
 * controls, search_L, search_K); rows of different widths are uploaded
 * zero-padded to the widest (distances are unchanged when every width is a
 * multiple of 32).  out[s].stats holds the whole batch's statistics; free
 * each out[s] with fk_graph_free.  Not combinable with FK_FLAG_SCREEN. */
fk_status fk_build_batch(fk_handle *h, fk_pg_type type, uint32_t n_sets, const uint64_t *set_off,
                         const fk_params *p, fk_graph *out);

/* Squared L2 between resident rows a[k] and b[k], k < n, in the reference's
 * float_l2sqr_avx summation order (bit-identical results). */
fk_status fk_pair_distances(fk_handle *h, const uint32_t *a, const uint32_t *b, size_t n,
                            float *out);

/* Gram-form (MFMA) squared L2 of resident rows a[k], b[k], k < n, in the
 * operand format a build with these flags (FK_FLAG_GRAM_*) would use; *mode
 * receives it.  lo[k] <= (canonical FP32 distance) <= hi[k] is guaranteed;
 * in FK_GRAM_EXACT mode out = lo = hi = the canonical distance.  Test and
 * measurement entry point of the FK_FLAG_GRAM join; no reference
 * counterpart (the reference has only the FP32 distance, src/metric.cpp:45-78). */
fk_status fk_gram_pair_distances(fk_handle *h, uint32_t flags, const uint32_t *a, const uint32_t *b, size_t n,
                                 float *out, float *lo, float *hi, uint32_t *mode);

/* Learned screen, test entry points (no reference counterpart).  Weights are
 * fk_screen_param_count(d) floats: P^T [dpad][32] | W1 [64][32] | b1 [64] |
 * w2 [64] | b2, dpad = d rounded up to 8.  fk_screen_grad runs the forward and
 * backward kernels of one training batch (n <= 65536 pairs of resident rows,
 * targets dist, inv_mu = 1 / distance scale) and returns the gradient of
 * loss = mean ((y - t) / t)^2, t = max(dist * inv_mu, 0.01), and the loss;
 * fk_screen_predict returns y (the predicted distance times inv_mu) in FP32. */
size_t fk_screen_param_count(uint32_t d);
fk_status fk_screen_grad(fk_handle *h, const float *weights, const uint32_t *a, const uint32_t *b,
                         const float *dist, uint32_t n, float inv_mu, float *grad, double *loss);
fk_status fk_screen_predict(fk_handle *h, const float *weights, const uint32_t *a, const uint32_t *b,
                            uint32_t n, float inv_mu, float *out);

/* Local-join distances.  For node t < n_nodes, its sample list is
 * ids[list_off[t] .. list_off[t+1]), whose first n_new[t] entries are "new" and
 * the rest "old".  Pairs are (x, y) with x < n_new[t] and x < y, enumerated
 * row-major in x, then y; out[pair_off[t] + k] receives the k-th pair's
 * distance, pair_off being the exclusive prefix sum of
 * n_new*(n_new-1)/2 + n_new*n_old.  *n_pairs receives the total. */
fk_status fk_join_distances(fk_handle *h, const uint32_t *ids, const uint64_t *list_off,
                            const uint32_t *n_new, uint32_t n_nodes, float *out,
                            uint64_t *n_pairs);

/* Occlusion prune of n_tasks candidate pools.  Pool t holds pool_len[t]
 * packed entries (float distance bits << 32 | id), sorted ascending, at
 * pools[t * pool_stride].  q[t] is the pool's owner node.  reversed != 0 runs
 * the *_reversed_prune form (no search_K cap, no id < size() guard).
 * Output: out_ids/out_dists[t * out_stride + k] for k < out_len[t] (at most
 * nsg_R, out_stride >= nsg_R), *cc = distances computed.  Rows are the resident
 * dataset; oracle_size is the reference's oracle.size(). */
fk_status fk_prune(fk_handle *h, fk_pg_type type, const uint64_t *pools, const uint32_t *pool_len,
                   uint32_t pool_stride, const uint32_t *q, uint32_t n_tasks, uint32_t oracle_size,
                   const fk_params *p, int reversed, uint32_t *out_ids, float *out_dists,
                   uint32_t out_stride, uint32_t *out_len, uint64_t *cc);

/* fk_prune without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * fk_prune; the staged pools, lengths and owners, the outputs and the
 * distance counter live in a static device scratch area of
 * FK_PRUNE_NOALLOC_SCRATCH_BYTES per GPU, held from the first fk_open on that
 * GPU until the process ends.  A call needing more than that (about
 * n_tasks * (8 pool_stride + 8 out_stride + 12) bytes) fails with
 * FK_E_INVALID.  Every call on one GPU shares that GPU's area: calls that
 * overlap on one GPU (several threads or handles) race.  With FK_FLAG_GRAM
 * the dataset's Gram form must already be prepared for the same flags (an
 * earlier fk_prune or fk_build with them), else FK_E_INVALID. */
#define FK_PRUNE_NOALLOC_SCRATCH_BYTES (64u << 20)
fk_status fk_prune_noalloc(fk_handle *h, fk_pg_type type, const uint64_t *pools, const uint32_t *pool_len,
                           uint32_t pool_stride, const uint32_t *q, uint32_t n_tasks, uint32_t oracle_size,
                           const fk_params *p, int reversed, uint32_t *out_ids, float *out_dists,
                           uint32_t out_stride, uint32_t *out_len, uint64_t *cc);

/* NN-descent only (the KGraphConstructor loop) over the first n_used rows.
 * pools receives n_used * pool_stride packed entries (pool_stride >= L), each
 * node's first pool_len[i] sorted ascending; stats may be NULL. */
fk_status fk_nndescent(fk_handle *h, uint32_t n_used, const fk_params *p, uint64_t *pools,
                       uint32_t *pool_len, uint32_t pool_stride, fk_stats *stats);

/* Exact K nearest rows (among the first n_used) of n_q query vectors given
 * in host memory (n_q x d, contiguous), ties by ascending id; exclude_self
 * != 0 treats query i as row self_ids[i] and skips it (LinearSearch). */
fk_status fk_knn_bruteforce(fk_handle *h, const float *queries, uint32_t n_q, uint32_t n_used,
                            uint32_t K, const uint32_t *self_ids, uint32_t *out_ids,
                            float *out_dists);

/* ---------------------------------------------------------------------------
 * Step entry points: one stateful step of the build's loop run on a state the
 * caller supplies, for parity tests of each step against the oracle.  The
 * reference runs these steps only inside KGraphImpl::buildFast*; each entry
 * point names the reference code it replaces.
 *
 * State of n nodes in host memory.  Node i's candidate pool is
 * pool[i * pool_cap + k], k < pool_len[i]: packed keys (float distance bits
 * << 32 | id) sorted ascending, with pool_flags[i * pool_cap + k] (bit 0:
 * flag, not yet expanded; bit 1: isnew).  Its out-list is out_id / out_dist /
 * out_isnew [i * out_cap + k], k < out_len[i]; the previous loop's out-list
 * (the reference's old_nsg) is prev_id[i * out_cap + k], k < prev_len[i].
 * loopcount = prune loops completed (the reference's loopcount).  Each entry
 * point reads and writes only the fields it names; rows are the resident
 * dataset, the state covers its first n rows (the oracle's size()).
 * --------------------------------------------------------------------------- */
typedef struct {
    uint32_t n, pool_cap, out_cap, loopcount;
    uint64_t *pool;
    uint32_t *pool_len;
    uint8_t *pool_flags;
    uint32_t *out_id;
    float *out_dist;
    uint8_t *out_isnew;
    uint32_t *out_len;
    uint32_t *prev_id;
    uint32_t *prev_len;
} fk_state;

/* One Link step: nsg_Link / taumg_Link / alphapg_Link
 * (src/kgraph.cpp:2087-2143, 3085-3141, 3424-3480) with the full prune rule
 * for every node (*_sync_prune, :1809-1851, 2833-2876, 3157-3203), the
 * reverse edges of InterInsert (:2051-2084) and the reversed prune of lists
 * above nsg_R (:2012-2048, 3038-3083, 3374-3422).  Reads pool, pool_len;
 * writes out_id, out_dist, out_len (out_cap >= nsg_R).  A list that fits
 * nsg_R is the kept prefix followed by its new reverse edges in (dist, id)
 * order.  st (optional) receives d_prune, d_prune_rev and prune_cc[0]. */
fk_status fk_step_link(fk_handle *h, fk_pg_type type, const fk_params *p, fk_state *s, fk_stats *st);

/* UpdateNhoods (src/kgraph.cpp:3537-3584): out_isnew[k] = 1 when loopcount is
 * 0 or out_id[k] is not in the previous out-list, else 0; then the previous
 * out-list becomes the out-list.  Reads out_id, out_len, prev_id, prev_len;
 * writes out_isnew, prev_id, prev_len (prev capacity out_cap). */
fk_status fk_step_update_nhoods(fk_handle *h, fk_state *s);

/* One BridgeView pass (src/kgraph.cpp:2478-2663) at s->loopcount >= 1: the
 * window initialisation of BridgeView_init, then update / join rounds until
 * the 50 % rule ends them.  Reads out_id, out_isnew, out_len; updates pool,
 * pool_len and pool_flags (pool_cap must equal max(L, K, search_L, search_K)
 * after the reference's shrinking for n nodes).  st (optional) receives d_bv
 * (every pair the reference evaluates), d_bv_computed, bv_rounds[0] and
 * search_cc[0]. */
fk_status fk_step_bridgeview(fk_handle *h, const fk_params *p, fk_state *s, fk_stats *st);

/* The NN-descent stop rule's statistics (src/kgraph.cpp:1365-1428,
 * EvaluateDelta :194-206, EvaluateRecall :113-169) of the state's pools:
 * *delta = mean over nodes of (flagged entries among the first min(K,
 * pool_len) entries) / K; *recall = mean over the n_ctrl controls of
 * |ids of pool[0, min(K, pool_len)) ∩ gt[c * gt_stride + (0 .. K))| / K, with
 * gt the exact neighbours of control ctrl_ids[c] (gt_stride >= K).  The build
 * stops NN-descent when delta <= p->delta or recall >= p->recall. */
fk_status fk_step_stop_stats(fk_handle *h, const fk_params *p, const fk_state *s, uint32_t n_ctrl,
                             const uint32_t *ctrl_ids, const uint32_t *gt, uint32_t gt_stride, float *delta,
                             float *recall);

/* ComputeEp (src/kgraph.cpp:1528-1536): the centroid of all resident rows
 * (MatrixOracle::Calcenter, src/kgraph-data.h:413-428: float sums in row
 * order divided by the row count), then get_neighbors (:1560-1623) of the
 * centroid with beam search_L over the out-lists (out_id, out_len) starting
 * from the pool of init_node (pool, pool_len; the reference draws init_node
 * with rand()); *ep = the best node found.  When init_node's pool has fewer
 * than search_L entries the start set is filled with random nodes. */
fk_status fk_step_entry_point(fk_handle *h, const fk_params *p, const fk_state *s, uint32_t init_node,
                              uint32_t *ep);

/* The connectivity pass (KGraphConstructor::tree_grow / findroot / DFS,
 * src/kgraph.cpp:2145-2260) on the state's out-lists (out_id, out_dist,
 * out_len) and pools (pool, pool_len) from entry point ep: nodes unreachable
 * from ep are attached in ascending order of the first unreached id, each
 * to the first entry of its pool that is reached and has fewer than nsg_R
 * out-edges (else a random reached node, under a bound that grows with the
 * draws); the edge and its distance are appended to that node's out-list
 * (out_cap > nsg_R).  st (optional) receives n_attached and d_conn. */
fk_status fk_step_connectivity(fk_handle *h, const fk_params *p, fk_state *s, uint32_t ep, fk_stats *st);

/* Beam search of n_q host queries over a CSR graph (the layout fk_build
 * returns) from entry point ep with beam width search_L; writes the best K
 * ids per query.  The best-first loop of get_neighbors (src/kgraph.cpp:
 * 1560-1623): the start set is ep, then ep's out-list in order, then ids
 * (hash of the query index and a counter) % n until search_L distinct ids;
 * expand the best unexpanded entry, evaluate each unvisited neighbour once,
 * keep the search_L best in (dist, id) order, resume from the smallest
 * position an insert reached.  *n_dist receives the distances computed. */
fk_status fk_search(fk_handle *h, const uint64_t *offsets, const uint32_t *ids, uint32_t n,
                    uint32_t ep, const float *queries, uint32_t n_q, uint32_t search_L,
                    uint32_t K, uint32_t *out_ids, uint64_t *n_dist);

#ifdef __cplusplus
}
#endif

#endif /* FK_API_H */
