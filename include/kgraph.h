// kgraph.h -- KGraph build API of the MI355X engine.
//
// Drop-in for the reference's public header (src/kgraph.h): the same
// namespace, class names, IndexParams fields and defaults, virtual method
// signatures and exception types, so code written against FastKCNA compiles
// unchanged.  The build methods run on the GPU through the C ABI of
// fk_api.h; members that no build path of the reference uses (legacy
// KNN-graph search, prune, get_nn, reverse) are declared for source
// compatibility and throw kgraph::runtime_error.
#ifndef FK_KGRAPH_H
#define FK_KGRAPH_H

#include <stddef.h>
#include <stdint.h>

#include <stdexcept>
#include <vector>

namespace kgraph {
// src/kgraph.h:12-40
static unsigned const default_iterations = 30;
static unsigned const default_L = 100;
static unsigned const default_K = 25;
static unsigned const default_P = 100;
static unsigned const default_M = 0;
static unsigned const default_T = 1;
static unsigned const default_S = 10;
static unsigned const default_R = 100;
static unsigned const default_controls = 100;
static unsigned const default_seed = 2024;
static float const default_delta = 0.002f;
static float const default_recall = 0.95f;
static float const default_epsilon = 1e30f;
static unsigned const default_verbosity = 1;
static unsigned const default_nthreads = 40;
static unsigned const default_num0 = 100;
static unsigned const default_iter = 3;
static float const default_incr_r = 10.0f;
static unsigned const default_loop_i = 2;
static unsigned const default_nsgr = 50;
static unsigned const default_searchl = 40;
static unsigned const default_searchk = 500;
static unsigned const default_step = 10;
static unsigned const default_massqS = 10;
static float const default_tau = 0;
static float const default_angle = 60;

enum { PRUNE_LEVEL_1 = 1, PRUNE_LEVEL_2 = 2 };
enum { REVERSE_AUTO = -1, REVERSE_NONE = 0 };
static unsigned const default_prune = 0;
static int const default_reverse = REVERSE_NONE;

/// Verbosity: 0 silences the per-phase log on stderr.
extern unsigned verbosity;

/// Index oracle: the user-supplied distance plugin (src/kgraph.h:63-81).
/// The GPU build reads the rows of a MatrixOracle<float, metric::l2sqr>
/// directly; other oracles are rejected with invalid_argument.
class IndexOracle {
public:
    virtual ~IndexOracle() {}
    virtual unsigned size() const = 0;
    virtual unsigned dim() const = 0;
    virtual float operator()(unsigned i, unsigned j) const = 0;
    virtual float operator()(unsigned i, unsigned j, unsigned offset, unsigned dim) const = 0;
    virtual float operator()(const float *query, unsigned i) const = 0;
    virtual float computeDot(unsigned i, unsigned j) const = 0;
    virtual void computeSelfDotPtr() = 0;
    virtual float *Calcenter() const = 0;
    virtual unsigned data_size() const = 0;
    virtual void set_size(unsigned mysize) = 0;
};

/// Search oracle (src/kgraph.h:88-108); brute-force search is provided.
class SearchOracle {
public:
    virtual ~SearchOracle() {}
    virtual unsigned size() const = 0;
    virtual float operator()(unsigned i) const = 0;
    unsigned search(unsigned K, float epsilon, unsigned *ids, float *dists = nullptr) const;
};

class KGraph {
public:
    struct IndexParams {
        unsigned nsg_R;
        unsigned search_L;
        unsigned search_K;
        unsigned step;
        float tau;
        float angle;
        unsigned nt;
        unsigned massq_S;
        unsigned loop_i;
        unsigned iterations;
        unsigned L;
        unsigned K;
        unsigned S;
        unsigned R;
        unsigned controls;
        unsigned seed;
        unsigned nthreads;
        float delta;
        float recall;
        unsigned prune;
        int reverse;
        unsigned num0;
        float incr_r;
        unsigned sl;
        unsigned iter;
        unsigned qid;
        unsigned Lcon;
        unsigned sidx;
        unsigned B;
        unsigned P;
        /// GPU ordinal used by the build (not in the reference; default 0).
        int device;
        /// GPU-only build options (not in the reference; default 0): the
        /// FK_FLAG_GRAM* and FK_FLAG_SCREEN* bits of fk_api.h.
        unsigned gpu_flags;
        IndexParams()
            : nsg_R(default_nsgr), search_L(default_searchl), search_K(default_searchk), step(default_step),
              tau(default_tau), angle(default_angle), nt(0), massq_S(default_massqS), loop_i(default_loop_i),
              iterations(default_iterations), L(default_L), K(default_K), S(default_S), R(default_R),
              controls(default_controls), seed(default_seed), nthreads(default_nthreads), delta(default_delta),
              recall(default_recall), prune(default_prune), reverse(default_reverse), num0(default_num0),
              incr_r(default_incr_r), sl(0), iter(default_iter), qid(0), Lcon(0), sidx(0), B(0), P(default_P),
              device(0), gpu_flags(0) {}
    };

    struct SearchParams {
        unsigned K, M, P, S, T;
        float epsilon;
        unsigned seed, init;
        SearchParams()
            : K(default_K), M(default_M), P(default_P), S(default_S), T(default_T), epsilon(default_epsilon),
              seed(1998), init(0) {}
    };

    enum { FORMAT_DEFAULT = 0, FORMAT_NO_DIST = 1 };

    struct IndexInfo {
        enum StopCondition { ITERATION = 0, DELTA, RECALL } stop_condition;
        unsigned iterations;
        float cost;
        float recall;
        float accuracy;
        float delta;
        float M;
    };

    struct SearchInfo {
        float cost;
        unsigned updates;
    };

    virtual ~KGraph() {}
    virtual void load(char const *path) = 0;
    virtual void save(char const *path, int format = FORMAT_DEFAULT) const = 0;
    virtual void buildFastNSG(IndexOracle const &oracle, IndexParams const &param, IndexInfo *info = 0,
                              const char *nsg_name = 0, const char *res_name = 0) = 0;
    virtual void buildFastTauMNG(IndexOracle const &oracle, IndexParams const &param, IndexInfo *info = 0,
                                 const char *nsg_name = 0, const char *res_name = 0) = 0;
    virtual void buildFastAlphaPG(IndexOracle const &oracle, IndexParams const &param, IndexInfo *info = 0,
                                  const char *nsg_name = 0, const char *res_name = 0) = 0;
    virtual void buildFastHNSW(IndexOracle const &oracle, IndexParams const &param, IndexInfo *info,
                               const char *nsg_name, const char *res_name, char *&data_level0_memory_,
                               char **linkLists_ = 0, unsigned level = 0, unsigned *ep_ptr = 0,
                               size_t size_data_per_element_ = 0) = 0;
    virtual void FatherSwapHNSWLayer0(IndexOracle const &oracle, IndexParams const &params,
                                      char *&data_level0_memory_) = 0;
    virtual void prune(IndexOracle const &oracle, unsigned level) = 0;
    unsigned search(SearchOracle const &oracle, SearchParams const &params, unsigned *ids,
                    SearchInfo *info = 0) const {
        return search(oracle, params, ids, nullptr, info);
    }
    virtual unsigned search(SearchOracle const &oracle, SearchParams const &params, unsigned *ids, float *dists,
                            SearchInfo *info) const = 0;
    static KGraph *create();
    static char const *version();
    virtual void get_nn(unsigned id, unsigned *nns, unsigned *M, unsigned *L) const {
        get_nn(id, nns, nullptr, M, L);
    }
    virtual void get_nn(unsigned id, unsigned *nns, float *dists, unsigned *M, unsigned *L) const = 0;
    virtual void reverse(int) = 0;
};

class invalid_argument : public std::invalid_argument {
public:
    using std::invalid_argument::invalid_argument;
};

class runtime_error : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

class io_error : public runtime_error {
public:
    using runtime_error::runtime_error;
};

}  // namespace kgraph

#endif  // FK_KGRAPH_H
