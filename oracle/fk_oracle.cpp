// fk_oracle.cpp -- CPU restatement of FastKCNA's graph-construction path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the
// timed CPU baseline; nothing in the product (libfastkcna.so, the C++ host
// layer, the command-line tools) links, loads or calls it.  Only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
//
// It restates, in this project's own code, the algorithm of the reference
// sources (paths relative to the FastKCNA tree):
//   distance        src/metric.cpp:45-78 (float_l2sqr_avx as g++ -O3
//                   -march=native compiles it: fused multiply-add per lane)
//   pool insert     src/kgraph.cpp:261-298 (InsertIntoPool), :948-967
//   NN-descent      src/kgraph.cpp:1001-1044 (init), :1097-1117 (join),
//                   :1118-1202 (update), :1281-1459 (driver, stop rule)
//   controls        src/kgraph.cpp:309-330, :363-387
//   pruners         src/kgraph.cpp:1809-2048 (NSG), :2833-3083 (tau-MNG),
//                   :3157-3422 (alpha-PG): *_sync_prune, *_sync_prune2,
//                   *_reversed_prune; nsg_Link / InterInsert :2051-2143
//   entry point     src/kgraph.cpp:1528-1536, :1560-1623
//   connectivity    src/kgraph.cpp:2145-2270
//   UpdateNhoods    src/kgraph.cpp:3537-3584
//   BridgeView      src/kgraph.cpp:2478-2663
//   drivers         src/kgraph.cpp:3720-3966
// with the same 12-byte Neighbor pools, spinlocks, OpenMP schedules and
// parameters.  Global rand() of the reference is replaced by a seeded
// std::mt19937 (the reference's own calls are unseeded, so no run of it is
// reproducible either).
//
// Parity pins (tests/test_oracle_*.py): the distance against an AVX
// intrinsics kernel built the reference's way and against FP64; the pool
// insert against brute force; the pruners against closed forms (tau = 0 and
// angle special cases equal the NSG rule, identity when nothing occludes) and
// brute-force re-checks of every kept/dropped decision; NN-descent and the
// full build by recall against brute force and graph invariants; the
// stateful steps (BridgeView, Link with InterInsert, UpdateNhoods, the stop
// rule, ComputeEp, the connectivity pass's attach order) on hand-built
// states whose results follow from the reference's definitions
// (tests/test_oracle_steps.py).  InterInsert gathers its reverse edges before
// appending them (Builder::inter_insert says why the edges are the
// reference's).
#include <omp.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <mutex>
#include <random>
#include <stack>
#include <vector>

#include "fk_api.h"

namespace {

using std::vector;

// ---------------------------------------------------------------------------
// Distances
// ---------------------------------------------------------------------------
// float_l2sqr_avx order: D = dim rounded up to 8 (rows are zero padded);
// the D % 32 tail blocks are accumulated first (offsets +16, +8, +0 past the
// last full 32-block), then the 8-blocks from 0 upward, each lane k of the
// 256-bit accumulator taking fma(d, d, acc[k]); finally the 8 lanes are
// added left to right.
float l2_ref(const float *a, const float *b, unsigned dpad) {
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    auto step = [&](unsigned off) {
        for (int k = 0; k < 8; ++k) {
            const float t = a[off + k] - b[off + k];
            acc[k] = std::fma(t, t, acc[k]);
        }
    };
    const unsigned dr = dpad % 32, dd = dpad - dr;
    if (dr == 24) { step(dd + 16); step(dd + 8); step(dd); }
    else if (dr == 16) { step(dd + 8); step(dd); }
    else if (dr == 8) { step(dd); }
    for (unsigned i = 0; i < dd; i += 8) step(i);
    float r = acc[0] + acc[1];
    r += acc[2]; r += acc[3]; r += acc[4]; r += acc[5]; r += acc[6]; r += acc[7];
    return r;
}

double l2_f64(const float *a, const float *b, unsigned d) {
    double s = 0;
    for (unsigned i = 0; i < d; ++i) {
        const double t = (double)a[i] - (double)b[i];
        s += t * t;
    }
    return s;
}

// Row store with the reference Matrix layout: rows zero padded to a
// multiple of 8 floats.
struct Rows {
    vector<float> buf;
    unsigned n = 0, d = 0, dpad = 0;
    unsigned oracle_n = 0;  // oracle.size() (prefix); n is data_size()
    const float *row(unsigned i) const { return buf.data() + (size_t)i * dpad; }
    float dist(unsigned i, unsigned j) const { return l2_ref(row(i), row(j), dpad); }
    float qdist(const float *q, unsigned i) const { return l2_ref(q, row(i), dpad); }
    void load(const float *src, size_t stride_bytes, unsigned n_, unsigned d_) {
        n = n_; d = d_; dpad = (d + 7) & ~7u;
        buf.assign((size_t)n * dpad, 0.f);
        for (unsigned i = 0; i < n; ++i)
            memcpy(buf.data() + (size_t)i * dpad, (const char *)src + (size_t)i * stride_bytes, (size_t)d * 4);
        oracle_n = n;
    }
};

// ---------------------------------------------------------------------------
// Pool entries and the sorted insert
// ---------------------------------------------------------------------------
struct Nb {
    uint32_t id;
    float dist;
    bool flag;
    bool isnew;
    int16_t pid;
    Nb() : id(0xFFFFFFFFu), dist(std::numeric_limits<float>::max()), flag(false), isnew(true), pid(-1) {}
    Nb(uint32_t i, float d, bool f = true) : id(i), dist(d), flag(f), isnew(true), pid(-1) {}
};
static_assert(sizeof(Nb) == 12, "Neighbor is 12 bytes");
inline bool operator<(const Nb &a, const Nb &b) { return a.dist < b.dist; }

// InsertIntoPool: addr[0..K) sorted by dist, addr[K] is the overflow slot.
// Returns the insert position, K when the entry lands in the overflow slot,
// or K + 1 when an equal-distance entry has the same id.
int insert_into_pool(Nb *addr, unsigned K, Nb nn) {
    int left = 0, right = (int)K - 1;
    if (addr[left].dist > nn.dist) {
        memmove(&addr[left + 1], &addr[left], K * sizeof(Nb));
        addr[left] = nn;
        return left;
    }
    if (addr[right].dist < nn.dist) {
        addr[K] = nn;
        return (int)K;
    }
    while (left < right - 1) {
        const int mid = (left + right) / 2;
        if (addr[mid].dist > nn.dist) right = mid;
        else left = mid;
    }
    while (left > 0) {
        if (addr[left].dist < nn.dist) break;
        if (addr[left].id == nn.id) return (int)K + 1;
        --left;
    }
    if (addr[left].id == nn.id || addr[right].id == nn.id) return (int)K + 1;
    memmove(&addr[right + 1], &addr[right], (K - right) * sizeof(Nb));
    addr[right] = nn;
    return right;
}

// LinearSearch's UpdateKnnList (src/kgraph.cpp:227-258): insertion after
// equal distances, so ties keep scan (id) order.
unsigned update_knn_list(Nb *addr, unsigned K, Nb nn) {
    unsigned i = K;
    while (i > 0) {
        const unsigned j = i - 1;
        if (addr[j].dist <= nn.dist) break;
        i = j;
    }
    unsigned l = i;
    while (l > 0) {
        const unsigned j = l - 1;
        if (addr[j].dist < nn.dist) break;
        if (addr[j].id == nn.id) return K + 1;
        l = j;
    }
    unsigned j = K;
    while (j > i) { addr[j] = addr[j - 1]; --j; }
    addr[i] = nn;
    return i;
}

struct Spin {
    std::atomic_flag f = ATOMIC_FLAG_INIT;
    void lock() { while (f.test_and_set(std::memory_order_acquire)) {} }
    void unlock() { f.clear(std::memory_order_release); }
};

struct Nhood {
    float radius = 0, radiusM = 0;
    unsigned L = 0, M = 0;
    bool found = false;
    vector<Nb> pool;
    vector<Nb> nn_nsg;
    vector<unsigned> nn_old, nn_new, rnn_old, rnn_new, old_nsg;
    vector<bool> rnn_new_flag;
    Spin lock;

    unsigned try_insert(unsigned id, float dist) {
        if (dist > radius) return (unsigned)pool.size();
        std::lock_guard<Spin> g(lock);
        const unsigned l = (unsigned)insert_into_pool(pool.data(), L, Nb(id, dist, true));
        if (l <= L) {
            if (L + 1 < pool.size()) ++L;
            else radius = pool[L - 1].dist;
        }
        return l;
    }
};

// GenRandom (src/kgraph.cpp:82-110)
template <typename RNG>
void gen_random(RNG &rng, unsigned *addr, unsigned size, unsigned N) {
    if (N == size) {
        for (unsigned i = 0; i < size; ++i) addr[i] = i;
        return;
    }
    for (unsigned i = 0; i < size; ++i) addr[i] = rng() % (N - size);
    std::sort(addr, addr + size);
    for (unsigned i = 1; i < size; ++i)
        if (addr[i] <= addr[i - 1]) addr[i] = addr[i - 1] + 1;
    const unsigned off = rng() % N;
    for (unsigned i = 0; i < size; ++i) addr[i] = (addr[i] + off) % N;
}

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum Rule { RULE_NSG = 0, RULE_TAU = 1, RULE_ALPHA = 2 };

struct RuleP {
    Rule rule;
    float tau;
    float thr;
    bool occl(float djk, float pdist, float dqt) const {
        if (rule == RULE_NSG) return djk < pdist;
        if (rule == RULE_TAU) return djk < pdist - 3 * tau;
        const float cos_ij = (dqt + djk - pdist) / 2 / std::sqrt(dqt * djk);
        return djk < pdist && cos_ij < thr;
    }
};

float alpha_thr(float angle) {
    const double kPi = 3.14159265358979323846264;
    return (float)std::cos(angle / 180 * kPi);
}

Rule rule_of(int t) { return t == FK_PG_TAUMNG ? RULE_TAU : (t == FK_PG_ALPHAPG ? RULE_ALPHA : RULE_NSG); }

// ---------------------------------------------------------------------------
// The builder (KGraphConstructor restated)
// ---------------------------------------------------------------------------
struct Control {
    unsigned id;
    vector<Nb> nbrs;
};

struct Builder {
    const Rows &o;
    fk_params P;
    vector<Nhood> nh;
    vector<Control> controls;
    unsigned MAXL = 0;
    unsigned ep = 0, width = 0;
    unsigned loopcount = 0;
    size_t n_comps = 0;
    vector<char> pruneAgain;
    std::mt19937 rnd;        // stands in for the reference's global rand()
    fk_stats *st;
    bool full_rule_only = false;   // never use *_sync_prune2
    uint64_t prune2_diff = 0;      // nodes where sync_prune2 != sync_prune
    uint64_t prune2_nodes = 0;     // nodes where both were compared
    vector<unsigned> kept_len;     // per node: out-list length after the prune, before InterInsert
    bool check_prune2 = false;
    RuleP rp{};
    bool verbose = false;

    Builder(const Rows &rows, const fk_params &p, fk_stats *s) : o(rows), P(p), nh(rows.oracle_n), rnd(p.seed), st(s) {}
    unsigned N() const { return o.oracle_n; }
    unsigned R_() { return (unsigned)(rnd() & 0x7FFFFFFF); }

    void shrink() {
        const unsigned N_ = N();
        if (N_ <= P.controls) P.controls = N_ - 1;
        if (N_ <= P.L) P.L = N_ - 1;
        if (N_ <= P.S) P.S = N_ - 1;
        if (N_ <= P.search_L) P.search_L = N_ - 1;
        if (N_ <= P.search_K) P.search_K = N_ - 1;
        if (N_ <= P.K) { P.K = N_ - 1; P.loop_i = 0; P.S = N_ - 1; P.iterations = 0; }
        MAXL = std::max(std::max(P.L, P.K), std::max(P.search_L, P.search_K));
    }

    void linear_search(unsigned i, unsigned K, vector<Nb> &out) {
        vector<Nb> nns(K + 1);
        unsigned k = 0;
        for (unsigned id = 0; id < N(); ++id) {
            if (id == i) continue;
            update_knn_list(nns.data(), k, Nb(id, o.dist(i, id), true));
            if (k < K) ++k;
        }
        nns.resize(K);
        out.swap(nns);
    }

    void generate_control() {
        const double t0 = now();
        unsigned C = (P.flags & FK_FLAG_NO_CONTROLS) ? 0 : P.controls;
        vector<unsigned> idx(N());
        for (unsigned i = 0; i < N(); ++i) idx[i] = i;
        for (unsigned i = 0; i < C; ++i) std::swap(idx[i], idx[i + R_() % (N() - i)]);
        controls.assign(C, Control());
#pragma omp parallel for
        for (unsigned c = 0; c < C; ++c) {
            controls[c].id = idx[c];
            linear_search(idx[c], MAXL, controls[c].nbrs);
        }
        st->d_ctrl += (uint64_t)C * (N() - 1);
        st->t_ctrl += now() - t0;
    }

    void init() {
        const unsigned N_ = N();
        for (auto &h : nh) {
            h.nn_new.resize(P.S * 2);
            h.pool.resize(P.L + 1);
            h.radius = std::numeric_limits<float>::max();
        }
#pragma omp parallel
        {
            std::mt19937 rng(P.seed ^ omp_get_thread_num());
            vector<unsigned> random(P.S + 1);
#pragma omp for
            for (unsigned n = 0; n < N_; ++n) {
                auto &h = nh[n];
                gen_random(rng, h.nn_new.data(), (unsigned)h.nn_new.size(), N_);
                gen_random(rng, random.data(), (unsigned)random.size(), N_);
                h.L = P.S;
                h.M = P.S;
                unsigned i = 0;
                for (unsigned l = 0; l < h.L; ++l) {
                    if (random[i] == n) ++i;
                    auto &nn = h.pool[l];
                    nn.id = random[i++];
                    nn.dist = o.dist(nn.id, n);
                    nn.flag = true;
                }
                std::sort(h.pool.begin(), h.pool.begin() + h.L);
            }
        }
        st->d_init += (uint64_t)N_ * P.S;
    }

    void join() {
        size_t cc = 0;
        const unsigned K = P.K;
#pragma omp parallel for default(shared) schedule(dynamic, 100) reduction(+ : cc)
        for (unsigned n = 0; n < N(); ++n) {
            size_t uu = 0;
            nh[n].found = false;
            auto cb = [&](unsigned i, unsigned j) {
                const float d = o.dist(i, j);
                ++cc;
                if (nh[i].try_insert(j, d) < K) ++uu;
                if (nh[j].try_insert(i, d) < K) ++uu;
            };
            for (unsigned i : nh[n].nn_new) {
                for (unsigned j : nh[n].nn_new)
                    if (i < j) cb(i, j);
                for (unsigned j : nh[n].nn_old) cb(i, j);
            }
            nh[n].found = uu > 0;
        }
        n_comps += cc;
        st->d_join += cc;
    }

    void update() {
        const unsigned N_ = N();
        for (auto &h : nh) {
            h.nn_new.clear(); h.nn_old.clear(); h.rnn_new.clear(); h.rnn_old.clear();
            h.radius = h.pool.back().dist;
        }
#pragma omp parallel for
        for (unsigned n = 0; n < N_; ++n) {
            auto &h = nh[n];
            if (h.found) {
                const unsigned maxl = std::min(h.M + P.S, h.L);
                unsigned c = 0, l = 0;
                while (l < maxl && c < P.S) {
                    if (h.pool[l].flag) ++c;
                    ++l;
                }
                h.M = l;
            }
            h.radiusM = h.pool[h.M - 1].dist;
        }
#pragma omp parallel for
        for (unsigned n = 0; n < N_; ++n) {
            auto &h = nh[n];
            for (unsigned l = 0; l < h.M; ++l) {
                auto &nn = h.pool[l];
                auto &other = nh[nn.id];
                if (nn.flag) {
                    h.nn_new.push_back(nn.id);
                    if (nn.dist > other.radiusM) {
                        std::lock_guard<Spin> g(other.lock);
                        other.rnn_new.push_back(n);
                    }
                    nn.flag = false;
                } else {
                    h.nn_old.push_back(nn.id);
                    if (nn.dist > other.radiusM) {
                        std::lock_guard<Spin> g(other.lock);
                        other.rnn_old.push_back(n);
                    }
                }
            }
        }
        auto shuffle = [&](vector<unsigned> &v) {
            for (size_t i = v.size(); i > 1; --i) std::swap(v[i - 1], v[R_() % i]);
        };
        for (unsigned i = 0; i < N_; ++i) {
            auto &h = nh[i];
            if (P.R && h.rnn_new.size() > P.R) { shuffle(h.rnn_new); h.rnn_new.resize(P.R); }
            h.nn_new.insert(h.nn_new.end(), h.rnn_new.begin(), h.rnn_new.end());
            if (P.R && h.rnn_old.size() > P.R) { shuffle(h.rnn_old); h.rnn_old.resize(P.R); }
            h.nn_old.insert(h.nn_old.end(), h.rnn_old.begin(), h.rnn_old.end());
        }
    }

    float eval_recall(unsigned K) {
        if (controls.empty()) return 0.f;
        double s = 0;
        for (auto &c : controls) {
            const auto &pool = nh[c.id].pool;
            unsigned found = 0;
            const unsigned lim = std::min(nh[c.id].L, K);
            for (unsigned i = 0; i < lim; ++i)
                for (unsigned j = 0; j < K && j < c.nbrs.size(); ++j)
                    if (pool[i].id == c.nbrs[j].id) { ++found; break; }
            s += (double)found / K;
        }
        return (float)(s / controls.size());
    }

    void nndescent() {
        const double t0 = now();
        init();
        const double tk = now();
        for (unsigned it = 0; it < P.iterations; ++it) {
            st->nnd_iterations++;
            join();
            double dsum = 0;
            for (auto &h : nh) {
                unsigned c = 0;
                const unsigned lim = std::min<unsigned>(P.K, (unsigned)h.pool.size());
                for (unsigned i = 0; i < lim; ++i) if (h.pool[i].flag) ++c;
                dsum += (double)c / P.K;
            }
            const float delta = (float)(dsum / N());
            const float recall = eval_recall(P.K);
            st->nnd_delta = delta;
            st->nnd_recall = recall;
            if (st->n_kgraph < FK_MAX_LOOPS + 64) st->kgraph_time[st->n_kgraph++] = now() - tk;
            if (st->n_recall < FK_MAX_LOOPS + 64) st->recall[st->n_recall++] = recall;
            if (verbose) fprintf(stderr, "[oracle] iteration %u recall %g delta %g\n", it + 1, recall, delta);
            if (delta <= P.delta) { st->nnd_stop = 1; break; }
            if (!controls.empty() && recall >= P.recall) { st->nnd_stop = 2; break; }
            if (it + 1 >= P.iterations) break;
            update();
        }
        st->t_nnd += now() - t0;
    }

    void init_nhoods() {
        for (auto &h : nh) {
            vector<unsigned>().swap(h.nn_new);
            vector<unsigned>().swap(h.nn_old);
            vector<unsigned>().swap(h.rnn_new);
            vector<unsigned>().swap(h.rnn_old);
            h.nn_nsg.reserve(P.nsg_R / 2);
            h.pool.resize(MAXL + 1);
        }
    }

    // *_sync_prune: the full rule (src/kgraph.cpp:1809-1851 and families).
    uint64_t sync_prune(unsigned q, vector<Nb> &out) {
        const unsigned range = P.nsg_R, maxc = P.search_K;
        width = range;
        uint64_t cc = 0;
        unsigned start = 0;
        auto &h = nh[q];
        out.clear();
        if (h.pool[start].id == q) ++start;
        out.push_back(h.pool[start]);
        while (out.size() < range && (++start) < h.L && start < maxc) {
            auto &p = h.pool[start];
            if (p.id >= N()) continue;
            bool occlude = false;
            for (unsigned t = 0; t < out.size(); ++t) {
                if (p.id == out[t].id) { occlude = true; break; }
                ++cc;
                const float djk = o.dist(out[t].id, p.id);
                if (rp.occl(djk, p.dist, out[t].dist)) { occlude = true; p.pid = (int16_t)t; break; }
            }
            if (!occlude) out.push_back(p);
        }
        for (auto &e : out) e.flag = false;
        return cc;
    }

    // *_sync_prune2: the incremental "Opt2" rule (src/kgraph.cpp:1853-2010).
    uint64_t sync_prune2(unsigned q) {
        const unsigned range = P.nsg_R, maxc = P.search_K;
        width = range;
        uint64_t cc = 0;
        unsigned start = 0;
        auto &h = nh[q];
        vector<unsigned> temp_nsg;
        for (auto &e : h.nn_nsg)
            if (!e.flag) temp_nsg.push_back(e.id);
        auto in_temp = [&](unsigned id) { return std::find(temp_nsg.begin(), temp_nsg.end(), id) != temp_nsg.end(); };
        h.nn_nsg.clear();
        if (h.pool[start].id == q) ++start;
        h.nn_nsg.push_back(h.pool[start]);
        h.nn_nsg.back().isnew = !in_temp(h.nn_nsg.back().id);
        auto full_check = [&](Nb &p, bool only_new) -> bool {
            for (unsigned t = 0; t < h.nn_nsg.size(); ++t) {
                if (p.id == h.nn_nsg[t].id) return true;
                if (only_new && !h.nn_nsg[t].isnew) continue;
                ++cc;
                const float djk = o.dist(h.nn_nsg[t].id, p.id);
                if (rp.occl(djk, p.dist, h.nn_nsg[t].dist)) { p.pid = (int16_t)t; return true; }
            }
            return false;
        };
        while (h.nn_nsg.size() < range && (++start) < h.L && start < maxc) {
            auto &p = h.pool[start];
            if (p.id >= N()) continue;
            if (p.isnew) {
                if (!full_check(p, false)) { h.nn_nsg.push_back(p); h.nn_nsg.back().isnew = true; }
                continue;
            }
            const bool was_kept = in_temp(p.id);
            if (!was_kept && p.pid >= 0 && (size_t)p.pid < temp_nsg.size()) {
                const unsigned prune_p = temp_nsg[p.pid];
                auto it = std::find_if(h.nn_nsg.begin(), h.nn_nsg.end(), [&](const Nb &e) { return e.id == prune_p; });
                if (it != h.nn_nsg.end()) {
                    p.pid = (int16_t)std::distance(h.nn_nsg.begin(), it);
                    continue;
                }
                if (!full_check(p, false)) { h.nn_nsg.push_back(p); h.nn_nsg.back().isnew = true; }
            } else if (!was_kept) {
                if (!full_check(p, false)) { h.nn_nsg.push_back(p); h.nn_nsg.back().isnew = true; }
            } else {
                if (!full_check(p, true)) { h.nn_nsg.push_back(p); h.nn_nsg.back().isnew = false; }
            }
        }
        for (auto &e : h.nn_nsg) e.flag = false;
        return cc;
    }

    // *_reversed_prune (src/kgraph.cpp:2012-2048): sort the list, full rule
    // without search_K cap; the canonical tie order is (dist, id).
    uint64_t reversed_prune(unsigned q) {
        const unsigned range = P.nsg_R;
        width = range;
        uint64_t cc = 0;
        unsigned start = 0;
        auto &h = nh[q];
        auto tmp = h.nn_nsg;
        std::sort(tmp.begin(), tmp.end(), [](const Nb &a, const Nb &b) {
            return a.dist < b.dist || (a.dist == b.dist && a.id < b.id);
        });
        h.nn_nsg.clear();
        if (tmp[start].id == q) ++start;
        h.nn_nsg.push_back(tmp[start]);
        while (h.nn_nsg.size() < range && (++start) < tmp.size()) {
            auto &p = tmp[start];
            bool occlude = false;
            for (unsigned t = 0; t < h.nn_nsg.size(); ++t) {
                if (p.id == h.nn_nsg[t].id) { occlude = true; break; }
                ++cc;
                const float djk = o.dist(h.nn_nsg[t].id, p.id);
                if (rp.occl(djk, p.dist, h.nn_nsg[t].dist)) { occlude = true; break; }
            }
            if (!occlude) h.nn_nsg.push_back(p);
        }
        return cc;
    }

    // InterInsert (src/kgraph.cpp:2051-2084) for node n: the reverse edges
    // (des, n) it adds, gathered without touching any list.  The reference
    // appends n to des's list, under des's lock, unless that list already
    // holds n; the list it scans is des's kept list plus reverse edges other
    // nodes appended, none of which is n (a node appends itself to a list at
    // most once: its kept list has distinct ids), so the check reduces to the
    // kept list.  Its loop also runs over the reverse edges appended to n's
    // own list meanwhile; each of those, m, has n in its kept list (that is
    // why m was appended), so they add nothing.  Gathering first and
    // appending afterwards gives the same edges without the reference's
    // unsynchronised read of a list other threads append to.
    void inter_insert(unsigned n, vector<std::pair<unsigned, Nb>> &add) const {
        for (unsigned i = 0; i < kept_len[n]; ++i) {
            const size_t des = nh[n].nn_nsg[i].id;
            if (des >= N()) continue;
            bool dup = false;
            for (unsigned k = 0; k < kept_len[des]; ++k)
                if (nh[des].nn_nsg[k].id == n) { dup = true; break; }
            if (dup) continue;
            add.emplace_back((unsigned)des, Nb(n, o.dist(n, (unsigned)des), true));
        }
    }

    void link(unsigned li) {
        const unsigned N_ = N(), range = P.nsg_R;
        const double t0 = now();
        uint64_t cc = 0, diff = 0, compared = 0;
#pragma omp parallel for schedule(dynamic, 100) reduction(+ : cc, diff, compared)
        for (unsigned n = 0; n < N_; ++n) {
            if (pruneAgain[n] || full_rule_only) {
                cc += sync_prune(n, nh[n].nn_nsg);
            } else if (check_prune2) {
                vector<Nb> full;
                // the full rule must see the same pool state: pids it writes
                // are restored before the incremental rule runs
                vector<int16_t> pids(nh[n].pool.size());
                for (size_t k = 0; k < pids.size(); ++k) pids[k] = nh[n].pool[k].pid;
                sync_prune(n, full);
                for (size_t k = 0; k < pids.size(); ++k) nh[n].pool[k].pid = pids[k];
                cc += sync_prune2(n);
                bool same = full.size() == nh[n].nn_nsg.size();
                for (size_t k = 0; same && k < full.size(); ++k) same = full[k].id == nh[n].nn_nsg[k].id;
                if (!same) ++diff;
                ++compared;
            } else {
                cc += sync_prune2(n);
            }
        }
        prune2_diff += diff;
        prune2_nodes += compared;
        st->d_prune += cc;
        const double t1 = now();
        if (li < FK_MAX_LOOPS) {
            st->prune_cc[li] = (float)((double)cc / ((double)N_ * (N_ - 1) / 2.0));
            st->prune_time[li] = t1 - t0;
        }
        st->t_prune += t1 - t0;
        uint64_t inter_d = 0;
        kept_len.resize(N_);
        for (unsigned n = 0; n < N_; ++n) {
            kept_len[n] = (unsigned)nh[n].nn_nsg.size();
            inter_d += nh[n].nn_nsg.size();
        }
        {
            // reverse edges gathered in parallel and appended; a list left
            // within nsg_R has them in (dist, id) order, the build's
            // canonical order (the reference's is its threads' order), an
            // over-full one is sorted whole by reversed_prune
            vector<vector<std::pair<unsigned, Nb>>> add(N_);
#pragma omp parallel for schedule(dynamic, 100)
            for (unsigned n = 0; n < N_; ++n) inter_insert(n, add[n]);
            for (unsigned n = 0; n < N_; ++n)
                for (auto &e : add[n]) nh[e.first].nn_nsg.push_back(e.second);
#pragma omp parallel for schedule(dynamic, 100)
            for (unsigned n = 0; n < N_; ++n) {
                auto &out = nh[n].nn_nsg;
                if (out.size() <= range)
                    std::sort(out.begin() + kept_len[n], out.end(), [](const Nb &x, const Nb &y) {
                        return x.dist < y.dist || (x.dist == y.dist && x.id < y.id);
                    });
            }
        }
        std::fill(pruneAgain.begin(), pruneAgain.end(), 0);
        uint64_t rc = 0;
#pragma omp parallel for schedule(dynamic, 100) reduction(+ : rc)
        for (unsigned n = 0; n < N_; ++n) {
            if (nh[n].nn_nsg.size() > range) {
                pruneAgain[n] = 1;
                rc += reversed_prune(n);
            }
        }
        st->d_prune_rev += rc;
        st->d_inter += inter_d;
        const double t2 = now();
        if (li < FK_MAX_LOOPS) st->inter_time[li] = t2 - t1;
        st->t_inter += t2 - t1;
    }

    // get_neighbors(const float *query, ...) (src/kgraph.cpp:1560-1623)
    void get_neighbors(const float *query, unsigned init_point, unsigned L, vector<Nb> &pool) {
        const unsigned nd = N();
        pool.assign(L + 1, Nb());
        vector<unsigned> init_ids(L);
        vector<char> flags(nd, 0);
        L = 0;
        for (unsigned i = 0; i < init_ids.size() && i < nh[init_point].pool.size(); ++i) {
            init_ids[i] = nh[init_point].pool[i].id;
            if (init_ids[i] < nd) flags[init_ids[i]] = 1;
            ++L;
        }
        while (L < init_ids.size()) {
            const unsigned id = R_() % nd;
            if (flags[id]) continue;
            init_ids[L++] = id;
            flags[id] = 1;
        }
        L = 0;
        for (unsigned i = 0; i < init_ids.size(); ++i) {
            const unsigned id = init_ids[i];
            if (id >= nd) continue;
            pool[i] = Nb(id, o.qdist(query, id), true);
            ++L;
            st->d_ep++;
        }
        std::sort(pool.begin(), pool.begin() + L);
        int k = 0;
        while (k < (int)L) {
            int nk = (int)L;
            if (pool[k].flag) {
                pool[k].flag = false;
                const unsigned n = pool[k].id;
                for (unsigned m = 0; m < nh[n].nn_nsg.size(); ++m) {
                    const unsigned id = nh[n].nn_nsg[m].id;
                    if (flags[id]) continue;
                    flags[id] = 1;
                    const float d = o.qdist(query, id);
                    st->d_ep++;
                    if (d >= pool[L - 1].dist) continue;
                    const int r = insert_into_pool(pool.data(), L, Nb(id, d, true));
                    if (L + 1 < pool.size()) ++L;
                    if (r < nk) nk = r;
                }
            }
            if (nk <= k) k = nk;
            else ++k;
        }
    }

    vector<float> center;
    void compute_center() {
        if (!center.empty()) return;
        // Calcenter over all rows (data_size), float accumulation in row order
        center.assign(o.dpad, 0.f);
        for (unsigned i = 0; i < o.n; ++i)
            for (unsigned j = 0; j < o.d; ++j) center[j] += o.row(i)[j];
        for (unsigned j = 0; j < o.d; ++j) center[j] /= o.n;
    }

    // ComputeEp (src/kgraph.cpp:1528-1536); init_point = rand() % N unless given
    void compute_ep(int64_t given_init = -1) {
        const double t0 = now();
        compute_center();
        vector<Nb> pool;
        const unsigned init_point = given_init >= 0 ? (unsigned)given_init : R_() % N();
        get_neighbors(center.data(), init_point, P.search_L, pool);
        ep = pool[0].id;
        st->t_ep += now() - t0;
    }

    void dfs(vector<char> &flag, unsigned root, unsigned &cnt) {
        const unsigned nd = N();
        unsigned tmp = root;
        std::stack<unsigned> s;
        s.push(root);
        if (!flag[root]) ++cnt;
        flag[root] = 1;
        while (!s.empty()) {
            unsigned next = nd + 1;
            for (auto &e : nh[tmp].nn_nsg)
                if (!flag[e.id]) { next = e.id; break; }
            if (next == nd + 1) {
                s.pop();
                if (s.empty()) break;
                tmp = s.top();
                continue;
            }
            tmp = next;
            flag[tmp] = 1;
            s.push(tmp);
            ++cnt;
        }
    }

    void findroot(vector<char> &flag, unsigned &root, unsigned &start) {
        const unsigned nd = N();
        unsigned id = nd;
        for (unsigned i = start; i < nd; ++i)
            if (!flag[i]) { id = i; start = i; break; }
        if (id == nd) return;
        bool found = false;
        for (unsigned i = 0; i < nh[id].L; ++i) {
            const unsigned pid = nh[id].pool[i].id;
            if (pid < nd && flag[pid] && nh[pid].nn_nsg.size() < P.nsg_R) { root = pid; found = true; break; }
        }
        if (!found) {
            uint64_t cnt = 0;
            while (true) {
                ++cnt;
                const uint64_t factor = cnt / nd + 1;
                const unsigned rid = R_() % nd;
                if (flag[rid] && nh[rid].nn_nsg.size() < P.nsg_R * factor) { root = rid; break; }
            }
        }
        nh[root].nn_nsg.push_back(Nb(id, o.dist(root, id), true));
        st->n_attached++;
        st->d_conn++;
    }

    void tree_grow() {
        const double t0 = now();
        const unsigned nd = N();
        unsigned root = ep, cnt = 0, start = 0;
        vector<char> flags(nd, 0);
        while (cnt < nd) {
            dfs(flags, root, cnt);
            if (cnt >= nd) break;
            findroot(flags, root, start);
        }
        for (unsigned i = 0; i < nd; ++i) width = std::max<unsigned>(width, (unsigned)nh[i].nn_nsg.size());
        const double dt = now() - t0;
        if (loopcount < FK_MAX_LOOPS) st->tree_grow_time[loopcount] = dt;
        st->t_conn += dt;
    }

    void update_nhoods() {
        const unsigned nd = N();
#pragma omp parallel for schedule(dynamic, 100)
        for (unsigned i = 0; i < nd; ++i) {
            auto &h = nh[i];
            for (auto &e : h.nn_nsg)
                e.isnew = loopcount == 0 ? true
                                         : std::find(h.old_nsg.begin(), h.old_nsg.end(), e.id) == h.old_nsg.end();
            h.old_nsg.clear();
            for (auto &e : h.nn_nsg) h.old_nsg.push_back(e.id);
        }
    }

    // BridgeView (src/kgraph.cpp:2478-2663)
    void bv_init() {
        const unsigned nd = N(), S_ = P.massq_S, L = MAXL;
        const unsigned cur_loop = loopcount - 1;
        const unsigned cur_L = std::min(P.search_L + cur_loop * P.step, MAXL);
#pragma omp parallel for
        for (unsigned i = 0; i < nd; ++i) {
            auto &h = nh[i];
            if (cur_loop == 0) {
                for (unsigned j = 0; j < h.L; ++j) { h.pool[j].flag = true; h.pool[j].isnew = false; }
            } else {
                const unsigned pre_L = std::min(P.search_L + (cur_loop - 1) * P.step, MAXL);
                for (unsigned j = 0; j < pre_L && j < h.L; ++j) { h.pool[j].flag = true; h.pool[j].isnew = false; }
                for (unsigned j = pre_L; j < cur_L && j < h.L; ++j) {
                    const bool e = h.pool[j].flag;
                    h.pool[j].isnew = e;
                    h.pool[j].flag = true;
                }
                for (unsigned j = cur_L; j < h.L; ++j) { h.pool[j].flag = true; h.pool[j].isnew = false; }
            }
            h.pool.resize(L + 1);
            h.radius = h.pool.back().dist;
            h.rnn_new.clear();
            h.rnn_new_flag.clear();
        }
        (void)S_;
    }

    unsigned bv_update(bool capped) {
        unsigned count = 0;
        const unsigned nd = N(), S_ = P.massq_S;
        const unsigned L = std::min(P.search_L + (loopcount - 1) * P.step, MAXL);
        if (capped) {
#pragma omp parallel for
            for (unsigned i = 0; i < nd; ++i) { nh[i].rnn_new.clear(); nh[i].rnn_new_flag.clear(); }
        }
#pragma omp parallel for reduction(+ : count)
        for (unsigned i = 0; i < nd; ++i) {
            auto &h = nh[i];
            unsigned cnt = 0;
            for (unsigned j = 0; j < L && j < h.L && (cnt < S_ || !capped); ++j) {
                if (h.pool[j].flag) {
                    const unsigned id = h.pool[j].id;
                    {
                        std::lock_guard<Spin> g(nh[id].lock);
                        nh[id].rnn_new.push_back(i);
                        nh[id].rnn_new_flag.push_back(h.pool[j].isnew);
                    }
                    h.pool[j].flag = false;
                    ++cnt;
                }
            }
            count += cnt;
        }
        return count;
    }

    unsigned parallel_insert(Nhood &h, unsigned id, float dist) {
        if (dist > h.radius) return (unsigned)h.pool.size();
        std::lock_guard<Spin> g(h.lock);
        Nb nn(id, dist, true);
        nn.isnew = true;
        const unsigned l = (unsigned)insert_into_pool(h.pool.data(), h.L, nn);
        if (l <= h.L) {
            if (h.L + 1 < h.pool.size()) ++h.L;
            else h.radius = h.pool[h.L - 1].dist;
        }
        return l;
    }

    uint64_t bv_join() {
        uint64_t cc = 0;
        const unsigned nd = N();
#pragma omp parallel for default(shared) schedule(dynamic, 100) reduction(+ : cc)
        for (unsigned i = 0; i < nd; ++i) {
            const unsigned q = nd - i - 1;
            auto &h = nh[q];
            for (size_t k = 0; k < h.rnn_new.size(); ++k) {
                const unsigned id = h.rnn_new[k];
                const bool ifnew = h.rnn_new_flag[k];
                for (auto &e : h.nn_nsg) {
                    if (id == e.id) continue;
                    if (!ifnew && !e.isnew) continue;
                    ++cc;
                    parallel_insert(nh[id], e.id, o.dist(id, e.id));
                }
            }
        }
        return cc;
    }

    void bridgeview(unsigned li) {
        const double t0 = now();
        const unsigned nd = N(), S_ = P.massq_S;
        uint64_t cc = 0;
        unsigned rounds = 0;
        bv_init();
        while (true) {
            unsigned cnt = bv_update(true);
            float delta = (float)cnt / nd / S_;
            bool last = false;
            if (delta < 0.5f) {
                cnt += bv_update(false);
                last = true;
            }
            cc += bv_join();
            ++rounds;
            if (last) break;
        }
        for (auto &h : nh) vector<unsigned>().swap(h.rnn_new);
        if (li < FK_MAX_LOOPS) {
            st->search_cc[li] = (float)((double)cc / ((double)nd * (nd - 1) / 2.0));
            st->bv_rounds[li] = rounds;
        }
        st->d_bv += cc;
        st->d_bv_computed += cc;  // the restatement computes every pair it counts
        st->t_bv += now() - t0;
    }

    void calculate_nsg() {
        unsigned mx = 0, mn = 1000000, avg = 0;
        for (unsigned i = 0; i < N(); ++i) {
            const unsigned s = (unsigned)nh[i].nn_nsg.size();
            mx = std::max(mx, s);
            mn = std::min(mn, s);
            avg += s;
        }
        avg = (unsigned)(avg / (1.0 * N()));
        st->aod = avg;
        st->max_degree = mx;
        st->min_degree = mn;
    }

    void build(int type) {
        const double t0 = now();
        rp.rule = rule_of(type);
        rp.tau = P.tau;
        rp.thr = alpha_thr(P.angle);
        shrink();
        generate_control();
        nndescent();
        const float rec = eval_recall(MAXL);
        if (st->n_recall < FK_MAX_LOOPS + 64) st->recall[st->n_recall++] = rec;
        if (st->n_kgraph < FK_MAX_LOOPS + 64) st->kgraph_time[st->n_kgraph++] = now() - t0;
        pruneAgain.assign(N(), 1);
        init_nhoods();
        loopcount = 0;
        const unsigned iters = P.loop_i + 1;
        for (unsigned iter = 1; iter <= iters; ++iter) {
            const unsigned li = iter - 1;
            const double tb = now();
            link(li);
            compute_ep();
            if (type != FK_PG_HNSW_LAYER) tree_grow();
            update_nhoods();
            if (li < FK_MAX_LOOPS) {
                st->buildnsg_time[li] = now() - tb;
                st->con_time[li] = now() - t0;
            }
            ++loopcount;
            st->n_loops = iter;
            if (iter == iters) {
                calculate_nsg();
                break;
            }
            const double ts = now();
            bridgeview(li);
            if (li < FK_MAX_LOOPS) st->search_time[li] = now() - ts;
            const float r2 = eval_recall(MAXL);
            if (st->n_recall < FK_MAX_LOOPS + 64) st->recall[st->n_recall++] = r2;
            if (st->n_kgraph < FK_MAX_LOOPS + 64) st->kgraph_time[st->n_kgraph++] = now() - t0;
        }
        st->t_total = now() - t0;
    }
};

uint64_t pack(float d, uint32_t id) {
    uint32_t b;
    memcpy(&b, &d, 4);
    return ((uint64_t)b << 32) | id;
}

}  // namespace

extern "C" {

// Canonical (reference-order) squared L2 and its FP64 counterpart.
float fko_l2(const float *a, const float *b, unsigned dpad) { return l2_ref(a, b, dpad); }
double fko_l2_f64(const float *a, const float *b, unsigned d) { return l2_f64(a, b, d); }

void fko_pair_distances(const float *rows, size_t stride_bytes, unsigned n, unsigned d, const uint32_t *a,
                        const uint32_t *b, size_t m, float *out, double *out64) {
    Rows r;
    r.load(rows, stride_bytes, n, d);
#pragma omp parallel for
    for (size_t k = 0; k < m; ++k) {
        out[k] = r.dist(a[k], b[k]);
        if (out64) out64[k] = l2_f64(r.row(a[k]), r.row(b[k]), d);
    }
}

// Sequence of InsertIntoPool calls on one pool of capacity L (+1 overflow
// slot) with parallel_try_insert's bookkeeping; returns the final count and
// writes the pool.
unsigned fko_pool_insert_sequence(const uint32_t *ids, const float *dists, unsigned n, unsigned L,
                                  uint32_t *out_ids, float *out_d, int *ret) {
    Nhood h;
    h.pool.assign(L + 1, Nb());
    h.L = 0;
    h.radius = std::numeric_limits<float>::max();
    for (unsigned k = 0; k < n; ++k) {
        if (h.L == 0) {
            h.pool[0] = Nb(ids[k], dists[k], true);
            h.L = 1;
            if (ret) ret[k] = 0;
            continue;
        }
        const unsigned r = h.try_insert(ids[k], dists[k]);
        if (ret) ret[k] = (int)r;
    }
    for (unsigned k = 0; k < h.L; ++k) { out_ids[k] = h.pool[k].id; out_d[k] = h.pool[k].dist; }
    return h.L;
}

// The full-rule pruner (or its reversed form) over given pools.
void fko_prune(int type, const float *rows, size_t stride_bytes, unsigned n, unsigned d, const uint64_t *pools,
               const uint32_t *pool_len, unsigned pool_stride, const uint32_t *q, unsigned n_tasks,
               unsigned oracle_size, const fk_params *p, int reversed, uint32_t *out_ids, float *out_d,
               unsigned out_stride, uint32_t *out_len, uint64_t *cc_out) {
    Rows r;
    r.load(rows, stride_bytes, n, d);
    r.oracle_n = std::min(oracle_size, n);
    fk_stats st;
    memset(&st, 0, sizeof(st));
    uint64_t cc = 0;
    RuleP rp{rule_of(type), p->tau, alpha_thr(p->angle)};
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : cc)
    for (unsigned t = 0; t < n_tasks; ++t) {
        const unsigned qq = q ? q[t] : t;
        const unsigned len = pool_len[t];
        vector<Nb> pool(len);
        for (unsigned k = 0; k < len; ++k) {
            const uint64_t e = pools[(size_t)t * pool_stride + k];
            const uint32_t bits = (uint32_t)(e >> 32);
            float dd;
            memcpy(&dd, &bits, 4);
            pool[k] = Nb((uint32_t)e, dd, true);
        }
        vector<Nb> out;
        if (len == 0) { out_len[t] = 0; continue; }
        unsigned start = 0;
        if (pool[0].id == qq) ++start;
        if (start >= len) { out_len[t] = 0; continue; }
        out.push_back(pool[start]);
        const unsigned maxc = reversed ? 0xFFFFFFFFu : p->search_K;
        const unsigned lim_id = reversed ? 0xFFFFFFFFu : r.oracle_n;
        while (out.size() < p->nsg_R && (++start) < len && start < maxc) {
            auto &e = pool[start];
            if (e.id >= lim_id) continue;
            bool occl = false;
            for (unsigned k = 0; k < out.size(); ++k) {
                if (e.id == out[k].id) { occl = true; break; }
                ++cc;
                const float djk = r.dist(out[k].id, e.id);
                if (rp.occl(djk, e.dist, out[k].dist)) { occl = true; break; }
            }
            if (!occl) out.push_back(e);
        }
        out_len[t] = (uint32_t)out.size();
        for (size_t k = 0; k < out.size(); ++k) {
            out_ids[(size_t)t * out_stride + k] = out[k].id;
            if (out_d) out_d[(size_t)t * out_stride + k] = out[k].dist;
        }
    }
    if (cc_out) *cc_out = cc;
}

// Exact kNN (LinearSearch semantics: ties in id order), self excluded when
// self_ids is given.
void fko_knn(const float *rows, size_t stride_bytes, unsigned n, unsigned d, const float *queries, unsigned n_q,
             unsigned K, const uint32_t *self_ids, uint32_t *out_ids, float *out_d) {
    Rows r;
    r.load(rows, stride_bytes, n, d);
    vector<float> qbuf((size_t)n_q * r.dpad, 0.f);
    for (unsigned i = 0; i < n_q; ++i) memcpy(qbuf.data() + (size_t)i * r.dpad, queries + (size_t)i * d, (size_t)d * 4);
#pragma omp parallel for schedule(dynamic, 1)
    for (unsigned qi = 0; qi < n_q; ++qi) {
        vector<Nb> nns(K + 1);
        unsigned k = 0;
        const float *qv = qbuf.data() + (size_t)qi * r.dpad;
        for (unsigned id = 0; id < n; ++id) {
            if (self_ids && self_ids[qi] == id) continue;
            update_knn_list(nns.data(), k, Nb(id, r.qdist(qv, id), true));
            if (k < K) ++k;
        }
        for (unsigned j = 0; j < K; ++j) {
            out_ids[(size_t)qi * K + j] = j < k ? nns[j].id : 0xFFFFFFFFu;
            if (out_d) out_d[(size_t)qi * K + j] = j < k ? nns[j].dist : INFINITY;
        }
    }
}

struct fko_build_opts {
    int threads;          // <= 0: all
    int full_rule_only;   // never use *_sync_prune2
    int check_prune2;     // run both rules where prune2 applies, count diffs
    int verbose;
};

// Full CPU build; the output graph is a malloc'd CSR (free with fko_free).
int fko_build(int type, const float *rows, size_t stride_bytes, unsigned n_total, unsigned n_used, unsigned d,
              const fk_params *p, const fko_build_opts *opt, uint64_t **offsets, uint32_t **ids, uint32_t *ep,
              uint32_t *width, fk_stats *st, uint64_t *prune2) {
    if (opt && opt->threads > 0) omp_set_num_threads(opt->threads);
    Rows r;
    r.load(rows, stride_bytes, n_total, d);
    r.oracle_n = std::min(n_used, n_total);
    memset(st, 0, sizeof(*st));
    Builder b(r, *p, st);
    b.full_rule_only = opt && opt->full_rule_only;
    b.check_prune2 = opt && opt->check_prune2;
    b.verbose = opt && opt->verbose;
    b.build(type);
    const unsigned N = r.oracle_n;
    uint64_t *off = (uint64_t *)malloc(((size_t)N + 1) * 8);
    off[0] = 0;
    for (unsigned i = 0; i < N; ++i) off[i + 1] = off[i] + b.nh[i].nn_nsg.size();
    uint32_t *out = (uint32_t *)malloc(std::max<uint64_t>(off[N], 1) * 4);
    for (unsigned i = 0; i < N; ++i)
        for (size_t k = 0; k < b.nh[i].nn_nsg.size(); ++k) out[off[i] + k] = b.nh[i].nn_nsg[k].id;
    *offsets = off;
    *ids = out;
    *ep = b.ep;
    *width = b.width;
    if (prune2) {  // nodes where the incremental rule differed, nodes compared
        prune2[0] = b.prune2_diff;
        prune2[1] = b.prune2_nodes;
    }
    return 0;
}

// NN-descent only; pools are written packed (dist bits << 32 | id).
int fko_nndescent(const float *rows, size_t stride_bytes, unsigned n, unsigned d, const fk_params *p, int threads,
                  uint64_t *pools, uint32_t *pool_len, unsigned pool_stride, fk_stats *st) {
    if (threads > 0) omp_set_num_threads(threads);
    Rows r;
    r.load(rows, stride_bytes, n, d);
    memset(st, 0, sizeof(*st));
    Builder b(r, *p, st);
    b.shrink();
    b.generate_control();
    b.nndescent();
    for (unsigned i = 0; i < n; ++i) {
        pool_len[i] = std::min(b.nh[i].L, pool_stride);
        for (unsigned k = 0; k < pool_len[i]; ++k) pools[(size_t)i * pool_stride + k] = pack(b.nh[i].pool[k].dist, b.nh[i].pool[k].id);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// The query evaluator's search (the product's fk_search) restated on the
// CPU: the best-first loop of get_neighbors (src/kgraph.cpp:1560-1623) over a
// CSR graph, one neighbour at a time with InsertIntoPool-style bookkeeping,
// in the build's canonical (dist, id) order.  Its start set is the
// evaluator's own: ep, ep's out-list, then (rand32(query, k) % n) fill; the
// fill hash is the evaluator's counter-based generator (a splitmix64
// finaliser), written out here so both sides draw the same ids.
// ---------------------------------------------------------------------------
static uint64_t mix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static uint32_t rand32(uint64_t seed, uint64_t stream, uint64_t k) {
    return (uint32_t)(mix64(seed * 0xD1B54A32D192ED03ull ^ mix64(stream * 0x8CB92BA72F3D8DD7ull + k)) >> 32);
}

void fko_search(const float *rows, size_t stride_bytes, unsigned n_total, unsigned d, const uint64_t *offsets,
                const uint32_t *ids, unsigned n, unsigned ep, const float *queries, unsigned n_q, unsigned L,
                unsigned K, uint64_t seed, uint32_t *out_ids, uint64_t *n_dist) {
    Rows r;
    r.load(rows, stride_bytes, n_total, d);
    uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : total)
    for (unsigned qi = 0; qi < n_q; ++qi) {
        vector<float> q(r.dpad, 0.f);
        memcpy(q.data(), queries + (size_t)qi * d, (size_t)d * 4);
        vector<char> seen(n, 0);
        vector<unsigned> start;
        const unsigned want = std::min(L, n);
        auto take = [&](unsigned id) {
            if (id < n && !seen[id] && start.size() < want) { seen[id] = 1; start.push_back(id); }
        };
        take(ep);
        for (uint64_t e = offsets[ep]; e < offsets[ep + 1]; ++e) take(ids[e]);
        for (uint64_t k = 0; start.size() < want; ++k) take(rand32(seed, qi, k) % n);
        struct E { float d; uint32_t id; bool open; };
        auto before = [](const E &a, const E &b) { return a.d < b.d || (a.d == b.d && a.id < b.id); };
        vector<E> pool;
        for (unsigned id : start) pool.push_back({r.qdist(q.data(), id), id, true});
        total += pool.size();
        std::sort(pool.begin(), pool.end(), before);
        size_t k = 0;
        while (k < pool.size()) {
            size_t nk = pool.size();
            if (pool[k].open) {
                pool[k].open = false;
                const unsigned u = pool[k].id;
                // neighbours in groups of 64 list positions, each group filtered
                // by the worst distance kept when it starts (the evaluator's
                // wavefront width); the kept set is the best L either way
                for (uint64_t e0 = offsets[u]; e0 < offsets[u + 1]; e0 += 64) {
                    const float worst = pool.back().d;
                    for (uint64_t e = e0; e < std::min(e0 + 64, offsets[u + 1]); ++e) {
                        const unsigned v = ids[e];
                        if (v >= n || seen[v]) continue;
                        seen[v] = 1;
                        ++total;
                        E c{r.qdist(q.data(), v), v, true};
                        if (!(c.d < worst)) continue;  // `dist >= pool[L-1].dist` skips
                        const size_t at =
                            (size_t)(std::lower_bound(pool.begin(), pool.end(), c, before) - pool.begin());
                        if (at >= L) continue;
                        pool.insert(pool.begin() + at, c);
                        if (pool.size() > L) pool.pop_back();
                        nk = std::min(nk, at);
                    }
                }
            }
            k = nk <= k ? nk : k + 1;
        }
        for (unsigned j = 0; j < K; ++j) out_ids[(size_t)qi * K + j] = j < pool.size() ? pool[j].id : 0xFFFFFFFFu;
    }
    if (n_dist) *n_dist = total;
}

// ---------------------------------------------------------------------------
// Step entry points: one step of the build's loop on a supplied state (the
// layout of fk_state in fk_api.h), run by the restated code above.
// ---------------------------------------------------------------------------
static Nb entry_of(uint64_t key, uint8_t flags) {
    const uint32_t bits = (uint32_t)(key >> 32);
    float d;
    memcpy(&d, &bits, 4);
    Nb e((uint32_t)key, d, (flags & 1u) != 0);
    e.isnew = (flags & 2u) != 0;
    return e;
}

// Pools of the state into the builder's Nhoods, each sized cap + 1 (the
// overflow slot and unfilled slots default-constructed: id ~0, FLT_MAX).
static void load_pools(Builder &b, const fk_state *s, unsigned cap) {
    for (unsigned i = 0; i < s->n; ++i) {
        auto &h = b.nh[i];
        const unsigned len = s->pool_len[i];
        h.pool.assign(std::max(cap, len) + 1, Nb());
        for (unsigned k = 0; k < len; ++k)
            h.pool[k] = entry_of(s->pool[(size_t)i * s->pool_cap + k],
                                 s->pool_flags ? s->pool_flags[(size_t)i * s->pool_cap + k] : (uint8_t)1);
        h.L = len;
    }
}

// Link step: the full rule for every node, InterInsert, reversed prune (a
// list left at or below nsg_R has its appended reverse edges in (dist, id)
// order, the build's canonical order).
int fko_step_link(int type, const float *rows, size_t stride_bytes, unsigned n_total, unsigned d,
                  const fk_params *p, fk_state *s, fk_stats *st_out) {
    Rows r;
    r.load(rows, stride_bytes, n_total, d);
    r.oracle_n = s->n;
    fk_stats local;
    fk_stats *st = st_out ? st_out : &local;
    memset(st, 0, sizeof(*st));
    Builder b(r, *p, st);
    b.rp = RuleP{rule_of(type), p->tau, alpha_thr(p->angle)};
    b.shrink();
    b.full_rule_only = true;
    b.pruneAgain.assign(s->n, 1);
    load_pools(b, s, s->pool_cap);
    b.link(0);
    for (unsigned i = 0; i < s->n; ++i) {
        auto &out = b.nh[i].nn_nsg;
        if (out.size() > s->out_cap) return 1;
        s->out_len[i] = (uint32_t)out.size();
        for (size_t k = 0; k < out.size(); ++k) {
            s->out_id[(size_t)i * s->out_cap + k] = out[k].id;
            s->out_dist[(size_t)i * s->out_cap + k] = out[k].dist;
        }
    }
    return 0;
}

// UpdateNhoods on the state's out-lists and previous out-lists.
int fko_step_update_nhoods(fk_state *s) {
    Rows r;  // UpdateNhoods reads no rows
    r.oracle_n = s->n;
    fk_params p;
    memset(&p, 0, sizeof(p));
    fk_stats st;
    memset(&st, 0, sizeof(st));
    Builder b(r, p, &st);
    b.loopcount = s->loopcount;
    for (unsigned i = 0; i < s->n; ++i) {
        auto &h = b.nh[i];
        for (unsigned k = 0; k < s->out_len[i]; ++k) h.nn_nsg.push_back(Nb(s->out_id[(size_t)i * s->out_cap + k], 0.f));
        for (unsigned k = 0; k < s->prev_len[i]; ++k) h.old_nsg.push_back(s->prev_id[(size_t)i * s->out_cap + k]);
    }
    b.update_nhoods();
    for (unsigned i = 0; i < s->n; ++i) {
        const auto &h = b.nh[i];
        for (size_t k = 0; k < h.nn_nsg.size(); ++k) s->out_isnew[(size_t)i * s->out_cap + k] = h.nn_nsg[k].isnew;
        if (h.old_nsg.size() > s->out_cap) return 1;
        s->prev_len[i] = (uint32_t)h.old_nsg.size();
        for (size_t k = 0; k < h.old_nsg.size(); ++k) s->prev_id[(size_t)i * s->out_cap + k] = h.old_nsg[k];
    }
    return 0;
}

// One BridgeView pass at s->loopcount.  Returns 1 when pool_cap is not the
// build's MAXL.
int fko_step_bridgeview(const float *rows, size_t stride_bytes, unsigned n_total, unsigned d, const fk_params *p,
                        fk_state *s, fk_stats *st_out) {
    Rows r;
    r.load(rows, stride_bytes, n_total, d);
    r.oracle_n = s->n;
    fk_stats local;
    fk_stats *st = st_out ? st_out : &local;
    memset(st, 0, sizeof(*st));
    Builder b(r, *p, st);
    b.shrink();
    if (s->pool_cap != b.MAXL) return 1;
    load_pools(b, s, b.MAXL);
    for (unsigned i = 0; i < s->n; ++i) {
        auto &h = b.nh[i];
        for (unsigned k = 0; k < s->out_len[i]; ++k) {
            Nb e(s->out_id[(size_t)i * s->out_cap + k], 0.f, false);
            e.isnew = s->out_isnew[(size_t)i * s->out_cap + k] != 0;
            h.nn_nsg.push_back(e);
        }
    }
    b.loopcount = s->loopcount;
    b.bridgeview(0);
    for (unsigned i = 0; i < s->n; ++i) {
        const auto &h = b.nh[i];
        s->pool_len[i] = h.L;
        for (unsigned k = 0; k < h.L; ++k) {
            s->pool[(size_t)i * s->pool_cap + k] = pack(h.pool[k].dist, h.pool[k].id);
            s->pool_flags[(size_t)i * s->pool_cap + k] = (uint8_t)((h.pool[k].flag ? 1u : 0u) | (h.pool[k].isnew ? 2u : 0u));
        }
    }
    return 0;
}

// The connectivity pass (KGraphConstructor::tree_grow / findroot / DFS,
// src/kgraph.cpp:2145-2260) on the state's out-lists and pools from entry
// point ep: the attached edges are appended to the out-lists (with their
// distances).  Returns 1 when an out-list would pass out_cap.
int fko_step_connectivity(const float *rows, size_t stride_bytes, unsigned n_total, unsigned d,
                          const fk_params *p, fk_state *s, unsigned ep, fk_stats *st_out) {
    Rows r;
    r.load(rows, stride_bytes, n_total, d);
    r.oracle_n = s->n;
    fk_stats local;
    fk_stats *st = st_out ? st_out : &local;
    memset(st, 0, sizeof(*st));
    Builder b(r, *p, st);
    b.shrink();
    load_pools(b, s, s->pool_cap);
    for (unsigned i = 0; i < s->n; ++i)
        for (unsigned k = 0; k < s->out_len[i]; ++k)
            b.nh[i].nn_nsg.push_back(
                Nb(s->out_id[(size_t)i * s->out_cap + k], s->out_dist[(size_t)i * s->out_cap + k], true));
    b.ep = ep;
    b.tree_grow();
    for (unsigned i = 0; i < s->n; ++i) {
        const auto &out = b.nh[i].nn_nsg;
        if (out.size() > s->out_cap) return 1;
        s->out_len[i] = (uint32_t)out.size();
        for (size_t k = 0; k < out.size(); ++k) {
            s->out_id[(size_t)i * s->out_cap + k] = out[k].id;
            s->out_dist[(size_t)i * s->out_cap + k] = out[k].dist;
        }
    }
    return 0;
}

// ComputeEp from init_node's pool over the state's out-lists; returns ep.
unsigned fko_step_entry_point(const float *rows, size_t stride_bytes, unsigned n_total, unsigned d,
                              const fk_params *p, const fk_state *s, unsigned init_node) {
    Rows r;
    r.load(rows, stride_bytes, n_total, d);
    r.oracle_n = s->n;
    fk_stats st;
    memset(&st, 0, sizeof(st));
    Builder b(r, *p, &st);
    b.shrink();
    load_pools(b, s, std::max(s->pool_cap, b.MAXL));
    for (unsigned i = 0; i < s->n; ++i)
        for (unsigned k = 0; k < s->out_len[i]; ++k)
            b.nh[i].nn_nsg.push_back(Nb(s->out_id[(size_t)i * s->out_cap + k], 0.f));
    b.compute_ep((int64_t)init_node);
    return b.ep;
}

// The stop rule's delta and recall of the state's pools (the expressions of
// Builder::nndescent and Builder::eval_recall; delta counts the flags of the
// first min(K, pool_len) entries).
void fko_step_stop_stats(const fk_params *p, const fk_state *s, unsigned n_ctrl, const uint32_t *ctrl_ids,
                         const uint32_t *gt, unsigned gt_stride, float *delta, float *recall) {
    Rows r;
    r.oracle_n = s->n;
    fk_stats st;
    memset(&st, 0, sizeof(st));
    Builder b(r, *p, &st);
    b.shrink();
    load_pools(b, s, s->pool_cap);
    const unsigned K = b.P.K;
    double dsum = 0;
    for (auto &h : b.nh) {
        unsigned c = 0;
        const unsigned lim = std::min(K, h.L);
        for (unsigned i = 0; i < lim; ++i) if (h.pool[i].flag) ++c;
        dsum += (double)c / K;
    }
    *delta = (float)(dsum / s->n);
    b.controls.assign(n_ctrl, Control());
    for (unsigned c = 0; c < n_ctrl; ++c) {
        b.controls[c].id = ctrl_ids[c];
        for (unsigned k = 0; k < K; ++k) b.controls[c].nbrs.push_back(Nb(gt[(size_t)c * gt_stride + k], 0.f));
    }
    *recall = b.eval_recall(K);
}

void fko_free(void *p) { free(p); }

}  // extern "C"
