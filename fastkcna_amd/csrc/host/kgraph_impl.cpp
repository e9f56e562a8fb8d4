// KGraphImpl: the reference's KGraph build API (src/kgraph.cpp:3720-3971)
// as a thin adapter over the C ABI.  The GPU does the whole build; this file
// keeps the reference's host-side side effects byte-compatible: the NSG index
// (SaveNsg, src/kgraph.cpp:1704-1724), the appended CSV log (SaveResult,
// :1725-1807), the hnswlib link-list swaps (SwapHNSWLayer :3598-3614,
// FatherSwapHNSWLayer0 :499-520) and the (empty) KNNGRAPH save (:463-493).
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "fk_api.h"
#include "kgraph-data.h"
#include "kgraph.h"

namespace kgraph {

unsigned verbosity = default_verbosity;

unsigned SearchOracle::search(unsigned K, float epsilon, unsigned *ids, float *dists) const {
    // brute force, ascending distance, ties in id order
    std::vector<std::pair<float, unsigned>> v;
    const unsigned N = size();
    v.reserve(N);
    for (unsigned k = 0; k < N; ++k) {
        const float d = operator()(k);
        if (d <= epsilon) v.emplace_back(d, k);
    }
    const unsigned L = std::min<unsigned>(K, (unsigned)v.size());
    std::partial_sort(v.begin(), v.begin() + L, v.end());
    for (unsigned k = 0; k < L; ++k) {
        if (ids) ids[k] = v[k].second;
        if (dists) dists[k] = v[k].first;
    }
    return L;
}

namespace {

[[noreturn]] void raise(fk_status st, const std::string &what) {
    switch (st) {
        case FK_E_INVALID: throw invalid_argument(what);
        case FK_E_IO: throw io_error(what);
        default: throw runtime_error(what);
    }
}

fk_params to_fk(KGraph::IndexParams const &p) {
    fk_params f;
    memset(&f, 0, sizeof(f));
    f.K = p.K; f.L = p.L; f.S = p.S; f.R = p.R; f.iterations = p.iterations; f.controls = p.controls;
    f.seed = p.seed; f.delta = p.delta; f.recall = p.recall; f.nsg_R = p.nsg_R; f.search_L = p.search_L;
    f.search_K = p.search_K; f.step = p.step; f.massq_S = p.massq_S; f.loop_i = p.loop_i; f.tau = p.tau;
    f.angle = p.angle;
    f.flags = verbosity > 1 ? FK_FLAG_VERBOSE : 0;
    f.flags |= p.gpu_flags & (FK_FLAG_GRAM | FK_FLAG_GRAM_BF16 | FK_FLAG_GRAM_F16 | FK_FLAG_SCREEN |
                              FK_FLAG_SCREEN_NOTRAIN);
    return f;
}

}  // namespace

class KGraphImpl : public KGraph {
    fk_handle *h_ = nullptr;
    int dev_ = -1;
    // What the resident rows were uploaded from: the Matrix address, shape
    // and a fingerprint of its contents, so that a matrix rewritten in place
    // or reallocated at the same address is uploaded again.
    const void *up_rows_ = nullptr;
    unsigned up_n_ = 0, up_dim_ = 0;
    size_t up_stride_ = 0;
    uint64_t up_sum_ = 0;

    // FNV-1a over the bytes of 64 rows spread over the matrix (first and
    // last included): O(64 d), negligible next to an upload.
    static uint64_t fingerprint(const MatrixOracle<float, metric::l2sqr> *mo) {
        uint64_t f = 1469598103934665603ull;
        const unsigned n = mo->proxy_size();
        const size_t bytes = (size_t)mo->dim() * sizeof(float);
        for (unsigned k = 0; k < 64 && n > 0; ++k) {
            const unsigned r = (unsigned)((uint64_t)k * (n - 1) / 63);
            const unsigned char *b = (const unsigned char *)mo->rows() + (size_t)r * mo->stride();
            for (size_t i = 0; i < bytes; ++i) f = (f ^ b[i]) * 1099511628211ull;
        }
        return f;
    }
    std::vector<std::vector<unsigned>> temp_graph_;  // level-0 adjacency (HNSW)

    fk_handle *handle(int device) {
        if (h_ && dev_ == device) return h_;
        if (h_) fk_close(h_);
        h_ = nullptr;
        up_rows_ = nullptr;
        const fk_status st = fk_open(device, &h_);
        if (st != FK_OK) raise(st, "fk_open: no usable gfx950 device " + std::to_string(device));
        dev_ = device;
        return h_;
    }

    void check(fk_status st, const char *what) {
        if (st != FK_OK) raise(st, std::string(what) + ": " + fk_last_error(h_));
    }

    // The GPU reads rows straight from a MatrixOracle<float, l2sqr>.
    void upload(IndexOracle const &oracle, int device) {
        auto mo = dynamic_cast<MatrixOracle<float, metric::l2sqr> const *>(&oracle);
        if (!mo)
            throw invalid_argument(
                "the GPU build needs a MatrixOracle<float, metric::l2sqr> (squared L2 over float rows)");
        fk_handle *h = handle(device);
        const uint64_t sum = fingerprint(mo);
        if (up_rows_ == mo->rows() && up_n_ == mo->proxy_size() && up_stride_ == mo->stride() &&
            up_dim_ == mo->dim() && up_sum_ == sum)
            return;
        up_rows_ = nullptr;
        check(fk_dataset_upload(h, mo->rows(), mo->stride(), mo->proxy_size(), mo->dim()), "fk_dataset_upload");
        up_rows_ = mo->rows();
        up_n_ = mo->proxy_size();
        up_stride_ = mo->stride();
        up_dim_ = mo->dim();
        up_sum_ = sum;
    }

    struct Graph {
        fk_graph g;
        Graph() { memset(&g, 0, sizeof(g)); }
        ~Graph() { fk_graph_free(&g); }
    };

    void run(fk_pg_type t, IndexOracle const &oracle, IndexParams const &p, Graph &out) {
        upload(oracle, p.device);
        const fk_params fp = to_fk(p);
        check(fk_build(h_, t, oracle.size(), &fp, &out.g), "fk_build");
        if (verbosity) {
            const fk_stats &s = out.g.stats;
            std::cout << "construct time: " << s.t_total << "\n";
            std::cout << "Degree Statistics: Max = " << s.max_degree << ", Min = " << s.min_degree
                      << ", Avg = " << s.aod << std::endl;
        }
    }

    static void save_nsg(const char *path, const fk_graph &g) {
        if (!path || !*path) return;
        std::ofstream out(path, std::ios::binary);
        if (!out) throw io_error(path);
        out.write((const char *)&g.width, 4);
        out.write((const char *)&g.ep, 4);
        for (uint32_t i = 0; i < g.n; ++i) {
            const uint32_t k = (uint32_t)(g.offsets[i + 1] - g.offsets[i]);
            out.write((const char *)&k, 4);
            out.write((const char *)(g.ids + g.offsets[i]), (std::streamsize)k * 4);
        }
        if (!out) throw io_error(path);
    }

    static void save_result(const char *path, int pg_type, IndexParams const &p, const fk_stats &s) {
        if (!path || !*path) return;
        std::ofstream ofs(path, std::ios::out | std::ios::app);
        if (!ofs.is_open()) return;
        const unsigned nl = s.n_loops, nb = nl ? nl - 1 : 0;
        ofs << "PG_type,K,L,S,R,I,searchL,searchK,nsgR,massqS,tau,angle,step" << std::endl;
        ofs << pg_type << "," << p.K << "," << p.L << "," << p.S << "," << p.R << "," << p.iterations << ","
            << p.search_L << "," << p.search_K << "," << p.nsg_R << "," << p.massq_S << "," << p.tau << ","
            << p.angle << "," << p.step << "," << std::endl;
        auto row = [&](const char *name, auto *v, unsigned n) {
            ofs << name << ",";
            for (unsigned i = 0; i < n; ++i) ofs << v[i] << ",";
            ofs << std::endl;
        };
        row("kgraph time", s.kgraph_time, s.n_kgraph);
        row("recall", s.recall, s.n_recall);
        row("prune time", s.prune_time, nl);
        row("inter time", s.inter_time, nl);
        row("tree_grow time", s.tree_grow_time, pg_type == FK_PG_HNSW_LAYER ? 0u : nl);
        row("buildnsg time", s.buildnsg_time, nl);
        row("search time", s.search_time, nb);
        row("con time", s.con_time, nl);
        row("prune cc", s.prune_cc, nl);
        row("search cc", s.search_cc, nb);
        ofs << "AOD" << "," << s.aod << std::endl;
        ofs << "final_prune_time" << "," << -1.0 << std::endl;
    }

    void build_flat(fk_pg_type t, IndexOracle const &oracle, IndexParams const &p, const char *nsg_name,
                    const char *res_name) {
        Graph g;
        run(t, oracle, p, g);
        save_nsg(nsg_name, g.g);
        save_result(res_name, (int)t, p, g.g.stats);
    }

public:
    ~KGraphImpl() override {
        if (h_) fk_close(h_);
    }

    void load(char const *path) override {
        std::ifstream is(path, std::ios::binary);
        char magic[8];
        is.read(magic, 8);
        if (!is || memcmp(magic, "KNNGRAPH", 8) != 0) throw io_error(path);
    }

    // The reference's save() writes the KNNGRAPH header of an index that no
    // buildFast* fills (src/kgraph.cpp:463-493): N = 0.
    void save(char const *path, int) const override {
        std::ofstream os(path, std::ios::binary);
        if (!os) throw io_error(path);
        const uint32_t version = 2, cap = 0, N = 0;
        os.write("KNNGRAPH", 8);
        os.write((const char *)&version, 4);
        os.write((const char *)&cap, 4);
        os.write((const char *)&N, 4);
    }

    void buildFastNSG(IndexOracle const &oracle, IndexParams const &param, IndexInfo *, const char *nsg_name,
                      const char *res_name) override {
        build_flat(FK_PG_NSG, oracle, param, nsg_name, res_name);
    }
    void buildFastTauMNG(IndexOracle const &oracle, IndexParams const &param, IndexInfo *, const char *nsg_name,
                         const char *res_name) override {
        build_flat(FK_PG_TAUMNG, oracle, param, nsg_name, res_name);
    }
    void buildFastAlphaPG(IndexOracle const &oracle, IndexParams const &param, IndexInfo *, const char *nsg_name,
                          const char *res_name) override {
        build_flat(FK_PG_ALPHAPG, oracle, param, nsg_name, res_name);
    }

    void buildFastHNSW(IndexOracle const &oracle, IndexParams const &param, IndexInfo *, const char *,
                       const char *res_name, char *&, char **linkLists_, unsigned level, unsigned *ep_ptr,
                       size_t) override {
        Graph g;
        run(FK_PG_HNSW_LAYER, oracle, param, g);
        if (ep_ptr) *ep_ptr = g.g.ep;
        if (level > 0) {
            // SwapHNSWLayer: [u16 count][u16 0][nsg_R ids] per level block
            const size_t per = (size_t)param.nsg_R * 4 + 4;
            for (uint32_t i = 0; i < g.g.n; ++i) {
                unsigned *cll = (unsigned *)(linkLists_[i] + (level - 1) * per);
                const uint32_t k = (uint32_t)std::min<uint64_t>(g.g.offsets[i + 1] - g.g.offsets[i], param.nsg_R);
                *(unsigned short *)cll = (unsigned short)k;
                memcpy(cll + 1, g.g.ids + g.g.offsets[i], (size_t)k * 4);
            }
        } else {
            temp_graph_.assign(g.g.n, {});
            for (uint32_t i = 0; i < g.g.n; ++i)
                temp_graph_[i].assign(g.g.ids + g.g.offsets[i], g.g.ids + g.g.offsets[i + 1]);
        }
        save_result(res_name, FK_PG_HNSW_LAYER, param, g.g.stats);
    }

    void FatherSwapHNSWLayer0(IndexOracle const &oracle, IndexParams const &params,
                              char *&data_level0_memory_) override {
        const size_t links = (size_t)params.nsg_R * 4 + 4;
        const size_t per = links + sizeof(float) * oracle.dim() + sizeof(size_t);
        for (unsigned i = 0; i < oracle.size() && i < temp_graph_.size(); i++) {
            unsigned *cll = (unsigned *)(data_level0_memory_ + i * per);
            const size_t k = std::min<size_t>(temp_graph_[i].size(), params.nsg_R);
            *(unsigned short *)cll = (unsigned short)k;
            memcpy(cll + 1, temp_graph_[i].data(), k * 4);
        }
    }

    void prune(IndexOracle const &, unsigned) override {
        throw runtime_error("KGraph::prune is not part of the build path");
    }
    unsigned search(SearchOracle const &, SearchParams const &, unsigned *, float *, SearchInfo *) const override {
        throw runtime_error("KGraph::search is not part of the build path");
    }
    void get_nn(unsigned, unsigned *, float *, unsigned *, unsigned *) const override {
        throw runtime_error("KGraph::get_nn is not part of the build path");
    }
    void reverse(int) override { throw runtime_error("KGraph::reverse is not part of the build path"); }
};

KGraph *KGraph::create() { return new KGraphImpl; }
char const *KGraph::version() { return fk_version(); }

}  // namespace kgraph
