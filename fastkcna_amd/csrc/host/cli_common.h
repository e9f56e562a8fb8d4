// Command-line front end shared by fastnsg / fasttaumng / fastalphapg /
// fasthnsw.  Same flags, defaults, positional arguments and files as the
// reference tools (src/fastnsg.cpp:40-209, src/fasthnsw.cpp:46-358), parsed
// without Boost: --name value, --name=value, -X value, -Xvalue; positional
// DATA and OUTPUT.  GPU-only additions: --device N (HIP ordinal, default 0),
// --gram (FK_FLAG_GRAM) and --screen (FK_FLAG_SCREEN).
#pragma once
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/resource.h>

#include <chrono>
#include <cmath>
#include <fstream>
#include <functional>
#include <memory>
#include <iostream>
#include <random>
#include <string>
#include <type_traits>
#include <vector>

#include "fk_api.h"
#include "kgraph-data.h"
#include "kgraph.h"

namespace fkcli {

using namespace kgraph;

enum Mode { NSG = 1, HNSW = 2, TAUMNG = 3, ALPHAPG = 4 };

struct Option {
    std::vector<std::string> names;  // "--long" and/or "-s"
    bool flag;
    std::string help;
    std::function<void(const std::string &)> set;
};

template <typename T>
std::function<void(const std::string &)> setter(T &dst) {
    return [&dst](const std::string &v) {
        char *end = nullptr;
        if (std::is_floating_point<T>::value) dst = (T)strtod(v.c_str(), &end);
        else if (std::is_signed<T>::value) dst = (T)strtol(v.c_str(), &end, 10);
        else dst = (T)strtoul(v.c_str(), &end, 10);
        if (!end || *end) throw invalid_argument("bad value '" + v + "'");
    };
}

inline long peak_memory_mb() {
    struct rusage r;
    getrusage(RUSAGE_SELF, &r);
    return r.ru_maxrss / 1024;
}

template <typename T>
static void write_pod(std::ostream &out, const T &v) {
    out.write((const char *)&v, sizeof(T));
}

inline int run_cli(int argc, char **argv, Mode mode) {
    std::string data_path, output_path, res_path;
    KGraph::IndexParams params;
    unsigned D = 0, skip = 0, gap = 0, synthetic = 0;
    float noise = 0;
    bool help = false, version = false, raw = false, l2norm = false, have_dim = false;
    std::vector<Option> opts = {
        {{"--help", "-h"}, true, "produce help message.", [&](const std::string &) { help = true; }},
        {{"--version", "-v"}, true, "print version information.", [&](const std::string &) { version = true; }},
        {{"--data"}, false, "data name", [&](const std::string &v) { data_path = v; }},
        {{"--output"}, false, "output path", [&](const std::string &v) { output_path = v; }},
        {{"--result"}, false, "result path", [&](const std::string &v) { res_path = v; }},
        {{"-K"}, false, "number of nearest neighbor", setter(params.K)},
        {{"--nthreads", "-nt"}, false, "number of threads", setter(params.nthreads)},
        {{"--controls", "-C"}, false, "number of control points", setter(params.controls)},
        {{"--nsg_R"}, false, "", setter(params.nsg_R)},
        {{"--loop_i"}, false, "", setter(params.loop_i)},
        {{"--search_L"}, false, "", setter(params.search_L)},
        {{"--search_K"}, false, "", setter(params.search_K)},
        {{"--massq_S"}, false, "", setter(params.massq_S)},
        {{"--angle"}, false, "", setter(params.angle)},
        {{"--step"}, false, "", setter(params.step)},
        {{"--tau"}, false, "", setter(params.tau)},
        {{"--iterations", "-I"}, false, "", setter(params.iterations)},
        {{"-S"}, false, "", setter(params.S)},
        {{"-R"}, false, "", setter(params.R)},
        {{"-L"}, false, "", setter(params.L)},
        {{"--delta"}, false, "", setter(params.delta)},
        {{"--recall"}, false, "", setter(params.recall)},
        {{"--prune"}, false, "", setter(params.prune)},
        {{"--reverse"}, false, "", setter(params.reverse)},
        {{"--noise"}, false, "noise", setter(noise)},
        {{"--seed"}, false, "", setter(params.seed)},
        {{"--dim", "-D"}, false, "dimension, see format",
         [&](const std::string &v) { setter(D)(v); have_dim = true; }},
        {{"--skip"}, false, "see format", setter(skip)},
        {{"--gap"}, false, "see format", setter(gap)},
        {{"--raw"}, true, "read raw binary file, need to specify D.", [&](const std::string &) { raw = true; }},
        {{"--synthetic"}, false, "generate synthetic data, for performance evaluation only", setter(synthetic)},
        {{"--l2norm"}, true, "l2-normalize data", [&](const std::string &) { l2norm = true; }},
        {{"--device"}, false, "GPU ordinal (HIP)", setter(params.device)},
        {{"--gram"}, true, "local join in Gram form on the MFMA units (same graph)",
         [&](const std::string &) { params.gpu_flags |= FK_FLAG_GRAM; }},
        {{"--screen"}, true, "learned distance screen in the local join",
         [&](const std::string &) { params.gpu_flags |= FK_FLAG_SCREEN; }},
    };
    auto find = [&](const std::string &name) -> Option * {
        for (auto &o : opts)
            for (auto &n : o.names)
                if (n == name) return &o;
        return nullptr;
    };
    std::vector<std::string> positional;
    try {
        for (int i = 1; i < argc; ++i) {
            std::string a = argv[i];
            if (a.size() > 1 && a[0] == '-' && !(a.size() > 1 && (isdigit((unsigned char)a[1]) || a[1] == '.'))) {
                std::string name = a, val;
                bool has_val = false;
                const size_t eq = a.find('=');
                if (a.compare(0, 2, "--") == 0 && eq != std::string::npos) {
                    name = a.substr(0, eq);
                    val = a.substr(eq + 1);
                    has_val = true;
                }
                Option *o = find(name);
                if (!o && a.compare(0, 2, "--") != 0 && a.size() > 2) {
                    // -K500 style
                    o = find(a.substr(0, 2));
                    if (o && !o->flag) { val = a.substr(2); has_val = true; }
                    else o = nullptr;
                }
                if (!o) throw invalid_argument("unrecognised option '" + a + "'");
                if (o->flag) {
                    o->set("");
                } else {
                    if (!has_val) {
                        if (i + 1 >= argc) throw invalid_argument("missing value for '" + a + "'");
                        val = argv[++i];
                    }
                    o->set(val);
                }
            } else {
                positional.push_back(a);
            }
        }
        if (positional.size() > 0 && data_path.empty()) data_path = positional[0];
        if (positional.size() > 1 && output_path.empty()) output_path = positional[1];
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    if (version) {
        std::cout << "KGraph version " << KGraph::version() << std::endl;
        return 0;
    }
    const bool lshkit = !raw;
    if (help || (synthetic && (!have_dim || !data_path.empty())) ||
        (!synthetic && (data_path.empty() || (!have_dim && !lshkit)))) {
        std::cout << "Usage: index [OTHER OPTIONS]... INPUT [OUTPUT]" << std::endl;
        for (auto &o : opts) {
            std::string n;
            for (auto &x : o.names) n += (n.empty() ? "" : ", ") + x;
            std::cout << "  " << n << (o.flag ? "" : " arg") << "\t" << o.help << std::endl;
        }
        return 0;
    }
    if (params.S == 0) params.S = params.K;
    try {
        if (lshkit && synthetic == 0) {
            unsigned header[3];
            std::ifstream is(data_path.c_str(), std::ios::binary);
            is.read((char *)header, sizeof header);
            if (!is || header[0] != sizeof(float)) throw io_error(data_path);
            D = header[2];
            skip = 3 * sizeof(unsigned);
            gap = 0;
        }
        Matrix<float> data;
        std::vector<int> labels, points_num;
        double get_level_time = 0;
        if (synthetic) {
            data.resize(synthetic, D);
            std::default_random_engine rng(params.seed);
            std::uniform_real_distribution<double> distribution(-1.0, 1.0);
            data.zero();
            for (unsigned i = 0; i < synthetic; ++i)
                for (unsigned j = 0; j < D; ++j) data[i][j] = (float)distribution(rng);
        } else if (mode == HNSW) {
            get_level_time = data.hnsw_dataload(data_path, D, skip, gap, labels, points_num, params.nsg_R, params.seed);
        } else {
            data.load(data_path, D, skip, gap);
        }
        if (mode == HNSW && synthetic) {
            // synthetic HNSW data keeps file order: levels over the generated rows
            std::vector<int> f2d;
            data.getlevel(data.size(), labels, f2d, points_num, params.nsg_R, params.seed);
            Matrix<float> re(data.size(), data.dim());
            re.zero();
            for (unsigned i = 0; i < data.size(); ++i) memcpy(re[i], data[labels[i]], sizeof(float) * data.dim());
            for (unsigned i = 0; i < data.size(); ++i) memcpy(data[i], re[i], sizeof(float) * data.dim());
        }
        if (noise != 0) {
            // Gaussian noise of noise x sigma(data); std::mt19937_64 here
            // (the reference draws from Boost's ranlux64_base_01)
            double sum = 0, sum2 = 0;
            for (unsigned i = 0; i < data.size(); ++i)
                for (unsigned j = 0; j < data.dim(); ++j) {
                    const double v = data[i][j];
                    sum += v;
                    sum2 += v * v;
                }
            const double total = double(data.size()) * data.dim();
            const double dev = std::sqrt(sum2 / total - (sum / total) * (sum / total));
            std::mt19937_64 rng(params.seed);
            std::normal_distribution<double> g(0, noise * dev);
            for (unsigned i = 0; i < data.size(); ++i)
                for (unsigned j = 0; j < data.dim(); ++j) data[i][j] += (float)g(rng);
        }
        if (l2norm) data.normalize2();
        if (verbosity) printf("Size1: %d\n", data.size());
        MatrixOracle<float, metric::l2sqr> oracle(data);
        KGraph::IndexInfo info;
        KGraph *kgraph = KGraph::create();
        std::unique_ptr<KGraph> guard(kgraph);
        if (mode != HNSW) {
            if (mode == NSG) kgraph->buildFastNSG(oracle, params, &info, output_path.c_str(), res_path.c_str());
            else if (mode == TAUMNG) kgraph->buildFastTauMNG(oracle, params, &info, output_path.c_str(), res_path.c_str());
            else kgraph->buildFastAlphaPG(oracle, params, &info, output_path.c_str(), res_path.c_str());
            if (verbosity) std::cout << "peakmemory: " << peak_memory_mb() << std::endl;
            return 0;
        }
        // ---- fasthnsw (src/fasthnsw.cpp:202-342) ----
        {
            std::ofstream ss(res_path, std::ios::out | std::ios::app);
            if (ss.is_open()) {
                ss << "start*************************************************************************************************************" << std::endl;
                ss << "get_level_time" << "," << get_level_time << std::endl;
            }
        }
        auto s = std::chrono::high_resolution_clock::now();
        std::vector<int> levels;
        for (int i = (int)points_num.size() - 1; i >= 0; i--)
            for (int j = 0; j < points_num[i]; j++) levels.push_back(i);
        const size_t N = data.size();
        const size_t M_ = params.nsg_R, maxM_ = params.nsg_R, maxM0_ = 2 * params.nsg_R;
        const size_t size_links_per_element_ = maxM_ * sizeof(unsigned) + sizeof(unsigned);
        const size_t size_links_level0_ = maxM0_ * sizeof(unsigned) + sizeof(unsigned);
        const size_t size_data_per_element_ = size_links_level0_ + sizeof(float) * data.dim() + sizeof(size_t);
        const size_t offsetData_ = size_links_level0_;
        const size_t label_offset_ = size_links_level0_ + sizeof(float) * data.dim();
        const size_t offsetLevel0_ = 0;
        char *data_level0_memory_ = nullptr;
        std::vector<char *> linkLists_(N, nullptr);
        const int maxlevel_ = levels[0];
        const double mult_ = 1 / std::log(1.0 * M_);
        const size_t ef_construction_ = params.search_K;
        unsigned int enterpoint_node_ = 0;
        for (size_t c = 0; c < N && levels[c] > 0; ++c)
            linkLists_[c] = (char *)calloc(size_links_per_element_ * levels[c] + 1, 1);
        unsigned temp_ep = 0;
        int size = 0;
        for (int level = (int)points_num.size() - 1; level >= 0; level--) {
            size += points_num[level];
            oracle.set_size(size);
            if (size == 1) {
                // a layer of one node: no links; it is the entry point
                if (level > 0) *(unsigned short *)(linkLists_[0] + (level - 1) * size_links_per_element_) = 0;
                if (level == (int)points_num.size() - 1) enterpoint_node_ = 0;
                if (level > 0) continue;
            }
            if (level > 0) {
                kgraph->buildFastHNSW(oracle, params, &info, output_path.c_str(), res_path.c_str(), data_level0_memory_,
                                      linkLists_.data(), level, &temp_ep);
                if (level == (int)points_num.size() - 1) enterpoint_node_ = temp_ep;
            } else {
                params.nsg_R *= 2;
                data_level0_memory_ = (char *)calloc(oracle.size() * size_data_per_element_, 1);
                if (size > 1) {
                    kgraph->buildFastHNSW(oracle, params, &info, output_path.c_str(), res_path.c_str(),
                                          data_level0_memory_, linkLists_.data(), level, &temp_ep,
                                          size_data_per_element_);
                    kgraph->FatherSwapHNSWLayer0(oracle, params, data_level0_memory_);
                }
                if (level == (int)points_num.size() - 1) enterpoint_node_ = temp_ep;
            }
        }
        for (size_t internal_id = 0; internal_id < labels.size(); internal_id++) {
            const size_t label = (size_t)labels[internal_id];
            memcpy(data_level0_memory_ + internal_id * size_data_per_element_ + offsetData_, data.getdata(internal_id),
                   sizeof(float) * data.dim());
            memcpy(data_level0_memory_ + internal_id * size_data_per_element_ + label_offset_, &label, sizeof(label));
        }
        std::chrono::duration<double> diff = std::chrono::high_resolution_clock::now() - s;
        const double time = diff.count() + get_level_time;
        if (verbosity) std::cout << "fasthnsw indexing time: " << time << "\n";
        {
            std::ofstream sr(res_path, std::ios::out | std::ios::app);
            if (sr.is_open()) {
                sr << "fasthnsw indexing time" << "," << time << std::endl;
                sr << "end***************************************************************************************************************" << std::endl << std::endl;
            }
        }
        std::ofstream output(output_path, std::ios::binary);
        if (!output) throw io_error(output_path);
        write_pod(output, offsetLevel0_);
        write_pod(output, N);
        write_pod(output, N);
        write_pod(output, size_data_per_element_);
        write_pod(output, label_offset_);
        write_pod(output, offsetData_);
        write_pod(output, maxlevel_);
        write_pod(output, enterpoint_node_);
        write_pod(output, maxM_);
        write_pod(output, maxM0_);
        write_pod(output, M_);
        write_pod(output, mult_);
        write_pod(output, ef_construction_);
        output.write(data_level0_memory_, N * size_data_per_element_);
        for (size_t i = 0; i < N; i++) {
            const unsigned int linkListSize = levels[i] > 0 ? (unsigned)(size_links_per_element_ * levels[i]) : 0;
            write_pod(output, linkListSize);
            if (linkListSize) output.write(linkLists_[i], linkListSize);
        }
        output.close();
        free(data_level0_memory_);
        for (auto p : linkLists_) free(p);
        if (verbosity) std::cout << "fasthnsw peakmemory: " << peak_memory_mb() << std::endl;
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}

}  // namespace fkcli
