// kgraph-data.h -- row container, metrics and the matrix oracle.
//
// Drop-in for the reference's src/kgraph-data.h: Matrix<T> (rows padded to
// 32 bytes, lshkit / raw loaders, HNSW level assignment), MatrixProxy,
// MatrixOracle<DATA, DIST> and metric::l2sqr keep their names and behaviour.
// Two additions: MatrixOracle exposes its row storage (rows(), stride(),
// proxy_size()) so the GPU build can upload it, and metric::l2sqr::apply<float>
// evaluates in the reference's float_l2sqr_avx order (8 fused lanes, tail
// blocks first) so host-side distances equal the GPU's and the reference's.
#ifndef FK_KGRAPH_DATA_H
#define FK_KGRAPH_DATA_H

#include <malloc.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <cmath>
#include <fstream>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "kgraph.h"

#define KGRAPH_MATRIX_ALIGN 32

namespace kgraph {

using std::vector;

/// Squared L2 in the reference's AVX order (src/metric.cpp:45-78).
inline float float_l2sqr_ordered(float const *t1, float const *t2, unsigned dim) {
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const unsigned D = (dim + 7) & ~7u, dr = D % 32, dd = D - dr;
    auto step = [&](unsigned off) {
        for (unsigned k = 0; k < 8; ++k) {
            const float a = off + k < dim ? t1[off + k] : 0.f;
            const float b = off + k < dim ? t2[off + k] : 0.f;
            const float t = a - b;
            acc[k] = std::fma(t, t, acc[k]);
        }
    };
    if (dr == 24) { step(dd + 16); step(dd + 8); step(dd); }
    else if (dr == 16) { step(dd + 8); step(dd); }
    else if (dr == 8) { step(dd); }
    for (unsigned i = 0; i < dd; i += 8) step(i);
    float r = acc[0] + acc[1];
    for (int k = 2; k < 8; ++k) r += acc[k];
    return r;
}

namespace metric {
struct l2sqr {
    template <typename T>
    static float apply(T const *t1, T const *t2, unsigned dim) {
        float r = 0;
        for (unsigned i = 0; i < dim; ++i) {
            float v = float(t1[i]) - float(t2[i]);
            v *= v;
            r += v;
        }
        return r;
    }
    template <typename T>
    static float dot(T const *t1, T const *t2, unsigned dim) {
        float r = 0;
        for (unsigned i = 0; i < dim; ++i) r += float(t1[i]) * float(t2[i]);
        return r;
    }
    template <typename T>
    static float norm2(T const *t1, unsigned dim) {
        float r = 0;
        for (unsigned i = 0; i < dim; ++i) {
            float v = float(t1[i]);
            v *= v;
            r += v;
        }
        return r;
    }
};
template <>
inline float l2sqr::apply<float>(float const *t1, float const *t2, unsigned dim) {
    return float_l2sqr_ordered(t1, t2, dim);
}

struct l2 {
    template <typename T>
    static float apply(T const *t1, T const *t2, unsigned dim) {
        return std::sqrt(l2sqr::apply<T>(t1, t2, dim));
    }
};
}  // namespace metric

/// Row matrix with rows padded to A bytes (src/kgraph-data.h:97-288).
template <typename T, unsigned A = KGRAPH_MATRIX_ALIGN>
class Matrix {
    unsigned col;
    unsigned row;
    size_t stride;
    char *data;

    void reset(unsigned r, unsigned c) {
        row = r;
        col = c;
        stride = (sizeof(T) * c + A - 1) / A * A;
        if (data) free(data);
        data = (char *)memalign(A, (size_t)row * stride + A);
        if (!data) throw runtime_error("memalign");
    }

public:
    Matrix() : col(0), row(0), stride(0), data(0) {}
    Matrix(unsigned r, unsigned c) : data(0) { reset(r, c); }
    Matrix(const Matrix &) = delete;
    Matrix &operator=(const Matrix &) = delete;
    ~Matrix() {
        if (data) free(data);
    }
    unsigned size() const { return row; }
    unsigned dim() const { return col; }
    size_t step() const { return stride; }
    void resize(unsigned r, unsigned c) { reset(r, c); }
    char *getdata(unsigned i) { return data + stride * i; }
    T const *operator[](unsigned i) const { return reinterpret_cast<T const *>(&data[stride * i]); }
    T *operator[](unsigned i) { return reinterpret_cast<T *>(&data[stride * i]); }
    void zero() { memset(data, 0, (size_t)row * stride); }

    /// Scales every row to unit L2 norm; the norm is the float sum of
    /// squares, taken in double for the square root and the division.
    void normalize2() {
#pragma omp parallel for
        for (unsigned i = 0; i < row; ++i) {
            T *r = operator[](i);
            const double len = std::sqrt((double)metric::l2sqr::norm2(r, col));
            for (T *v = r; v != r + col; ++v) *v = (T)(*v / len);
        }
    }

private:
    /// A headerless file of fixed-size records: `skip` bytes of preamble,
    /// then records of `dim` values each followed by `gap` bytes.
    struct RecordFile {
        std::ifstream in;
        size_t skip, gap, value_bytes;
        unsigned count = 0;
        RecordFile(const std::string &path, unsigned dim, size_t skip_, size_t gap_)
            : in(path.c_str(), std::ios::binary), skip(skip_), gap(gap_), value_bytes(sizeof(T) * (size_t)dim) {
            if (!in) throw io_error(path);
            in.seekg(0, std::ios::end);
            const size_t total = (size_t)in.tellg();
            count = total > skip ? (unsigned)((total - skip) / (value_bytes + gap)) : 0u;
            in.seekg((std::streamoff)skip, std::ios::beg);
        }
        // Record k of the file into dst; records are read in file order.
        void next(char *dst) {
            in.read(dst, (std::streamsize)value_bytes);
            if (gap) in.seekg((std::streamoff)gap, std::ios::cur);
        }
    };

    /// Fills the matrix from a record file; record k lands in row slot(k).
    template <typename SLOT>
    void fill_from(RecordFile &f, const std::string &path, SLOT slot) {
        for (unsigned k = 0; k < f.count; ++k) f.next(getdata(slot(k)));
        if (!f.in) throw io_error(path);
    }

public:
    void load(const std::string &path, unsigned dim, unsigned skip = 0, unsigned gap = 0) {
        RecordFile f(path, dim, skip, gap);
        reset(f.count, dim);
        zero();
        fill_from(f, path, [](unsigned k) { return k; });
    }

    // HNSW level of one point: (int)(-ln U(0,1) * reverse_size)
    int getRandomLevel(double reverse_size, std::default_random_engine &level_generator_) {
        std::uniform_real_distribution<double> distribution(0.0, 1.0);
        double r = -std::log(distribution(level_generator_)) * reverse_size;
        return (int)r;
    }

    // Levels in point order from default_random_engine(random_seed), then an
    // order of rows by level, highest first (src/kgraph-data.h:200-222).
    void getlevel(unsigned N, vector<int> &data2file, vector<int> &file2data, vector<int> &points_num, unsigned M_,
                  unsigned random_seed) {
        std::default_random_engine level_generator_;
        const double mult = 1 / std::log(1.0 * M_);
        level_generator_.seed(random_seed);
        data2file.clear();
        data2file.reserve(N);
        vector<int> levels(N);
        int max_level = 0;
        for (unsigned i = 0; i < N; i++) {
            levels[i] = getRandomLevel(mult, level_generator_);
            if (levels[i] > max_level) max_level = levels[i];
        }
        points_num.assign(max_level + 1, 0);
        vector<vector<int>> by_level(max_level + 1);
        for (unsigned j = 0; j < N; j++) by_level[levels[j]].push_back((int)j);
        for (int i = max_level; i >= 0; i--) {
            for (int j : by_level[i]) data2file.push_back(j);
            points_num[i] = (int)by_level[i].size();
        }
        file2data.assign(N, 0);
        for (unsigned i = 0; i < N; i++) file2data[data2file[i]] = (int)i;
    }

    /// Loads the records in level order (highest level first, the reference's
    /// hnsw_dataload): record k goes to row file2data[k]; returns the seconds
    /// the level draw took.
    double hnsw_dataload(const std::string &path, unsigned dim, unsigned skip, unsigned gap, vector<int> &data2file,
                         vector<int> &points_num, unsigned M_ = 50, unsigned random_seed = 2024) {
        RecordFile f(path, dim, skip, gap);
        vector<int> file2data;
        const auto t0 = std::chrono::steady_clock::now();
        getlevel(f.count, data2file, file2data, points_num, M_, random_seed);
        const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (verbosity) std::cout << "get level time: " << secs << "\n";
        reset(f.count, dim);
        zero();
        fill_from(f, path, [&](unsigned k) { return (unsigned)file2data[k]; });
        return secs;
    }

    /// lshkit file: a header {value size, rows, dim} of three u32, then the
    /// rows back to back.
    void load_lshkit(std::string const &path) {
        uint32_t header[3] = {0, 0, 0};
        std::ifstream in(path.c_str(), std::ios::binary);
        if (!in.read(reinterpret_cast<char *>(header), sizeof header) || header[0] != sizeof(T)) throw io_error(path);
        in.close();
        load(path, header[2], (unsigned)sizeof header, 0);
    }

    void save_lshkit(std::string const &path) {
        std::ofstream os(path.c_str(), std::ios::binary);
        unsigned header[3] = {(unsigned)sizeof(T), row, col};
        os.write((const char *)header, sizeof(header));
        for (unsigned i = 0; i < row; ++i) os.write(&data[stride * i], sizeof(T) * col);
    }
};

/// Non-owning view of a row matrix (src/kgraph-data.h:291-347).
template <typename DATA_TYPE, unsigned A = KGRAPH_MATRIX_ALIGN>
class MatrixProxy {
    unsigned rows_;
    unsigned cols_;
    size_t stride_;
    uint8_t const *data_;

public:
    MatrixProxy(Matrix<DATA_TYPE> const &m)
        : rows_(m.size()), cols_(m.dim()), stride_(m.step()), data_(reinterpret_cast<uint8_t const *>(m[0])) {}
    MatrixProxy(DATA_TYPE const *data, unsigned rows, unsigned cols, size_t stride_bytes)
        : rows_(rows), cols_(cols), stride_(stride_bytes), data_(reinterpret_cast<uint8_t const *>(data)) {}
    unsigned size() const { return rows_; }
    unsigned dim() const { return cols_; }
    size_t step() const { return stride_; }
    DATA_TYPE const *operator[](unsigned i) const {
        return reinterpret_cast<DATA_TYPE const *>(data_ + stride_ * i);
    }
};

/// Oracle over a Matrix / MatrixProxy (src/kgraph-data.h:349-433).
template <typename DATA_TYPE, typename DIST_TYPE>
class MatrixOracle : public kgraph::IndexOracle {
    MatrixProxy<DATA_TYPE> proxy;

public:
    float *selfdot = nullptr;
    unsigned oracle_size;
    class SearchOracle : public kgraph::SearchOracle {
        MatrixProxy<DATA_TYPE> proxy;
        DATA_TYPE const *query;

    public:
        SearchOracle(MatrixProxy<DATA_TYPE> const &p, DATA_TYPE const *q) : proxy(p), query(q) {}
        virtual unsigned size() const { return proxy.size(); }
        virtual float operator()(unsigned i) const { return DIST_TYPE::apply(proxy[i], query, proxy.dim()); }
    };
    template <typename MATRIX_TYPE>
    MatrixOracle(MATRIX_TYPE const &m) : proxy(m) {
        oracle_size = proxy.size();
    }
    ~MatrixOracle() { delete[] selfdot; }
    virtual unsigned size() const { return oracle_size; }
    virtual unsigned data_size() const { return proxy.size(); }
    virtual void set_size(unsigned mysize) { oracle_size = mysize > proxy.size() ? proxy.size() : mysize; }
    virtual unsigned dim() const { return proxy.dim(); }
    virtual float operator()(unsigned i, unsigned j) const { return DIST_TYPE::apply(proxy[i], proxy[j], proxy.dim()); }
    virtual float operator()(unsigned i, unsigned j, unsigned offset, unsigned dim) const {
        return DIST_TYPE::apply(proxy[i] + offset, proxy[j] + offset, dim);
    }
    virtual float operator()(const float *query, unsigned i) const {
        return DIST_TYPE::apply(query, proxy[i], proxy.dim());
    }
    virtual float computeDot(unsigned i, unsigned j) const { return DIST_TYPE::dot(proxy[i], proxy[j], proxy.dim()); }
    virtual void computeSelfDotPtr() {
        delete[] selfdot;
        selfdot = new float[proxy.size()];
        for (unsigned i = 0; i < proxy.size(); i++) selfdot[i] = DIST_TYPE::dot(proxy[i], proxy[i], proxy.dim());
    }
    /// Mean of all rows (data_size(), not the size() prefix), float sums in
    /// row order; the caller deletes[] the result.
    virtual float *Calcenter() const {
        const unsigned d = proxy.dim(), n = proxy.size();
        float *mean = new float[d]();
        for (unsigned r = 0; r < n; ++r) {  // running float sums, one row at a time
            DATA_TYPE const *x = proxy[r];
            for (float *m = mean; m != mean + d; ++m, ++x) *m += (float)*x;
        }
        const float count = (float)n;
        for (float *m = mean; m != mean + d; ++m) *m /= count;
        return mean;
    }
    SearchOracle query(DATA_TYPE const *q) const { return SearchOracle(proxy, q); }
    // Row storage for the GPU upload (not in the reference).
    DATA_TYPE const *rows() const { return proxy[0]; }
    size_t stride() const { return proxy.step(); }
    unsigned proxy_size() const { return proxy.size(); }
};

}  // namespace kgraph

#endif  // FK_KGRAPH_DATA_H
