// kgraph-data.h -- dense matrices, metrics and matrix oracles (the reference's
// kgraph-data.h:1-369), for use with the GPU index of libkgraph_amd.so.
//
// Same class templates and constructors as the reference: metric::l2sqr /
// metric::l2, Matrix<T>, MatrixProxy<T>, MatrixOracle<T, DIST>, AverageRecall.
// MatrixOracle<float, metric::l2sqr> additionally reports its rows through
// kgraph::DenseOracleInfo, which is what the GPU build and search read.
//
// Host-side distances use the canonical FP32 order of the GPU kernels (16
// interleaved fmaf partials, the AVX-512 reduce tree of xsimd
// (arch/xsimd_avx512f.hpp:1393-1403), then a sequential fmaf tail), so a
// distance computed here equals the one the GPU stores, bit for bit.
#ifndef WDONG_KGRAPH_DATA
#define WDONG_KGRAPH_DATA

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "kgraph.h"

namespace kgraph {

static constexpr size_t KGRAPH_MATRIX_ALIGN = 64;

namespace metric {
namespace detail {
template <typename T>
inline T reduce16(T const *p) {
    T u[4];
    for (int e = 0; e < 4; ++e) u[e] = (p[e] + p[e + 4]) + (p[e + 8] + p[e + 12]);
    return (u[0] + u[1]) + (u[2] + u[3]);
}
}  // namespace detail

/// Squared Euclidean distance (kgraph-data.h:38-104).
struct l2sqr {
    static constexpr int dense_id = 0;
    template <typename T>
    static T apply(T const *t1, T const *t2, unsigned size) {
        T p[16] = {0};
        unsigned const n16 = size & ~15u;
        for (unsigned i = 0; i < n16; i += 16)
            for (int t = 0; t < 16; ++t) {
                T a = t1[i + t] - t2[i + t];
                p[t] = std::fma(a, a, p[t]);
            }
        T acc = detail::reduce16(p);
        for (unsigned i = n16; i < size; ++i) {
            T a = t1[i] - t2[i];
            acc = std::fma(a, a, acc);
        }
        return acc;
    }
    template <typename T>
    static T dot(T const *t1, T const *t2, unsigned size) {
        T p[16] = {0};
        unsigned const n16 = size & ~15u;
        for (unsigned i = 0; i < n16; i += 16)
            for (int t = 0; t < 16; ++t) p[t] = std::fma(t1[i + t], t2[i + t], p[t]);
        T acc = detail::reduce16(p);
        for (unsigned i = n16; i < size; ++i) acc = std::fma(t1[i], t2[i], acc);
        return acc;
    }
    template <typename T>
    static float norm2(T const *t1, unsigned size) {
        return dot(t1, t1, size);
    }
};

/// Euclidean distance (kgraph-data.h:106-111); no GPU form.
struct l2 {
    static constexpr int dense_id = -1;
    template <typename T>
    static T apply(T const *t1, T const *t2, unsigned dim) {
        return std::sqrt(l2sqr::apply<T>(t1, t2, dim));
    }
};
}  // namespace metric

/// Row-major matrix with rows padded to A bytes (kgraph-data.h:114-219).
template <typename T, unsigned A = KGRAPH_MATRIX_ALIGN>
class Matrix {
    unsigned col;
    unsigned row;
    size_t stride;
    char *data;

protected:
    void reset(unsigned r, unsigned c) {
        row = r;
        col = c;
        stride = (sizeof(T) * c + A - 1) / A * A;
        if (data) free(data);
        data = static_cast<char *>(aligned_alloc(A, std::max<size_t>(row * stride, A)));
        if (!data) throw runtime_error("memalign");
    }

public:
    Matrix() : col(0), row(0), stride(0), data(0) {}
    Matrix(unsigned r, unsigned c) : data(0) { reset(r, c); }
    Matrix(Matrix const &) = delete;
    Matrix &operator=(Matrix const &) = delete;
    ~Matrix() {
        if (data) free(data);
    }
    unsigned size() const { return row; }
    unsigned dim() const { return col; }
    size_t step() const { return stride; }
    void resize(unsigned r, unsigned c) { reset(r, c); }
    T const *operator[](unsigned i) const { return reinterpret_cast<T const *>(&data[stride * i]); }
    T *operator[](unsigned i) { return reinterpret_cast<T *>(&data[stride * i]); }
    void zero() { memset(data, 0, row * stride); }
    void normalize2() {
        for (unsigned i = 0; i < row; ++i) {
            T *p = operator[](i);
            double sum = std::sqrt(double(metric::l2sqr::norm2(p, col)));
            for (unsigned j = 0; j < col; ++j) p[j] /= sum;
        }
    }
    /// Rows of `dim` values each, after a `skip`-byte header, with `gap`
    /// bytes after every row (the raw layout of kgraph-data.h:175-191): as
    /// many whole rows as the file holds.
    void load(std::string const &path, unsigned dim, unsigned skip = 0, unsigned gap = 0) {
        std::ifstream in(path.c_str(), std::ios::binary | std::ios::ate);
        if (!in) throw io_error(path);
        std::streamoff const bytes = in.tellg();
        size_t const rowbytes = sizeof(T) * size_t(dim);
        size_t const period = rowbytes + gap;
        size_t const body = bytes > std::streamoff(skip) ? size_t(bytes) - skip : 0;
        size_t const n = period ? body / period : 0;
        reset(unsigned(n), dim);
        zero();
        in.seekg(skip);
        // one read of the whole body when rows are packed back to back in
        // both places, else a row at a time
        if (gap == 0 && stride == rowbytes) {
            in.read(data, std::streamsize(n * rowbytes));
        } else {
            std::vector<char> buf(period);
            for (size_t r = 0; r < n && in; ++r) {
                in.read(buf.data(), std::streamsize(r + 1 < n ? period : rowbytes));
                std::memcpy(data + stride * r, buf.data(), rowbytes);
            }
        }
        if (!in) throw io_error(path);
    }
    /// lshkit files: a header of three 32-bit words {bytes per value, rows,
    /// columns}, then the rows packed (kgraph-data.h:193-218).
    void load_lshkit(std::string const &path) {
        uint32_t head[3] = {0, 0, 0};
        {
            std::ifstream in(path.c_str(), std::ios::binary);
            if (!in.read(reinterpret_cast<char *>(head), sizeof(head)) || head[0] != sizeof(T)) throw io_error(path);
        }
        load(path, head[2], unsigned(sizeof(head)), 0);
    }
    void save_lshkit(std::string const &path) {
        std::ofstream out(path.c_str(), std::ios::binary);
        uint32_t const head[3] = {uint32_t(sizeof(T)), row, col};
        out.write(reinterpret_cast<char const *>(head), sizeof(head));
        size_t const rowbytes = sizeof(T) * size_t(col);
        if (stride == rowbytes) {
            out.write(data, std::streamsize(rowbytes * row));
        } else {
            for (unsigned r = 0; r < row; ++r) out.write(data + stride * r, std::streamsize(rowbytes));
        }
        if (!out) throw io_error(path);
    }
};

/// Non-owning view of a row-major matrix (kgraph-data.h:221-296).
template <typename DATA_TYPE, unsigned A = KGRAPH_MATRIX_ALIGN>
class MatrixProxy {
    unsigned rows;
    unsigned cols;
    size_t stride;  // bytes
    uint8_t const *data;

public:
    MatrixProxy(Matrix<DATA_TYPE> const &m)
        : rows(m.size()), cols(m.dim()), stride(m.step()), data(reinterpret_cast<uint8_t const *>(m[0])) {}
    /// Rows of `c` elements `stride_bytes` apart.
    MatrixProxy(DATA_TYPE const *p, unsigned r, unsigned c, size_t stride_bytes)
        : rows(r), cols(c), stride(stride_bytes), data(reinterpret_cast<uint8_t const *>(p)) {
        if (stride < cols * sizeof(DATA_TYPE)) throw invalid_argument("bad stride");
    }
    unsigned size() const { return rows; }
    unsigned dim() const { return cols; }
    size_t step() const { return stride; }
    DATA_TYPE const *operator[](unsigned i) const { return reinterpret_cast<DATA_TYPE const *>(data + stride * i); }
};

/// Oracle over a Matrix / MatrixProxy (kgraph-data.h:298-331).  With float
/// rows and metric::l2sqr it also exposes the rows to the GPU index.
template <typename DATA_TYPE, typename DIST_TYPE>
class MatrixOracle : public IndexOracle, public DenseOracleInfo {
    MatrixProxy<DATA_TYPE> proxy;

public:
    class SearchOracle : public kgraph::SearchOracle, public DenseQueryInfo {
        MatrixOracle const *parent;
        DATA_TYPE const *query;

    public:
        SearchOracle(MatrixOracle const *p, DATA_TYPE const *q) : parent(p), query(q) {}
        virtual unsigned size() const { return parent->proxy.size(); }
        virtual float operator()(unsigned i) const {
            return DIST_TYPE::apply(parent->proxy[i], query, parent->proxy.dim());
        }
        virtual DenseOracleInfo const &dense_data() const { return *parent; }
        virtual float const *dense_query() const { return reinterpret_cast<float const *>(query); }
    };
    template <typename MATRIX_TYPE>
    MatrixOracle(MATRIX_TYPE const &m) : proxy(m) {}
    virtual unsigned size() const { return proxy.size(); }
    virtual float operator()(unsigned i, unsigned j) const { return DIST_TYPE::apply(proxy[i], proxy[j], proxy.dim()); }
    SearchOracle query(DATA_TYPE const *q) const { return SearchOracle(this, q); }

    virtual float const *dense_rows() const { return reinterpret_cast<float const *>(proxy[0]); }
    virtual unsigned dense_dim() const { return proxy.dim(); }
    virtual size_t dense_stride() const { return proxy.step() / sizeof(DATA_TYPE); }
    virtual int dense_metric() const {
        return (sizeof(DATA_TYPE) == sizeof(float) && proxy.step() % sizeof(float) == 0) ? DIST_TYPE::dense_id : -1;
    }
};

/// Mean recall@K of result rows against ground-truth rows (kgraph-data.h:
/// 333-364): each row holds distances sorted ascending; a result distance
/// counts as found when it equals a ground-truth distance not yet matched, so
/// ties never count as misses.  A result distance below every unmatched
/// ground-truth one means the two disagree on distances: runtime_error
/// "distance is unstable", as the reference throws.
namespace detail {
inline unsigned matched_prefix(float const *truth, float const *found, unsigned K) {
    unsigned hits = 0;
    float const *t = truth, *te = truth + K, *f = found, *fe = found + K;
    while (t != te && f != fe) {
        if (*f < *t) throw runtime_error("distance is unstable");
        if (*f == *t) {
            ++hits;
            ++f;
        }
        ++t;
    }
    return hits;
}
}  // namespace detail

inline float AverageRecall(Matrix<float> const &gs, Matrix<float> const &result, unsigned K = 0) {
    if (K == 0) K = result.dim();
    if (!(gs.dim() >= K)) throw invalid_argument("gs.dim() >= K");
    if (!(result.dim() >= K)) throw invalid_argument("result.dim() >= K");
    if (!(gs.size() >= result.size())) throw invalid_argument("gs.size() > result.size()");
    double total = 0;
    for (unsigned r = 0; r < result.size(); ++r) total += detail::matched_prefix(gs[r], result[r], K);
    return float(total / K / result.size());
}

}  // namespace kgraph

#endif
