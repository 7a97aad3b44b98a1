// la_matrix.hpp -- C++ host mirror of rust-la's hot-path API over la_hip.h.
//
// Same names, argument meaning and error behaviour as the reference:
//   la::Matrix<T>                 src/matrix/mod.rs:26-30 (row-major {no_rows, data})
//   operator*(Matrix, Matrix)     src/matrix/mod.rs:957-998 (impl Mul)
//   Matrix::mmul                  src/matrix/mmatrix.rs:82-98
//   Matrix::{det, solve, inverse, is_singular, is_non_singular}
//                                 src/matrix/mod.rs:1025-1047
//   la::LUDecomposition<T>        src/decomp/lu.rs:95-279
//   la::CholeskyDecomposition<T>  src/decomp/cholesky.rs:51-145
//   la::QRDecomposition<T>        src/decomp/qr.rs:20-237
//   Matrix::pinverse              src/matrix/mod.rs:1049-1057
// Reference panics (assert!) throw la::Panic; Option::None is an empty
// std::optional.  Every f64 factorisation, solve and product runs on the
// GPU through the C ABI; when no device or library is present the call
// throws la::DeviceError (there is no CPU path for f64).  Products of other
// element types (the reference's Mul is generic, e.g. its integer KATs) use
// the reference's loop, since the GPU path is f64-only.
#pragma once

#include <cstdint>
#include <cstring>
#include <optional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "la_hip.h"

namespace la {

struct Panic : std::logic_error {
    using std::logic_error::logic_error;
};
struct DeviceError : std::runtime_error {
    int status;
    DeviceError(int s, const std::string &what) : std::runtime_error(what), status(s) {}
};

// One context per host thread (the C ABI's threading rule).
inline la_ctx *thread_ctx() {
    thread_local struct Holder {
        la_ctx *ctx = nullptr;
        ~Holder() { if (ctx) la_ctx_destroy(ctx); }
    } h;
    if (!h.ctx) {
        int st = la_ctx_create(0, &h.ctx);
        if (st != LA_OK) throw DeviceError(st, std::string("la_ctx_create: ") + la_status_string(st));
    }
    return h.ctx;
}

inline void check(int st, const char *what) {
    if (st == LA_OK) return;
    if (st == LA_ERR_SHAPE) throw Panic(what);
    throw DeviceError(st, std::string(what) + ": " + la_status_string(st) + " (" +
                              la_ctx_last_error(thread_ctx()) + ")");
}

template <typename T>
class Matrix {
public:
    // Matrix::new (mod.rs:207-211)
    Matrix(size_t no_rows, size_t no_cols, std::vector<T> data) : no_rows_(no_rows), data_(std::move(data)) {
        if (no_rows * no_cols != data_.size()) throw Panic("Matrix::new: rows*cols != data.len()");
        if (!(no_rows > 0 && no_cols > 0)) throw Panic("Matrix::new: empty matrix");
    }
    static Matrix dirty(size_t r, size_t c) { return Matrix(r, c, std::vector<T>(r * c)); }
    static Matrix vector(std::vector<T> d) { size_t n = d.size(); return Matrix(n, 1, std::move(d)); }
    static Matrix row_vector(std::vector<T> d) { size_t n = d.size(); return Matrix(1, n, std::move(d)); }
    static Matrix zero(size_t r, size_t c) { return Matrix(r, c, std::vector<T>(r * c, T(0))); }
    // Matrix::id (mod.rs:416-427)
    static Matrix id(size_t m, size_t n) {
        std::vector<T> d(m * n, T(0));
        for (size_t i = 0; i < std::min(m, n); ++i) d[i * n + i] = T(1);
        return Matrix(m, n, std::move(d));
    }

    size_t rows() const { return no_rows_; }
    size_t cols() const { return data_.size() / no_rows_; }
    const std::vector<T> &get_data() const { return data_; }
    std::vector<T> &get_mut_data() { return data_; }
    // get (mod.rs:557-560): bounds assert
    T get(size_t r, size_t c) const {
        if (!(r < no_rows_ && c < cols())) throw Panic("Matrix::get: index out of bounds");
        return data_[r * cols() + c];
    }
    bool operator==(const Matrix &o) const { return no_rows_ == o.no_rows_ && data_ == o.data_; }
    bool operator!=(const Matrix &o) const { return !(*this == o); }

    Matrix t() const {
        const size_t r = rows(), c = cols();
        std::vector<T> d(r * c);
        for (size_t i = 0; i < r; ++i)
            for (size_t j = 0; j < c; ++j) d[j * r + i] = data_[i * c + j];
        return Matrix(c, r, std::move(d));
    }
    // permute_rows (mod.rs:757-759): row i of the result is row rows[i]
    Matrix permute_rows(const std::vector<size_t> &idx) const {
        const size_t c = cols();
        std::vector<T> d;
        d.reserve(idx.size() * c);
        for (size_t i : idx) {
            if (i >= no_rows_) throw Panic("permute_rows: index out of bounds");
            d.insert(d.end(), data_.begin() + i * c, data_.begin() + (i + 1) * c);
        }
        return Matrix(idx.size(), c, std::move(d));
    }

    // mmul (mmatrix.rs:82-98)
    Matrix &mmul(const Matrix &m, Matrix &dst) const;

    // Matrix wrappers (mod.rs:1025-1047), Float types only
    T det() const;
    std::optional<Matrix> solve(const Matrix &b) const;
    std::optional<Matrix> inverse() const;
    bool is_singular() const { return !is_non_singular(); }
    bool is_non_singular() const;
    // pinverse (mod.rs:1049-1057), f64
    Matrix pinverse() const;

private:
    size_t no_rows_;
    std::vector<T> data_;
};

// ---- Mul (mod.rs:957-980) ---------------------------------------------
template <typename T>
Matrix<T> operator*(const Matrix<T> &a, const Matrix<T> &b) {
    if (a.cols() != b.rows()) throw Panic("Mul: self.cols() != m.rows()");
    const size_t m = a.rows(), k = a.cols(), p = b.cols();
    std::vector<T> c(m * p);
    if constexpr (std::is_same<T, double>::value) {
        check(la_dgemm_rm(thread_ctx(), (int64_t)m, (int64_t)k, (int64_t)b.rows(), (int64_t)p,
                          a.get_data().data(), b.get_data().data(), c.data()),
              "Mul");
    } else {
        // generic element types keep the reference loop (mod.rs:965-973)
        const auto &A = a.get_data();
        const auto &B = b.get_data();
        for (size_t row = 0; row < m; ++row)
            for (size_t col = 0; col < p; ++col) {
                T res = T(0);
                for (size_t idx = 0; idx < k; ++idx) res = res + A[row * k + idx] * B[idx * p + col];
                c[row * p + col] = res;
            }
    }
    return Matrix<T>(m, p, std::move(c));
}

template <typename T>
Matrix<T> &Matrix<T>::mmul(const Matrix &m, Matrix &dst) const {
    if (cols() != m.rows()) throw Panic("mmul: self.cols() != m.rows()");
    if (dst.rows() != rows()) throw Panic("mmul: dst.rows() != self.rows()");
    if (dst.cols() != m.cols()) throw Panic("mmul: dst.cols() != m.cols()");
    if constexpr (std::is_same<T, double>::value) {
        check(la_dmmul_rm(thread_ctx(), (int64_t)rows(), (int64_t)cols(), (int64_t)m.rows(),
                          (int64_t)m.cols(), (int64_t)dst.rows(), (int64_t)dst.cols(),
                          data_.data(), m.get_data().data(), dst.get_mut_data().data()),
              "mmul");
    } else {
        dst = (*this) * m;
    }
    return dst;
}

// ---- LUDecomposition (lu.rs:95-279) -------------------------------------
template <typename T>
class LUDecomposition {
    static_assert(std::is_same<T, double>::value, "the GPU LU path is f64");

public:
    explicit LUDecomposition(const Matrix<T> &a) : lu_(a) {
        const size_t m = a.rows(), n = a.cols();
        std::vector<uint64_t> piv(m);
        int sign = 1;
        check(la_dgetrf_rm(thread_ctx(), (int64_t)m, (int64_t)n, a.get_data().data(),
                           lu_.get_mut_data().data(), piv.data(), &sign),
              "LUDecomposition::new");
        piv_.assign(piv.begin(), piv.end());
        pospivsign_ = sign != 0;
    }
    // is_non_singular (lu.rs:174-182): indexes lu[j*n+j] for j < cols
    bool is_non_singular() const {
        const size_t n = lu_.cols();
        for (size_t j = 0; j < n; ++j) {
            if (j * n + j >= lu_.get_data().size()) throw Panic("is_non_singular: index out of bounds");
            if (lu_.get_data()[j * n + j] == T(0)) return false;
        }
        return true;
    }
    bool is_singular() const { return !is_non_singular(); }
    // get_l (lu.rs:184-202)
    Matrix<T> get_l() const {
        const size_t m = lu_.rows(), n = std::min(lu_.rows(), lu_.cols()), c = lu_.cols();
        std::vector<T> d(m * n);
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j)
                d[i * n + j] = i > j ? lu_.get_data()[i * c + j] : (i == j ? T(1) : T(0));
        return Matrix<T>(m, n, std::move(d));
    }
    // get_u (lu.rs:204-215)
    Matrix<T> get_u() const {
        const size_t m = std::min(lu_.rows(), lu_.cols()), n = lu_.cols();
        std::vector<T> d(m * n);
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) d[i * n + j] = i <= j ? lu_.get_data()[i * n + j] : T(0);
        return Matrix<T>(m, n, std::move(d));
    }
    // get_p (lu.rs:217-220)
    Matrix<T> get_p() const {
        const size_t len = piv_.size();
        return Matrix<T>::id(len, len).permute_rows(piv_);
    }
    const std::vector<size_t> &get_piv() const { return piv_; }
    // det (lu.rs:224-232)
    T det() const {
        if (lu_.rows() != lu_.cols()) throw Panic("det: not square");
        const size_t n = lu_.cols();
        T d = pospivsign_ ? T(1) : T(-1);
        for (size_t j = 0; j < n; ++j) d = d * lu_.get_data()[j * n + j];
        return d;
    }
    // solve (lu.rs:237-278)
    std::optional<Matrix<T>> solve(const Matrix<T> &b) const {
        const size_t m = lu_.rows(), n = lu_.cols(), nx = b.cols();
        if (b.rows() != m) throw Panic("LUDecomposition::solve: b.rows() != lu.rows()");
        if (!is_non_singular()) return std::nullopt;
        std::vector<uint64_t> piv(piv_.begin(), piv_.end());
        std::vector<T> x(m * nx);
        const int st = la_dgetrs_rm(thread_ctx(), (int64_t)m, (int64_t)n, lu_.get_data().data(),
                                    piv.data(), (int64_t)b.rows(), (int64_t)nx,
                                    b.get_data().data(), x.data());
        if (st == LA_SINGULAR) return std::nullopt;
        check(st, "LUDecomposition::solve");
        return Matrix<T>(m, nx, std::move(x));
    }

private:
    Matrix<T> lu_;
    bool pospivsign_ = true;
    std::vector<size_t> piv_;
};

// ---- CholeskyDecomposition (cholesky.rs:51-145) ---------------------------
template <typename T>
class CholeskyDecomposition {
    static_assert(std::is_same<T, double>::value, "the GPU Cholesky path is f64");

public:
    // CholeskyDecomposition::new -> Option<Self>
    static std::optional<CholeskyDecomposition> create(const Matrix<T> &m) {
        if (m.rows() != m.cols()) return std::nullopt;
        const size_t n = m.rows();
        std::vector<T> l(n * n);
        const int st = la_dpotrf_rm(thread_ctx(), (int64_t)m.rows(), (int64_t)m.cols(),
                                    m.get_data().data(), l.data());
        if (st == LA_NOT_SQUARE || st == LA_NOT_SYMMETRIC || st == LA_NOT_SPD) return std::nullopt;
        check(st, "CholeskyDecomposition::new");
        return CholeskyDecomposition(Matrix<T>(n, n, std::move(l)));
    }
    const Matrix<T> &get_l() const { return l_; }
    // solve (cholesky.rs:116-144)
    Matrix<T> solve(const Matrix<T> &b) const {
        if (l_.rows() != b.rows()) throw Panic("CholeskyDecomposition::solve: l.rows() != b.rows()");
        const size_t n = l_.rows(), nx = b.cols();
        std::vector<T> x(n * nx);
        check(la_dpotrs_rm(thread_ctx(), (int64_t)n, l_.get_data().data(), (int64_t)b.rows(),
                           (int64_t)nx, b.get_data().data(), x.data()),
              "CholeskyDecomposition::solve");
        return Matrix<T>(n, nx, std::move(x));
    }

private:
    explicit CholeskyDecomposition(Matrix<T> l) : l_(std::move(l)) {}
    Matrix<T> l_;
};

// ---- QRDecomposition (qr.rs:20-237) ---------------------------------------
template <typename T>
class QRDecomposition {
    static_assert(std::is_same<T, double>::value, "the GPU QR path is f64");

public:
    // QRDecomposition::new (qr.rs:26-43)
    explicit QRDecomposition(const Matrix<T> &a) : qr_(a) {
        const size_t m = a.rows(), n = a.cols();
        std::vector<T> qr(m * n);
        rdiag_.resize(std::min(m, n));
        check(la_dgeqrf_rm(thread_ctx(), (int64_t)m, (int64_t)n, a.get_data().data(), qr.data(),
                           rdiag_.data()),
              "QRDecomposition::new");
        qr_ = Matrix<T>(m, n, std::move(qr));
    }
    // is_full_rank (qr.rs:108-115): rdiag is indexed for every column
    bool is_full_rank() const {
        for (size_t j = 0; j < qr_.cols(); ++j) {
            if (j >= rdiag_.size()) throw Panic("is_full_rank: index out of bounds");
            if (rdiag_[j] == T(0)) return false;
        }
        return true;
    }
    // get_h (qr.rs:117-132): the lower trapezoid of qr
    Matrix<T> get_h() const {
        const size_t m = qr_.rows(), n = qr_.cols();
        std::vector<T> h(m * n);
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j) h[i * n + j] = i >= j ? qr_.get_data()[i * n + j] : T(0);
        return Matrix<T>(m, n, std::move(h));
    }
    // get_r (qr.rs:134-148)
    Matrix<T> get_r() const {
        const size_t m = qr_.rows(), n = qr_.cols();
        std::vector<T> r(m * n);
        for (size_t i = 0; i < m; ++i)
            for (size_t j = 0; j < n; ++j)
                r[i * n + j] = i < j ? qr_.get_data()[i * n + j] : (i == j ? rdiag_[i] : T(0));
        return Matrix<T>(m, n, std::move(r));
    }
    // get_q (qr.rs:150-194), m x m
    Matrix<T> get_q() const {
        const size_t m = qr_.rows(), n = qr_.cols();
        std::vector<T> q(m * m);
        check(la_dqr_q_rm(thread_ctx(), (int64_t)m, (int64_t)n, qr_.get_data().data(),
                          rdiag_.data(), q.data()),
              "QRDecomposition::get_q");
        return Matrix<T>(m, m, std::move(q));
    }
    // solve (qr.rs:196-237) as a least-squares solve: n x nx, None when not
    // of full rank; the reference's panics (b.rows() != rows, cols > rows)
    // throw Panic.  Where the reference's formula is not the reflection, the
    // reflections are applied (la_hip.h, la_dqrs_rm).
    std::optional<Matrix<T>> solve(const Matrix<T> &b) const {
        if (b.rows() != qr_.rows()) throw Panic("QRDecomposition::solve: b.rows() != rows");
        if (!is_full_rank()) return std::nullopt;
        const size_t m = qr_.rows(), n = qr_.cols(), nx = b.cols();
        std::vector<T> x(n * nx);
        const int st = la_dqrs_rm(thread_ctx(), (int64_t)m, (int64_t)n, qr_.get_data().data(),
                                  rdiag_.data(), (int64_t)b.rows(), (int64_t)nx,
                                  b.get_data().data(), x.data());
        if (st == LA_SINGULAR) return std::nullopt;
        check(st, "QRDecomposition::solve");
        return Matrix<T>(n, nx, std::move(x));
    }
    const std::vector<T> &rdiag() const { return rdiag_; }

private:
    Matrix<T> qr_;
    std::vector<T> rdiag_;
};

// ---- Matrix wrappers (mod.rs:1025-1047) ---------------------------------
template <typename T>
T Matrix<T>::det() const {
    if (cols() != no_rows_) throw Panic("det: not square");
    return LUDecomposition<T>(*this).det();
}
template <typename T>
std::optional<Matrix<T>> Matrix<T>::solve(const Matrix &b) const {
    return LUDecomposition<T>(*this).solve(b);
}
template <typename T>
std::optional<Matrix<T>> Matrix<T>::inverse() const {
    if (no_rows_ != cols()) throw Panic("inverse: not square");
    if constexpr (std::is_same_v<T, double>) {
        // one device round trip: LU, P I formed on the device, MFMA TRSMs
        std::vector<double> out(no_rows_ * no_rows_);
        const int st = la_dinverse_rm(thread_ctx(), (int64_t)no_rows_, (int64_t)no_rows_,
                                      data_.data(), out.data());
        if (st == LA_SINGULAR) return std::nullopt;
        check(st, "inverse");
        return Matrix(no_rows_, no_rows_, std::move(out));
    } else {
        return LUDecomposition<T>(*this).solve(Matrix::id(no_rows_, no_rows_));
    }
}
template <typename T>
Matrix<T> Matrix<T>::pinverse() const {
    static_assert(std::is_same<T, double>::value, "the GPU pinverse path is f64");
    const size_t m = no_rows_, n = cols();
    std::vector<T> p(n * m);
    const int st = la_dpinverse_rm(thread_ctx(), (int64_t)m, (int64_t)n, data_.data(), p.data());
    if (st == LA_SINGULAR) throw Panic("pinverse: inverse().unwrap() on None");
    check(st, "pinverse");
    return Matrix(n, m, std::move(p));
}
template <typename T>
bool Matrix<T>::is_non_singular() const {
    if (no_rows_ != cols()) throw Panic("is_non_singular: not square");
    return LUDecomposition<T>(*this).is_non_singular();
}

// m!(a, b; c, d)-style literal (src/macros.rs:39-42)
template <typename T>
Matrix<T> m(std::initializer_list<std::initializer_list<T>> rows) {
    std::vector<T> d;
    size_t r = 0, c = 0;
    for (auto &row : rows) {
        if (r == 0) c = row.size();
        else if (row.size() != c) throw Panic("m!: ragged rows");
        d.insert(d.end(), row.begin(), row.end());
        ++r;
    }
    return Matrix<T>(r, c, std::move(d));
}

}  // namespace la
