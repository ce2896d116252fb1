// Fixed-size row-major dense matrix, the value type of the DG state blocks.
// API-compatible subset of the reference's include/Matrix/DenseMatrix.h:9-257
// (operator()(row,col), operator[], element-wise + - * / with matrices and
// scalars, multiply(), transpose(), inverse() of a 3x3 by cofactors
// (DenseMatrix.h:521-538), SubMat).  Plain scalar loops: this type is only used
// on the host for setup and for the caller-facing API; the residual itself
// runs in the HIP kernels.
#pragma once
#include <initializer_list>
#include <ostream>
#include "base/Type.h"

template <uInt M, uInt N>
class DenseMatrix {
public:
    static constexpr uInt Rows = M, Cols = N, Size = M * N;
    std::array<Scalar, M * N> data{};

    DenseMatrix() = default;
    DenseMatrix(Scalar v) { data.fill(v); }
    DenseMatrix(std::initializer_list<Scalar> il) {
        uInt i = 0;
        for (Scalar v : il) {
            if (i < Size) data[i++] = v;
        }
    }
    DenseMatrix(const std::array<Scalar, Size>& a) : data(a) {}
    // rows of a 3x3 given as three vectors (reference: DenseMatrix(std::array<vector3f,3>))
    template <uInt MM = M, uInt NN = N, typename = std::enable_if_t<MM == 3 && NN == 3>>
    DenseMatrix(const std::array<vector3f, 3>& rows) {
        for (uInt i = 0; i < 3; ++i)
            for (uInt j = 0; j < 3; ++j) data[i * 3 + j] = rows[i][j];
    }
    template <uInt MM = M, uInt NN = N, typename = std::enable_if_t<NN == 1 && MM == 3>>
    DenseMatrix(const vector3f& v, int = 0) {
        for (uInt i = 0; i < 3; ++i) data[i] = v[i];
    }

    static DenseMatrix Zeros() { return DenseMatrix(0.0); }
    static DenseMatrix Ones() { return DenseMatrix(1.0); }
    static DenseMatrix Identity() {
        DenseMatrix r(0.0);
        for (uInt i = 0; i < std::min(M, N); ++i) r(i, i) = 1.0;
        return r;
    }
    // square matrix with d on the diagonal (reference use: examples/Eular2D/Sedov/Sedov.cpp:164)
    static DenseMatrix Diag(const DenseMatrix<M, 1>& d) {
        static_assert(M == N, "Diag builds a square matrix");
        DenseMatrix r(0.0);
        for (uInt i = 0; i < M; ++i) r(i, i) = d[i];
        return r;
    }

    Scalar& operator()(uInt r, uInt c) { return data[r * N + c]; }
    const Scalar& operator()(uInt r, uInt c) const { return data[r * N + c]; }
    Scalar& operator[](uInt i) { return data[i]; }
    const Scalar& operator[](uInt i) const { return data[i]; }
    Scalar* data_ptr() { return data.data(); }
    const Scalar* data_ptr() const { return data.data(); }
    constexpr uInt size() const { return Size; }
    DenseMatrix& operator=(Scalar v) {
        data.fill(v);
        return *this;
    }
    operator std::array<Scalar, Size>() const { return data; }

#define FVDG_EWISE(op)                                                                  \
    DenseMatrix& operator op##=(const DenseMatrix& r) {                                 \
        for (uInt i = 0; i < Size; ++i) data[i] op## = r.data[i];                       \
        return *this;                                                                   \
    }                                                                                   \
    DenseMatrix operator op(const DenseMatrix& r) const {                               \
        DenseMatrix t(*this);                                                           \
        return t op## = r;                                                              \
    }                                                                                   \
    DenseMatrix& operator op##=(Scalar v) {                                             \
        for (auto& d : data) d op## = v;                                                \
        return *this;                                                                   \
    }                                                                                   \
    DenseMatrix operator op(Scalar v) const {                                           \
        DenseMatrix t(*this);                                                           \
        return t op## = v;                                                              \
    }                                                                                   \
    friend DenseMatrix operator op(Scalar v, const DenseMatrix& r) {                    \
        DenseMatrix t(r);                                                               \
        for (auto& d : t.data) d = v op d;                                              \
        return t;                                                                       \
    }
    FVDG_EWISE(+)
    FVDG_EWISE(-)
    FVDG_EWISE(*)
    FVDG_EWISE(/)
#undef FVDG_EWISE
    DenseMatrix operator-() const {
        DenseMatrix t;
        for (uInt i = 0; i < Size; ++i) t.data[i] = -data[i];
        return t;
    }

    template <uInt K>
    DenseMatrix<M, K> multiply(const DenseMatrix<N, K>& r) const {
        DenseMatrix<M, K> out(0.0);
        for (uInt i = 0; i < M; ++i)
            for (uInt k = 0; k < N; ++k)
                for (uInt j = 0; j < K; ++j) out(i, j) += (*this)(i, k) * r(k, j);
        return out;
    }
    DenseMatrix<M, 1> multiply(const std::array<Scalar, N>& r) const {
        DenseMatrix<M, 1> out(0.0);
        for (uInt i = 0; i < M; ++i)
            for (uInt k = 0; k < N; ++k) out(i, 0) += (*this)(i, k) * r[k];
        return out;
    }
    DenseMatrix<N, M> transpose() const {
        DenseMatrix<N, M> t;
        for (uInt i = 0; i < M; ++i)
            for (uInt j = 0; j < N; ++j) t(j, i) = (*this)(i, j);
        return t;
    }
    Scalar dot(const DenseMatrix& r) const {
        Scalar s = 0;
        for (uInt i = 0; i < Size; ++i) s += data[i] * r.data[i];
        return s;
    }
    Scalar norm() const { return std::sqrt(dot(*this)); }

    template <uInt SM, uInt SN>
    DenseMatrix<SM, SN> SubMat(uInt r0, uInt c0) const {
        DenseMatrix<SM, SN> s;
        for (uInt i = 0; i < SM; ++i)
            for (uInt j = 0; j < SN; ++j) s(i, j) = (*this)(r0 + i, c0 + j);
        return s;
    }

    // 3x3 inverse by cofactors divided by the determinant expanded along the
    // first column, same operation order as the reference (DenseMatrix.h:521-538).
    DenseMatrix inverse() const {
        static_assert(M == 3 && N == 3, "inverse() is provided for 3x3 only");
        DenseMatrix r;
        const DenseMatrix& A = *this;
        r(0, 0) = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1);
        r(1, 0) = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2);
        r(2, 0) = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
        r(0, 1) = A(2, 1) * A(0, 2) - A(2, 2) * A(0, 1);
        r(1, 1) = A(2, 2) * A(0, 0) - A(2, 0) * A(0, 2);
        r(2, 1) = A(2, 0) * A(0, 1) - A(2, 1) * A(0, 0);
        r(0, 2) = A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1);
        r(1, 2) = A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2);
        r(2, 2) = A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0);
        Scalar det = 0;
        det += A(0, 0) * r(0, 0);
        det += A(1, 0) * r(0, 1);
        det += A(2, 0) * r(0, 2);
        r /= det;
        return r;
    }

    friend std::ostream& operator<<(std::ostream& os, const DenseMatrix& m) {
        for (uInt i = 0; i < M; ++i) {
            for (uInt j = 0; j < N; ++j) os << m(i, j) << (j + 1 < N ? " " : "");
            os << "\n";
        }
        return os;
    }
};

template <uInt M, uInt N>
inline DenseMatrix<M, N> pow(const DenseMatrix<M, N>& a, Scalar e) {
    DenseMatrix<M, N> r;
    for (uInt i = 0; i < M * N; ++i) r[i] = std::pow(a[i], e);
    return r;
}
