// Orthogonal modal basis on the unit tetrahedron {x,y,z >= 0, x+y+z <= 1}.
//
// The reference hard-codes this basis as sympy-generated polynomials
// phi_0..phi_285 (include/DG/DG_Basis/DG_Basis_Func.h, e.g. phi_1 = 4x-1,
// phi_2 = x+3y-1, phi_3 = x+y+2z-1 at :25-80) and evaluates them through
// DGBasisEvaluator<Order>::eval_all / grad_all (DG_Basis_Evaluator.h:14-31).
// Those polynomials are the unnormalised Dubiner (collapsed-coordinate Jacobi)
// products
//
//   phi_{ijk} = [s^k P_k^{(0,0)}(2z/s-1)] * [r^j P_j^{(2k+1,0)}(2y/r-1)]
//               * P_i^{(2j+2k+2,0)}(2x-1),      s = 1-x-y,  r = 1-x,
//
// with the classical Jacobi normalisation P_n^{(a,0)}(1) = binom(n+a, n),
// enumerated by total degree n, then k = 0..n, then j = 0..n-k (i = n-j-k).
// This file evaluates them with the homogeneous form of the Jacobi three-term
// recurrence (no division by s or r, so it is exact at the collapsed vertex),
// and derivatives by forward-mode differentiation of the same recurrence.
#pragma once
#include <type_traits>
#include "GaussQuadrature/GaussLegendre.h"
#include "Matrix/DenseMatrix.h"

namespace fvdg_basis_detail {

// value and gradient (d/dx, d/dy, d/dz) carried together
template <typename T>
struct Dual {
    T v, dx, dy, dz;
};
template <typename T>
inline Dual<T> operator+(const Dual<T>& a, const Dual<T>& b) { return {a.v + b.v, a.dx + b.dx, a.dy + b.dy, a.dz + b.dz}; }
template <typename T>
inline Dual<T> operator-(const Dual<T>& a, const Dual<T>& b) { return {a.v - b.v, a.dx - b.dx, a.dy - b.dy, a.dz - b.dz}; }
template <typename T>
inline Dual<T> operator*(const Dual<T>& a, const Dual<T>& b) {
    return {a.v * b.v, a.dx * b.v + a.v * b.dx, a.dy * b.v + a.v * b.dy, a.dz * b.v + a.v * b.dz};
}
template <typename T>
inline Dual<T> operator*(T c, const Dual<T>& a) { return {c * a.v, c * a.dx, c * a.dy, c * a.dz}; }

// Q_n(z; s) = s^n P_n^{(alpha,0)}((2z - s)/s) for n = 0..nmax, homogeneous recurrence.
template <typename T>
inline void jacobi_homog(int nmax, int alpha, const Dual<T>& z, const Dual<T>& s, Dual<T>* q) {
    q[0] = {T(1), T(0), T(0), T(0)};
    if (nmax < 1) return;
    const T a = T(alpha);
    q[1] = (a + T(1)) * s + (a + T(2)) * (z - s);
    const Dual<T> t2 = T(2) * z - s;
    const Dual<T> s2 = s * s;
    for (int n = 2; n <= nmax; ++n) {
        const T N = T(n);
        const T a1 = T(2) * N * (N + a) * (T(2) * N + a - T(2));
        const T a2 = (T(2) * N + a - T(1)) * a * a;
        const T a3 = (T(2) * N + a - T(2)) * (T(2) * N + a - T(1)) * (T(2) * N + a);
        const T a4 = T(2) * (N + a - T(1)) * (N - T(1)) * (T(2) * N + a);
        const Dual<T> lin = a2 * s + a3 * t2;
        const Dual<T> num = lin * q[n - 1] - a4 * (s2 * q[n - 2]);
        const T inv = T(1) / a1;
        q[n] = inv * num;
    }
}

// Evaluate all basis functions up to total degree `order` (values and gradients).
template <typename T>
inline void eval_dubiner(int order, T x, T y, T z, T* val, T (*grad)[3]) {
    const Dual<T> X{x, T(1), T(0), T(0)}, Y{y, T(0), T(1), T(0)}, Z{z, T(0), T(0), T(1)};
    const Dual<T> one{T(1), T(0), T(0), T(0)};
    const Dual<T> s = one - X - Y;  // 1 - x - y
    const Dual<T> r = one - X;      // 1 - x
    Dual<T> zq[16], yq[16], xq[16];
    jacobi_homog(order, 0, Z, s, zq);
    int idx = 0;
    for (int n = 0; n <= order; ++n) {
        for (int k = 0; k <= n; ++k) {
            for (int j = 0; j <= n - k; ++j) {
                const int i = n - j - k;
                jacobi_homog(j, 2 * k + 1, Y, r, yq);
                jacobi_homog(i, 2 * j + 2 * k + 2, X, one, xq);
                const Dual<T> phi = zq[k] * yq[j] * xq[i];
                if (val) val[idx] = phi.v;
                if (grad) {
                    grad[idx][0] = phi.dx;
                    grad[idx][1] = phi.dy;
                    grad[idx][2] = phi.dz;
                }
                ++idx;
            }
        }
    }
}

}  // namespace fvdg_basis_detail

constexpr inline int dg_num_basis(int order) { return (order + 1) * (order + 2) * (order + 3) / 6; }

// Run-time-order entry points (values: out[num_basis]; gradients: out[num_basis][3]).
inline void dg_basis_eval(int order, Scalar x, Scalar y, Scalar z, Scalar* out) {
    fvdg_basis_detail::eval_dubiner<Scalar>(order, x, y, z, out, nullptr);
}
inline void dg_basis_grad(int order, Scalar x, Scalar y, Scalar z, Scalar (*out)[3]) {
    fvdg_basis_detail::eval_dubiner<Scalar>(order, x, y, z, nullptr, out);
}

template <uInt Order>
class DGBasisEvaluator {
public:
    static constexpr uInt OrderBasis = Order;
    static constexpr uInt NumBasis = (Order + 3) * (Order + 2) * (Order + 1) / 6;

    template <typename Type>
    static std::array<Type, NumBasis> eval_all(Type x, Type y, Type z) {
        std::array<Type, NumBasis> v{};
        fvdg_basis_detail::eval_dubiner<Type>((int)Order, x, y, z, v.data(), nullptr);
        return v;
    }
    template <typename Type>
    static std::array<std::array<Type, 3>, NumBasis> grad_all(Type x, Type y, Type z) {
        std::array<std::array<Type, 3>, NumBasis> g{};
        Type tmp[NumBasis][3];
        fvdg_basis_detail::eval_dubiner<Type>((int)Order, x, y, z, nullptr, tmp);
        for (uInt k = 0; k < NumBasis; ++k) g[k] = {tmp[k][0], tmp[k][1], tmp[k][2]};
        return g;
    }

    // Diagonal L2 projection on the order's volume rule, basis and mass diagonal in
    // long double (reference DG_Basis_Evaluator.h:33-59).  func(vector3f xi) may return
    // Scalar or DenseMatrix<M,1>.
    template <typename Func>
    static auto func2coef(const Func& func) {
        using QuadC = typename AutoQuadSelector<Order, GaussLegendreTet::Auto>::type;
        using R = std::decay_t<decltype(func(QuadC::points[0]))>;
        std::array<R, NumBasis> result;
        std::array<std::array<long double, NumBasis>, QuadC::num_points> phi;
        for (int g = 0; g < QuadC::num_points; ++g) {
            const auto& p = QuadC::points[g];
            phi[g] = eval_all<long double>(p[0], p[1], p[2]);
        }
        std::array<R, QuadC::num_points> fv;
        for (int g = 0; g < QuadC::num_points; ++g) fv[g] = func(QuadC::points[g]);
        for (uInt k = 0; k < NumBasis; ++k) {
            long double diag = 0.0L;
            R rhs = R(0.0);
            for (int g = 0; g < QuadC::num_points; ++g) {
                diag += phi[g][k] * phi[g][k] * (long double)QuadC::weights[g];
                rhs += fv[g] * (Scalar)(phi[g][k] * (long double)QuadC::weights[g]);
            }
            result[k] = rhs / (Scalar)diag;
        }
        return result;
    }
};
