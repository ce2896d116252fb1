// A reference-style caller of the C++ drop-in headers: it builds a mesh with
// OrthHexMesh + split_hex6_scan + rebuild_cell_topology + ComputingMesh (the reference's
// mesh; periodic cases: split_hex6_periodic + make_periodic), tags boundaries like the vortex
// example (examples/Vortex/tools.h:22-31), defines the exact-solution hooks of
// base/exact.h, projects the vortex with DGBasisEvaluator::func2coef, and calls
// ExplicitConvection<P>::eval / eval_cells / eval_internals / eval_boundarys, the
// ExplicitSSPRK3 driver, PositiveLimiter (constructMinMax, apply, apply_1, apply_2), and one
// backward-Euler solve with ImplicitConvection + BlockSparseMatrix + EigenSparseSolver.  Everything it computes is written to a binary file that
// tests/test_cpp_dropin.py checks against the CPU oracle.
//
//   explicit_convection_check OUT.bin
// exit status: 0 ok, 2 a libfvdg call failed (message on stderr), 1 bad usage.
#include <cstdio>
#include <cstring>
#include <fstream>

#include "DG/DG_Schemes/ExplicitConvection.h"
#include "DG/DG_Schemes/ExplicitSSPRK3.h"
#include "DG/DG_Schemes/ImplicitConvection.h"
#include "DG/DG_Schemes/PositiveLimiter.h"
#include "EigenSolver/EigenSparseSolver.h"

// ---- the vortex of examples/Vortex/problem.h:7-110 (centre (4, 4.5), free stream (1, 0.5, 0), strength 5)
namespace {
constexpr Scalar kGamma = 1.4, kStrength = 5.0, kX0 = 4.0, kY0 = 4.5, kU = 1.0, kV = 0.5;
constexpr Scalar kPi = 3.14159265358979323846;

void vortex(const vector3f& x, Scalar t, Scalar U[5]) {
    const Scalar dx = x[0] - (kX0 + kU * t), dy = x[1] - (kY0 + kV * t);
    const Scalar r2 = dx * dx + dy * dy;
    const Scalar T = 1 - kStrength * kStrength * (kGamma - 1) / (16 * kPi * kPi * kGamma) * std::exp(2 * (1 - r2));
    const Scalar rho = std::pow(T, 1 / (kGamma - 1));
    const Scalar u = kU - kStrength * dy / (2 * kPi) * std::exp(1 - r2);
    const Scalar v = kV + kStrength * dx / (2 * kPi) * std::exp(1 - r2);
    const Scalar p = std::pow(rho, kGamma);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = 0;
    U[4] = p / (kGamma - 1) + 0.5 * rho * (u * u + v * v);
}
Scalar component(const CompTetrahedron& cell, const vector3f& Xi, Scalar t, int k) {
    Scalar U[5];
    vortex(cell.transform_to_physical(Xi), t, U);
    return U[k];
}
}  // namespace

Scalar rho_Xi(const CompTetrahedron& cell, const vector3f& Xi, Scalar t) { return component(cell, Xi, t, 0); }
Scalar rhou_Xi(const CompTetrahedron& cell, const vector3f& Xi, Scalar t) { return component(cell, Xi, t, 1); }
Scalar rhov_Xi(const CompTetrahedron& cell, const vector3f& Xi, Scalar t) { return component(cell, Xi, t, 2); }
Scalar rhow_Xi(const CompTetrahedron& cell, const vector3f& Xi, Scalar t) { return component(cell, Xi, t, 3); }
Scalar rhoe_Xi(const CompTetrahedron& cell, const vector3f& Xi, Scalar t) { return component(cell, Xi, t, 4); }

namespace {

struct Writer {
    std::ofstream f;
    explicit Writer(const char* path) : f(path, std::ios::binary) {}
    void i64(int64_t v) { f.write(reinterpret_cast<const char*>(&v), sizeof v); }
    void f64(const double* p, size_t n) { f.write(reinterpret_cast<const char*>(p), n * sizeof(double)); }
    void mesh(const ComputingMesh& m) {
        i64(m.m_points.size());
        i64(m.m_cells.size());
        i64(m.m_faces.size());
        for (auto& p : m.m_points) f64(p.data(), 3);
        for (auto& c : m.m_cells)
            for (int a = 0; a < 4; ++a) i64(c.m_nodes[a]);
        for (auto& F : m.m_faces) {
            for (int k = 0; k < 3; ++k) i64(F.m_nodes[k]);
            for (int k = 0; k < 3; ++k) i64(F.m_nodes_r[k]);
            i64(F.m_neighbor_cells[0]);
            i64(F.is_boundary() ? -1 : (int64_t)F.m_neighbor_cells[1]);
        }
        for (auto b : m.m_boundaryTypes) i64((int64_t)b);
    }
    template <uInt B>
    void vec(const LongVector<B>& v) {
        f64(v.data(), v.size() * B);
    }
};

template <uInt P>
LongVector<5 * DGBasisEvaluator<P>::NumBasis> project(const ComputingMesh& m, Scalar t) {
    using Basis = DGBasisEvaluator<P>;
    LongVector<5 * Basis::NumBasis> U(m.m_cells.size());
    for (size_t e = 0; e < m.m_cells.size(); ++e) {
        const auto& cell = m.m_cells[e];
        const auto c = Basis::func2coef([&](const vector3f& xi) {
            Scalar s[5];
            vortex(cell.transform_to_physical(xi), t, s);
            return DenseMatrix<5, 1>({s[0], s[1], s[2], s[3], s[4]});
        });
        for (uInt k = 0; k < Basis::NumBasis; ++k)
            for (int v = 0; v < 5; ++v) U[e][5 * k + v] = c[k][v];
    }
    return U;
}

// Dirichlet / Pseudo3DZ box: residual parts at t = 0.3 and two SSP-RK3 steps.
template <uInt P>
void boundary_case(Writer& w, int flux) {
    OrthHexMesh g({0, 0, 0}, {10, 10, 2.5}, {4, 4, 2});
    g.split_hex6_scan();
    g.rebuild_cell_topology();
    ComputingMesh cm(g);
    for (size_t f = 0; f < cm.m_faces.size(); ++f)
        if (cm.m_faces[f].is_boundary())
            cm.m_boundaryTypes[f] =
                std::abs(cm.m_faces[f].m_normal[2]) > 0.5 ? BoundaryType::Pseudo3DZ : BoundaryType::Dirichlet;
    const Scalar t = 0.3;
    const auto U = project<P>(cm, t);
    ExplicitConvection<P> conv;
    conv.set_riemann_solver(flux);
    const auto R = conv.eval(cm, U, t);
    const auto Rc = conv.eval_cells(cm, U, t);
    const auto Ri = conv.eval_internals(cm, U, t);
    const auto Rb = conv.eval_boundarys(cm, U, t);
    ExplicitSSPRK3<P> rk(cm);
    rk.convection().set_riemann_solver(flux);
    rk.set_solution(U);
    const Scalar dt = rk.stable_dt(0.3);
    rk.advance(t, dt, 2);
    const auto U2 = rk.solution();
    const auto M = rk.mass();
    w.i64(P);
    w.i64(flux);
    w.mesh(cm);
    w.vec(U);
    w.vec(R);
    w.vec(Rc);
    w.vec(Ri);
    w.vec(Rb);
    w.f64(&dt, 1);
    w.vec(U2);
    w.vec(M);
}

// periodic jittered cube: residual and one step (no boundary faces)
template <uInt P>
void periodic_case(Writer& w, int flux) {
    OrthHexMesh g({0, 0, 0}, {10, 10, 10}, {3, 3, 3});
    g.split_hex6_periodic();
    g.make_periodic({0, 0, 0}, {10, 10, 10}, {true, true, true});
    g.jitter_points(0.1, 7, {0, 0, 0}, {10, 10, 10}, {10.0 / 3, 10.0 / 3, 10.0 / 3}, {true, true, true});
    ComputingMesh cm(g);
    const auto U = project<P>(cm, 0.0);
    ExplicitConvection<P, MonatomicFluxC> conv;
    conv.set_riemann_solver(flux);
    const auto R = conv.eval(cm, U);
    ExplicitSSPRK3<P, MonatomicFluxC> rk(cm);
    rk.set_solution(U);
    rk.convection().set_riemann_solver(flux);  // a scheme change after the upload keeps the state
    const Scalar dt = 0.01;
    rk.step(0.0, dt);
    const auto U1 = rk.solution();
    w.i64(P);
    w.i64(flux);
    w.mesh(cm);
    w.vec(U);
    w.vec(R);
    w.f64(&dt, 1);
    w.vec(U1);
}

// limiter on a periodic jittered cube, on a sharpened vortex: point bounds with apply, and
// neighbour-average bounds with apply_1 then apply_2 (as examples/Eular2D/Sedov/Sedov.cpp:92,188-189)
template <uInt P>
void limiter_case(Writer& w) {
    constexpr uInt N = DGBasisEvaluator<P>::NumBasis;
    OrthHexMesh g({0, 0, 0}, {10, 10, 10}, {3, 3, 3});
    g.split_hex6_periodic();
    g.make_periodic({0, 0, 0}, {10, 10, 10}, {true, true, true});
    g.jitter_points(0.1, 9, {0, 0, 0}, {10, 10, 10}, {10.0 / 3, 10.0 / 3, 10.0 / 3}, {true, true, true});
    ComputingMesh cm(g);
    const auto Uprev = project<P>(cm, 0.0);
    auto U = Uprev;
    for (size_t e = 0; e < U.size(); ++e)
        for (uInt k = 5; k < 5 * N; ++k) U[e][k] *= 6.0 * (1.0 + 0.5 * std::sin(0.7 * e + k));
    auto A = U;
    PositiveLimiter<P> lim(cm, 1.4);
    lim.constructMinMax(A, Uprev);
    lim.apply(A, Uprev);
    auto B = U;
    using QuadC = typename AutoQuadSelector<P, GaussLegendreTet::Auto>::type;
    PositiveLimiter<P, QuadC, true> lim2(cm, 1.4);
    lim2.constructMinMax(B, Uprev);
    lim2.apply_1(B, Uprev);
    lim2.apply_2(B, Uprev);
    w.i64(P);
    w.mesh(cm);
    w.vec(Uprev);
    w.vec(U);
    w.vec(A);
    w.vec(B);
}

// one backward-Euler system as the reference's implicit drivers build it
// (examples/Eular2D/Sedov/Sedov.cpp:153-180): assemble at U, add mass/dt blocks and mass/dt U to the
// right-hand side, finalize, DGMRES from U
template <uInt P>
void implicit_case(Writer& w) {
    constexpr uInt NB = 5 * DGBasisEvaluator<P>::NumBasis;
    OrthHexMesh g({0, 0, 0}, {10, 10, 10}, {3, 3, 3});
    g.split_hex6_periodic();
    g.make_periodic({0, 0, 0}, {10, 10, 10}, {true, true, true});
    g.jitter_points(0.1, 11, {0, 0, 0}, {10, 10, 10}, {10.0 / 3, 10.0 / 3, 10.0 / 3}, {true, true, true});
    ComputingMesh cm(g);
    const auto U = project<P>(cm, 0.0);
    const Scalar dt = 0.05;
    ExplicitSSPRK3<P> rk(cm);
    const auto mass = rk.mass();
    ImplicitConvection<P> convection;
    BlockSparseMatrix<NB, NB> sparse_mat;
    LongVector<NB> rhs(U.size());
    convection.assemble(cm, U, 0.0, sparse_mat, rhs);
    for (uInt e = 0; e < cm.m_cells.size(); ++e) {
        sparse_mat.add_block(e, e, DenseMatrix<NB, NB>::Diag(mass[e] / dt));
        rhs[e] += mass[e] / dt * U[e];
    }
    sparse_mat.finalize();
    EigenSparseSolver<NB, NB> solver(sparse_mat, rhs);
    solver.set_tol(1e-13);
    uInt iter = 0;
    Scalar residual = 0;
    const auto x = solver.DGMRES(U, iter, residual);
    const auto Ax = sparse_mat.multiply(x);
    const int64_t it = iter;
    w.i64(P);
    w.mesh(cm);
    w.vec(U);
    w.f64(&dt, 1);
    w.vec(mass);
    w.vec(rhs);
    w.vec(x);
    w.vec(Ax);
    w.i64(it);
    w.f64(&residual, 1);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s OUT.bin\n", argv[0]);
        return 1;
    }
    try {
        Writer w(argv[1]);
        w.i64(5);  // number of cases
        boundary_case<1>(w, FVDG_FLUX_LF);
        boundary_case<3>(w, FVDG_FLUX_ROE);
        periodic_case<2>(w, FVDG_FLUX_ROE);
        limiter_case<2>(w);
        implicit_case<1>(w);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "explicit_convection_check: %s\n", e.what());
        return 2;
    }
    std::printf("explicit_convection_check: done\n");
    return 0;
}
