// ref_driver.cpp -- TEST INFRASTRUCTURE: extern "C" entry points into the reference's
// own compiled code, used to pin the oracle (oracle/dg_oracle.c) and the library's
// tables against the reference.  Built by oracle/Makefile into oracle/_ref/libref.so
// from the reference's sources where they lie (never copied here); only in a
// container that has the reference tree.
//
// Reference code reached from here (nothing else of the reference is needed):
//   DGBasisEvaluator<P>::eval_all / grad_all      include/DG/DG_Basis/DG_Basis_Evaluator.h:14-31
//   AutoQuadSelector<P, GaussLegendreTet/Tri::Auto>  include/GaussQuadrature/*.h
//   EulerPhysicalFlux<7,5>, <5,3>                  src/DG/DG_Flux/EulerPhysicalFlux.cpp
//   DenseMatrix<3,3>::inverse                      include/Matrix/DenseMatrix.h:521-538
//   OrthHexMesh + GeneralMesh::split_hex5/6_scan   include/Mesh/OrthHexMesh.h, src/Mesh/GeneralMesh*.cpp
// ExplicitConvection.cpp itself is not built: its mesh header includes
// Mesh/MeshGenerator.h, which needs CGAL and Boost (absent).
#include <cstring>

#include "DG/DG_Basis/DG_Basis.h"
#include "DG/DG_Flux/EulerPhysicalFlux.h"
#include "Mesh/OrthHexMesh.h"

namespace {
template <uInt P>
void basis_p(double x, double y, double z, double* val, double* grad) {
    const auto v = DGBasisEvaluator<P>::eval_all(x, y, z);
    const auto g = DGBasisEvaluator<P>::grad_all(x, y, z);
    for (uInt k = 0; k < DGBasisEvaluator<P>::NumBasis; ++k) {
        if (val) val[k] = v[k];
        if (grad)
            for (int d = 0; d < 3; ++d) grad[3 * k + d] = g[k][d];
    }
}
template <uInt P>
int quad_p(int face, double* pts, double* w) {
    if (face) {
        using R = typename AutoQuadSelector<P, GaussLegendreTri::Auto>::type;
        for (int i = 0; i < R::num_points; ++i) {
            if (pts) pts[2 * i] = R::points[i][0], pts[2 * i + 1] = R::points[i][1];
            if (w) w[i] = R::weights[i];
        }
        return R::num_points;
    }
    using R = typename AutoQuadSelector<P, GaussLegendreTet::Auto>::type;
    for (int i = 0; i < R::num_points; ++i) {
        if (pts)
            for (int d = 0; d < 3; ++d) pts[3 * i + d] = R::points[i][d];
        if (w) w[i] = R::weights[i];
    }
    return R::num_points;
}
template <typename Flux>
void flux_t(const double* U, double* F) {
    DenseMatrix<5, 1> u;
    for (int v = 0; v < 5; ++v) u[v] = U[v];
    const auto f = Flux::computeFlux(u);
    for (int i = 0; i < 15; ++i) F[i] = f[i];
}
template <typename Flux>
double wave_t(const double* UL, const double* UR) {
    DenseMatrix<5, 1> a, b;
    for (int v = 0; v < 5; ++v) a[v] = UL[v], b[v] = UR[v];
    return Flux::computeWaveSpeed(a, b);
}

GeneralMesh g_mesh;  // last mesh built by ref_box_mesh
}  // namespace

extern "C" {

int ref_basis(int order, double x, double y, double z, double* val, double* grad) {
    switch (order) {
        case 0: basis_p<0>(x, y, z, val, grad); return 1;
        case 1: basis_p<1>(x, y, z, val, grad); return 4;
        case 2: basis_p<2>(x, y, z, val, grad); return 10;
        case 3: basis_p<3>(x, y, z, val, grad); return 20;
        case 4: basis_p<4>(x, y, z, val, grad); return 35;
        case 5: basis_p<5>(x, y, z, val, grad); return 56;
        default: return 0;
    }
}

int ref_quadrature(int order, int face, double* pts, double* w) {
    switch (order) {
        case 0: return quad_p<0>(face, pts, w);
        case 1: return quad_p<1>(face, pts, w);
        case 2: return quad_p<2>(face, pts, w);
        case 3: return quad_p<3>(face, pts, w);
        case 4: return quad_p<4>(face, pts, w);
        case 5: return quad_p<5>(face, pts, w);
        default: return 0;
    }
}

// gamma_kind 0: EulerPhysicalFlux<7,5> (AirFluxC), 1: <5,3> (MonatomicFluxC)
void ref_flux(int gamma_kind, const double* U, double* F) {
    if (gamma_kind)
        flux_t<MonatomicFluxC>(U, F);
    else
        flux_t<AirFluxC>(U, F);
}
double ref_wave_speed(int gamma_kind, const double* UL, const double* UR) {
    return gamma_kind ? wave_t<MonatomicFluxC>(UL, UR) : wave_t<AirFluxC>(UL, UR);
}

void ref_inverse3(const double* A, double* out) {
    DenseMatrix<3, 3> m;
    for (int i = 0; i < 9; ++i) m[i] = A[i];
    const auto r = m.inverse();
    for (int i = 0; i < 9; ++i) out[i] = r[i];
}

// Build OrthHexMesh(lb, ub, n) + split_hex5_scan/split_hex6_scan + rebuild + validate,
// keep it, return the counts; ref_box_mesh_export copies it out.
int ref_box_mesh(const double* lb, const double* ub, const int64_t* n, int split, int64_t* np, int64_t* nc,
                 int64_t* nf) {
    try {
        g_mesh = OrthHexMesh({lb[0], lb[1], lb[2]}, {ub[0], ub[1], ub[2]}, {(uInt)n[0], (uInt)n[1], (uInt)n[2]});
        if (split == 5)
            g_mesh.split_hex5_scan();
        else
            g_mesh.split_hex6_scan();
        g_mesh.rebuild_cell_topology();
        g_mesh.validate_mesh();
    } catch (...) {
        return -1;
    }
    int64_t ntet = 0, ntri = 0;
    for (const auto& c : g_mesh.m_cells) ntet += std::holds_alternative<Tetrahedron>(c);
    for (const auto& f : g_mesh.m_faces) ntri += std::holds_alternative<TriangleFace>(f);
    if (ntet != (int64_t)g_mesh.m_cells.size() || ntri != (int64_t)g_mesh.m_faces.size()) return -2;
    *np = g_mesh.m_points.size();
    *nc = ntet;
    *nf = ntri;
    return 0;
}

void ref_box_mesh_export(double* pts, int64_t* cells, int64_t* fnodes, int64_t* fcells) {
    for (size_t p = 0; p < g_mesh.m_points.size(); ++p)
        for (int d = 0; d < 3; ++d) pts[3 * p + d] = g_mesh.m_points[p][d];
    for (size_t c = 0; c < g_mesh.m_cells.size(); ++c) {
        const auto& t = std::get<Tetrahedron>(g_mesh.m_cells[c]);
        for (int a = 0; a < 4; ++a) cells[4 * c + a] = (int64_t)t.nodes[a];
    }
    for (size_t f = 0; f < g_mesh.m_faces.size(); ++f) {
        const auto& t = std::get<TriangleFace>(g_mesh.m_faces[f]);
        for (int k = 0; k < 3; ++k) fnodes[3 * f + k] = (int64_t)t.nodes[k];
        for (int s = 0; s < 2; ++s) fcells[2 * f + s] = t.neighbor_cells[s] == uInt(-1) ? -1 : (int64_t)t.neighbor_cells[s];
    }
}

}  // extern "C"
