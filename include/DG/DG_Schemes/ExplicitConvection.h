// Explicit DG residual of the 3D compressible Euler equations on tetrahedra, on the GPU.
//
// Same class template, method names and signatures as the reference's
// include/DG/DG_Schemes/ExplicitConvection.h:9-45, so that callers written against
// it compile unchanged:
//
//     ExplicitConvection<3> convection;
//     LongVector<5*N> R = convection.eval(cmesh, U, t);   // R = cells + internals + boundarys
//
// with the reference's sign convention (src/DG/DG_Schemes/ExplicitConvection.cpp):
//   eval_cells      R[e][5j+v] = - sum_q w_q |J_e| (F(U_e(xi_q)) . A_e^{-1} grad phi_j(xi_q))_v
//   eval_internals  R[L][5j+v] += sum_g w_g 2A_f Fhat_v phi_j(xi_L(g)),  R[R][..] -= ... phi_j(xi_R(g))
//   eval_boundarys  the same on boundary faces with a ghost state: Dirichlet 2 U_D - U_in with
//                   U_D from the application's rho_Xi .. rhoe_Xi hooks (base/exact.h),
//                   Pseudo3DZ the mirror of rho w; other boundary types contribute nothing.
// The semi-discrete system is M dU/dt = -R with the diagonal mass of ExplicitSSPRK3.h.
//
// Every call runs on the GPU through the C ABI of fvdg.h: the object owns a device
// context (created on first use on device FVDG_DEVICE, default 0), binds the mesh
// the first time it sees it (and again when a different mesh object or a changed
// boundary-type array is passed; call rebind() after changing a mesh in place),
// uploads the state, and downloads the result.  Copies of the object share the context.
//
// Differences from the reference, all opt-in: set_riemann_solver(FVDG_FLUX_ROE)
// selects the Roe flux (the reference has Lax-Friedrichs only); set_precision()
// selects the contraction precision; set_boundary_function() replaces the link-time
// hooks.  Only the reference's automatic quadrature rules are available on the GPU.
#pragma once
#include <functional>
#include <type_traits>

#include "DG/DG_Basis/DG_Basis.h"
#include "DG/DG_Flux/EulerPhysicalFlux.h"
#include "DG/DG_Schemes/DeviceSession.h"
#include "GaussQuadrature/GaussLegendre.h"
#include "Matrix/LongVector.h"
#include "Mesh/Mesh.h"
#ifndef FVDG_NO_EXACT_HOOKS
#include "base/exact.h"
#endif

// Ghost-state source of Dirichlet faces: conserved state U[5] at reference point Xi
// of `cell` at time t.
using BoundaryStateFunction = std::function<void(const CompTetrahedron& cell, const vector3f& Xi, Scalar t, Scalar U[5])>;

template <uInt Order = 3, typename Flux = AirFluxC, typename GaussQuadCell = GaussLegendreTet::Auto,
          typename GaussQuadFace = GaussLegendreTri::Auto>
class ExplicitConvection {
private:
    using Basis = DGBasisEvaluator<Order>;
    using QuadC = typename std::conditional_t<std::is_same_v<GaussQuadCell, GaussLegendreTet::Auto>,
                                              typename AutoQuadSelector<Order, GaussLegendreTet::Auto>::type,
                                              GaussQuadCell>;
    using QuadF = typename std::conditional_t<std::is_same_v<GaussQuadFace, GaussLegendreTri::Auto>,
                                              typename AutoQuadSelector<Order, GaussLegendreTri::Auto>::type,
                                              GaussQuadFace>;
    static_assert(Order <= 5, "orders 0..5 are available");
    static_assert(std::is_same_v<QuadC, typename AutoQuadSelector<Order, GaussLegendreTet::Auto>::type> &&
                      std::is_same_v<QuadF, typename AutoQuadSelector<Order, GaussLegendreTri::Auto>::type>,
                  "the GPU residual uses the automatic quadrature rules of each order");

public:
    static constexpr uInt N = Basis::NumBasis;

    LongVector<5 * N> eval(const ComputingMesh& mesh, const LongVector<5 * N>& old_solution,
                           const Scalar curr_time = 0.0) {
        return run(mesh, old_solution, curr_time, FVDG_PART_ALL);
    }
    LongVector<5 * N> eval_cells(const ComputingMesh& mesh, const LongVector<5 * N>& old_solution,
                                 const Scalar curr_time = 0.0) {
        return run(mesh, old_solution, curr_time, FVDG_PART_CELLS);
    }
    LongVector<5 * N> eval_internals(const ComputingMesh& mesh, const LongVector<5 * N>& old_solution,
                                     const Scalar curr_time = 0.0) {
        return run(mesh, old_solution, curr_time, FVDG_PART_INTERNALS);
    }
    LongVector<5 * N> eval_boundarys(const ComputingMesh& mesh, const LongVector<5 * N>& old_solution,
                                     const Scalar curr_time = 0.0) {
        return run(mesh, old_solution, curr_time, FVDG_PART_BOUNDARYS);
    }

    // reference-element point of face point uv seen from side `side` (0 or 1):
    // nc[0](1-u-v) + nc[1]u + nc[2]v (ExplicitConvection.cpp:129-134)
    vector3f transform_to_cell(const CompTriangleFace& face, const vector2f& uv, uInt side) const {
        const auto& nc = face.m_natural_coords[side];
        vector3f xi;
        for (int d = 0; d < 3; ++d) xi[d] = nc[0][d] * (1 - uv[0] - uv[1]) + nc[1][d] * uv[0] + nc[2][d] * uv[1];
        return xi;
    }
    // max over L,R of |u| + c (EulerPhysicalFlux::computeWaveSpeed)
    Scalar compute_max_wave_speed(const DenseMatrix<5, 1>& UL, const DenseMatrix<5, 1>& UR) const {
        return Flux::computeWaveSpeed(UL, UR);
    }

    // ---------------- additions
    void set_riemann_solver(int flux) { flux_ = flux; }
    void set_precision(int precision) { precision_ = precision; }
    void set_boundary_function(BoundaryStateFunction f) { bfun_ = std::move(f); }
    void rebind() {
        if (s_) s_->rebind();
    }
    // the device context (created and bound to `mesh` on demand)
    fvdg::Session& session(const ComputingMesh& mesh) {
        if (!s_) s_ = std::make_shared<fvdg::Session>(fvdg::default_device());
        s_->bind(mesh, (int)Order, flux_, Flux::gamma, precision_);
        return *s_;
    }
    // Evaluate the Dirichlet states of the bound mesh at time t and send them to the device.
    void update_boundary_states(const ComputingMesh& mesh, Scalar t) {
        fvdg::Session& s = session(mesh);
        const auto& faces = s.dirichlet_faces();
        if (faces.empty()) return;
        constexpr int Qf = QuadF::num_points;
        std::vector<double> UD(faces.size() * Qf * 5);
        for (size_t i = 0; i < faces.size(); ++i) {
            const auto& face = mesh.m_faces[faces[i]];
            const auto& cell = mesh.m_cells[face.m_neighbor_cells[0]];
            for (int g = 0; g < Qf; ++g) {
                const vector3f xi = transform_to_cell(face, {QuadF::points[g][0], QuadF::points[g][1]}, 0);
                boundary_state(cell, xi, t, &UD[(i * Qf + g) * 5]);
            }
        }
        s.set_dirichlet_states(UD.data());
    }

private:
    LongVector<5 * N> run(const ComputingMesh& mesh, const LongVector<5 * N>& U, Scalar t, int parts) {
        if (U.size() != mesh.m_cells.size()) throw std::runtime_error("ExplicitConvection: state size != cell count");
        fvdg::Session& s = session(mesh);
        if (parts & FVDG_PART_BOUNDARYS) update_boundary_states(mesh, t);
        s.upload(U.data());
        LongVector<5 * N> R(U.size());
        s.residual(t, parts, R.data());
        return R;
    }

    void boundary_state(const CompTetrahedron& cell, const vector3f& xi, Scalar t, Scalar* out) const {
        if (bfun_) {
            bfun_(cell, xi, t, out);
            return;
        }
#ifndef FVDG_NO_EXACT_HOOKS
        out[0] = rho_Xi(cell, xi, t);
        out[1] = rhou_Xi(cell, xi, t);
        out[2] = rhov_Xi(cell, xi, t);
        out[3] = rhow_Xi(cell, xi, t);
        out[4] = rhoe_Xi(cell, xi, t);
#else
        (void)cell;
        (void)xi;
        (void)t;
        (void)out;
        throw std::runtime_error("ExplicitConvection: Dirichlet faces need set_boundary_function()");
#endif
    }

    std::shared_ptr<fvdg::Session> s_;
    int flux_ = FVDG_FLUX_LF;
    int precision_ = FVDG_PREC_FP64;
    BoundaryStateFunction bfun_;
};
