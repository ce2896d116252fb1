// SSP-RK3 time stepper for the explicit DG Euler scheme, state resident on the GPU.
//
// The reference advances M dU/dt = -R(U) in its example drivers with
// x += dt * (-1/b * convection.eval(mesh, x, t)) stages (RK2 midpoint in
// examples/Vortex/123/VortexBatch.cpp:183-208), b being the diagonal mass
// b[e][5k+v] = |J_e| sum_q w_q phi_k(xi_q)^2 (VortexBatch.cpp:153-166).  This driver
// runs the three-stage strong-stability-preserving scheme of Shu & Osher (1988),
//     U1 = U + dt L(U),  U2 = 3/4 U + 1/4 (U1 + dt L(U1)),  U+ = 1/3 U + 2/3 (U2 + dt L(U2)),
// with L = -M^{-1} R, on the device: each stage is the volume kernel plus the face
// kernels with the inverse mass and the stage combination fused into the last one.
// Dirichlet states are re-evaluated at the stage times t, t + dt and t + dt/2 when the
// mesh has Dirichlet faces; otherwise whole steps run without host round trips.
//
//     ExplicitSSPRK3<3> rk(cmesh);
//     rk.set_solution(U0);
//     for (...) { Scalar dt = rk.stable_dt(0.3); rk.step(t, dt); t += dt; }
//     LongVector<5*N> U = rk.solution();
#pragma once
#include "DG/DG_Schemes/ExplicitConvection.h"

template <uInt Order = 3, typename Flux = AirFluxC>
class ExplicitSSPRK3 {
public:
    using Convection = ExplicitConvection<Order, Flux>;
    static constexpr uInt N = Convection::N;

    // `mesh` must outlive the stepper.
    explicit ExplicitSSPRK3(const ComputingMesh& mesh) : mesh_(mesh) {}

    // the residual object (Riemann solver, precision, boundary function); changing the Riemann
    // solver or the precision after set_solution keeps the uploaded state
    Convection& convection() { return conv_; }

    void set_solution(const LongVector<5 * N>& U) {
        if (U.size() != mesh_.m_cells.size()) throw std::runtime_error("ExplicitSSPRK3: state size != cell count");
        session().upload(U.data());
    }
    LongVector<5 * N> solution() {
        LongVector<5 * N> U(mesh_.m_cells.size());
        session().download(U.data());
        return U;
    }
    // diagonal mass in the state layout (the reference drivers' b)
    LongVector<5 * N> mass() {
        LongVector<5 * N> M(mesh_.m_cells.size());
        session().mass(M.data());
        return M;
    }
    // limit every stage result with PositiveLimiter's steps (FVDG_LIMIT_BOUNDS: apply_1,
    // FVDG_LIMIT_POSITIVITY: apply_2; 0: off), bounds from the stage's input state
    void set_limiter(int parts, bool neighbour_averages = false) { session().set_stage_limiter(parts, neighbour_averages); }
    // cfl * h_min / ((2p+1) lambda_max) for the current device state
    Scalar stable_dt(Scalar cfl = 0.3) { return session().stable_dt(cfl); }

    // one step from time t
    void step(Scalar t, Scalar dt) {
        fvdg::Session& s = session();
        if (s.dirichlet_faces().empty()) {
            s.steps(dt, 1);
            return;
        }
        const Scalar stage_time[3] = {t, t + dt, t + 0.5 * dt};
        for (int k = 0; k < 3; ++k) {
            conv_.update_boundary_states(mesh_, stage_time[k]);
            s.stage(dt, k);
        }
    }
    // n steps of size dt from time t
    void advance(Scalar t, Scalar dt, uInt n) {
        fvdg::Session& s = session();
        if (s.dirichlet_faces().empty()) {
            s.steps(dt, (int)n);
            return;
        }
        for (uInt i = 0; i < n; ++i) step(t + i * dt, dt);
    }
    // sum over cells of the cell integrals of the conserved variables
    std::array<Scalar, 5> conservation_sums() {
        std::array<Scalar, 5> out{};
        session().conservation_sums(out.data());
        return out;
    }

private:
    fvdg::Session& session() { return conv_.session(mesh_); }
    const ComputingMesh& mesh_;
    Convection conv_;
};
