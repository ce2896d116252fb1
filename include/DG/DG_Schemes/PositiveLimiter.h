// Bound-preserving then positivity-preserving limiter, on the GPU.
//
// Same class template, constructor and method names as the reference's
// include/DG/DG_Schemes/PositiveLimiter.h:12-261, so that driver code such as
//
//     PositiveLimiter<Basis::OrderBasis, QuadC, true> positivelimiter(cmesh, param_gamma);
//     positivelimiter.constructMinMax(ddx, dx);
//     positivelimiter.apply(ddx, dx);
//
// (examples/Eular2D/Sedov/Sedov.cpp:92,188-189) compiles unchanged.  What it computes:
//   constructMinMax(cur, prev)  per-cell bounds from `prev`: the extremes of the values at
//                               the rule's points of the cell and its face neighbours, or of
//                               their averages when OnlyNeigbAvg; the lower bound keeps the
//                               reference's rule min(max_U, min(val1, val2)) (:98), i.e. the
//                               minimum at the rule's last point;
//   apply_1(cur, prev)          per variable, the higher modes scaled so that the point
//                               values fit the bounds (:135-170);
//   apply_2(cur, prev)          the density's higher modes scaled so that rho >= 1e-16 at the
//                               points, then all higher modes so that p >= 0, the pressure
//                               scaling found by 100 bisection steps per point (:171-259);
//   apply(cur, prev)            apply_1 then apply_2.
// Cell averages (mode 0) are never changed.  Every call runs on the GPU through the C ABI
// of fvdg.h (fvdg_limiter_bounds / fvdg_limiter_apply); `cur` is limited in place.
//
// Differences from the reference: QuadType must be the order's automatic tetrahedron rule
// (the one its drivers pass); apply / apply_1 before any constructMinMax throw instead of
// limiting against zero bounds.  For a limiter inside the time stepper see
// ExplicitSSPRK3::set_limiter.
#pragma once
#include <memory>
#include <type_traits>

#include "DG/DG_Basis/DG_Basis.h"
#include "DG/DG_Schemes/DeviceSession.h"
#include "GaussQuadrature/GaussLegendre.h"
#include "Matrix/LongVector.h"
#include "Mesh/Mesh.h"

template <uInt Order, typename QuadType = typename AutoQuadSelector<Order, GaussLegendreTet::Auto>::type,
          bool OnlyNeigbAvg = false>
class PositiveLimiter {
private:
    using Basis = DGBasisEvaluator<Order>;
    static_assert(Order <= 5, "orders 0..5 are available");
    static_assert(std::is_same_v<QuadType, typename AutoQuadSelector<Order, GaussLegendreTet::Auto>::type>,
                  "the GPU limiter samples the order's automatic tetrahedron rule");
    static constexpr uInt NV = 5 * Basis::NumBasis;

public:
    PositiveLimiter(const ComputingMesh& mesh, Scalar gamma = 1.4) : mesh_(mesh), gamma_(gamma) {}

    void apply(LongVector<NV>& current_coeffs, const LongVector<NV>& previous_coeffs) {
        run(current_coeffs, previous_coeffs, FVDG_LIMIT_BOUNDS | FVDG_LIMIT_POSITIVITY);
    }
    void apply_1(LongVector<NV>& current_coeffs, const LongVector<NV>& previous_coeffs) {
        run(current_coeffs, previous_coeffs, FVDG_LIMIT_BOUNDS);
    }
    void apply_2(LongVector<NV>& current_coeffs, const LongVector<NV>& previous_coeffs) {
        run(current_coeffs, previous_coeffs, FVDG_LIMIT_POSITIVITY);
    }
    void constructMinMax(LongVector<NV>& current_coeffs, const LongVector<NV>& previous_coeffs) {
        (void)current_coeffs;
        check_size(previous_coeffs);
        session().limiter_bounds(previous_coeffs.data(), OnlyNeigbAvg);
    }

private:
    void run(LongVector<NV>& cur, const LongVector<NV>& prev, int parts) {
        (void)prev;
        check_size(cur);
        session().limiter_apply(cur.data(), parts);
    }
    void check_size(const LongVector<NV>& U) const {
        if (U.size() != mesh_.m_cells.size()) throw std::runtime_error("PositiveLimiter: state size != cell count");
    }
    fvdg::Session& session() {
        if (!s_) s_ = std::make_shared<fvdg::Session>(fvdg::default_device());
        s_->bind(mesh_, (int)Order, FVDG_FLUX_LF, gamma_, FVDG_PREC_FP64);
        return *s_;
    }

    const ComputingMesh& mesh_;
    Scalar gamma_ = 1.4;
    std::shared_ptr<fvdg::Session> s_;
};
