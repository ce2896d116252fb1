// Physical flux of the 3D compressible Euler equations, host side.
// Same class template, member names and arithmetic as the reference's
// include/DG/DG_Flux/EulerPhysicalFlux.h:6-29 and
// src/DG/DG_Flux/EulerPhysicalFlux.cpp:7-89:
//   rho = max(U0, 1e-16); u = U1/rho, v = U2/rho, w = U3/rho, e = U4/rho
//   p   = (gamma-1) * (rho*e - 0.5*rho*(u^2+v^2+w^2))
//   F   = [rho u, rho v, rho w; rho u u + p, rho v u, rho w u; ...;
//          u(rho e + p), v(rho e + p), w(rho e + p)]   (5x3, row = variable)
//   wave speed = max over L,R of sqrt(u^2+v^2+w^2) + sqrt(gamma*max(p/rho, 1e-16))
// The GPU kernels restate these formulas in device code (fve_dg_amd/csrc).
#pragma once
#include "Matrix/DenseMatrix.h"

template <uInt GammaNumerator = 7, uInt GammaDenominator = 5>
class EulerPhysicalFlux {
public:
    static constexpr Scalar gamma = static_cast<Scalar>(GammaNumerator) / static_cast<Scalar>(GammaDenominator);
    static constexpr Scalar epslion = 1e-16;

    static DenseMatrix<5, 3> computeFlux(const DenseMatrix<5, 1>& U) {
        const Scalar rho = positive(U[0]);
        const Scalar u = U[1] / rho, v = U[2] / rho, w = U[3] / rho, e = U[4] / rho;
        const Scalar p = computePressure(U);
        return {rho * u, rho * v, rho * w,
                rho * u * u + p, rho * v * u, rho * w * u,
                rho * u * v, rho * v * v + p, rho * w * v,
                rho * u * w, rho * v * w, rho * w * w + p,
                u * (rho * e + p), v * (rho * e + p), w * (rho * e + p)};
    }
    static Scalar computeWaveSpeed(const DenseMatrix<5, 1>& UL, const DenseMatrix<5, 1>& UR) {
        return std::max(waveSpeedMagnitude(UL), waveSpeedMagnitude(UR));
    }
    static Scalar computePressure(const DenseMatrix<5, 1>& U) {
        const Scalar rho = positive(U[0]);
        const Scalar u = U[1] / rho, v = U[2] / rho, w = U[3] / rho, e = U[4] / rho;
        return (gamma - 1) * (rho * e - 0.5 * rho * (u * u + v * v + w * w));
    }

private:
    static Scalar positive(Scalar val) { return std::max(val, epslion); }
    static Scalar soundSpeed(const DenseMatrix<5, 1>& U) {
        return std::sqrt(gamma * positive(computePressure(U) / positive(U[0])));
    }
    static Scalar waveSpeedMagnitude(const DenseMatrix<5, 1>& U) {
        const Scalar rho = positive(U[0]);
        const Scalar u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
        return std::sqrt(u * u + v * v + w * w) + soundSpeed(U);
    }
};

using MonatomicFluxC = EulerPhysicalFlux<5, 3>;  // gamma = 5/3
using AirFluxC = EulerPhysicalFlux<7, 5>;        // gamma = 7/5
