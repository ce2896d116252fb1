// gfx950 kernel of the flow diagnostics of the device state: integrals over the mesh at the
// volume quadrature points (weights w_q |det A_e|, the reference's err_integral measure,
// examples/Vortex/VortexBatch3.cpp:349-366) of
//   1, rho, kinetic energy rho |u|^2 / 2, enstrophy rho |curl u|^2 / 2, total energy rho E,
// and the minimum density, minimum pressure and maximum Mach number over the points.
// Velocity gradients come from the conserved ones, grad u_i = (grad(rho u_i) - u_i grad rho) / rho,
// with grad_x = A^{-1} grad_xi.  The kinetic-energy and enstrophy series are the standard
// diagnostics of the Taylor-Green vortex (normalised by rho_0 |Omega| by the caller).
//
// Work split as in the limiter: LPE lanes per element (one quadrature point each), the block's
// coefficients staged in LDS, one partial record of 8 doubles per block (summed on the host,
// in block order, so the result does not depend on scheduling).
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>
#include <stdint.h>

namespace fvdg_diag {

constexpr int kNumDiag = 8;  // volume, mass, kinetic energy, enstrophy, total energy, min rho, min p, max Mach

template <int LPE>
__global__ __launch_bounds__(256) void k_flow_diagnostics(const double* __restrict__ U, const double* __restrict__ cgeo,
                                                          const double* __restrict__ V, const double* __restrict__ G,
                                                          const double* __restrict__ wq, int64_t E, int N, int Q,
                                                          double gamma, double* __restrict__ part) {
    constexpr int EB = 256 / LPE;
    extern __shared__ double lds[];
    const int NV = 5 * N;
    double* sC = lds;              // EB * NV
    double* sR = lds + EB * NV;    // kNumDiag * 256
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int nel = (int)(E - e0 < EB ? E - e0 : EB);
    for (int i = threadIdx.x; i < nel * NV; i += 256) sC[i] = U[e0 * NV + i];
    __syncthreads();
    const int el = threadIdx.x / LPE, q = threadIdx.x % LPE;
    double r[kNumDiag] = {0, 0, 0, 0, 0, DBL_MAX, DBL_MAX, -DBL_MAX};
    if (el < nel && q < Q) {
        const double* c = sC + el * NV;
        const double* geo = cgeo + (e0 + el) * 10;
        double u[5] = {0, 0, 0, 0, 0}, gr[3][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};
        for (int l = 0; l < N; ++l) {
            const double phi = V[q * N + l];
            const double g0 = G[(q * 3) * N + l], g1 = G[(q * 3 + 1) * N + l], g2 = G[(q * 3 + 2) * N + l];
            for (int k = 0; k < 5; ++k) {
                const double ck = c[k + 5 * l];
                u[k] += phi * ck;
                gr[0][k] += g0 * ck;
                gr[1][k] += g1 * ck;
                gr[2][k] += g2 * ck;
            }
        }
        double gx[3][5];  // physical gradients: A^{-1} (row-major) times the reference gradients
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < 5; ++k) gx[d][k] = geo[3 * d] * gr[0][k] + geo[3 * d + 1] * gr[1][k] + geo[3 * d + 2] * gr[2][k];
        const double w = wq[q] * geo[9];
        const double rho = u[0], irho = 1.0 / rho;
        const double vel[3] = {u[1] * irho, u[2] * irho, u[3] * irho};
        double du[3][3];  // du[i][d] = d u_i / d x_d
        for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d) du[i][d] = (gx[d][1 + i] - vel[i] * gx[d][0]) * irho;
        const double om[3] = {du[2][1] - du[1][2], du[0][2] - du[2][0], du[1][0] - du[0][1]};
        const double v2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
        const double p = (gamma - 1.0) * (u[4] - 0.5 * rho * v2);
        r[0] = w;
        r[1] = w * rho;
        r[2] = w * 0.5 * rho * v2;
        r[3] = w * 0.5 * rho * (om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        r[4] = w * u[4];
        r[5] = rho;
        r[6] = p;
        r[7] = sqrt(v2) / sqrt(fmax(gamma * p * irho, 1e-300));
    }
    for (int i = 0; i < kNumDiag; ++i) sR[i * 256 + threadIdx.x] = r[i];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const int a = threadIdx.x, b = threadIdx.x + s;
            for (int i = 0; i < 5; ++i) sR[i * 256 + a] += sR[i * 256 + b];
            sR[5 * 256 + a] = fmin(sR[5 * 256 + a], sR[5 * 256 + b]);
            sR[6 * 256 + a] = fmin(sR[6 * 256 + a], sR[6 * 256 + b]);
            sR[7 * 256 + a] = fmax(sR[7 * 256 + a], sR[7 * 256 + b]);
        }
        __syncthreads();
    }
    if (threadIdx.x < kNumDiag) part[(int64_t)blockIdx.x * kNumDiag + threadIdx.x] = sR[threadIdx.x * 256];
}

}  // namespace fvdg_diag
