// gfx950 kernels of the explicit DG Euler residual and SSP-RK3 update (FP64).
//
// Residual R(U) = R_cells + R_internals + R_boundarys, as in
// src/DG/DG_Schemes/ExplicitConvection.cpp:
//   R_cells  [e,j,v] = - sum_q w_q |J_e| (F(U_e(xi_q)) . (A_e^{-1} grad phi_j(xi_q)))_v   (:144-198)
//   R_faces  [L,j,v] += sum_g w_g 2A_f Fhat_v(U_L,U_R,n) phi_j(xi_L(g))                 (:201-272)
//            [R,j,v] -= sum_g w_g 2A_f Fhat_v(U_L,U_R,n) phi_j(xi_R(g))
//   boundary faces use a ghost state: Dirichlet 2 U_D - U_in, Pseudo3DZ (1,1,1,-1,1) U_in  (:21-128)
// The kernels in this file are the plain FP64 path ("simple"): pointwise fluxes go
// through scratch buffers (Fq: E*Q*15, Fh: F*Qf*5) and every contraction is a
// thread-private loop, so each output has one fixed summation order.  They serve
// as the GPU cross-check of the tuned kernels (fvdg_kernels_fast.hpp) at sizes the
// CPU oracle cannot reach.
// FVDG_DEVICE_HELPERS_ONLY (defined by other translation units): the device helpers without the
// non-template kernels, which are defined in the solver's translation unit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fvdg.h"

struct KernelArgs {
    int64_t E = 0, F = 0;
    int N = 0, Q = 0, Qf = 0;
    double gamma = 1.4;
    int flux = 0, parts = 7;
    const double* U = nullptr;
    double* R = nullptr;
    const double* cgeo = nullptr;     // E*10: A^{-1} row-major, |det A|
    const int32_t* cfaces = nullptr;  // E*4: face*2 + side
    const int32_t* fcell = nullptr;   // F*2
    const uint8_t* fcombo = nullptr;  // F*2
    const double* fgeo = nullptr;     // F*4: n, 2*area
    const uint8_t* fbc = nullptr;     // F
    const int32_t* fdir = nullptr;    // F
    const double* UD = nullptr;       // Dirichlet states
    const double* V = nullptr;        // Q*N
    const double* G = nullptr;        // Q*3*N
    const double* wq = nullptr;       // Q
    const double* Phi = nullptr;      // 24*Qf*N
    const double* wf = nullptr;       // Qf
    double* Fq = nullptr;             // E*Q*15
    double* Fh = nullptr;             // F*Qf*5
};

// element e, coefficient kv (= 5k+v): element-major, the reference's LongVector layout
__device__ __forceinline__ int64_t cidx(int64_t e, int kv, int NV) { return e * NV + kv; }

// ------------------------------------------------------------------ Euler algebra
// The formulas of src/DG/DG_Flux/EulerPhysicalFlux.cpp:7-18,46-55,70-89
// (rho = max(U0, 1e-16), u = U1/rho, ..., p = (gamma-1)(rho e - rho |u|^2 / 2),
// c = sqrt(gamma max(p / max(U0,1e-16), 1e-16))), evaluated once per state with
// one reciprocal 1/rho instead of the reference's repeated divisions (each
// quotient is then within about one rounding of the reference's).
__device__ __forceinline__ double pos_eps(double x) { return fmax(x, 1e-16); }

// Reciprocal and square root of a positive normal FP64 value (every caller clamps its argument
// to >= 1e-16): the hardware approximations v_rcp_f64 / v_rsq_f64 (about 2^-23 relative) refined
// by two Newton-Raphson steps / one Goldschmidt step and one residual correction, to within about
// one ulp.  The correctly rounded IEEE sequences add argument scaling and fix-ups and cost about
// twice the instructions; the face-flux kernel is bound by VALU issue.  A NaN argument gives NaN.
__device__ __forceinline__ double rcp_pos(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}
__device__ __forceinline__ double sqrt_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    double s = x * y, h = 0.5 * y;
    const double r = fma(-h, s, 0.5);
    s = fma(s, r, s);
    h = fma(h, r, h);
    const double d = fma(-s, s, x);
    return fma(d, h, s);
}
// x >= 0 (a sum of squares): 0 stays 0, NaN stays NaN
__device__ __forceinline__ double sqrt_nonneg(double x) { return x > 0.0 ? sqrt_pos(x) : x; }

// maximum that keeps a NaN operand (fmax drops it)
__device__ __forceinline__ double max_keep_nan(double a, double b) { return (a != a || b != b) ? a + b : fmax(a, b); }

// stage-output check of the SSP-RK update: a non-finite coefficient, or a non-positive mean
// density (kv == 0 is mode 0 of rho; phi_0 = 1, so that coefficient is the cell mean)
__device__ __forceinline__ bool stage_value_bad(double x, int kv) {
    return !(x - x == 0.0) || (kv == 0 && !(x > 0.0));
}

struct Prim {
    double rho, ri, u, v, w, e, p;  // ri = 1/rho
};

__device__ __forceinline__ Prim euler_prim(const double U[5], double gamma) {
    Prim q;
    q.rho = pos_eps(U[0]);
    const double ri = rcp_pos(q.rho);
    q.ri = ri;
    q.u = U[1] * ri;
    q.v = U[2] * ri;
    q.w = U[3] * ri;
    q.e = U[4] * ri;
    q.p = (gamma - 1) * (q.rho * q.e - 0.5 * q.rho * (q.u * q.u + q.v * q.v + q.w * q.w));
    return q;
}

__device__ __forceinline__ double euler_pressure(const double U[5], double gamma) { return euler_prim(U, gamma).p; }

__device__ __forceinline__ void euler_flux_prim(const Prim& q, double F[5][3]) {
    const double rho = q.rho, u = q.u, v = q.v, w = q.w, p = q.p;
    const double re = rho * q.e + p;
    F[0][0] = rho * u;         F[0][1] = rho * v;         F[0][2] = rho * w;
    F[1][0] = rho * u * u + p; F[1][1] = rho * v * u;     F[1][2] = rho * w * u;
    F[2][0] = rho * u * v;     F[2][1] = rho * v * v + p; F[2][2] = rho * w * v;
    F[3][0] = rho * u * w;     F[3][1] = rho * v * w;     F[3][2] = rho * w * w + p;
    F[4][0] = u * re;          F[4][1] = v * re;          F[4][2] = w * re;
}

__device__ __forceinline__ void euler_flux(const double U[5], double gamma, double F[5][3]) {
    euler_flux_prim(euler_prim(U, gamma), F);
}

// normal flux F(U) n from primitives
__device__ __forceinline__ void euler_flux_n(const Prim& q, const double n[3], double out[5]) {
    const double un = q.u * n[0] + q.v * n[1] + q.w * n[2];
    const double mn = q.rho * un;
    out[0] = mn;
    out[1] = mn * q.u + q.p * n[0];
    out[2] = mn * q.v + q.p * n[1];
    out[3] = mn * q.w + q.p * n[2];
    out[4] = un * (q.rho * q.e + q.p);
}

// |u| + c; c uses p / max(U0, 1e-16) = p / rho
__device__ __forceinline__ double euler_wave_speed_prim(const Prim& q, double gamma) {
    const double c = sqrt_pos(gamma * pos_eps(q.p * q.ri));
    return sqrt_nonneg(q.u * q.u + q.v * q.v + q.w * q.w) + c;
}

__device__ __forceinline__ double euler_wave_speed(const double U[5], double gamma) {
    return euler_wave_speed_prim(euler_prim(U, gamma), gamma);
}

// |u| + c for the time-step reductions: NaN when any conserved value is not finite (the clamps of
// euler_prim would otherwise turn a NaN energy into a finite sound speed)
__device__ __forceinline__ double wave_speed_checked(const double U[5], double gamma) {
    const double s = U[0] + U[1] + U[2] + U[3] + U[4];
    return (s - s == 0.0) ? euler_wave_speed(U, gamma) : __builtin_nan("");
}

// Lax-Friedrichs: 0.5 (F(U_L) n + F(U_R) n + lambda (U_L - U_R)), lambda = max(|u|+c)
__device__ __forceinline__ void flux_lf(const double UL[5], const double UR[5], const double n[3], double gamma,
                                        double out[5]) {
    const Prim L = euler_prim(UL, gamma), R = euler_prim(UR, gamma);
    double fl[5], fr[5];
    euler_flux_n(L, n, fl);
    euler_flux_n(R, n, fr);
    const double lam = fmax(euler_wave_speed_prim(L, gamma), euler_wave_speed_prim(R, gamma));
    for (int v = 0; v < 5; ++v) out[v] = 0.5 * (fl[v] + fr[v] + lam * (UL[v] - UR[v]));
}

// Roe (1981) flux with Harten's entropy fix on the acoustic waves, delta = 0.1 * a_roe:
//   Fhat = 0.5 (F_L n + F_R n) - 0.5 sum_k |lambda_k| alpha_k r_k
// The Roe average of two physical states (rho > 0, p > 0) has a positive sound speed;
// for a non-physical trace (an overshooting high-order polynomial) it may not, and the
// wave strengths dp / a^2 blow up, so such points take the Lax-Friedrichs flux instead.
__device__ __forceinline__ void flux_roe(const double UL[5], const double UR[5], const double n[3], double gamma,
                                         double out[5]) {
    const Prim L = euler_prim(UL, gamma), R = euler_prim(UR, gamma);
    if (!(UL[0] > 0.0 && UR[0] > 0.0 && L.p > 0.0 && R.p > 0.0)) {
        flux_lf(UL, UR, n, gamma, out);
        return;
    }
    double fl[5], fr[5];
    euler_flux_n(L, n, fl);
    euler_flux_n(R, n, fr);
    const double uL[3] = {L.u, L.v, L.w}, uR[3] = {R.u, R.v, R.w};
    const double HL = (UL[4] + L.p) * L.ri, HR = (UR[4] + R.p) * R.ri;
    // Roe averages with the weight ratio k = sqrt(rho_R / rho_L): (x_L + k x_R) / (1 + k) equals
    // (sqrt(rho_L) x_L + sqrt(rho_R) x_R) / (sqrt(rho_L) + sqrt(rho_R)), and sqrt(rho_L rho_R) = k rho_L
    const double k = sqrt_pos(R.rho * L.ri), inv = rcp_pos(1.0 + k);
    double ut[3], q2 = 0, qn = 0;
    for (int d = 0; d < 3; ++d) {
        ut[d] = (uL[d] + k * uR[d]) * inv;
        q2 += ut[d] * ut[d];
        qn += ut[d] * n[d];
    }
    const double Ht = (HL + k * HR) * inv;
    const double a2 = fmax((gamma - 1) * (Ht - 0.5 * q2), 1e-16);
    const double a = sqrt_pos(a2), ia2 = rcp_pos(a2);
    const double rt = k * L.rho;
    const double dr = R.rho - L.rho, dp = R.p - L.p;
    double du[3], dqn = 0;
    for (int d = 0; d < 3; ++d) {
        du[d] = uR[d] - uL[d];
        dqn += du[d] * n[d];
    }
    double dut[3], ut_dut = 0;
    for (int d = 0; d < 3; ++d) {
        dut[d] = du[d] - dqn * n[d];
        ut_dut += ut[d] * dut[d];
    }
    const double a1 = (dp - rt * a * dqn) * (0.5 * ia2);
    const double a5 = (dp + rt * a * dqn) * (0.5 * ia2);
    const double a_2 = dr - dp * ia2;
    const double delta = 0.1 * a, inv_2delta = 5.0 * (a * ia2);  // 1 / (2 delta) = 5 / a = 5 a / a^2
    auto fix = [delta, inv_2delta](double lam) {
        const double al = fabs(lam);
        return al < delta ? (lam * lam + delta * delta) * inv_2delta : al;
    };
    const double l1 = fix(qn - a), l2 = fabs(qn), l5 = fix(qn + a);
    const double w1 = l1 * a1, w2 = l2 * a_2, w5 = l5 * a5, ws = l2 * rt;
    double d[5];
    d[0] = w1 + w2 + w5;
    for (int k = 0; k < 3; ++k) d[1 + k] = w1 * (ut[k] - a * n[k]) + w2 * ut[k] + ws * dut[k] + w5 * (ut[k] + a * n[k]);
    d[4] = w1 * (Ht - a * qn) + w2 * 0.5 * q2 + ws * ut_dut + w5 * (Ht + a * qn);
    for (int v = 0; v < 5; ++v) out[v] = 0.5 * (fl[v] + fr[v]) - 0.5 * d[v];
}

__device__ __forceinline__ void num_flux(int kind, const double UL[5], const double UR[5], const double n[3],
                                         double gamma, double out[5]) {
    if (kind == FVDG_FLUX_ROE)
        flux_roe(UL, UR, n, gamma, out);
    else
        flux_lf(UL, UR, n, gamma, out);
}

// ------------------------------------------------------------------ mixed-precision numerical fluxes
// The mixed path keeps FP64 where the flux value itself is formed -- the primitives (1/rho), the
// pressures and positivity tests, the physical normal fluxes F(U_L) n and F(U_R) n and the jumps
// U_R - U_L -- and evaluates the upwind dissipation, which is proportional to the jumps, in FP32
// (Roe averages, sound speed, eigenvalues, wave strengths; hardware square root and reciprocal).
// Its FP32 rounding is relative to the jump, not to the flux.
__device__ __forceinline__ float sqrt_f32(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float rcp_f32(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ __forceinline__ void flux_lf_mx(const double UL[5], const double UR[5], const double n[3], double gamma,
                                           double out[5]) {
    const Prim L = euler_prim(UL, gamma), R = euler_prim(UR, gamma);
    double fl[5], fr[5];
    euler_flux_n(L, n, fl);
    euler_flux_n(R, n, fr);
    const float g = (float)gamma;
    auto speed = [g](const Prim& q) {
        const float u = (float)q.u, v = (float)q.v, w = (float)q.w;
        const float c = sqrt_f32(g * fmaxf((float)(q.p * q.ri), 1e-16f));
        return sqrt_f32(u * u + v * v + w * w) + c;
    };
    const double lam = (double)fmaxf(speed(L), speed(R));
    for (int v = 0; v < 5; ++v) out[v] = 0.5 * (fl[v] + fr[v] + lam * (UL[v] - UR[v]));
}

__device__ __forceinline__ void flux_roe_mx(const double UL[5], const double UR[5], const double n[3], double gamma,
                                            double out[5]) {
    const Prim L = euler_prim(UL, gamma), R = euler_prim(UR, gamma);
    if (!(UL[0] > 0.0 && UR[0] > 0.0 && L.p > 0.0 && R.p > 0.0)) {
        flux_lf_mx(UL, UR, n, gamma, out);
        return;
    }
    double fl[5], fr[5];
    euler_flux_n(L, n, fl);
    euler_flux_n(R, n, fr);
    const float g = (float)gamma;
    const float nf[3] = {(float)n[0], (float)n[1], (float)n[2]};
    const float uL[3] = {(float)L.u, (float)L.v, (float)L.w}, uR[3] = {(float)R.u, (float)R.v, (float)R.w};
    const float HL = (float)((UL[4] + L.p) * L.ri), HR = (float)((UR[4] + R.p) * R.ri);
    const float sL = sqrt_f32((float)L.rho), sR = sqrt_f32((float)R.rho), inv = rcp_f32(sL + sR);
    float ut[3], q2 = 0, qn = 0;
    for (int d = 0; d < 3; ++d) {
        ut[d] = (sL * uL[d] + sR * uR[d]) * inv;
        q2 += ut[d] * ut[d];
        qn += ut[d] * nf[d];
    }
    const float Ht = (sL * HL + sR * HR) * inv;
    const float a2 = fmaxf((g - 1) * (Ht - 0.5f * q2), 1e-16f);
    const float a = sqrt_f32(a2), ia2 = rcp_f32(a2);
    const float rt = sL * sR;
    const float dr = (float)(R.rho - L.rho), dp = (float)(R.p - L.p);
    const float du[3] = {(float)(R.u - L.u), (float)(R.v - L.v), (float)(R.w - L.w)};
    const float dqn = du[0] * nf[0] + du[1] * nf[1] + du[2] * nf[2];
    float dut[3], ut_dut = 0;
    for (int d = 0; d < 3; ++d) {
        dut[d] = du[d] - dqn * nf[d];
        ut_dut += ut[d] * dut[d];
    }
    const float a1 = (dp - rt * a * dqn) * (0.5f * ia2);
    const float a5 = (dp + rt * a * dqn) * (0.5f * ia2);
    const float a_2 = dr - dp * ia2;
    const float delta = 0.1f * a;
    auto fix = [delta](float lam) {
        const float al = fabsf(lam);
        return al < delta ? (lam * lam + delta * delta) * rcp_f32(2 * delta) : al;
    };
    const float l1 = fix(qn - a), l2 = fabsf(qn), l5 = fix(qn + a);
    const float w1 = l1 * a1, w2 = l2 * a_2, w5 = l5 * a5, ws = l2 * rt;
    float d[5];
    d[0] = w1 + w2 + w5;
    for (int k = 0; k < 3; ++k) d[1 + k] = w1 * (ut[k] - a * nf[k]) + w2 * ut[k] + ws * dut[k] + w5 * (ut[k] + a * nf[k]);
    d[4] = w1 * (Ht - a * qn) + w2 * 0.5f * q2 + ws * ut_dut + w5 * (Ht + a * qn);
    for (int v = 0; v < 5; ++v) out[v] = 0.5 * (fl[v] + fr[v]) - 0.5 * (double)d[v];
}

__device__ __forceinline__ void num_flux_mx(int kind, const double UL[5], const double UR[5], const double n[3],
                                            double gamma, double out[5]) {
    if (kind == FVDG_FLUX_ROE)
        flux_roe_mx(UL, UR, n, gamma, out);
    else
        flux_lf_mx(UL, UR, n, gamma, out);
}

// ------------------------------------------------------------------ volume term
// one thread per (element, volume point): trace of U and its physical flux
#ifndef FVDG_DEVICE_HELPERS_ONLY
__global__ void k_vol_flux(KernelArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.E * a.Q) return;
    const int64_t e = t / a.Q;
    const int q = (int)(t - e * a.Q);
    const int NV = 5 * a.N;
    double U[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < a.N; ++k) {
        const double phi = a.V[q * a.N + k];
        for (int v = 0; v < 5; ++v) U[v] += phi * a.U[cidx(e, 5 * k + v, NV)];
    }
    double F[5][3];
    euler_flux(U, a.gamma, F);
    double* out = a.Fq + t * 15;
    for (int v = 0; v < 5; ++v)
        for (int d = 0; d < 3; ++d) out[v * 3 + d] = F[v][d];
}
#endif

// one thread per (element, test function j): R = -sum_q (F . A^{-1} grad phi_j) w_q |J|
#ifndef FVDG_DEVICE_HELPERS_ONLY
__global__ void k_vol_test(KernelArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.E * a.N) return;
    const int64_t e = t / a.N;
    const int j = (int)(t - e * a.N);
    const int NV = 5 * a.N;
    const double* geo = a.cgeo + e * 10;
    double Ai[9];
    for (int i = 0; i < 9; ++i) Ai[i] = geo[i];
    const double det = geo[9];
    double r[5] = {0, 0, 0, 0, 0};
    for (int q = 0; q < a.Q; ++q) {
        const double g0 = a.G[(q * 3 + 0) * a.N + j], g1 = a.G[(q * 3 + 1) * a.N + j], g2 = a.G[(q * 3 + 2) * a.N + j];
        double gp[3];
        for (int d = 0; d < 3; ++d) gp[d] = Ai[3 * d] * g0 + Ai[3 * d + 1] * g1 + Ai[3 * d + 2] * g2;
        const double jw = a.wq[q] * det;
        const double* F = a.Fq + (e * a.Q + q) * 15;
        for (int v = 0; v < 5; ++v) r[v] -= (F[3 * v] * gp[0] + F[3 * v + 1] * gp[1] + F[3 * v + 2] * gp[2]) * jw;
    }
    for (int v = 0; v < 5; ++v) a.R[cidx(e, 5 * j + v, NV)] = r[v];
}
#endif

// ------------------------------------------------------------------ face term
__device__ __forceinline__ void face_trace(const KernelArgs& a, int64_t cell, int combo, int g, double U[5]) {
    const int NV = 5 * a.N;
    for (int v = 0; v < 5; ++v) U[v] = 0;
    const double* phi = a.Phi + ((int64_t)combo * a.Qf + g) * a.N;
    for (int k = 0; k < a.N; ++k)
        for (int v = 0; v < 5; ++v) U[v] += phi[k] * a.U[cidx(cell, 5 * k + v, NV)];
}

// ghost state of a boundary face point; false when the boundary type contributes nothing
__device__ __forceinline__ bool ghost_state(const KernelArgs& a, int64_t f, int g, const double Uin[5],
                                            double Ug[5]) {
    const int bc = a.fbc[f];
    if (bc == FVDG_BC_DIRICHLET) {
        const double* ud = a.UD + ((int64_t)a.fdir[f] * a.Qf + g) * 5;
        for (int v = 0; v < 5; ++v) Ug[v] = ud[v] + (ud[v] - Uin[v]);
        return true;
    }
    if (bc == FVDG_BC_PSEUDO3DZ) {
        for (int v = 0; v < 5; ++v) Ug[v] = Uin[v];
        Ug[3] = -Uin[3];
        return true;
    }
    return false;
}

// one thread per (face, face point): the numerical flux Fhat(U_L, U_R, n)
#ifndef FVDG_DEVICE_HELPERS_ONLY
__global__ void k_face_flux(KernelArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.F * a.Qf) return;
    const int64_t f = t / a.Qf;
    const int g = (int)(t - f * a.Qf);
    const int64_t L = a.fcell[2 * f], Rc = a.fcell[2 * f + 1];
    const double n[3] = {a.fgeo[4 * f], a.fgeo[4 * f + 1], a.fgeo[4 * f + 2]};
    double UL[5], UR[5], out[5] = {0, 0, 0, 0, 0};
    face_trace(a, L, a.fcombo[2 * f], g, UL);
    bool active = true;
    if (Rc >= 0)
        face_trace(a, Rc, a.fcombo[2 * f + 1], g, UR);
    else
        active = ghost_state(a, f, g, UL, UR);
    if (active) num_flux(a.flux, UL, UR, n, a.gamma, out);
    for (int v = 0; v < 5; ++v) a.Fh[t * 5 + v] = out[v];
}
#endif

// one thread per (element, test function j): add the four faces' contributions
#ifndef FVDG_DEVICE_HELPERS_ONLY
__global__ void k_face_test(KernelArgs a) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.E * a.N) return;
    const int64_t e = t / a.N;
    const int j = (int)(t - e * a.N);
    const int NV = 5 * a.N;
    const bool do_int = (a.parts & FVDG_PART_INTERNALS) != 0, do_bnd = (a.parts & FVDG_PART_BOUNDARYS) != 0;
    double r[5] = {0, 0, 0, 0, 0};
    for (int l = 0; l < 4; ++l) {
        const int32_t slot = a.cfaces[e * 4 + l];
        const int64_t f = slot >> 1;
        const int side = slot & 1;
        const bool bnd = a.fcell[2 * f + 1] < 0;
        if ((bnd && !do_bnd) || (!bnd && !do_int)) continue;
        const int combo = a.fcombo[2 * f + side];
        const double J = a.fgeo[4 * f + 3];
        for (int g = 0; g < a.Qf; ++g) {
            const double phi = a.Phi[((int64_t)combo * a.Qf + g) * a.N + j];
            const double jw = a.wf[g] * J;
            const double* fh = a.Fh + (f * a.Qf + g) * 5;
            if (side == 0)
                for (int v = 0; v < 5; ++v) r[v] += fh[v] * phi * jw;
            else
                for (int v = 0; v < 5; ++v) r[v] -= fh[v] * phi * jw;
        }
    }
    const bool acc = (a.parts & FVDG_PART_CELLS) != 0;
    for (int v = 0; v < 5; ++v) {
        const int64_t i = cidx(e, 5 * j + v, NV);
        a.R[i] = acc ? a.R[i] + r[v] : r[v];
    }
}
#endif

// ------------------------------------------------------------------ SSP-RK stage update
// out = a0*U0 + b0*(Us + dt * (-R / M)),  M = |det A_e| * mass_hat[k]
#ifndef FVDG_DEVICE_HELPERS_ONLY
__global__ void k_rk_update(const double* __restrict__ U0, const double* __restrict__ Us,
                            const double* __restrict__ R, double* __restrict__ out, const double* __restrict__ cgeo,
                            const double* __restrict__ mhat, int64_t E, int N, double a0, double b0, double dt,
                            const double* __restrict__ edt, int* __restrict__ bad) {
    const int NV = 5 * N;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E * NV) return;
    const int64_t e = i / NV;
    const int kv = (int)(i - e * NV);
    const double M = cgeo[e * 10 + 9] * mhat[kv / 5];
    const double L = -R[i] / M;
    const double s = Us[i] + (edt ? edt[e] : dt) * L;
    const double o = (a0 == 0.0) ? b0 * s : a0 * U0[i] + b0 * s;
    out[i] = o;
    if (bad && stage_value_bad(o, kv)) bad[0] = 1;
}
#endif

// ------------------------------------------------------------------ reductions
#ifndef FVDG_DEVICE_HELPERS_ONLY
__global__ void k_max_wave_speed(KernelArgs a, double* __restrict__ part) {
    __shared__ double sm[256];
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (t < a.F * a.Qf) {
        const int64_t f = t / a.Qf;
        const int g = (int)(t - f * a.Qf);
        double U[5];
        face_trace(a, a.fcell[2 * f], a.fcombo[2 * f], g, U);
        m = wave_speed_checked(U, a.gamma);
        if (a.fcell[2 * f + 1] >= 0) {
            face_trace(a, a.fcell[2 * f + 1], a.fcombo[2 * f + 1], g, U);
            m = max_keep_nan(m, wave_speed_checked(U, a.gamma));
        }
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = max_keep_nan(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}
#endif

// max over face points of max(|u|+c) in a grid-stride loop: one partial per block (any mesh size fits
// a fixed number of partials), then k_max_partials folds them into out[0]
#ifndef FVDG_DEVICE_HELPERS_ONLY
__global__ void k_max_wave_speed_gs(KernelArgs a, double* __restrict__ part) {
    __shared__ double sm[256];
    const int64_t n = a.F * a.Qf;
    double m = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = t / a.Qf;
        const int g = (int)(t - f * a.Qf);
        double U[5];
        face_trace(a, a.fcell[2 * f], a.fcombo[2 * f], g, U);
        m = max_keep_nan(m, wave_speed_checked(U, a.gamma));
        if (a.fcell[2 * f + 1] >= 0) {
            face_trace(a, a.fcell[2 * f + 1], a.fcombo[2 * f + 1], g, U);
            m = max_keep_nan(m, wave_speed_checked(U, a.gamma));
        }
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = max_keep_nan(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// out[0] = NaN-keeping maximum of part[0..n) (one block of 256 threads)
__global__ void k_max_partials(const double* __restrict__ part, int n, double* __restrict__ out) {
    __shared__ double sm[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) m = max_keep_nan(m, part[i]);
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sm[threadIdx.x] = max_keep_nan(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sm[0];
}
#endif

// per face: max over its points of max(|u|+c) of the traces (one thread per face)
#ifndef FVDG_DEVICE_HELPERS_ONLY
__global__ void k_face_wave_speed(KernelArgs a, double* __restrict__ out) {
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= a.F) return;
    double m = 0.0;
    for (int g = 0; g < a.Qf; ++g) {
        double U[5];
        face_trace(a, a.fcell[2 * f], a.fcombo[2 * f], g, U);
        m = max_keep_nan(m, wave_speed_checked(U, a.gamma));
        if (a.fcell[2 * f + 1] >= 0) {
            face_trace(a, a.fcell[2 * f + 1], a.fcombo[2 * f + 1], g, U);
            m = max_keep_nan(m, wave_speed_checked(U, a.gamma));
        }
    }
    out[f] = m;
}
#endif

// integral of each conserved variable: |det A_e| * c_{e,0,v} * (1/6), phi_0 = 1
#ifndef FVDG_DEVICE_HELPERS_ONLY
__global__ void k_conservation(const double* __restrict__ U, const double* __restrict__ cgeo, int64_t E, int N,
                               double* __restrict__ part) {
    __shared__ double sm[5][256];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int NV = 5 * N;
    for (int v = 0; v < 5; ++v) sm[v][threadIdx.x] = (e < E) ? cgeo[e * 10 + 9] * U[cidx(e, v, NV)] / 6.0 : 0.0;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int v = 0; v < 5; ++v) sm[v][threadIdx.x] += sm[v][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int v = 0; v < 5; ++v) part[blockIdx.x * 5 + v] = sm[v][0];
}
#endif
