// gfx950 kernels of the implicit path: the Picard-linearised block system of the reference's
// ImplicitConvection::assemble (src/DG/DG_Schemes/ImplicitConvection.cpp) and the pieces of a
// block-Jacobi preconditioned GMRES solve on it.
//
// System.  R(U) ~ A(U_lin) U - rhs with, per element e (NB = 5N unknowns, row 5j+v = test
// function j / variable v, column 5i+u = trial function i / variable u):
//   diagonal block   -sum_q w_q |det| phi_i (sum_d J_d(U_q) d phi_j / d x_d)      (:266-325)
//                    + sum over the element's faces of its own trace products     (:329-420, :20-167)
//   neighbour block  across each interior face: products of its own and the neighbour's traces,
//                    weighted by 1/2 (F_n(U_L) + lambda I) or 1/2 (F_n(U_R) - lambda I)
//   rhs              the boundary faces' explicit remainders
// with J_d = dF_d/dU (EulerPhysicalFlux.cpp:21-42), F_n = sum_d J_d n_d, lambda = max(|u|+c) of
// the two traces; boundary faces linearise around the ghost states of :20-167 (Dirichlet, WallTD,
// WallTN with data, Pseudo3DX/Y/Z mirrors; other types contribute nothing).
//
// Storage ("block rows"): each block row has a fixed number of slots (5 for assembled systems:
// the diagonal, then the neighbours across local faces 0..3), a column index per slot (-1 = no
// block) and NB x NB blocks stored column-major, so that the rows of a block are contiguous for
// the matrix-vector product (one thread per row, coalesced).  Assembly is owner-computes: one
// workgroup per element builds its diagonal block and its four neighbour blocks (each face's
// point matrices are formed by both neighbours), so no block is written twice and there are no
// atomics.  Reductions of the solver (dot products) are per-block partial sums added in block
// order on the host: results do not depend on scheduling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef FVDG_DEVICE_HELPERS_ONLY
#define FVDG_DEVICE_HELPERS_ONLY  // the plain kernels belong to the solver's translation unit
#endif
#include "fvdg_kernels_fast.hpp"  // Ord<ORDER>, Euler helpers

namespace fvdg_imp {

constexpr int kSlots = 5;   // diagonal + 4 faces
constexpr int kMaxCoef = 64;  // linear-combination coefficients passed by value

struct Coefs {
    double c[kMaxCoef];
};

// dF_d/dU of the Euler flux (EulerPhysicalFlux.cpp:21-42; u = U1/U0, e = U4/U0), J[d][5 r + c]
__device__ __forceinline__ void euler_jacobian(const double U[5], double gamma, double J[3][25]) {
    const double ri = 1.0 / U[0];
    const double vel[3] = {U[1] * ri, U[2] * ri, U[3] * ri};
    const double e = U[4] * ri, g1 = gamma - 1.0;
    const double q2 = vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double* A = J[d];
#pragma unroll
        for (int i = 0; i < 25; ++i) A[i] = 0.0;
        const double a = vel[d];
        A[1 + d] = 1.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            double* row = A + 5 * (1 + m);
            row[0] = -vel[m] * a;
            row[1 + m] += a;
            row[1 + d] += vel[m];
        }
        double* rd = A + 5 * (1 + d);
        rd[0] += 0.5 * g1 * q2;
#pragma unroll
        for (int k = 0; k < 3; ++k) rd[1 + k] -= g1 * vel[k];
        rd[4] = g1;
        double* er = A + 20;
        er[0] = a * (-gamma * e + g1 * q2);
#pragma unroll
        for (int k = 0; k < 3; ++k) er[1 + k] = -g1 * a * vel[k];
        er[1 + d] = gamma * e - 0.5 * g1 * (2.0 * a * a + q2);
        er[4] = gamma * a;
    }
}

// sum_d n_d dF_d/dU, one direction's Jacobian at a time
__device__ __forceinline__ void jacobian_normal(const double U[5], double gamma, const double n[3], double Fn[25]) {
    double J[3][25];
    euler_jacobian(U, gamma, J);
#pragma unroll
    for (int r = 0; r < 25; ++r) Fn[r] = J[0][r] * n[0] + J[1][r] * n[1] + J[2][r] * n[2];
}

__device__ __forceinline__ double wave_speed2(const double UL[5], const double UR[5], double gamma) {
    const double a = euler_wave_speed(UL, gamma), b = euler_wave_speed(UR, gamma);
    return a > b ? a : b;
}

// boundary point: dF multiplies the interior trace, drhs is the explicit remainder
// (ImplicitConvection.cpp:20-167)
__device__ __forceinline__ void implicit_boundary(int bc, const double Uin[5], const double* UD, const double n[3],
                                                  double gamma, double dF[25], double drhs[5]) {
#pragma unroll
    for (int i = 0; i < 25; ++i) dF[i] = 0.0;
#pragma unroll
    for (int v = 0; v < 5; ++v) drhs[v] = 0.0;
    double Ug[5], Jg[3][25];
    if ((bc == FVDG_BC_DIRICHLET || bc == 9 || bc == 10) && UD) {  // Dirichlet, WallTD, WallTN
#pragma unroll
        for (int v = 0; v < 5; ++v) Ug[v] = UD[v] + (UD[v] - Uin[v]);
        if (bc != FVDG_BC_DIRICHLET) {
            Ug[0] = Uin[0];
            const double cvT_in = Uin[4] - 0.5 * (Uin[1] * Uin[1] + Uin[2] * Uin[2] + Uin[3] * Uin[3]) / Uin[0];
            const double cvT_g = (bc == 9) ? 2 * 300 - cvT_in : cvT_in;
            Ug[4] = cvT_g + 0.5 * (Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3]) / Ug[0];
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) Ug[v] = 0.5 * (Uin[v] + Ug[v]);
        euler_jacobian(Ug, gamma, Jg);
    } else if (bc == 3 || bc == 4 || bc == FVDG_BC_PSEUDO3DZ) {  // Pseudo3DX / Y / Z
#pragma unroll
        for (int v = 0; v < 5; ++v) Ug[v] = Uin[v];
        Ug[bc - 2] = -Uin[bc - 2];
        double Um[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) Um[v] = 0.5 * (Uin[v] + Ug[v]);
        euler_jacobian(Um, gamma, Jg);
    } else {
        return;
    }
    const double lam = wave_speed2(Uin, Ug, gamma);
#pragma unroll
    for (int v = 0; v < 5; ++v) Ug[v] = 0.5 * (Uin[v] + Ug[v]);
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        double fu = 0.0;
#pragma unroll
        for (int c = 0; c < 5; ++c) {
            const double fn = Jg[0][5 * r + c] * n[0] + Jg[1][5 * r + c] * n[1] + Jg[2][5 * r + c] * n[2];
            dF[5 * r + c] = 0.5 * (fn + (r == c ? lam : 0.0));
            fu += fn * Ug[c];
        }
        drhs[r] = 0.5 * (fu - lam * Ug[r]);
    }
}

struct AssembleArgs {
    int64_t E;
    double gamma;
    int parts;              // FVDG_PART_CELLS / INTERNALS / BOUNDARYS
    const double* U;        // E*NB linearisation state
    const double* cgeo;     // E*10: A^{-1} row-major, |det A|
    const int32_t* cfaces;  // E*4: face*2 + side
    const int32_t* fcell;   // F*2
    const uint8_t* fcombo;  // F*2
    const double* fgeo;     // F*4: normal, 2*area
    const uint8_t* fbc;     // F
    const int32_t* fdir;    // F: boundary-data slot or -1
    const double* UD;       // boundary data
    const double* V;        // Q*N
    const double* G;        // Q*3*N reference gradients
    const double* wq;       // Q
    const double* Phi;      // 24*QF*N face tables by (local face, vertex order) combo
    const double* wf;       // QF
    double* A;              // E*kSlots blocks, NB*NB column-major
    int32_t* col;           // E*kSlots
    double* rhs;            // E*NB
};

template <int ORDER>
struct ImpShape {
    static constexpr int N = Ord<ORDER>::N, Q = Ord<ORDER>::Q, QF = Ord<ORDER>::QF, NB = 5 * N;
    // LDS: coefficients of the element and its 4 neighbours, volume point matrices (Q x 3 x 25),
    // face point matrices (4 x QF x (25 own + 25 neighbour + 5 remainder))
    static constexpr int LDS_DOUBLES = 5 * NB + Q * 75 + 4 * QF * 55;
};

template <int ORDER>
__global__ __launch_bounds__(256) void k_assemble(AssembleArgs a) {
    using S = ImpShape<ORDER>;
    constexpr int N = S::N, Q = S::Q, QF = S::QF, NB = S::NB;
    extern __shared__ double lds[];
    double* sC = lds;                 // 5 * NB
    double* sJ = sC + 5 * NB;         // Q * 75: -w_q |det| sum_d A^{-1}[d][d'] J_d (d' reference direction)
    double* sDs = sJ + Q * 75;        // 4 * QF * 25: own-trace point matrices (weighted, signed)
    double* sDo = sDs + 4 * QF * 25;  // 4 * QF * 25: neighbour-trace point matrices
    double* sR = sDo + 4 * QF * 25;   // 4 * QF * 5: boundary remainders
    __shared__ int32_t sF[4], sSide[4], sNb[4], sCs[4], sCo[4], sBc[4];
    const int64_t e = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid < 4) {
        const int32_t cf = a.cfaces[e * 4 + tid];
        const int f = cf >> 1, s = cf & 1;
        sF[tid] = f;
        sSide[tid] = s;
        sNb[tid] = a.fcell[2 * f + 1 - s];
        sCs[tid] = a.fcombo[2 * f + s];
        sCo[tid] = a.fcombo[2 * f + 1 - s];
        sBc[tid] = a.fbc[f];
    }
    __syncthreads();
    for (int i = tid; i < NB; i += 256) sC[i] = a.U[e * NB + i];
    for (int l = 0; l < 4; ++l)
        if (sNb[l] >= 0)
            for (int i = tid; i < NB; i += 256) sC[(1 + l) * NB + i] = a.U[(int64_t)sNb[l] * NB + i];
    __syncthreads();

    // A1. volume point matrices
    const double* geo = a.cgeo + e * 10;
    for (int q = tid; q < Q; q += 256) {
        double Uq[5] = {0, 0, 0, 0, 0}, J[3][25];
        for (int k = 0; k < N; ++k) {
            const double phi = a.V[q * N + k];
#pragma unroll
            for (int v = 0; v < 5; ++v) Uq[v] += phi * sC[5 * k + v];
        }
        euler_jacobian(Uq, a.gamma, J);
        const double s = (a.parts & FVDG_PART_CELLS) ? -a.wq[q] * geo[9] : 0.0;
        for (int dp = 0; dp < 3; ++dp)
            for (int r = 0; r < 25; ++r)
                sJ[q * 75 + dp * 25 + r] = s * (geo[dp] * J[0][r] + geo[3 + dp] * J[1][r] + geo[6 + dp] * J[2][r]);
    }
    // A2. face point matrices
    for (int idx = tid; idx < 4 * QF; idx += 256) {
        const int l = idx / QF, g = idx % QF;
        const int f = sF[l], nb = sNb[l];
        const double* fg = a.fgeo + (int64_t)f * 4;
        const double n[3] = {fg[0], fg[1], fg[2]};
        const double wgt = a.wf[g] * fg[3];
        double Us[5] = {0, 0, 0, 0, 0};
        const double* ps = a.Phi + ((int64_t)sCs[l] * QF + g) * N;
        for (int k = 0; k < N; ++k)
#pragma unroll
            for (int v = 0; v < 5; ++v) Us[v] += ps[k] * sC[5 * k + v];
        double ds[25], dof[25], dr[5] = {0, 0, 0, 0, 0};
        if (nb >= 0) {
            double Un[5] = {0, 0, 0, 0, 0};
            const double* pn = a.Phi + ((int64_t)sCo[l] * QF + g) * N;
            for (int k = 0; k < N; ++k)
#pragma unroll
                for (int v = 0; v < 5; ++v) Un[v] += pn[k] * sC[(1 + l) * NB + 5 * k + v];
            const bool own_left = sSide[l] == 0;
            const double lam = wave_speed2(Us, Un, a.gamma);  // symmetric in the two traces
            const double w = (a.parts & FVDG_PART_INTERNALS) ? wgt : 0.0;
            // dF_L = 1/2 (F_n(U_L) + lam I), dF_R = 1/2 (F_n(U_R) - lam I); rows of L take +dF_L (own)
            // and +dF_R (neighbour), rows of R -dF_R (own) and -dF_L (neighbour)
            double Fn[25];
            jacobian_normal(Us, a.gamma, n, Fn);
#pragma unroll
            for (int r = 0; r < 25; ++r) {
                const double id = (r % 6 == 0) ? lam : 0.0;  // diagonal entries of the 5x5
                ds[r] = own_left ? w * 0.5 * (Fn[r] + id) : -w * 0.5 * (Fn[r] - id);
            }
            jacobian_normal(Un, a.gamma, n, Fn);
#pragma unroll
            for (int r = 0; r < 25; ++r) {
                const double id = (r % 6 == 0) ? lam : 0.0;
                dof[r] = own_left ? w * 0.5 * (Fn[r] - id) : -w * 0.5 * (Fn[r] + id);
            }
        } else {
            const double* ud = a.fdir[f] >= 0 ? a.UD + ((int64_t)a.fdir[f] * QF + g) * 5 : nullptr;
            double dF[25];
            implicit_boundary(sBc[l], Us, ud, n, a.gamma, dF, dr);
            const double w = (a.parts & FVDG_PART_BOUNDARYS) ? wgt : 0.0;
#pragma unroll
            for (int r = 0; r < 25; ++r) {
                ds[r] = w * dF[r];
                dof[r] = 0.0;
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) dr[v] *= w;
        }
        for (int r = 0; r < 25; ++r) {
            sDs[idx * 25 + r] = ds[r];
            sDo[idx * 25 + r] = dof[r];
        }
#pragma unroll
        for (int v = 0; v < 5; ++v) sR[idx * 5 + v] = dr[v];
    }
    __syncthreads();

    // B. blocks: one (test j, trial i) pair per thread and round, 25 entries each; consecutive
    // threads take consecutive test functions j, i.e. rows 5 j + v of the column-major blocks, so
    // a wavefront's stores of one entry fall in a few cache lines
    double* Ae = a.A + e * kSlots * (int64_t)NB * NB;
    for (int p = tid; p < N * N; p += 256) {
        const int i = p / N, j = p % N;
        double dg[25];
#pragma unroll
        for (int r = 0; r < 25; ++r) dg[r] = 0.0;
        for (int q = 0; q < Q; ++q) {
            const double vi = a.V[q * N + i];
#pragma unroll
            for (int dp = 0; dp < 3; ++dp) {
                const double c = a.G[(q * 3 + dp) * N + j] * vi;
                const double* m = sJ + q * 75 + dp * 25;
#pragma unroll
                for (int r = 0; r < 25; ++r) dg[r] += c * m[r];
            }
        }
        for (int l = 0; l < 4; ++l) {
            const double* ps = a.Phi + (int64_t)sCs[l] * QF * N;
            for (int g = 0; g < QF; ++g) {
                const double c = ps[g * N + j] * ps[g * N + i];
                const double* m = sDs + (l * QF + g) * 25;
#pragma unroll
                for (int r = 0; r < 25; ++r) dg[r] += c * m[r];
            }
        }
#pragma unroll
        for (int v = 0; v < 5; ++v)
#pragma unroll
            for (int u = 0; u < 5; ++u) Ae[(int64_t)(5 * i + u) * NB + 5 * j + v] = dg[5 * v + u];
        for (int l = 0; l < 4; ++l) {
            if (sNb[l] < 0) continue;
            const double* ps = a.Phi + (int64_t)sCs[l] * QF * N;
            const double* pn = a.Phi + (int64_t)sCo[l] * QF * N;
#pragma unroll
            for (int r = 0; r < 25; ++r) dg[r] = 0.0;
            for (int g = 0; g < QF; ++g) {
                const double c = ps[g * N + j] * pn[g * N + i];
                const double* m = sDo + (l * QF + g) * 25;
#pragma unroll
                for (int r = 0; r < 25; ++r) dg[r] += c * m[r];
            }
            double* B = Ae + (int64_t)(1 + l) * NB * NB;
#pragma unroll
            for (int v = 0; v < 5; ++v)
#pragma unroll
                for (int u = 0; u < 5; ++u) B[(int64_t)(5 * i + u) * NB + 5 * j + v] = dg[5 * v + u];
        }
    }
    // C. right-hand side (boundary remainders) and the slot columns
    for (int p = tid; p < NB; p += 256) {
        const int j = p / 5, v = p % 5;
        double r = 0.0;
        for (int l = 0; l < 4; ++l) {
            if (sNb[l] >= 0) continue;
            const double* ps = a.Phi + (int64_t)sCs[l] * QF * N;
            for (int g = 0; g < QF; ++g) r -= ps[g * N + j] * sR[(l * QF + g) * 5 + v];
        }
        a.rhs[e * NB + p] = r;
    }
    if (tid < kSlots) a.col[e * kSlots + tid] = tid == 0 ? (int32_t)e : sNb[tid - 1];
}

// out = rhs + s M x_old (M: the diagonal mass per unknown; the backward-Euler right-hand side)
__global__ void k_rhs_shift(double* __restrict__ out, const double* __restrict__ rhs, const double* __restrict__ Md,
                            const double* __restrict__ x_old, double s, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = rhs[t] + s * Md[t] * x_old[t];
}

// y = (A + shift diag(Md)) x over block rows with `slots` blocks each (column -1: no block);
// one thread per row
__global__ void k_bsr_matvec(const double* __restrict__ A, const int32_t* __restrict__ col, int slots,
                             const double* __restrict__ Md, double shift, const double* __restrict__ x,
                             double* __restrict__ y, int64_t nrows, int NB) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows * NB) return;
    const int64_t b = t / NB;
    const int r = (int)(t % NB);
    double acc = 0.0;
    for (int s = 0; s < slots; ++s) {
        const int32_t c = col[b * slots + s];
        if (c < 0) continue;
        const double* blk = A + (b * slots + s) * (int64_t)NB * NB + r;
        const double* xc = x + (int64_t)c * NB;
        for (int k = 0; k < NB; ++k) acc += blk[(int64_t)k * NB] * xc[k];
    }
    if (Md) acc += shift * Md[t] * x[t];
    y[t] = acc;
}

// k_bsr_matvec with the block size known at compile time: the inner product over a block column
// is unrolled, so every thread keeps many independent loads in flight
template <int NB>
__global__ __launch_bounds__(256) void k_bsr_matvec_nb(const double* __restrict__ A, const int32_t* __restrict__ col,
                                                       int slots, const double* __restrict__ Md, double shift,
                                                       const double* __restrict__ x, double* __restrict__ y,
                                                       int64_t nrows) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows * NB) return;
    const int64_t b = t / NB;
    const int r = (int)(t - b * NB);
    double acc = 0.0;
    for (int s = 0; s < slots; ++s) {
        const int32_t c = col[b * slots + s];
        if (c < 0) continue;
        const double* blk = A + (b * slots + s) * (int64_t)NB * NB + r;
        const double* xc = x + (int64_t)c * NB;
        double part[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k = 0; k < NB; ++k) part[k & 3] += blk[(int64_t)k * NB] * xc[k];
        acc += (part[0] + part[1]) + (part[2] + part[3]);
    }
    if (Md) acc += shift * Md[t] * x[t];
    y[t] = acc;
}

// In-place LU with partial pivoting of one NB x NB column-major block per workgroup: the block
// row's diagonal block plus shift diag(Md) (identity where a row has no diagonal block)
__global__ __launch_bounds__(256) void k_block_lu(const double* __restrict__ A, const int32_t* __restrict__ diag_slot,
                                                  int slots, const double* __restrict__ Md, double shift,
                                                  double* __restrict__ LU, int32_t* __restrict__ piv, int NB,
                                                  int32_t* __restrict__ singular) {
    __shared__ double sv[256];
    __shared__ int si[256];
    const int64_t b = blockIdx.x;
    const int tid = threadIdx.x;
    double* M = LU + b * (int64_t)NB * NB;
    const int ds = diag_slot[b];
    const double* src = ds >= 0 ? A + (b * slots + ds) * (int64_t)NB * NB : nullptr;
    for (int64_t i = tid; i < (int64_t)NB * NB; i += 256) M[i] = src ? src[i] : ((i / NB) == (i % NB) ? 1.0 : 0.0);
    __syncthreads();
    if (Md && src)
        for (int i = tid; i < NB; i += 256) M[(int64_t)i * NB + i] += shift * Md[b * NB + i];
    __syncthreads();
    for (int k = 0; k < NB; ++k) {
        double best = -1.0;
        int bi = k;
        for (int i = k + tid; i < NB; i += 256) {
            const double x = fabs(M[(int64_t)k * NB + i]);
            if (x > best) {
                best = x;
                bi = i;
            }
        }
        sv[tid] = best;
        si[tid] = bi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (tid < s && (sv[tid + s] > sv[tid] || (sv[tid + s] == sv[tid] && si[tid + s] < si[tid]))) {
                sv[tid] = sv[tid + s];
                si[tid] = si[tid + s];
            }
            __syncthreads();
        }
        const int p = si[0];
        if (tid == 0) {
            piv[b * NB + k] = p;
            if (!(sv[0] > 0.0)) singular[0] = 1;  // benign race: every writer stores 1
        }
        if (p != k)
            for (int j = tid; j < NB; j += 256) {
                const double t = M[(int64_t)j * NB + k];
                M[(int64_t)j * NB + k] = M[(int64_t)j * NB + p];
                M[(int64_t)j * NB + p] = t;
            }
        __syncthreads();
        const double pk = M[(int64_t)k * NB + k];
        for (int i = k + 1 + tid; i < NB; i += 256) M[(int64_t)k * NB + i] /= pk;
        __syncthreads();
        const int m = NB - k - 1;
        for (int64_t t = tid; t < (int64_t)m * m; t += 256) {
            const int j = k + 1 + (int)(t / m), i = k + 1 + (int)(t % m);
            M[(int64_t)j * NB + i] -= M[(int64_t)k * NB + i] * M[(int64_t)j * NB + k];
        }
        __syncthreads();
    }
}

// Explicit inverse of one NB x NB block per workgroup (NB <= kMaxInvBlock), by in-place Gauss-Jordan
// elimination with partial pivoting in LDS (row-major there), of the block row's diagonal block
// plus shift diag(Md); written column-major to Inv so that applying it is a block product
// (k_bsr_matvec with one slot).  Row interchanges are undone on the columns at the end.
constexpr int kMaxInvBlock = 100;
__global__ __launch_bounds__(256) void k_block_inverse(const double* __restrict__ A, const int32_t* __restrict__ diag_slot,
                                                       int slots, const double* __restrict__ Md, double shift,
                                                       double* __restrict__ Inv, int NB, int32_t* __restrict__ singular) {
    extern __shared__ double sA[];  // NB * NB row-major, then NB column factors
    __shared__ double sv[256];
    __shared__ int si[256];
    __shared__ int spiv[kMaxInvBlock];
    double* sf = sA + NB * NB;
    const int64_t b = blockIdx.x;
    const int tid = threadIdx.x;
    const int ds = diag_slot[b];
    const double* src = ds >= 0 ? A + (b * slots + ds) * (int64_t)NB * NB : nullptr;
    for (int t = tid; t < NB * NB; t += 256) {  // src column-major: t = col * NB + row
        const int c = t / NB, r = t % NB;
        double x = src ? src[t] : (r == c ? 1.0 : 0.0);
        if (Md && src && r == c) x += shift * Md[b * NB + r];
        sA[r * NB + c] = x;
    }
    __syncthreads();
    for (int k = 0; k < NB; ++k) {
        double best = -1.0;
        int bi = k;
        for (int i = k + tid; i < NB; i += 256) {
            const double x = fabs(sA[i * NB + k]);
            if (x > best) {
                best = x;
                bi = i;
            }
        }
        sv[tid] = best;
        si[tid] = bi;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (tid < h && (sv[tid + h] > sv[tid] || (sv[tid + h] == sv[tid] && si[tid + h] < si[tid]))) {
                sv[tid] = sv[tid + h];
                si[tid] = si[tid + h];
            }
            __syncthreads();
        }
        const int p = si[0];
        if (tid == 0) {
            spiv[k] = p;
            if (!(sv[0] > 0.0)) singular[0] = 1;  // benign race: every writer stores 1
        }
        if (p != k)
            for (int j = tid; j < NB; j += 256) {
                const double t = sA[k * NB + j];
                sA[k * NB + j] = sA[p * NB + j];
                sA[p * NB + j] = t;
            }
        __syncthreads();
        const double pinv = 1.0 / sA[k * NB + k];
        for (int i = tid; i < NB; i += 256) sf[i] = (i == k) ? 0.0 : sA[i * NB + k];
        __syncthreads();
        if (tid == 0) sA[k * NB + k] = 1.0;
        __syncthreads();
        for (int j = tid; j < NB; j += 256) sA[k * NB + j] *= pinv;
        for (int i = tid; i < NB; i += 256)
            if (i != k) sA[i * NB + k] = 0.0;
        __syncthreads();
        for (int t = tid; t < NB * NB; t += 256) {
            const int i = t / NB, j = t % NB;
            if (i != k) sA[t] -= sf[i] * sA[k * NB + j];
        }
        __syncthreads();
    }
    for (int k = NB - 1; k >= 0; --k) {
        const int p = spiv[k];
        if (p != k)
            for (int i = tid; i < NB; i += 256) {
                const double t = sA[i * NB + k];
                sA[i * NB + k] = sA[i * NB + p];
                sA[i * NB + p] = t;
            }
        __syncthreads();
    }
    double* out = Inv + b * (int64_t)NB * NB;
    for (int t = tid; t < NB * NB; t += 256) out[t] = sA[(t % NB) * NB + t / NB];
}

// The same Gauss-Jordan inversion (partial pivoting, row swaps, column swaps undone at the end, the
// inverse written column-major) for blocks of at most 64 rows, one wavefront per block: each lane
// holds one row in registers, so a step needs no barrier.  Row swaps are virtual: every lane keeps
// the position its row has in the swapped order (the pivot is the largest |a_ik| over positions
// >= k, the smallest position on ties, as in k_block_inverse), and the pivot row reaches every lane
// through readlane (the pivot lane is wave-uniform).  The entry of column k (runtime) is picked from
// the registers by compare-and-select.  The arithmetic is that of k_block_inverse.
__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(x), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(x), lane);
    return __hiloint2double(hi, lo);
}

template <int NB>
__global__ __launch_bounds__(64) void k_block_inverse_wave(const double* __restrict__ A,
                                                           const int32_t* __restrict__ diag_slot, int slots,
                                                           const double* __restrict__ Md, double shift,
                                                           double* __restrict__ Inv, int32_t* __restrict__ singular) {
    static_assert(NB <= 64, "one row per lane");
    __shared__ double sZ[NB * NB];
    __shared__ int spiv[NB], ssig[NB];
    const int64_t b = blockIdx.x;
    const int lane = threadIdx.x;
    const bool own = lane < NB;
    const int ds = diag_slot[b];
    const double* src = ds >= 0 ? A + (b * slots + ds) * (int64_t)NB * NB : nullptr;
    double a[NB];
    if (src) {  // wave-uniform; src column-major: entry (lane, j) at j * NB + lane
        const int rl = own ? lane : 0;  // lanes past the block read row 0 and drop it
        const double dshift = (Md && own) ? shift * Md[b * NB + lane] : 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j) a[j] = src[j * NB + rl];
#pragma unroll
        for (int j = 0; j < NB; ++j) a[j] = own ? a[j] + (lane == j ? dshift : 0.0) : 0.0;
    } else {
#pragma unroll
        for (int j = 0; j < NB; ++j) a[j] = (own && lane == j) ? 1.0 : 0.0;
    }
    int pos = own ? lane : NB + lane;  // position of this lane's row in the swapped order
    bool bad = false;
#pragma unroll 1
    for (int k = 0; k < NB; ++k) {
        double ak = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j) ak = (j == k) ? a[j] : ak;
        double best = (own && pos >= k) ? fabs(ak) : -1.0;
        int bp = pos, bl = lane;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ob = __shfl_xor(best, off);
            const int op = __shfl_xor(bp, off), ol = __shfl_xor(bl, off);
            if (ob > best || (ob == best && op < bp)) {
                best = ob;
                bp = op;
                bl = ol;
            }
        }
        const int p = __builtin_amdgcn_readfirstlane(bp), q = __builtin_amdgcn_readfirstlane(bl);
        bad = bad || !(best > 0.0);
        if (lane == 0) spiv[k] = p;
        // the row at position k moves to the pivot's position, the pivot row to position k
        pos = lane == q ? k : (pos == k ? p : pos);
        const double pinv = 1.0 / readlane_f64(ak, q);
        const bool isq = lane == q;
        const double colk = isq ? pinv : -ak * pinv;  // the new column k
#pragma unroll
        for (int j = 0; j < NB; ++j) {  // branch-free: selects only
            const double kj = readlane_f64(a[j], q) * pinv;
            const double v = isq ? kj : a[j] - ak * kj;
            a[j] = (j == k) ? colk : v;
        }
    }
    if (bad && lane == 0) singular[0] = 1;  // benign race: every writer stores 1
    if (own)
#pragma unroll
        for (int j = 0; j < NB; ++j) sZ[pos * NB + j] = a[j];
    __syncthreads();
    if (lane == 0) {  // the column swaps undone in reverse order, as one permutation
        for (int c = 0; c < NB; ++c) ssig[c] = c;
        for (int k = NB - 1; k >= 0; --k) {
            const int q = spiv[k], t = ssig[k];
            ssig[k] = ssig[q];
            ssig[q] = t;
        }
    }
    __syncthreads();
    double* out = Inv + b * (int64_t)NB * NB;  // column-major: entry (r, c) at c * NB + r
    for (int t = lane; t < NB * NB; t += 64) out[t] = sZ[(t % NB) * NB + ssig[t / NB]];
}

// z = P^{-1} r per block row (forward and back substitution with the pivots of k_block_lu)
__global__ __launch_bounds__(256) void k_block_lu_solve(const double* __restrict__ LU, const int32_t* __restrict__ piv,
                                                        const double* __restrict__ r, double* __restrict__ z, int NB) {
    extern __shared__ double sz[];
    const int64_t b = blockIdx.x;
    const int tid = threadIdx.x;
    const double* M = LU + b * (int64_t)NB * NB;
    for (int i = tid; i < NB; i += 256) sz[i] = r[b * NB + i];
    __syncthreads();
    if (tid == 0)
        for (int k = 0; k < NB; ++k) {
            const int p = piv[b * NB + k];
            if (p != k) {
                const double t = sz[k];
                sz[k] = sz[p];
                sz[p] = t;
            }
        }
    __syncthreads();
    for (int k = 0; k < NB; ++k) {  // unit lower triangle
        const double zk = sz[k];
        for (int i = k + 1 + tid; i < NB; i += 256) sz[i] -= M[(int64_t)k * NB + i] * zk;
        __syncthreads();
    }
    for (int k = NB - 1; k >= 0; --k) {  // upper triangle
        if (tid == 0) sz[k] /= M[(int64_t)k * NB + k];
        __syncthreads();
        const double zk = sz[k];
        for (int i = tid; i < k; i += 256) sz[i] -= M[(int64_t)k * NB + i] * zk;
        __syncthreads();
    }
    for (int i = tid; i < NB; i += 256) z[b * NB + i] = sz[i];
}

// part[blk * nv + i] = partial sum over this block's range of V_i . w  (V: nv vectors of length n)
__global__ __launch_bounds__(256) void k_multi_dot(const double* __restrict__ V, int nv, int64_t n,
                                                   const double* __restrict__ w, double* __restrict__ part) {
    __shared__ double red[256];
    const int tid = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * 256 * 8;
    double wl[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int64_t i = i0 + k * 256 + tid;
        wl[k] = i < n ? w[i] : 0.0;
    }
    for (int v = 0; v < nv; ++v) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int64_t i = i0 + k * 256 + tid;
            if (i < n) s += V[(int64_t)v * n + i] * wl[k];
        }
        red[tid] = s;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
            if (tid < h) red[tid] += red[tid + h];
            __syncthreads();
        }
        if (tid == 0) part[(int64_t)blockIdx.x * nv + v] = red[0];
        __syncthreads();
    }
}

// y = beta y + sum_i c_i V_i
__global__ void k_lincomb(double* __restrict__ y, double beta, const double* __restrict__ V, int nv, int64_t n,
                          Coefs c) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double acc = beta == 0.0 ? 0.0 : beta * y[t];
    for (int v = 0; v < nv; ++v) acc += c.c[v] * V[(int64_t)v * n + t];
    y[t] = acc;
}

// y = a x + b y (x may be y; with b = 0 the old y is not read)
__global__ void k_axpby(double* y, double a, const double* x, double b, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = b == 0.0 ? a * x[t] : a * x[t] + b * y[t];
}

}  // namespace fvdg_imp
