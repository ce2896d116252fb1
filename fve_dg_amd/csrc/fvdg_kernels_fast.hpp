// Tuned gfx950 kernels of the DG Euler residual and SSP-RK3 stage (FP64).
//
// k_vol_mfma<ORDER>: volume term on the FP64 matrix cores.  One wavefront owns
//   16 consecutive elements.  Both contractions are v_mfma_f64_16x16x4 products:
//     interpolation  U_v[q][e] = sum_k V[q][k] C_v[k][e]            (M=q, N=e, K=k)
//     test           R_v[j][e] = sum_(q,d) G[q][d][j] B_v[(q,d)][e]   (M=j, N=e, K=(q,d))
//   with B_v[(q,d)][e] = -w_q |J_e| (F(U_e(xi_q)) A_e^{-1})_(v,d), the pointwise
//   Euler flux computed in FP64 by the lane that already holds U at that point
//   (lane l: element l%16, points q = 16t + 4r + l/16, r = 0..3: the result
//   layout of v_mfma_f64_16x16x4 puts row 4r + l/16, column l%16 in element r
//   of lane l).  The K dimension of the test product is ordered so that lane group
//   l/16 feeds exactly those points: no data movement between the two products.
//   The 16-element coefficient tile (contiguous in the element-major layout) is
//   staged once through LDS with full-width coalesced loads, and the result goes
//   back the same way.  Accumulators live in AGPRs.
//
// k_surf_tile<ORDER,FUSE>: face terms, owner-computes (each element evaluates its
//   own four faces, so no two blocks write the same element: no atomics, fixed
//   summation order).  A block owns 8 elements (4 at P4):
//     0. face metadata of the 32 (element, local face) slots
//     1. stage the 8 own coefficient blocks and the 32 neighbour blocks in LDS
//        (contiguous 5N-double copies)
//     2. lanes = (element, face, 3 face points): both traces, ghost states,
//        LF/Roe flux, scaled by +-w_g 2A_f, into LDS
//     3. lanes = (element, test function j): sum over 4 faces x Qf points
//   With FUSE the SSP-RK stage update (inverse diagonal mass + Shu-Osher
//   combination) is applied right there and the residual never goes to memory.
#pragma once
#include "fvdg_kernels.hpp"

typedef double v4d __attribute__((ext_vector_type(4)));

template <int ORDER>
struct Ord;
template <> struct Ord<0> { static constexpr int N = 1, Q = 15, QF = 7; };
template <> struct Ord<1> { static constexpr int N = 4, Q = 15, QF = 7; };
template <> struct Ord<2> { static constexpr int N = 10, Q = 15, QF = 24; };
template <> struct Ord<3> { static constexpr int N = 20, Q = 31, QF = 24; };
template <> struct Ord<4> { static constexpr int N = 35, Q = 57, QF = 24; };

template <int ORDER>
struct VolShape {
    static constexpr int N = Ord<ORDER>::N, Q = Ord<ORDER>::Q, NV = 5 * N;
    static constexpr int QT = (Q + 15) / 16;  // point tiles of the interpolation product
    static constexpr int KS = (N + 3) / 4;    // K steps of the interpolation product
    static constexpr int JT = (N + 15) / 16;  // test-function tiles of the test product
    static constexpr int QPAD = 16 * QT, KPAD = 4 * KS, JPAD = 16 * JT;
    static constexpr int STRIDE = NV | 1;     // odd LDS row stride (doubles): spreads banks
};

struct VolArgs {
    int64_t E;
    double gamma;
    const double* U;      // E*NV
    double* R;            // E*NV (written)
    const double* cgeo;   // E*10
    const double* VA;     // QPAD*KPAD   V[q][k], zero padded
    const double* GA;     // QPAD*3*JPAD G[q][d][j], zero padded
    const double* wq;     // QPAD, zero padded
};

template <int ORDER>
__global__ __launch_bounds__(64) void k_vol_mfma(VolArgs a) {
    using S = VolShape<ORDER>;
    constexpr int N = S::N, Q = S::Q, NV = S::NV, QT = S::QT, KS = S::KS, JT = S::JT;
    constexpr int KPAD = S::KPAD, JPAD = S::JPAD, STR = S::STRIDE;
    __shared__ double sC[16 * STR];
    const int l = threadIdx.x, col = l & 15, grp = l >> 4;
    const int64_t e0 = (int64_t)blockIdx.x * 16;
    const int64_t base = e0 * NV, limit = a.E * NV;

    // 1. stage the tile (16*NV contiguous doubles) in LDS
    for (int i = l; i < 16 * NV; i += 64) {
        const int ee = i / NV, kv = i - ee * NV;
        sC[ee * STR + kv] = (base + i < limit) ? a.U[base + i] : 0.0;
    }
    __syncthreads();

    // 2. geometry of this lane's element (zeros past the end)
    const int64_t e = e0 + col;
    double Ai[9], det = 0.0;
    if (e < a.E) {
        const double* g = a.cgeo + e * 10;
#pragma unroll
        for (int i = 0; i < 9; ++i) Ai[i] = g[i];
        det = g[9];
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) Ai[i] = 0.0;
    }

    v4d racc[5][JT];
#pragma unroll
    for (int v = 0; v < 5; ++v)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) racc[v][jt] = (v4d){0.0, 0.0, 0.0, 0.0};

#pragma unroll 1
    for (int t = 0; t < QT; ++t) {
        // 3. interpolation to the 16 points of tile t
        v4d uacc[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) uacc[v] = (v4d){0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const double av = a.VA[(16 * t + col) * KPAD + 4 * s + grp];
            const int k = 4 * s + grp;
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const double bv = (k < N) ? sC[col * STR + 5 * k + v] : 0.0;
                uacc[v] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, uacc[v], 0, 0, 0);
            }
        }
        // 4. pointwise flux at this lane's 4 points, fed straight into the test product
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = 16 * t + 4 * r + grp;
            double B[5][3];
            if (q < Q) {
                const double Uq[5] = {uacc[0][r], uacc[1][r], uacc[2][r], uacc[3][r], uacc[4][r]};
                double F[5][3];
                euler_flux(Uq, a.gamma, F);
                const double jw = -a.wq[q] * det;
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        B[v][d] = (F[v][0] * Ai[d] + F[v][1] * Ai[3 + d] + F[v][2] * Ai[6 + d]) * jw;
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d) B[v][d] = 0.0;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) {
                    const double av = a.GA[(q * 3 + d) * JPAD + 16 * jt + col];
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        racc[v][jt] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, B[v][d], racc[v][jt], 0, 0, 0);
                }
            }
        }
    }

    // 5. result R_v[j = 16 jt + 4 r + grp][element col] -> LDS -> coalesced store
    __syncthreads();
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = 16 * jt + 4 * r + grp;
            if (j < N)
#pragma unroll
                for (int v = 0; v < 5; ++v) sC[col * STR + 5 * j + v] = racc[v][jt][r];
        }
    __syncthreads();
    for (int i = l; i < 16 * NV; i += 64) {
        const int ee = i / NV, kv = i - ee * NV;
        if (base + i < limit) a.R[base + i] = sC[ee * STR + kv];
    }
}

// ------------------------------------------------------------------ surface
struct SurfArgs {
    int64_t E;
    double gamma;
    int flux, parts, accumulate;  // accumulate: R += faces (else R = faces)
    const double* U;              // stage input state
    double* R;                    // residual (volume part in, total out) -- unused by FUSE
    const double* cgeo;
    const int32_t* cfaces;
    const int32_t* fcell;
    const uint8_t* fcombo;
    const double* fgeo;
    const uint8_t* fbc;
    const int32_t* fdir;
    const double* UD;
    const double* Phi;   // 24*QF*N  (c, g, k)
    const double* PhiT;  // 24*N*QF  (c, k, g)
    const double* wf;
    // fused SSP-RK stage: out = a0*U0 + b0*(U + dt*(-R/M)), M = |det| * mhat[k]
    const double* U0;
    double* out;
    const double* mhat;
    double a0, b0, dt;
};

// elements per block: 8 (256 threads) up to P3, 4 (128 threads) at P4 to keep the
// LDS footprint (own + 4 neighbour coefficient blocks + face fluxes) near 45-65 KB
template <int ORDER>
struct SurfShape {
    static constexpr int TE = Ord<ORDER>::N > 20 ? 4 : 8;
    static constexpr int THREADS = TE * 32;
};

template <int ORDER, bool FUSE>
__global__ __launch_bounds__(SurfShape<ORDER>::THREADS) void k_surf_tile(SurfArgs a) {
    constexpr int N = Ord<ORDER>::N, QF = Ord<ORDER>::QF, NV = 5 * N;
    constexpr int TE = SurfShape<ORDER>::TE, NT = SurfShape<ORDER>::THREADS, GPT = (QF + 7) / 8;
    __shared__ double sSelf[TE * NV];
    __shared__ double sNbr[TE * 4 * NV];
    __shared__ double sFh[TE * 4 * QF * 5];
    __shared__ int sFace[TE * 4], sNb[TE * 4], sCs[TE * 4], sCn[TE * 4], sKind[TE * 4];
    __shared__ double sScale[TE * 4];

    const int t = threadIdx.x;
    const int64_t e_base = (int64_t)blockIdx.x * TE;
    const bool do_int = (a.parts & FVDG_PART_INTERNALS) != 0, do_bnd = (a.parts & FVDG_PART_BOUNDARYS) != 0;

    // 0. face metadata: kind 0 = skip, 1 = interior, 2 = boundary ghost
    if (t < TE * 4) {
        const int64_t e = e_base + t / 4;
        int kind = 0, f = 0, nb = -1, cs = 0, cn = 0;
        double scale = 0.0;
        if (e < a.E) {
            const int32_t slot = a.cfaces[e * 4 + (t & 3)];
            f = slot >> 1;
            const int side = slot & 1;
            const int32_t other = a.fcell[2 * f + 1 - side];
            const bool bnd = a.fcell[2 * f + 1] < 0;
            cs = a.fcombo[2 * f + side];
            cn = a.fcombo[2 * f + 1 - side];
            if (!bnd && do_int) {
                kind = 1;
                nb = other;
            } else if (bnd && do_bnd && (a.fbc[f] == FVDG_BC_DIRICHLET || a.fbc[f] == FVDG_BC_PSEUDO3DZ)) {
                kind = 2;
            }
            scale = (side == 0 ? 1.0 : -1.0) * a.fgeo[4 * f + 3];
        }
        sFace[t] = f;
        sNb[t] = nb;
        sCs[t] = cs;
        sCn[t] = cn;
        sKind[t] = kind;
        sScale[t] = scale;
    }
    // 1. own coefficient blocks (contiguous)
    for (int i = t; i < TE * NV; i += NT) sSelf[i] = (e_base * NV + i < a.E * NV) ? a.U[e_base * NV + i] : 0.0;
    __syncthreads();
    //    neighbour blocks
    for (int i = t; i < TE * 4 * NV; i += NT) {
        const int s = i / NV, kv = i - s * NV;
        const int nb = sNb[s];
        sNbr[i] = nb >= 0 ? a.U[(int64_t)nb * NV + kv] : 0.0;
    }
    __syncthreads();

    // 2. traces, ghost states and numerical flux at (element, face, GPT points)
    {
        const int s = t / 8;  // slot = element*4 + face
        const int gg = t & 7;
        const int kind = sKind[s];
        const int side_is0 = sScale[s] >= 0.0;
        double Us[GPT][5], Un[GPT][5];
#pragma unroll
        for (int i = 0; i < GPT; ++i)
#pragma unroll
            for (int v = 0; v < 5; ++v) Us[i][v] = Un[i][v] = 0.0;
        if (kind != 0) {
            const double* Ps = a.PhiT + (int64_t)sCs[s] * N * QF;
            const double* Pn = a.PhiT + (int64_t)sCn[s] * N * QF;
            const double* cs = sSelf + (s >> 2) * NV;
            const double* cn = sNbr + s * NV;
#pragma unroll 4
            for (int k = 0; k < N; ++k) {
                double c[5], d[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    c[v] = cs[5 * k + v];
                    d[v] = cn[5 * k + v];
                }
#pragma unroll
                for (int i = 0; i < GPT; ++i) {
                    const int g = gg * GPT + i;
                    if (g < QF) {
                        const double ps = Ps[k * QF + g];
                        const double pn = Pn[k * QF + g];
#pragma unroll
                        for (int v = 0; v < 5; ++v) {
                            Us[i][v] += ps * c[v];
                            Un[i][v] += pn * d[v];
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < GPT; ++i) {
            const int g = gg * GPT + i;
            if (g >= QF) continue;
            double out[5] = {0, 0, 0, 0, 0};
            if (kind != 0) {
                const int f = sFace[s];
                const double n[3] = {a.fgeo[4 * f], a.fgeo[4 * f + 1], a.fgeo[4 * f + 2]};
                double Ug[5];
                if (kind == 2) {
                    if (a.fbc[f] == FVDG_BC_DIRICHLET) {
                        const double* ud = a.UD + ((int64_t)a.fdir[f] * QF + g) * 5;
#pragma unroll
                        for (int v = 0; v < 5; ++v) Ug[v] = ud[v] + (ud[v] - Us[i][v]);
                    } else {
#pragma unroll
                        for (int v = 0; v < 5; ++v) Ug[v] = Us[i][v];
                        Ug[3] = -Us[i][3];
                    }
                    num_flux(a.flux, Us[i], Ug, n, a.gamma, out);
                } else if (side_is0) {
                    num_flux(a.flux, Us[i], Un[i], n, a.gamma, out);
                } else {
                    num_flux(a.flux, Un[i], Us[i], n, a.gamma, out);
                }
                const double sc = a.wf[g] * sScale[s];
#pragma unroll
                for (int v = 0; v < 5; ++v) out[v] *= sc;
            }
#pragma unroll
            for (int v = 0; v < 5; ++v) sFh[(s * QF + g) * 5 + v] = out[v];
        }
    }
    __syncthreads();

    // 3. test functions: r[j][v] = sum_faces sum_g phi_j(xi(g)) * Fh_scaled(g)
    for (int item = t; item < TE * N; item += NT) {
        const int el = item / N, j = item - el * N;
        const int64_t e = e_base + el;
        if (e >= a.E) continue;
        double r[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int lf = 0; lf < 4; ++lf) {
            const int s = el * 4 + lf;
            const double* P = a.Phi + (int64_t)sCs[s] * QF * N + j;
            const double* fh = sFh + s * QF * 5;
#pragma unroll 4
            for (int g = 0; g < QF; ++g) {
                const double p = P[g * N];
#pragma unroll
                for (int v = 0; v < 5; ++v) r[v] += p * fh[g * 5 + v];
            }
        }
        const int64_t i0 = e * NV + 5 * j;
        if (FUSE) {
            const double M = a.cgeo[e * 10 + 9] * a.mhat[j];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const double Rt = a.R[i0 + v] + r[v];
                const double s = a.U[i0 + v] + a.dt * (-Rt / M);
                a.out[i0 + v] = (a.a0 == 0.0) ? a.b0 * s : a.a0 * a.U0[i0 + v] + a.b0 * s;
            }
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) a.R[i0 + v] = a.accumulate ? a.R[i0 + v] + r[v] : r[v];
        }
    }
}

// ------------------------------------------------------------------ surface, persistent
// k_surf_lds<ORDER,FUSE>: same mathematics as k_surf_tile, organised so that no
// inner loop touches global memory:
//   * the face tables live in LDS for the whole kernel.  The face rules are
//     invariant under cyclic rotation of the face vertices, so the 24
//     (local face x vertex order) tables are 8 canonical ones (vertex order
//     starting at the smallest reference vertex) read through a point
//     permutation: phi_k(xi_c(g)) = PhiC[canon(c)][perm[rot(c)][g]][k];
//   * blocks are persistent (one LDS table load per block) and walk tiles of TE
//     elements; the four local faces are processed in rounds, each round staging
//     one neighbour coefficient block per element, computing traces + Riemann
//     flux for GP points per lane into LDS, then adding the test-function sums
//     into per-lane registers.
struct SurfLdsArgs {
    SurfArgs s;
    const double* PhiC;      // 8*QF*N canonical face tables
    const uint8_t* canon;    // 24: canonical index * 3 + rotation, per combo
    const int32_t* perm;     // 3*QF point permutations (rotation 0, 1, 2)
    int64_t ntiles;
};

template <int ORDER>
struct SurfLdsShape {
    static constexpr int N = Ord<ORDER>::N, QF = Ord<ORDER>::QF;
    static constexpr int TE = N > 20 ? 8 : 16;  // elements per tile
    static constexpr int GP = 2;                // face points per lane in the flux phase
    static constexpr int NPG = (QF + GP - 1) / GP;
    static constexpr int THREADS = 256;
    static constexpr int ITEMS = (TE * N + THREADS - 1) / THREADS;  // test items per lane
};

template <int ORDER, bool FUSE>
__global__ __launch_bounds__(256) void k_surf_lds(SurfLdsArgs A) {
    using SH = SurfLdsShape<ORDER>;
    constexpr int N = SH::N, QF = SH::QF, NV = 5 * N, TE = SH::TE, GP = SH::GP, NPG = SH::NPG;
    constexpr int NT = SH::THREADS, ITEMS = SH::ITEMS;
    const SurfArgs& a = A.s;
    __shared__ double sPhi[8 * QF * N];
    __shared__ double sSelf[TE * NV];
    __shared__ double sNbr[TE * NV];
    __shared__ double sFh[TE * QF * 5];
    __shared__ int sPerm[3 * QF];
    __shared__ int sCanon[24];
    __shared__ int sFace[TE * 4], sNb[TE * 4], sCs[TE * 4], sCn[TE * 4], sKind[TE * 4];
    __shared__ double sScale[TE * 4];

    const int t = threadIdx.x;
    for (int i = t; i < 8 * QF * N; i += NT) sPhi[i] = A.PhiC[i];
    for (int i = t; i < 3 * QF; i += NT) sPerm[i] = A.perm[i];
    if (t < 24) sCanon[t] = A.canon[t];
    const bool do_int = (a.parts & FVDG_PART_INTERNALS) != 0, do_bnd = (a.parts & FVDG_PART_BOUNDARYS) != 0;

    for (int64_t tile = blockIdx.x; tile < A.ntiles; tile += gridDim.x) {
        const int64_t e_base = tile * TE;
        __syncthreads();  // previous tile's readers are done with the shared buffers
        if (t < TE * 4) {
            const int64_t e = e_base + t / 4;
            int kind = 0, f = 0, nb = -1, cs = 0, cn = 0;
            double scale = 0.0;
            if (e < a.E) {
                const int32_t slot = a.cfaces[e * 4 + (t & 3)];
                f = slot >> 1;
                const int side = slot & 1;
                const bool bnd = a.fcell[2 * f + 1] < 0;
                cs = a.fcombo[2 * f + side];
                cn = a.fcombo[2 * f + 1 - side];
                if (!bnd && do_int) {
                    kind = 1;
                    nb = a.fcell[2 * f + 1 - side];
                } else if (bnd && do_bnd && (a.fbc[f] == FVDG_BC_DIRICHLET || a.fbc[f] == FVDG_BC_PSEUDO3DZ)) {
                    kind = 2;
                }
                scale = (side == 0 ? 1.0 : -1.0) * a.fgeo[4 * f + 3];
            }
            sFace[t] = f;
            sNb[t] = nb;
            sCs[t] = cs;
            sCn[t] = cn;
            sKind[t] = kind;
            sScale[t] = scale;
        }
        for (int i = t; i < TE * NV; i += NT) sSelf[i] = (e_base * NV + i < a.E * NV) ? a.U[e_base * NV + i] : 0.0;
        double r[ITEMS][5];
#pragma unroll
        for (int it = 0; it < ITEMS; ++it)
#pragma unroll
            for (int v = 0; v < 5; ++v) r[it][v] = 0.0;

#pragma unroll 1
        for (int lf = 0; lf < 4; ++lf) {
            __syncthreads();  // metadata ready / previous round's readers done
            for (int i = t; i < TE * NV; i += NT) {
                const int el = i / NV, kv = i - el * NV;
                const int nb = sNb[el * 4 + lf];
                sNbr[i] = nb >= 0 ? a.U[(int64_t)nb * NV + kv] : 0.0;
            }
            __syncthreads();
            // traces + numerical flux, GP points per lane
            if (t < TE * NPG) {
                const int el = t / NPG, pg = t - el * NPG;
                const int s = el * 4 + lf;
                const int kind = sKind[s];
                double Us[GP][5], Un[GP][5];
                int gs[GP], gn[GP];
                const int cs = sCanon[sCs[s]], cn = sCanon[sCn[s]];
                const double* Ps = sPhi + (cs / 3) * QF * N;
                const double* Pn = sPhi + (cn / 3) * QF * N;
#pragma unroll
                for (int i = 0; i < GP; ++i) {
                    const int g = min(pg * GP + i, QF - 1);
                    gs[i] = sPerm[(cs % 3) * QF + g];
                    gn[i] = sPerm[(cn % 3) * QF + g];
#pragma unroll
                    for (int v = 0; v < 5; ++v) Us[i][v] = Un[i][v] = 0.0;
                }
                if (kind != 0) {
                    const double* c0 = sSelf + el * NV;
                    const double* c1 = sNbr + el * NV;
#pragma unroll 4
                    for (int k = 0; k < N; ++k) {
                        double c[5], d[5];
#pragma unroll
                        for (int v = 0; v < 5; ++v) {
                            c[v] = c0[5 * k + v];
                            d[v] = c1[5 * k + v];
                        }
#pragma unroll
                        for (int i = 0; i < GP; ++i) {
                            const double ps = Ps[gs[i] * N + k], pn = Pn[gn[i] * N + k];
#pragma unroll
                            for (int v = 0; v < 5; ++v) {
                                Us[i][v] += ps * c[v];
                                Un[i][v] += pn * d[v];
                            }
                        }
                    }
                }
#pragma unroll
                for (int i = 0; i < GP; ++i) {
                    const int g = pg * GP + i;
                    if (g >= QF) continue;
                    double out[5] = {0, 0, 0, 0, 0};
                    if (kind != 0) {
                        const int f = sFace[s];
                        const double n[3] = {a.fgeo[4 * f], a.fgeo[4 * f + 1], a.fgeo[4 * f + 2]};
                        if (kind == 2) {
                            double Ug[5];
                            if (a.fbc[f] == FVDG_BC_DIRICHLET) {
                                const double* ud = a.UD + ((int64_t)a.fdir[f] * QF + g) * 5;
#pragma unroll
                                for (int v = 0; v < 5; ++v) Ug[v] = ud[v] + (ud[v] - Us[i][v]);
                            } else {
#pragma unroll
                                for (int v = 0; v < 5; ++v) Ug[v] = Us[i][v];
                                Ug[3] = -Us[i][3];
                            }
                            num_flux(a.flux, Us[i], Ug, n, a.gamma, out);
                        } else if (sScale[s] >= 0.0) {
                            num_flux(a.flux, Us[i], Un[i], n, a.gamma, out);
                        } else {
                            num_flux(a.flux, Un[i], Us[i], n, a.gamma, out);
                        }
                        const double sc = a.wf[g] * sScale[s];
#pragma unroll
                        for (int v = 0; v < 5; ++v) out[v] *= sc;
                    }
#pragma unroll
                    for (int v = 0; v < 5; ++v) sFh[(el * QF + g) * 5 + v] = out[v];
                }
            }
            __syncthreads();
            // test functions of this round's face
#pragma unroll
            for (int it = 0; it < ITEMS; ++it) {
                const int item = t + it * NT;
                if (item >= TE * N) continue;
                const int el = item / N, j = item - el * N;
                const int s = el * 4 + lf;
                if (sKind[s] == 0) continue;
                const int cs = sCanon[sCs[s]];
                const double* P = sPhi + (cs / 3) * QF * N + j;
                const int* pm = sPerm + (cs % 3) * QF;
                const double* fh = sFh + el * QF * 5;
#pragma unroll 4
                for (int g = 0; g < QF; ++g) {
                    const double p = P[pm[g] * N];
#pragma unroll
                    for (int v = 0; v < 5; ++v) r[it][v] += p * fh[g * 5 + v];
                }
            }
        }
        // write back (or fused SSP-RK stage update)
#pragma unroll
        for (int it = 0; it < ITEMS; ++it) {
            const int item = t + it * NT;
            if (item >= TE * N) continue;
            const int el = item / N, j = item - el * N;
            const int64_t e = e_base + el;
            if (e >= a.E) continue;
            const int64_t i0 = e * NV + 5 * j;
            if (FUSE) {
                const double M = a.cgeo[e * 10 + 9] * a.mhat[j];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    const double Rt = a.R[i0 + v] + r[it][v];
                    const double sv = a.U[i0 + v] + a.dt * (-Rt / M);
                    a.out[i0 + v] = (a.a0 == 0.0) ? a.b0 * sv : a.a0 * a.U0[i0 + v] + a.b0 * sv;
                }
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) a.R[i0 + v] = a.accumulate ? a.R[i0 + v] + r[it][v] : r[it][v];
            }
        }
    }
}
