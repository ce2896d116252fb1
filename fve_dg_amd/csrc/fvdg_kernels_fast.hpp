// Tuned gfx950 kernels of the DG Euler residual and SSP-RK3 stage.
//
// Every kernel is a template on the contraction type T: T = double is the FP64 path
// (the drop-in), T = float the mixed-precision path, which rounds coefficients and
// tables to FP32 for the volume and face contractions (FP32 matrix cores / FP32 FMA,
// FP32 accumulation) and keeps FP64 for the pointwise Euler and Riemann fluxes
// (pressure, positivity clamps, wave speeds), the stage update and the state.
//
// k_vol_mfma<ORDER,T>: volume term on the matrix cores.  One wavefront owns
//   16 consecutive elements.  Both contractions are 16x16x4 MFMA products:
//     interpolation  U_v[q][e] = sum_k V[q][k] C_v[k][e]            (M=q, N=e, K=k)
//     test           R_v[j][e] = sum_(q,d) G[q][d][j] B_v[(q,d)][e]   (M=j, N=e, K=(q,d))
//   with B_v[(q,d)][e] = -w_q |J_e| (F(U_e(xi_q)) A_e^{-1})_(v,d), the pointwise
//   Euler flux computed in FP64 by the lane that already holds U at that point
//   (lane l: element l%16, the 4 points q = 16t + row(r, l/16), r = 0..3, where
//   row() is the accumulator layout of the MFMA form: 4r + l/16 for FP64,
//   4(l/16) + r for FP32).  The K dimension of the test product is ordered so that
//   lane group l/16 feeds exactly those points: no data movement between the two
//   products.  The 16-element coefficient tile (contiguous in the element-major
//   layout) is staged once through LDS with full-width coalesced loads, and the
//   result goes back the same way.  Accumulators live in AGPRs.  The EXTREMA form
//   also writes the limiter's per-element extremes of U (max over the points and the
//   last point's values), read off the interpolated point values it already holds.
//
// Faces (no atomics anywhere; every output has one fixed summation order):
//   k_face_flux_dpp or k_face_flux_lds, then k_face_test_mfma   two passes (default): each
//     face's traces and Riemann flux once (FP64 P2-P4: one face per 16-lane row with the
//     coefficients broadcast from registers; elsewhere staged in LDS), then the test functions
//     per element on the matrix cores (FP64 up to P4, FP32 up to P3; k_face_test_lds, lanes =
//     (element, test function), elsewhere);
//   k_surf_lds                                one pass, owner-computes: each element
//     evaluates its own four faces (both traces and the flux are computed by both
//     neighbours).
//   All finish with the SSP-RK stage update (inverse diagonal mass + Shu-Osher
//   combination) when used inside the time stepper, so the total residual never goes
//   to memory.
#pragma once
#include <float.h>
#include <type_traits>
#include <utility>

#include "fvdg_kernels.hpp"

typedef double v4d __attribute__((ext_vector_type(4)));

// Cooperative global -> LDS copy of TOTAL values by NT threads: every load of the
// thread is issued before the first LDS store, so the block pays one memory
// latency per staging round instead of one per element copied.
template <int TOTAL, int NT, typename Src, typename Dst>
__device__ __forceinline__ void stage_to_lds(int t, Src src, Dst dst) {
    using V = decltype(src(0));
    constexpr int PF = (TOTAL + NT - 1) / NT;
    V tmp[PF];
#pragma unroll
    for (int m = 0; m < PF; ++m) {
        const int i = t + m * NT;
        tmp[m] = (i < TOTAL) ? src(i) : V(0);
    }
#pragma unroll
    for (int m = 0; m < PF; ++m) {
        const int i = t + m * NT;
        if (i < TOTAL) dst(i, tmp[m]);
    }
}

template <int ORDER>
struct Ord;
template <> struct Ord<0> { static constexpr int N = 1, Q = 15, QF = 7; };
template <> struct Ord<1> { static constexpr int N = 4, Q = 15, QF = 7; };
template <> struct Ord<2> { static constexpr int N = 10, Q = 15, QF = 24; };
template <> struct Ord<3> { static constexpr int N = 20, Q = 31, QF = 24; };
template <> struct Ord<4> { static constexpr int N = 35, Q = 57, QF = 24; };
template <> struct Ord<5> { static constexpr int N = 56, Q = 94, QF = 36; };

template <int ORDER>
struct VolShape {
    static constexpr int N = Ord<ORDER>::N, Q = Ord<ORDER>::Q, NV = 5 * N;
    static constexpr int QT = (Q + 15) / 16;  // point tiles of the interpolation product
    static constexpr int KS = (N + 3) / 4;    // K steps of the interpolation product
    static constexpr int JT = (N + 15) / 16;  // test-function tiles of the test product
    static constexpr int QPAD = 16 * QT, KPAD = 4 * KS, JPAD = 16 * JT;
    static constexpr int STRIDE = NV | 1;     // odd LDS row stride: spreads banks
    // waves per SIMD the register allocation is held to: P4 fits two (measured faster than one)
    static constexpr int MIN_WAVES = N == 35 ? 2 : 1;
};

// The 16x16x4 matrix-core product in the contraction type T.  Operand lanes are the
// same for both forms (A: row l%16, k l/16; B: k l/16, column l%16); the accumulator
// element r of lane l holds column l%16 and row `row(r, l/16)`, which differs.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    __device__ static acc_t mma(double a, double b, acc_t c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static constexpr int row(int r, int grp) { return 4 * r + grp; }
};
template <>
struct Mfma<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    __device__ static acc_t mma(float a, float b, acc_t c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static constexpr int row(int r, int grp) { return 4 * grp + r; }
};

template <typename T>
struct VolArgs {
    int64_t E;
    double gamma;
    const double* U;     // E*NV state (FP64)
    double* R;           // E*NV volume residual (written, FP64)
    const double* cgeo;  // E*10
    const T* VA;         // QPAD*KPAD   V[q][k], zero padded
    const T* GA;         // QPAD*3*JPAD G[q][d][j], zero padded
    const double* wq;    // QPAD, zero padded
    const double* Gs;    // N*3: sum_q w_q d phi_j / d xi_d (FP32 path only)
    double* ext = nullptr;  // EXTREMA: E*10 limiter extremes of U (max over points, last point)
};

// k_vol_mfma<ORDER,T>: T = double is the FP64 path.  T = float splits off the parts
// that cancel against the face terms and keeps them in FP64: the interpolation adds
// the mode-0 coefficient (phi_0 = 1) in FP64 to an FP32 product over modes >= 1, and
// the test product runs on F(U_q) - F(c_0) (c_0 = cell mean) in FP32, while the mean
// flux's exact contribution -|J| (F(c_0) A^{-1}) . sum_q w_q grad phi_j is added in
// FP64.  FP32 rounding then scales with the variation of the state inside the cell,
// and a uniform state has an FP64-exact residual.
template <int ORDER, typename T, bool EXTREMA = false>
__global__ __launch_bounds__(64, VolShape<ORDER>::MIN_WAVES) void k_vol_mfma(VolArgs<T> a) {
    using S = VolShape<ORDER>;
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    static_assert(!(SPLIT && EXTREMA), "limiter extremes need the FP64 point values");
    constexpr int N = S::N, Q = S::Q, NV = S::NV, QT = S::QT, KS = S::KS, JT = S::JT;
    constexpr int KPAD = S::KPAD, JPAD = S::JPAD, STR = S::STRIDE;
    // the tile's coefficients, and (split) later the FP64 result in the same LDS
    __shared__ double sBuf[16 * STR];
    T* sC = reinterpret_cast<T*>(sBuf);
    double* sOut = sBuf;
    const int l = threadIdx.x, col = l & 15, grp = l >> 4;
    const int64_t e0 = (int64_t)blockIdx.x * 16;
    const int64_t base = e0 * NV, limit = a.E * NV;

    // 1. stage the tile (16*NV contiguous doubles) in LDS (split: mode 0 left out)
    stage_to_lds<16 * NV, 64>(
        l,
        [&](int i) { return (base + i < limit && (!SPLIT || i % NV >= 5)) ? (T)a.U[base + i] : (T)0; },
        [&](int i, T x) { sC[(i / NV) * STR + (i % NV)] = x; });
    __syncthreads();

    // 2. geometry of this lane's element (zeros past the end); split: mean state and flux
    const int64_t e = e0 + col;
    double Ai[9], det = 0.0, c0[5] = {0, 0, 0, 0, 0}, FrA[5][3];
    if (e < a.E) {
        const double* g = a.cgeo + e * 10;
#pragma unroll
        for (int i = 0; i < 9; ++i) Ai[i] = g[i];
        det = g[9];
        if (SPLIT)
#pragma unroll
            for (int v = 0; v < 5; ++v) c0[v] = a.U[e * NV + v];
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) Ai[i] = 0.0;
    }
    if (SPLIT) {
        double F[5][3];
        euler_flux(c0, a.gamma, F);
#pragma unroll
        for (int v = 0; v < 5; ++v)
#pragma unroll
            for (int d = 0; d < 3; ++d) FrA[v][d] = F[v][0] * Ai[d] + F[v][1] * Ai[3 + d] + F[v][2] * Ai[6 + d];
    }

    acc_t racc[5][JT];
#pragma unroll
    for (int v = 0; v < 5; ++v)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) racc[v][jt] = (acc_t){0, 0, 0, 0};
    // EXTREMA: this lane's running maximum over its points (the last point's values are
    // stored as soon as they are known)
    double emax[5] = {-DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};

#pragma unroll 1
    for (int t = 0; t < QT; ++t) {
        // 3. interpolation to the 16 points of tile t
        acc_t uacc[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) uacc[v] = (acc_t){0, 0, 0, 0};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const T av = a.VA[(16 * t + col) * KPAD + 4 * s + grp];
            const int k = 4 * s + grp;
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const T bv = (k < N) ? sC[col * STR + 5 * k + v] : (T)0;
                uacc[v] = M::mma(av, bv, uacc[v]);
            }
        }
        // 4. pointwise flux at this lane's 4 points, fed straight into the test product
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = 16 * t + M::row(r, grp);
            T B[5][3];
            if (q < Q) {
                double Uq[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) Uq[v] = c0[v] + (double)uacc[v][r];
                if (EXTREMA) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) emax[v] = Uq[v] > emax[v] ? Uq[v] : emax[v];
                    if (q == Q - 1 && e < a.E)
#pragma unroll
                        for (int v = 0; v < 5; ++v) a.ext[e * 10 + 5 + v] = Uq[v];
                }
                double F[5][3];
                euler_flux(Uq, a.gamma, F);
                const double jw = -a.wq[q] * det;
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        double fa = F[v][0] * Ai[d] + F[v][1] * Ai[3 + d] + F[v][2] * Ai[6 + d];
                        if (SPLIT) fa -= FrA[v][d];
                        B[v][d] = (T)(fa * jw);
                    }
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d) B[v][d] = (T)0;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) {
                    const T av = a.GA[(q * 3 + d) * JPAD + 16 * jt + col];
#pragma unroll
                    for (int v = 0; v < 5; ++v) racc[v][jt] = M::mma(av, B[v][d], racc[v][jt]);
                }
            }
        }
    }

    if (EXTREMA) {  // the 4 lanes of an element (grp 0..3) hold its 16-point tiles' rows
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            double x = __shfl_xor(emax[v], 16, 64);
            emax[v] = x > emax[v] ? x : emax[v];
            x = __shfl_xor(emax[v], 32, 64);
            emax[v] = x > emax[v] ? x : emax[v];
        }
        if (e < a.E && grp == 0)
#pragma unroll
            for (int v = 0; v < 5; ++v) a.ext[e * 10 + v] = emax[v];
    }

    // 5. result R_v[j = 16 jt + row(r, grp)][element col] -> LDS -> coalesced store
    __syncthreads();
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = 16 * jt + M::row(r, grp);
            if (j < N) {
                if (SPLIT) {
                    const double g0 = a.Gs[3 * j], g1 = a.Gs[3 * j + 1], g2 = a.Gs[3 * j + 2];
#pragma unroll
                    for (int v = 0; v < 5; ++v)
                        sOut[col * STR + 5 * j + v] =
                            (double)racc[v][jt][r] - det * (FrA[v][0] * g0 + FrA[v][1] * g1 + FrA[v][2] * g2);
                } else {
#pragma unroll
                    for (int v = 0; v < 5; ++v) sC[col * STR + 5 * j + v] = racc[v][jt][r];
                }
            }
        }
    __syncthreads();
    for (int i = l; i < 16 * NV; i += 64) {
        const int ee = i / NV, kv = i - ee * NV;
        if (base + i < limit) a.R[base + i] = SPLIT ? sOut[ee * STR + kv] : (double)sC[ee * STR + kv];
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
    // fused SSP-RK stage: out = a0*U0 + b0*(U + dt_e*(-R/M)), M = |det| * mhat[k],
    // dt_e = edt[e] when edt is set (problems of an ensemble), else dt
    const double* U0;
    double* out;
    const double* mhat;
    double a0, b0, dt;
    const double* edt;
    int* bad;  // set to 1 when a stage output is non-finite or a mean density <= 0 (nullable)
};

// ------------------------------------------------------------------ surface, persistent
// k_surf_lds<ORDER,FUSE>: owner-computes faces, organised so that no inner loop
// touches global memory:
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
    static constexpr int TE = N > 35 ? 4 : (N > 20 ? 8 : 16);  // elements per tile
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
        stage_to_lds<TE * NV, NT>(
            t, [&](int i) { return (e_base * NV + i < a.E * NV) ? a.U[e_base * NV + i] : 0.0; },
            [&](int i, double x) { sSelf[i] = x; });
        double r[ITEMS][5];
#pragma unroll
        for (int it = 0; it < ITEMS; ++it)
#pragma unroll
            for (int v = 0; v < 5; ++v) r[it][v] = 0.0;

#pragma unroll 1
        for (int lf = 0; lf < 4; ++lf) {
            __syncthreads();  // metadata ready / previous round's readers done
            stage_to_lds<TE * NV, NT>(
                t,
                [&](int i) {
                    const int nb = sNb[(i / NV) * 4 + lf];
                    return nb >= 0 ? a.U[(int64_t)nb * NV + (i % NV)] : 0.0;
                },
                [&](int i, double x) { sNbr[i] = x; });
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
                bool badl = false;
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    const double Rt = a.R[i0 + v] + r[it][v];
                    const double sv = a.U[i0 + v] + (a.edt ? a.edt[e] : a.dt) * (-Rt / M);
                    const double o = (a.a0 == 0.0) ? a.b0 * sv : a.a0 * a.U0[i0 + v] + a.b0 * sv;
                    a.out[i0 + v] = o;
                    badl |= stage_value_bad(o, 5 * j + v);
                }
                if (badl && a.bad) a.bad[0] = 1;
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) a.R[i0 + v] = a.accumulate ? a.R[i0 + v] + r[it][v] : r[it][v];
            }
        }
    }
}

// ------------------------------------------------------------------ faces in two passes
// k_face_flux_lds<ORDER,T>: every face once (face-based, no duplicated work): lanes =
//   (face, face point) (two points at P5 in FP64); both traces from LDS-staged coefficient blocks and the
//   canonical LDS tables (contraction in T), ghost state on boundary faces, LF/Roe
//   flux in FP64, scaled by w_g * 2A_f and written to Fh[f][g][v] (type T,
//   contiguous per face).
// k_face_test_lds<ORDER,FUSE,T>: lanes = (element, test function j); each local face's
//   Fh of the tile's elements is staged in LDS with the side's sign, then
//   sum_faces sum_g phi_j(xi_side(g)) Fh[g] (in T) is added to the volume part Rv
//   and written to R (FP64), or fed to the SSP-RK stage update (FP64).
// With T = float the parts that cancel between neighbours and against the volume term
// stay FP64: traces add the mode-0 coefficient in FP64 to an FP32 product over modes
// >= 1 (k_face_flux_dpp keeps the whole trace in FP64), and each face stores a reference flux Fref (its point-0 flux, FP64, times 2A)
// plus the FP32 differences Fh[g] = (Fhat_g - Fref) w_g 2A; the test pass adds
// Fref 2A sum_g w_g phi_j(xi(g)) in FP64 (sum_g w_g phi_j depends only on the
// canonical face table, table IC).  A uniform flux then cancels exactly in FP64.
template <typename T>
struct FaceFluxArgs {
    int64_t F;
    double gamma;
    int flux;
    const double* U;
    const int32_t* fcell;
    const uint8_t* fcombo;
    const double* fgeo;
    const uint8_t* fbc;
    const int32_t* fdir;
    const double* UD;
    const double* wf;
    const T* PhiC;
    const uint8_t* canon;
    const int32_t* perm;
    T* Fh;         // F*QF*5
    double* Fref;  // F*5 (FP32 path only)
    int64_t ntiles;
};

template <int ORDER, typename T = double>
struct FaceFluxShape {
    static constexpr int N = Ord<ORDER>::N, QF = Ord<ORDER>::QF;
    static constexpr bool F64 = std::is_same<T, double>::value;
    // face points per lane: one, except at P5 in FP64 (two: the trace products reuse each
    // staged coefficient twice, and the 64-thread tile leaves the registers for it)
    static constexpr int GP = F64 && N > 35 ? 2 : 1, LPF = (QF + GP - 1) / GP;  // lanes per face
    // P5's table (8 x 36 x 57 doubles, 131 KB) leaves room for only a small tile; elsewhere the
    // wavefronts of a block share one copy of the tables.  FP64: P4 in one 12-wave block per CU
    // (147 KB of LDS; 3 waves per SIMD, a few registers spilled), P0-P3 in 4-wave blocks with the
    // registers held to 3 waves per SIMD at P2 and P3 (P3: 3 blocks, 12 waves per CU; P1 would spill)
    static constexpr int THREADS = N > 35 ? 64 : (F64 && N == 35 ? 768 : 256);
    static constexpr int MIN_WAVES = N > 35 || !F64 ? 1 : (N == 35 || N >= 10 ? 3 : 2);
    static constexpr int FB = THREADS / LPF;                // faces per tile
};

template <int ORDER, typename T>
__global__ __launch_bounds__((FaceFluxShape<ORDER, T>::THREADS), (FaceFluxShape<ORDER, T>::MIN_WAVES)) void k_face_flux_lds(FaceFluxArgs<T> a) {
    using SH = FaceFluxShape<ORDER, T>;
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    constexpr int N = SH::N, QF = SH::QF, NV = 5 * N, GP = SH::GP, LPF = SH::LPF, FB = SH::FB;
    // tables transposed to [table][mode][point] with an odd point stride: at one mode the lanes of
    // a face read within one row, and the eight tables start on well-separated LDS banks
    constexpr int NT = SH::THREADS, QS = QF | 1;
    __shared__ T sPhi[8 * N * QS];
    __shared__ T sC[2 * FB * NV];
    __shared__ double sC0[SPLIT ? 2 * FB * 5 : 1], sRef[SPLIT ? FB * 5 : 1];
    __shared__ int sPerm[3 * QF];
    __shared__ int sCanon[24];
    __shared__ int sCell[2 * FB], sCb[2 * FB], sKind[FB];

    const int t = threadIdx.x;
    for (int i = t; i < 8 * QF * N; i += NT) {  // PhiC[c][g][k] -> sPhi[c][k][g]
        const int c = i / (QF * N), g = (i / N) % QF, kk = i % N;
        sPhi[(c * N + kk) * QS + g] = a.PhiC[i];
    }
    for (int i = t; i < 3 * QF; i += NT) sPerm[i] = a.perm[i];
    if (t < 24) sCanon[t] = a.canon[t];
    const int fl = t / LPF, pg = t - fl * LPF;

    for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const int64_t f0 = tile * FB;
        __syncthreads();
        if (t < FB) {
            const int64_t f = f0 + t;
            int kind = 0, L = -1, R = -1, cL = 0, cR = 0;
            if (f < a.F) {
                L = a.fcell[2 * f];
                R = a.fcell[2 * f + 1];
                cL = sCanon[a.fcombo[2 * f]];
                cR = sCanon[a.fcombo[2 * f + 1]];
                if (R >= 0)
                    kind = 1;
                else if (a.fbc[f] == FVDG_BC_DIRICHLET || a.fbc[f] == FVDG_BC_PSEUDO3DZ)
                    kind = 2;
            }
            sCell[2 * t] = (kind != 0) ? L : -1;
            sCell[2 * t + 1] = (kind == 1) ? R : -1;
            sCb[2 * t] = cL;
            sCb[2 * t + 1] = cR;
            sKind[t] = kind;
        }
        __syncthreads();
        stage_to_lds<2 * FB * NV, NT>(
            t,
            [&](int i) {
                const int c = sCell[i / NV];
                return (c >= 0 && (!SPLIT || i % NV >= 5)) ? (T)a.U[(int64_t)c * NV + (i % NV)] : (T)0;
            },
            [&](int i, T x) { sC[i] = x; });
        if (SPLIT)
            for (int i = t; i < 2 * FB * 5; i += NT) {
                const int c = sCell[i / 5];
                sC0[i] = c >= 0 ? a.U[(int64_t)c * NV + i % 5] : 0.0;
            }
        __syncthreads();
        const int64_t f = f0 + fl;
        const bool lane_on = t < FB * LPF && f < a.F;
        double fo[GP][5];
        double J = 0.0;
        if (lane_on) {
            const int kind = sKind[fl];
            T UL[GP][5], UR[GP][5];
            int gl[GP], gr[GP];
            const int cL = sCb[2 * fl], cR = sCb[2 * fl + 1];
            const T* PL = sPhi + (cL / 3) * N * QS;
            const T* PR = sPhi + (cR / 3) * N * QS;
#pragma unroll
            for (int i = 0; i < GP; ++i) {
                const int g = min(pg * GP + i, QF - 1);
                gl[i] = sPerm[(cL % 3) * QF + g];
                gr[i] = sPerm[(cR % 3) * QF + g];
#pragma unroll
                for (int v = 0; v < 5; ++v) UL[i][v] = UR[i][v] = (T)0;
            }
            if (kind != 0) {
                const T* c0 = sC + (2 * fl) * NV;
                const T* c1 = sC + (2 * fl + 1) * NV;
#pragma unroll 4
                for (int k = SPLIT ? 1 : 0; k < N; ++k) {
                    T c[5], d[5];
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        c[v] = c0[5 * k + v];
                        d[v] = c1[5 * k + v];
                    }
#pragma unroll
                    for (int i = 0; i < GP; ++i) {
                        const T pl = PL[k * QS + gl[i]], pr = PR[k * QS + gr[i]];
#pragma unroll
                        for (int v = 0; v < 5; ++v) {
                            UL[i][v] += pl * c[v];
                            UR[i][v] += pr * d[v];
                        }
                    }
                }
            }
            const double n[3] = {a.fgeo[4 * f], a.fgeo[4 * f + 1], a.fgeo[4 * f + 2]};
            J = a.fgeo[4 * f + 3];
#pragma unroll
            for (int i = 0; i < GP; ++i) {
                const int g = pg * GP + i;
                double uL[5], uR[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    fo[i][v] = 0.0;
                    uL[v] = (SPLIT ? sC0[(2 * fl) * 5 + v] : 0.0) + (double)UL[i][v];
                    uR[v] = (SPLIT ? sC0[(2 * fl + 1) * 5 + v] : 0.0) + (double)UR[i][v];
                }
                if (g >= QF) continue;
                if (kind == 1) {
                    num_flux(a.flux, uL, uR, n, a.gamma, fo[i]);
                } else if (kind == 2) {
                    double Ug[5];
                    if (a.fbc[f] == FVDG_BC_DIRICHLET) {
                        const double* ud = a.UD + ((int64_t)a.fdir[f] * QF + g) * 5;
#pragma unroll
                        for (int v = 0; v < 5; ++v) Ug[v] = ud[v] + (ud[v] - uL[v]);
                    } else {
#pragma unroll
                        for (int v = 0; v < 5; ++v) Ug[v] = uL[v];
                        Ug[3] = -uL[3];
                    }
                    num_flux(a.flux, uL, Ug, n, a.gamma, fo[i]);
                }
            }
        }
        if (SPLIT) {
            // the face's reference flux: its point 0, held by lane pg = 0 as i = 0
            if (lane_on && pg == 0)
#pragma unroll
                for (int v = 0; v < 5; ++v) sRef[fl * 5 + v] = fo[0][v];
            __syncthreads();
            if (lane_on && pg == 0)
#pragma unroll
                for (int v = 0; v < 5; ++v) a.Fref[f * 5 + v] = fo[0][v] * J;
        }
        if (lane_on) {
#pragma unroll
            for (int i = 0; i < GP; ++i) {
                const int g = pg * GP + i;
                if (g >= QF) continue;
                const double sc = a.wf[g] * J;
                T* o = a.Fh + (f * QF + g) * 5;
#pragma unroll
                for (int v = 0; v < 5; ++v) o[v] = (T)((fo[i][v] - (SPLIT ? sRef[fl * 5 + v] : 0.0)) * sc);
            }
        }
    }
}

// k_face_flux_dpp<ORDER>: the FP64 face pass with one face per 16-lane row of a wavefront
//   (QF <= 32: up to 16 lanes of GP face points).  The two neighbours' coefficient blocks are
//   not staged in LDS: lane m of the row holds entries m, m + 16, ... of each block in registers
//   (loaded straight from U, 128 contiguous bytes per row and load), and each entry reaches the
//   row's lanes through a DPP row broadcast (v_mov_b64_dpp row_newbcast) at the step of the
//   trace sum that uses it.  LDS holds only the canonical face tables, so a lane reads GP table
//   values per mode and side instead of the 10 coefficients the LDS kernel re-reads per mode.
//   Every lane of a row runs the trace sum (a DPP broadcast needs its source lane active);
//   lanes past the face's points compute a clamped point and discard it.  Flux and output as in
//   k_face_flux_lds; with T = float (the mixed-precision stage) the traces and fluxes stay FP64
//   and the face's reference flux (its point 0, held by lane 0 of the row) reaches the row by one
//   more broadcast: Fref = F_0 2A in FP64, Fh[g] = (F_g - F_0) w_g 2A in FP32.
template <int ORDER, typename T = double, int NTHREADS = 256>
struct FaceFluxDppShape {
    static constexpr int N = Ord<ORDER>::N, QF = Ord<ORDER>::QF, NV = 5 * N;
    static constexpr int GP = (QF + 15) / 16, LPF = (QF + GP - 1) / GP;  // points per lane, lanes per face
    static constexpr int CR = (NV + 15) / 16;                           // coefficient registers per side
    // faces per block and tile: one per 16-lane row (the ensemble autotuner also tries 128- and
    // 512-thread blocks: the tables are loaded once per block)
    static constexpr int THREADS = NTHREADS, FB = THREADS / 16;
    // FP64: registers held to 3 waves per SIMD up to P3 (measured faster than 2 waves; P4's tables,
    // 56 KB, allow 2 blocks per CU anyway), each tile's coefficients loaded at its start.  FP32 traces
    // (mixed): the next tile's face and coefficients are loaded during the current tile's Riemann
    // solves, with 2 waves per SIMD (P3 196k: 0.290 ms vs 0.322 at 3 waves with spills and 0.378
    // without the prefetch; the FP64 kernel measured 0.312 without and 0.321 with it)
    static constexpr bool PREFETCH = !std::is_same<T, double>::value;
    static constexpr int MIN_WAVES = (N <= 20 && !PREFETCH) ? 3 : 2;
    static constexpr bool SUPPORTED = QF <= 32;
};

template <typename F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int COUNT, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_impl(f, std::make_integer_sequence<int, COUNT>{});
}

// every lane of a 16-lane row receives the value lane SRC of the row holds (DPP row_newbcast,
// one 32-bit move per half: the DPP builtins move 32-bit values)
template <int SRC>
__device__ __forceinline__ double row_bcast(double x) {
    const unsigned long u = __builtin_bit_cast(unsigned long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, 0x150 + SRC, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x150 + SRC, 0xF, 0xF, true);
    return __builtin_bit_cast(double, ((unsigned long)(unsigned)hi << 32) | (unsigned)lo);
}

// One mode k of a trace sum for two face points of one side:
//   acc[i][v] += c_v(broadcast from lane L_v = (5k + v) % 16 of the row) * t[i],  i = 0, 1, v = 0..4,
// as ten v_fmac_f64_dpp / v_fmac_f32_dpp row_newbcast instructions: the broadcast is the DPP source
// operand of the FMA itself (the compiler does not fuse a DPP move into a multiply-add, and
// spends two 32-bit moves per double on the broadcast otherwise).  DPP reads of a VGPR need two
// wait states after a VALU write of it: the block opens with s_nop 1, and no instruction of the
// block reads a register another one of it writes.  Ten accumulators, five coefficient registers
// (c_v lives in register (5k + v) / 16 of the side's coefficient set) and two table values.
// The lane pattern repeats with period 16 in k: dpp_mode<k % 16>.
#define FVDG_DPP_I(OP, ACC, C, T, L) OP " %" #ACC ", %" #C ", %" #T " row_newbcast:" #L " row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
#define FVDG_DPP_BODY(OP, L0, L1, L2, L3, L4)                                                       \
    "s_nop 1\n" FVDG_DPP_I(OP, 0, 10, 15, L0) FVDG_DPP_I(OP, 5, 10, 16, L0) FVDG_DPP_I(OP, 1, 11, 15, L1) \
        FVDG_DPP_I(OP, 6, 11, 16, L1) FVDG_DPP_I(OP, 2, 12, 15, L2) FVDG_DPP_I(OP, 7, 12, 16, L2)       \
            FVDG_DPP_I(OP, 3, 13, 15, L3) FVDG_DPP_I(OP, 8, 13, 16, L3) FVDG_DPP_I(OP, 4, 14, 15, L4)   \
                FVDG_DPP_I(OP, 9, 14, 16, L4)
template <int KM, typename TR>
__device__ __forceinline__ void dpp_mode(TR (&a)[2][5], TR c0, TR c1, TR c2, TR c3, TR c4, TR t0, TR t1);
#define FVDG_DPP_SPEC(KM, TR, OP, L0, L1, L2, L3, L4)                                                          \
    template <>                                                                                             \
    __device__ __forceinline__ void dpp_mode<KM, TR>(TR(&a)[2][5], TR c0, TR c1, TR c2, TR c3, TR c4, TR t0, TR t1) { \
        asm volatile(FVDG_DPP_BODY(OP, L0, L1, L2, L3, L4)                                                  \
                     : "+v"(a[0][0]), "+v"(a[0][1]), "+v"(a[0][2]), "+v"(a[0][3]), "+v"(a[0][4]), "+v"(a[1][0]),  \
                       "+v"(a[1][1]), "+v"(a[1][2]), "+v"(a[1][3]), "+v"(a[1][4])                           \
                     : "v"(c0), "v"(c1), "v"(c2), "v"(c3), "v"(c4), "v"(t0), "v"(t1));                      \
    }
#define FVDG_DPP_SPEC2(KM, L0, L1, L2, L3, L4)                            \
    FVDG_DPP_SPEC(KM, double, "v_fmac_f64_dpp", L0, L1, L2, L3, L4) \
    FVDG_DPP_SPEC(KM, float, "v_fmac_f32_dpp", L0, L1, L2, L3, L4)
FVDG_DPP_SPEC2(0, 0, 1, 2, 3, 4)
FVDG_DPP_SPEC2(1, 5, 6, 7, 8, 9)
FVDG_DPP_SPEC2(2, 10, 11, 12, 13, 14)
FVDG_DPP_SPEC2(3, 15, 0, 1, 2, 3)
FVDG_DPP_SPEC2(4, 4, 5, 6, 7, 8)
FVDG_DPP_SPEC2(5, 9, 10, 11, 12, 13)
FVDG_DPP_SPEC2(6, 14, 15, 0, 1, 2)
FVDG_DPP_SPEC2(7, 3, 4, 5, 6, 7)
FVDG_DPP_SPEC2(8, 8, 9, 10, 11, 12)
FVDG_DPP_SPEC2(9, 13, 14, 15, 0, 1)
FVDG_DPP_SPEC2(10, 2, 3, 4, 5, 6)
FVDG_DPP_SPEC2(11, 7, 8, 9, 10, 11)
FVDG_DPP_SPEC2(12, 12, 13, 14, 15, 0)
FVDG_DPP_SPEC2(13, 1, 2, 3, 4, 5)
FVDG_DPP_SPEC2(14, 6, 7, 8, 9, 10)
FVDG_DPP_SPEC2(15, 11, 12, 13, 14, 15)
#undef FVDG_DPP_SPEC2
#undef FVDG_DPP_SPEC
#undef FVDG_DPP_BODY
#undef FVDG_DPP_I

// k_face_flux_dpp<ORDER,T>: the trace arithmetic is T.  T = double: FP64 traces (the drop-in).
// T = float (mixed precision): the coefficients of modes >= 1 and the tables in FP32, FP32
// accumulation, the cell means (mode 0, phi_0 = 1) added in FP64; Riemann flux in FP64.
template <int ORDER, typename T = double, int NTHREADS = 256>
__global__ __launch_bounds__(NTHREADS, (FaceFluxDppShape<ORDER, T, NTHREADS>::MIN_WAVES)) void k_face_flux_dpp(FaceFluxArgs<T> a) {
    using SH = FaceFluxDppShape<ORDER, T, NTHREADS>;
    constexpr int N = SH::N, QF = SH::QF, NV = SH::NV, GP = SH::GP, LPF = SH::LPF, CR = SH::CR, FB = SH::FB;
    constexpr int NT = SH::THREADS, QS = QF | 1;
    static_assert(SH::SUPPORTED && GP == 2, "two face points per lane of a 16-lane row");
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    __shared__ T sPhi[8 * N * QS];  // [table][mode][point], odd point stride
    __shared__ int sPerm[3 * QF];
    __shared__ int sCanon[24];
    const int t = threadIdx.x;
    for (int i = t; i < 8 * QF * N; i += NT) {
        const int c = i / (QF * N), g = (i / N) % QF, kk = i % N;
        sPhi[(c * N + kk) * QS + g] = a.PhiC[i];
    }
    for (int i = t; i < 3 * QF; i += NT) sPerm[i] = a.perm[i];
    if (t < 24) sCanon[t] = a.canon[t];
    __syncthreads();
    const int row = t >> 4, m = t & 15;

    // the face of this row in a tile: its kind (0 none, 1 interior, 2 Dirichlet / Pseudo3DZ boundary),
    // canonical tables of both sides and the raw FP64 coefficients (lane m holds entries m, m + 16,
    // ... of each block).  The next tile's are loaded while the current tile's Riemann solves run.
    int64_t f = 0;
    int kind = 0, cL = 0, cR = 0;
    double ql[CR], qr[CR];
    auto fetch = [&](int64_t tile) __attribute__((always_inline)) {
        f = tile * FB + row;
        kind = 0;
        cL = cR = 0;
        int64_t L = -1, R = -1;
        if (tile < a.ntiles && f < a.F) {
            L = a.fcell[2 * f];
            R = a.fcell[2 * f + 1];
            cL = sCanon[a.fcombo[2 * f]];
            cR = sCanon[a.fcombo[2 * f + 1]];
            if (R >= 0)
                kind = 1;
            else if (a.fbc[f] == FVDG_BC_DIRICHLET || a.fbc[f] == FVDG_BC_PSEUDO3DZ)
                kind = 2;
        }
        const int64_t Lc = kind != 0 ? L : -1, Rc = kind == 1 ? R : -1;
#pragma unroll
        for (int r = 0; r < CR; ++r) {
            const int i = r * 16 + m;
            ql[r] = (Lc >= 0 && i < NV) ? a.U[Lc * NV + i] : 0.0;
            qr[r] = (Rc >= 0 && i < NV) ? a.U[Rc * NV + i] : 0.0;
        }
    };
    if constexpr (SH::PREFETCH) fetch(blockIdx.x);
    for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        if constexpr (!SH::PREFETCH) fetch(tile);
        const int64_t fc = f;
        const int kc = kind;
        const bool on = fc < a.F && m < LPF;
        // this tile's face geometry and weights, loaded before the next tile's prefetch
        double n[3] = {0, 0, 0}, J = 0, wg[GP];
        if (on) {
            n[0] = a.fgeo[4 * fc];
            n[1] = a.fgeo[4 * fc + 1];
            n[2] = a.fgeo[4 * fc + 2];
            J = a.fgeo[4 * fc + 3];
        }
#pragma unroll
        for (int i = 0; i < GP; ++i) wg[i] = on ? a.wf[min(m * GP + i, QF - 1)] : 0.0;
        // split: mode 0 (the cell means, entries 0..4 in register 0 of lanes 0..4) in FP64 to every
        // lane of the row, the rest rounded to FP32
        double c0l[5] = {0, 0, 0, 0, 0}, c0r[5] = {0, 0, 0, 0, 0};
        if constexpr (SPLIT) {
            static_for<5>([&](auto VC) __attribute__((always_inline)) {
                constexpr int v = decltype(VC)::value;
                c0l[v] = row_bcast<v>(ql[0]);
                c0r[v] = row_bcast<v>(qr[0]);
            });
        }
        T cl[CR], cr[CR];
#pragma unroll
        for (int r = 0; r < CR; ++r) {
            const bool keep = !SPLIT || r * 16 + m >= 5;
            cl[r] = keep ? (T)ql[r] : (T)0;
            cr[r] = keep ? (T)qr[r] : (T)0;
        }
        const T* PL = sPhi + (cL / 3) * N * QS;
        const T* PR = sPhi + (cR / 3) * N * QS;
        int gl[GP], gr[GP];
        T UL[GP][5], UR[GP][5], pl[GP], pr[GP];
#pragma unroll
        for (int i = 0; i < GP; ++i) {
            const int g = min(m * GP + i, QF - 1);
            gl[i] = sPerm[(cL % 3) * QF + g];
            gr[i] = sPerm[(cR % 3) * QF + g];
#pragma unroll
            for (int v = 0; v < 5; ++v) UL[i][v] = UR[i][v] = (T)0;
        }
        // trace sums U_side(g)[v] = sum_k phi_side[k](g) c_side[k][v]: entry 5k + v of each block
        // comes from lane (5k + v) % 16, register (5k + v) / 16, as the DPP operand of the FMA
        static_for<N>([&](auto KC) __attribute__((always_inline)) {
            constexpr int k = decltype(KC)::value;
            if constexpr (!SPLIT || k > 0) {
                // empty asm statements redefining this mode's table offsets: its table reads cannot
                // be hoisted above the previous mode's products (hoisting all of them spills)
#pragma unroll
                for (int i = 0; i < GP; ++i) asm volatile("" : "+v"(gl[i]), "+v"(gr[i]));
#pragma unroll
                for (int i = 0; i < GP; ++i) {
                    pl[i] = PL[k * QS + gl[i]];
                    pr[i] = PR[k * QS + gr[i]];
                }
                constexpr int r0 = (5 * k) / 16, r1 = (5 * k + 1) / 16, r2 = (5 * k + 2) / 16;
                constexpr int r3 = (5 * k + 3) / 16, r4 = (5 * k + 4) / 16;
                dpp_mode<k % 16, T>(UL, cl[r0], cl[r1], cl[r2], cl[r3], cl[r4], pl[0], pl[1]);
                dpp_mode<k % 16, T>(UR, cr[r0], cr[r1], cr[r2], cr[r3], cr[r4], pr[0], pr[1]);
            }
        });
        // the next tile's face and coefficients (into ql / qr, free now)
        if constexpr (SH::PREFETCH) fetch(tile + gridDim.x);
        double fo[GP][5];
#pragma unroll
        for (int i = 0; i < GP; ++i)
#pragma unroll
            for (int v = 0; v < 5; ++v) fo[i][v] = 0.0;
        if (on) {
#pragma unroll
            for (int i = 0; i < GP; ++i) {
                const int g = m * GP + i;
                if (g >= QF) continue;
                double uL[5], Ub[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    uL[v] = c0l[v] + (double)UL[i][v];
                    Ub[v] = c0r[v] + (double)UR[i][v];
                }
                // outer state: the neighbour's trace, or the ghost state of a Dirichlet (2 U_D - U)
                // or Pseudo3DZ (mirrored z momentum) face; one Riemann solve per point
                if (kc == 2) {
                    if (a.fbc[fc] == FVDG_BC_DIRICHLET) {
                        const double* ud = a.UD + ((int64_t)a.fdir[fc] * QF + g) * 5;
#pragma unroll
                        for (int v = 0; v < 5; ++v) Ub[v] = ud[v] + (ud[v] - uL[v]);
                    } else {
#pragma unroll
                        for (int v = 0; v < 5; ++v) Ub[v] = uL[v];
                        Ub[3] = -uL[3];
                    }
                }
                if (kc != 0) {
                    if (SPLIT)
                        num_flux_mx(a.flux, uL, Ub, n, a.gamma, fo[i]);
                    else
                        num_flux(a.flux, uL, Ub, n, a.gamma, fo[i]);
                }
            }
        }
        // FP32 output: every lane of the row takes point 0's flux from lane 0 (all lanes active)
        double ref[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if constexpr (SPLIT) {
#pragma unroll
            for (int v = 0; v < 5; ++v) ref[v] = row_bcast<0>(fo[0][v]);
        }
        if (on) {
            if (SPLIT && m == 0)
#pragma unroll
                for (int v = 0; v < 5; ++v) a.Fref[fc * 5 + v] = ref[v] * J;
#pragma unroll
            for (int i = 0; i < GP; ++i) {
                const int g = m * GP + i;
                if (g >= QF) continue;
                const double sc = wg[i] * J;
                T* o = a.Fh + (fc * QF + g) * 5;
#pragma unroll
                for (int v = 0; v < 5; ++v) o[v] = (T)((fo[i][v] - ref[v]) * sc);
            }
        }
    }
}

template <typename T>
struct FaceTestArgs {
    int64_t E;
    int parts;
    const double* U;      // stage input state
    const double* Rv;     // volume part of the residual (nullable: none)
    double* R;            // residual output (FUSE = false)
    const double* cgeo;
    const int32_t* cfaces;
    const int32_t* fcell;
    const uint8_t* fcombo;
    // fused SSP-RK stage: out = a0*U0 + b0*(U + dt_e*(-R/M)), M = |det| * mhat[k],
    // dt_e = edt[e] when edt is set, else dt
    const double* U0;
    double* out;
    const double* mhat;
    double a0, b0, dt;
    const double* edt;
    const T* Fh;
    const double* Fref;   // F*5 (FP32 path only)
    const double* IC;     // 8*N: sum_g w_g PhiC[c][g][j] (FP32 path only)
    const T* PhiC;
    const uint8_t* canon;
    const int32_t* perm;
    int64_t ntiles;
    const uint8_t* cmeta;  // E*4: vertex combination of the slot's side | 0x80 on boundary faces
    const T* PhiB;         // k_face_test_mfma: canonical tables as per-lane MFMA B operands
    int* bad;              // FUSE: set to 1 when an output is non-finite or a mean density <= 0 (nullable)
};

template <int ORDER>
struct FaceTestShape {
    static constexpr int N = Ord<ORDER>::N, QF = Ord<ORDER>::QF;
    static constexpr int THREADS = 256;
    static constexpr int TE = THREADS / N > 32 ? 32 : THREADS / N;  // elements per tile
};

// k_face_test_lds<ORDER,FUSE,T>: the test pass.  The four local faces are staged one
// round at a time (TE*Qf*5 values of LDS): measured faster than staging all four
// at once (smaller footprint, more blocks per compute unit).
template <int ORDER, bool FUSE, typename T>
__global__ __launch_bounds__(256) void k_face_test_lds(FaceTestArgs<T> a) {
    using SH = FaceTestShape<ORDER>;
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    constexpr int N = SH::N, QF = SH::QF, NV = 5 * N, TE = SH::TE, NT = SH::THREADS;
    __shared__ T sPhi[8 * QF * N];
    __shared__ T sF[TE * QF * 5];
    __shared__ double sIC[SPLIT ? 8 * N : 1], sRefF[SPLIT ? TE * 5 : 1];
    __shared__ int sPerm[3 * QF];
    __shared__ int sCanon[24];
    __shared__ int sFace[TE * 4], sCb[TE * 4];

    const int t = threadIdx.x;
    for (int i = t; i < 8 * QF * N; i += NT) sPhi[i] = a.PhiC[i];
    for (int i = t; i < 3 * QF; i += NT) sPerm[i] = a.perm[i];
    if (SPLIT)
        for (int i = t; i < 8 * N; i += NT) sIC[i] = a.IC[i];
    if (t < 24) sCanon[t] = a.canon[t];
    const bool do_int = (a.parts & FVDG_PART_INTERNALS) != 0, do_bnd = (a.parts & FVDG_PART_BOUNDARYS) != 0;

    for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        const int64_t e_base = tile * TE;
        __syncthreads();
        if (t < TE * 4) {
            const int64_t e = e_base + t / 4;
            int f = -1, cb = 0;
            if (e < a.E) {
                const int32_t slot = a.cfaces[e * 4 + (t & 3)];
                const int ff = slot >> 1, side = slot & 1;
                const bool bnd = a.fcell[2 * ff + 1] < 0;
                if ((bnd && do_bnd) || (!bnd && do_int)) f = slot;
                cb = sCanon[a.fcombo[2 * ff + side]];
            }
            sFace[t] = f;
            sCb[t] = cb;
        }
        const int el = t / N, j = t - el * N;
        const bool active = t < TE * N && e_base + el < a.E;
        T r[5] = {0, 0, 0, 0, 0};
        double rc[5] = {0, 0, 0, 0, 0};
#pragma unroll 1
        for (int lf = 0; lf < 4; ++lf) {
            __syncthreads();
            stage_to_lds<TE * QF * 5, NT>(
                t,
                [&](int i) {
                    const int e2 = i / (QF * 5), gv = i - e2 * QF * 5;
                    const int slot = sFace[e2 * 4 + lf];
                    if (slot < 0) return (T)0;
                    const T x = a.Fh[(int64_t)(slot >> 1) * QF * 5 + gv];
                    return (slot & 1) ? -x : x;
                },
                [&](int i, T x) { sF[i] = x; });
            if (SPLIT)
                for (int i = t; i < TE * 5; i += NT) {
                    const int slot = sFace[(i / 5) * 4 + lf];
                    double x = 0.0;
                    if (slot >= 0) {
                        x = a.Fref[(int64_t)(slot >> 1) * 5 + i % 5];
                        if (slot & 1) x = -x;
                    }
                    sRefF[i] = x;
                }
            __syncthreads();
            if (active && sFace[el * 4 + lf] >= 0) {
                const int cb = sCb[el * 4 + lf];
                const T* P = sPhi + (cb / 3) * QF * N + j;
                const int* pm = sPerm + (cb % 3) * QF;
                const T* fh = sF + el * QF * 5;
#pragma unroll 4
                for (int g = 0; g < QF; ++g) {
                    const T p = P[pm[g] * N];
#pragma unroll
                    for (int v = 0; v < 5; ++v) r[v] += p * fh[g * 5 + v];
                }
                if (SPLIT) {
                    const double ic = sIC[(cb / 3) * N + j];
#pragma unroll
                    for (int v = 0; v < 5; ++v) rc[v] += ic * sRefF[el * 5 + v];
                }
            }
        }
        if (active) {
            const int64_t e = e_base + el;
            const int64_t i0 = e * NV + 5 * j;
            if (FUSE) {
                const double M = a.cgeo[e * 10 + 9] * a.mhat[j];
                bool badl = false;
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    const double Rt = a.Rv[i0 + v] + (rc[v] + (double)r[v]);
                    const double sv = a.U[i0 + v] + (a.edt ? a.edt[e] : a.dt) * (-Rt / M);
                    const double o = (a.a0 == 0.0) ? a.b0 * sv : a.a0 * a.U0[i0 + v] + a.b0 * sv;
                    a.out[i0 + v] = o;
                    badl |= stage_value_bad(o, 5 * j + v);
                }
                if (badl && a.bad) a.bad[0] = 1;
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    const double rf = rc[v] + (double)r[v];
                    a.R[i0 + v] = a.Rv ? a.Rv[i0 + v] + rf : rf;
                }
            }
        }
    }
}

// ------------------------------------------------------------------ face test on the matrix cores
// k_face_test_mfma<ORDER,FUSE,T>: the test pass of k_face_test_lds as 16x16x4 MFMA products in T.
// One wavefront = 16 elements (the A rows); the K index is the canonical face point g' of
// one canonical table c, the columns are test functions j:
//   R[e][j][v] += sum_slots sum_g' Fh_slot[pinv_m(g')][v] * PhiC[c][g'][j]
// where slot s of element e sees its face through table c(e,s) and rotation m(e,s), and pinv_m
// is the inverse point permutation of that rotation (face point -> canonical point is perm_m).
// Per slot the tile's elements use few tables (two on a conforming tet mesh: the face opposite
// vertex s in either orientation), so for each table in use (wave-uniform ballot) the product
// runs with the A rows of elements on another table set to zero.  The B operands (tables, zero
// padded to whole 16x4 tiles) are stored per lane, PhiB[c][ks][jt][lane]: one coalesced load per
// product, served by the caches.  The A operands come straight from the face-flux array (each
// lane's 5 variables of a point are contiguous), with the side's sign.  The products accumulate
// in registers; the wavefront then writes its 16 x 5N block to LDS in element-major order and
// the epilogue (R_vol added, fused SSP-RK stage update exactly as k_face_test_lds) reads and
// writes the state arrays in contiguous runs.  With T = float the faces' FP64 reference fluxes
// enter in the epilogue in FP64, sum_slots IC[c][j] * (+-Fref[v]), from a per-tile LDS table.
template <int ORDER, typename T = double>
struct FaceTestMfmaShape {
    static constexpr int N = Ord<ORDER>::N, QF = Ord<ORDER>::QF, NV = 5 * N;
    static constexpr int KS = (QF + 3) / 4, JT = (N + 15) / 16;
    // the epilogue block (16 x 5N values per wavefront) sets the LDS per block: from P4 on, one
    // wavefront per block lets more of them share a compute unit
    static constexpr int WAVES = N > 20 ? 1 : 4, THREADS = 64 * WAVES, TE = 16;
    static constexpr int OUT = TE / 2 * NV;  // values of LDS per wavefront: half a tile at a time
    static constexpr int MIN_WAVES = N == 20 ? 4 : 3;  // waves per SIMD the registers are held to (P3: 4, measured)
    static constexpr bool SPLIT = !std::is_same<T, double>::value;
    static constexpr int REF = SPLIT ? TE * 4 * 5 : 1;  // reference fluxes per wavefront
    // P5's accumulators (4 column tiles x 5 variables) do not fit the registers: P <= 4 in FP64;
    // in FP32 up to P3 (at P4 the LDS kernel measured faster: 1.2 vs 2.1 ms at 384k tets)
    static constexpr bool SUPPORTED =
        N <= (SPLIT ? 20 : 35) && WAVES * (OUT * sizeof(T) + REF * 8 + TE * 4 * 4) <= 96 * 1024;
};

// LDS ordering between the lanes of one wavefront (the waves of a block run different tile
// counts, so a block barrier cannot be used inside the tile loop)
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int ORDER, bool FUSE, typename T>
__global__ __launch_bounds__(256, (FaceTestMfmaShape<ORDER, T>::MIN_WAVES)) void k_face_test_mfma(FaceTestArgs<T> a) {
    using SH = FaceTestMfmaShape<ORDER, T>;
    using M = Mfma<T>;
    typedef typename M::acc_t acc_t;
    constexpr int N = SH::N, QF = SH::QF, NV = SH::NV, KS = SH::KS, JT = SH::JT;
    constexpr bool SPLIT = SH::SPLIT;
    __shared__ T sOut[SH::WAVES * SH::OUT];
    __shared__ double sRef[SH::WAVES * SH::REF];
    __shared__ int sSlotC[SPLIT ? SH::WAVES * SH::TE * 4 : 1];
    __shared__ int sPinv[3 * QF];
    __shared__ int sCanon[24];
    const int t = threadIdx.x, l = t & 63, wave = t >> 6, row = l & 15, kq = l >> 4;
    for (int i = t; i < 3 * QF; i += SH::THREADS) {
        const int m = i / QF, g = i - m * QF;
        sPinv[m * QF + a.perm[i]] = g;
    }
    if (t < 24) sCanon[t] = a.canon[t];
    __syncthreads();
    const bool do_int = (a.parts & FVDG_PART_INTERNALS) != 0, do_bnd = (a.parts & FVDG_PART_BOUNDARYS) != 0;
    T* so = sOut + wave * SH::OUT;
    double* sr = sRef + wave * SH::REF;
    int* sc = sSlotC + (SPLIT ? wave * SH::TE * 4 : 0);

    bool badl = false;  // FUSE: a stage output of this lane failed stage_value_bad
    for (int64_t tile = (int64_t)blockIdx.x * SH::WAVES + wave; tile < a.ntiles;
         tile += (int64_t)gridDim.x * SH::WAVES) {
        const int64_t e0 = tile * SH::TE, er = e0 + row;
        acc_t acc[5][JT];
#pragma unroll
        for (int v = 0; v < 5; ++v)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) acc[v][jt] = (acc_t){0, 0, 0, 0};
        // the four slots of this lane's A-row element (one 16-byte and one 4-byte load)
        int4 cf4 = make_int4(-1, -1, -1, -1);
        uchar4 cm4 = make_uchar4(0, 0, 0, 0);
        if (er < a.E) {
            cf4 = *reinterpret_cast<const int4*>(a.cfaces + er * 4);
            cm4 = *reinterpret_cast<const uchar4*>(a.cmeta + er * 4);
        }
        if (SPLIT) wave_sync();  // the previous tile's epilogue reads of sr / sc are done
#pragma unroll 1
        for (int s = 0; s < 4; ++s) {
            // this lane's A row: element er, slot s
            int cl = -1, m = 0;
            int64_t f = 0;
            T sign = 1;
            if (er < a.E) {
                const int32_t slot = s == 0 ? cf4.x : s == 1 ? cf4.y : s == 2 ? cf4.z : cf4.w;
                const int meta = s == 0 ? cm4.x : s == 1 ? cm4.y : s == 2 ? cm4.z : cm4.w;
                const bool bnd = (meta & 0x80) != 0;
                if ((bnd && do_bnd) || (!bnd && do_int)) {
                    const int cb = sCanon[meta & 0x1f];
                    cl = cb / 3;
                    m = cb - 3 * cl;
                    f = slot >> 1;
                    sign = (slot & 1) ? -1 : 1;
                }
            }
            if (SPLIT && kq == 0) {
                // this slot's FP64 reference flux (signed) and table, for the epilogue
                sc[row * 4 + s] = cl;
#pragma unroll
                for (int v = 0; v < 5; ++v) sr[(row * 4 + s) * 5 + v] = cl >= 0 ? (double)sign * a.Fref[f * 5 + v] : 0.0;
            }
            // tables in use by the tile's elements at this slot (wave-uniform)
            unsigned used = 0;
#pragma unroll
            for (int c = 0; c < 8; ++c) used |= (__ballot(cl == c) != 0 ? 1u : 0u) << c;
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int gc = 4 * ks + kq;
                T x[5];
                if (cl >= 0 && gc < QF) {
                    const T* p = a.Fh + (f * QF + sPinv[m * QF + gc]) * 5;
#pragma unroll
                    for (int v = 0; v < 5; ++v) x[v] = sign * p[v];
                } else {
#pragma unroll
                    for (int v = 0; v < 5; ++v) x[v] = 0;
                }
                for (unsigned u = used; u != 0; u &= u - 1) {
                    const int c = __builtin_ctz(u);
                    const bool mine = cl == c;
                    T b[JT];
#pragma unroll
                    for (int jt = 0; jt < JT; ++jt) b[jt] = a.PhiB[((c * KS + ks) * JT + jt) * 64 + l];
#pragma unroll
                    for (int v = 0; v < 5; ++v) {
                        const T av = mine ? x[v] : (T)0;
#pragma unroll
                        for (int jt = 0; jt < JT; ++jt) acc[v][jt] = M::mma(av, b[jt], acc[v][jt]);
                    }
                }
            }
        }
        // accumulator element r of this lane: tile element row(r, kq), test function 16 jt + (l & 15);
        // the epilogue runs on 8 elements at a time (half the LDS per wavefront)
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            wave_sync();  // the previous half's (or tile's) epilogue reads are done
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int rw = M::row(r, kq) - 8 * h;
                if (rw < 0 || rw >= 8) continue;
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) {
                    const int j = 16 * jt + row;
                    if (j < N)
#pragma unroll
                        for (int v = 0; v < 5; ++v) so[rw * NV + 5 * j + v] = acc[v][jt][r];
                }
            }
            wave_sync();
            const int64_t eh = e0 + 8 * h;
            const int64_t nvalid = max<int64_t>(0, min<int64_t>(8, a.E - eh)) * NV;
            for (int i = l; i < nvalid; i += 64) {
                const int ee = i / NV, kv = i - ee * NV, j = kv / 5;
                const int64_t e = eh + ee, gi = eh * NV + i;
                double rf = (double)so[i];
                if (SPLIT) {
                    const int v = kv - 5 * j, et = 8 * h + ee;
#pragma unroll
                    for (int s = 0; s < 4; ++s) {
                        const int c = sc[et * 4 + s];
                        if (c >= 0) rf += a.IC[c * N + j] * sr[(et * 4 + s) * 5 + v];
                    }
                }
                if (FUSE) {
                    const double Mj = a.cgeo[e * 10 + 9] * a.mhat[j];
                    const double dte = a.edt ? a.edt[e] : a.dt;
                    const double Rt = a.Rv[gi] + rf;
                    const double sv = a.U[gi] + dte * (-Rt / Mj);
                    const double o = (a.a0 == 0.0) ? a.b0 * sv : a.a0 * a.U0[gi] + a.b0 * sv;
                    a.out[gi] = o;
                    badl |= stage_value_bad(o, kv);
                } else {
                    a.R[gi] = a.Rv ? a.Rv[gi] + rf : rf;
                }
            }
        }
    }
    if (FUSE && badl && a.bad) a.bad[0] = 1;
}
