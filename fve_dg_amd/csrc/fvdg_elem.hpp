// k_elem_mfma<ORDER,T,FUSE>: one kernel per element tile for everything after the face-flux pass:
// the volume term, the face test functions and (FUSE) the SSP-RK stage update.
//
// One wavefront owns 16 consecutive elements.  All contractions accumulate into ONE set of
// matrix-core accumulators R_v[j][e] (rows = test functions j, columns = the tile's elements):
//   volume interpolation  U_v[q][e] = sum_k V[q][k] C_v[k][e]                 (16x16x4 products)
//   volume test           R_v[j][e] += sum_(q,d) G[q][d][j] B_v[(q,d)][e]
//                         B_v[(q,d)][e] = -w_q |J_e| (F(U_e(xi_q)) A_e^{-1})_(v,d), pointwise, FP64
//   face test             R_v[j][e] += sum_slots sum_g' PhiC_c[g'][j] Fh_slot[pinv_m(g')][v]
// The face test has the canonical face table as the A operand (per-lane layout PhiB[c][ks][jt][lane]:
// table point g' = 4 ks + lane/16, test function j = 16 jt + lane%16) and the face fluxes of the
// tile's elements as the B operand (column = element, K = canonical point g'), so its products land
// in the volume term's accumulator layout.  Each element slot sees its face through one of the eight
// canonical tables c and a point rotation m; per slot the product runs once per table in use by the
// tile (wave-uniform ballot) with the B columns of elements on another table zero.
// The volume residual therefore never goes to memory, and the stage input U is read once: its
// tile stays in LDS (FP64) for the interpolation and the update.  Each lane then turns its
// accumulators (test function j, element e) into the stage value U + dt (-R / M), M = |det A_e|
// mhat[j] (FUSE; the residual itself otherwise), in place in the LDS tile, after adding the FP64
// parts of the mixed path; a last pass over the tile in contiguous runs writes
// out = a0 U0 + b0 (stage value) (or R) with the stage-output check (non-finite value, mean
// density <= 0).
//
// T = double: FP64 matrix cores (the drop-in).  T = float (mixed): FP32 matrix cores with the
// splits of the mixed path kept in FP64 -- the interpolation adds the cell mean (mode 0) in FP64,
// the volume test runs on F(U_q) - F(c_0) with the mean flux's term -|J| (F(c_0) A^{-1}) .
// sum_q w_q grad phi_j added in FP64, and each face's FP64 reference flux enters as
// sum_slots (+-Fref) sum_g w_g PhiC_c[g][j] (table IC).
#pragma once
#include "fvdg_bf16x3.hpp"
#include "fvdg_kernels_fast.hpp"

template <typename T>
struct ElemArgs {
    int64_t E;
    double gamma;
    int parts;
    const double* U;      // stage input (FP64)
    double* R;            // residual output (FUSE = false)
    const double* cgeo;   // E*10: A^{-1}, |det A|
    const int32_t* cfaces;
    const uint8_t* cmeta;
    const T* VA;          // volume tables, zero padded MFMA operands
    const T* GA;
    const double* wq;
    const double* Gs;     // N*3 (mixed)
    const T* Fh;          // face fluxes [F][Qf][5] (mixed: differences from Fref)
    const double* Fref;   // [F][5] (mixed)
    const double* IC;     // [8][N] (mixed)
    const T* PhiB;        // canonical face tables as per-lane MFMA A operands
    const uint8_t* canon;
    const int32_t* perm;
    // fused SSP-RK stage
    const double* U0;
    double* out;
    const double* mhat;
    double a0, b0, dt;
    const double* edt;
    int* bad;
    const T* Cf;          // CF: per (face, side) test contributions [F][2][5][NP] of k_face_tile
    double phi0;          // CF, split: the constant value of phi_0 (MF in Fref: the faces' flux integrals)
    const __bf16* VB3;    // BF3: volume operand tables of k_vol_bf16x3, 3 parts each
    const __bf16* GB3;
};

template <int ORDER, typename T>
struct ElemShape {
    static constexpr int N = Ord<ORDER>::N;
    // waves per SIMD the registers are held to: the tile's LDS (16 x 5N doubles, 12.9 KB at P3) allows
    // three; the FP64 form would spill 456 bytes per lane at three (two), the FP32 form 20 (three:
    // P3 196k 0.434 vs 0.457 ms)
    static constexpr int MIN_WAVES = (N <= 10 && !std::is_same<T, double>::value) ? 3 : 2;
    static constexpr bool SUPPORTED = N <= 35;
};

// F(U0 + d) - F(U0) for the Euler flux, from the primitives of U0 (pb: rho0, u0 (3), E0 = rho0 e0,
// p0) and the perturbation d = (d rho, d m (3), d E) = entry r of the accumulators u[v], in FP32
// with every term proportional to d (no difference of two large fluxes):
//   du = (dm - u0 d rho) / rho,  u = u0 + du,  dp = (gamma - 1) (dE - (dm . u + rho0 u0 . du) / 2)
//   mass: dm_d;  momentum: dm_i u_d + rho0 u0_i du_d + dp [i = d];  energy: (dE + dp) u_d + (E0 + p0) du_d
template <typename A4>
__device__ __forceinline__ void flux_delta_f32(const float pb[6], const A4 (&u)[5], int r, float gm1, float F[5][3]) {
    const float dr = u[0][r], dm[3] = {u[1][r], u[2][r], u[3][r]}, dE = u[4][r];
    const float r0 = pb[0], u0[3] = {pb[1], pb[2], pb[3]};
    const float ri = __builtin_amdgcn_rcpf(r0 + dr);
    float du[3], uu[3], dmu = 0.0f, u0du = 0.0f;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        du[d] = (dm[d] - u0[d] * dr) * ri;
        uu[d] = u0[d] + du[d];
        dmu += dm[d] * uu[d];
        u0du += u0[d] * du[d];
    }
    const float dp = gm1 * (dE - 0.5f * (dmu + r0 * u0du));
    const float dEp = dE + dp, Ep0 = pb[4] + pb[5];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        F[0][d] = dm[d];
#pragma unroll
        for (int i = 0; i < 3; ++i) F[1 + i][d] = dm[i] * uu[d] + r0 * u0[i] * du[d] + (i == d ? dp : 0.0f);
        F[4][d] = dEp * uu[d] + Ep0 * du[d];
    }
}

// BF3 (T = float, CF): both volume contractions as the six bf16 part products of k_vol_bf16x3
// (three-part splits of the tile's modes >= 1 and of the FP32 perturbation fluxes, FP32
// accumulation); its accumulator layout is that of the FP32 16x16x4 form, so the face
// contributions and the update are unchanged.
template <int ORDER, typename T, bool FUSE, bool CF = false, bool BF3 = false>
__global__ __launch_bounds__(64, (ElemShape<ORDER, T>::MIN_WAVES)) void k_elem_mfma(ElemArgs<T> a) {
    using S = VolShape<ORDER>;
    using M = Mfma<T>;
    using acc_t = typename M::acc_t;
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    constexpr int N = S::N, Q = S::Q, NV = S::NV, QT = S::QT, KS = S::KS, JT = S::JT;
    constexpr int KPAD = S::KPAD, JPAD = S::JPAD, STR = S::STRIDE;
    constexpr int QF = Ord<ORDER>::QF, FKS = (QF + 3) / 4;
    __shared__ double sU[16 * STR];  // the tile's coefficients (FP64), later its stage values (or residual)
    __shared__ int sPinv[CF ? 1 : 3 * QF];
    __shared__ int sCanon[24];
    __shared__ double sW[16 * QT + N];  // volume weights (zero padded), then mhat
    const int l = threadIdx.x, col = l & 15, grp = l >> 4;
    const int64_t e0 = (int64_t)blockIdx.x * 16;
    const int64_t base = e0 * NV, limit = a.E * NV;
    const bool do_cells = (a.parts & FVDG_PART_CELLS) != 0;
    const bool do_int = (a.parts & FVDG_PART_INTERNALS) != 0, do_bnd = (a.parts & FVDG_PART_BOUNDARYS) != 0;
    const int64_t e = e0 + col;
    const bool live = e < a.E;
    const int64_t ec = live ? e : a.E - 1;  // clamped: every load below is unconditional

    // 1. everything the tile reads up front, issued together: the lane's element geometry and face
    //    slots, the small tables (LDS), the tile's coefficients (LDS)
    double Ai[9], det, c0[5] = {0, 0, 0, 0, 0};
    {
        const double* g = a.cgeo + ec * 10;
#pragma unroll
        for (int i = 0; i < 9; ++i) Ai[i] = live ? g[i] : 0.0;
        det = live ? g[9] : 0.0;
        if (SPLIT)
#pragma unroll
            for (int v = 0; v < 5; ++v) c0[v] = live ? a.U[ec * NV + v] : 0.0;
    }
    const int4 cf4r = *reinterpret_cast<const int4*>(a.cfaces + ec * 4);
    const uchar4 cm4r = *reinterpret_cast<const uchar4*>(a.cmeta + ec * 4);
    {
        constexpr int NW = 16 * QT + N;
        const double w0 = a.wq[min(l, 16 * QT - 1)];
        const double w1 = l + 64 < NW ? (l + 64 < 16 * QT ? a.wq[l + 64] : a.mhat[max(l + 64 - 16 * QT, 0)]) : 0.0;
        const double m0v = a.mhat[min(max(l - 16 * QT, 0), N - 1)];
        if (l < 16 * QT) sW[l] = w0;
        else if (l < NW) sW[l] = m0v;
        if (l + 64 < NW) sW[l + 64] = w1;
        if (!CF) {
            for (int i = l; i < 3 * QF; i += 64) {
                const int m = i / QF, g = i - m * QF;
                sPinv[m * QF + a.perm[i]] = g;
            }
        }
        if (l < 24) sCanon[l] = a.canon[l];
    }
    stage_to_lds<16 * NV, 64>(
        l, [&](int i) { return (base + i < limit) ? a.U[base + i] : 0.0; },
        [&](int i, double x) { sU[(i / NV) * STR + (i % NV)] = x; });
    __syncthreads();

    // 2. split: the mean state's flux times A^{-1}; CF: the mean state's primitives in FP32 for the
    //    perturbation flux (PF: every point's flux is formed as its difference from the mean's)
    constexpr bool PF = CF && SPLIT;
    double FrA[5][3];
    float pb[6] = {1, 0, 0, 0, 0, 0}, Af[9], detf = 0.0f;  // rho0, u0 (3), E0, p0
    if (PF) {
        const Prim q0 = euler_prim(c0, a.gamma);
        if (live) {
            pb[0] = (float)q0.rho;
            pb[1] = (float)q0.u;
            pb[2] = (float)q0.v;
            pb[3] = (float)q0.w;
            pb[4] = (float)c0[4];
            pb[5] = (float)q0.p;
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) Af[i] = (float)Ai[i];
        detf = (float)det;
    } else if (SPLIT) {
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

    // 3. volume term
    static_assert(!BF3 || (CF && !std::is_same<T, double>::value), "BF3: the FP32 face-tile form only");
    if constexpr (BF3) {
        using S3 = fvdg_bf3::Bf3Shape<ORDER>;
        using fvdg_bf3::bf16x8;
        constexpr int KS3 = S3::KS;
        const float gm1 = (float)(a.gamma - 1.0);
#pragma unroll 1
        for (int t = 0; t < QT; ++t) {
            // interpolation of modes >= 1: U_v[q][e] - c0_v = sum_k V[q][k] C_v[k][e]
            acc_t uacc[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) uacc[v] = (acc_t){0, 0, 0, 0};
#pragma unroll
            for (int ks = 0; ks < KS3; ++ks) {
                bf16x8 av[3];
#pragma unroll
                for (int p = 0; p < 3; ++p)
                    av[p] = *reinterpret_cast<const bf16x8*>(a.VB3 + (size_t)p * S3::VB + ((t * KS3 + ks) * 64 + l) * 8);
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    bf16x8 bv[3];
                    float xs[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const int k = 32 * ks + 8 * grp + i, kc = min(k, N - 1);
                        const double x = sU[col * STR + 5 * kc + v];
                        xs[i] = (k < N && k > 0) ? (float)x : 0.0f;
                    }
                    fvdg_bf3::split3x8(xs, bv);
                    uacc[v] = fvdg_bf3::mma6(av, bv, uacc[v]);
                    asm volatile("" ::: "memory");  // one variable's LDS reads and splits at a time
                }
            }
            // F(U_q) - F(c_0) at the lane's four points (FP32 perturbation form), times A^{-1} and -w |J|;
            // K slice 0 takes pairs 0..7 (points 0, 1 and two directions of point 2), slice 1 pairs
            // 8..11: points 0-2 are formed before slice 0's products and point 3 after them (registers)
            float Bq[5][12];
            auto point = [&](auto RC) __attribute__((always_inline)) {
                constexpr int r = decltype(RC)::value;
                const int q = 16 * t + 4 * grp + r;
                float dF[5][3];
                flux_delta_f32(pb, uacc, r, gm1, dF);
                const float jw = -(float)sW[q] * detf;  // zero weight on padding points
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
                        Bq[v][3 * r + d] = (dF[v][0] * Af[d] + dF[v][1] * Af[3 + d] + dF[v][2] * Af[6 + d]) * jw;
            };
            // test: K slices sl of (point, direction) pairs 8 sl + i of each lane group; the slice's
            // table operands once, then one variable's split fluxes at a time (registers)
            auto slice = [&](auto SL) __attribute__((always_inline)) {
                constexpr int sl = decltype(SL)::value;
                bf16x8 av[JT][3];
#pragma unroll
                for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                    for (int p = 0; p < 3; ++p)
                        av[jt][p] = *reinterpret_cast<const bf16x8*>(a.GB3 + (size_t)p * S3::GB +
                                                                      (((t * 2 + sl) * JT + jt) * 64 + l) * 8);
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    bf16x8 bv[3];
                    float xs[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) xs[i] = 8 * sl + i < 12 ? Bq[v][8 * sl + i] : 0.0f;
                    fvdg_bf3::split3x8(xs, bv);
#pragma unroll
                    for (int jt = 0; jt < JT; ++jt) racc[v][jt] = fvdg_bf3::mma6(av[jt], bv, racc[v][jt]);
                    __builtin_amdgcn_sched_barrier(0);  // one variable's splits at a time
                }
            };
            point(std::integral_constant<int, 0>{});
            point(std::integral_constant<int, 1>{});
            point(std::integral_constant<int, 2>{});
            slice(std::integral_constant<int, 0>{});
            point(std::integral_constant<int, 3>{});
            slice(std::integral_constant<int, 1>{});
        }
    } else if (do_cells) {
#pragma unroll 1
        for (int t = 0; t < QT; ++t) {
            // this point tile's interpolation operands and the first point row's test operands
            T va[KS], ga[3][JT];
#pragma unroll
            for (int s = 0; s < KS; ++s) va[s] = a.VA[(16 * t + col) * KPAD + 4 * s + grp];
#pragma unroll
            for (int d = 0; d < 3; ++d)
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) ga[d][jt] = a.GA[((16 * t + M::row(0, grp)) * 3 + d) * JPAD + 16 * jt + col];
            acc_t uacc[5];
#pragma unroll
            for (int v = 0; v < 5; ++v) uacc[v] = (acc_t){0, 0, 0, 0};
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = 4 * s + grp, kc = min(k, N - 1);
                const bool use_k = k < N && (!SPLIT || k > 0);  // split: mode 0 is added in FP64
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    const double x = sU[col * STR + 5 * kc + v];
                    uacc[v] = M::mma(va[s], use_k ? (T)x : (T)0, uacc[v]);
                }
                asm volatile("" ::: "memory");  // one K step's LDS reads in flight at a time
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = 16 * t + M::row(r, grp);
                // the next point row's test operands, in flight during this row's flux
                T gn[3][JT];
                if (r < 3) {
                    const int qn = 16 * t + M::row(r + 1, grp);
#pragma unroll
                    for (int d = 0; d < 3; ++d)
#pragma unroll
                        for (int jt = 0; jt < JT; ++jt) gn[d][jt] = a.GA[(qn * 3 + d) * JPAD + 16 * jt + col];
                }
                T B[5][3];
                if constexpr (PF) {
                    // F(U_q) - F(c_0) from the perturbation d = U_q - c_0 (the FP32 interpolation of
                    // modes >= 1) and the mean's primitives, without cancellation: all FP32
                    float dF[5][3];
                    flux_delta_f32(pb, uacc, r, (float)(a.gamma - 1.0), dF);
                    const float jw = -(float)sW[q] * detf;
#pragma unroll
                    for (int v = 0; v < 5; ++v)
#pragma unroll
                        for (int d = 0; d < 3; ++d)
                            B[v][d] = (T)((dF[v][0] * Af[d] + dF[v][1] * Af[3 + d] + dF[v][2] * Af[6 + d]) * jw);
                } else {
                    // every lane evaluates (rows q >= Q have zero weight and zero test operands)
                    double Uq[5];
#pragma unroll
                    for (int v = 0; v < 5; ++v) Uq[v] = c0[v] + (double)uacc[v][r];
                    double F[5][3];
                    euler_flux(Uq, a.gamma, F);
                    const double jw = -sW[q] * det;
#pragma unroll
                    for (int v = 0; v < 5; ++v)
#pragma unroll
                        for (int d = 0; d < 3; ++d) {
                            double fa = F[v][0] * Ai[d] + F[v][1] * Ai[3 + d] + F[v][2] * Ai[6 + d];
                            if (SPLIT) fa -= FrA[v][d];
                            B[v][d] = (T)(fa * jw);
                        }
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) {
#pragma unroll
                    for (int jt = 0; jt < JT; ++jt) {
#pragma unroll
                        for (int v = 0; v < 5; ++v) racc[v][jt] = M::mma(ga[d][jt], B[v][d], racc[v][jt]);
                    }
                }
                if (r < 3)
#pragma unroll
                    for (int d = 0; d < 3; ++d)
#pragma unroll
                        for (int jt = 0; jt < JT; ++jt) ga[d][jt] = gn[d][jt];
            }
        }
    }

    // 4. face test: the four slots of this lane's element (column col)
    const int4 cf4 = live ? cf4r : make_int4(-1, -1, -1, -1);
    const uchar4 cm4 = live ? cm4r : make_uchar4(0, 0, 0, 0);
    int slot_c[4] = {-1, -1, -1, -1};  // split: each slot's table and face (-f-1: side 1), for step 5
    int64_t slot_f[4] = {0, 0, 0, 0};
    // CF, split: the cell-mean row (j = 0, entry 0 of lane group 0) from the faces' FP64 flux
    // integrals MF (sides with opposite signs: the cell means stay conserved in FP64)
    double mf[5] = {0, 0, 0, 0, 0};
    if constexpr (CF) {
        // the four (face, side) contribution blocks of k_face_tile: lane (element col, grp) adds the
        // entries of its accumulator rows (FP32: 4 contiguous test functions per load).  Loads are
        // unconditional (face 0 for unused slots, rows clamped into the block) and their values
        // selected: a divergent branch would wait for each load.  Two slots in flight at a time.
        constexpr int NP = (N + 3) / 4 * 4;
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            if (s == 2) asm volatile("" ::: "memory");
            const int32_t slot = s == 0 ? cf4.x : s == 1 ? cf4.y : s == 2 ? cf4.z : cf4.w;
            const int meta = s == 0 ? cm4.x : s == 1 ? cm4.y : s == 2 ? cm4.z : cm4.w;
            const bool bnd = (meta & 0x80) != 0;
            const bool use = e < a.E && ((bnd && do_bnd) || (!bnd && do_int));
            const int32_t sl = use ? slot : 0;
            const T* p = a.Cf + (int64_t)sl * 5 * NP;
            if constexpr (SPLIT) {
                const double* q = a.Fref + (int64_t)(sl >> 1) * 5;
                const double sg = !use ? 0.0 : (sl & 1) ? -1.0 : 1.0;
#pragma unroll
                for (int v = 0; v < 5; ++v) mf[v] = fma(sg, q[v], mf[v]);
            }
#pragma unroll
            for (int v = 0; v < 5; ++v)
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) {
                    if constexpr (SPLIT) {
                        const int j0 = 16 * jt + 4 * grp;
                        if (16 * jt < NP) {  // rows j0 >= NP: an in-block load, discarded
                            const float4 x = *reinterpret_cast<const float4*>(p + v * NP + min(j0, NP - 4));
                            const bool in = use && j0 < NP;
                            racc[v][jt][0] += (in && j0 > 0) ? x.x : 0.0f;  // j = 0: from mf
                            racc[v][jt][1] += in ? x.y : 0.0f;
                            racc[v][jt][2] += in ? x.z : 0.0f;
                            racc[v][jt][3] += in ? x.w : 0.0f;
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int j = 16 * jt + M::row(r, grp);
                            const T y = p[v * NP + min(j, N - 1)];
                            racc[v][jt][r] += (use && j < N) ? y : (T)0;
                        }
                    }
                }
            // the sums are formed here, not sunk into the epilogue (which would keep every
            // loaded block alive until then)
            if (s & 1)
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int jt = 0; jt < JT; ++jt) asm volatile("" : "+v"(racc[v][jt]));
        }
    }
#pragma unroll 1
    for (int s = 0; s < (CF ? 0 : 4); ++s) {
        int cl = -1, m = 0;
        int64_t f = 0;
        T sign = 1;
        if (e < a.E) {
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
        if (SPLIT) {
            slot_c[s] = cl;
            slot_f[s] = sign < 0 ? -(f + 1) : f;
        }
        unsigned used = 0;
#pragma unroll
        for (int c = 0; c < 8; ++c) used |= (__ballot(cl == c) != 0 ? 1u : 0u) << c;
        // the first table's operands (all lanes), then the lane's face fluxes of this slot, loaded
        // once and unconditionally (clamped point, face 0 for an unused slot): all of the slot's
        // loads in flight together; the products below select the fluxes
        const int cfirst = used != 0 ? __builtin_ctz(used) : 0;
        T b[FKS][JT];
#pragma unroll
        for (int ks = 0; ks < FKS; ++ks)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) b[ks][jt] = a.PhiB[((cfirst * FKS + ks) * JT + jt) * 64 + l];
        T xs[FKS][5];
#pragma unroll
        for (int ks = 0; ks < FKS; ++ks) {
            const T* p = a.Fh + (f * QF + sPinv[m * QF + min(4 * ks + grp, QF - 1)]) * 5;
#pragma unroll
            for (int v = 0; v < 5; ++v) xs[ks][v] = sign * p[v];
        }
        // per table in use (the first already loaded): the products with the fluxes of the lanes
        // whose elements use that table (zero elsewhere)
        for (unsigned u = used; u != 0; u &= u - 1) {
            const int c = __builtin_ctz(u);
            const bool mine = cl == c;
            if (c != cfirst)
#pragma unroll
                for (int ks = 0; ks < FKS; ++ks)
#pragma unroll
                    for (int jt = 0; jt < JT; ++jt) b[ks][jt] = a.PhiB[((c * FKS + ks) * JT + jt) * 64 + l];
#pragma unroll
            for (int ks = 0; ks < FKS; ++ks)
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int jt = 0; jt < JT; ++jt)
                        racc[v][jt] = M::mma(b[ks][jt], (mine && 4 * ks + grp < QF) ? xs[ks][v] : (T)0,
                                             racc[v][jt]);
        }
    }

    // 5. lane (j = 16 jt + row(r, grp), element col): R_v[j] from the accumulators, plus (split) the
    //    mean flux's volume term and the faces' reference fluxes in FP64; FUSE: the stage value
    //    U + dt (-R / M) replaces U in the LDS tile (the same lane reads and writes the entry)
    double fref[4][5];
    int fc[4] = {-1, -1, -1, -1};
    if (SPLIT && !CF) {
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            fc[s] = slot_c[s];
            const int64_t fs = slot_f[s] < 0 ? -slot_f[s] - 1 : slot_f[s];
            const double sg = slot_f[s] < 0 ? -1.0 : 1.0;
#pragma unroll
            for (int v = 0; v < 5; ++v) fref[s][v] = fc[s] >= 0 ? sg * a.Fref[fs * 5 + v] : 0.0;
        }
    }
    const double dte = FUSE ? (a.edt ? (e < a.E ? a.edt[e] : 0.0) : a.dt) : 0.0;
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = 16 * jt + M::row(r, grp);
            if (j < N) {
                double rr[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) rr[v] = (double)racc[v][jt][r];
                if (CF && SPLIT && j == 0)
#pragma unroll
                    for (int v = 0; v < 5; ++v) rr[v] += a.phi0 * mf[v];
                if (SPLIT && !CF) {
                    if (do_cells) {
                        const double g0 = a.Gs[3 * j], g1 = a.Gs[3 * j + 1], g2 = a.Gs[3 * j + 2];
#pragma unroll
                        for (int v = 0; v < 5; ++v) rr[v] -= det * (FrA[v][0] * g0 + FrA[v][1] * g1 + FrA[v][2] * g2);
                    }
#pragma unroll
                    for (int s = 0; s < 4; ++s)
                        if (fc[s] >= 0) {
                            const double ic = a.IC[fc[s] * N + j];
#pragma unroll
                            for (int v = 0; v < 5; ++v) rr[v] += ic * fref[s][v];
                        }
                }
                double* p = sU + col * STR + 5 * j;
                if (FUSE) {
                    const double inv_m = e < a.E ? rcp_pos(det * sW[16 * QT + j]) : 0.0;
#pragma unroll
                    for (int v = 0; v < 5; ++v) p[v] = p[v] + dte * (-rr[v] * inv_m);
                } else {
#pragma unroll
                    for (int v = 0; v < 5; ++v) p[v] = rr[v];
                }
            }
        }
    __syncthreads();

    // 6. out = a0 U0 + b0 (stage value), or R, over the tile in contiguous runs (U^n, for
    //    a0 != 0: the loads of CH runs issued together, clamped to the tile, then their outputs)
    constexpr int NIT = (16 * NV + 63) / 64, CH = 8;
    // the lane index redefined here: the unrolled runs' index arithmetic is not hoisted above the
    // barrier (where it would hold registers through the whole kernel)
    int lo = l;
    asm volatile("" : "+v"(lo));
    const int64_t nvalid = max<int64_t>(0, min<int64_t>(16, a.E - e0)) * NV;
    const bool need_u0 = FUSE && a.a0 != 0.0;
    bool badl = false;
#pragma unroll
    for (int i0 = 0; i0 < NIT; i0 += CH) {
        double u0v[CH];
        if (need_u0)
#pragma unroll
            for (int k = 0; k < CH && i0 + k < NIT; ++k)
                u0v[k] = a.U0[base + min<int64_t>(lo + 64 * (i0 + k), nvalid - 1)];
#pragma unroll
        for (int k = 0; k < CH && i0 + k < NIT; ++k) {
            const int i = lo + 64 * (i0 + k);
            if (i < nvalid) {
                const int ee = i / NV, kv = i - ee * NV;
                const int64_t gi = base + i;
                const double x = sU[ee * STR + kv];
                if (FUSE) {
                    const double o = need_u0 ? a.a0 * u0v[k] + a.b0 * x : a.b0 * x;
                    a.out[gi] = o;
                    badl |= stage_value_bad(o, kv);
                } else {
                    a.R[gi] = x;
                }
            }
        }
    }
    if (FUSE && badl && a.bad) a.bad[0] = 1;
}
