// k_face_tile<ORDER,T>: the face pass with both face contractions on the matrix cores.
//
// The faces are grouped by class = (vertex combination of side 0) x (vertex combination of
// side 1, or "boundary"): within a class every face sees both neighbours through the same two
// reference tables Phi_c[g][k] = phi_k(xi_c(g)) (face point g, mode k), so a tile of 16 faces of
// one class is a pair of dense products with a wave-uniform table operand.  One wavefront owns a
// tile (16 faces, columns of every product):
//   1. the two neighbours' coefficient blocks are staged in LDS (each 16-lane row loads 128
//      contiguous bytes of one block per load, every load issued before the first store;
//      T = float: modes >= 1 rounded to FP32, the cell means kept in FP64 beside them);
//   2. traces   U_s[g][f] = sum_k Phi_s[g][k] c_s(f)[k]           16x16x4 products, M = points
//      (two 16-row blocks for the 24 points, rows ordered so that every lane holds six valid
//      points: the 4 + 2 accumulator entries of slot s = 0..5 below);
//   3. one Riemann solve per point (LF or Roe; FP64 Euler algebra; T = float: the cell means are
//      added in FP64 and the upwind dissipation runs in FP32 as in flux_roe_mx), ghost states
//      on Dirichlet / Pseudo3DZ faces, and dF_g = Fhat_g w_g 2A (T = double) or
//      dF_g = (Fhat_g - F(m_0) n) w_g 2A (T = float: the difference from the normal flux of
//      side 0's cell mean m_0, which is what FP32 rounds);
//   4. test     C_s[j][f] = +- sum_g Phi_s[g][j] dF_g[f]            16x16x4 products, M = test
//      functions, K = the lanes' own six points (lane group grp supplies the point of its
//      slot ks at K step ks, so the Riemann results feed the product without data movement);
//      T = float, side 1: its differences are from its own mean, F(m_1) n, which adds
//      (F(m_0) - F(m_1)) n 2A sum_g w_g Phi_1[g][j] (per-combination table ICc);
//   5. C_s is stored per (face, side) as [f][side][v][NP] (NP = N rounded up to 4), side 1
//      negated; T = float also stores MF[f][v] = sum_g Fhat_g w_g 2A in FP64.
// The element kernel (k_elem_mfma with CF = true) adds the four (face, side) blocks of each
// element.  T = float: with every difference taken from the cell's own mean flux F(m), the terms
// left out -- |J| (F(m) A^{-1}) . sum_q w_q grad phi_j in the volume, sum_faces F(m) n 2A
// sum_g w_g phi_j on the faces -- are the two sides of the divergence theorem for the constant
// flux F(m) and cancel (to the accuracy of the quadrature rules, which integrate both
// polynomials exactly); the cell-mean row (j = 0, phi_0 constant) takes the FP64 face sums
// +-phi_0 MF instead of its FP32 entries, so that the cell means stay conserved in FP64.
#pragma once
#include "fvdg_bf16x3.hpp"
#include "fvdg_kernels_fast.hpp"

// Lane layouts of the face tile.  Lane l = (col = l % 16: face, grp = l / 16) holds six face
// points, slot s = 0..5: slots 0..3 are accumulator entries r = s of the first point block,
// slots 4, 5 entries r = 0, 1 of the second.  row_point(mb, i): face point of accumulator row i
// of point block mb (-1: padding row).  jrow(r, grp): test-function row of accumulator entry r.
template <typename T>
struct TileLayout;
template <>
struct TileLayout<float> {  // accumulator entry r of lane group grp is row 4 grp + r
    __host__ __device__ static constexpr int point(int s, int grp) { return s < 4 ? 4 * grp + s : 16 + 2 * grp + (s - 4); }
    __host__ __device__ static constexpr int row_point(int mb, int i) {
        return mb == 0 ? i : (i % 4 < 2 ? 16 + 2 * (i / 4) + i % 4 : -1);
    }
    __host__ __device__ static constexpr int jrow(int r, int grp) { return 4 * grp + r; }
};
template <>
struct TileLayout<double> {  // accumulator entry r of lane group grp is row 4 r + grp
    __host__ __device__ static constexpr int point(int s, int grp) { return 4 * s + grp; }
    __host__ __device__ static constexpr int row_point(int mb, int i) { return mb == 0 ? i : (i < 8 ? 16 + i : -1); }
    __host__ __device__ static constexpr int jrow(int r, int grp) { return 4 * r + grp; }
};

template <int ORDER>
struct TileShape {
    static constexpr int N = Ord<ORDER>::N, QF = Ord<ORDER>::QF, NV = 5 * N;
    static constexpr int KS = (N + 3) / 4;    // K steps of the trace product
    static constexpr int JT = (N + 15) / 16;  // test-function blocks of the test product
    static constexpr int NP = (N + 3) / 4 * 4;
    static constexpr int STR = NV | 1;        // odd LDS row stride of a staged block
    static constexpr int CR = (NV + 15) / 16; // 16-value loads per block and row
    static constexpr bool SUPPORTED = QF == 24 && N >= 10 && N <= 35;
    static constexpr int CLASSES = 24 * 25;   // side-0 combination x (side-1 combination or boundary)
};

template <typename T>
struct FaceTileArgs {
    int64_t ntiles;
    double gamma;
    int flux;
    const double* U;
    const int4* trec;       // ntiles*16 column records: face (-1 padding), side-0 cell, side-1 cell (-1 on
                            // the boundary), kind | pseudo3dz << 2 | (Dirichlet slot + 1) << 3
    const int32_t* tclass;  // ntiles: comb0 * 25 + (comb1, or 24 on boundary faces)
    const double* fgeo;
    const double* UD;
    const double* wf;
    const T* At;      // [24][2][KS][64] trace operands: Phi_c[row_point(mb, lane%16)][4 ks + lane/16]
    const T* Ae;      // [24][JT][6][64] test operands: Phi_c[point(ks, lane/16)][16 jt + lane%16]
    const T* ICc;     // [24][NP] sum_g w_g Phi_c[g][j] (T = float)
    T* Cf;            // [F][2][5][NP]
    double* MF;       // [F][5] sum_g Fhat_g w_g 2A (T = float)
    const __bf16* At3;  // BF3 faces: [3 parts][24][2][KS3][64][8] trace operands
    const __bf16* Ae3;  // BF3 faces: [3 parts][24][JT][64][8] test operands
};

// BF3 (T = float): both face contractions as the six bf16 part products of three-part splits
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation; the accumulator layout is that of the FP32
// 16x16x4 form, so the Riemann solves and the stores are unchanged): the traces with K = the
// lane group's 8 modes of each 32-mode step, the test with K = the lane group's six point slots
// (and two zeros), each lane's six differences split once for both sides.
template <int ORDER, typename T, bool BF3>
__device__ __forceinline__ void face_tile_body(const FaceTileArgs<T>& a) {
    using SH = TileShape<ORDER>;
    using M = Mfma<T>;
    using LY = TileLayout<T>;
    using acc_t = typename M::acc_t;
    constexpr bool SPLIT = !std::is_same<T, double>::value;
    constexpr int N = SH::N, QF = SH::QF, NV = SH::NV, KS = SH::KS, JT = SH::JT, NP = SH::NP, STR = SH::STR;
    constexpr int CR = SH::CR;
    static_assert(SH::SUPPORTED, "face tiles need the 24-point face rule");
    __shared__ T sC[32 * STR];                   // block b = side * 16 + column
    __shared__ double sC0[SPLIT ? 32 * 5 : 1];   // T = float: the cell means in FP64
    __shared__ int32_t sCell[32];
    const int l = threadIdx.x, col = l & 15, grp = l >> 4;
    const int64_t tile = blockIdx.x;
    const int cls = a.tclass[tile];
    const int cb0 = cls / 25, cb1 = cls - 25 * cb0;
    const bool two = cb1 < 24;  // wave-uniform: interior tiles have two sides

    // 0. this column's face record: kind 0 (padding, or a boundary type without face flux),
    //    1 interior, 2 Dirichlet / Pseudo3DZ boundary
    const int4 rec = a.trec[tile * 16 + col];
    const int32_t f = rec.x;
    const int kind = rec.w & 3;
    if (grp == 0) {
        // side 0's cell of every face (a boundary face without a flux still takes its cell's mean
        // flux reference, which the element kernel leaves out of the cell's other terms)
        sCell[col] = f >= 0 ? rec.y : -1;
        sCell[16 + col] = kind == 1 ? rec.z : -1;
    }
    // geometry and weights of the lane's face and points, needed after the traces
    const int64_t fs = f >= 0 ? f : 0;
    const double2 g01 = *reinterpret_cast<const double2*>(a.fgeo + 4 * fs);
    const double2 g23 = *reinterpret_cast<const double2*>(a.fgeo + 4 * fs + 2);
    double wg[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) wg[s] = a.wf[LY::point(s, grp)];
    // the trace operands of both sides (boundary tiles: side 1 unused)
    T at[2][2][KS];
    if constexpr (!BF3) {
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int ks = 0; ks < KS; ++ks)
                    at[s][mb][ks] = a.At[((size_t)(s ? (two ? cb1 : cb0) : cb0) * 2 * KS + mb * KS + ks) * 64 + l];
    }
    __syncthreads();

    // 1. coefficient blocks into LDS: row grp loads blocks 4 i + grp, lane col entries 16 r + col;
    //    every load is issued (clamped to a valid address) before the first store
    //    (P4: two batches of four blocks, for the registers)
    constexpr int IB = CR <= 7 ? 8 : 4;
    int32_t cells[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) cells[i] = sCell[4 * i + grp];
#pragma unroll
    for (int i0 = 0; i0 < 8; i0 += IB) {
        double x[IB][CR];
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            const double* src = a.U + (int64_t)(cells[i0 + i] >= 0 ? cells[i0 + i] : 0) * NV;
#pragma unroll
            for (int r = 0; r < CR; ++r) x[i][r] = src[min(16 * r + col, NV - 1)];
        }
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            const int b = 4 * (i0 + i) + grp;
            const bool ok = cells[i0 + i] >= 0;
#pragma unroll
            for (int r = 0; r < CR; ++r) {
            const int kv = 16 * r + col;
            if (kv < NV) {
                const double xv = ok ? x[i][r] : 0.0;
                if (SPLIT && kv < 5) {
                    sC0[b * 5 + kv] = xv;
                    sC[b * STR + kv] = (T)0;
                } else {
                    sC[b * STR + kv] = (T)xv;
                }
            }
            }
        }
    }
    __syncthreads();

    // 2. traces: tr[s][v][mb] (rows: face points row_point(mb, .), columns: faces)
    acc_t tr[2][5][2];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int v = 0; v < 5; ++v)
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) tr[s][v][mb] = (acc_t){0, 0, 0, 0};
    constexpr int KS3 = (N + 31) / 32;
    static_assert(!BF3 || SPLIT, "BF3 face contractions: the FP32 form only");
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (s == 1 && !two) break;
        const T* B = sC + (s * 16 + col) * STR;
        if constexpr (BF3) {
            const int cbs = s ? (two ? cb1 : cb0) : cb0;
#pragma unroll
            for (int ks = 0; ks < KS3; ++ks) {
                fvdg_bf3::bf16x8 a3[2][3];
#pragma unroll
                for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                    for (int p = 0; p < 3; ++p)
                        a3[mb][p] = *reinterpret_cast<const fvdg_bf3::bf16x8*>(
                            a.At3 + (size_t)p * 24 * 2 * KS3 * 512 + ((((size_t)cbs * 2 + mb) * KS3 + ks) * 64 + l) * 8);
#pragma unroll
                for (int v = 0; v < 5; ++v) {
                    fvdg_bf3::bf16x8 bv[3];
                    float xs[8];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const int k = 32 * ks + 8 * grp + i, kc = min(k, N - 1);
                        const T bx = B[5 * kc + v];
                        xs[i] = k < N ? bx : 0.0f;
                    }
                    fvdg_bf3::split3x8(xs, bv);
                    tr[s][v][0] = fvdg_bf3::mma6(a3[0], bv, tr[s][v][0]);
                    tr[s][v][1] = fvdg_bf3::mma6(a3[1], bv, tr[s][v][1]);
                    asm volatile("" ::: "memory");  // one variable's LDS reads and splits at a time
                }
            }
            continue;
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int k = 4 * ks + grp, kc = min(k, N - 1);  // rows k >= N: zero operand (branch-free)
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const T bx = B[5 * kc + v];
                const T bv = k < N ? bx : (T)0;
                tr[s][v][0] = M::mma(at[s][0][ks], bv, tr[s][v][0]);
                tr[s][v][1] = M::mma(at[s][1][ks], bv, tr[s][v][1]);
            }
        }
    }

    // 3. Riemann solves at the lane's six points
    const double n[3] = {g01.x, g01.y, g23.x};
    const double J = f >= 0 ? g23.y : 0.0;
    const bool pseudo = (rec.w & 4) != 0;
    const int32_t dslot = (rec.w >> 3) - 1;
    double m0[5] = {0, 0, 0, 0, 0}, m1[5] = {0, 0, 0, 0, 0};
    if (SPLIT) {
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            m0[v] = sC0[col * 5 + v];
            m1[v] = two ? sC0[(16 + col) * 5 + v] : 0.0;
        }
    }
    auto point_flux = [&](auto SC, double fo[5]) __attribute__((always_inline)) {
        constexpr int s = decltype(SC)::value, mb = s < 4 ? 0 : 1, r = s < 4 ? s : s - 4;
        const int g = LY::point(s, grp);
        double uL[5], uR[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            uL[v] = m0[v] + (double)tr[0][v][mb][r];
            uR[v] = m1[v] + (double)tr[1][v][mb][r];
        }
        if (kind == 2) {
            if (!pseudo) {
                const double* ud = a.UD + ((int64_t)dslot * QF + g) * 5;
#pragma unroll
                for (int v = 0; v < 5; ++v) uR[v] = ud[v] + (ud[v] - uL[v]);
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v) uR[v] = uL[v];
                uR[3] = -uL[3];
            }
        }
        if (kind != 0) {
            if (SPLIT)
                num_flux_mx(a.flux, uL, uR, n, a.gamma, fo);
            else
                num_flux(a.flux, uL, uR, n, a.gamma, fo);
        } else {
#pragma unroll
            for (int v = 0; v < 5; ++v) fo[v] = 0.0;
        }
    };
    // T = float: the normal fluxes of the two cell means (FP64), the reference of each side
    double ref0[5] = {0, 0, 0, 0, 0}, ref1[5] = {0, 0, 0, 0, 0}, mf[5] = {0, 0, 0, 0, 0};
    if (SPLIT) {
        euler_flux_n(euler_prim(m0, a.gamma), n, ref0);
        if (two) euler_flux_n(euler_prim(m1, a.gamma), n, ref1);
        if (f < 0)
#pragma unroll
            for (int v = 0; v < 5; ++v) ref0[v] = 0.0;
    }
    // the test operands of both sides and side 1's mean-row correction, in flight during the solves
    // (FP32 at P2, P3; otherwise each side's operands are loaded at its products: registers)
    constexpr bool PRE = SPLIT && JT <= 2 && !BF3;
    const int cbs1 = two ? cb1 : cb0;
    T ae[2][JT][6], icv[JT][4];
    auto load_ae = [&](int s) __attribute__((always_inline)) {
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
#pragma unroll
            for (int ks = 0; ks < 6; ++ks) ae[s][jt][ks] = a.Ae[((size_t)(s ? cbs1 : cb0) * JT * 6 + jt * 6 + ks) * 64 + l];
    };
    if (PRE) {
        load_ae(0);
        load_ae(1);
    }
    if (SPLIT)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt)
#pragma unroll
            for (int r = 0; r < 4; ++r) icv[jt][r] = a.ICc[(size_t)cbs1 * NP + min(16 * jt + LY::jrow(r, grp), NP - 1)];
    T dF[6][5];
    static_for<6>([&](auto SC) __attribute__((always_inline)) {
        constexpr int s = decltype(SC)::value;
        double fo[5];
        point_flux(SC, fo);
        const double sc = wg[s] * J;
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            dF[s][v] = (T)((fo[v] - ref0[v]) * sc);
            if (SPLIT) mf[v] = fma(fo[v], sc, mf[v]);
        }
    });
    if (SPLIT) {
        // the face's FP64 flux integral over its 24 points: the four lane groups' partial sums
        // (every lane gets the same bits: the additions commute pairwise)
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            mf[v] += __shfl_xor(mf[v], 16, 64);
            mf[v] += __shfl_xor(mf[v], 32, 64);
        }
        if (grp == 0 && f >= 0)
#pragma unroll
            for (int v = 0; v < 5; ++v) a.MF[(int64_t)f * 5 + v] = mf[v];
    }

    // 4./5. test products per side and the stores
    // BF3: the lane's six differences of each variable split once (K entries 8 grp + slot)
    fvdg_bf3::bf16x8 d3[BF3 ? 5 : 1][3];
    if constexpr (BF3)
#pragma unroll
        for (int v = 0; v < 5; ++v) {
            float xs[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) xs[i] = i < 6 ? (float)dF[i < 6 ? i : 0][v] : 0.0f;
            fvdg_bf3::split3x8(xs, d3[v]);
        }
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (s == 1 && !two) break;
        acc_t te[5][JT];
#pragma unroll
        for (int v = 0; v < 5; ++v)
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) te[v][jt] = (acc_t){0, 0, 0, 0};
        if constexpr (BF3) {
            const int cbs = s ? cbs1 : cb0;
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                fvdg_bf3::bf16x8 a3[3];
#pragma unroll
                for (int p = 0; p < 3; ++p)
                    a3[p] = *reinterpret_cast<const fvdg_bf3::bf16x8*>(
                        a.Ae3 + (size_t)p * 24 * JT * 512 + (((size_t)cbs * JT + jt) * 64 + l) * 8);
#pragma unroll
                for (int v = 0; v < 5; ++v) te[v][jt] = fvdg_bf3::mma6(a3, d3[v], te[v][jt]);
            }
        } else {
            if (!PRE) load_ae(s);
#pragma unroll
            for (int ks = 0; ks < 6; ++ks)
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) te[v][jt] = M::mma(ae[s][jt][ks], dF[ks][v], te[v][jt]);
                }
        }
        if (SPLIT && s == 1) {
            // side 1's differences from its own mean flux: + (F(m_0) - F(m_1)) n 2A ICc[cb1][j]
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                const T d = (T)((ref0[v] - ref1[v]) * J);
#pragma unroll
                for (int jt = 0; jt < JT; ++jt)
#pragma unroll
                    for (int r = 0; r < 4; ++r) te[v][jt][r] += d * icv[jt][r];
            }
        }
        if (f >= 0) {
            T* o = a.Cf + ((int64_t)(2 * f + s) * 5) * NP;
            const T sg = s ? (T)-1 : (T)1;
#pragma unroll
            for (int v = 0; v < 5; ++v)
#pragma unroll
                for (int jt = 0; jt < JT; ++jt) {
                    if constexpr (SPLIT) {
                        const int j0 = 16 * jt + 4 * grp;  // entries r = 0..3: rows j0 .. j0 + 3
                        if (j0 < NP) {
                            float4 w;
                            w.x = sg * te[v][jt][0];
                            w.y = sg * te[v][jt][1];
                            w.z = sg * te[v][jt][2];
                            w.w = sg * te[v][jt][3];
                            *reinterpret_cast<float4*>(o + v * NP + j0) = w;
                        }
                    } else {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            const int j = 16 * jt + LY::jrow(r, grp);
                            if (j < NP) o[v * NP + j] = sg * te[v][jt][r];
                        }
                    }
                }
        }
    }
}

template <int ORDER, typename T>
__global__ __launch_bounds__(64, 2) void k_face_tile(FaceTileArgs<T> a) {
    face_tile_body<ORDER, T, false>(a);
}

// the face tile of the BF16x3 precision: both face contractions as BF16x3 products
template <int ORDER>
__global__ __launch_bounds__(64, 2) void k_face_tile_bf3(FaceTileArgs<float> a) {
    face_tile_body<ORDER, float, true>(a);
}
