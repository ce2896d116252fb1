// gfx950 volume kernel of the BF16x3 precision (FVDG_PREC_BF16X3): the volume term of the FP32
// path (k_vol_mfma<ORDER, float>: mode 0 and the mean flux F(c0) kept in FP64, the variation in
// FP32) with both contractions on the bf16 matrix cores.  Every FP32 operand x is split into three
// bf16 parts, x = x0 + x1 + x2 (24 significant bits together: the host rounds the tables' parts,
// the kernels truncate the operands' parts, split3x8), and a product a.b is formed from the six part products
// a_i.b_j with i + j <= 2 (the dropped terms are below FP32 rounding), accumulated in FP32 by
// v_mfma_f32_16x16x32_bf16.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (checked on the device): lane l holds A[l%16][8(l/16)+i]
// and B[8(l/16)+i][l%16] for i < 8; accumulator element r of lane l is D[4(l/16)+r][l%16], the layout
// of the FP32 16x16x4 form, so a lane again holds 4 points of one element after the
// interpolation.  The test product's K index is therefore laid out over (lane group, (point r,
// direction d) pair): K slice s covers the pairs p = 8s + i (p = 3r + d < 12), so each lane feeds
// its own 12 flux values.  The tables (A operands) are prepared on the host in exactly this
// per-lane order, already split into parts.
#pragma once
#include <hip/hip_runtime.h>

#include "fvdg_kernels_fast.hpp"

namespace fvdg_bf3 {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int ORDER>
struct Bf3Shape {
    static constexpr int N = Ord<ORDER>::N, Q = Ord<ORDER>::Q, NV = 5 * N;
    static constexpr int QT = (Q + 15) / 16;  // 16-point tiles
    static constexpr int KS = (N + 31) / 32;  // K steps of the interpolation (modes)
    static constexpr int JT = (N + 15) / 16;  // test-function tiles
    static constexpr int STR = NV | 1;
    // table sizes in bf16 values per part
    static constexpr int VB = QT * KS * 64 * 8;      // interpolation A operands
    static constexpr int GB = QT * 2 * JT * 64 * 8;  // test A operands (2 K slices per tile)
    static constexpr int MIN_WAVES = N == 35 ? 2 : 1;  // P4: registers held to 2 waves per SIMD
};

struct Bf3Args {
    int64_t E;
    double gamma;
    const double* U;
    double* R;
    const double* cgeo;
    const __bf16* VB;  // 3 parts x Bf3Shape::VB
    const __bf16* GB;  // 3 parts x Bf3Shape::GB
    const double* wq;  // Q (padded)
    const double* Gs;  // N*3: sum_q w_q d phi_j / d xi_d
};

// the three-part splits of eight values into b[0..2] (element i of each part vector): truncating
// splits, x0 = x with its low 16 bits cleared, x1 = the same of x - x0, x2 = the top 16 bits of
// x - x0 - x1 (both subtractions exact; x0 + x1 + x2 within 2^-24 |x| of x), two values packed per
// 32-bit word by one byte permute (the top halves of both words)
__device__ __forceinline__ void split3x8(const float (&x)[8], bf16x8 (&b)[3]) {
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    u32x4 w0, w1, w2;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t a = __float_as_uint(x[2 * k]), c = __float_as_uint(x[2 * k + 1]);
        const float ra = x[2 * k] - __uint_as_float(a & 0xffff0000u);
        const float rc = x[2 * k + 1] - __uint_as_float(c & 0xffff0000u);
        const uint32_t ua = __float_as_uint(ra), uc = __float_as_uint(rc);
        const float ta = ra - __uint_as_float(ua & 0xffff0000u);
        const float tc = rc - __uint_as_float(uc & 0xffff0000u);
        w0[k] = __builtin_amdgcn_perm(c, a, 0x07060302u);
        w1[k] = __builtin_amdgcn_perm(uc, ua, 0x07060302u);
        w2[k] = __builtin_amdgcn_perm(__float_as_uint(tc), __float_as_uint(ta), 0x07060302u);
    }
    b[0] = __builtin_bit_cast(bf16x8, w0);
    b[1] = __builtin_bit_cast(bf16x8, w1);
    b[2] = __builtin_bit_cast(bf16x8, w2);
}

// sum over the six part products with i + j <= 2
__device__ __forceinline__ f32x4 mma6(const bf16x8 (&a)[3], const bf16x8 (&b)[3], f32x4 c) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], c, 0, 0, 0);
    return c;
}

template <int ORDER>
__global__ __launch_bounds__(64, Bf3Shape<ORDER>::MIN_WAVES) void k_vol_bf16x3(Bf3Args a) {
    using S = Bf3Shape<ORDER>;
    constexpr int N = S::N, Q = S::Q, NV = S::NV, QT = S::QT, KS = S::KS, JT = S::JT, STR = S::STR;
    // the tile's coefficients (FP32), later the FP64 result in the same LDS
    __shared__ double sBuf[16 * STR];
    float* sC = reinterpret_cast<float*>(sBuf);
    double* sOut = sBuf;
    const int l = threadIdx.x, col = l & 15, grp = l >> 4;
    const int64_t e0 = (int64_t)blockIdx.x * 16;
    const int64_t base = e0 * NV, limit = a.E * NV;

    // 1. the tile's coefficients of modes >= 1 in LDS (FP32)
    for (int i = l; i < 16 * NV; i += 64)
        sC[(i / NV) * STR + (i % NV)] = (base + i < limit && i % NV >= 5) ? (float)a.U[base + i] : 0.0f;
    __syncthreads();

    // 2. this lane's element: geometry, mean state and its flux (FP64)
    const int64_t e = e0 + col;
    double Ai[9], det = 0.0, c0[5] = {0, 0, 0, 0, 0}, FrA[5][3];
    if (e < a.E) {
        const double* g = a.cgeo + e * 10;
#pragma unroll
        for (int i = 0; i < 9; ++i) Ai[i] = g[i];
        det = g[9];
#pragma unroll
        for (int v = 0; v < 5; ++v) c0[v] = a.U[e * NV + v];
    } else {
#pragma unroll
        for (int i = 0; i < 9; ++i) Ai[i] = 0.0;
        c0[0] = 1.0;  // a harmless state for the padding lanes
        c0[4] = 2.5;
    }
    {
        double F[5][3];
        euler_flux(c0, a.gamma, F);
#pragma unroll
        for (int v = 0; v < 5; ++v)
#pragma unroll
            for (int d = 0; d < 3; ++d) FrA[v][d] = F[v][0] * Ai[d] + F[v][1] * Ai[3 + d] + F[v][2] * Ai[6 + d];
    }

    f32x4 racc[5][JT];
#pragma unroll
    for (int v = 0; v < 5; ++v)
#pragma unroll
        for (int jt = 0; jt < JT; ++jt) racc[v][jt] = (f32x4){0, 0, 0, 0};

#pragma unroll 1
    for (int t = 0; t < QT; ++t) {
        // 3. interpolation: U_v[q][e] - c0_v = sum_k V[q][k] C_v[k][e] over modes k >= 1
        f32x4 uacc[5];
#pragma unroll
        for (int v = 0; v < 5; ++v) uacc[v] = (f32x4){0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            bf16x8 av[3];
#pragma unroll
            for (int p = 0; p < 3; ++p)
                av[p] = *reinterpret_cast<const bf16x8*>(a.VB + (size_t)p * S::VB + ((t * KS + ks) * 64 + l) * 8);
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                bf16x8 bv[3];
                float xs[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int k = 32 * ks + 8 * grp + i;
                    xs[i] = k < N ? sC[col * STR + 5 * k + v] : 0.0f;
                }
                split3x8(xs, bv);
                uacc[v] = mma6(av, bv, uacc[v]);
            }
        }
        // 4. pointwise flux minus the mean flux at this lane's 4 points, 3 directions
        float B[5][12];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int q = 16 * t + 4 * grp + r;
            if (q < Q) {
                double Uq[5];
#pragma unroll
                for (int v = 0; v < 5; ++v) Uq[v] = c0[v] + (double)uacc[v][r];
                double F[5][3];
                euler_flux(Uq, a.gamma, F);
                const double jw = -a.wq[q] * det;
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        const double fa = F[v][0] * Ai[d] + F[v][1] * Ai[3 + d] + F[v][2] * Ai[6 + d] - FrA[v][d];
                        B[v][3 * r + d] = (float)(fa * jw);
                    }
            } else {
#pragma unroll
                for (int v = 0; v < 5; ++v)
#pragma unroll
                    for (int d = 0; d < 3; ++d) B[v][3 * r + d] = 0.0f;
            }
        }
        // 5. test: R_v[j][e] += sum over (point, direction) of G[q][d][j] B_v[(q,d)][e], two K slices
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            bf16x8 bv[5][3];
#pragma unroll
            for (int v = 0; v < 5; ++v) {
                float xs[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) xs[i] = 8 * s + i < 12 ? B[v][8 * s + i] : 0.0f;
                split3x8(xs, bv[v]);
            }
#pragma unroll
            for (int jt = 0; jt < JT; ++jt) {
                bf16x8 av[3];
#pragma unroll
                for (int p = 0; p < 3; ++p)
                    av[p] = *reinterpret_cast<const bf16x8*>(a.GB + (size_t)p * S::GB +
                                                              (((t * 2 + s) * JT + jt) * 64 + l) * 8);
#pragma unroll
                for (int v = 0; v < 5; ++v) racc[v][jt] = mma6(av, bv[v], racc[v][jt]);
            }
        }
    }

    // 6. R = test product + the mean flux's exact part, in FP64, then a coalesced store
    __syncthreads();  // every lane is done with the coefficients in sBuf
#pragma unroll
    for (int jt = 0; jt < JT; ++jt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int j = 16 * jt + 4 * grp + r;
            if (j < N) {
                const double g0 = a.Gs[3 * j], g1 = a.Gs[3 * j + 1], g2 = a.Gs[3 * j + 2];
#pragma unroll
                for (int v = 0; v < 5; ++v)
                    sOut[col * STR + 5 * j + v] =
                        (double)racc[v][jt][r] - det * (FrA[v][0] * g0 + FrA[v][1] * g1 + FrA[v][2] * g2);
            }
        }
    __syncthreads();
    for (int i = l; i < 16 * NV; i += 64) {
        const int ee = i / NV, kv = i - ee * NV;
        if (base + i < limit) a.R[base + i] = sOut[ee * STR + kv];
    }
}

}  // namespace fvdg_bf3
