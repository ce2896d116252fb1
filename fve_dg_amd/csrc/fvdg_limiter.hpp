// gfx950 kernels of the bound-preserving and positivity-preserving limiter of
// include/DG/DG_Schemes/PositiveLimiter.h (constructMinMax :62-119, apply_1 :135-170,
// apply_21 :171-198, apply_22 :199-259), on element-major FP64 coefficients.
//
// The limiter samples each element at the points of its order's tetrahedron rule (the
// rule the reference's drivers instantiate it with).  Work split: LPE lanes per element
// (LPE = the power of two >= Q, the rule's point count; lane q owns point q and keeps the
// basis values at its point in registers), 256 / LPE elements at a time, four such groups
// per block: the block's coefficients are staged in LDS up front with coalesced loads (more
// bytes in flight per block) and read back as broadcasts.
// Kernels are instantiated per order (compile-time modes and points: unrolled evaluations).
//
// k_limit_extrema (constructMinMax, first half): per element of the previous state,
//   ext[e][k]     = max over the rule's points of variable k,
//   ext[e][5 + k] = its value at the rule's LAST point,
// or both the average (mode 0) with neighbour averages.  The reference folds per-face
// extremes of the two neighbouring cells into both cells; a cell's bounds are therefore
// the max of ext[.][k] and the min of ext[.][5+k] over the cell and its face neighbours.
// (The reference's face minimum is min(max_U[k], min(val1, val2)), PositiveLimiter.h:98,
// which keeps only the last point's minimum; the second half of ext reproduces that.)
// k_limit_apply gathers those bounds and limits each element in LDS: apply_1 per
// variable, then the density and pressure scalings of apply_2, and writes back.
// Neither kernel reads a coefficient of another element, so U may be updated in place.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "fvdg_kernels_fast.hpp"  // Ord<ORDER>: modes and rule sizes per order

namespace fvdg_lim {

template <int ORDER>
struct LimShape {
    static constexpr int N = Ord<ORDER>::N, Q = Ord<ORDER>::Q, NV = 5 * N;
    static constexpr int LPE = Q <= 16 ? 16 : Q <= 32 ? 32 : Q <= 64 ? 64 : 128;
    static constexpr int EB = 256 / LPE;  // elements processed together (one per lane group)
    static constexpr int ROUNDS = 4;      // element groups per block, all loaded up front
    static constexpr int ET = EB * ROUNDS;  // elements per block
    static_assert(N <= LPE, "one higher mode per lane when scaling");
};

// Element-group extremes of M values at once: xor butterflies inside the wave, then (128-lane
// groups) the two waves' results combined through LDS (sx: 4 * 2M doubles).  Every thread of
// the block calls this together.
template <int LPE, int M>
__device__ __forceinline__ void grp_minmax(double (&mn)[M], double (&mx)[M], double* sx) {
    constexpr int W = LPE < 64 ? LPE : 64;
#pragma unroll
    for (int o = 1; o < W; o <<= 1) {
        double a[M], b[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            a[i] = __shfl_xor(mn[i], o, 64);
            b[i] = __shfl_xor(mx[i], o, 64);
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
            mn[i] = a[i] < mn[i] ? a[i] : mn[i];
            mx[i] = b[i] > mx[i] ? b[i] : mx[i];
        }
    }
    if constexpr (LPE > 64) {
        const int wv = threadIdx.x >> 6;
        __syncthreads();
        if ((threadIdx.x & 63) == 0)
            for (int i = 0; i < M; ++i) {
                sx[wv * 2 * M + i] = mn[i];
                sx[wv * 2 * M + M + i] = mx[i];
            }
        __syncthreads();
        for (int i = 0; i < M; ++i) {
            const double a = sx[(wv ^ 1) * 2 * M + i], b = sx[(wv ^ 1) * 2 * M + M + i];
            mn[i] = a < mn[i] ? a : mn[i];
            mx[i] = b > mx[i] ? b : mx[i];
        }
    }
}

// all 5 variables at the lane's point: u_k = sum_l phi_l c[k + 5l], l ascending
template <int N>
__device__ __forceinline__ void eval5(const double (&phi)[N], const double* c, double (&u)[5]) {
#pragma unroll
    for (int k = 0; k < 5; ++k) u[k] = 0.0;
#pragma unroll
    for (int l = 0; l < N; ++l)
#pragma unroll
        for (int k = 0; k < 5; ++k) u[k] += phi[l] * c[5 * l + k];
}

// PositiveLimiter.h:121-132
__device__ __forceinline__ double lim_pressure(const double U[5], double gamma) {
    if (U[0] <= 1e-16) return -1.0;
    const double ke = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / (2.0 * U[0]);
    return (gamma - 1.0) * (U[4] - ke);
}

template <int ORDER>
__global__ __launch_bounds__(256) void k_limit_extrema(const double* __restrict__ U, const double* __restrict__ V,
                                                       int64_t E, int only_avg, double* __restrict__ ext) {
    using S = LimShape<ORDER>;
    constexpr int N = S::N, Q = S::Q, NV = S::NV, LPE = S::LPE, EB = S::EB, ET = S::ET;
    __shared__ double sC[ET * NV];
    __shared__ double sx[4 * 10];
    const int64_t e0 = (int64_t)blockIdx.x * ET;
    const int nel = (int)(E - e0 < ET ? E - e0 : ET);
    if (only_avg) {
        for (int i = threadIdx.x; i < nel * 5; i += 256) {
            const double a = U[(e0 + i / 5) * NV + i % 5];
            ext[(e0 + i / 5) * 10 + i % 5] = a;
            ext[(e0 + i / 5) * 10 + 5 + i % 5] = a;
        }
        return;
    }
    for (int i = threadIdx.x; i < nel * NV; i += 256) sC[i] = U[e0 * NV + i];
    const int g = threadIdx.x / LPE, q = threadIdx.x % LPE;
    double phi[N];
#pragma unroll
    for (int l = 0; l < N; ++l) phi[l] = q < Q ? V[q * N + l] : 0.0;
    __syncthreads();
    for (int r = 0; r < S::ROUNDS; ++r) {
        const int el = r * EB + g;
        const bool act = el < nel && q < Q;
        double u[5], mn[5], mx[5];
        eval5<N>(phi, sC + el * NV, u);
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            mn[k] = 0.0;  // unused
            mx[k] = act ? u[k] : -DBL_MAX;
        }
        grp_minmax<LPE, 5>(mn, mx, sx);
        if (act && q == 0)
            for (int k = 0; k < 5; ++k) ext[(e0 + el) * 10 + k] = mx[k];
        if (act && q == Q - 1)
            for (int k = 0; k < 5; ++k) ext[(e0 + el) * 10 + 5 + k] = u[k];
    }
}

template <int ORDER>
__global__ __launch_bounds__(256) void k_limit_apply(double* __restrict__ U, const double* __restrict__ ext,
                                                     const int32_t* __restrict__ cnb, const double* __restrict__ V,
                                                     int64_t E, double gamma, int parts) {
    using S = LimShape<ORDER>;
    constexpr int N = S::N, Q = S::Q, NV = S::NV, LPE = S::LPE, EB = S::EB, ET = S::ET;
    __shared__ double sC[ET * NV];
    __shared__ double sB[ET * 10];  // per element: upper bounds (5), lower bounds (5)
    __shared__ double sx[4 * 10];
    const int64_t e0 = (int64_t)blockIdx.x * ET;
    const int nel = (int)(E - e0 < ET ? E - e0 : ET);
    for (int i = threadIdx.x; i < nel * NV; i += 256) sC[i] = U[e0 * NV + i];
    if (parts & 1)
        for (int i = threadIdx.x; i < nel * 5; i += 256) {
            const int64_t e = e0 + i / 5;
            const int k = i % 5;
            double hi = ext[e * 10 + k], lo = ext[e * 10 + 5 + k];
            const int4 nb4 = *reinterpret_cast<const int4*>(cnb + e * 4);  // the 4 face neighbours
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const int32_t nb = f == 0 ? nb4.x : f == 1 ? nb4.y : f == 2 ? nb4.z : nb4.w;
                if (nb >= 0) {
                    const double h = ext[(int64_t)nb * 10 + k], l = ext[(int64_t)nb * 10 + 5 + k];
                    hi = h > hi ? h : hi;
                    lo = l < lo ? l : lo;
                }
            }
            sB[(i / 5) * 10 + k] = hi;
            sB[(i / 5) * 10 + 5 + k] = lo;
        }
    const int g = threadIdx.x / LPE, q = threadIdx.x % LPE;
    double phi[N];
#pragma unroll
    for (int l = 0; l < N; ++l) phi[l] = q < Q ? V[q * N + l] : 0.0;
    __syncthreads();
    const int l1 = 1 + q;  // the higher mode this lane scales

    for (int r = 0; r < S::ROUNDS; ++r) {
        const int el = r * EB + g;
        const bool act = el < nel && q < Q;
        const bool owns = el < nel && l1 < N;
        double* c = sC + el * NV;
        // the element's values at this lane's point, evaluated once: scaling the higher
        // modes of variable k by t maps a point value u to avg + t (u - avg) (phi_0 = 1),
        // which is what the reference obtains by evaluating the scaled coefficients again
        double u[5], avg[5];
        eval5<N>(phi, c, u);
#pragma unroll
        for (int k = 0; k < 5; ++k) avg[k] = c[k];
        double th1[5] = {1.0, 1.0, 1.0, 1.0, 1.0}, thr = 1.0, thp = 1.0;
        bool lim1[5] = {false, false, false, false, false}, limr = false;
        if (parts & 1) {  // apply_1: per variable, into [lower, upper]
            double mn[5], mx[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                mn[k] = act ? u[k] : DBL_MAX;
                mx[k] = act ? u[k] : -DBL_MAX;
            }
            grp_minmax<LPE, 5>(mn, mx, sx);
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                if (el < nel) {
                    const double gmax = sB[el * 10 + k], gmin = sB[el * 10 + 5 + k];
                    if (mn[k] < gmin || mx[k] > gmax) {
                        const double tmin = (avg[k] != mn[k]) ? (avg[k] - gmin) / (avg[k] - mn[k] + 1e-32) : 0.0;
                        const double tmax = (avg[k] != mx[k]) ? (gmax - avg[k]) / (mx[k] - avg[k] + 1e-32) : 0.0;
                        const double t = tmin < tmax ? tmin : tmax;
                        th1[k] = 1.0 < t ? 1.0 : t;
                        lim1[k] = true;
                        u[k] = avg[k] + th1[k] * (u[k] - avg[k]);
                    }
                }
            }
        }
        if (parts & 2) {  // apply_21: density
            const double eps = 1e-16;
            double mn[1] = {act ? u[0] : DBL_MAX}, mx[1] = {0.0};
            grp_minmax<LPE, 1>(mn, mx, sx);
            if (el < nel && mn[0] < eps) {
                const double ravg = avg[0] > eps ? avg[0] : eps;
                const double num = ravg - eps, den = ravg - mn[0];
                thr = (den < 1e-32) ? 0.0 : (num / den < 1.0 ? num / den : 1.0);
                limr = true;
                u[0] = avg[0] + thr * (u[0] - avg[0]);
            }
            // apply_22: pressure, by bisection towards the average at each point with p < 0
            double t = 1.0;
            if (act && !(lim_pressure(u, gamma) >= 0)) {
                double lo = 0.0, hi = 1.0, mid = 0.5;
                for (int it = 0; it < 100; ++it) {
                    mid = 0.5 * (lo + hi);
                    double Um[5];
                    for (int k = 0; k < 5; ++k) Um[k] = (1.0 - mid) * avg[k] + mid * u[k];
                    if (lim_pressure(Um, gamma) < 0) hi = mid;
                    else lo = mid;
                }
                t = mid;
            }
            double tm[1] = {t}, unused[1] = {0.0};
            grp_minmax<LPE, 1>(tm, unused, sx);
            thp = tm[0];
        }
        __syncthreads();  // every lane has read c before it is scaled
        // the reference's successive in-place scalings of mode l1: apply_1, apply_21, apply_22
        if (owns)
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                double x = c[k + 5 * l1];
                if (lim1[k]) x *= th1[k];
                if (k == 0 && limr) x *= thr;
                if (parts & 2) x *= thp;
                c[k + 5 * l1] = x;
            }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nel * NV; i += 256) U[e0 * NV + i] = sC[i];
}

}  // namespace fvdg_lim
