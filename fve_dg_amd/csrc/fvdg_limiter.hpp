// gfx950 kernels of the bound-preserving and positivity-preserving limiter of
// include/DG/DG_Schemes/PositiveLimiter.h (constructMinMax :62-119, apply_1 :135-170,
// apply_21 :171-198, apply_22 :199-259), on element-major FP64 coefficients.
//
// The limiter samples each element at the points of its order's tetrahedron rule (the
// rule the reference's drivers instantiate it with).  Work split: LPE lanes per element
// (LPE = the power of two >= Q, the rule's point count; lane q owns point q), 256 / LPE
// elements per block, the block's coefficients staged once in LDS with coalesced loads.
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

namespace fvdg_lim {

// max / min over the LPE lanes of an element group (all 256 threads of the block must
// call these together: the 128-lane form combines the two waves through LDS)
template <int LPE>
__device__ __forceinline__ double grp_max(double v, double* sx) {
    constexpr int W = LPE < 64 ? LPE : 64;
    for (int o = 1; o < W; o <<= 1) {
        const double u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    if constexpr (LPE > 64) {
        __syncthreads();
        sx[threadIdx.x] = v;
        __syncthreads();
        const double u = sx[threadIdx.x ^ 64];
        v = u > v ? u : v;
    }
    return v;
}
template <int LPE>
__device__ __forceinline__ double grp_min(double v, double* sx) {
    constexpr int W = LPE < 64 ? LPE : 64;
    for (int o = 1; o < W; o <<= 1) {
        const double u = __shfl_xor(v, o, 64);
        v = u < v ? u : v;
    }
    if constexpr (LPE > 64) {
        __syncthreads();
        sx[threadIdx.x] = v;
        __syncthreads();
        const double u = sx[threadIdx.x ^ 64];
        v = u < v ? u : v;
    }
    return v;
}

// value of variable k at point q: sum_l phi_l(x_q) c[k + 5l], l ascending
__device__ __forceinline__ double point_value(const double* __restrict__ V, const double* c, int N, int q, int k) {
    double v = 0.0;
    for (int l = 0; l < N; ++l) v += V[q * N + l] * c[k + 5 * l];
    return v;
}

// PositiveLimiter.h:121-132
__device__ __forceinline__ double lim_pressure(const double U[5], double gamma) {
    if (U[0] <= 1e-16) return -1.0;
    const double ke = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / (2.0 * U[0]);
    return (gamma - 1.0) * (U[4] - ke);
}

template <int LPE>
__global__ __launch_bounds__(256) void k_limit_extrema(const double* __restrict__ U, const double* __restrict__ V,
                                                       int64_t E, int N, int Q, int only_avg,
                                                       double* __restrict__ ext) {
    constexpr int EB = 256 / LPE;
    extern __shared__ double lds[];
    const int NV = 5 * N;
    double* sC = lds;
    double* sx = lds + EB * NV;
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int nel = (int)(E - e0 < EB ? E - e0 : EB);
    const int el = threadIdx.x / LPE, q = threadIdx.x % LPE;
    if (only_avg) {
        if (el < nel && q < 5) {
            const double a = U[(e0 + el) * NV + q];
            ext[(e0 + el) * 10 + q] = a;
            ext[(e0 + el) * 10 + 5 + q] = a;
        }
        return;
    }
    for (int i = threadIdx.x; i < nel * NV; i += 256) sC[i] = U[e0 * NV + i];
    __syncthreads();
    const bool act = el < nel && q < Q;
    const double* c = sC + el * NV;
    for (int k = 0; k < 5; ++k) {
        const double v = act ? point_value(V, c, N, q, k) : -DBL_MAX;
        const double mx = grp_max<LPE>(v, sx);
        if (act && q == 0) ext[(e0 + el) * 10 + k] = mx;
        if (act && q == Q - 1) ext[(e0 + el) * 10 + 5 + k] = v;
    }
}

template <int LPE>
__global__ __launch_bounds__(256) void k_limit_apply(double* __restrict__ U, const double* __restrict__ ext,
                                                     const int32_t* __restrict__ cfaces,
                                                     const int32_t* __restrict__ fcell, const double* __restrict__ V,
                                                     int64_t E, int N, int Q, double gamma, int parts) {
    constexpr int EB = 256 / LPE;
    extern __shared__ double lds[];
    const int NV = 5 * N;
    double* sC = lds;             // EB * NV coefficients
    double* sB = sC + EB * NV;    // EB * 10: upper bounds (5), lower bounds (5)
    double* sx = sB + EB * 10;    // 256: cross-wave reduction scratch
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    const int nel = (int)(E - e0 < EB ? E - e0 : EB);
    const int el = threadIdx.x / LPE, q = threadIdx.x % LPE;
    for (int i = threadIdx.x; i < nel * NV; i += 256) sC[i] = U[e0 * NV + i];
    if ((parts & 1) && el < nel && q < 5) {
        const int64_t e = e0 + el;
        double hi = ext[e * 10 + q], lo = ext[e * 10 + 5 + q];
        for (int i = 0; i < 4; ++i) {
            const int32_t cf = cfaces[e * 4 + i];
            const int32_t nb = fcell[(cf >> 1) * 2 + (1 - (cf & 1))];
            if (nb >= 0) {
                const double h = ext[(int64_t)nb * 10 + q], l = ext[(int64_t)nb * 10 + 5 + q];
                hi = h > hi ? h : hi;
                lo = l < lo ? l : lo;
            }
        }
        sB[el * 10 + q] = hi;
        sB[el * 10 + 5 + q] = lo;
    }
    __syncthreads();
    const bool act = el < nel && q < Q;
    double* c = sC + el * NV;

    if (parts & 1) {  // apply_1: per variable, into [lower, upper]
        double th[5];
        bool lim[5];
        for (int k = 0; k < 5; ++k) {
            const double v = act ? point_value(V, c, N, q, k) : 0.0;
            const double vmin = grp_min<LPE>(act ? v : DBL_MAX, sx);
            const double vmax = grp_max<LPE>(act ? v : -DBL_MAX, sx);
            th[k] = 1.0;
            lim[k] = false;
            if (el < nel) {
                const double avg = c[k], gmax = sB[el * 10 + k], gmin = sB[el * 10 + 5 + k];
                if (vmin < gmin || vmax > gmax) {
                    const double tmin = (avg != vmin) ? (avg - gmin) / (avg - vmin + 1e-32) : 0.0;
                    const double tmax = (avg != vmax) ? (gmax - avg) / (vmax - avg + 1e-32) : 0.0;
                    double t = tmin < tmax ? tmin : tmax;
                    th[k] = 1.0 < t ? 1.0 : t;
                    lim[k] = true;
                }
            }
        }
        __syncthreads();  // every lane has read c before it is scaled
        if (el < nel)
            for (int l = 1 + q; l < N; l += LPE)
                for (int k = 0; k < 5; ++k)
                    if (lim[k]) c[k + 5 * l] *= th[k];
        __syncthreads();
    }
    if (parts & 2) {  // apply_21: density
        const double eps = 1e-16;
        const double r = act ? point_value(V, c, N, q, 0) : DBL_MAX;
        const double rmin = grp_min<LPE>(r, sx);
        double th = 1.0;
        bool lim = false;
        if (el < nel && rmin < eps) {
            const double ravg = c[0] > eps ? c[0] : eps;
            const double num = ravg - eps, den = ravg - rmin;
            th = (den < 1e-32) ? 0.0 : (num / den < 1.0 ? num / den : 1.0);
            lim = true;
        }
        __syncthreads();
        if (lim)
            for (int l = 1 + q; l < N; l += LPE) c[5 * l] *= th;
        __syncthreads();
        // apply_22: pressure, by bisection towards the average at each point with p < 0
        double t = 1.0;
        if (act) {
            double P[5], avg[5];
            for (int k = 0; k < 5; ++k) {
                P[k] = point_value(V, c, N, q, k);
                avg[k] = c[k];
            }
            if (!(lim_pressure(P, gamma) >= 0)) {
                double lo = 0.0, hi = 1.0, mid = 0.5;
                for (int it = 0; it < 100; ++it) {
                    mid = 0.5 * (lo + hi);
                    double Um[5];
                    for (int k = 0; k < 5; ++k) Um[k] = (1.0 - mid) * avg[k] + mid * P[k];
                    if (lim_pressure(Um, gamma) < 0) hi = mid;
                    else lo = mid;
                }
                t = mid;
            }
        }
        const double thp = grp_min<LPE>(t, sx);
        __syncthreads();
        if (el < nel)
            for (int l = 1 + q; l < N; l += LPE)
                for (int k = 0; k < 5; ++k) c[k + 5 * l] *= thp;
        __syncthreads();
    }
    for (int i = threadIdx.x; i < nel * NV; i += 256) U[e0 * NV + i] = sC[i];
}

}  // namespace fvdg_lim
