// Implicit path of libfvdg.so: GPU assembly of the Picard-linearised block system of the
// reference's ImplicitConvection (src/DG/DG_Schemes/ImplicitConvection.cpp), a restarted GMRES
// with a block-Jacobi preconditioner on the device (the role of EigenSparseSolver::DGMRES,
// src/EigenSolver/EigenSparseSolver.cpp:163-195), and the backward-Euler Picard step of the
// reference's implicit drivers (examples/Eular2D/Sedov/Sedov.cpp:137-226).  Kernels:
// fvdg_implicit.hpp.
//
// GMRES (Saad & Schultz 1986), right preconditioned: the Krylov basis V lives on the device,
// orthogonalised by classical Gram-Schmidt with one re-orthogonalisation (each pass one fused
// multi-dot kernel and one fused update kernel), the Hessenberg matrix and its Givens rotations
// on the host.  Dot products are per-block partial sums added in block order on the host, so a
// solve is reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "fvdg_device.hpp"
#include "fvdg_implicit.hpp"

using namespace fvdg;
using namespace fvdg_imp;

namespace {

inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

int ready(fvdg_ctx* c, bool need_state) {
    if (!c) return fail(FVDG_ERR_ARG, "null context");
    if (c->order < 0) return fail(FVDG_ERR_STATE, "fvdg_set_scheme has not been called");
    if (!c->have_mesh) return fail(FVDG_ERR_STATE, "fvdg_set_mesh has not been called");
    if (need_state && !c->have_state) return fail(FVDG_ERR_STATE, "no state uploaded");
    return FVDG_OK;
}

// a block-row system on one stream (matrix, preconditioner factors and workspace)
struct Sys {
    hipStream_t s = nullptr;
    int64_t nrows = 0;
    int NB = 0, slots = 0;
    const double* A = nullptr;
    const int32_t* col = nullptr;
    const int32_t* dslot = nullptr;  // per block row: slot of the diagonal block, -1 none
    const double* Md = nullptr;      // per-unknown mass of the shift (nullable)
    double shift = 0.0;
    double* LU = nullptr;            // LU factors (NB > kMaxInvBlock) or explicit inverses
    int32_t* piv = nullptr;
    int32_t* flag = nullptr;
    const int32_t* self = nullptr;   // per block row: its own index (one-slot product with the inverses)
    bool use_inverse() const { return NB <= kMaxInvBlock; }
    DevBuf<double>* V = nullptr;  // (restart + 1) * n
    DevBuf<double>* w = nullptr;  // n
    DevBuf<double>* z = nullptr;  // n
    DevBuf<double>* part = nullptr;
    int64_t n() const { return nrows * NB; }
};

int sys_factor(const Sys& S) {
    HIP_TRY(hipMemsetAsync(S.flag, 0, sizeof(int32_t), S.s));
    if (S.use_inverse() && (S.NB == 5 || S.NB == 20 || S.NB == 50)) {
        // P0..P2 blocks fit one row per lane: registers, no barriers
        const auto kern = S.NB == 5 ? k_block_inverse_wave<5> : S.NB == 20 ? k_block_inverse_wave<20>
                                                                            : k_block_inverse_wave<50>;
        hipLaunchKernelGGL(kern, dim3((unsigned)S.nrows), dim3(64), 0, S.s, S.A, S.dslot, S.slots, S.Md, S.shift,
                           S.LU, S.flag);
    } else if (S.use_inverse()) {
        const size_t lds = sizeof(double) * ((size_t)S.NB * S.NB + S.NB);
        HIP_TRY(hipFuncSetAttribute((const void*)k_block_inverse, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_block_inverse, dim3((unsigned)S.nrows), dim3(256), lds, S.s, S.A, S.dslot, S.slots, S.Md,
                           S.shift, S.LU, S.NB, S.flag);
    } else {
        hipLaunchKernelGGL(k_block_lu, dim3((unsigned)S.nrows), dim3(256), 0, S.s, S.A, S.dslot, S.slots, S.Md,
                           S.shift, S.LU, S.piv, S.NB, S.flag);
    }
    HIP_TRY(hipGetLastError());
    int32_t f = 0;
    HIP_TRY(hipMemcpyAsync(&f, S.flag, sizeof(int32_t), hipMemcpyDeviceToHost, S.s));
    HIP_TRY(hipStreamSynchronize(S.s));
    if (f) return fail(FVDG_ERR_NONFINITE, "a diagonal block of the implicit system is singular");
    return FVDG_OK;
}

// y = (A + shift diag(Md)) x over `slots` blocks per block row
void bsr_product(hipStream_t st, const double* A, const int32_t* col, int slots, const double* Md, double shift,
                 const double* x, double* y, int64_t nrows, int NB) {
    const dim3 grid(nblk(nrows * NB, 256)), block(256);
    switch (NB) {
        case 20: hipLaunchKernelGGL(k_bsr_matvec_nb<20>, grid, block, 0, st, A, col, slots, Md, shift, x, y, nrows); break;
        case 50: hipLaunchKernelGGL(k_bsr_matvec_nb<50>, grid, block, 0, st, A, col, slots, Md, shift, x, y, nrows); break;
        case 100: hipLaunchKernelGGL(k_bsr_matvec_nb<100>, grid, block, 0, st, A, col, slots, Md, shift, x, y, nrows); break;
        default: hipLaunchKernelGGL(k_bsr_matvec, grid, block, 0, st, A, col, slots, Md, shift, x, y, nrows, NB);
    }
}

void sys_matvec(const Sys& S, const double* x, double* y) {
    bsr_product(S.s, S.A, S.col, S.slots, S.Md, S.shift, x, y, S.nrows, S.NB);
}

void sys_precond(const Sys& S, const double* r, double* z) {
    if (S.use_inverse())
        bsr_product(S.s, S.LU, S.self, 1, nullptr, 0.0, r, z, S.nrows, S.NB);
    else
        hipLaunchKernelGGL(k_block_lu_solve, dim3((unsigned)S.nrows), dim3(256), sizeof(double) * S.NB, S.s, S.LU,
                           S.piv, r, z, S.NB);
}

// h[i] = V_i . w for i < nv (V_i consecutive vectors of length n starting at V)
int sys_dots(const Sys& S, const double* V, int nv, const double* w, double* h) {
    const unsigned nb = nblk(S.n(), 2048);
    if (S.part->n < (size_t)nb * nv) HIP_TRY(S.part->alloc((size_t)nb * nv));
    hipLaunchKernelGGL(k_multi_dot, dim3(nb), dim3(256), 0, S.s, V, nv, S.n(), w, S.part->p);
    HIP_TRY(hipGetLastError());
    std::vector<double> p((size_t)nb * nv);
    HIP_TRY(hipMemcpyAsync(p.data(), S.part->p, p.size() * sizeof(double), hipMemcpyDeviceToHost, S.s));
    HIP_TRY(hipStreamSynchronize(S.s));
    for (int v = 0; v < nv; ++v) {
        double acc = 0.0;
        for (unsigned b = 0; b < nb; ++b) acc += p[(size_t)b * nv + v];
        h[v] = acc;
    }
    return FVDG_OK;
}

void sys_lincomb(const Sys& S, double* y, double beta, const double* V, int nv, const double* c) {
    Coefs k{};
    for (int i = 0; i < nv; ++i) k.c[i] = c[i];
    hipLaunchKernelGGL(k_lincomb, dim3(nblk(S.n(), 256)), dim3(256), 0, S.s, y, beta, V, nv, S.n(), k);
}

void sys_axpby(const Sys& S, double* y, double a, const double* x, double b) {
    hipLaunchKernelGGL(k_axpby, dim3(nblk(S.n(), 256)), dim3(256), 0, S.s, y, a, x, b, S.n());
}

// Solve (A + shift M) x = b, x holding the initial guess; the factors must be current.
int gmres(const Sys& S, const double* b, double* x, double tol, int m, int max_it, int* iters, double* relres) {
    const int64_t n = S.n();
    if (S.V->n < (size_t)(m + 1) * n) HIP_TRY(S.V->alloc((size_t)(m + 1) * n));
    if (S.w->n < (size_t)n) HIP_TRY(S.w->alloc(n));
    if (S.z->n < (size_t)n) HIP_TRY(S.z->alloc(n));
    double* V = S.V->p;
    double* w = S.w->p;
    double* z = S.z->p;
    double bn2 = 0.0;
    int rc = sys_dots(S, b, 1, b, &bn2);
    if (rc) return rc;
    const double bnorm = std::sqrt(bn2);
    if (!std::isfinite(bnorm)) return fail(FVDG_ERR_NONFINITE, "right-hand side is not finite");
    if (bnorm == 0.0) {
        HIP_TRY(hipMemsetAsync(x, 0, n * sizeof(double), S.s));
        HIP_TRY(hipStreamSynchronize(S.s));
        if (iters) *iters = 0;
        if (relres) *relres = 0.0;
        return FVDG_OK;
    }
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), h(m + 1), h2(m + 1);
    int total = 0;
    double est = INFINITY;
    for (;;) {
        // r = b - A x, into V_0
        sys_matvec(S, x, w);
        HIP_TRY(hipMemcpyAsync(V, b, n * sizeof(double), hipMemcpyDeviceToDevice, S.s));
        sys_axpby(S, V, -1.0, w, 1.0);
        double beta2 = 0.0;
        rc = sys_dots(S, V, 1, V, &beta2);
        if (rc) return rc;
        const double beta = std::sqrt(beta2);
        est = beta / bnorm;
        if (!std::isfinite(beta)) return fail(FVDG_ERR_NONFINITE, "GMRES residual is not finite");
        if (est < tol || total >= max_it) break;
        sys_axpby(S, V, 0.0, V, 1.0 / beta);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int jend = 0;
        for (int j = 0; j < m; ++j) {
            sys_precond(S, V + (int64_t)j * n, z);
            sys_matvec(S, z, w);
            // classical Gram-Schmidt, twice
            rc = sys_dots(S, V, j + 1, w, h.data());
            if (rc) return rc;
            for (int i = 0; i <= j; ++i) h2[i] = -h[i];
            sys_lincomb(S, w, 1.0, V, j + 1, h2.data());
            rc = sys_dots(S, V, j + 1, w, h2.data());
            if (rc) return rc;
            for (int i = 0; i <= j; ++i) h[i] += h2[i];
            for (int i = 0; i <= j; ++i) h2[i] = -h2[i];
            sys_lincomb(S, w, 1.0, V, j + 1, h2.data());
            double hn2 = 0.0;
            rc = sys_dots(S, w, 1, w, &hn2);
            if (rc) return rc;
            double hn = std::sqrt(hn2);
            if (hn > 0.0) sys_axpby(S, V + (int64_t)(j + 1) * n, 1.0 / hn, w, 0.0);
            // Givens rotations
            for (int i = 0; i < j; ++i) {
                const double t = cs[i] * h[i] + sn[i] * h[i + 1];
                h[i + 1] = -sn[i] * h[i] + cs[i] * h[i + 1];
                h[i] = t;
            }
            const double r = std::hypot(h[j], hn);
            cs[j] = r > 0.0 ? h[j] / r : 1.0;
            sn[j] = r > 0.0 ? hn / r : 0.0;
            h[j] = r;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            for (int i = 0; i <= j; ++i) H[(size_t)i * m + j] = h[i];
            ++total;
            jend = j + 1;
            est = std::fabs(g[j + 1]) / bnorm;
            if (!std::isfinite(est)) return fail(FVDG_ERR_NONFINITE, "GMRES iteration is not finite");
            if (est < tol || total >= max_it || hn == 0.0) break;
        }
        // y = H^{-1} g (upper triangular), x += P^{-1} V y
        for (int i = jend - 1; i >= 0; --i) {
            double s = g[i];
            for (int k = i + 1; k < jend; ++k) s -= H[(size_t)i * m + k] * y[k];
            y[i] = s / H[(size_t)i * m + i];
        }
        sys_lincomb(S, w, 0.0, V, jend, y.data());
        sys_precond(S, w, z);
        sys_axpby(S, x, 1.0, z, 1.0);
        if (total >= max_it) {
            // report the true residual of the final iterate
            sys_matvec(S, x, w);
            HIP_TRY(hipMemcpyAsync(V, b, n * sizeof(double), hipMemcpyDeviceToDevice, S.s));
            sys_axpby(S, V, -1.0, w, 1.0);
            rc = sys_dots(S, V, 1, V, &beta2);
            if (rc) return rc;
            est = std::sqrt(beta2) / bnorm;
            break;
        }
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(S.s));
    if (iters) *iters = total;
    if (relres) *relres = est;
    return FVDG_OK;
}

// ---------------------------------------------------------------- assembled systems of a context
int prepare(fvdg_ctx* c) {
    const int64_t E = c->E;
    const int NB = 5 * c->tab.N;
    const size_t nbb = (size_t)NB * NB;
    if (c->imp_A.n != (size_t)E * kSlots * nbb) {
        HIP_TRY(c->imp_A.alloc((size_t)E * kSlots * nbb));
        HIP_TRY(c->imp_col.alloc((size_t)E * kSlots));
        HIP_TRY(c->imp_rhs.alloc((size_t)E * NB));
        HIP_TRY(c->imp_b.alloc((size_t)E * NB));
        HIP_TRY(c->imp_LU.alloc((size_t)E * nbb));
        HIP_TRY(c->imp_piv.alloc((size_t)E * NB));
        HIP_TRY(c->imp_flag.alloc(1));
        HIP_TRY(c->imp_xold.alloc((size_t)E * NB));
        HIP_TRY(c->imp_dx.alloc((size_t)E * NB));
        // diagonal slot 0 in every row, each row's own index; per-unknown mass |det| mhat_k
        HIP_TRY(c->imp_dslot.upload(std::vector<int32_t>(E, 0), c->stream));
        std::vector<int32_t> self(E);
        for (int64_t e = 0; e < E; ++e) self[e] = (int32_t)e;
        HIP_TRY(c->imp_self.upload(self, c->stream));
        std::vector<double> Md((size_t)E * NB);
        for (int64_t e = 0; e < E; ++e)
            for (int r = 0; r < NB; ++r) Md[(size_t)e * NB + r] = c->cell_det_host[e] * c->tab.mass_hat[r / 5];
        HIP_TRY(c->imp_Md.upload(Md, c->stream));
        c->imp_have = false;
    }
    return FVDG_OK;
}

template <int P>
int assemble_t(fvdg_ctx* c, const double* U, int parts) {
    using S = ImpShape<P>;
    const size_t lds = sizeof(double) * S::LDS_DOUBLES;
    HIP_TRY(hipFuncSetAttribute((const void*)k_assemble<P>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    AssembleArgs a{c->E,       c->gamma,  parts,      U,           c->cgeo.p,   c->cfaces.p, c->fcell.p,
                   c->fcombo.p, c->fgeo.p, c->fbc.p,   c->fdir.p,   c->UD.p,     c->tV.p,     c->tG.p,
                   c->tWq.p,   c->tPhi.p, c->tWf.p,   c->imp_A.p,  c->imp_col.p, c->imp_rhs.p};
    hipLaunchKernelGGL(k_assemble<P>, dim3((unsigned)c->E), dim3(256), lds, c->stream, a);
    HIP_TRY(hipGetLastError());
    return FVDG_OK;
}

int assemble(fvdg_ctx* c, const double* U, int parts) {
    int rc = prepare(c);
    if (rc) return rc;
    switch (c->order) {
        case 0: rc = assemble_t<0>(c, U, parts); break;
        case 1: rc = assemble_t<1>(c, U, parts); break;
        case 2: rc = assemble_t<2>(c, U, parts); break;
        case 3: rc = assemble_t<3>(c, U, parts); break;
        case 4: rc = assemble_t<4>(c, U, parts); break;
        default: rc = assemble_t<5>(c, U, parts); break;
    }
    if (rc) return rc;
    c->imp_have = true;
    return FVDG_OK;
}

Sys context_sys(fvdg_ctx* c, double shift) {
    Sys S;
    S.s = c->stream;
    S.nrows = c->E;
    S.NB = 5 * c->tab.N;
    S.slots = kSlots;
    S.A = c->imp_A.p;
    S.col = c->imp_col.p;
    S.dslot = c->imp_dslot.p;
    S.Md = c->imp_Md.p;
    S.shift = shift;
    S.LU = c->imp_LU.p;
    S.piv = c->imp_piv.p;
    S.flag = c->imp_flag.p;
    S.self = c->imp_self.p;
    S.V = &c->imp_V;
    S.w = &c->imp_w;
    S.z = &c->imp_z;
    S.part = &c->imp_part;
    return S;
}

int check_gmres_args(double tol, int restart, int max_iters) {
    if (!(tol > 0.0)) return fail(FVDG_ERR_ARG, "tol must be positive");
    if (restart < 1 || restart >= kMaxCoef) return fail(FVDG_ERR_ARG, "restart must be in 1..63");
    if (max_iters < 1) return fail(FVDG_ERR_ARG, "max_iters must be positive");
    return FVDG_OK;
}

}  // namespace

extern "C" {

int fvdg_implicit_assemble(fvdg_ctx* c, int parts) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (parts < 1 || parts > 7) return fail(FVDG_ERR_ARG, "parts must combine FVDG_PART_CELLS/INTERNALS/BOUNDARYS");
    HIP_TRY(hipSetDevice(c->device));
    rc = assemble(c, c->U.p, parts);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FVDG_OK;
}

int fvdg_implicit_blocks(fvdg_ctx* c, double* diag, double* off, int32_t* off_cols, double* rhs) {
    int rc = ready(c, false);
    if (rc) return rc;
    if (!c->imp_have) return fail(FVDG_ERR_STATE, "fvdg_implicit_assemble has not been called");
    HIP_TRY(hipSetDevice(c->device));
    const int64_t E = c->E;
    const int NB = 5 * c->tab.N;
    const size_t nbb = (size_t)NB * NB;
    if (diag || off) {
        std::vector<double> A((size_t)E * kSlots * nbb);
        HIP_TRY(hipMemcpyAsync(A.data(), c->imp_A.p, A.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int64_t e = 0; e < E; ++e)
            for (int s = 0; s < kSlots; ++s) {
                double* dst = s == 0 ? (diag ? diag + e * nbb : nullptr) : (off ? off + (e * 4 + s - 1) * nbb : nullptr);
                if (!dst) continue;
                const double* src = A.data() + (e * kSlots + s) * nbb;
                for (int r = 0; r < NB; ++r)
                    for (int k = 0; k < NB; ++k) dst[(size_t)r * NB + k] = src[(size_t)k * NB + r];
            }
    }
    if (off_cols) {
        std::vector<int32_t> col((size_t)E * kSlots);
        HIP_TRY(hipMemcpyAsync(col.data(), c->imp_col.p, col.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int64_t e = 0; e < E; ++e)
            for (int l = 0; l < 4; ++l) off_cols[e * 4 + l] = col[e * kSlots + 1 + l];
    }
    if (rhs) {
        HIP_TRY(hipMemcpyAsync(rhs, c->imp_rhs.p, (size_t)E * NB * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return FVDG_OK;
}

int fvdg_implicit_matvec(fvdg_ctx* c, double shift, const double* x, double* y) {
    int rc = ready(c, false);
    if (rc) return rc;
    if (!x || !y) return fail(FVDG_ERR_ARG, "null vector");
    if (!c->imp_have) return fail(FVDG_ERR_STATE, "fvdg_implicit_assemble has not been called");
    HIP_TRY(hipSetDevice(c->device));
    Sys S = context_sys(c, shift);
    const int64_t n = S.n();
    if (c->imp_w.n < (size_t)n) HIP_TRY(c->imp_w.alloc(n));
    HIP_TRY(hipMemcpyAsync(c->imp_b.p, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    sys_matvec(S, c->imp_b.p, c->imp_w.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(y, c->imp_w.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FVDG_OK;
}

int fvdg_implicit_solve(fvdg_ctx* c, double shift, const double* b, double* x, double tol, int restart,
                        int max_iters, int* iters, double* rel_res) {
    int rc = ready(c, false);
    if (rc) return rc;
    if (!x) return fail(FVDG_ERR_ARG, "null solution vector");
    if (!c->imp_have) return fail(FVDG_ERR_STATE, "fvdg_implicit_assemble has not been called");
    rc = check_gmres_args(tol, restart, max_iters);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    Sys S = context_sys(c, shift);
    const int64_t n = S.n();
    if (b) HIP_TRY(hipMemcpyAsync(c->imp_b.p, b, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    else HIP_TRY(hipMemcpyAsync(c->imp_b.p, c->imp_rhs.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->imp_dx.p, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    rc = sys_factor(S);
    if (rc) return rc;
    rc = gmres(S, c->imp_b.p, c->imp_dx.p, tol, restart, max_iters, iters, rel_res);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(x, c->imp_dx.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FVDG_OK;
}

int fvdg_implicit_euler_step(fvdg_ctx* c, double dt, int max_picard, double picard_tol, double tol, int restart,
                             int max_iters, int* picard_iters, int* linear_iters, double* delta_out) {
    int rc = ready(c, true);
    if (rc) return rc;
    if (!(dt > 0.0) || max_picard < 1 || !(picard_tol > 0.0)) return fail(FVDG_ERR_ARG, "bad dt or Picard settings");
    rc = check_gmres_args(tol, restart, max_iters);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    rc = prepare(c);
    if (rc) return rc;
    Sys S = context_sys(c, 1.0 / dt);
    const int64_t n = S.n();
    HIP_TRY(hipMemcpyAsync(c->imp_xold.p, c->U.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->imp_dx.p, c->U.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    double init_delta = 0.0, prev_delta = 0.0, delta = 0.0;
    int it = 0, lin = 0;
    for (; it < max_picard;) {
        rc = assemble(c, c->imp_dx.p, FVDG_PART_ALL);
        if (rc) return rc;
        // b = rhs + M/dt x_old; the new iterate starts from dx (in U's buffer)
        hipLaunchKernelGGL(k_rhs_shift, dim3(nblk(n, 256)), dim3(256), 0, c->stream, c->imp_b.p, c->imp_rhs.p,
                           c->imp_Md.p, c->imp_xold.p, 1.0 / dt, n);
        HIP_TRY(hipMemcpyAsync(c->U.p, c->imp_dx.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        rc = sys_factor(S);
        if (rc) return rc;
        int li = 0;
        rc = gmres(S, c->imp_b.p, c->U.p, tol, restart, max_iters, &li, nullptr);
        if (rc) return rc;
        lin += li;
        if (c->lim_parts) fvdg::limit_state(c, c->imp_dx.p, c->U.p, c->lim_parts, c->lim_avg);
        // delta = |x - dx|; dx = x
        sys_axpby(S, c->imp_dx.p, 1.0, c->U.p, -1.0);
        double d2 = 0.0;
        rc = sys_dots(S, c->imp_dx.p, 1, c->imp_dx.p, &d2);
        if (rc) return rc;
        delta = std::sqrt(d2);
        HIP_TRY(hipMemcpyAsync(c->imp_dx.p, c->U.p, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        if (!std::isfinite(delta)) return fail(FVDG_ERR_NONFINITE, "Picard iterate is not finite");
        if (it == 0) init_delta = prev_delta = delta;
        const double rate = init_delta > 0.0 ? delta / init_delta : 0.0;
        const double rel = prev_delta > 0.0 ? delta / prev_delta : 0.0;
        prev_delta = delta;
        ++it;
        if (delta < picard_tol || rel < picard_tol || (it > 11 && rate > 0.95)) break;
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (picard_iters) *picard_iters = it;
    if (linear_iters) *linear_iters = lin;
    if (delta_out) *delta_out = delta;
    return FVDG_OK;
}

int fvdg_bsr_solve(int device, int64_t n_rows, int block, const int64_t* row_ptr, const int64_t* cols,
                   const double* blocks, const double* b, double* x, double tol, int restart, int max_iters,
                   int* iters, double* rel_res) {
    if (n_rows <= 0 || block <= 0 || !row_ptr || !cols || !blocks || !b || !x)
        return fail(FVDG_ERR_ARG, "bad block-sparse system");
    int rc = check_gmres_args(tol, restart, max_iters);
    if (rc) return rc;
    if (n_rows > (int64_t(1) << 30)) return fail(FVDG_ERR_UNSUPPORTED, "too many block rows");
    int slots = 0;
    for (int64_t r = 0; r < n_rows; ++r) {
        if (row_ptr[r + 1] < row_ptr[r]) return fail(FVDG_ERR_ARG, "row_ptr is not non-decreasing");
        slots = std::max<int>(slots, (int)(row_ptr[r + 1] - row_ptr[r]));
    }
    slots = std::max(slots, 1);
    const size_t nbb = (size_t)block * block;
    std::vector<double> A((size_t)n_rows * slots * nbb, 0.0);
    std::vector<int32_t> col((size_t)n_rows * slots, -1), dslot(n_rows, -1);
    for (int64_t r = 0; r < n_rows; ++r)
        for (int64_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            const int s = (int)(k - row_ptr[r]);
            if (cols[k] < 0 || cols[k] >= n_rows) return fail(FVDG_ERR_ARG, "column index out of range");
            col[r * slots + s] = (int32_t)cols[k];
            if (cols[k] == r && dslot[r] < 0) dslot[r] = s;
            const double* src = blocks + k * nbb;
            double* dst = A.data() + (r * slots + s) * nbb;
            for (int i = 0; i < block; ++i)
                for (int j = 0; j < block; ++j) dst[(size_t)j * block + i] = src[(size_t)i * block + j];
        }
    HIP_TRY(hipSetDevice(device));
    hipStream_t s = nullptr;
    HIP_TRY(hipStreamCreate(&s));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    DevBuf<double> dA, dLU, dV, dw, dz, dpart, db, dx;
    DevBuf<int32_t> dcol, ddslot, dpiv, dflag, dself;
    HIP_TRY(dA.upload(A, s));
    HIP_TRY(dcol.upload(col, s));
    HIP_TRY(ddslot.upload(dslot, s));
    std::vector<int32_t> self(n_rows);
    for (int64_t r = 0; r < n_rows; ++r) self[r] = (int32_t)r;
    HIP_TRY(dself.upload(self, s));
    const int64_t n = n_rows * block;
    HIP_TRY(dLU.alloc((size_t)n_rows * nbb));
    HIP_TRY(dpiv.alloc((size_t)n));
    HIP_TRY(dflag.alloc(1));
    HIP_TRY(db.upload(std::vector<double>(b, b + n), s));
    HIP_TRY(dx.upload(std::vector<double>(x, x + n), s));
    Sys S;
    S.s = s;
    S.nrows = n_rows;
    S.NB = block;
    S.slots = slots;
    S.A = dA.p;
    S.col = dcol.p;
    S.dslot = ddslot.p;
    S.LU = dLU.p;
    S.piv = dpiv.p;
    S.flag = dflag.p;
    S.self = dself.p;
    S.V = &dV;
    S.w = &dw;
    S.z = &dz;
    S.part = &dpart;
    rc = sys_factor(S);
    if (rc) return rc;
    rc = gmres(S, db.p, dx.p, tol, restart, max_iters, iters, rel_res);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(x, dx.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return FVDG_OK;
}

}  // extern "C"
