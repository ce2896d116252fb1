/*
 * dg_oracle.c -- CPU ORACLE for the explicit DG Euler residual (TEST INFRASTRUCTURE).
 *
 * This file is a plain-C restatement of the reference's ExplicitConvection path
 * and is used ONLY as a checker: by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg.  Nothing in the product (fve_dg_amd/, include/)
 * links, loads or calls it.
 *
 * What it restates (reference file:line):
 *   or_basis           DGBasis<k>::eval/grad, include/DG/DG_Basis/DG_Basis_Func.h:7-670, via the
 *                      explicit Jacobi sum s^n P_n^{(a,0)}((2z-s)/s) = sum_m C(n+a,n-m) C(n,m) (z-s)^m z^(n-m)
 *                      (a different formula from the library's recurrence)
 *   or_flux, or_wave_speed, or_lf
 *                      src/DG/DG_Flux/EulerPhysicalFlux.cpp:7-18, 46-55, 70-89;
 *                      LF form of src/DG/DG_Schemes/ExplicitConvection.cpp:232-240
 *   or_roe             Roe (1981) with Harten's entropy fix -- no counterpart in the reference
 *   or_inverse3        include/Matrix/DenseMatrix.h:521-538
 *   or_geometry        src/Mesh/ComputingMesh.cpp:40-76 (faces), :113-125 and :145-164 (tets)
 *   or_residual        src/DG/DG_Schemes/ExplicitConvection.cpp:21-272 (cells, internals,
 *                      boundarys), same loop nest: per point basis evaluation, a fresh 3x3
 *                      inverse per (q, j), and with threads>1 OpenMP dynamic scheduling with
 *                      atomic scatter as in the reference
 *   or_mass            examples/Vortex/123/VortexBatch.cpp:153-166
 *   or_ssprk3          Shu & Osher (1988) SSP-RK3 on dU/dt = -M^{-1} R (driver form
 *                      examples/Vortex/123/VortexBatch.cpp:189-197 generalised to 3 stages)
 *   or_vortex_state, or_project
 *                      examples/Vortex/problem.h:17-110, include/DG/DG_Basis/DG_Basis_Evaluator.h:33-59
 *   or_jacobian        src/DG/DG_Flux/EulerPhysicalFlux.cpp:21-42
 *   or_implicit_assemble
 *                      src/DG/DG_Schemes/ImplicitConvection.cpp:20-167 (boundarys), :266-325 (cells),
 *                      :329-420 (internals): the Picard blocks and right-hand side
 *   or_limiter_bounds, or_limiter_apply
 *                      include/DG/DG_Schemes/PositiveLimiter.h:62-119 (constructMinMax), :135-170
 *                      (apply_1), :171-198 (apply_21), :199-259 (apply_22); that header cannot be
 *                      compiled here (Mesh/Mesh.h reaches CGAL and Boost), so these are pinned by
 *                      the limiter's invariants in tests/test_limiter.py
 *
 * Layouts: state U[e*5N + 5k + v] (reference LongVector<5N>); cells[4e+a]; faces[3f+k];
 * face_cells[2f+s] with -1 for a boundary side; face_nodes_r[3f+k] = nodes seen from side 1.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/GaussQuadrature/quad_rules.inc"

/* ------------------------------------------------------------------ quadrature */
static const double TET15_P[][3] = FVDG_TET_DEGREE5POINTS15_PTS, TET15_W[] = FVDG_TET_DEGREE5POINTS15_W;
static const double TET31_P[][3] = FVDG_TET_DEGREE7POINTS31_PTS, TET31_W[] = FVDG_TET_DEGREE7POINTS31_W;
static const double TET57_P[][3] = FVDG_TET_DEGREE9POINTS57_PTS, TET57_W[] = FVDG_TET_DEGREE9POINTS57_W;
static const double TET94_P[][3] = FVDG_TET_DEGREE11POINTS94_PTS, TET94_W[] = FVDG_TET_DEGREE11POINTS94_W;
static const double TRI7_P[][2] = FVDG_TRI_DEGREE4POINTS7_PTS, TRI7_W[] = FVDG_TRI_DEGREE4POINTS7_W;
static const double TRI24_P[][2] = FVDG_TRI_DEGREE10POINTS24_PTS, TRI24_W[] = FVDG_TRI_DEGREE10POINTS24_W;
static const double TRI36_P[][2] = FVDG_TRI_DEGREE13POINTS36_PTS, TRI36_W[] = FVDG_TRI_DEGREE13POINTS36_W;

/* AutoQuadSelector: include/GaussQuadrature/GaussLegendre.h:84-111, AutoQuadSelectorTri.h:1-22 */
int or_tet_rule(int order, const double** pts, const double** w) {
    switch (order) {
        case 0: case 1: case 2: *pts = &TET15_P[0][0]; *w = TET15_W; return 15;
        case 3: *pts = &TET31_P[0][0]; *w = TET31_W; return 31;
        case 4: *pts = &TET57_P[0][0]; *w = TET57_W; return 57;
        case 5: *pts = &TET94_P[0][0]; *w = TET94_W; return 94;
        default: return 0;
    }
}
int or_tri_rule(int order, const double** pts, const double** w) {
    switch (order) {
        case 0: case 1: *pts = &TRI7_P[0][0]; *w = TRI7_W; return 7;
        case 2: case 3: case 4: *pts = &TRI24_P[0][0]; *w = TRI24_W; return 24;
        case 5: *pts = &TRI36_P[0][0]; *w = TRI36_W; return 36;
        default: return 0;
    }
}
/* copy-out form for tests */
int or_quadrature(int order, int face, double* pts, double* w) {
    const double *p, *ww;
    int n = face ? or_tri_rule(order, &p, &ww) : or_tet_rule(order, &p, &ww);
    if (pts) memcpy(pts, p, sizeof(double) * n * (face ? 2 : 3));
    if (w) memcpy(w, ww, sizeof(double) * n);
    return n;
}

int or_num_basis(int order) { return (order + 1) * (order + 2) * (order + 3) / 6; }

/* ------------------------------------------------------------------ basis */
typedef struct { double v, x, y, z; } dual;
static dual d_const(double c) { dual r = {c, 0, 0, 0}; return r; }
static dual d_add(dual a, dual b) { dual r = {a.v + b.v, a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static dual d_sub(dual a, dual b) { dual r = {a.v - b.v, a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static dual d_mul(dual a, dual b) {
    dual r = {a.v * b.v, a.x * b.v + a.v * b.x, a.y * b.v + a.v * b.y, a.z * b.v + a.v * b.z};
    return r;
}
static dual d_scale(double c, dual a) { dual r = {c * a.v, c * a.x, c * a.y, c * a.z}; return r; }
static dual d_pow(dual a, int m) {
    dual r = d_const(1.0);
    for (int i = 0; i < m; ++i) r = d_mul(r, a);
    return r;
}
static double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    double r = 1.0;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}
/* s^n P_n^{(alpha,0)}((2z - s)/s) = sum_m C(n+alpha, n-m) C(n, m) (z - s)^m z^(n-m) */
static dual jacobi_sum(int n, int alpha, dual z, dual s) {
    dual acc = d_const(0.0), zs = d_sub(z, s);
    for (int m = 0; m <= n; ++m)
        acc = d_add(acc, d_scale(binom(n + alpha, n - m) * binom(n, m), d_mul(d_pow(zs, m), d_pow(z, n - m))));
    return acc;
}

/* values val[N] and gradients grad[3N] (may be NULL) of the order-`order` basis */
void or_basis(int order, double x, double y, double z, double* val, double* grad) {
    dual X = {x, 1, 0, 0}, Y = {y, 0, 1, 0}, Z = {z, 0, 0, 1}, one = d_const(1.0);
    dual s = d_sub(d_sub(one, X), Y), r = d_sub(one, X);
    int idx = 0;
    for (int n = 0; n <= order; ++n)
        for (int k = 0; k <= n; ++k)
            for (int j = 0; j <= n - k; ++j) {
                int i = n - j - k;
                dual phi = d_mul(d_mul(jacobi_sum(k, 0, Z, s), jacobi_sum(j, 2 * k + 1, Y, r)),
                                 jacobi_sum(i, 2 * j + 2 * k + 2, X, one));
                if (val) val[idx] = phi.v;
                if (grad) {
                    grad[3 * idx] = phi.x;
                    grad[3 * idx + 1] = phi.y;
                    grad[3 * idx + 2] = phi.z;
                }
                ++idx;
            }
}

/* ------------------------------------------------------------------ Euler algebra */
static double positive(double v) { return v > 1e-16 ? v : 1e-16; }
double or_pressure(const double U[5], double gamma) {
    double rho = positive(U[0]);
    double u = U[1] / rho, v = U[2] / rho, w = U[3] / rho, e = U[4] / rho;
    return (gamma - 1) * (rho * e - 0.5 * rho * (u * u + v * v + w * w));
}
/* F[3v+d], row v = conserved variable, column d = direction */
void or_flux(const double U[5], double gamma, double F[15]) {
    double rho = positive(U[0]);
    double u = U[1] / rho, v = U[2] / rho, w = U[3] / rho, e = U[4] / rho;
    double p = or_pressure(U, gamma);
    F[0] = rho * u;              F[1] = rho * v;              F[2] = rho * w;
    F[3] = rho * u * u + p;      F[4] = rho * v * u;          F[5] = rho * w * u;
    F[6] = rho * u * v;          F[7] = rho * v * v + p;      F[8] = rho * w * v;
    F[9] = rho * u * w;          F[10] = rho * v * w;         F[11] = rho * w * w + p;
    F[12] = u * (rho * e + p);   F[13] = v * (rho * e + p);   F[14] = w * (rho * e + p);
}
static double wave_mag(const double U[5], double gamma) {
    double rho = positive(U[0]);
    double u = U[1] / rho, v = U[2] / rho, w = U[3] / rho;
    double c = sqrt(gamma * positive(or_pressure(U, gamma) / positive(U[0])));
    return sqrt(u * u + v * v + w * w) + c;
}
double or_wave_speed(const double UL[5], const double UR[5], double gamma) {
    double a = wave_mag(UL, gamma), b = wave_mag(UR, gamma);
    return a > b ? a : b;
}
static void flux_n(const double U[5], const double n[3], double gamma, double out[5]) {
    double F[15];
    or_flux(U, gamma, F);
    for (int v = 0; v < 5; ++v) out[v] = F[3 * v] * n[0] + F[3 * v + 1] * n[1] + F[3 * v + 2] * n[2];
}
void or_lf(const double UL[5], const double UR[5], const double n[3], double gamma, double out[5]) {
    double fl[5], fr[5];
    flux_n(UL, n, gamma, fl);
    flux_n(UR, n, gamma, fr);
    double lam = or_wave_speed(UL, UR, gamma);
    for (int v = 0; v < 5; ++v) out[v] = 0.5 * (fl[v] + fr[v] + lam * (UL[v] - UR[v]));
}
/* Roe flux 0.5(F_L n + F_R n) - 0.5 sum_k |l_k| alpha_k r_k with Harten's fix (delta = 0.1 a)
 * on the acoustic waves.  signed_eigs != 0 returns instead sum_k l_k alpha_k r_k (no abs, no
 * fix), which equals (F_R - F_L) n exactly for the Roe average (tests pin the decomposition).
 * A non-physical state (rho <= 0 or p <= 0) takes the Lax-Friedrichs flux: the Roe average
 * is only guaranteed a positive sound speed for physical states. */
void or_roe(const double UL[5], const double UR[5], const double n[3], double gamma, double out[5],
            int signed_eigs) {
    double rL = positive(UL[0]), rR = positive(UR[0]);
    double pL = or_pressure(UL, gamma), pR = or_pressure(UR, gamma);
    if (!signed_eigs && !(UL[0] > 0 && UR[0] > 0 && pL > 0 && pR > 0)) {
        or_lf(UL, UR, n, gamma, out);
        return;
    }
    double uL[3] = {UL[1] / rL, UL[2] / rL, UL[3] / rL}, uR[3] = {UR[1] / rR, UR[2] / rR, UR[3] / rR};
    double HL = (UL[4] + pL) / rL, HR = (UR[4] + pR) / rR;
    double sL = sqrt(rL), sR = sqrt(rR);
    double ut[3], q2 = 0, qn = 0;
    for (int d = 0; d < 3; ++d) {
        ut[d] = (sL * uL[d] + sR * uR[d]) / (sL + sR);
        q2 += ut[d] * ut[d];
        qn += ut[d] * n[d];
    }
    double Ht = (sL * HL + sR * HR) / (sL + sR);
    double a2 = (gamma - 1) * (Ht - 0.5 * q2);
    if (a2 < 1e-16) a2 = 1e-16;
    double a = sqrt(a2), rt = sL * sR;
    double dr = rR - rL, dp = pR - pL, du[3], dqn = 0;
    for (int d = 0; d < 3; ++d) {
        du[d] = uR[d] - uL[d];
        dqn += du[d] * n[d];
    }
    double dut[3], ut_dut = 0;
    for (int d = 0; d < 3; ++d) {
        dut[d] = du[d] - dqn * n[d];
        ut_dut += ut[d] * dut[d];
    }
    double al1 = (dp - rt * a * dqn) / (2 * a2), al5 = (dp + rt * a * dqn) / (2 * a2), al2 = dr - dp / a2;
    double l1 = qn - a, l2 = qn, l5 = qn + a;
    if (!signed_eigs) {
        double delta = 0.1 * a;
        l1 = fabs(l1) < delta ? (l1 * l1 + delta * delta) / (2 * delta) : fabs(l1);
        l5 = fabs(l5) < delta ? (l5 * l5 + delta * delta) / (2 * delta) : fabs(l5);
        l2 = fabs(l2);
    }
    double dd[5];
    dd[0] = l1 * al1 + l2 * al2 + l5 * al5;
    for (int k = 0; k < 3; ++k)
        dd[1 + k] = l1 * al1 * (ut[k] - a * n[k]) + l2 * al2 * ut[k] + l2 * rt * dut[k] + l5 * al5 * (ut[k] + a * n[k]);
    dd[4] = l1 * al1 * (Ht - a * qn) + l2 * al2 * 0.5 * q2 + l2 * rt * ut_dut + l5 * al5 * (Ht + a * qn);
    if (signed_eigs) {
        for (int v = 0; v < 5; ++v) out[v] = dd[v];
        return;
    }
    double fl[5], fr[5];
    flux_n(UL, n, gamma, fl);
    flux_n(UR, n, gamma, fr);
    for (int v = 0; v < 5; ++v) out[v] = 0.5 * (fl[v] + fr[v]) - 0.5 * dd[v];
}
static void num_flux(int kind, const double UL[5], const double UR[5], const double n[3], double gamma, double out[5]) {
    if (kind == 1)
        or_roe(UL, UR, n, gamma, out, 0);
    else
        or_lf(UL, UR, n, gamma, out);
}

/* ------------------------------------------------------------------ geometry */
void or_inverse3(const double A[9], double R[9]) {
#define a(i, j) A[3 * (i) + (j)]
#define r(i, j) R[3 * (i) + (j)]
    r(0, 0) = a(1, 1) * a(2, 2) - a(1, 2) * a(2, 1);
    r(1, 0) = a(1, 2) * a(2, 0) - a(1, 0) * a(2, 2);
    r(2, 0) = a(1, 0) * a(2, 1) - a(1, 1) * a(2, 0);
    r(0, 1) = a(2, 1) * a(0, 2) - a(2, 2) * a(0, 1);
    r(1, 1) = a(2, 2) * a(0, 0) - a(2, 0) * a(0, 2);
    r(2, 1) = a(2, 0) * a(0, 1) - a(2, 1) * a(0, 0);
    r(0, 2) = a(0, 1) * a(1, 2) - a(0, 2) * a(1, 1);
    r(1, 2) = a(0, 2) * a(1, 0) - a(0, 0) * a(1, 2);
    r(2, 2) = a(0, 0) * a(1, 1) - a(0, 1) * a(1, 0);
    double det = 0;
    det += a(0, 0) * r(0, 0);
    det += a(1, 0) * r(0, 1);
    det += a(2, 0) * r(0, 2);
    for (int i = 0; i < 9; ++i) R[i] /= det;
#undef a
#undef r
}
static void sub3(const double* a, const double* b, double* o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static void cross3(const double* a, const double* b, double* o) {
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* tet: A (rows p1-p0, p2-p0, p3-p0), |det| = 6 * (|(p1-p0).((p2-p0)x(p3-p0))| / 6).
 * face: unit normal from cell 0 towards cell 1, area, and for each side the reference
 * vertex (0..3) of each face node (natural coordinates). */
void or_geometry(int64_t np, const double* pts, int64_t nc, const int64_t* cells, int64_t nf,
                 const int64_t* fnodes, const int64_t* fcells, const int64_t* fnodes_r, double* cA, double* cdet,
                 double* fnormal, double* farea, int* fvert) {
    (void)np;
    for (int64_t e = 0; e < nc; ++e) {
        const double *p0 = pts + 3 * cells[4 * e], *p1 = pts + 3 * cells[4 * e + 1], *p2 = pts + 3 * cells[4 * e + 2],
                     *p3 = pts + 3 * cells[4 * e + 3];
        double v1[3], v2[3], v3[3], c[3];
        sub3(p1, p0, v1);
        sub3(p2, p0, v2);
        sub3(p3, p0, v3);
        cross3(v2, v3, c);
        double vol = fabs(dot3(v1, c)) / 6.0;
        if (cdet) cdet[e] = vol * 6.0;
        if (cA)
            for (int d = 0; d < 3; ++d) {
                cA[9 * e + d] = v1[d];
                cA[9 * e + 3 + d] = v2[d];
                cA[9 * e + 6 + d] = v3[d];
            }
    }
    for (int64_t f = 0; f < nf; ++f) {
        const int64_t* fn = fnodes + 3 * f;
        const int64_t* fr = fnodes_r ? fnodes_r + 3 * f : fn;
        const double *p0 = pts + 3 * fn[0], *p1 = pts + 3 * fn[1], *p2 = pts + 3 * fn[2];
        double v1[3], v2[3], n[3];
        sub3(p1, p0, v1);
        sub3(p2, p0, v2);
        cross3(v1, v2, n);
        int64_t c0 = fcells[2 * f], c1 = fcells[2 * f + 1];
        int periodic = c1 >= 0 && (fr[0] != fn[0] || fr[1] != fn[1] || fr[2] != fn[2]);
        if (c1 < 0 || periodic) {
            double p3[3] = {0, 0, 0}, v3[3], cc[3];
            for (int a = 0; a < 4; ++a) {
                int64_t v = cells[4 * c0 + a];
                if (v != fn[0] && v != fn[1] && v != fn[2]) memcpy(p3, pts + 3 * v, sizeof p3);
            }
            sub3(p3, p0, v3);
            cross3(v1, v2, cc);
            if (dot3(cc, v3) * dot3(cc, n) > 0) n[0] = -n[0], n[1] = -n[1], n[2] = -n[2];
        } else {
            double g0[3] = {0, 0, 0}, g1[3] = {0, 0, 0}, dg[3];
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) {
                    g0[d] += pts[3 * cells[4 * c0 + a] + d];
                    g1[d] += pts[3 * cells[4 * c1 + a] + d];
                }
            sub3(g1, g0, dg);
            if (dot3(n, dg) < 0) n[0] = -n[0], n[1] = -n[1], n[2] = -n[2];
        }
        double len = sqrt(dot3(n, n));
        if (farea) farea[f] = len * 0.5;
        if (fnormal)
            for (int d = 0; d < 3; ++d) fnormal[3 * f + d] = n[d] / len;
        if (fvert)
            for (int s = 0; s < 2; ++s)
                for (int k = 0; k < 3; ++k) {
                    int64_t c = fcells[2 * f + s], node = s ? fr[k] : fn[k];
                    int vtx = -1;
                    if (c >= 0)
                        for (int a = 0; a < 4; ++a)
                            if (cells[4 * c + a] == node) vtx = a;
                    fvert[6 * f + 3 * s + k] = vtx;
                }
    }
}

/* face point (u,v) of a face whose nodes are reference vertices vt[0..2] -> reference coords,
 * nc0*(1-u-v) + nc1*u + nc2*v (ExplicitConvection.cpp:129-134) */
static void face_to_cell(const int* vt, double u, double v, double xi[3]) {
    double nc[3][3] = {{0}};
    for (int k = 0; k < 3; ++k)
        if (vt[k] > 0) nc[k][vt[k] - 1] = 1.0;
    for (int d = 0; d < 3; ++d) xi[d] = nc[0][d] * (1 - u - v) + nc[1][d] * u + nc[2][d] * v;
}

/* ------------------------------------------------------------------ residual */
typedef struct {
    int order, N, Q, Qf, flux, parts, threads;
    double gamma;
    const double *qp, *qw, *fp, *fw;
} scheme_t;

static void add_scaled(double* dst, const double* src, double s, int n, int atomic_mode) {
    for (int i = 0; i < n; ++i) {
        if (atomic_mode) {
#pragma omp atomic update
            dst[i] += src[i] * s;
        } else {
            dst[i] += src[i] * s;
        }
    }
}

/*
 * R = cells + internals + boundarys for the selected parts (1, 2, 4), reference form.
 * bc[f] for boundary faces: 0 Dirichlet (ghost 2 U_D - U_in, U_D from UD[(i*Qf+g)*5+v]
 * for the i-th Dirichlet face in face order), 5 Pseudo3DZ (mirror rho w), others nothing.
 * flux: 0 LF, 1 Roe.  threads > 1: OpenMP dynamic schedule + atomics (reference mode).
 */
int or_residual(int order, int flux, double gamma, int64_t np, const double* pts, int64_t nc, const int64_t* cells,
                int64_t nf, const int64_t* fnodes, const int64_t* fcells, const int64_t* fnodes_r, const uint8_t* bc,
                const double* UD, const double* U, int parts, double* R, int threads) {
    scheme_t S;
    S.order = order;
    S.N = or_num_basis(order);
    S.Q = or_tet_rule(order, &S.qp, &S.qw);
    S.Qf = or_tri_rule(order, &S.fp, &S.fw);
    if (!S.Q || !S.Qf) return -1;
    const int N = S.N, NV = 5 * N;
    double* cA = (double*)malloc(sizeof(double) * 9 * nc);
    double* cdet = (double*)malloc(sizeof(double) * nc);
    double* fn = (double*)malloc(sizeof(double) * 3 * nf);
    double* fa = (double*)malloc(sizeof(double) * nf);
    int* fv = (int*)malloc(sizeof(int) * 6 * nf);
    int64_t* dslot = (int64_t*)malloc(sizeof(int64_t) * nf);
    or_geometry(np, pts, nc, cells, nf, fnodes, fcells, fnodes_r, cA, cdet, fn, fa, fv);
    int64_t nd = 0;
    for (int64_t f = 0; f < nf; ++f) dslot[f] = (fcells[2 * f + 1] < 0 && bc && bc[f] == 0) ? nd++ : -1;
    double* Rc = (double*)calloc((size_t)nc * NV, sizeof(double));
    double* Ri = (double*)calloc((size_t)nc * NV, sizeof(double));
    double* Rb = (double*)calloc((size_t)nc * NV, sizeof(double));
    const int am = threads > 1;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif

    /* eval_cells: ExplicitConvection.cpp:144-198 */
    if (parts & 1) {
        double* vb = (double*)malloc(sizeof(double) * S.Q * N);
        double* vg = (double*)malloc(sizeof(double) * S.Q * 3 * N);
        for (int q = 0; q < S.Q; ++q) or_basis(order, S.qp[3 * q], S.qp[3 * q + 1], S.qp[3 * q + 2], vb + q * N, vg + q * 3 * N);
#pragma omp parallel for schedule(dynamic) if (am)
        for (int64_t e = 0; e < nc; ++e) {
            const double* coef = U + e * NV;
            for (int q = 0; q < S.Q; ++q) {
                double Uq[5] = {0, 0, 0, 0, 0}, F[15];
                for (int k = 0; k < N; ++k)
                    for (int v = 0; v < 5; ++v) Uq[v] += vb[q * N + k] * coef[5 * k + v];
                or_flux(Uq, gamma, F);
                const double jw = S.qw[q] * cdet[e];
                for (int j = 0; j < N; ++j) {
                    double J[9], gp[3], fl[5];
                    or_inverse3(cA + 9 * e, J);
                    const double* gr = vg + (q * N + j) * 3;
                    for (int d = 0; d < 3; ++d) gp[d] = J[3 * d] * gr[0] + J[3 * d + 1] * gr[1] + J[3 * d + 2] * gr[2];
                    for (int v = 0; v < 5; ++v) fl[v] = F[3 * v] * gp[0] + F[3 * v + 1] * gp[1] + F[3 * v + 2] * gp[2];
                    add_scaled(Rc + e * NV + 5 * j, fl, -jw, 5, am);
                }
            }
        }
        free(vb);
        free(vg);
    }

    /* eval_internals (:201-272) and eval_boundarys (:21-128) */
    if (parts & 6) {
#pragma omp parallel for schedule(dynamic) if (am)
        for (int64_t f = 0; f < nf; ++f) {
            const int64_t c0 = fcells[2 * f], c1 = fcells[2 * f + 1];
            const int bnd = c1 < 0;
            if (bnd && !(parts & 4)) continue;
            if (!bnd && !(parts & 2)) continue;
            if (bnd && (!bc || (bc[f] != 0 && bc[f] != 5))) continue;
            double* phiL = (double*)malloc(sizeof(double) * N);
            double* phiR = (double*)malloc(sizeof(double) * N);
            const double jf = fa[f] * 2.0;
            for (int g = 0; g < S.Qf; ++g) {
                const double jw = S.fw[g] * jf;
                double xiL[3], UL[5] = {0, 0, 0, 0, 0}, UR[5] = {0, 0, 0, 0, 0}, Fh[5], t[5];
                face_to_cell(fv + 6 * f, S.fp[2 * g], S.fp[2 * g + 1], xiL);
                or_basis(order, xiL[0], xiL[1], xiL[2], phiL, NULL);
                for (int k = 0; k < N; ++k)
                    for (int v = 0; v < 5; ++v) UL[v] += phiL[k] * U[c0 * NV + 5 * k + v];
                if (!bnd) {
                    double xiR[3];
                    face_to_cell(fv + 6 * f + 3, S.fp[2 * g], S.fp[2 * g + 1], xiR);
                    or_basis(order, xiR[0], xiR[1], xiR[2], phiR, NULL);
                    for (int k = 0; k < N; ++k)
                        for (int v = 0; v < 5; ++v) UR[v] += phiR[k] * U[c1 * NV + 5 * k + v];
                } else if (bc[f] == 0) {
                    const double* ud = UD + (dslot[f] * S.Qf + g) * 5;
                    for (int v = 0; v < 5; ++v) UR[v] = ud[v] + (ud[v] - UL[v]);
                } else {
                    for (int v = 0; v < 5; ++v) UR[v] = UL[v];
                    UR[3] = -UL[3];
                }
                num_flux(flux, UL, UR, fn + 3 * f, gamma, Fh);
                for (int j = 0; j < N; ++j) {
                    for (int v = 0; v < 5; ++v) t[v] = Fh[v] * phiL[j];
                    add_scaled((bnd ? Rb : Ri) + c0 * NV + 5 * j, t, jw, 5, am);
                    if (!bnd) {
                        for (int v = 0; v < 5; ++v) t[v] = Fh[v] * phiR[j];
                        add_scaled(Ri + c1 * NV + 5 * j, t, -jw, 5, am);
                    }
                }
            }
            free(phiL);
            free(phiR);
        }
    }
    for (int64_t i = 0; i < nc * NV; ++i) R[i] = Rc[i] + Ri[i] + Rb[i];
    free(Rc);
    free(Ri);
    free(Rb);
    free(cA);
    free(cdet);
    free(fn);
    free(fa);
    free(fv);
    free(dslot);
    return 0;
}

/* diagonal mass M[e*5N+5k+v] = sum_q phi_k^2 w_q |det| (VortexBatch.cpp:155-163) */
int or_mass(int order, int64_t np, const double* pts, int64_t nc, const int64_t* cells, double* M) {
    const double *qp, *qw;
    const int Q = or_tet_rule(order, &qp, &qw), N = or_num_basis(order);
    double* det = (double*)malloc(sizeof(double) * nc);
    or_geometry(np, pts, nc, cells, 0, NULL, NULL, NULL, NULL, det, NULL, NULL, NULL);
    double* phi = (double*)malloc(sizeof(double) * Q * N);
    for (int q = 0; q < Q; ++q) or_basis(order, qp[3 * q], qp[3 * q + 1], qp[3 * q + 2], phi + q * N, NULL);
    for (int64_t e = 0; e < nc; ++e)
        for (int k = 0; k < N; ++k) {
            double s = 0;
            for (int q = 0; q < Q; ++q) s += phi[q * N + k] * phi[q * N + k] * qw[q] * det[e];
            for (int v = 0; v < 5; ++v) M[e * 5 * N + 5 * k + v] = s;
        }
    free(det);
    free(phi);
    return 0;
}

/* n_steps of Shu-Osher SSP-RK3 on dU/dt = -M^{-1} R(U), in place on U (no Dirichlet faces) */
int or_ssprk3(int order, int flux, double gamma, int64_t np, const double* pts, int64_t nc, const int64_t* cells,
              int64_t nf, const int64_t* fnodes, const int64_t* fcells, const int64_t* fnodes_r, const uint8_t* bc,
              double* U, double dt, int n_steps, int threads) {
    const int64_t n = nc * 5 * or_num_basis(order);
    double *M = malloc(sizeof(double) * n), *R = malloc(sizeof(double) * n), *U1 = malloc(sizeof(double) * n);
    or_mass(order, np, pts, nc, cells, M);
    static const double A[3] = {0.0, 0.75, 1.0 / 3.0}, B[3] = {1.0, 0.25, 2.0 / 3.0};
    for (int s = 0; s < n_steps; ++s) {
        for (int st = 0; st < 3; ++st) {
            const double* src = st == 0 ? U : U1;
            or_residual(order, flux, gamma, np, pts, nc, cells, nf, fnodes, fcells, fnodes_r, bc, NULL, src, 7, R,
                        threads);
            double* dst = st == 2 ? U : U1;
            for (int64_t i = 0; i < n; ++i) {
                double x = src[i] + dt * (-R[i] / M[i]);
                dst[i] = st == 0 ? B[0] * x : A[st] * U[i] + B[st] * x;
            }
        }
    }
    free(M);
    free(R);
    free(U1);
    return 0;
}

/* ------------------------------------------------------------------ initial data */
/* isentropic vortex of examples/Vortex/problem.h:17-110 with rho_inf = p_inf = 1;
 * P = gamma, strength, x0, y0, u_inf, v_inf, Lx, Ly (periods for the nearest image; 0 = none) */
void or_vortex_state(const double* P, double t, const double x[3], double U[5]) {
    const double g = P[0], phi = P[1];
    double dx = x[0] - (P[2] + P[4] * t), dy = x[1] - (P[3] + P[5] * t);
    if (P[6] > 0) dx -= P[6] * floor(dx / P[6] + 0.5);
    if (P[7] > 0) dy -= P[7] * floor(dy / P[7] + 0.5);
    const double r2 = dx * dx + dy * dy;
    const double slope = (phi * phi * (g - 1)) / (16 * M_PI * M_PI * g);
    const double rho = pow(1 - slope * exp(2 * (1 - r2)), 1.0 / (g - 1));
    const double u = P[4] - phi * dy / (2 * M_PI) * exp(1 - r2);
    const double v = P[5] + phi * dx / (2 * M_PI) * exp(1 - r2);
    const double p = pow(rho, g);
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = 0;
    U[4] = rho * (p / rho / (g - 1) + 0.5 * (u * u + v * v));
}

/* func2coef of the vortex on every cell (diagonal projection, long double diagonal) */
int or_project_vortex(int order, const double* P, double t, int64_t np, const double* pts, int64_t nc,
                      const int64_t* cells, double* U) {
    (void)np;
    const double *qp, *qw;
    const int Q = or_tet_rule(order, &qp, &qw), N = or_num_basis(order);
    double* phi = (double*)malloc(sizeof(double) * Q * N);
    long double* diag = (long double*)calloc(N, sizeof(long double));
    for (int q = 0; q < Q; ++q) or_basis(order, qp[3 * q], qp[3 * q + 1], qp[3 * q + 2], phi + q * N, NULL);
    for (int k = 0; k < N; ++k)
        for (int q = 0; q < Q; ++q) diag[k] += (long double)phi[q * N + k] * phi[q * N + k] * qw[q];
    for (int64_t e = 0; e < nc; ++e) {
        const double *p0 = pts + 3 * cells[4 * e], *p1 = pts + 3 * cells[4 * e + 1], *p2 = pts + 3 * cells[4 * e + 2],
                     *p3 = pts + 3 * cells[4 * e + 3];
        for (int k = 0; k < N; ++k) {
            double rhs[5] = {0, 0, 0, 0, 0};
            for (int q = 0; q < Q; ++q) {
                const double* xi = qp + 3 * q;
                double x[3], Us[5];
                for (int d = 0; d < 3; ++d)
                    x[d] = p0[d] * (1 - xi[0] - xi[1] - xi[2]) + p1[d] * xi[0] + p2[d] * xi[1] + p3[d] * xi[2];
                or_vortex_state(P, t, x, Us);
                for (int v = 0; v < 5; ++v) rhs[v] += Us[v] * (phi[q * N + k] * qw[q]);
            }
            for (int v = 0; v < 5; ++v) U[e * 5 * N + 5 * k + v] = rhs[v] / (double)diag[k];
        }
    }
    free(phi);
    free(diag);
    return 0;
}

/* ------------------------------------------------------------------ limiter
 * PositiveLimiter<Order, QuadC, OnlyNeigbAvg> of include/DG/DG_Schemes/PositiveLimiter.h,
 * with QuadC the order's automatic tetrahedron rule (the rule the reference's drivers pass,
 * examples/Eular2D/Sedov/Sedov.cpp:59,92).  Point values use the tetrahedron rule's points. */

/* values of the 5 variables of one element at the rule's points: out[q*5 + k] */
static void lim_point_values(int order, const double* c, double* out) {
    const double* pts;
    const double* w;
    const int Q = or_tet_rule(order, &pts, &w), N = or_num_basis(order);
    double phi[64];
    for (int q = 0; q < Q; ++q) {
        or_basis(order, pts[3 * q], pts[3 * q + 1], pts[3 * q + 2], phi, NULL);
        for (int k = 0; k < 5; ++k) {
            double v = 0.0;
            for (int l = 0; l < N; ++l) v += phi[l] * c[k + 5 * l];
            out[q * 5 + k] = v;
        }
    }
}

/* constructMinMax (PositiveLimiter.h:62-119): per face, the extremes of both neighbours'
 * (one neighbour's for a boundary face) values of Uprev at the rule's points, or of their
 * averages (only_avg), folded into per-cell bounds.  The face's lower bound follows the
 * statement min_U[k] = min(max_U[k], min(val1, val2)) of :98, which leaves the minimum over
 * the two cells at the LAST point of the rule. */
int or_limiter_bounds(int order, int64_t nc, int64_t nf, const int64_t* face_cells, const double* Uprev, int only_avg,
                      double* cmax, double* cmin) {
    const double* pts;
    const double* w;
    const int Q = or_tet_rule(order, &pts, &w), NV = 5 * or_num_basis(order);
    if (Q == 0) return -1;
    double* vals = (double*)malloc(sizeof(double) * (size_t)nc * Q * 5);
    if (!vals) return -2;
    for (int64_t e = 0; e < nc; ++e) lim_point_values(order, Uprev + e * NV, vals + e * Q * 5);
    for (int64_t i = 0; i < nc * 5; ++i) cmax[i] = -DBL_MAX, cmin[i] = DBL_MAX;
    for (int64_t f = 0; f < nf; ++f) {
        const int64_t c1 = face_cells[2 * f], c2 = face_cells[2 * f + 1] < 0 ? c1 : face_cells[2 * f + 1];
        double mx[5], mn[5];
        for (int k = 0; k < 5; ++k) mx[k] = -DBL_MAX, mn[k] = DBL_MAX;
        if (only_avg) {
            for (int k = 0; k < 5; ++k) {
                const double a = Uprev[c1 * NV + k], b = Uprev[c2 * NV + k];
                mx[k] = a > b ? a : b;
                mn[k] = a < b ? a : b;
            }
        } else {
            for (int q = 0; q < Q; ++q)
                for (int k = 0; k < 5; ++k) {
                    const double a = vals[(c1 * Q + q) * 5 + k], b = vals[(c2 * Q + q) * 5 + k];
                    const double hi = a > b ? a : b, lo = a < b ? a : b;
                    mx[k] = mx[k] > hi ? mx[k] : hi;
                    mn[k] = mx[k] < lo ? mx[k] : lo;
                }
        }
        for (int k = 0; k < 5; ++k) {
            if (mx[k] > cmax[c1 * 5 + k]) cmax[c1 * 5 + k] = mx[k];
            if (mn[k] < cmin[c1 * 5 + k]) cmin[c1 * 5 + k] = mn[k];
            if (face_cells[2 * f + 1] >= 0) {
                if (mx[k] > cmax[c2 * 5 + k]) cmax[c2 * 5 + k] = mx[k];
                if (mn[k] < cmin[c2 * 5 + k]) cmin[c2 * 5 + k] = mn[k];
            }
        }
    }
    free(vals);
    return 0;
}

static double lim_pressure(const double U[5], double gamma) { /* PositiveLimiter.h:121-132 */
    if (U[0] <= 1e-16) return -1.0;
    const double ke = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / (2.0 * U[0]);
    return (gamma - 1.0) * (U[4] - ke);
}

/* apply_1 (bounds, :135-170) when parts & 1, then apply_21 (density, :171-198) and apply_22
 * (pressure, :199-259) when parts & 2, cell by cell on U in place.  Cell averages are the
 * mode-0 coefficients (phi_0 = 1, the other modes have zero mean). */
int or_limiter_apply(int order, int64_t nc, const double* cmax, const double* cmin, double gamma, int parts,
                     double* U) {
    const double* pts;
    const double* w;
    const int Q = or_tet_rule(order, &pts, &w), N = or_num_basis(order), NV = 5 * N;
    if (Q == 0) return -1;
    double vals[128 * 5];
    for (int64_t e = 0; e < nc; ++e) {
        double* c = U + e * NV;
        if (parts & 1) {
            lim_point_values(order, c, vals);
            for (int k = 0; k < 5; ++k) {
                const double avg = c[k], gmin = cmin[e * 5 + k], gmax = cmax[e * 5 + k];
                double vmin = vals[k], vmax = vals[k];
                for (int q = 1; q < Q; ++q) {
                    const double v = vals[q * 5 + k];
                    if (v < vmin) vmin = v;
                    if (v > vmax) vmax = v;
                }
                if (vmin < gmin || vmax > gmax) {
                    const double tmin = (avg != vmin) ? (avg - gmin) / (avg - vmin + 1e-32) : 0.0;
                    const double tmax = (avg != vmax) ? (gmax - avg) / (vmax - avg + 1e-32) : 0.0;
                    double th = tmin < tmax ? tmin : tmax;
                    if (1.0 < th) th = 1.0;
                    for (int l = 1; l < N; ++l) c[k + 5 * l] *= th;
                }
            }
        }
        if (parts & 2) {
            const double eps = 1e-16;
            lim_point_values(order, c, vals);
            double rmin = vals[0];
            for (int q = 1; q < Q; ++q)
                if (vals[q * 5] < rmin) rmin = vals[q * 5];
            if (rmin < eps) {
                const double ravg = c[0] > eps ? c[0] : eps;
                const double num = ravg - eps, den = ravg - rmin;
                double th = (den < 1e-32) ? 0.0 : (num / den < 1.0 ? num / den : 1.0);
                for (int l = 1; l < N; ++l) c[5 * l] *= th;
            }
            lim_point_values(order, c, vals);
            double avg[5], thp = 1.0;
            for (int k = 0; k < 5; ++k) avg[k] = c[k];
            for (int q = 0; q < Q; ++q) {
                const double* P = vals + q * 5;
                if (lim_pressure(P, gamma) >= 0) continue;
                double lo = 0.0, hi = 1.0, mid = 0.5;
                for (int it = 0; it < 100; ++it) {
                    mid = 0.5 * (lo + hi);
                    double Um[5];
                    for (int k = 0; k < 5; ++k) Um[k] = (1.0 - mid) * avg[k] + mid * P[k];
                    if (lim_pressure(Um, gamma) < 0) hi = mid;
                    else lo = mid;
                }
                if (mid < thp) thp = mid;
            }
            for (int k = 0; k < 5; ++k)
                for (int l = 1; l < N; ++l) c[k + 5 * l] *= thp;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ implicit (Picard) blocks */

/* dF_d/dU, d = x, y, z (EulerPhysicalFlux.cpp:21-42): J[d][5 r + c] */
void or_jacobian(const double U[5], double gamma, double J[3][25]) {
    const double u = U[1] / U[0], v = U[2] / U[0], w = U[3] / U[0], e = U[4] / U[0];
    const double q2 = u * u + v * v + w * w, g1 = gamma - 1.0;
    const double vel[3] = {u, v, w};
    for (int d = 0; d < 3; ++d) {
        double* A = J[d];
        for (int i = 0; i < 25; ++i) A[i] = 0.0;
        const double a = vel[d];
        A[1 + d] = 1.0;  /* mass row: the momentum component of direction d */
        for (int m = 0; m < 3; ++m) {  /* momentum rows 1..3 */
            double* row = A + 5 * (1 + m);
            row[0] = -vel[m] * a + (m == d ? 0.5 * g1 * q2 : 0.0);
            row[1 + m] += a;
            row[1 + d] += vel[m];
            if (m == d)
                for (int k = 0; k < 3; ++k) row[1 + k] -= g1 * vel[k];
            if (m == d) row[4] = g1;
        }
        double* er = A + 20;  /* energy row */
        er[0] = a * (-gamma * e + g1 * q2);
        for (int k = 0; k < 3; ++k) er[1 + k] = -g1 * a * vel[k];
        er[1 + d] = gamma * e - 0.5 * g1 * (2.0 * a * a + q2);
        er[4] = gamma * a;
    }
}

static void normal_jacobian(const double J[3][25], const double n[3], double Fn[25]) {
    for (int i = 0; i < 25; ++i) Fn[i] = J[0][i] * n[0] + J[1][i] * n[1] + J[2][i] * n[2];
}

/* ghost-state linearisation of one boundary point (ImplicitConvection.cpp:20-167):
 * dF (5x5) multiplies the interior trace, drhs is the explicit remainder */
static void implicit_boundary(int bc, const double Uin[5], const double* UD, const double n[3], double gamma,
                              double dF[25], double drhs[5]) {
    double Ug[5], Jg[3][25], Fn[25];
    for (int i = 0; i < 25; ++i) dF[i] = 0.0;
    for (int v = 0; v < 5; ++v) drhs[v] = 0.0;
    if (bc == 0 || bc == 9 || bc == 10) { /* Dirichlet, WallTD, WallTN */
        for (int v = 0; v < 5; ++v) Ug[v] = UD[v] + (UD[v] - Uin[v]);
        if (bc != 0) {
            Ug[0] = Uin[0];
            const double cvT_in = Uin[4] - 0.5 * (Uin[1] * Uin[1] + Uin[2] * Uin[2] + Uin[3] * Uin[3]) / Uin[0];
            const double cvT_g = (bc == 9) ? 2 * 300 - cvT_in : cvT_in;
            Ug[4] = cvT_g + 0.5 * (Ug[1] * Ug[1] + Ug[2] * Ug[2] + Ug[3] * Ug[3]) / Ug[0];
        }
        for (int v = 0; v < 5; ++v) Ug[v] = 0.5 * (Uin[v] + Ug[v]);
        or_jacobian(Ug, gamma, Jg);
    } else if (bc == 3 || bc == 4 || bc == 5) { /* Pseudo3DX / Y / Z: mirror of that momentum */
        for (int v = 0; v < 5; ++v) Ug[v] = Uin[v];
        Ug[bc - 2] = -Uin[bc - 2];
        double Um[5];
        for (int v = 0; v < 5; ++v) Um[v] = 0.5 * (Uin[v] + Ug[v]);
        or_jacobian(Um, gamma, Jg);
    } else {
        return;
    }
    normal_jacobian(Jg, n, Fn);
    const double lam = or_wave_speed(Uin, Ug, gamma);
    for (int v = 0; v < 5; ++v) Ug[v] = 0.5 * (Uin[v] + Ug[v]);
    for (int r = 0; r < 5; ++r) {
        double fu = 0.0;
        for (int c = 0; c < 5; ++c) {
            dF[5 * r + c] = 0.5 * (Fn[5 * r + c] + (r == c ? lam : 0.0));
            fu += Fn[5 * r + c] * Ug[c];
        }
        drhs[r] = 0.5 * (fu - lam * Ug[r]);
    }
}

/* ImplicitConvection<Order>::assemble (src/DG/DG_Schemes/ImplicitConvection.cpp:169-448): the Picard
 * blocks of R(U) ~ A(U_lin) U - rhs, linearised at U, accumulated into (zeroed by the caller)
 *   diag[e]    (NB x NB, NB = 5N, row 5j+v = test function j / variable v, column 5i+u),
 *   off[2f]    block (L, R) of interior face f, off[2f+1] block (R, L),
 *   rhs[e]     (NB).
 * Boundary data UD: per face of type Dirichlet / WallTD / WallTN in face order, Qf x 5 each. */
int or_implicit_assemble(int order, double gamma, int64_t np, const double* pts, int64_t nc, const int64_t* cells,
                         int64_t nf, const int64_t* fnodes, const int64_t* fcells, const int64_t* fnodes_r,
                         const uint8_t* bc, const double* UD, const double* U, double* diag, double* off,
                         double* rhs) {
    scheme_t S;
    S.order = order;
    S.N = or_num_basis(order);
    S.Q = or_tet_rule(order, &S.qp, &S.qw);
    S.Qf = or_tri_rule(order, &S.fp, &S.fw);
    if (!S.Q || !S.Qf) return -1;
    const int N = S.N, NB = 5 * N;
    const size_t BB = (size_t)NB * NB;
    double* cA = (double*)malloc(sizeof(double) * 9 * nc);
    double* cdet = (double*)malloc(sizeof(double) * nc);
    double* fn = (double*)malloc(sizeof(double) * 3 * nf);
    double* fa = (double*)malloc(sizeof(double) * nf);
    int* fv = (int*)malloc(sizeof(int) * 6 * nf);
    or_geometry(np, pts, nc, cells, nf, fnodes, fcells, fnodes_r, cA, cdet, fn, fa, fv);
    double* vb = (double*)malloc(sizeof(double) * S.Q * N);
    double* vg = (double*)malloc(sizeof(double) * S.Q * 3 * N);
    for (int q = 0; q < S.Q; ++q) or_basis(order, S.qp[3 * q], S.qp[3 * q + 1], S.qp[3 * q + 2], vb + q * N, vg + q * 3 * N);
    double phiL[64], phiR[64];

    /* assemble_cells (:266-325): block (j, i) -= sum_q w |det| phi_i (sum_d J_d(U_q) d phi_j / d x_d) */
    for (int64_t e = 0; e < nc; ++e) {
        const double* coef = U + e * NB;
        double Ai[9];
        or_inverse3(cA + 9 * e, Ai);
        for (int q = 0; q < S.Q; ++q) {
            double Uq[5] = {0, 0, 0, 0, 0}, J[3][25];
            for (int k = 0; k < N; ++k)
                for (int v = 0; v < 5; ++v) Uq[v] += vb[q * N + k] * coef[5 * k + v];
            or_jacobian(Uq, gamma, J);
            const double jw = S.qw[q] * cdet[e];
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) {
                    const double* gr = vg + (q * N + j) * 3;
                    double gp[3];
                    for (int d = 0; d < 3; ++d) gp[d] = Ai[3 * d] * gr[0] + Ai[3 * d + 1] * gr[1] + Ai[3 * d + 2] * gr[2];
                    const double s = vb[q * N + i] * jw;
                    for (int v = 0; v < 5; ++v)
                        for (int u = 0; u < 5; ++u)
                            diag[e * BB + (size_t)(5 * j + v) * NB + 5 * i + u] -=
                                (J[0][5 * v + u] * gp[0] + J[1][5 * v + u] * gp[1] + J[2][5 * v + u] * gp[2]) * s;
                }
        }
    }

    /* assemble_internals (:329-420) and assemble_boundarys (:20-167) */
    int64_t nd = 0;
    for (int64_t f = 0; f < nf; ++f) {
        const int64_t c0 = fcells[2 * f], c1 = fcells[2 * f + 1];
        const double jf = fa[f] * 2.0;
        const double* n = fn + 3 * f;
        int64_t slot = -1;
        if (c1 < 0 && bc && (bc[f] == 0 || bc[f] == 9 || bc[f] == 10)) slot = nd++;
        for (int g = 0; g < S.Qf; ++g) {
            const double wgt = S.fw[g] * jf;
            double xi[3], UL[5] = {0, 0, 0, 0, 0}, UR[5] = {0, 0, 0, 0, 0};
            face_to_cell(fv + 6 * f, S.fp[2 * g], S.fp[2 * g + 1], xi);
            or_basis(order, xi[0], xi[1], xi[2], phiL, NULL);
            for (int k = 0; k < N; ++k)
                for (int v = 0; v < 5; ++v) UL[v] += phiL[k] * U[c0 * NB + 5 * k + v];
            if (c1 >= 0) {
                face_to_cell(fv + 6 * f + 3, S.fp[2 * g], S.fp[2 * g + 1], xi);
                or_basis(order, xi[0], xi[1], xi[2], phiR, NULL);
                for (int k = 0; k < N; ++k)
                    for (int v = 0; v < 5; ++v) UR[v] += phiR[k] * U[c1 * NB + 5 * k + v];
                double JL[3][25], JR[3][25], FnL[25], FnR[25], dFL[25], dFR[25];
                or_jacobian(UL, gamma, JL);
                or_jacobian(UR, gamma, JR);
                normal_jacobian(JL, n, FnL);
                normal_jacobian(JR, n, FnR);
                const double lam = or_wave_speed(UL, UR, gamma);
                for (int r = 0; r < 5; ++r)
                    for (int c = 0; c < 5; ++c) {
                        dFL[5 * r + c] = 0.5 * (FnL[5 * r + c] + (r == c ? lam : 0.0));
                        dFR[5 * r + c] = 0.5 * (FnR[5 * r + c] - (r == c ? lam : 0.0));
                    }
                for (int i = 0; i < N; ++i)
                    for (int j = 0; j < N; ++j)
                        for (int v = 0; v < 5; ++v)
                            for (int u = 0; u < 5; ++u) {
                                const size_t ix = (size_t)(5 * j + v) * NB + 5 * i + u;
                                diag[c0 * BB + ix] += dFL[5 * v + u] * (phiL[i] * phiL[j]) * wgt;
                                off[(2 * f) * BB + ix] += dFR[5 * v + u] * (phiR[i] * phiL[j]) * wgt;
                                off[(2 * f + 1) * BB + ix] -= dFL[5 * v + u] * (phiL[i] * phiR[j]) * wgt;
                                diag[c1 * BB + ix] -= dFR[5 * v + u] * (phiR[i] * phiR[j]) * wgt;
                            }
            } else if (bc) {
                double dF[25], drhs[5];
                implicit_boundary(bc[f], UL, slot >= 0 ? UD + (slot * S.Qf + g) * 5 : NULL, n, gamma, dF, drhs);
                for (int i = 0; i < N; ++i)
                    for (int j = 0; j < N; ++j)
                        for (int v = 0; v < 5; ++v)
                            for (int u = 0; u < 5; ++u)
                                diag[c0 * BB + (size_t)(5 * j + v) * NB + 5 * i + u] += dF[5 * v + u] * (phiL[i] * phiL[j]) * wgt;
                for (int j = 0; j < N; ++j)
                    for (int v = 0; v < 5; ++v) rhs[c0 * NB + 5 * j + v] -= drhs[v] * phiL[j] * wgt;
            }
        }
    }
    free(cA);
    free(cdet);
    free(fn);
    free(fa);
    free(fv);
    free(vb);
    free(vg);
    return 0;
}

/* threads the OpenMP runtime would use (for the baseline record) */
int or_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
