// Device half of libfvdg.so: the solver context (one GPU, one stream), device
// copies of mesh geometry and reference tables, and the residual / SSP-RK3
// kernels for gfx950.
//
// Device state layout: element-major, coefficient kv = 5k+v of element e at
// U[e*5N + kv] -- the reference's LongVector layout, so upload/download are plain
// copies, one element's 5N coefficients (800 B at P3) are one contiguous block for
// the neighbour gathers of the face kernel, and the matrix-core volume kernel
// stages its 16-element tiles through LDS with full-width coalesced loads.
//
// Kernels (fvdg_kernels_fast.hpp): k_vol_mfma (volume, FP64 MFMA), then the faces in
// two passes: k_face_flux_dpp / k_face_flux_lds (every face once: traces + Riemann flux) and
// k_face_test_mfma / k_face_test_lds (test functions per element, fused with the SSP-RK3
// stage update inside the time stepper).  Environment switches: FVDG_FACE_FLUX=lds and
// FVDG_FACE_TEST=lds select the LDS forms of the two passes; FVDG_FACES=owner selects the
// one-pass owner-computes face kernel k_surf_lds; FVDG_KERNELS=simple selects the
// plain per-point kernels of fvdg_kernels.hpp (the GPU cross-check).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>

#include "fvdg_device.hpp"
#include "fvdg_internal.hpp"
#include "fvdg_kernels.hpp"
#include "fvdg_kernels_fast.hpp"
#include "fvdg_limiter.hpp"
#include "fvdg_diagnostics.hpp"
#include "fvdg_bf16x3.hpp"
#include "fvdg_elem.hpp"
#include "fvdg_facetile.hpp"

using namespace fvdg;

// Scratch memory of fvdg_stable_dt_noalloc: the partial maxima of the wave-speed reduction
// (kScratchPartials, one per block of a grid-stride launch) and the folded maximum after them.
// One copy per GPU for the life of the process (32 KiB); host code uses the per-device address
// from hipGetSymbolAddress (g_scratch), never the array's name.  Calls that overlap on one GPU
// share it and race.
constexpr int kScratchPartials = 4096;
static __device__ __align__(256) double scratch[kScratchPartials + 1];

namespace {

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mutex;

// device address of `scratch` on the current device `device` (set once per device)
int scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(FVDG_ERR_UNSUPPORTED, "device index beyond 64");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch[device]) return FVDG_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(FVDG_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<double*>(p);
    return FVDG_OK;
}

int check_ready(fvdg_ctx* c, bool need_state) {
    if (!c) return fail(FVDG_ERR_ARG, "null context");
    if (c->order < 0) return fail(FVDG_ERR_STATE, "fvdg_set_scheme has not been called");
    if (!c->have_mesh) return fail(FVDG_ERR_STATE, "fvdg_set_mesh has not been called");
    if (need_state && !c->have_state) return fail(FVDG_ERR_STATE, "no state uploaded");
    return FVDG_OK;
}

inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// resident blocks per compute unit of a kernel (at least 1), queried once per kernel: the
// persistent face kernels size their grids with it on every stage
template <auto Kern, int THREADS>
int resident_blocks() {
    static const int n = [] {
        int occ = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, Kern, THREADS, 0) != hipSuccess || occ < 1) occ = 1;
        return occ;
    }();
    return n;
}

// bf16 bits of a float, rounded to nearest even
uint16_t bf16_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
double bf16_value(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
// x = x0 + x1 + x2 in bf16 parts, written to parts[p * stride + i]
void split_bf16x3(double x, std::vector<uint16_t>& parts, size_t stride, size_t i) {
    for (int p = 0; p < 3; ++p) {
        const uint16_t b = bf16_bits((float)x);
        parts[p * stride + i] = b;
        x -= bf16_value(b);
    }
}

// BF16x3 operand tables of the volume kernel, in the per-lane order of fvdg_bf3::k_vol_bf16x3
template <int P>
int upload_bf16x3_tables_t(fvdg_ctx* c) {
    using S = fvdg_bf3::Bf3Shape<P>;
    const SchemeTables& T = c->tab;
    const int N = T.N, Q = T.Q;
    std::vector<uint16_t> VB(3 * (size_t)S::VB, 0), GB(3 * (size_t)S::GB, 0);
    for (int t = 0; t < S::QT; ++t)
        for (int ks = 0; ks < S::KS; ++ks)
            for (int l = 0; l < 64; ++l)
                for (int i = 0; i < 8; ++i) {
                    const int q = 16 * t + l % 16, k = 32 * ks + 8 * (l / 16) + i;
                    const double x = (q < Q && k < N) ? T.V[q * N + k] : 0.0;
                    split_bf16x3(x, VB, S::VB, (((size_t)t * S::KS + ks) * 64 + l) * 8 + i);
                }
    for (int t = 0; t < S::QT; ++t)
        for (int s = 0; s < 2; ++s)
            for (int jt = 0; jt < S::JT; ++jt)
                for (int l = 0; l < 64; ++l)
                    for (int i = 0; i < 8; ++i) {
                        const int pr = 8 * s + i, r = pr / 3, d = pr % 3;
                        const int j = 16 * jt + l % 16, q = 16 * t + 4 * (l / 16) + r;
                        const double x = (pr < 12 && j < N && q < Q) ? T.G[(q * 3 + d) * N + j] : 0.0;
                        split_bf16x3(x, GB, S::GB, ((((size_t)t * 2 + s) * S::JT + jt) * 64 + l) * 8 + i);
                    }
    HIP_TRY(c->tVB3.upload(VB, c->stream));
    HIP_TRY(c->tGB3.upload(GB, c->stream));
    return FVDG_OK;
}
int upload_bf16x3_tables(fvdg_ctx* c) {
    switch (c->order) {
        case 0: return upload_bf16x3_tables_t<0>(c);
        case 1: return upload_bf16x3_tables_t<1>(c);
        case 2: return upload_bf16x3_tables_t<2>(c);
        case 3: return upload_bf16x3_tables_t<3>(c);
        case 4: return upload_bf16x3_tables_t<4>(c);
        default: return upload_bf16x3_tables_t<5>(c);
    }
}

// Per vertex combination cb (24) the face-tile operand tables of k_face_tile in the lane layouts of
// TileLayout<T> (zero outside the table):
//   At[cb][mb][ks][lane] = Phi_cb[row_point(mb, lane % 16)][4 ks + lane / 16]       (traces)
//   Ae[cb][jt][ks][lane] = Phi_cb[point(ks, lane / 16)][16 jt + lane % 16]         (test)
// Orders whose face rule has not 24 points get empty tables (the face-tile path is not used).
template <typename T>
int upload_face_tile_tables(fvdg_ctx* c, DevBuf<T>& At, DevBuf<T>& Ae, DevBuf<T>* ICc) {
    const SchemeTables& S = c->tab;
    if (S.Qf != 24 || S.N < 10 || S.N > 35) {
        At.release();
        Ae.release();
        if (ICc) ICc->release();
        return FVDG_OK;
    }
    using LY = TileLayout<T>;
    const int N = S.N, KS = (N + 3) / 4, JT = (N + 15) / 16;
    auto phi = [&](int cb, int g, int k) { return S.Phi[((size_t)cb * S.Qf + g) * N + k]; };
    std::vector<T> at((size_t)24 * 2 * KS * 64, (T)0), ae((size_t)24 * JT * 6 * 64, (T)0);
    for (int cb = 0; cb < 24; ++cb) {
        for (int mb = 0; mb < 2; ++mb)
            for (int ks = 0; ks < KS; ++ks)
                for (int ln = 0; ln < 64; ++ln) {
                    const int g = LY::row_point(mb, ln % 16), k = 4 * ks + ln / 16;
                    if (g >= 0 && k < N) at[(((size_t)cb * 2 + mb) * KS + ks) * 64 + ln] = (T)phi(cb, g, k);
                }
        for (int jt = 0; jt < JT; ++jt)
            for (int ks = 0; ks < 6; ++ks)
                for (int ln = 0; ln < 64; ++ln) {
                    const int g = LY::point(ks, ln / 16), j = 16 * jt + ln % 16;
                    if (j < N) ae[(((size_t)cb * JT + jt) * 6 + ks) * 64 + ln] = (T)phi(cb, g, j);
                }
    }
    HIP_TRY(At.upload(at, c->stream));
    HIP_TRY(Ae.upload(ae, c->stream));
    if (ICc) {
        // ICc[cb][j] = sum_g w_g Phi_cb[g][j], rows padded to NP
        const int NP = (N + 3) / 4 * 4;
        std::vector<T> ic((size_t)24 * NP, (T)0);
        for (int cb = 0; cb < 24; ++cb)
            for (int j = 0; j < N; ++j) {
                double sum = 0.0;
                for (int g = 0; g < S.Qf; ++g) sum += S.fq_w[g] * phi(cb, g, j);
                ic[(size_t)cb * NP + j] = (T)sum;
            }
        HIP_TRY(ICc->upload(ic, c->stream));
    }
    return FVDG_OK;
}

// BF16x3 operand tables of k_face_tile_bf3 (three bf16 parts, part stride = one table):
//   At3[cb][mb][ks][lane][i] = Phi_cb[row_point(mb, lane % 16)][32 ks + 8 (lane / 16) + i]   (traces)
//   Ae3[cb][jt][lane][i]     = Phi_cb[point(i, lane / 16)][16 jt + lane % 16], i < 6; 0 for i = 6, 7 (test)
int upload_face_tile_bf3_tables(fvdg_ctx* c) {
    const SchemeTables& S = c->tab;
    if (S.Qf != 24 || S.N < 10 || S.N > 35) {
        c->tAt3.release();
        c->tAe3.release();
        return FVDG_OK;
    }
    using LY = TileLayout<float>;
    const int N = S.N, KS3 = (N + 31) / 32, JT = (N + 15) / 16;
    auto phi = [&](int cb, int g, int k) { return S.Phi[((size_t)cb * S.Qf + g) * N + k]; };
    const size_t nat = (size_t)24 * 2 * KS3 * 512, nae = (size_t)24 * JT * 512;
    std::vector<uint16_t> at(3 * nat, 0), ae(3 * nae, 0);
    for (int cb = 0; cb < 24; ++cb) {
        for (int mb = 0; mb < 2; ++mb)
            for (int ks = 0; ks < KS3; ++ks)
                for (int ln = 0; ln < 64; ++ln)
                    for (int i = 0; i < 8; ++i) {
                        const int g = LY::row_point(mb, ln % 16), k = 32 * ks + 8 * (ln / 16) + i;
                        const double x = (g >= 0 && k < N) ? phi(cb, g, k) : 0.0;
                        split_bf16x3(x, at, nat, ((((size_t)cb * 2 + mb) * KS3 + ks) * 64 + ln) * 8 + i);
                    }
        for (int jt = 0; jt < JT; ++jt)
            for (int ln = 0; ln < 64; ++ln)
                for (int i = 0; i < 8; ++i) {
                    const int j = 16 * jt + ln % 16;
                    const double x = (i < 6 && j < N) ? phi(cb, LY::point(i, ln / 16), j) : 0.0;
                    split_bf16x3(x, ae, nae, (((size_t)cb * JT + jt) * 64 + ln) * 8 + i);
                }
    }
    HIP_TRY(c->tAt3.upload(at, c->stream));
    HIP_TRY(c->tAe3.upload(ae, c->stream));
    return FVDG_OK;
}

int upload_tables(fvdg_ctx* c) {
    const SchemeTables& T = c->tab;
    HIP_TRY(c->tV.upload(T.V, c->stream));
    HIP_TRY(c->tG.upload(T.G, c->stream));
    HIP_TRY(c->tWq.upload(T.vq_w, c->stream));
    HIP_TRY(c->tPhi.upload(T.Phi, c->stream));
    HIP_TRY(c->tWf.upload(T.fq_w, c->stream));
    HIP_TRY(c->tMhat.upload(T.mass_hat, c->stream));
    // MFMA operand tables, zero padded to whole 16x4 tiles
    const int QPAD = (T.Q + 15) / 16 * 16, KPAD = (T.N + 3) / 4 * 4, JPAD = (T.N + 15) / 16 * 16;
    std::vector<double> VA(QPAD * KPAD, 0.0), GA(QPAD * 3 * JPAD, 0.0), WA(QPAD, 0.0);
    for (int q = 0; q < T.Q; ++q) {
        WA[q] = T.vq_w[q];
        for (int k = 0; k < T.N; ++k) VA[q * KPAD + k] = T.V[q * T.N + k];
        for (int d = 0; d < 3; ++d)
            for (int j = 0; j < T.N; ++j) GA[(q * 3 + d) * JPAD + j] = T.G[(q * 3 + d) * T.N + j];
    }
    std::vector<double> PT(T.Phi.size());
    for (int cb = 0; cb < 24; ++cb)
        for (int g = 0; g < T.Qf; ++g)
            for (int k = 0; k < T.N; ++k) PT[((size_t)cb * T.N + k) * T.Qf + g] = T.Phi[((size_t)cb * T.Qf + g) * T.N + k];
    HIP_TRY(c->tVA.upload(VA, c->stream));
    HIP_TRY(c->tGA.upload(GA, c->stream));
    HIP_TRY(c->tWqa.upload(WA, c->stream));
    HIP_TRY(c->tPhiT.upload(PT, c->stream));
    std::vector<double> PC;
    std::vector<uint8_t> canon;
    std::vector<int32_t> perm;
    const double dev = build_canonical_face_tables(T, PC, canon, perm);
    double scale = 1.0;
    for (double x : T.Phi) scale = std::max(scale, std::abs(x));
    if (!(dev <= 1e-13 * scale))
        return fail(FVDG_ERR_UNSUPPORTED, "face rule is not invariant under vertex rotation (deviation " +
                                              std::to_string(dev) + ")");
    HIP_TRY(c->tPhiC.upload(PC, c->stream));
    {
        // canonical tables as per-lane B operands of k_face_test_mfma: [c][ks][jt][lane] holds
        // PC[c][g = 4 ks + lane / 16][j = 16 jt + lane % 16], zero outside the table
        const int KS = (T.Qf + 3) / 4, JT = (T.N + 15) / 16;
        std::vector<double> PB((size_t)8 * KS * JT * 64, 0.0);
        for (int cc = 0; cc < 8; ++cc)
            for (int ks = 0; ks < KS; ++ks)
                for (int jt = 0; jt < JT; ++jt)
                    for (int ln = 0; ln < 64; ++ln) {
                        const int g = 4 * ks + ln / 16, j = 16 * jt + ln % 16;
                        if (g < T.Qf && j < T.N)
                            PB[(((size_t)cc * KS + ks) * JT + jt) * 64 + ln] = PC[((size_t)cc * T.Qf + g) * T.N + j];
                    }
        HIP_TRY(c->tPhiB.upload(PB, c->stream));
        HIP_TRY(c->tPhiB32.upload(std::vector<float>(PB.begin(), PB.end()), c->stream));
    }
    HIP_TRY(c->tVA32.upload(std::vector<float>(VA.begin(), VA.end()), c->stream));
    HIP_TRY(c->tGA32.upload(std::vector<float>(GA.begin(), GA.end()), c->stream));
    HIP_TRY(c->tPhiC32.upload(std::vector<float>(PC.begin(), PC.end()), c->stream));
    std::vector<double> Gs(T.N * 3, 0.0), IC(8 * T.N, 0.0);
    for (int j = 0; j < T.N; ++j)
        for (int d = 0; d < 3; ++d)
            for (int q = 0; q < T.Q; ++q) Gs[j * 3 + d] += T.vq_w[q] * T.G[(q * 3 + d) * T.N + j];
    for (int cc = 0; cc < 8; ++cc)
        for (int j = 0; j < T.N; ++j)
            for (int g = 0; g < T.Qf; ++g) IC[cc * T.N + j] += T.fq_w[g] * PC[((size_t)cc * T.Qf + g) * T.N + j];
    HIP_TRY(c->tGs.upload(Gs, c->stream));
    HIP_TRY(c->tIC.upload(IC, c->stream));
    {
        int rc = upload_face_tile_tables<double>(c, c->tAt, c->tAe, nullptr);
        if (rc) return rc;
        rc = upload_face_tile_tables<float>(c, c->tAt32, c->tAe32, &c->tICc32);
        if (rc) return rc;
        rc = upload_face_tile_bf3_tables(c);
        if (rc) return rc;
    }
    HIP_TRY(c->tCanon.upload(canon, c->stream));
    HIP_TRY(c->tPerm.upload(perm, c->stream));
    return upload_bf16x3_tables(c);
}

// the face-tile path (k_face_tile + k_elem_mfma with contributions): P2-P4 (24-point face rule),
// FP32 contractions by default, FP64 too with FVDG_FACE_TILE=all
bool use_face_tile(const fvdg_ctx* c) {
    if (c->face_tile == 0 || c->simple || c->owner_faces || c->lds_face_test || c->lds_face_flux || !c->elem_fused)
        return false;
    if (c->order < 2 || c->order > 4 || c->tab.Qf != 24 || c->n_face_tiles == 0) return false;
    if (c->precision == FVDG_PREC_FP32 || c->precision == FVDG_PREC_BF16X3) return true;
    return c->precision == FVDG_PREC_FP64 && c->face_tile == 2;
}

// BF16x3 (unless FVDG_BF16X3_FACES=0): the face tiles' contractions as bf16 part products too
bool bf3_face_tiles(const fvdg_ctx* c) {
    return c->precision == FVDG_PREC_BF16X3 && c->bf3_faces && c->tAt3.n > 0 && c->tAe3.n > 0;
}

// Scratch buffers of the contraction precision (face fluxes, and for the mixed path the
// FP32 volume residual); the state buffers are not touched.
int precision_buffers(fvdg_ctx* c) {
    if (c->order < 0 || !c->have_mesh) return FVDG_OK;
    const size_t nfh = (size_t)c->F * c->tab.Qf * 5;
    if (c->precision == FVDG_PREC_FP64 || c->simple) {
        if (c->Fh.n != nfh) HIP_TRY(c->Fh.alloc(nfh));
    } else {
        if (c->Fh32.n != nfh) HIP_TRY(c->Fh32.alloc(nfh));
        if (c->Fref.n != (size_t)c->F * 5) HIP_TRY(c->Fref.alloc((size_t)c->F * 5));
    }
    // face-tile contributions, allocated when the path can serve this precision (an ensemble's
    // autotuner switches the path on and off per order group)
    const size_t ncf = (size_t)c->F * 2 * 5 * ((c->tab.N + 3) / 4 * 4);
    const int ft = c->face_tile;
    if (c->precision != FVDG_PREC_FP64 && c->face_tile == 0) c->face_tile = 1;
    const bool tiles = use_face_tile(c);
    c->face_tile = ft;
    if (tiles) {
        if (c->precision == FVDG_PREC_FP64) {
            if (c->Cf.n != ncf) HIP_TRY(c->Cf.alloc(ncf));
        } else if (c->Cf32.n != ncf) {
            HIP_TRY(c->Cf32.alloc(ncf));
        }
    }
    return FVDG_OK;
}

// lanes per element of the per-point element kernels (limiter, diagnostics): the power of
// two >= Q; 256 / lanes elements per block
inline int limiter_lanes(int Q) { return Q <= 16 ? 16 : Q <= 32 ? 32 : Q <= 64 ? 64 : 128; }
inline unsigned point_blocks(const fvdg_ctx* c) { return nblocks(c->E, 256 / limiter_lanes(c->tab.Q)); }

int alloc_state(fvdg_ctx* c) {
    if (c->order < 0 || !c->have_mesh) return FVDG_OK;
    const int N = c->tab.N;
    const size_t ndof = (size_t)c->Epad * 5 * N;
    HIP_TRY(c->U.alloc(ndof));
    HIP_TRY(c->Ua.alloc(ndof));
    HIP_TRY(c->Ub.alloc(ndof));
    HIP_TRY(c->R.alloc(ndof));
    for (DevBuf<double>* b : {&c->U, &c->Ua, &c->Ub, &c->R})
        HIP_TRY(hipMemsetAsync(b->p, 0, ndof * sizeof(double), c->stream));
    HIP_TRY(c->bad.alloc(1));
    HIP_TRY(hipMemsetAsync(c->bad.p, 0, sizeof(int32_t), c->stream));
    if (c->simple) HIP_TRY(c->Fq.alloc((size_t)c->E * c->tab.Q * 15));
    int rc = precision_buffers(c);
    if (rc) return rc;
    HIP_TRY(c->UD.alloc((size_t)std::max<int64_t>(c->n_dir, 1) * c->tab.Qf * 5));
    HIP_TRY(hipMemsetAsync(c->UD.p, 0, c->UD.n * sizeof(double), c->stream));
    c->n_red = std::max<int64_t>(nblocks(std::max(c->E, c->F * c->tab.Qf), 256), 1);
    HIP_TRY(c->red.alloc((size_t)c->n_red * 5));
    HIP_TRY(c->lim_ext.alloc((size_t)c->E * 10));
    HIP_TRY(c->diag.alloc((size_t)point_blocks(c) * fvdg_diag::kNumDiag));
    c->lim_buf.release();
    c->lim_bounds = false;
    for (DevBuf<double>* b : {&c->imp_A, &c->imp_rhs, &c->imp_b, &c->imp_Md, &c->imp_LU, &c->imp_V, &c->imp_w,
                              &c->imp_z, &c->imp_part, &c->imp_xold, &c->imp_dx})
        b->release();
    for (DevBuf<int32_t>* b : {&c->imp_col, &c->imp_piv, &c->imp_dslot, &c->imp_flag, &c->imp_self}) b->release();
    c->imp_have = false;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->have_state = false;
    return FVDG_OK;
}

// ---------------------------------------------------------------- launches
KernelArgs simple_args(fvdg_ctx* c, const double* U, double* R, int parts) {
    const SchemeTables& T = c->tab;
    KernelArgs a;
    a.E = c->E;
    a.F = c->F;
    a.N = T.N;
    a.Q = T.Q;
    a.Qf = T.Qf;
    a.gamma = c->gamma;
    a.flux = c->flux;
    a.parts = parts;
    a.U = U;
    a.R = R;
    a.cgeo = c->cgeo.p;
    a.cfaces = c->cfaces.p;
    a.fcell = c->fcell.p;
    a.fcombo = c->fcombo.p;
    a.fgeo = c->fgeo.p;
    a.fbc = c->fbc.p;
    a.fdir = c->fdir.p;
    a.UD = c->UD.p;
    a.V = c->tV.p;
    a.G = c->tG.p;
    a.wq = c->tWq.p;
    a.Phi = c->tPhi.p;
    a.wf = c->tWf.p;
    a.Fq = c->Fq.p;
    a.Fh = c->Fh.p;
    return a;
}

SurfArgs surf_args(fvdg_ctx* c, const double* U, double* R, int parts) {
    SurfArgs a;
    a.E = c->E;
    a.gamma = c->gamma;
    a.flux = c->flux;
    a.parts = parts;
    a.accumulate = (parts & FVDG_PART_CELLS) != 0;
    a.U = U;
    a.R = R;
    a.cgeo = c->cgeo.p;
    a.cfaces = c->cfaces.p;
    a.fcell = c->fcell.p;
    a.fcombo = c->fcombo.p;
    a.fgeo = c->fgeo.p;
    a.fbc = c->fbc.p;
    a.fdir = c->fdir.p;
    a.UD = c->UD.p;
    a.Phi = c->tPhi.p;
    a.PhiT = c->tPhiT.p;
    a.wf = c->tWf.p;
    a.U0 = nullptr;
    a.out = nullptr;
    a.mhat = c->tMhat.p;
    a.a0 = a.b0 = a.dt = 0.0;
    a.edt = c->edt.n ? c->edt.p : nullptr;
    a.bad = c->bad.p;
    return a;
}

// Stream-ordered event pair for per-kernel timing (no-op when ms == nullptr).
struct Span {
    hipEvent_t a = nullptr, b = nullptr;
    double* acc;
    hipStream_t s;
    Span(double* ms, hipStream_t st) : acc(ms), s(st) {
        if (acc) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
            (void)hipEventRecord(a, s);
        }
    }
    void stop() {
        if (!acc) return;
        (void)hipEventRecord(b, s);
        (void)hipEventSynchronize(b);
        float t = 0;
        (void)hipEventElapsedTime(&t, a, b);
        *acc += t;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        acc = nullptr;
    }
};

// SSP-RK stage combination applied by the last face kernel: out = a0*U0 + b0*(U + dt*(-R/M))
struct StageUpdate {
    const double* U0;
    double* out;
    double a0, b0, dt;
};

// the FP32-contraction kernels (the plain cross-check kernels are FP64 only)
inline bool mixed(const fvdg_ctx* c) { return c->precision != FVDG_PREC_FP64 && !c->simple; }

template <int P, typename T>
void launch_vol_t(fvdg_ctx* c, const double* U, double* R, const T* VA, const T* GA, double* ext) {
    VolArgs<T> v{c->E, c->gamma, U, R, c->cgeo.p, VA, GA, c->tWqa.p, c->tGs.p};
    if constexpr (std::is_same<T, double>::value) {
        if (ext) {
            v.ext = ext;
            hipLaunchKernelGGL((k_vol_mfma<P, T, true>), dim3(nblocks(c->E, 16)), dim3(64), 0, c->stream, v);
            return;
        }
    }
    hipLaunchKernelGGL((k_vol_mfma<P, T>), dim3(nblocks(c->E, 16)), dim3(64), 0, c->stream, v);
}
template <int P>
void launch_vol_fast(fvdg_ctx* c, const double* U, double* ext) {
    if (c->precision == FVDG_PREC_BF16X3 && !c->simple) {
        fvdg_bf3::Bf3Args b{c->E, c->gamma, U, c->R.p, c->cgeo.p, reinterpret_cast<const __bf16*>(c->tVB3.p),
                            reinterpret_cast<const __bf16*>(c->tGB3.p), c->tWqa.p, c->tGs.p};
        hipLaunchKernelGGL((fvdg_bf3::k_vol_bf16x3<P>), dim3(nblocks(c->E, 16)), dim3(64), 0, c->stream, b);
    } else if (mixed(c))
        launch_vol_t<P, float>(c, U, c->R.p, c->tVA32.p, c->tGA32.p, nullptr);
    else
        launch_vol_t<P, double>(c, U, c->R.p, c->tVA.p, c->tGA.p, ext);
}
// volume part of the residual of U into c->R; ext (FP64 path only, else ignored): also the
// limiter's per-element extremes of U
void dispatch_vol(fvdg_ctx* c, const double* U, double* ext = nullptr) {
    switch (c->order) {
        case 0: launch_vol_fast<0>(c, U, ext); break;
        case 1: launch_vol_fast<1>(c, U, ext); break;
        case 2: launch_vol_fast<2>(c, U, ext); break;
        case 3: launch_vol_fast<3>(c, U, ext); break;
        case 4: launch_vol_fast<4>(c, U, ext); break;
        case 5: launch_vol_fast<5>(c, U, ext); break;
    }
}

template <int P>
void launch_surf_fast(fvdg_ctx* c, const SurfArgs& a, bool fuse) {
    constexpr int TE = SurfLdsShape<P>::TE, NT = SurfLdsShape<P>::THREADS;
    SurfLdsArgs A{a, c->tPhiC.p, c->tCanon.p, c->tPerm.p, (int64_t)nblocks(c->E, TE)};
    // persistent grid: a few blocks per compute unit, each walking tiles
    const unsigned grid = (unsigned)std::min<int64_t>(A.ntiles, (int64_t)c->n_cu * std::max(1, c->grid_mult / 2));
    if (fuse)
        hipLaunchKernelGGL((k_surf_lds<P, true>), dim3(grid), dim3(NT), 0, c->stream, A);
    else
        hipLaunchKernelGGL((k_surf_lds<P, false>), dim3(grid), dim3(NT), 0, c->stream, A);
}

// the row-broadcast face-flux kernel in NT-thread blocks, a persistent grid of the resident blocks
template <int P, typename T, int NT>
void launch_flux_dpp(fvdg_ctx* c, FaceFluxArgs<T> f) {
    constexpr int FBD = FaceFluxDppShape<P, T, NT>::FB;
    f.ntiles = (c->F + FBD - 1) / FBD;
    const int64_t res = resident_blocks<k_face_flux_dpp<P, T, NT>, NT>();
    const unsigned gd = (unsigned)std::min<int64_t>(f.ntiles, (int64_t)c->n_cu * res);
    hipLaunchKernelGGL((k_face_flux_dpp<P, T, NT>), dim3(gd), dim3(NT), 0, c->stream, f);
}

// the face-flux pass: every face's traces and numerical flux into Fh (mixed: + Fref)
template <int P, typename T>
void flux_pass(fvdg_ctx* c, const double* U, const T* PhiC, T* Fh, double* ms_flux) {
    FaceFluxArgs<T> f{c->F, c->gamma, c->flux, U, c->fcell.p, c->fcombo.p, c->fgeo.p, c->fbc.p, c->fdir.p,
                      c->UD.p, c->tWf.p, PhiC, c->tCanon.p, c->tPerm.p, Fh, c->Fref.p, 0};
    constexpr int FB = FaceFluxShape<P, T>::FB;
    f.ntiles = (c->F + FB - 1) / FB;
    // FP64: exactly the resident blocks (measured best at P3: 0.351 ms vs 0.364 with 8 per CU);
    // FP32: grid_mult blocks per compute unit (more than resident measured faster)
    int64_t per_cu = c->grid_mult;
    if constexpr (std::is_same<T, double>::value)
        per_cu = resident_blocks<k_face_flux_lds<P, T>, FaceFluxShape<P, T>::THREADS>();
    Span s1(ms_flux, c->stream);
    if constexpr (P >= 2 && FaceFluxDppShape<P, T>::SUPPORTED) {
        if (!c->lds_face_flux) {
            // P2-P4: one face per 16-lane row, coefficients broadcast from registers
            if (c->flux_threads == 128)
                launch_flux_dpp<P, T, 128>(c, f);
            else if (c->flux_threads == 512)
                launch_flux_dpp<P, T, 512>(c, f);
            else
                launch_flux_dpp<P, T, 256>(c, f);
        }
    }
    if (P < 2 || !FaceFluxDppShape<P, T>::SUPPORTED || c->lds_face_flux) {
        const unsigned g1 = (unsigned)std::min<int64_t>(f.ntiles, (int64_t)c->n_cu * per_cu);
        hipLaunchKernelGGL((k_face_flux_lds<P, T>), dim3(g1), dim3(FaceFluxShape<P, T>::THREADS), 0, c->stream, f);
    }
    s1.stop();
}

template <int P, typename T>
void faces_two_pass(fvdg_ctx* c, const double* U, int parts, const double* Rv, double* Rout, const StageUpdate* up,
                    const T* PhiC, T* Fh, double* ms_flux, double* ms_test) {
    flux_pass<P, T>(c, U, PhiC, Fh, ms_flux);
    constexpr int TE = FaceTestShape<P>::TE;
    FaceTestArgs<T> t{};
    t.E = c->E;
    t.parts = parts;
    t.U = U;
    t.Rv = Rv;
    t.R = Rout;
    t.cgeo = c->cgeo.p;
    t.cfaces = c->cfaces.p;
    t.cmeta = c->cmeta.p;
    if constexpr (std::is_same<T, double>::value)
        t.PhiB = c->tPhiB.p;
    else
        t.PhiB = c->tPhiB32.p;
    t.fcell = c->fcell.p;
    t.fcombo = c->fcombo.p;
    t.mhat = c->tMhat.p;
    t.edt = c->edt.n ? c->edt.p : nullptr;
    t.bad = c->bad.p;
    if (up) {
        t.U0 = up->U0;
        t.out = up->out;
        t.a0 = up->a0;
        t.b0 = up->b0;
        t.dt = up->dt;
    }
    t.Fh = Fh;
    t.Fref = c->Fref.p;
    t.IC = c->tIC.p;
    t.PhiC = PhiC;
    t.canon = c->tCanon.p;
    t.perm = c->tPerm.p;
    if constexpr (FaceTestMfmaShape<P, T>::SUPPORTED) {
        if (!c->lds_face_test) {
            using MS = FaceTestMfmaShape<P, T>;
            t.ntiles = (c->E + MS::TE - 1) / MS::TE;
            const auto kern = up ? k_face_test_mfma<P, true, T> : k_face_test_mfma<P, false, T>;
            const int per_cu = up ? resident_blocks<k_face_test_mfma<P, true, T>, MS::THREADS>()
                                  : resident_blocks<k_face_test_mfma<P, false, T>, MS::THREADS>();
            const unsigned g2 = (unsigned)std::min<int64_t>((t.ntiles + MS::WAVES - 1) / MS::WAVES,
                                                            (int64_t)c->n_cu * per_cu);
            Span s2(ms_test, c->stream);
            hipLaunchKernelGGL(kern, dim3(g2), dim3(MS::THREADS), 0, c->stream, t);
            s2.stop();
            return;
        }
    }
    t.ntiles = (c->E + TE - 1) / TE;
    const unsigned g2 = (unsigned)std::min<int64_t>(t.ntiles, (int64_t)c->n_cu * c->grid_mult);
    Span s2(ms_test, c->stream);
    if (up)
        hipLaunchKernelGGL((k_face_test_lds<P, true, T>), dim3(g2), dim3(FaceTestShape<P>::THREADS), 0, c->stream, t);
    else
        hipLaunchKernelGGL((k_face_test_lds<P, false, T>), dim3(g2), dim3(FaceTestShape<P>::THREADS), 0, c->stream, t);
    s2.stop();
}

template <int P>
void faces_fast(fvdg_ctx* c, const double* U, int parts, bool have_vol, const StageUpdate* up, double* ms_flux,
                double* ms_test) {
    if (mixed(c)) {
        faces_two_pass<P, float>(c, U, parts, have_vol ? c->R.p : nullptr, c->R.p, up, c->tPhiC32.p, c->Fh32.p,
                                 ms_flux, ms_test);
    } else if (!c->owner_faces) {
        faces_two_pass<P, double>(c, U, parts, have_vol ? c->R.p : nullptr, c->R.p, up, c->tPhiC.p, c->Fh.p, ms_flux,
                                  ms_test);
    } else {
        SurfArgs a = surf_args(c, U, c->R.p, parts);
        a.accumulate = have_vol;
        if (up) {
            a.U0 = up->U0;
            a.out = up->out;
            a.a0 = up->a0;
            a.b0 = up->b0;
            a.dt = up->dt;
        }
        a.edt = c->edt.n ? c->edt.p : nullptr;
        Span so(ms_flux, c->stream);
        launch_surf_fast<P>(c, a, up != nullptr);
        so.stop();
    }
}

// The fused element path: the face-flux pass, then k_elem_mfma (volume + face test + stage
// update, or the residual of the selected parts into c->R).  ms_flux / ms_elem: event times.
// the face-tile pass: every face's traces, Riemann flux and test contributions of both sides
template <int P, typename T>
void face_tile_pass(fvdg_ctx* c, const double* U, const T* At, const T* Ae, T* Cf, double* ms_flux) {
    if constexpr (TileShape<P>::SUPPORTED) {
        const T* ICc = nullptr;  // FP32 only (the FP64 tiles take no reference flux)
        if constexpr (!std::is_same<T, double>::value) ICc = c->tICc32.p;
        FaceTileArgs<T> f{c->n_face_tiles, c->gamma, c->flux, U, reinterpret_cast<const int4*>(c->trec.p),
                          c->tclass.p, c->fgeo.p, c->UD.p, c->tWf.p, At, Ae, ICc, Cf, c->Fref.p};
        Span s1(ms_flux, c->stream);
        if constexpr (!std::is_same<T, double>::value) {
            if (bf3_face_tiles(c)) {
                f.At3 = reinterpret_cast<const __bf16*>(c->tAt3.p);
                f.Ae3 = reinterpret_cast<const __bf16*>(c->tAe3.p);
                hipLaunchKernelGGL((k_face_tile_bf3<P>), dim3((unsigned)c->n_face_tiles), dim3(64), 0, c->stream, f);
                s1.stop();
                return;
            }
        }
        hipLaunchKernelGGL((k_face_tile<P, T>), dim3((unsigned)c->n_face_tiles), dim3(64), 0, c->stream, f);
        s1.stop();
    }
}

template <int P, typename T>
void elem_path_t(fvdg_ctx* c, const double* U, int parts, const StageUpdate* up, const T* PhiC, T* Fh,
                 const T* VA, const T* GA, const T* PhiB, double* ms_flux, double* ms_elem) {
    if constexpr (ElemShape<P, T>::SUPPORTED) {
        // FP32 face tiles take every difference from the cell's own mean flux and leave out the
        // terms that cancel between the volume and the faces: the whole residual only (single
        // parts take the per-face reference path)
        const bool tiles = TileShape<P>::SUPPORTED && use_face_tile(c) &&
                           (std::is_same<T, double>::value || parts == FVDG_PART_ALL);
        T* Cf = nullptr;
        if (tiles) {
            const T* At;
            const T* Ae;
            if constexpr (std::is_same<T, double>::value) {
                At = c->tAt.p, Ae = c->tAe.p, Cf = c->Cf.p;
            } else {
                At = c->tAt32.p, Ae = c->tAe32.p, Cf = c->Cf32.p;
            }
            if (parts & (FVDG_PART_INTERNALS | FVDG_PART_BOUNDARYS)) face_tile_pass<P, T>(c, U, At, Ae, Cf, ms_flux);
        } else if (parts & (FVDG_PART_INTERNALS | FVDG_PART_BOUNDARYS)) {
            flux_pass<P, T>(c, U, PhiC, Fh, ms_flux);
        }
        ElemArgs<T> a{};
        a.Cf = Cf;
        a.phi0 = c->tab.V[0];  // phi_0 is constant: its value at any point
        a.E = c->E;
        a.gamma = c->gamma;
        a.parts = parts;
        a.U = U;
        a.R = c->R.p;
        a.cgeo = c->cgeo.p;
        a.cfaces = c->cfaces.p;
        a.cmeta = c->cmeta.p;
        a.VA = VA;
        a.GA = GA;
        a.wq = c->tWqa.p;
        a.Gs = c->tGs.p;
        a.Fh = Fh;
        a.Fref = c->Fref.p;
        a.IC = c->tIC.p;
        a.PhiB = PhiB;
        a.canon = c->tCanon.p;
        a.perm = c->tPerm.p;
        a.mhat = c->tMhat.p;
        a.edt = c->edt.n ? c->edt.p : nullptr;
        a.bad = c->bad.p;
        if (up) {
            a.U0 = up->U0;
            a.out = up->out;
            a.a0 = up->a0;
            a.b0 = up->b0;
            a.dt = up->dt;
        }
        Span se(ms_elem, c->stream);
        const dim3 ge(nblocks(c->E, 16));
        if (tiles && c->precision == FVDG_PREC_BF16X3) {
            if constexpr (TileShape<P>::SUPPORTED && !std::is_same<T, double>::value) {
                a.VB3 = reinterpret_cast<const __bf16*>(c->tVB3.p);
                a.GB3 = reinterpret_cast<const __bf16*>(c->tGB3.p);
                if (up)
                    hipLaunchKernelGGL((k_elem_mfma<P, T, true, true, true>), ge, dim3(64), 0, c->stream, a);
                else
                    hipLaunchKernelGGL((k_elem_mfma<P, T, false, true, true>), ge, dim3(64), 0, c->stream, a);
            }
        } else if (tiles) {
            if (up)
                hipLaunchKernelGGL((k_elem_mfma<P, T, true, true>), ge, dim3(64), 0, c->stream, a);
            else
                hipLaunchKernelGGL((k_elem_mfma<P, T, false, true>), ge, dim3(64), 0, c->stream, a);
        } else if (up) {
            hipLaunchKernelGGL((k_elem_mfma<P, T, true>), ge, dim3(64), 0, c->stream, a);
        } else {
            hipLaunchKernelGGL((k_elem_mfma<P, T, false>), ge, dim3(64), 0, c->stream, a);
        }
        se.stop();
    }
}
template <int P>
void elem_path_p(fvdg_ctx* c, const double* U, int parts, const StageUpdate* up, double* ms_flux, double* ms_elem) {
    if (mixed(c))
        elem_path_t<P, float>(c, U, parts, up, c->tPhiC32.p, c->Fh32.p, c->tVA32.p, c->tGA32.p, c->tPhiB32.p,
                              ms_flux, ms_elem);
    else
        elem_path_t<P, double>(c, U, parts, up, c->tPhiC.p, c->Fh.p, c->tVA.p, c->tGA.p, c->tPhiB.p, ms_flux,
                               ms_elem);
}
// the fused element path serves P0-P4 in FP64 and FP32, and BF16x3 where the face tiles serve it
// (P2-P4, the whole residual); P5, BF16x3 otherwise, the alternative face kernels, the plain
// kernels and FVDG_ELEM=split take the volume kernel + face-test kernel path
bool use_elem(const fvdg_ctx* c, int parts) {
    return c->elem_fused && !c->simple && !c->owner_faces && !c->lds_face_test && c->order <= 4 &&
           (c->precision != FVDG_PREC_BF16X3 || (parts == FVDG_PART_ALL && use_face_tile(c)));
}
void elem_path(fvdg_ctx* c, const double* U, int parts, const StageUpdate* up, double* ms_flux, double* ms_elem) {
    switch (c->order) {
        case 0: elem_path_p<0>(c, U, parts, up, ms_flux, ms_elem); break;
        case 1: elem_path_p<1>(c, U, parts, up, ms_flux, ms_elem); break;
        case 2: elem_path_p<2>(c, U, parts, up, ms_flux, ms_elem); break;
        case 3: elem_path_p<3>(c, U, parts, up, ms_flux, ms_elem); break;
        case 4: elem_path_p<4>(c, U, parts, up, ms_flux, ms_elem); break;
    }
}

// Face part of the residual of U: added to the volume part (have_vol) and written to
// c->R, or, with `up`, fed to the SSP-RK stage update.  ms_flux / ms_test (nullable):
// event-timed face-flux and face-test(+update) kernels; the one-pass owner-computes
// kernel is accounted under ms_flux.
void dispatch_faces(fvdg_ctx* c, const double* U, int parts, bool have_vol, const StageUpdate* up,
                    double* ms_flux = nullptr, double* ms_test = nullptr) {
    switch (c->order) {
        case 0: faces_fast<0>(c, U, parts, have_vol, up, ms_flux, ms_test); break;
        case 1: faces_fast<1>(c, U, parts, have_vol, up, ms_flux, ms_test); break;
        case 2: faces_fast<2>(c, U, parts, have_vol, up, ms_flux, ms_test); break;
        case 3: faces_fast<3>(c, U, parts, have_vol, up, ms_flux, ms_test); break;
        case 4: faces_fast<4>(c, U, parts, have_vol, up, ms_flux, ms_test); break;
        case 5: faces_fast<5>(c, U, parts, have_vol, up, ms_flux, ms_test); break;
    }
}

// R = selected parts of the residual of U (device buffers; R is c->R)
int launch_residual(fvdg_ctx* c, const double* U, int parts, double* ms_vol, double* ms_surf) {
    const SchemeTables& T = c->tab;
    const int bs = 256;
    if (c->simple) {
        KernelArgs a = simple_args(c, U, c->R.p, parts);
        Span sv(ms_vol, c->stream);
        if (parts & FVDG_PART_CELLS) {
            hipLaunchKernelGGL(k_vol_flux, dim3(nblocks(c->E * T.Q, bs)), dim3(bs), 0, c->stream, a);
            hipLaunchKernelGGL(k_vol_test, dim3(nblocks(c->E * T.N, bs)), dim3(bs), 0, c->stream, a);
        }
        sv.stop();
        Span ss(ms_surf, c->stream);
        if (parts & (FVDG_PART_INTERNALS | FVDG_PART_BOUNDARYS))
            hipLaunchKernelGGL(k_face_flux, dim3(nblocks(c->F * T.Qf, bs)), dim3(bs), 0, c->stream, a);
        hipLaunchKernelGGL(k_face_test, dim3(nblocks(c->E * T.N, bs)), dim3(bs), 0, c->stream, a);
        ss.stop();
    } else if (use_elem(c, parts)) {
        Span ss(ms_surf, c->stream);
        elem_path(c, U, parts, nullptr, nullptr, nullptr);
        ss.stop();
    } else {
        const bool vol = (parts & FVDG_PART_CELLS) != 0;
        Span sv(ms_vol, c->stream);
        if (vol) dispatch_vol(c, U);
        sv.stop();
        Span ss(ms_surf, c->stream);
        dispatch_faces(c, U, parts, vol, nullptr);
        ss.stop();
    }
    HIP_TRY(hipGetLastError());
    return FVDG_OK;
}

// ---------------------------------------------------------------- limiter

// lim_ext <- per-element extremes of Uprev (constructMinMax)
template <int P>
void launch_limit_extrema_t(fvdg_ctx* c, const double* Uprev, int only_avg) {
    const dim3 grid(nblocks(c->E, fvdg_lim::LimShape<P>::ET));
    hipLaunchKernelGGL(fvdg_lim::k_limit_extrema<P>, grid, dim3(256), 0, c->stream, Uprev, c->tV.p, c->E, only_avg,
                       c->lim_ext.p);
}
void launch_limit_extrema(fvdg_ctx* c, const double* Uprev, int only_avg) {
    switch (c->order) {
        case 0: launch_limit_extrema_t<0>(c, Uprev, only_avg); break;
        case 1: launch_limit_extrema_t<1>(c, Uprev, only_avg); break;
        case 2: launch_limit_extrema_t<2>(c, Uprev, only_avg); break;
        case 3: launch_limit_extrema_t<3>(c, Uprev, only_avg); break;
        case 4: launch_limit_extrema_t<4>(c, Uprev, only_avg); break;
        default: launch_limit_extrema_t<5>(c, Uprev, only_avg); break;
    }
}

// U <- limited U (apply_1 when parts & 1 with the bounds of lim_ext, apply_2 when parts & 2)
template <int P>
void launch_limit_apply_t(fvdg_ctx* c, double* U, int parts) {
    const dim3 grid(nblocks(c->E, fvdg_lim::LimShape<P>::ET));
    hipLaunchKernelGGL(fvdg_lim::k_limit_apply<P>, grid, dim3(256), 0, c->stream, U, c->lim_ext.p, c->cnb.p, c->tV.p,
                       c->E, c->gamma, parts);
}
void launch_limit_apply(fvdg_ctx* c, double* U, int parts) {
    switch (c->order) {
        case 0: launch_limit_apply_t<0>(c, U, parts); break;
        case 1: launch_limit_apply_t<1>(c, U, parts); break;
        case 2: launch_limit_apply_t<2>(c, U, parts); break;
        case 3: launch_limit_apply_t<3>(c, U, parts); break;
        case 4: launch_limit_apply_t<4>(c, U, parts); break;
        default: launch_limit_apply_t<5>(c, U, parts); break;
    }
}

// One SSP-RK3 step on c->U (Shu & Osher 1988):
//   U1 = U + dt L(U);  U2 = 3/4 U + 1/4 (U1 + dt L(U1));  U <- 1/3 U + 2/3 (U2 + dt L(U2)),
// L = -M^{-1} R.  Stage inputs and outputs ping-pong between Ua and Ub so that no
// kernel writes a buffer whose neighbour values other blocks may still read; the
// last stage writes U in place (each thread reads U0[i] before writing out[i]).
// ms (nullable): [1] volume, [2] face flux, [3] face test + stage update, accumulated.
int ssprk3_stage(fvdg_ctx* c, double dt, int s, double* ms) {
    static const double A[3] = {0.0, 0.75, 1.0 / 3.0};
    static const double B[3] = {1.0, 0.25, 2.0 / 3.0};
    const double* src[3] = {c->U.p, c->Ua.p, c->Ub.p};
    double* dst[3] = {c->Ua.p, c->Ub.p, c->U.p};
    // point bounds of the stage input come out of the FP64 volume kernel itself
    const bool fused_extrema = (c->lim_parts & FVDG_LIMIT_BOUNDS) && !c->lim_avg && !c->simple && !mixed(c);
    {
        if (c->simple) {
            int rc = launch_residual(c, src[s], FVDG_PART_ALL, ms ? &ms[1] : nullptr, ms ? &ms[2] : nullptr);
            if (rc) return rc;
            Span su(ms ? &ms[3] : nullptr, c->stream);
            const int64_t n = c->E * 5 * c->tab.N;
            hipLaunchKernelGGL(k_rk_update, dim3(nblocks(n, 256)), dim3(256), 0, c->stream, c->U.p, src[s], c->R.p,
                               dst[s], c->cgeo.p, c->tMhat.p, c->E, c->tab.N, A[s], B[s], dt,
                               c->edt.n ? c->edt.p : nullptr, c->bad.p);
            su.stop();
        } else if (use_elem(c, FVDG_PART_ALL) && !fused_extrema) {
            const StageUpdate up{c->U.p, dst[s], A[s], B[s], dt};
            elem_path(c, src[s], FVDG_PART_ALL, &up, ms ? &ms[2] : nullptr, ms ? &ms[3] : nullptr);
        } else {
            Span sv(ms ? &ms[1] : nullptr, c->stream);
            dispatch_vol(c, src[s], fused_extrema ? c->lim_ext.p : nullptr);
            sv.stop();
            const StageUpdate up{c->U.p, dst[s], A[s], B[s], dt};
            dispatch_faces(c, src[s], FVDG_PART_ALL, true, &up, ms ? &ms[2] : nullptr, ms ? &ms[3] : nullptr);
        }
        if (c->lim_parts) {  // limit the stage result, bounds from the stage input
            if ((c->lim_parts & FVDG_LIMIT_BOUNDS) && !fused_extrema) launch_limit_extrema(c, src[s], c->lim_avg);
            launch_limit_apply(c, dst[s], c->lim_parts);
            c->lim_bounds = (c->lim_parts & FVDG_LIMIT_BOUNDS) != 0;
        }
        HIP_TRY(hipGetLastError());
    }
    return FVDG_OK;
}


// Names of the kernels one SSP-RK3 stage launches on the context's current scheme, as
// "volume=<name>;flux=<name>;test=<name>" (an empty name: that slot launches nothing)
std::string stage_kernel_names(const fvdg_ctx* c) {
    const int P = c->order;
    const bool mx = mixed(c);
    const std::string T = mx ? "float" : "double", ps = std::to_string(P);
    if (c->simple) return "volume=k_vol_flux+k_vol_test;flux=k_face_flux;test=k_face_test+k_rk_update";
    const bool fused_extrema = (c->lim_parts & FVDG_LIMIT_BOUNDS) && !c->lim_avg && !mx;
    const bool dpp = P >= 2 && P <= 4 && !c->lds_face_flux;
    const std::string flux = dpp ? "k_face_flux_dpp<" + ps + "," + T + "," + std::to_string(c->flux_threads) + ">"
                                 : "k_face_flux_lds<" + ps + "," + T + ">";
    if (use_elem(c, FVDG_PART_ALL) && !fused_extrema && use_face_tile(c))
        return "volume=;flux=" + (bf3_face_tiles(c) ? "k_face_tile_bf3<" + ps + ">" : "k_face_tile<" + ps + "," + T + ">") +
               ";test=k_elem_mfma<" + ps + "," + T + ",true,true" + (c->precision == FVDG_PREC_BF16X3 ? ",true>" : ">");
    if (use_elem(c, FVDG_PART_ALL) && !fused_extrema) return "volume=;flux=" + flux + ";test=k_elem_mfma<" + ps + "," + T + ",true>";
    std::string vol = c->precision == FVDG_PREC_BF16X3 ? "k_vol_bf16x3<" + ps + ">"
                                                       : "k_vol_mfma<" + ps + "," + T + (fused_extrema ? ",true>" : ">");
    if (c->owner_faces && !mx) return "volume=" + vol + ";flux=k_surf_lds<" + ps + ",true>;test=";
    const bool mfma_test = !c->lds_face_test && P <= (mx ? 3 : 4);
    const std::string test = (mfma_test ? "k_face_test_mfma<" : "k_face_test_lds<") + ps + ",true," + T + ">";
    return "volume=" + vol + ";flux=" + flux + ";test=" + test;
}

// Waits for the stream, then reports the stage-output flag the update kernels set
// (FVDG_ERR_NONFINITE; the state is left as computed).
int finish_checked(fvdg_ctx* c) {
    int32_t flag = 0;
    HIP_TRY(hipMemcpyAsync(&flag, c->bad.p, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (flag)
        return fail(FVDG_ERR_NONFINITE,
                    "an SSP-RK3 stage produced a non-finite coefficient or a non-positive mean density");
    return FVDG_OK;
}

}  // namespace

int fvdg::ssprk3_once(fvdg_ctx* c, double dt, double* ms) {
    for (int s = 0; s < 3; ++s) {
        const int rc = ssprk3_stage(c, dt, s, ms);
        if (rc) return rc;
    }
    return FVDG_OK;
}

void fvdg::limit_state(fvdg_ctx* c, const double* prev, double* cur, int parts, int neighbour_averages) {
    if (parts & FVDG_LIMIT_BOUNDS) launch_limit_extrema(c, prev, neighbour_averages);
    launch_limit_apply(c, cur, parts);
}

int fvdg::face_wave_speeds(fvdg_ctx* c, std::vector<double>& out) {
    int rc = check_ready(c, true);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    KernelArgs a = simple_args(c, c->U.p, c->R.p, FVDG_PART_ALL);
    DevBuf<double> d;
    HIP_TRY(d.alloc(c->F));
    hipLaunchKernelGGL(k_face_wave_speed, dim3(nblocks(c->F, 256)), dim3(256), 0, c->stream, a, d.p);
    HIP_TRY(hipGetLastError());
    out.resize(c->F);
    HIP_TRY(hipMemcpyAsync(out.data(), d.p, c->F * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FVDG_OK;
}

// ================================================================ C ABI (device part)
extern "C" {

int fvdg_device_count(int* n) {
    if (!n) return fail(FVDG_ERR_ARG, "null output");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) count = 0;  // no driver or no device: none
    *n = count;
    return FVDG_OK;
}

int fvdg_create(fvdg_ctx** out, int device) {
    if (!out) return fail(FVDG_ERR_ARG, "null output pointer");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(FVDG_ERR_ARG, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    std::unique_ptr<fvdg_ctx> c(new fvdg_ctx);
    c->device = device;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    c->n_cu = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    const char* kern = std::getenv("FVDG_KERNELS");
    c->simple = kern && std::strcmp(kern, "simple") == 0;
    const char* faces = std::getenv("FVDG_FACES");
    c->owner_faces = faces && std::strcmp(faces, "owner") == 0;
    const char* ftest = std::getenv("FVDG_FACE_TEST");
    c->lds_face_test = ftest && std::strcmp(ftest, "lds") == 0;
    const char* elem = std::getenv("FVDG_ELEM");
    c->elem_fused = !(elem && std::strcmp(elem, "split") == 0);
    const char* fflux = std::getenv("FVDG_FACE_FLUX");
    c->lds_face_flux = fflux && std::strcmp(fflux, "lds") == 0;
    if (const char* gm = std::getenv("FVDG_GRID_MULT")) c->grid_mult = std::max(1, std::atoi(gm));
    if (const char* to = std::getenv("FVDG_TILE_ORDER"))
        c->tile_order = std::strcmp(to, "class") == 0 ? 0 : std::strcmp(to, "space") == 0 ? 1 : 2;
    if (const char* bf = std::getenv("FVDG_BF16X3_FACES")) c->bf3_faces = std::atoi(bf) != 0;
    if (const char* ft = std::getenv("FVDG_FACE_TILE"))
        c->face_tile = std::strcmp(ft, "off") == 0 ? 0 : std::strcmp(ft, "all") == 0 ? 2 : 1;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    const int rc = scratch_init(device);  // the static scratch's address on this device
    if (rc) return rc;
    *out = c.release();
    return FVDG_OK;
}

void fvdg_destroy(fvdg_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
}

int fvdg_set_scheme(fvdg_ctx* c, int order, int flux, double gamma, int precision) {
    if (!c) return fail(FVDG_ERR_ARG, "null context");
    if (order < 0 || order > kMaxOrder) return fail(FVDG_ERR_UNSUPPORTED, "order must be 0..5");
    if (flux != FVDG_FLUX_LF && flux != FVDG_FLUX_ROE) return fail(FVDG_ERR_ARG, "unknown flux");
    if (!(gamma > 1.0)) return fail(FVDG_ERR_ARG, "gamma must exceed 1");
    if (precision != FVDG_PREC_FP64 && precision != FVDG_PREC_FP32 && precision != FVDG_PREC_BF16X3)
        return fail(FVDG_ERR_UNSUPPORTED, "precision must be FVDG_PREC_FP64, FVDG_PREC_FP32 or FVDG_PREC_BF16X3");
    HIP_TRY(hipSetDevice(c->device));
    const bool reorder = c->order != order;
    c->flux = flux;
    c->gamma = gamma;
    c->precision = precision;
    if (reorder) {
        c->order = order;
        c->tab = build_tables(order);
        int rc = upload_tables(c);
        if (rc) return rc;
        rc = alloc_state(c);
        if (rc) return rc;
    } else {
        const int rc = precision_buffers(c);  // keeps an uploaded state
        if (rc) return rc;
    }
    return FVDG_OK;
}

int fvdg_set_mesh(fvdg_ctx* c, const fvdg_mesh* m) {
    if (!c || !m) return fail(FVDG_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(c->device));
    const ComputingMesh& cm = m->cm;
    const int64_t E = cm.m_cells.size(), F = cm.m_faces.size();
    if (E == 0) return fail(FVDG_ERR_MESH, "empty mesh");
    if (E > (int64_t(1) << 30) || F > (int64_t(1) << 30)) return fail(FVDG_ERR_UNSUPPORTED, "mesh too large for 32-bit indices");
    std::vector<double> cgeo(E * 10);
    std::vector<int32_t> cfaces(E * 4);
    std::vector<uint8_t> cmeta(E * 4);
    std::vector<int32_t> cnb(E * 4);  // neighbour cell across each local face, -1 on the boundary
    std::vector<double> det(E);
    double h_min = INFINITY;
    for (int64_t e = 0; e < E; ++e) {
        const auto& t = cm.m_cells[e];
        h_min = std::min(h_min, 3.0 * t.m_h);
        for (int i = 0; i < 9; ++i) cgeo[e * 10 + i] = t.m_invJac[i];
        cgeo[e * 10 + 9] = t.compute_jacobian_det();
        det[e] = cgeo[e * 10 + 9];
    }
    std::vector<int32_t> fcell(F * 2);
    std::vector<uint8_t> fcombo(F * 2), fbc(F);
    std::vector<double> fgeo(F * 4);
    std::vector<int32_t> fdir(F, -1);
    std::vector<int64_t> dir_faces;
    for (int64_t f = 0; f < F; ++f) {
        const auto& Fc = cm.m_faces[f];
        const bool bnd = Fc.is_boundary();
        fcell[2 * f] = (int32_t)Fc.m_neighbor_cells[0];
        fcell[2 * f + 1] = bnd ? -1 : (int32_t)Fc.m_neighbor_cells[1];
        for (int s = 0; s < 2; ++s) {
            const int cb = (s == 1 && bnd) ? 0 : face_combo(Fc, s);
            if (cb < 0) return fail(FVDG_ERR_MESH, "face natural coordinates are not a vertex permutation");
            fcombo[2 * f + s] = (uint8_t)cb;
        }
        for (int d = 0; d < 3; ++d) fgeo[4 * f + d] = Fc.m_normal[d];
        fgeo[4 * f + 3] = Fc.compute_jacobian_det();
        const BoundaryType bt = f < (int64_t)cm.m_boundaryTypes.size() ? cm.m_boundaryTypes[f] : BoundaryType::Dirichlet;
        fbc[f] = bnd ? (uint8_t)bt : (uint8_t)255;
        // faces that take boundary data: Dirichlet (both paths), WallTD / WallTN (implicit path)
        if (bnd && (bt == BoundaryType::Dirichlet || bt == BoundaryType::WallTD || bt == BoundaryType::WallTN)) {
            fdir[f] = (int32_t)dir_faces.size();
            dir_faces.push_back(f);
        }
    }
    for (int64_t e = 0; e < E; ++e) {
        const auto& t = cm.m_cells[e];
        for (int l = 0; l < 4; ++l) {
            const uInt f = t.m_faces[l];
            const auto& Fc = cm.m_faces[f];
            int side = (Fc.m_neighbor_cells[0] == (uInt)e) ? 0 : 1;
            // a cell that is its own periodic neighbour: pick the side whose nodes are this local face
            if (!Fc.is_boundary() && Fc.m_neighbor_cells[0] == Fc.m_neighbor_cells[1]) {
                const vector3u& n0 = Fc.m_nodes;
                bool has_opp = (n0[0] == t.m_nodes[l] || n0[1] == t.m_nodes[l] || n0[2] == t.m_nodes[l]);
                side = has_opp ? 1 : 0;
            }
            cfaces[e * 4 + l] = (int32_t)(2 * f + side);
            cnb[e * 4 + l] = Fc.is_boundary() ? -1 : (int32_t)Fc.m_neighbor_cells[1 - side];
            cmeta[e * 4 + l] = (uint8_t)(fcombo[2 * f + side] | (Fc.is_boundary() ? 0x80 : 0));
        }
    }
    // face tiles of k_face_tile: faces grouped by class = side-0 combination * 25 + (side-1
    // combination, or 24 on the boundary), each class in face order, padded to whole tiles of 16
    std::vector<int32_t> tface, tclass, trec;
    {
        constexpr int NCL = 24 * 25;
        std::vector<int32_t> fcl(F);
        std::vector<int64_t> cnt(NCL, 0), start(NCL, 0);
        for (int64_t f = 0; f < F; ++f) {
            fcl[f] = fcombo[2 * f] * 25 + (fcell[2 * f + 1] >= 0 ? fcombo[2 * f + 1] : 24);
            ++cnt[fcl[f]];
        }
        int64_t nt = 0;
        for (int k = 0; k < NCL; ++k) {
            start[k] = nt * 16;
            nt += (cnt[k] + 15) / 16;
        }
        tface.assign(nt * 16, -1);
        tclass.assign(nt, 0);
        for (int k = 0; k < NCL; ++k)
            for (int64_t t = start[k] / 16; t < start[k] / 16 + (cnt[k] + 15) / 16; ++t) tclass[t] = k;
        for (int64_t f = 0; f < F; ++f) tface[start[fcl[f]]++] = (int32_t)f;
        if (c->tile_order > 0) {
            // launch order: tiles by their smallest side-0 cell, so that tiles of different classes
            // over the same cells run together and share the cells' coefficient blocks in L2;
            // tile_order 2 also splits that order into 8 contiguous ranges, one per XCD (block b
            // runs on XCD b % 8)
            std::vector<int64_t> key(nt, INT64_MAX), ord(nt);
            for (int64_t t = 0; t < nt; ++t) {
                ord[t] = t;
                for (int i = 0; i < 16; ++i)
                    if (tface[16 * t + i] >= 0) key[t] = std::min<int64_t>(key[t], fcell[2 * tface[16 * t + i]]);
            }
            std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return key[x] < key[y]; });
            const int64_t ntp = c->tile_order == 2 ? (nt + 7) / 8 * 8 : nt, chunk = ntp / 8;
            std::vector<int32_t> tf2(ntp * 16, -1), tc2(ntp, 0);
            for (int64_t p = 0; p < nt; ++p) {
                const int64_t b = c->tile_order == 2 ? (p % chunk) * 8 + p / chunk : p;
                tc2[b] = tclass[ord[p]];
                for (int i = 0; i < 16; ++i) tf2[16 * b + i] = tface[16 * ord[p] + i];
            }
            tface.swap(tf2);
            tclass.swap(tc2);
            nt = ntp;
        }
        c->n_face_tiles = nt;
        // column records: face, side-0 cell, side-1 cell, kind | pseudo3dz << 2 | (Dirichlet slot + 1) << 3
        trec.assign(nt * 16 * 4, -1);
        for (int64_t i = 0; i < nt * 16; ++i) {
            const int32_t f = tface[i];
            int32_t* r = &trec[4 * i];
            r[3] = 0;
            if (f < 0) continue;
            r[0] = f;
            r[1] = fcell[2 * f];
            r[2] = fcell[2 * f + 1];
            if (r[2] >= 0) {
                r[3] = 1;
            } else if (fbc[f] == FVDG_BC_DIRICHLET) {
                r[3] = 2 | ((fdir[f] + 1) << 3);
            } else if (fbc[f] == FVDG_BC_PSEUDO3DZ) {
                r[3] = 2 | 4;
            }
        }
    }
    c->E = E;
    c->Epad = (E + 63) / 64 * 64;
    c->F = F;
    c->n_dir = dir_faces.size();
    c->dir_faces = dir_faces;
    c->dir_xyz.clear();
    c->cell_det_host = det;
    c->h_min = h_min;
    HIP_TRY(c->cgeo.upload(cgeo, c->stream));
    HIP_TRY(c->cfaces.upload(cfaces, c->stream));
    HIP_TRY(c->cmeta.upload(cmeta, c->stream));
    HIP_TRY(c->cnb.upload(cnb, c->stream));
    HIP_TRY(c->fcell.upload(fcell, c->stream));
    HIP_TRY(c->fcombo.upload(fcombo, c->stream));
    HIP_TRY(c->fgeo.upload(fgeo, c->stream));
    HIP_TRY(c->fbc.upload(fbc, c->stream));
    HIP_TRY(c->fdir.upload(fdir, c->stream));
    HIP_TRY(c->trec.upload(trec, c->stream));
    HIP_TRY(c->tclass.upload(tclass, c->stream));
    // keep the Dirichlet faces' corner coordinates to place their quadrature points
    c->dir_xyz.resize(c->n_dir * 9);
    for (int64_t i = 0; i < c->n_dir; ++i)
        for (int k = 0; k < 3; ++k)
            for (int d = 0; d < 3; ++d) c->dir_xyz[i * 9 + 3 * k + d] = cm.m_faces[dir_faces[i]].m_coords[k][d];
    c->have_mesh = true;
    c->have_state = false;
    return alloc_state(c);
}

int fvdg_state_upload(fvdg_ctx* c, const double* Uh) {
    int rc = check_ready(c, false);
    if (rc) return rc;
    if (!Uh) return fail(FVDG_ERR_ARG, "null state");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(c->U.p, Uh, (size_t)c->E * 5 * c->tab.N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->have_state = true;
    return FVDG_OK;
}

int fvdg_state_download(fvdg_ctx* c, double* Uh) {
    int rc = check_ready(c, true);
    if (rc) return rc;
    if (!Uh) return fail(FVDG_ERR_ARG, "null state");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(Uh, c->U.p, (size_t)c->E * 5 * c->tab.N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FVDG_OK;
}

int fvdg_dirichlet_points(fvdg_ctx* c, int64_t* n_faces, int64_t* face_ids, double* xyz) {
    int rc = check_ready(c, false);
    if (rc) return rc;
    if (n_faces) *n_faces = c->n_dir;
    const int Qf = c->tab.Qf;
    for (int64_t i = 0; i < c->n_dir; ++i) {
        if (face_ids) face_ids[i] = c->dir_faces[i];
        if (xyz)
            for (int g = 0; g < Qf; ++g) {
                const double u = c->tab.fq_pts[2 * g], v = c->tab.fq_pts[2 * g + 1];
                for (int d = 0; d < 3; ++d) {
                    const double* p = &c->dir_xyz[i * 9];
                    xyz[(i * Qf + g) * 3 + d] = p[d] * (1 - u - v) + p[3 + d] * u + p[6 + d] * v;
                }
            }
    }
    return FVDG_OK;
}

int fvdg_set_dirichlet_states(fvdg_ctx* c, const double* UDh) {
    int rc = check_ready(c, false);
    if (rc) return rc;
    if (c->n_dir == 0) return FVDG_OK;
    if (!UDh) return fail(FVDG_ERR_ARG, "null Dirichlet states");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemcpyAsync(c->UD.p, UDh, (size_t)c->n_dir * c->tab.Qf * 5 * sizeof(double), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FVDG_OK;
}

int fvdg_residual(fvdg_ctx* c, double t, int parts, double* Rh) {
    (void)t;  // time enters only through the Dirichlet states, which the caller sets
    int rc = check_ready(c, true);
    if (rc) return rc;
    if (parts & ~FVDG_PART_ALL) return fail(FVDG_ERR_ARG, "unknown residual part bits");
    HIP_TRY(hipSetDevice(c->device));
    rc = launch_residual(c, c->U.p, parts, nullptr, nullptr);
    if (rc) return rc;
    if (Rh)
        HIP_TRY(hipMemcpyAsync(Rh, c->R.p, (size_t)c->E * 5 * c->tab.N * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FVDG_OK;
}

int fvdg_mass_diagonal(fvdg_ctx* c, double* M) {
    int rc = check_ready(c, false);
    if (rc) return rc;
    if (!M) return fail(FVDG_ERR_ARG, "null output");
    const int N = c->tab.N;
    // reference driver form b(5i+k) += phi_i^2 * w_q * |det| summed over q (VortexBatch.cpp:155-163)
    for (int64_t e = 0; e < c->E; ++e)
        for (int k = 0; k < N; ++k) {
            double s = 0;
            for (int q = 0; q < c->tab.Q; ++q) {
                const double ph = c->tab.V[q * N + k];
                s += ph * ph * c->tab.vq_w[q] * c->cell_det_host[e];
            }
            for (int v = 0; v < 5; ++v) M[e * 5 * N + 5 * k + v] = s;
        }
    return FVDG_OK;
}

int fvdg_ssprk3_step(fvdg_ctx* c, double dt, int n_steps) {
    int rc = check_ready(c, true);
    if (rc) return rc;
    if (n_steps < 0 || !(dt >= 0)) return fail(FVDG_ERR_ARG, "bad dt or step count");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemsetAsync(c->bad.p, 0, sizeof(int32_t), c->stream));
    for (int i = 0; i < n_steps; ++i) {
        rc = ssprk3_once(c, dt, nullptr);
        if (rc) return rc;
    }
    return finish_checked(c);
}

int fvdg_ssprk3_stage(fvdg_ctx* c, double dt, int stage) {
    int rc = check_ready(c, true);
    if (rc) return rc;
    if (stage < 0 || stage > 2 || !(dt >= 0)) return fail(FVDG_ERR_ARG, "bad dt or stage index");
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipMemsetAsync(c->bad.p, 0, sizeof(int32_t), c->stream));
    rc = ssprk3_stage(c, dt, stage, nullptr);
    if (rc) return rc;
    return finish_checked(c);
}

int fvdg_stable_dt(fvdg_ctx* c, double cfl, double* dt) {
    if (!dt || !(cfl > 0)) return fail(FVDG_ERR_ARG, "bad cfl or null output");
    double lam = 0;
    int rc = fvdg_max_wave_speed(c, &lam);
    if (rc) return rc;
    if (!(lam > 0) || !std::isfinite(lam)) return fail(FVDG_ERR_NONFINITE, "maximum wave speed is not positive and finite");
    *dt = cfl * c->h_min / ((2 * c->order + 1) * lam);
    return FVDG_OK;
}

int fvdg_stable_dt_noalloc(fvdg_ctx* c, double cfl, double* dt) {
    if (!dt || !(cfl > 0)) return fail(FVDG_ERR_ARG, "bad cfl or null output");
    int rc = check_ready(c, true);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    double* scr = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch[c->device] : nullptr;
    if (!scr) return fail(FVDG_ERR_STATE, "scratch address of this device not set (fvdg_create sets it)");
    KernelArgs a;
    a.E = c->E;
    a.F = c->F;
    a.N = c->tab.N;
    a.Qf = c->tab.Qf;
    a.gamma = c->gamma;
    a.U = c->U.p;
    a.fcell = c->fcell.p;
    a.fcombo = c->fcombo.p;
    a.Phi = c->tPhi.p;
    const int64_t n = c->F * c->tab.Qf;
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(nblocks(n, 256), kScratchPartials));
    hipLaunchKernelGGL(k_max_wave_speed_gs, dim3(nb), dim3(256), 0, c->stream, a, scr);
    hipLaunchKernelGGL(k_max_partials, dim3(1), dim3(256), 0, c->stream, scr, nb, scr + kScratchPartials);
    HIP_TRY(hipGetLastError());
    double lam = 0;
    HIP_TRY(hipMemcpyAsync(&lam, scr + kScratchPartials, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (!(lam > 0) || !std::isfinite(lam)) return fail(FVDG_ERR_NONFINITE, "maximum wave speed is not positive and finite");
    *dt = cfl * c->h_min / ((2 * c->order + 1) * lam);
    return FVDG_OK;
}

int fvdg_max_wave_speed(fvdg_ctx* c, double* out) {
    int rc = check_ready(c, true);
    if (rc) return rc;
    if (!out) return fail(FVDG_ERR_ARG, "null output");
    HIP_TRY(hipSetDevice(c->device));
    KernelArgs a;
    a.E = c->E;
    a.F = c->F;
    a.N = c->tab.N;
    a.Qf = c->tab.Qf;
    a.gamma = c->gamma;
    a.U = c->U.p;
    a.fcell = c->fcell.p;
    a.fcombo = c->fcombo.p;
    a.Phi = c->tPhi.p;
    const int64_t n = c->F * c->tab.Qf;
    const unsigned nb = nblocks(n, 256);
    hipLaunchKernelGGL(k_max_wave_speed, dim3(nb), dim3(256), 0, c->stream, a, c->red.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> part(nb);
    HIP_TRY(hipMemcpyAsync(part.data(), c->red.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double m = 0;
    for (double v : part) m = (v != v || v > m) ? v : m;  // a NaN partial stays NaN
    *out = m;
    return FVDG_OK;
}

int fvdg_conservation_sums(fvdg_ctx* c, double out[5]) {
    int rc = check_ready(c, true);
    if (rc) return rc;
    if (!out) return fail(FVDG_ERR_ARG, "null output");
    HIP_TRY(hipSetDevice(c->device));
    const unsigned nb = nblocks(c->E, 256);
    hipLaunchKernelGGL(k_conservation, dim3(nb), dim3(256), 0, c->stream, c->U.p, c->cgeo.p, c->E, c->tab.N, c->red.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> part(nb * 5);
    HIP_TRY(hipMemcpyAsync(part.data(), c->red.p, nb * 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int v = 0; v < 5; ++v) {
        double s = 0;
        for (unsigned b = 0; b < nb; ++b) s += part[b * 5 + v];
        out[v] = s;
    }
    return FVDG_OK;
}

int fvdg_limiter_bounds(fvdg_ctx* c, const double* U_prev, int neighbour_averages) {
    int rc = check_ready(c, U_prev == nullptr);
    if (rc) return rc;
    HIP_TRY(hipSetDevice(c->device));
    const double* src = c->U.p;
    if (U_prev) {
        const size_t n = (size_t)c->E * 5 * c->tab.N;
        if (c->lim_buf.n != n) HIP_TRY(c->lim_buf.alloc(n));
        HIP_TRY(hipMemcpyAsync(c->lim_buf.p, U_prev, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        src = c->lim_buf.p;
    }
    launch_limit_extrema(c, src, neighbour_averages != 0);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->lim_bounds = true;
    return FVDG_OK;
}

int fvdg_limiter_apply(fvdg_ctx* c, double* U, int parts) {
    int rc = check_ready(c, U == nullptr);
    if (rc) return rc;
    if (parts < 0 || parts > 3) return fail(FVDG_ERR_ARG, "parts must combine FVDG_LIMIT_BOUNDS and FVDG_LIMIT_POSITIVITY");
    if ((parts & FVDG_LIMIT_BOUNDS) && !c->lim_bounds)
        return fail(FVDG_ERR_STATE, "fvdg_limiter_bounds has not been called for this mesh");
    if (parts == 0) return FVDG_OK;
    HIP_TRY(hipSetDevice(c->device));
    double* dst = c->U.p;
    const size_t n = (size_t)c->E * 5 * c->tab.N;
    if (U) {
        if (c->lim_buf.n != n) HIP_TRY(c->lim_buf.alloc(n));
        HIP_TRY(hipMemcpyAsync(c->lim_buf.p, U, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        dst = c->lim_buf.p;
    }
    launch_limit_apply(c, dst, parts);
    HIP_TRY(hipGetLastError());
    if (U) HIP_TRY(hipMemcpyAsync(U, dst, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return FVDG_OK;
}

int fvdg_set_stage_limiter(fvdg_ctx* c, int parts, int neighbour_averages) {
    if (!c) return fail(FVDG_ERR_ARG, "null context");
    if (parts < 0 || parts > 3) return fail(FVDG_ERR_ARG, "parts must combine FVDG_LIMIT_BOUNDS and FVDG_LIMIT_POSITIVITY");
    c->lim_parts = parts;
    c->lim_avg = neighbour_averages != 0;
    return FVDG_OK;
}

int fvdg_flow_diagnostics(fvdg_ctx* c, double out[8]) {
    int rc = check_ready(c, true);
    if (rc) return rc;
    if (!out) return fail(FVDG_ERR_ARG, "null output");
    HIP_TRY(hipSetDevice(c->device));
    const int N = c->tab.N, Q = c->tab.Q, L = limiter_lanes(Q), EB = 256 / L;
    const unsigned nb = point_blocks(c);
    if ((size_t)nb * fvdg_diag::kNumDiag > c->diag.n) return fail(FVDG_ERR_STATE, "diagnostics buffer too small");
    const size_t lds = sizeof(double) * ((size_t)EB * 5 * N + (size_t)fvdg_diag::kNumDiag * 256);
    switch (L) {
        case 16: hipLaunchKernelGGL(fvdg_diag::k_flow_diagnostics<16>, dim3(nb), dim3(256), lds, c->stream, c->U.p, c->cgeo.p, c->tV.p, c->tG.p, c->tWq.p, c->E, N, Q, c->gamma, c->diag.p); break;
        case 32: hipLaunchKernelGGL(fvdg_diag::k_flow_diagnostics<32>, dim3(nb), dim3(256), lds, c->stream, c->U.p, c->cgeo.p, c->tV.p, c->tG.p, c->tWq.p, c->E, N, Q, c->gamma, c->diag.p); break;
        case 64: hipLaunchKernelGGL(fvdg_diag::k_flow_diagnostics<64>, dim3(nb), dim3(256), lds, c->stream, c->U.p, c->cgeo.p, c->tV.p, c->tG.p, c->tWq.p, c->E, N, Q, c->gamma, c->diag.p); break;
        default: hipLaunchKernelGGL(fvdg_diag::k_flow_diagnostics<128>, dim3(nb), dim3(256), lds, c->stream, c->U.p, c->cgeo.p, c->tV.p, c->tG.p, c->tWq.p, c->E, N, Q, c->gamma, c->diag.p); break;
    }
    HIP_TRY(hipGetLastError());
    std::vector<double> part((size_t)nb * fvdg_diag::kNumDiag);
    HIP_TRY(hipMemcpyAsync(part.data(), c->diag.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double r[8] = {0, 0, 0, 0, 0, part[5], part[6], part[7]};
    for (unsigned b = 0; b < nb; ++b) {
        const double* p = &part[(size_t)b * fvdg_diag::kNumDiag];
        for (int i = 0; i < 5; ++i) r[i] += p[i];
        r[5] = std::min(r[5], p[5]);
        r[6] = std::min(r[6], p[6]);
        r[7] = std::max(r[7], p[7]);
    }
    for (int i = 0; i < 8; ++i) out[i] = r[i];
    return FVDG_OK;
}

// fvdg_time_stages with its two events (created and destroyed by the caller)
static int time_stages_with(fvdg_ctx* c, double dt, int n_stages, double* ms, hipEvent_t e0, hipEvent_t e1) {
    // pass 1: whole stages, no per-kernel events inside (the throughput figure)
    HIP_TRY(hipEventRecord(e0, c->stream));
    for (int s = 0; s < n_stages / 3; ++s) {
        const int rc = ssprk3_once(c, dt, nullptr);
        if (rc) return rc;
    }
    HIP_TRY(hipEventRecord(e1, c->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, e0, e1));
    ms[0] = t / n_stages;
    // pass 2: per-kernel split (events between kernels)
    double split[4] = {0, 0, 0, 0};
    for (int s = 0; s < n_stages / 3; ++s) {
        const int rc = ssprk3_once(c, dt, split);
        if (rc) return rc;
    }
    for (int i = 1; i < 4; ++i) ms[i] = split[i] / n_stages;
    return FVDG_OK;
}

int fvdg_stage_kernels(fvdg_ctx* c, char* buf, int len) {
    int rc = check_ready(c, false);
    if (rc) return rc;
    if (!buf || len <= 0) return fail(FVDG_ERR_ARG, "null or empty buffer");
    const std::string n = stage_kernel_names(c);
    if ((int)n.size() >= len) return fail(FVDG_ERR_ARG, "buffer too small for the kernel names");
    std::memcpy(buf, n.c_str(), n.size() + 1);
    return FVDG_OK;
}

int fvdg_time_stages(fvdg_ctx* c, double dt, int n_stages, double* ms) {
    int rc = check_ready(c, true);
    if (rc) return rc;
    if (!ms || n_stages <= 0 || n_stages % 3) return fail(FVDG_ERR_ARG, "n_stages must be a positive multiple of 3");
    HIP_TRY(hipSetDevice(c->device));
    for (int i = 0; i < 4; ++i) ms[i] = 0;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(FVDG_ERR_HIP, "hipEventCreate failed");
    }
    rc = time_stages_with(c, dt, n_stages, ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

}  // extern "C"
