// Device-side internals shared by the solver (fvdg_solver.hip) and the ensemble
// (fvdg_ensemble.hip) translation units of libfvdg.so: the solver context and the
// stage driver.  Not installed; the public ABI is include/fvdg.h.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "fvdg_internal.hpp"

#define HIP_TRY(expr)                                                                         \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            return fail(FVDG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));     \
    } while (0)

namespace fvdg {

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc(&p, count * sizeof(T));
    }
    hipError_t upload(const std::vector<T>& v, hipStream_t s) {
        hipError_t e = alloc(v.size());
        if (e != hipSuccess || v.empty()) return e;
        e = hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return e;
        return hipStreamSynchronize(s);
    }
};

}  // namespace fvdg

using fvdg::DevBuf;
using fvdg::SchemeTables;

struct fvdg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = true;  // false: a stream lent by an ensemble
    int grid_mult = 8;        // persistent face kernels: blocks per compute unit (tuned by ensembles)
    int flux_threads = 256;   // block size of the row-broadcast face-flux kernel: 128, 256 or 512 (tuned by ensembles)
    // scheme
    int order = -1, flux = FVDG_FLUX_LF, precision = FVDG_PREC_FP64;
    double gamma = 1.4;
    SchemeTables tab;
    // mesh
    int64_t E = 0, Epad = 0, F = 0, n_dir = 0;
    bool have_mesh = false, have_state = false;
    std::vector<int64_t> dir_faces;       // Dirichlet face ids, host copy
    std::vector<double> dir_xyz;          // their face quadrature points
    std::vector<double> cell_det_host;    // for the mass diagonal
    double h_min = 0;                     // min over cells of 3 * volume / (sum of face areas)
    // device geometry
    DevBuf<double> cgeo;        // E*10: A^{-1} (9), |det A|
    DevBuf<int32_t> cfaces;     // E*4: face*2 + side
    DevBuf<uint8_t> cmeta;      // E*4: vertex combination of that side | 0x80 on boundary faces
    DevBuf<int32_t> cnb;        // E*4: neighbour cell across each local face, -1 on the boundary
    DevBuf<int32_t> fcell;      // F*2
    DevBuf<uint8_t> fcombo;     // F*2
    DevBuf<double> fgeo;        // F*4: normal (3), 2*area
    DevBuf<uint8_t> fbc;        // F: 255 interior, else BoundaryType
    DevBuf<int32_t> fdir;       // F: Dirichlet slot or -1
    DevBuf<double> UD;          // n_dir*Qf*5
    // device tables
    DevBuf<double> tV, tG, tWq, tPhi, tWf, tMhat;
    DevBuf<double> tVA, tGA, tWqa, tPhiT;  // zero-padded MFMA operand tables, transposed face table
    DevBuf<double> tPhiC;                 // canonical face tables (8 vertex orders)
    DevBuf<float> tVA32, tGA32, tPhiC32;  // FP32 copies for the mixed-precision kernels
    DevBuf<double> tGs, tIC;              // mixed: sum_q w_q grad phi_j (N*3), sum_g w_g PhiC (8*N)
    DevBuf<uint16_t> tVB3, tGB3;          // BF16x3: volume operand tables split into 3 bf16 parts
    DevBuf<uint8_t> tCanon;               // per combo: canonical index * 3 + rotation
    DevBuf<int32_t> tPerm;                // face-rule point permutations of the 3 rotations
    DevBuf<double> tPhiB;                 // canonical tables as MFMA B operands (face test)
    DevBuf<float> tPhiB32;
    // face tiles (k_face_tile, P2-P4): faces grouped by (side-0, side-1 or boundary) vertex
    // combination in tiles of 16, the per-combination trace and test operand tables, and the
    // per (face, side) test contributions
    int face_tile = 1;                    // 0 off, 1 mixed precision (FP32), 2 also FP64 (FVDG_FACE_TILE)
    int64_t n_face_tiles = 0;
    int tile_order = 2;                   // 0 class order, 1 spatial, 2 spatial in 8 XCD ranges (FVDG_TILE_ORDER)
    DevBuf<int32_t> trec, tclass;         // n_face_tiles*16 column records (int4: face or -1, cells, kind),
                                          // n_face_tiles classes
    DevBuf<double> tAt, tAe;              // FP64 operand layouts
    DevBuf<float> tAt32, tAe32;           // FP32 operand layouts
    DevBuf<float> tICc32;                 // [24][NP] sum_g w_g Phi_cb[g][j] (FP32 side-1 reference term)
    DevBuf<uint16_t> tAt3, tAe3;          // BF16x3 face-tile operand tables (k_face_tile_bf3)
    bool bf3_faces = true;                // BF16x3: the face contractions as bf16 part products too
                                          // (FVDG_BF16X3_FACES=0: FP32 face contractions)
    DevBuf<double> Cf;                    // FP64 contributions [F][2][5][NP]
    DevBuf<float> Cf32;                   // FP32 contributions
    int n_cu = 256;
    // device state: U (state), Ua/Ub (stage ping-pong), R (residual)
    DevBuf<double> U, Ua, Ub, R;
    DevBuf<double> Fq, Fh;           // Fq: simple-kernel scratch E*Q*15; Fh: face fluxes F*Qf*5
    DevBuf<float> Fh32;              // mixed precision: face flux differences F*Qf*5
    DevBuf<double> Fref;             // mixed precision: per-face reference flux F*5
    bool simple = false;             // FVDG_KERNELS=simple
    bool owner_faces = false;        // FVDG_FACES=owner: one-pass owner-computes face kernel
    bool lds_face_test = false;      // FVDG_FACE_TEST=lds: the LDS/VALU face test kernel
    bool lds_face_flux = false;      // FVDG_FACE_FLUX=lds: the LDS-staged face-flux kernel at P2-P4 (FP64)
    bool elem_fused = true;          // k_elem_mfma after the face-flux pass (FVDG_ELEM=split: volume kernel +
                                     // face-test kernel)
    DevBuf<double> red;              // reduction partials
    DevBuf<int32_t> bad;             // 1 word: set by the stage update when an output is non-finite or a
                                     // mean density <= 0 (cleared and read by fvdg_ssprk3_step/_stage)
    DevBuf<double> edt;              // per-element time step (ensembles); empty: the call's dt
    int64_t n_red = 0;
    // limiter (PositiveLimiter): per-element extremes of the previous state, staging of
    // host arrays, and the limiter applied after every SSP-RK3 stage (parts 0: none)
    DevBuf<double> lim_ext;          // E*10
    DevBuf<double> lim_buf;          // E*5N, allocated on first host-array call
    bool lim_bounds = false;         // lim_ext holds bounds of the current mesh/scheme
    int lim_parts = 0, lim_avg = 0;
    DevBuf<double> diag;             // flow diagnostics: 8 partials per block
    // implicit path (allocated on first use): assembled block rows (5 slots), right-hand side,
    // per-unknown mass, block-Jacobi factors, GMRES and Picard work vectors
    DevBuf<double> imp_A, imp_rhs, imp_b, imp_Md, imp_LU, imp_V, imp_w, imp_z, imp_part, imp_xold, imp_dx;
    DevBuf<int32_t> imp_col, imp_piv, imp_dslot, imp_flag, imp_self;
    bool imp_have = false;           // imp_A holds a system assembled on the current mesh and scheme

    ~fvdg_ctx() {
        if (stream && owns_stream) (void)hipStreamDestroy(stream);
    }
};

namespace fvdg {
// one SSP-RK3 step of the context's device state (stream-ordered, not synchronised);
// ms (nullable): per-kernel event times, accumulated
int ssprk3_once(fvdg_ctx* c, double dt, double* ms);
// per face: max over its points of the traces' |u| + c (host copy of F values)
int face_wave_speeds(fvdg_ctx* c, std::vector<double>& out);
// PositiveLimiter on device arrays: bounds from prev (constructMinMax; skipped when parts has
// no FVDG_LIMIT_BOUNDS), then apply to cur in place (stream-ordered)
void limit_state(fvdg_ctx* c, const double* prev, double* cur, int parts, int neighbour_averages);
}  // namespace fvdg
