// Batched ensembles: many independent DG Euler problems of uneven size and order,
// stepped together on one GPU.
//
// The problems of one order form one disjoint-union mesh (independent connected
// components), so a single launch of each stage kernel covers all of them; each
// problem keeps its own time step through a per-element step size.  Each order group
// has its own stream: one SSP-RK3 step forks from the ensemble's stream to every group
// (an event), runs the group's three stages (face-flux pass + element kernel each), and
// joins back, so groups of different orders run concurrently and a small group does not
// leave the device idle.  One step is captured once into a HIP graph (the fork/join
// becomes its branches) and replayed per step.
//
// The autotuner times, per order group on the ensemble's own data, the stage's kernel
// choices: for FP32 and BF16x3 groups at P2-P4 the face tiles (16 faces of one vertex-combination class per
// wavefront, both face contractions on the matrix cores); the fused element kernel or the volume +
// face-test kernels, with the row-broadcast face-flux kernel in 128-, 256- or 512-thread blocks
// (16 faces per 256 threads) or the LDS-staged one; keeping the fastest; the state is restored
// after every trial.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>

#include "fvdg_device.hpp"

using namespace fvdg;

struct fvdg_ensemble {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = FVDG_PREC_FP64;
    struct Group {
        int order = 0;
        std::vector<int> problems;           // problem ids in union order
        std::vector<int64_t> cell_off;       // per member: cell offset in the union (+ total)
        std::vector<int64_t> face_off;
        std::unique_ptr<fvdg_mesh> mesh;
        fvdg_ctx* ctx = nullptr;
    };
    std::vector<Group> groups;
    std::vector<int> prob_group, prob_member;
    std::vector<double> dt;  // per problem
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool serial = false;  // every group on the ensemble's stream, one after another (comparison)
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~fvdg_ensemble() {
        drop_graph();
        for (auto& g : groups) fvdg_destroy(g.ctx);
        if (fork) (void)hipEventDestroy(fork);
        for (hipEvent_t ev : joins) (void)hipEventDestroy(ev);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int check_problem(fvdg_ensemble* e, int p) {
    if (!e) return fail(FVDG_ERR_ARG, "null ensemble");
    if (p < 0 || p >= (int)e->prob_group.size()) return fail(FVDG_ERR_ARG, "problem index out of range");
    return FVDG_OK;
}

// per-element step sizes of every group from the problems' dt
int upload_dt(fvdg_ensemble* e) {
    for (auto& g : e->groups) {
        std::vector<double> edt(g.ctx->E);
        for (size_t m = 0; m < g.problems.size(); ++m)
            std::fill(edt.begin() + g.cell_off[m], edt.begin() + g.cell_off[m + 1], e->dt[g.problems[m]]);
        // copy into the existing buffer when there is one: captured graphs hold its address
        if (g.ctx->edt.n != edt.size()) HIP_TRY(g.ctx->edt.alloc(edt.size()));
        HIP_TRY(hipMemcpyAsync(g.ctx->edt.p, edt.data(), edt.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
        HIP_TRY(hipStreamSynchronize(e->stream));
    }
    return FVDG_OK;
}

// one step of every group: forked from and joined back into the ensemble's stream
int step_once(fvdg_ensemble* e) {
    if (e->serial) {
        for (auto& g : e->groups) {
            hipStream_t own = g.ctx->stream;
            g.ctx->stream = e->stream;
            const int rc = ssprk3_once(g.ctx, 0.0, nullptr);
            g.ctx->stream = own;
            if (rc) return rc;
        }
        return FVDG_OK;
    }
    HIP_TRY(hipEventRecord(e->fork, e->stream));
    for (size_t i = 0; i < e->groups.size(); ++i) {
        fvdg_ctx* c = e->groups[i].ctx;
        HIP_TRY(hipStreamWaitEvent(c->stream, e->fork, 0));
        const int rc = ssprk3_once(c, 0.0, nullptr);
        if (rc) return rc;
        HIP_TRY(hipEventRecord(e->joins[i], c->stream));
        HIP_TRY(hipStreamWaitEvent(e->stream, e->joins[i], 0));
    }
    return FVDG_OK;
}

// the stage kernels of one group, timed over `steps` steps on its own stream (after one warm step)
int time_group(fvdg_ctx* c, int steps, float* ms) {
    int rc = ssprk3_once(c, 0.0, nullptr);
    if (rc) return rc;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return fail(FVDG_ERR_HIP, "hipEventCreate failed");
    }
    hipError_t he = hipEventRecord(a, c->stream);
    for (int i = 0; i < steps && rc == FVDG_OK && he == hipSuccess; ++i) rc = ssprk3_once(c, 0.0, nullptr);
    if (he == hipSuccess) he = hipEventRecord(b, c->stream);
    if (he == hipSuccess) he = hipEventSynchronize(b);
    if (he == hipSuccess) he = hipEventElapsedTime(ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    if (he != hipSuccess) return fail(FVDG_ERR_HIP, std::string("group timing: ") + hipGetErrorString(he));
    return FVDG_OK;
}

}  // namespace

extern "C" {

int fvdg_ensemble_create(fvdg_ensemble** out, int device, int n, const fvdg_mesh* const* meshes, const int* orders,
                         int flux, double gamma, int precision) {
    if (!out || n <= 0 || !meshes || !orders) return fail(FVDG_ERR_ARG, "null or empty ensemble description");
    std::unique_ptr<fvdg_ensemble> e(new fvdg_ensemble);
    e->device = device;
    e->precision = precision;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&e->fork, hipEventDisableTiming));
    std::map<int, std::vector<int>> by_order;
    for (int p = 0; p < n; ++p) {
        if (!meshes[p]) return fail(FVDG_ERR_ARG, "null mesh in ensemble");
        by_order[orders[p]].push_back(p);
    }
    e->prob_group.assign(n, -1);
    e->prob_member.assign(n, -1);
    e->dt.assign(n, 0.0);
    for (const auto& kv : by_order) {
        fvdg_ensemble::Group g;
        g.order = kv.first;
        g.problems = kv.second;
        std::vector<const fvdg_mesh*> parts;
        for (int p : g.problems) parts.push_back(meshes[p]);
        try {
            g.mesh.reset(union_mesh(parts, g.cell_off, g.face_off));
        } catch (const std::exception& ex) {
            return fail(FVDG_ERR_MESH, ex.what());
        }
        int rc = fvdg_create(&g.ctx, device);  // the group works on its own stream
        if (rc) return rc;
        hipEvent_t join = nullptr;
        if (hipEventCreateWithFlags(&join, hipEventDisableTiming) != hipSuccess) {
            fvdg_destroy(g.ctx);
            return fail(FVDG_ERR_HIP, "hipEventCreateWithFlags failed");
        }
        e->joins.push_back(join);
        rc = fvdg_set_scheme(g.ctx, g.order, flux, gamma, precision);
        if (rc == FVDG_OK) rc = fvdg_set_mesh(g.ctx, g.mesh.get());
        if (rc) {
            fvdg_destroy(g.ctx);
            return rc;
        }
        const int gi = (int)e->groups.size();
        for (size_t m = 0; m < g.problems.size(); ++m) {
            e->prob_group[g.problems[m]] = gi;
            e->prob_member[g.problems[m]] = (int)m;
        }
        e->groups.push_back(std::move(g));
    }
    *out = e.release();
    return FVDG_OK;
}

void fvdg_ensemble_destroy(fvdg_ensemble* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    delete e;
}

int fvdg_ensemble_num_cells(fvdg_ensemble* e, int p, int64_t* n_cells) {
    int rc = check_problem(e, p);
    if (rc) return rc;
    const auto& g = e->groups[e->prob_group[p]];
    const int m = e->prob_member[p];
    if (n_cells) *n_cells = g.cell_off[m + 1] - g.cell_off[m];
    return FVDG_OK;
}

int fvdg_ensemble_upload(fvdg_ensemble* e, int p, const double* U) {
    int rc = check_problem(e, p);
    if (rc) return rc;
    if (!U) return fail(FVDG_ERR_ARG, "null state");
    auto& g = e->groups[e->prob_group[p]];
    HIP_TRY(hipStreamSynchronize(g.ctx->stream));
    const int m = e->prob_member[p];
    const size_t NV = 5 * (size_t)g.ctx->tab.N;
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemcpyAsync(g.ctx->U.p + g.cell_off[m] * NV, U, (g.cell_off[m + 1] - g.cell_off[m]) * NV * sizeof(double),
                           hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    g.ctx->have_state = true;
    return FVDG_OK;
}

int fvdg_ensemble_download(fvdg_ensemble* e, int p, double* U) {
    int rc = check_problem(e, p);
    if (rc) return rc;
    if (!U) return fail(FVDG_ERR_ARG, "null state");
    auto& g = e->groups[e->prob_group[p]];
    HIP_TRY(hipStreamSynchronize(g.ctx->stream));
    const int m = e->prob_member[p];
    const size_t NV = 5 * (size_t)g.ctx->tab.N;
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipMemcpyAsync(U, g.ctx->U.p + g.cell_off[m] * NV, (g.cell_off[m + 1] - g.cell_off[m]) * NV * sizeof(double),
                           hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FVDG_OK;
}

int fvdg_ensemble_max_wave_speeds(fvdg_ensemble* e, double* out) {
    if (!e || !out) return fail(FVDG_ERR_ARG, "null argument");
    HIP_TRY(hipSetDevice(e->device));
    for (auto& g : e->groups) {
        std::vector<double> fw;
        const int rc = face_wave_speeds(g.ctx, fw);
        if (rc) return rc;
        for (size_t m = 0; m < g.problems.size(); ++m)
            out[g.problems[m]] = *std::max_element(fw.begin() + g.face_off[m], fw.begin() + g.face_off[m + 1]);
    }
    return FVDG_OK;
}

int fvdg_ensemble_set_dt(fvdg_ensemble* e, const double* dt) {
    if (!e || !dt) return fail(FVDG_ERR_ARG, "null argument");
    for (size_t p = 0; p < e->dt.size(); ++p) {
        if (!(dt[p] >= 0)) return fail(FVDG_ERR_ARG, "negative or NaN time step");
        e->dt[p] = dt[p];
    }
    HIP_TRY(hipSetDevice(e->device));
    return upload_dt(e);  // device array contents change, its address does not: graphs stay valid
}

int fvdg_ensemble_step(fvdg_ensemble* e, int n_steps, int use_graph) {
    if (!e || n_steps < 0) return fail(FVDG_ERR_ARG, "bad arguments");
    for (auto& g : e->groups)
        if (!g.ctx->have_state) return fail(FVDG_ERR_STATE, "an order group has no uploaded state");
    for (auto& g : e->groups)
        if (g.ctx->edt.n != (size_t)g.ctx->E) return fail(FVDG_ERR_STATE, "fvdg_ensemble_set_dt has not been called");
    HIP_TRY(hipSetDevice(e->device));
    if (!use_graph) {
        for (int i = 0; i < n_steps; ++i) {
            const int rc = step_once(e);
            if (rc) return rc;
        }
    } else {
        if (!e->exec) {
            HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
            const int rc = step_once(e);
            hipGraph_t graph = nullptr;
            const hipError_t ce = hipStreamEndCapture(e->stream, &graph);
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            HIP_TRY(ce);
            e->graph = graph;
            HIP_TRY(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
        }
        for (int i = 0; i < n_steps; ++i) HIP_TRY(hipGraphLaunch(e->exec, e->stream));
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    return FVDG_OK;
}

int fvdg_ensemble_time_steps(fvdg_ensemble* e, int n_steps, int use_graph, double* ms_per_step) {
    if (!e || !ms_per_step || n_steps <= 0) return fail(FVDG_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(e->device));
    int rc = fvdg_ensemble_step(e, 1, use_graph);  // warm (and capture)
    if (rc) return rc;
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    HIP_TRY(hipEventRecord(a, e->stream));
    rc = fvdg_ensemble_step(e, n_steps, use_graph);
    HIP_TRY(hipEventRecord(b, e->stream));
    HIP_TRY(hipEventSynchronize(b));
    float t = 0;
    HIP_TRY(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms_per_step = t / n_steps;
    return rc;
}

int fvdg_ensemble_autotune(fvdg_ensemble* e, int trial_steps, int* choices) {
    if (!e || trial_steps <= 0) return fail(FVDG_ERR_ARG, "bad arguments");
    HIP_TRY(hipSetDevice(e->device));
    e->drop_graph();
    HIP_TRY(hipStreamSynchronize(e->stream));
    int gi = 0;
    for (auto& g : e->groups) {
        fvdg_ctx* c = g.ctx;
        if (c->edt.n != (size_t)c->E || !c->have_state) return fail(FVDG_ERR_STATE, "set states and dt before tuning");
        const size_t nd = (size_t)c->E * 5 * c->tab.N;
        DevBuf<double> save;
        HIP_TRY(save.alloc(nd));
        HIP_TRY(hipMemcpyAsync(save.p, c->U.p, nd * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        // candidates: (fused element kernel, face pass: -1 = face tiles, 0 = LDS-staged face flux, else
        // the row-broadcast face flux's block size)
        // (BF16x3: its fused element kernel runs only on its face tiles)
        const bool dpp_ok = c->order >= 2 && c->order <= 4;
        const bool bf3 = c->precision == FVDG_PREC_BF16X3;
        const bool tile_ok = c->order >= 2 && c->order <= 4 && c->precision != FVDG_PREC_FP64 &&
                             c->n_face_tiles > 0 && c->tab.Qf == 24;
        const bool elem_ok = c->order <= 4 && (!bf3 || tile_ok);
        double best = 1e300;
        int best_elem = c->elem_fused, best_flux = c->lds_face_flux ? 0 : c->flux_threads;
        const int tile0 = c->face_tile;
        for (int elem = elem_ok ? 1 : 0; elem >= 0; --elem)
            for (int flux : {-1, 256, 128, 512, 0}) {
                if (flux == -1 && (!tile_ok || elem == 0)) continue;
                if (bf3 && elem == 1 && flux != -1) continue;
                if (flux > 0 && !dpp_ok) continue;
                if (flux == 0 && dpp_ok && elem == 0) continue;  // the split path with the LDS flux: not competitive
                c->elem_fused = elem != 0;
                c->face_tile = flux == -1 ? 1 : 0;
                c->lds_face_flux = flux == 0;
                c->flux_threads = flux > 0 ? flux : 256;
                float t = 0;
                int rc = time_group(c, trial_steps, &t);
                if (rc) return rc;
                if (t < best) best = t, best_elem = elem, best_flux = flux;
                HIP_TRY(hipMemcpyAsync(c->U.p, save.p, nd * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
            }
        c->elem_fused = best_elem != 0;
        c->face_tile = best_flux == -1 ? 1 : (tile_ok ? 0 : tile0);
        c->lds_face_flux = best_flux == 0;
        c->flux_threads = best_flux > 0 ? best_flux : 256;
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (choices) {
            choices[2 * gi] = best_elem;
            choices[2 * gi + 1] = best_flux;
        }
        ++gi;
    }
    return FVDG_OK;
}

int fvdg_ensemble_set_serial(fvdg_ensemble* e, int serial) {
    if (!e) return fail(FVDG_ERR_ARG, "null ensemble");
    HIP_TRY(hipSetDevice(e->device));
    HIP_TRY(hipDeviceSynchronize());
    e->drop_graph();
    e->serial = serial != 0;
    return FVDG_OK;
}

int fvdg_ensemble_num_groups(fvdg_ensemble* e, int* n_groups, int* orders) {
    if (!e) return fail(FVDG_ERR_ARG, "null ensemble");
    if (n_groups) *n_groups = (int)e->groups.size();
    if (orders)
        for (size_t i = 0; i < e->groups.size(); ++i) orders[i] = e->groups[i].order;
    return FVDG_OK;
}

}  // extern "C"
