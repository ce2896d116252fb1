// C++ owner of one libfvdg device context and the mesh bound to it.
//
// Host C++ reaches the GPU only through the C ABI of fvdg.h; this header wraps it
// with RAII and turns nonzero status codes into std::runtime_error (the error
// style of the reference's mesh layer, src/Mesh/GeneralMesh.cpp:42-117).  It is
// the common base of ExplicitConvection.h (the residual) and ExplicitSSPRK3.h
// (the time stepper).
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "Mesh/ComputingMesh.h"
#include "fvdg.h"

namespace fvdg {

inline void check(int status, const char* what) {
    if (status != FVDG_OK)
        throw std::runtime_error(std::string(what) + ": " + fvdg_last_error() + " (status " + std::to_string(status) + ")");
}

// Host mesh handle built from a ComputingMesh's arrays (points, tet nodes, face
// nodes as seen from both sides, face cells, boundary types).
class MeshHandle {
public:
    explicit MeshHandle(const ComputingMesh& cm) {
        const int64_t np = cm.m_points.size(), nc = cm.m_cells.size(), nf = cm.m_faces.size();
        std::vector<double> pts(3 * np);
        std::vector<int64_t> cells(4 * nc), fnodes(3 * nf), fcells(2 * nf), fnodes_r(3 * nf);
        for (int64_t p = 0; p < np; ++p)
            for (int d = 0; d < 3; ++d) pts[3 * p + d] = cm.m_points[p][d];
        for (int64_t c = 0; c < nc; ++c)
            for (int a = 0; a < 4; ++a) cells[4 * c + a] = (int64_t)cm.m_cells[c].m_nodes[a];
        for (int64_t f = 0; f < nf; ++f) {
            const auto& F = cm.m_faces[f];
            for (int k = 0; k < 3; ++k) {
                fnodes[3 * f + k] = (int64_t)F.m_nodes[k];
                fnodes_r[3 * f + k] = (int64_t)F.m_nodes_r[k];
            }
            fcells[2 * f] = (int64_t)F.m_neighbor_cells[0];
            fcells[2 * f + 1] = F.is_boundary() ? -1 : (int64_t)F.m_neighbor_cells[1];
        }
        fvdg_mesh* m = nullptr;
        check(fvdg_mesh_create(&m, np, pts.data(), nc, cells.data(), nf, fnodes.data(), fcells.data(),
                               fnodes_r.data()),
              "fvdg_mesh_create");
        h_.reset(m);
        if (cm.m_boundaryTypes.size() == (size_t)nf) {
            std::vector<uint8_t> bc(nf);
            for (int64_t f = 0; f < nf; ++f) bc[f] = (uint8_t)cm.m_boundaryTypes[f];
            check(fvdg_mesh_set_boundary_types(m, bc.data()), "fvdg_mesh_set_boundary_types");
        }
    }
    fvdg_mesh* get() const { return h_.get(); }

private:
    struct Del {
        void operator()(fvdg_mesh* m) const { fvdg_mesh_destroy(m); }
    };
    std::unique_ptr<fvdg_mesh, Del> h_;
};

// One device context with a scheme and a mesh.  A scheme change calls only fvdg_set_scheme
// (which keeps an uploaded state unless the order changes); the mesh is rebound when the mesh
// object, its sizes or its boundary types change (bind() is cheap otherwise).
class Session {
public:
    explicit Session(int device = 0) {
        fvdg_ctx* c = nullptr;
        check(fvdg_create(&c, device), "fvdg_create");
        ctx_.reset(c);
    }

    void bind(const ComputingMesh& cm, int order, int flux, double gamma, int precision) {
        const bool scheme_changed = order != order_ || flux != flux_ || gamma != gamma_ || precision != precision_;
        if (scheme_changed) {
            check(fvdg_set_scheme(ctx_.get(), order, flux, gamma, precision), "fvdg_set_scheme");
            order_ = order;
            flux_ = flux;
            gamma_ = gamma;
            precision_ = precision;
        }
        if (&cm != mesh_key_ || cm.m_cells.size() != n_cells_ || cm.m_faces.size() != n_faces_ ||
            !same_boundary_types(cm)) {
            mesh_ = std::make_unique<MeshHandle>(cm);
            check(fvdg_set_mesh(ctx_.get(), mesh_->get()), "fvdg_set_mesh");
            mesh_key_ = &cm;
            n_cells_ = cm.m_cells.size();
            n_faces_ = cm.m_faces.size();
            bc_ = cm.m_boundaryTypes;
            int64_t nd = 0;
            check(fvdg_dirichlet_points(ctx_.get(), &nd, nullptr, nullptr), "fvdg_dirichlet_points");
            dir_faces_.assign(nd, 0);
            if (nd) check(fvdg_dirichlet_points(ctx_.get(), &nd, dir_faces_.data(), nullptr), "fvdg_dirichlet_points");
        }
    }
    // forget the bound mesh (call after modifying a mesh in place)
    void rebind() { mesh_key_ = nullptr; }

    fvdg_ctx* ctx() const { return ctx_.get(); }
    const std::vector<int64_t>& dirichlet_faces() const { return dir_faces_; }

    void upload(const double* U) { check(fvdg_state_upload(ctx(), U), "fvdg_state_upload"); }
    void download(double* U) { check(fvdg_state_download(ctx(), U), "fvdg_state_download"); }
    void set_dirichlet_states(const double* UD) { check(fvdg_set_dirichlet_states(ctx(), UD), "fvdg_set_dirichlet_states"); }
    void residual(double t, int parts, double* R) { check(fvdg_residual(ctx(), t, parts, R), "fvdg_residual"); }
    void stage(double dt, int s) { check(fvdg_ssprk3_stage(ctx(), dt, s), "fvdg_ssprk3_stage"); }
    void steps(double dt, int n) { check(fvdg_ssprk3_step(ctx(), dt, n), "fvdg_ssprk3_step"); }
    double stable_dt(double cfl) {
        double dt = 0;
        check(fvdg_stable_dt(ctx(), cfl, &dt), "fvdg_stable_dt");
        return dt;
    }
    void mass(double* M) { check(fvdg_mass_diagonal(ctx(), M), "fvdg_mass_diagonal"); }
    void conservation_sums(double out[5]) { check(fvdg_conservation_sums(ctx(), out), "fvdg_conservation_sums"); }
    void limiter_bounds(const double* U_prev, bool neighbour_averages) {
        check(fvdg_limiter_bounds(ctx(), U_prev, neighbour_averages ? 1 : 0), "fvdg_limiter_bounds");
    }
    void limiter_apply(double* U, int parts) { check(fvdg_limiter_apply(ctx(), U, parts), "fvdg_limiter_apply"); }
    void set_stage_limiter(int parts, bool neighbour_averages) {
        check(fvdg_set_stage_limiter(ctx(), parts, neighbour_averages ? 1 : 0), "fvdg_set_stage_limiter");
    }

private:
    bool same_boundary_types(const ComputingMesh& cm) const { return cm.m_boundaryTypes == bc_; }
    struct Del {
        void operator()(fvdg_ctx* c) const { fvdg_destroy(c); }
    };
    std::unique_ptr<fvdg_ctx, Del> ctx_;
    std::unique_ptr<MeshHandle> mesh_;
    const ComputingMesh* mesh_key_ = nullptr;
    size_t n_cells_ = 0, n_faces_ = 0;
    std::vector<BoundaryType> bc_;
    std::vector<int64_t> dir_faces_;
    int order_ = -1, flux_ = -1, precision_ = -1;
    double gamma_ = 0;
};

// Device index for new sessions: FVDG_DEVICE from the environment, else 0.
inline int default_device() {
    const char* s = std::getenv("FVDG_DEVICE");
    return s ? std::atoi(s) : 0;
}

}  // namespace fvdg
