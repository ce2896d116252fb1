// Host half of libfvdg.so: error reporting, reference-element tables, the mesh
// C ABI (box generation, periodic pairing, geometry export) and the analytic
// fields used for initial states and error norms.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <thread>

#include "Mesh/Mesh.h"
#include "fvdg_internal.hpp"

namespace fvdg {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

static const int (&combo_table())[24][3] {
    static int t[24][3];
    static bool init = false;
    if (!init) {
        int n = 0;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                for (int c = 0; c < 4; ++c)
                    if (a != b && a != c && b != c) {
                        t[n][0] = a;
                        t[n][1] = b;
                        t[n][2] = c;
                        ++n;
                    }
        init = true;
    }
    return t;
}

int combo_index(int a0, int a1, int a2) {
    const auto& t = combo_table();
    for (int i = 0; i < 24; ++i)
        if (t[i][0] == a0 && t[i][1] == a1 && t[i][2] == a2) return i;
    return -1;
}
void combo_vertices(int combo, int* a) {
    const auto& t = combo_table();
    for (int k = 0; k < 3; ++k) a[k] = t[combo][k];
}

static int ref_vertex_of(const vector3f& nc) {
    // natural coordinate of a reference vertex: 0 -> (0,0,0), 1 -> e_x, 2 -> e_y, 3 -> e_z
    if (nc[0] == 1) return 1;
    if (nc[1] == 1) return 2;
    if (nc[2] == 1) return 3;
    return 0;
}

int face_combo(const CompTriangleFace& F, int side) {
    const auto& nc = F.m_natural_coords[side];
    return combo_index(ref_vertex_of(nc[0]), ref_vertex_of(nc[1]), ref_vertex_of(nc[2]));
}

vector3f face_point_to_cell(int combo, double u, double v) {
    int a[3];
    combo_vertices(combo, a);
    vector3f nc[3];
    for (int k = 0; k < 3; ++k) {
        nc[k] = {0, 0, 0};
        if (a[k] > 0) nc[k][a[k] - 1] = 1;
    }
    return nc[0] * (1 - u - v) + nc[1] * u + nc[2] * v;
}

SchemeTables build_tables(int order) {
    SchemeTables T;
    T.order = order;
    T.N = dg_num_basis(order);
    const QuadRuleView vr = tet_rule_for_order(order);
    const QuadRuleView fr = tri_rule_for_order(order);
    T.Q = vr.num_points;
    T.Qf = fr.num_points;
    T.vq_pts.assign(vr.points, vr.points + 3 * T.Q);
    T.vq_w.assign(vr.weights, vr.weights + T.Q);
    T.fq_pts.assign(fr.points, fr.points + 2 * T.Qf);
    T.fq_w.assign(fr.weights, fr.weights + T.Qf);
    const int N = T.N;
    T.V.resize(T.Q * N);
    T.G.resize(T.Q * 3 * N);
    T.mass_hat.assign(N, 0.0);
    std::vector<double> val(N), grad(3 * N);
    for (int q = 0; q < T.Q; ++q) {
        const double* p = &T.vq_pts[3 * q];
        dg_basis_eval(order, p[0], p[1], p[2], val.data());
        dg_basis_grad(order, p[0], p[1], p[2], reinterpret_cast<double(*)[3]>(grad.data()));
        for (int k = 0; k < N; ++k) {
            T.V[q * N + k] = val[k];
            for (int d = 0; d < 3; ++d) T.G[(q * 3 + d) * N + k] = grad[3 * k + d];
        }
    }
    // mass diagonal in the reference driver's form: sum_q phi^2 * w (VortexBatch.cpp:157-162)
    for (int k = 0; k < N; ++k) {
        double s = 0;
        for (int q = 0; q < T.Q; ++q) s += T.V[q * N + k] * T.V[q * N + k] * T.vq_w[q];
        T.mass_hat[k] = s;
    }
    T.Phi.resize(24 * T.Qf * N);
    for (int c = 0; c < 24; ++c)
        for (int g = 0; g < T.Qf; ++g) {
            const vector3f xi = face_point_to_cell(c, T.fq_pts[2 * g], T.fq_pts[2 * g + 1]);
            dg_basis_eval(order, xi[0], xi[1], xi[2], val.data());
            for (int k = 0; k < N; ++k) T.Phi[(c * T.Qf + g) * N + k] = val[k];
        }
    return T;
}

// canonical orders: for each vertex subset {x<y<z} in lexicographic order: (x,y,z) then (x,z,y)
static void canonical_combos(int canon_combo[8]) {
    int n = 0;
    for (int x = 0; x < 4; ++x)
        for (int y = x + 1; y < 4; ++y)
            for (int z = y + 1; z < 4; ++z) {
                canon_combo[n++] = combo_index(x, y, z);
                canon_combo[n++] = combo_index(x, z, y);
            }
}

int canonical_of_combo(int combo) {
    int canon_combo[8];
    canonical_combos(canon_combo);
    int a[3];
    combo_vertices(combo, a);
    int m = 0;
    for (int i = 1; i < 3; ++i)
        if (a[i] < a[m]) m = i;
    const int cc = combo_index(a[m], a[(m + 1) % 3], a[(m + 2) % 3]);
    for (int i = 0; i < 8; ++i)
        if (canon_combo[i] == cc) return i * 3 + m;
    return -1;
}

double build_canonical_face_tables(const SchemeTables& T, std::vector<double>& PhiC, std::vector<uint8_t>& canon,
                                   std::vector<int32_t>& perm) {
    const int N = T.N, QF = T.Qf;
    int canon_combo[8];
    canonical_combos(canon_combo);
    PhiC.assign(8 * QF * N, 0.0);
    for (int c = 0; c < 8; ++c)
        for (int g = 0; g < QF; ++g)
            for (int k = 0; k < N; ++k) PhiC[(c * QF + g) * N + k] = T.Phi[((size_t)canon_combo[c] * QF + g) * N + k];
    // point permutations: rotation m maps barycentrics (l0,l1,l2) to (l_m, l_m+1, l_m+2)
    perm.assign(3 * QF, 0);
    double worst_match = 0.0;
    for (int m = 0; m < 3; ++m)
        for (int g = 0; g < QF; ++g) {
            const double u = T.fq_pts[2 * g], v = T.fq_pts[2 * g + 1];
            const double lam[3] = {1 - u - v, u, v};
            const double up = lam[(1 + m) % 3], vp = lam[(2 + m) % 3];
            int best = -1;
            double bd = 1e300;
            for (int h = 0; h < QF; ++h) {
                const double d = std::abs(T.fq_pts[2 * h] - up) + std::abs(T.fq_pts[2 * h + 1] - vp);
                if (d < bd) bd = d, best = h;
            }
            perm[m * QF + g] = best;
            worst_match = std::max(worst_match, bd);
        }
    if (worst_match > 1e-12) return INFINITY;  // the rule is not rotation invariant
    canon.assign(24, 0);
    double dev = 0.0;
    for (int c = 0; c < 24; ++c) {
        const int cm = canonical_of_combo(c), ci = cm / 3, m = cm % 3;
        canon[c] = (uint8_t)cm;
        for (int g = 0; g < QF; ++g)
            for (int k = 0; k < N; ++k)
                dev = std::max(dev, std::abs(T.Phi[((size_t)c * QF + g) * N + k] - PhiC[(ci * QF + perm[m * QF + g]) * N + k]));
    }
    return dev;
}

// ---------------------------------------------------------------- analytic fields
bool analytic_state(int field, const double* P, int np, double t, const vector3f& x, double U[5]) {
    if (field == FVDG_FIELD_VORTEX) {
        // isentropic vortex of examples/Vortex/problem.h:17-110 (rho_inf = p_inf = 1), translating
        // with (u_inf, v_inf); P = gamma, strength, x0, y0, u_inf, v_inf, Lx, Ly (L = 0: no wrap)
        if (np < 8) return false;
        const double g = P[0], phi = P[1];
        double dx = x[0] - (P[2] + P[4] * t), dy = x[1] - (P[3] + P[5] * t);
        if (P[6] > 0) dx -= P[6] * std::floor(dx / P[6] + 0.5);
        if (P[7] > 0) dy -= P[7] * std::floor(dy / P[7] + 0.5);
        const double r2 = dx * dx + dy * dy;
        const double slope = (phi * phi * (g - 1)) / (16 * M_PI * M_PI * g);
        const double rho = std::pow(1 - slope * std::exp(2 * (1 - r2)), 1.0 / (g - 1));
        const double u = P[4] - phi * dy / (2 * M_PI) * std::exp(1 - r2);
        const double v = P[5] + phi * dx / (2 * M_PI) * std::exp(1 - r2);
        const double p = std::pow(rho, g);
        U[0] = rho;
        U[1] = rho * u;
        U[2] = rho * v;
        U[3] = 0.0;
        U[4] = rho * (p / rho / (g - 1) + 0.5 * (u * u + v * v));
        return true;
    }
    if (field == FVDG_FIELD_TGV) {
        // compressible Taylor-Green vortex initial state (t ignored); P = gamma, rho0, V0, Mach
        if (np < 4) return false;
        const double g = P[0], rho0 = P[1], V0 = P[2], M = P[3];
        const double p0 = rho0 * V0 * V0 / (g * M * M);
        const double u = V0 * std::sin(x[0]) * std::cos(x[1]) * std::cos(x[2]);
        const double v = -V0 * std::cos(x[0]) * std::sin(x[1]) * std::cos(x[2]);
        const double p = p0 + rho0 * V0 * V0 / 16.0 * (std::cos(2 * x[0]) + std::cos(2 * x[1])) * (std::cos(2 * x[2]) + 2);
        U[0] = rho0;
        U[1] = rho0 * u;
        U[2] = rho0 * v;
        U[3] = 0.0;
        U[4] = p / (g - 1) + 0.5 * rho0 * (u * u + v * v);
        return true;
    }
    if (field == FVDG_FIELD_CONST) {
        if (np < 5) return false;
        for (int v = 0; v < 5; ++v) U[v] = P[v];
        return true;
    }
    return false;
}

template <typename Fn>
static void parallel_for(int64_t n, Fn fn) {
    int nt = (int)std::min<int64_t>(std::max(1u, std::thread::hardware_concurrency()), 32);
    if (n < 4096) nt = 1;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([=]() {
            for (int64_t i = (n * t) / nt; i < (n * (t + 1)) / nt; ++i) fn(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace fvdg

using namespace fvdg;

// ================================================================ C ABI (host part)
// Disjoint union of meshes (the problems of an ensemble): points, cells and faces of part i
// follow those of parts 0..i-1, every index shifted; boundary types are kept.
fvdg_mesh* fvdg::union_mesh(const std::vector<const fvdg_mesh*>& parts, std::vector<int64_t>& cell_off,
                            std::vector<int64_t>& face_off) {
    std::vector<vector3f> pts;
    std::vector<uInt> cn, fn, fc, fr;
    std::vector<BoundaryType> bc;
    cell_off.assign(1, 0);
    face_off.assign(1, 0);
    for (const fvdg_mesh* m : parts) {
        const auto& cm = m->cm;
        const uInt p0 = pts.size(), c0 = cell_off.back();
        pts.insert(pts.end(), cm.m_points.begin(), cm.m_points.end());
        for (const auto& c : cm.m_cells)
            for (int a = 0; a < 4; ++a) cn.push_back(c.m_nodes[a] + p0);
        for (const auto& F : cm.m_faces) {
            for (int k = 0; k < 3; ++k) {
                fn.push_back(F.m_nodes[k] + p0);
                fr.push_back(F.m_nodes_r[k] + p0);
            }
            fc.push_back(F.m_neighbor_cells[0] + c0);
            fc.push_back(F.is_boundary() ? uInt(-1) : F.m_neighbor_cells[1] + c0);
        }
        bc.insert(bc.end(), cm.m_boundaryTypes.begin(), cm.m_boundaryTypes.end());
        cell_off.push_back(c0 + cm.m_cells.size());
        face_off.push_back(face_off.back() + cm.m_faces.size());
    }
    auto* u = new fvdg_mesh{ComputingMesh::from_arrays(pts, cn, fn, fc, fr)};
    u->cm.m_boundaryTypes = bc;
    return u;
}

extern "C" {

const char* fvdg_version(void) { return "fvdg 0.1 (gfx950)"; }
const char* fvdg_last_error(void) { return g_last_error.c_str(); }
int fvdg_num_basis(int order) { return order < 0 ? 0 : dg_num_basis(order); }

int fvdg_basis_eval(int order, double x, double y, double z, double* val, double* grad) {
    if (order < 0 || order > 5) return fail(FVDG_ERR_ARG, "order must be 0..5");
    if (val) dg_basis_eval(order, x, y, z, val);
    if (grad) dg_basis_grad(order, x, y, z, reinterpret_cast<double(*)[3]>(grad));
    return dg_num_basis(order);
}

int fvdg_quadrature(int order, int* n_vol, double* vol_pts, double* vol_w, int* n_face, double* face_pts,
                    double* face_w) {
    if (order < 0 || order > 5) return fail(FVDG_ERR_ARG, "order must be 0..5");
    const QuadRuleView vr = tet_rule_for_order(order), fr = tri_rule_for_order(order);
    if (n_vol) *n_vol = vr.num_points;
    if (n_face) *n_face = fr.num_points;
    if (vol_pts) std::memcpy(vol_pts, vr.points, sizeof(double) * 3 * vr.num_points);
    if (vol_w) std::memcpy(vol_w, vr.weights, sizeof(double) * vr.num_points);
    if (face_pts) std::memcpy(face_pts, fr.points, sizeof(double) * 2 * fr.num_points);
    if (face_w) std::memcpy(face_w, fr.weights, sizeof(double) * fr.num_points);
    return FVDG_OK;
}

int fvdg_mesh_create_box(fvdg_mesh** out, const double lb[3], const double ub[3], const int64_t n[3], int split,
                         int periodic_mask, double jitter, uint64_t seed) {
    if (!out || !lb || !ub || !n) return fail(FVDG_ERR_ARG, "null argument");
    if (split != 5 && split != 6 && split != 15 && split != 16)
        return fail(FVDG_ERR_ARG, "split must be 5, 6 (periodic-capable) or 15, 16 (the reference's split)");
    if (split > 6 && (periodic_mask & 7))
        return fail(FVDG_ERR_ARG, "the reference's split (15, 16) has no periodic pairing: use split 5 or 6");
    for (int d = 0; d < 3; ++d) {
        if (n[d] < 1) return fail(FVDG_ERR_ARG, "n must be >= 1 along each axis");
        if (!(ub[d] > lb[d])) return fail(FVDG_ERR_ARG, "ub must exceed lb");
        if (split == 5 && ((periodic_mask >> d) & 1) && (n[d] & 1))
            return fail(FVDG_ERR_ARG, "5-tet split needs an even cell count along periodic axes");
    }
    if (jitter < 0 || jitter >= 0.25) return fail(FVDG_ERR_ARG, "jitter must be in [0, 0.25)");
    try {
        OrthHexMesh g({lb[0], lb[1], lb[2]}, {ub[0], ub[1], ub[2]}, {(uInt)n[0], (uInt)n[1], (uInt)n[2]});
        if (split == 6)
            g.split_hex6_periodic();
        else if (split == 5)
            g.split_hex5_periodic();
        else {
            // the reference's sequence (OrthHexMesh, split_hex*_scan, rebuild_cell_topology)
            if (split == 16)
                g.split_hex6_scan();
            else
                g.split_hex5_scan();
            g.rebuild_cell_topology();
        }
        const std::array<bool, 3> per = {(periodic_mask & 1) != 0, (periodic_mask & 2) != 0, (periodic_mask & 4) != 0};
        if (periodic_mask & 7) {
            const uInt merged = g.make_periodic({lb[0], lb[1], lb[2]}, {ub[0], ub[1], ub[2]}, per);
            uInt expect = 0;
            if (per[0]) expect += 2 * n[1] * n[2];
            if (per[1]) expect += 2 * n[0] * n[2];
            if (per[2]) expect += 2 * n[0] * n[1];
            if (merged != expect) return fail(FVDG_ERR_MESH, "periodic pairing matched an unexpected number of faces");
        }
        const vector3f h = {(ub[0] - lb[0]) / n[0], (ub[1] - lb[1]) / n[1], (ub[2] - lb[2]) / n[2]};
        g.jitter_points(jitter, seed, {lb[0], lb[1], lb[2]}, {ub[0], ub[1], ub[2]}, h, per);
        g.validate_mesh();
        auto* m = new fvdg_mesh{ComputingMesh(g)};
        *out = m;
        return FVDG_OK;
    } catch (const std::exception& e) {
        return fail(FVDG_ERR_MESH, e.what());
    }
}

int fvdg_mesh_create(fvdg_mesh** out, int64_t n_points, const double* points, int64_t n_cells,
                     const int64_t* cell_nodes, int64_t n_faces, const int64_t* face_nodes,
                     const int64_t* face_cells, const int64_t* face_nodes_r) {
    if (!out || n_points <= 0 || n_cells <= 0 || n_faces <= 0 || !points || !cell_nodes || !face_nodes || !face_cells)
        return fail(FVDG_ERR_ARG, "null or empty mesh arrays");
    try {
        std::vector<vector3f> pts(n_points);
        for (int64_t p = 0; p < n_points; ++p) pts[p] = {points[3 * p], points[3 * p + 1], points[3 * p + 2]};
        auto conv = [](const int64_t* a, int64_t n) {
            std::vector<uInt> v(n);
            for (int64_t i = 0; i < n; ++i) v[i] = a[i] < 0 ? uInt(-1) : (uInt)a[i];
            return v;
        };
        std::vector<uInt> fr;
        if (face_nodes_r) fr = conv(face_nodes_r, 3 * n_faces);
        auto* m = new fvdg_mesh{ComputingMesh::from_arrays(pts, conv(cell_nodes, 4 * n_cells), conv(face_nodes, 3 * n_faces),
                                                           conv(face_cells, 2 * n_faces), fr)};
        for (const auto& c : m->cm.m_cells)
            if (!(c.m_volume > 0)) {
                delete m;
                return fail(FVDG_ERR_MESH, "degenerate tetrahedron");
            }
        *out = m;
        return FVDG_OK;
    } catch (const std::exception& e) {
        return fail(FVDG_ERR_MESH, e.what());
    }
}

int fvdg_mesh_read_fluent(fvdg_mesh** out, const char* path) {
    if (!out || !path) return fail(FVDG_ERR_ARG, "null argument");
    try {
        FluentMesh fm;
        fm.read(path);
        std::vector<int32_t> fz;
        GeneralMesh g = fm.to_tets(fz);
        auto* m = new fvdg_mesh{ComputingMesh(g), fz, {}};
        for (const auto& kv : fm.zones) m->zones.push_back(kv.second);
        std::map<int, int> bc_of;
        for (const auto& z : m->zones) bc_of[z.id] = z.bc_type;
        for (size_t f = 0; f < m->cm.m_faces.size(); ++f)
            if (m->cm.m_faces[f].is_boundary())
                m->cm.m_boundaryTypes[f] = fz[f] >= 0 ? FluentMesh::boundary_type_of(bc_of[fz[f]]) : BoundaryType::Wall;
        for (const auto& c : m->cm.m_cells)
            if (!(c.m_volume > 0)) {
                delete m;
                return fail(FVDG_ERR_MESH, "degenerate tetrahedron after splitting");
            }
        *out = m;
        return FVDG_OK;
    } catch (const std::exception& e) {
        return fail(FVDG_ERR_MESH, e.what());
    }
}

int fvdg_mesh_face_zones(const fvdg_mesh* m, int32_t* zone) {
    if (!m || !zone) return fail(FVDG_ERR_ARG, "null argument");
    for (size_t f = 0; f < m->cm.m_faces.size(); ++f) zone[f] = f < m->face_zone.size() ? m->face_zone[f] : -1;
    return FVDG_OK;
}

int fvdg_mesh_zones(const fvdg_mesh* m, int* n, int32_t* ids, int32_t* bc_types, char* names, int name_len) {
    if (!m || !n) return fail(FVDG_ERR_ARG, "null argument");
    *n = (int)m->zones.size();
    for (size_t i = 0; i < m->zones.size(); ++i) {
        if (ids) ids[i] = m->zones[i].id;
        if (bc_types) bc_types[i] = m->zones[i].bc_type;
        if (names && name_len > 0) {
            std::string s = m->zones[i].name;
            if (s.empty()) s = m->zones[i].kind;
            s = s.substr(0, (size_t)name_len - 1);
            std::memcpy(names + i * name_len, s.c_str(), s.size() + 1);
        }
    }
    return FVDG_OK;
}

int fvdg_mesh_set_zone_boundary_type(fvdg_mesh* m, int32_t zone, uint8_t bc) {
    if (!m) return fail(FVDG_ERR_ARG, "null mesh");
    if (bc >= (uint8_t)BoundaryType::COUNT) return fail(FVDG_ERR_ARG, "unknown boundary type");
    int64_t hit = 0;
    for (size_t f = 0; f < m->cm.m_faces.size() && f < m->face_zone.size(); ++f)
        if (m->face_zone[f] == zone && m->cm.m_faces[f].is_boundary()) {
            m->cm.m_boundaryTypes[f] = (BoundaryType)bc;
            ++hit;
        }
    if (hit == 0) return fail(FVDG_ERR_ARG, "no boundary face in that zone");
    return FVDG_OK;
}

int fvdg_mesh_sizes(const fvdg_mesh* m, int64_t* np, int64_t* nc, int64_t* nf) {
    if (!m) return fail(FVDG_ERR_ARG, "null mesh");
    if (np) *np = m->cm.m_points.size();
    if (nc) *nc = m->cm.m_cells.size();
    if (nf) *nf = m->cm.m_faces.size();
    return FVDG_OK;
}

int fvdg_mesh_export(const fvdg_mesh* m, double* points, int64_t* cell_nodes, int64_t* face_nodes,
                     int64_t* face_cells, int64_t* face_nodes_r) {
    if (!m) return fail(FVDG_ERR_ARG, "null mesh");
    const auto& cm = m->cm;
    auto I = [](uInt v) { return v == uInt(-1) ? int64_t(-1) : int64_t(v); };
    if (points)
        for (size_t p = 0; p < cm.m_points.size(); ++p)
            for (int d = 0; d < 3; ++d) points[3 * p + d] = cm.m_points[p][d];
    if (cell_nodes)
        for (size_t c = 0; c < cm.m_cells.size(); ++c)
            for (int a = 0; a < 4; ++a) cell_nodes[4 * c + a] = I(cm.m_cells[c].m_nodes[a]);
    for (size_t f = 0; f < cm.m_faces.size(); ++f) {
        const auto& F = cm.m_faces[f];
        for (int k = 0; k < 3; ++k) {
            if (face_nodes) face_nodes[3 * f + k] = I(F.m_nodes[k]);
            if (face_nodes_r) face_nodes_r[3 * f + k] = I(F.m_nodes_r[k]);
        }
        if (face_cells) {
            face_cells[2 * f] = I(F.m_neighbor_cells[0]);
            face_cells[2 * f + 1] = I(F.m_neighbor_cells[1]);
        }
    }
    return FVDG_OK;
}

int fvdg_mesh_geometry(const fvdg_mesh* m, double* inv_jac, double* det, double* h, double* normal, double* area) {
    if (!m) return fail(FVDG_ERR_ARG, "null mesh");
    const auto& cm = m->cm;
    for (size_t c = 0; c < cm.m_cells.size(); ++c) {
        const auto& t = cm.m_cells[c];
        if (inv_jac)
            for (int i = 0; i < 9; ++i) inv_jac[9 * c + i] = t.m_invJac[i];
        if (det) det[c] = t.compute_jacobian_det();
        if (h) h[c] = t.m_h;
    }
    for (size_t f = 0; f < cm.m_faces.size(); ++f) {
        if (normal)
            for (int d = 0; d < 3; ++d) normal[3 * f + d] = cm.m_faces[f].m_normal[d];
        if (area) area[f] = cm.m_faces[f].m_area;
    }
    return FVDG_OK;
}

int fvdg_mesh_set_boundary_types(fvdg_mesh* m, const uint8_t* bc) {
    if (!m || !bc) return fail(FVDG_ERR_ARG, "null argument");
    for (size_t f = 0; f < m->cm.m_faces.size(); ++f) {
        if (bc[f] >= (uint8_t)BoundaryType::COUNT) return fail(FVDG_ERR_ARG, "boundary type out of range");
        m->cm.m_boundaryTypes[f] = (BoundaryType)bc[f];
    }
    return FVDG_OK;
}

int fvdg_mesh_get_boundary_types(const fvdg_mesh* m, uint8_t* bc) {
    if (!m || !bc) return fail(FVDG_ERR_ARG, "null argument");
    for (size_t f = 0; f < m->cm.m_faces.size(); ++f) bc[f] = (uint8_t)m->cm.m_boundaryTypes[f];
    return FVDG_OK;
}

int fvdg_mesh_tag_boundaries_pseudo3dz(fvdg_mesh* m) {
    if (!m) return fail(FVDG_ERR_ARG, "null mesh");
    for (size_t f = 0; f < m->cm.m_faces.size(); ++f) {
        const auto& F = m->cm.m_faces[f];
        if (!F.is_boundary()) continue;
        m->cm.m_boundaryTypes[f] = std::abs(F.m_normal[2]) > 0.5 ? BoundaryType::Pseudo3DZ : BoundaryType::Dirichlet;
    }
    return FVDG_OK;
}

void fvdg_mesh_destroy(fvdg_mesh* m) { delete m; }

int fvdg_project_analytic(const fvdg_mesh* m, int order, int field, const double* params, int n_params, double t,
                          double* U) {
    if (!m || !U) return fail(FVDG_ERR_ARG, "null argument");
    if (order < 0 || order > 5) return fail(FVDG_ERR_ARG, "order must be 0..5");
    double probe[5];
    if (!analytic_state(field, params, n_params, t, {0, 0, 0}, probe)) return fail(FVDG_ERR_ARG, "unknown field or too few parameters");
    const SchemeTables T = build_tables(order);
    const int N = T.N, Q = T.Q;
    // diagonal of the reference mass in long double (DG_Basis_Evaluator.h:48-56)
    std::vector<long double> diag(N, 0.0L);
    std::vector<double> phiw(Q * N);
    for (int k = 0; k < N; ++k)
        for (int q = 0; q < Q; ++q) {
            long double phi[64];
            fvdg_basis_detail::eval_dubiner<long double>(order, T.vq_pts[3 * q], T.vq_pts[3 * q + 1], T.vq_pts[3 * q + 2],
                                                         phi, nullptr);
            diag[k] += phi[k] * phi[k] * (long double)T.vq_w[q];
            phiw[q * N + k] = (double)(phi[k] * (long double)T.vq_w[q]);
        }
    const auto& cm = m->cm;
    parallel_for((int64_t)cm.m_cells.size(), [&](int64_t e) {
        const auto& cell = cm.m_cells[e];
        double f[128][5];
        for (int q = 0; q < Q; ++q) {
            const vector3f x = cell.transform_to_physical(vector3f{T.vq_pts[3 * q], T.vq_pts[3 * q + 1], T.vq_pts[3 * q + 2]});
            analytic_state(field, params, n_params, t, x, f[q]);
        }
        double* out = U + e * 5 * N;
        for (int k = 0; k < N; ++k)
            for (int v = 0; v < 5; ++v) {
                double rhs = 0;
                for (int q = 0; q < Q; ++q) rhs += f[q][v] * phiw[q * N + k];
                out[5 * k + v] = rhs / (double)diag[k];
            }
    });
    return FVDG_OK;
}

int fvdg_solution_write(const fvdg_mesh* m, int order, const double* U, int field, const double* params,
                        int n_params, double t, const char* path) {
    if (!m || !U || !path) return fail(FVDG_ERR_ARG, "null argument");
    if (order < 0 || order > 5) return fail(FVDG_ERR_ARG, "order must be 0..5");
    double probe[5];
    if (field >= 0 && !analytic_state(field, params, n_params, t, {0, 0, 0}, probe))
        return fail(FVDG_ERR_ARG, "unknown field or too few parameters");
    std::ofstream fp(path);
    if (!fp) return fail(FVDG_ERR_ARG, std::string("cannot write ") + path);
    const SchemeTables T = build_tables(order);
    const int N = T.N, Q = T.Q;
    const auto& cm = m->cm;
    // one line per (cell, volume point): x y z, then (U_h, U_s) per variable, 16 digits
    // (examples/Vortex/VortexBatch3.cpp:402-415)
    for (size_t e = 0; e < cm.m_cells.size(); ++e) {
        const auto& cell = cm.m_cells[e];
        const double* c = U + e * 5 * N;
        for (int q = 0; q < Q; ++q) {
            const vector3f xi = {T.vq_pts[3 * q], T.vq_pts[3 * q + 1], T.vq_pts[3 * q + 2]};
            const vector3f pos = cell.transform_to_physical(xi);
            double uh[5] = {0, 0, 0, 0, 0}, us[5] = {0, 0, 0, 0, 0};
            for (int k = 0; k < N; ++k)
                for (int v = 0; v < 5; ++v) uh[v] += T.V[q * N + k] * c[5 * k + v];
            if (field >= 0) analytic_state(field, params, n_params, t, pos, us);
            fp << std::setprecision(16) << pos[0] << "  " << pos[1] << "  " << pos[2];
            for (int v = 0; v < 5; ++v) fp << "  " << uh[v] << "  " << us[v];
            fp << "\n";
        }
    }
    fp.close();
    if (!fp) return fail(FVDG_ERR_ARG, std::string("write failed: ") + path);
    return FVDG_OK;
}

int fvdg_solution_read(const fvdg_mesh* m, int order, const char* path, double* U) {
    if (!m || !U || !path) return fail(FVDG_ERR_ARG, "null argument");
    if (order < 0 || order > 5) return fail(FVDG_ERR_ARG, "order must be 0..5");
    std::ifstream in(path);
    if (!in) return fail(FVDG_ERR_ARG, std::string("cannot open ") + path);
    const SchemeTables T = build_tables(order);
    const int N = T.N, Q = T.Q;
    const size_t E = m->cm.m_cells.size();
    // U_h at the volume points (x y z, then U_h U_s pairs), then per mode the diagonal
    // projection sum_q U_h phi_k w / sum_q phi_k^2 w (examples/Vortex/tools.h:72-129)
    std::vector<double> uh((size_t)E * Q * 5);
    std::string line;
    size_t row = 0;
    while (row < E * Q && std::getline(in, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream iss(line);
        double x, y, z, a, b;
        if (!(iss >> x >> y >> z)) return fail(FVDG_ERR_ARG, "malformed solution line " + std::to_string(row + 1));
        for (int v = 0; v < 5; ++v) {
            if (!(iss >> a >> b)) return fail(FVDG_ERR_ARG, "malformed solution line " + std::to_string(row + 1));
            uh[row * 5 + v] = a;
        }
        ++row;
    }
    if (row != E * Q) return fail(FVDG_ERR_ARG, "solution file has fewer lines than cells x volume points");
    std::vector<double> den(N, 0.0);
    for (int k = 0; k < N; ++k)
        for (int q = 0; q < Q; ++q) den[k] += T.V[q * N + k] * T.V[q * N + k] * T.vq_w[q];
    parallel_for(64, [&](int64_t ch) {
        for (size_t e = E * ch / 64; e < E * (ch + 1) / 64; ++e)
            for (int k = 0; k < N; ++k)
                for (int v = 0; v < 5; ++v) {
                    double s = 0.0;
                    for (int q = 0; q < Q; ++q) s += uh[(e * Q + q) * 5 + v] * T.V[q * N + k] * T.vq_w[q];
                    U[e * 5 * N + 5 * k + v] = s / den[k];
                }
    });
    return FVDG_OK;
}

// ---------------------------------------------------------------- binary checkpoints
// Header of 56 bytes, then the raw coefficients: magic "FVDGCKP1", u32 version, i32 order, i64 cells,
// i64 values (cells * 5N), f64 time, i64 step, u64 FNV-1a hash of the coefficient bytes.
namespace {
constexpr char kCkptMagic[8] = {'F', 'V', 'D', 'G', 'C', 'K', 'P', '1'};
struct CkptHeader {
    char magic[8];
    uint32_t version;
    int32_t order;
    int64_t cells, values;
    double time;
    int64_t step;
    uint64_t hash;
};
static_assert(sizeof(CkptHeader) == 56, "checkpoint header layout");
uint64_t fnv1a(const unsigned char* p, size_t n) {
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
    return h;
}
}  // namespace

int fvdg_checkpoint_write(const char* path, int order, int64_t n_cells, const double* U, double t, int64_t step) {
    if (!path || !U) return fail(FVDG_ERR_ARG, "null argument");
    if (order < 0 || order > 5 || n_cells <= 0) return fail(FVDG_ERR_ARG, "order must be 0..5 and cells positive");
    const int64_t nv = n_cells * 5 * ((order + 1) * (order + 2) * (order + 3) / 6);
    CkptHeader h;
    std::memcpy(h.magic, kCkptMagic, 8);
    h.version = 1;
    h.order = order;
    h.cells = n_cells;
    h.values = nv;
    h.time = t;
    h.step = step;
    h.hash = fnv1a(reinterpret_cast<const unsigned char*>(U), (size_t)nv * sizeof(double));
    // written to a temporary file and renamed: a crash mid-write leaves the previous checkpoint
    const std::string tmp = std::string(path) + ".tmp";
    {
        std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
        if (!out) return fail(FVDG_ERR_ARG, "cannot open " + tmp);
        out.write(reinterpret_cast<const char*>(&h), sizeof h);
        out.write(reinterpret_cast<const char*>(U), (std::streamsize)(nv * sizeof(double)));
        out.close();
        if (!out) return fail(FVDG_ERR_ARG, "write failed: " + tmp);
    }
    if (std::rename(tmp.c_str(), path) != 0) return fail(FVDG_ERR_ARG, std::string("cannot rename to ") + path);
    return FVDG_OK;
}

int fvdg_checkpoint_read(const char* path, int* order, int64_t* n_cells, double* U, double* t, int64_t* step) {
    if (!path) return fail(FVDG_ERR_ARG, "null path");
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(FVDG_ERR_ARG, std::string("cannot open ") + path);
    CkptHeader h;
    if (!in.read(reinterpret_cast<char*>(&h), sizeof h)) return fail(FVDG_ERR_STATE, "truncated checkpoint header");
    if (std::memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != 1)
        return fail(FVDG_ERR_ARG, "not a version-1 fvdg checkpoint");
    if (h.order < 0 || h.order > 5 || h.cells <= 0 ||
        h.values != h.cells * 5 * ((h.order + 1) * (h.order + 2) * (h.order + 3) / 6))
        return fail(FVDG_ERR_STATE, "inconsistent checkpoint header");
    if (order) *order = h.order;
    if (n_cells) *n_cells = h.cells;
    if (t) *t = h.time;
    if (step) *step = h.step;
    if (!U) return FVDG_OK;
    if (!in.read(reinterpret_cast<char*>(U), (std::streamsize)(h.values * sizeof(double))))
        return fail(FVDG_ERR_STATE, "truncated checkpoint data");
    if (fnv1a(reinterpret_cast<const unsigned char*>(U), (size_t)h.values * sizeof(double)) != h.hash)
        return fail(FVDG_ERR_STATE, "checkpoint data does not match its checksum");
    return FVDG_OK;
}

int fvdg_analytic_error(const fvdg_mesh* m, int order, int field, const double* params, int n_params, double t,
                        const double* U, double err[5]) {
    if (!m || !U || !err) return fail(FVDG_ERR_ARG, "null argument");
    if (order < 0 || order > 5) return fail(FVDG_ERR_ARG, "order must be 0..5");
    double probe[5];
    if (!analytic_state(field, params, n_params, t, {0, 0, 0}, probe)) return fail(FVDG_ERR_ARG, "unknown field or too few parameters");
    const SchemeTables T = build_tables(order);
    const int N = T.N, Q = T.Q;
    const auto& cm = m->cm;
    const int64_t E = cm.m_cells.size();
    const int nchunk = 64;
    std::vector<double> num(nchunk * 5, 0.0), den(nchunk * 5, 0.0);
    parallel_for(nchunk, [&](int64_t ch) {
        for (int64_t e = E * ch / nchunk; e < E * (ch + 1) / nchunk; ++e) {
            const auto& cell = cm.m_cells[e];
            const double* c = U + e * 5 * N;
            for (int q = 0; q < Q; ++q) {
                const vector3f xi = {T.vq_pts[3 * q], T.vq_pts[3 * q + 1], T.vq_pts[3 * q + 2]};
                double us[5], uh[5] = {0, 0, 0, 0, 0};
                analytic_state(field, params, n_params, t, cell.transform_to_physical(xi), us);
                for (int k = 0; k < N; ++k)
                    for (int v = 0; v < 5; ++v) uh[v] += T.V[q * N + k] * c[5 * k + v];
                const double w = T.vq_w[q] * cell.compute_jacobian_det();
                for (int v = 0; v < 5; ++v) {
                    num[ch * 5 + v] += (uh[v] - us[v]) * (uh[v] - us[v]) * w;
                    den[ch * 5 + v] += us[v] * us[v] * w;
                }
            }
        }
    });
    for (int v = 0; v < 5; ++v) {
        double a = 0, b = 1e-47;
        for (int ch = 0; ch < nchunk; ++ch) {
            a += num[ch * 5 + v];
            b += den[ch * 5 + v];
        }
        err[v] = std::sqrt(a / b);
    }
    return FVDG_OK;
}

}  // extern "C"
