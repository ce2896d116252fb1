// Internal declarations shared by the host part (fvdg_host.cpp) and the device
// part (fvdg_solver.hip) of libfvdg.so.  Not installed; the public ABI is include/fvdg.h.
#pragma once
#include <string>
#include <vector>
#include "DG/DG_Basis/DG_Basis.h"
#include "GaussQuadrature/GaussLegendre.h"
#include "Mesh/ComputingMesh.h"
#include "Mesh/FluentMesh.h"
#include "fvdg.h"

struct fvdg_mesh {
    ComputingMesh cm;
    // meshes read from a file: per face the file's zone (-1: none) and the zones
    std::vector<int32_t> face_zone;
    std::vector<FluentZone> zones;
};

namespace fvdg {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

constexpr int kMaxOrder = 5;  // device kernels are instantiated for orders 0..5

// The 24 ordered triples (a0,a1,a2) of distinct reference-tet vertices, in
// lexicographic order.  A face seen from one of its cells is one such triple:
// a_k is the cell-local vertex of face node k.
int combo_index(int a0, int a1, int a2);
void combo_vertices(int combo, int* a);
// combo of face `F` seen from side `side` (0 or 1), from its natural coordinates
int face_combo(const CompTriangleFace& F, int side);

// Reference-element tables of one order, row-major, FP64.
struct SchemeTables {
    int order = 0, N = 0, Q = 0, Qf = 0;
    std::vector<double> vq_pts, vq_w;  // Q*3, Q
    std::vector<double> fq_pts, fq_w;  // Qf*2, Qf
    std::vector<double> V;             // Q*N        phi_k(xi_q)
    std::vector<double> G;             // Q*3*N      d phi_k / d xi_d at xi_q, index (q*3+d)*N+k
    std::vector<double> Phi;           // 24*Qf*N    phi_k(xi(combo, uv_g)), index (c*Qf+g)*N+k
    std::vector<double> mass_hat;      // N          sum_q w_q phi_k(xi_q)^2
};
SchemeTables build_tables(int order);

// Face tables reduced by the rotation symmetry of the face rule: 8 canonical
// vertex orders (starting at the smallest reference vertex), per combo its
// canonical index*3 + rotation, and the 3 point permutations of the rule.
// Returns the largest deviation |Phi[c][g][k] - PhiC[canon][perm[rot][g]][k]|.
// canonical index * 3 + rotation of a combo (canonical index ci: vertex subset ci/2 in
// lexicographic order {0,1,2},{0,1,3},{0,2,3},{1,2,3}, i.e. the face opposite vertex
// 3 - ci/2; ci%2 = 0 ascending order, 1 the other rotation class)
int canonical_of_combo(int combo);
double build_canonical_face_tables(const SchemeTables& T, std::vector<double>& PhiC, std::vector<uint8_t>& canon,
                                   std::vector<int32_t>& perm);

// Map face point g of a face seen through `combo` to reference coordinates, with
// the reference's operation order nc0*(1-u-v) + nc1*u + nc2*v.
vector3f face_point_to_cell(int combo, double u, double v);

// Disjoint union of meshes; cell_off / face_off get the n+1 prefix offsets of the parts.
fvdg_mesh* union_mesh(const std::vector<const fvdg_mesh*>& parts, std::vector<int64_t>& cell_off,
                      std::vector<int64_t>& face_off);

// Analytic fields (host): conserved state at a physical point.
bool analytic_state(int field, const double* params, int n_params, double t, const vector3f& x, double U[5]);

}  // namespace fvdg
