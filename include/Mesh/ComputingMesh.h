// Geometric tetrahedral mesh consumed by the DG scheme.
//
// Class names, members and conventions follow the reference's
// include/Mesh/ComputingMesh.h:17-109 and src/Mesh/ComputingMesh.cpp:27-263:
//   * tet: A = rows (p1-p0, p2-p0, p3-p0) ("compute_jacobian_mat", the transpose
//     of dx/dxi), m_invJac = A^{-1} by cofactors, m_volume = |det A|/6,
//     compute_jacobian_det() = 6*m_volume, physical gradient = A^{-1} * reference gradient;
//   * face: normal = unit(cross(p1-p0, p2-p0)) oriented from cell 0 towards cell 1
//     (interior: by the centroid difference; boundary: away from cell 0's fourth
//     vertex), m_area = |cross|/2, compute_jacobian_det() = 2*m_area;
//   * m_natural_coords[side][k] = reference-tet vertex (0, e_x, e_y or e_z) of
//     face node k inside cell `side`, so that a face point (u,v) maps to
//     xi = nc[0](1-u-v) + nc[1]u + nc[2]v in either neighbour;
//   * m_h = volume / (sum of the 4 face areas).
// Differences: elements store their own vertex coordinates instead of a
// back-reference to the parent mesh (so the mesh can be moved and returned
// by value), faces carry m_nodes_r (the node numbering seen from cell 1, which
// differs across a periodic boundary), and a mesh can be built from raw arrays.
#pragma once
#include <cstring>
#include "Matrix/DenseMatrix.h"
#include "Mesh/GeneralMesh.h"

enum class BoundaryType : uint8_t {
    Dirichlet,
    Neumann,
    Robin,
    Pseudo3DX,
    Pseudo3DY,
    Pseudo3DZ,
    Inflow,
    Outflow,
    Wall,
    WallTD,
    WallTN,
    WallTR,
    COUNT
};

class CompTriangleFace {
public:
    vector3u m_nodes;
    vector2u m_neighbor_cells;
    vector3f m_normal;
    Scalar m_area = 0;
    std::array<std::array<vector3f, 3>, 2> m_natural_coords;
    vector3u m_nodes_r;                  // m_nodes as numbered in cell 1
    std::array<vector3f, 3> m_coords;    // positions of m_nodes

    vector3f compute_physical_coord(const vector2f& uv) const {
        const Scalar u = uv[0], v = uv[1];
        return m_coords[0] * (1 - u - v) + m_coords[1] * u + m_coords[2] * v;
    }
    Scalar compute_jacobian_det() const { return m_area * 2.0; }
    bool is_boundary() const { return m_neighbor_cells[1] == uInt(-1); }
    bool is_periodic() const { return !is_boundary() && m_nodes_r != m_nodes; }
};

class CompTetrahedron {
public:
    vector4u m_nodes;
    vector3f m_centroid;
    vector4u m_faces;      // m_faces[l] is the face opposite m_nodes[l]
    vector4u m_neighbors;  // cell across m_faces[l], uInt(-1) on a boundary
    Scalar m_volume = 0;
    Scalar m_h = 0;
    DenseMatrix<3, 3> m_JacMat;
    DenseMatrix<3, 3> m_invJac;
    std::array<vector3f, 4> m_coords;  // positions of m_nodes

    void compute_geometry() {
        const auto &p0 = m_coords[0], &p1 = m_coords[1], &p2 = m_coords[2], &p3 = m_coords[3];
        m_centroid = (p0 + p1 + p2 + p3) / 4.0;
        m_volume = std::abs(vec_dot(p1 - p0, vec_cross(p2 - p0, p3 - p0))) / 6.0;
        m_JacMat = DenseMatrix<3, 3>(compute_jacobian_mat()).transpose();
        m_invJac = DenseMatrix<3, 3>(compute_jacobian_mat()).inverse();
    }
    std::array<vector3f, 3> compute_jacobian_mat() const {
        return {m_coords[1] - m_coords[0], m_coords[2] - m_coords[0], m_coords[3] - m_coords[0]};
    }
    Scalar compute_jacobian_det() const { return m_volume * 6.0; }
    DenseMatrix<3, 3> get_JacMat() const { return m_JacMat; }
    DenseMatrix<3, 3> get_invJacMat() const { return m_invJac; }
    vector3f transform_to_physical(const vector3f& xi) const {
        return m_coords[0] * (1 - xi[0] - xi[1] - xi[2]) + m_coords[1] * xi[0] + m_coords[2] * xi[1] +
               m_coords[3] * xi[2];
    }
    vector3f transform_to_physical(const vector4f& lam) const {
        return m_coords[0] * lam[0] + m_coords[1] * lam[1] + m_coords[2] * lam[2] + m_coords[3] * lam[3];
    }
};

class ComputingMesh {
public:
    std::vector<vector3f> m_points;
    std::vector<CompTriangleFace> m_faces;
    std::vector<CompTetrahedron> m_cells;
    std::vector<BoundaryType> m_boundaryTypes;

    ComputingMesh() = default;

    // From a pure tet GeneralMesh (after split_hex*_scan / make_periodic / jitter_points).
    explicit ComputingMesh(const GeneralMesh& g) {
        const uInt nc = g.m_cells.size(), nf = g.m_faces.size();
        std::vector<uInt> cn(4 * nc), fn(3 * nf), fc(2 * nf), fr(3 * nf);
        for (uInt c = 0; c < nc; ++c) {
            const auto* t = std::get_if<Tetrahedron>(&g.m_cells[c]);
            if (!t) throw std::runtime_error("ComputingMesh: every cell must be a tetrahedron");
            for (int a = 0; a < 4; ++a) cn[4 * c + a] = t->nodes[a];
        }
        for (uInt f = 0; f < nf; ++f) {
            const auto* t = std::get_if<TriangleFace>(&g.m_faces[f]);
            if (!t) throw std::runtime_error("ComputingMesh: every face must be a triangle");
            for (int k = 0; k < 3; ++k) {
                fn[3 * f + k] = t->nodes[k];
                fr[3 * f + k] = t->side_nodes(1)[k];
            }
            fc[2 * f] = t->neighbor_cells[0];
            fc[2 * f + 1] = t->neighbor_cells[1];
        }
        build(g.m_points, cn, fn, fc, fr);
    }

    // From raw arrays: cell_nodes[4*nc], face_nodes[3*nf] (side-0 numbering),
    // face_cells[2*nf] (uInt(-1) = boundary), face_nodes_r[3*nf] (side-1
    // numbering; empty = same as face_nodes).  Cell -> face links are derived.
    static ComputingMesh from_arrays(const std::vector<vector3f>& points, const std::vector<uInt>& cell_nodes,
                                     const std::vector<uInt>& face_nodes, const std::vector<uInt>& face_cells,
                                     const std::vector<uInt>& face_nodes_r = {}) {
        ComputingMesh m;
        m.build(points, cell_nodes, face_nodes, face_cells, face_nodes_r.empty() ? face_nodes : face_nodes_r);
        return m;
    }

private:
    void build(const std::vector<vector3f>& pts, const std::vector<uInt>& cn, const std::vector<uInt>& fn,
               const std::vector<uInt>& fc, const std::vector<uInt>& fr) {
        const uInt nc = cn.size() / 4, nf = fn.size() / 3, np = pts.size();
        if (cn.size() != 4 * nc || fn.size() != 3 * nf || fc.size() != 2 * nf || fr.size() != 3 * nf)
            throw std::runtime_error("ComputingMesh: inconsistent array sizes");
        m_points = pts;
        m_cells.resize(nc);
        for (uInt c = 0; c < nc; ++c) {
            auto& t = m_cells[c];
            for (int a = 0; a < 4; ++a) {
                t.m_nodes[a] = cn[4 * c + a];
                if (t.m_nodes[a] >= np) throw std::runtime_error("ComputingMesh: cell node out of range");
                t.m_coords[a] = pts[t.m_nodes[a]];
                t.m_faces[a] = uInt(-1);
                t.m_neighbors[a] = uInt(-1);
            }
            t.compute_geometry();
        }
        m_faces.resize(nf);
        for (uInt f = 0; f < nf; ++f) {
            auto& F = m_faces[f];
            for (int k = 0; k < 3; ++k) {
                F.m_nodes[k] = fn[3 * f + k];
                F.m_nodes_r[k] = fr[3 * f + k];
                if (F.m_nodes[k] >= np || F.m_nodes_r[k] >= np) throw std::runtime_error("ComputingMesh: face node out of range");
                F.m_coords[k] = pts[F.m_nodes[k]];
            }
            F.m_neighbor_cells = {fc[2 * f], fc[2 * f + 1]};
            if (F.m_neighbor_cells[0] >= nc) throw std::runtime_error("ComputingMesh: face without a side-0 cell");
            if (F.m_neighbor_cells[1] != uInt(-1) && F.m_neighbor_cells[1] >= nc)
                throw std::runtime_error("ComputingMesh: face cell out of range");
            face_geometry(F);
            // cell -> face links: the local face is the one whose opposite vertex is not on the face
            for (int side = 0; side < 2; ++side) {
                const uInt c = F.m_neighbor_cells[side];
                if (c == uInt(-1)) continue;
                const vector3u& nodes = side ? F.m_nodes_r : F.m_nodes;
                auto& t = m_cells[c];
                int opp = -1, hits = 0;
                for (int a = 0; a < 4; ++a) {
                    const bool on = t.m_nodes[a] == nodes[0] || t.m_nodes[a] == nodes[1] || t.m_nodes[a] == nodes[2];
                    hits += on;
                    if (!on) opp = a;
                }
                if (hits != 3 || opp < 0) throw std::runtime_error("ComputingMesh: face nodes are not a face of its cell");
                if (t.m_faces[opp] != uInt(-1)) throw std::runtime_error("ComputingMesh: cell face listed twice");
                t.m_faces[opp] = f;
                t.m_neighbors[opp] = F.m_neighbor_cells[1 - side];
            }
        }
        for (uInt c = 0; c < nc; ++c) {
            auto& t = m_cells[c];
            Scalar s = 0.0;
            for (int a = 0; a < 4; ++a) {
                if (t.m_faces[a] == uInt(-1)) throw std::runtime_error("ComputingMesh: cell with a missing face");
                s += m_faces[t.m_faces[a]].m_area;
            }
            t.m_h = t.m_volume / s;
        }
        m_boundaryTypes.assign(nf, BoundaryType::Dirichlet);
    }

    void face_geometry(CompTriangleFace& F) const {
        const auto &p0 = F.m_coords[0], &p1 = F.m_coords[1], &p2 = F.m_coords[2];
        vector3f n = vec_cross(p1 - p0, p2 - p0);
        const auto& c0 = m_cells[F.m_neighbor_cells[0]];
        if (F.m_neighbor_cells[1] == uInt(-1) || F.m_nodes_r != F.m_nodes) {
            // outward from cell 0: away from its vertex that is not on the face
            vector3f p3 = {0, 0, 0};
            for (int a = 0; a < 4; ++a) {
                const uInt v = c0.m_nodes[a];
                if (v != F.m_nodes[0] && v != F.m_nodes[1] && v != F.m_nodes[2]) p3 = c0.m_coords[a];
            }
            const auto v1 = p1 - p0, v2 = p2 - p0, v3 = p3 - p0;
            if (vec_dot(vec_cross(v1, v2), v3) * vec_dot(vec_cross(v1, v2), n) > 0) n *= -1.0;
        } else if (vec_dot(n, m_cells[F.m_neighbor_cells[1]].m_centroid - c0.m_centroid) < 0) {
            n *= -1.0;
        }
        F.m_area = vec_length(n) * 0.5;
        F.m_normal = vec_unit(n);
        for (int side = 0; side < 2; ++side) {
            for (int k = 0; k < 3; ++k) F.m_natural_coords[side][k] = {0, 0, 0};
            const uInt c = F.m_neighbor_cells[side];
            if (c == uInt(-1)) continue;
            const vector3u& nodes = side ? F.m_nodes_r : F.m_nodes;
            for (int k = 0; k < 3; ++k)
                for (int ki = 0; ki < 3; ++ki)
                    if (nodes[k] == m_cells[c].m_nodes[ki + 1]) F.m_natural_coords[side][k][ki] = 1;
        }
    }
};
