// Topological mesh: points, faces and cells, hex -> tet splitting, face
// rebuilding, periodic face pairing and vertex jitter.
//
// Class and method names follow the reference's include/Mesh/GeneralMesh.h:9-25
// (m_points / m_faces / m_cells, split_hex5_scan, split_hex6_scan,
// rebuild_cell_topology, validate_mesh) and the cell/face structs of
// include/Mesh/{CellType.h,FaceType.h}.
//
// Two splittings:
//   * split_hex5_scan / split_hex6_scan reproduce the reference's (HexSplitScan.h):
//     followed by rebuild_cell_topology they give the reference's tetrahedra,
//     vertex orders and face numbering for the hexahedra of an OrthHexMesh.
//   * split_hex5_periodic / split_hex6_periodic are this library's own: every
//     hexahedron is split by a fixed, translation-invariant pattern (6 tets around
//     the v0-v6 diagonal, or 5 tets alternating with (i+j+k) parity), which keeps
//     opposite boundary faces exact translates of each other, so that periodic
//     pairing (make_periodic) works.  The reference's diagonals depend on visiting
//     order and rounding and do not pair up across opposite planes.
// Tet nodes are oriented positively as in the reference's Tetrahedron::reorder
// (CellType.h:14-23).
//
// Additions with no counterpart in the reference (its BoundaryType,
// include/Mesh/ComputingMesh.h:71-86, has no periodic value):
//   * TriangleFace::nodes_r     -- the face's nodes as numbered in cell 1 (periodic image)
//   * GeneralMesh::make_periodic -- merges matching boundary faces on opposite planes
//   * GeneralMesh::jitter_points -- seeded vertex perturbation, consistent across periodic images
#pragma once
#include <map>
#include <random>
#include <unordered_map>
#include <unordered_set>
#include <variant>
#include "Mesh/HexSplitScan.h"
#include "base/Type.h"

//---------------- faces ----------------
struct TriangleFace {
    vector3u nodes;
    vector2u neighbor_cells = {uInt(-1), uInt(-1)};
    // nodes[k] as numbered in neighbor_cells[1]; differs from nodes only across a periodic boundary
    vector3u nodes_r = {uInt(-1), uInt(-1), uInt(-1)};
    const vector3u& side_nodes(int side) const { return (side == 1 && nodes_r[0] != uInt(-1)) ? nodes_r : nodes; }
};
struct QuadFace {
    vector4u nodes;
    vector2u neighbor_cells = {uInt(-1), uInt(-1)};
    vector2u diagonal_nodes = {uInt(-1), uInt(-1)};  // cutting diagonal, once the split decides it
};
using GeneralFace = std::variant<TriangleFace, QuadFace>;

//---------------- cells ----------------
struct Tetrahedron {
    vector4u nodes;
    // the cell's faces: faces[l] opposite nodes[l] when this library builds the faces, in
    // ascending face order after rebuild_cell_topology on existing faces (as the reference)
    vector4u faces = {uInt(-1), uInt(-1), uInt(-1), uInt(-1)};
    // positive orientation: swap nodes 1 and 2 if (p1-p0, p2-p0, p3-p0) is left-handed
    void reorder(const std::vector<vector3f>& pts) {
        const auto v1 = pts[nodes[1]] - pts[nodes[0]];
        const auto v2 = pts[nodes[2]] - pts[nodes[0]];
        const auto v3 = pts[nodes[3]] - pts[nodes[0]];
        if (vec_dot(vec_cross(v1, v2), v3) < 0) std::swap(nodes[1], nodes[2]);
    }
};
// Hexahedron corners: v0=(0,0,0) v1=(1,0,0) v2=(1,1,0) v3=(0,1,0), v4..v7 the same at z=1.
struct Hexahedron {
    vector8u nodes;
    // quadrilateral faces in ascending face order; opposite pairs (0,1), (2,3), (4,5)
    std::array<uInt, 6> faces = {uInt(-1), uInt(-1), uInt(-1), uInt(-1), uInt(-1), uInt(-1)};
    uInt parity = 0;  // (i+j+k) mod 2 of a structured hex; selects the periodic 5-tet pattern
};
using GeneralCell = std::variant<Hexahedron, Tetrahedron>;

class GeneralMesh {
public:
    std::vector<vector3f> m_points;
    std::vector<GeneralFace> m_faces;
    std::vector<GeneralCell> m_cells;

    // the reference's splittings (GeneralMesh.h:24-25); call rebuild_cell_topology after
    void split_hex5_scan() { split_scan(true); }
    void split_hex6_scan() { split_scan(false); }
    // this library's translation-invariant splittings (faces rebuilt, tets oriented)
    void split_hex5_periodic() { split_hex(true); }
    void split_hex6_periodic() { split_hex(false); }

    // Pure tet mesh.  With faces present (after split_hex*_scan), as the reference's
    // rebuild_cell_topology (GeneralMesh.cpp:4-52): each cell's faces in ascending
    // order, its nodes in the iteration order of a std::unordered_set<uint64_t> filled
    // with those faces' nodes in that order, then oriented positively.  Without faces,
    // builds them: one face per distinct node triple, numbered in order of first
    // appearance over (cell, local face); face nodes in the first cell's local order;
    // side-0 cell is the first cell seen.  A no-op while hexahedra remain.
    void rebuild_cell_topology() {
        for (const auto& c : m_cells)
            if (!std::holds_alternative<Tetrahedron>(c)) return;
        if (!m_faces.empty()) {
            reorder_from_faces();
            return;
        }
        std::unordered_map<Key3, uInt, Key3Hash> index;
        index.reserve(m_cells.size() * 3);
        std::vector<GeneralFace> faces;
        faces.reserve(m_cells.size() * 2 + 16);
        for (uInt c = 0; c < m_cells.size(); ++c) {
            auto& t = std::get<Tetrahedron>(m_cells[c]);
            t.reorder(m_points);
            for (int l = 0; l < 4; ++l) {
                vector3u fn;
                int m = 0;
                for (int a = 0; a < 4; ++a)
                    if (a != l) fn[m++] = t.nodes[a];
                const Key3 key = sorted_key(fn);
                auto it = index.find(key);
                if (it == index.end()) {
                    TriangleFace f;
                    f.nodes = fn;
                    f.neighbor_cells = {c, uInt(-1)};
                    index.emplace(key, faces.size());
                    t.faces[l] = faces.size();
                    faces.push_back(f);
                } else {
                    auto& f = std::get<TriangleFace>(faces[it->second]);
                    if (f.neighbor_cells[1] != uInt(-1)) throw std::runtime_error("face shared by more than two cells");
                    f.neighbor_cells[1] = c;
                    t.faces[l] = it->second;
                }
            }
        }
        m_faces.swap(faces);
    }

    // Every face has at least one cell, every tet face is one of its own faces,
    // and every tet has positive volume (reference: GeneralMesh.cpp:54-74).
    void validate_mesh() const {
        for (const auto& fv : m_faces) {
            const vector2u nc = std::visit([](auto&& f) { return f.neighbor_cells; }, fv);
            if (nc[0] == uInt(-1) && nc[1] == uInt(-1)) throw std::runtime_error("Orphan face detected");
            for (uInt c : nc)
                if (c != uInt(-1) && c >= m_cells.size()) throw std::runtime_error("face references a missing cell");
        }
        for (uInt c = 0; c < m_cells.size(); ++c) {
            if (const auto* t = std::get_if<Tetrahedron>(&m_cells[c])) {
                const auto v1 = m_points[t->nodes[1]] - m_points[t->nodes[0]];
                const auto v2 = m_points[t->nodes[2]] - m_points[t->nodes[0]];
                const auto v3 = m_points[t->nodes[3]] - m_points[t->nodes[0]];
                if (!(vec_dot(vec_cross(v1, v2), v3) > 0)) throw std::runtime_error("degenerate or inverted tetrahedron");
                for (uInt f : t->faces) {
                    if (f == uInt(-1)) continue;
                    if (f >= m_faces.size()) throw std::runtime_error("cell references a missing face");
                    const vector2u nc = std::visit([](auto&& x) { return x.neighbor_cells; }, m_faces[f]);
                    if (nc[0] != c && nc[1] != c) throw std::runtime_error("cell-face incidence is not symmetric");
                }
            }
        }
    }

    // Pair boundary faces on opposite planes x=lb/ub (per periodic axis) whose
    // nodes coincide after the translation, and merge each pair into one interior
    // face: side 0 keeps the first face's cell and nodes, side 1 is the other
    // cell with nodes_r = the matching image nodes.  Unmatched boundary faces
    // stay boundary faces.  Returns the number of merged pairs.
    uInt make_periodic(const vector3f& lb, const vector3f& ub, std::array<bool, 3> axes) {
        const std::vector<uInt> cls = periodic_classes(lb, ub, axes);
        const Scalar tol = plane_tolerance(lb, ub), q = min_edge() / 64;
        // Faces are matched by their wrapped centroid: on a grid with two cells along a
        // periodic axis, distinct faces of one plane can share all three node classes.
        std::map<std::array<int64_t, 3>, uInt> first;
        std::vector<uInt> remap(m_faces.size(), uInt(-1));
        std::vector<char> drop(m_faces.size(), 0);
        uInt merged = 0;
        for (uInt f = 0; f < m_faces.size(); ++f) {
            auto* tf = std::get_if<TriangleFace>(&m_faces[f]);
            if (!tf || tf->neighbor_cells[1] != uInt(-1)) continue;
            vector3f c = (m_points[tf->nodes[0]] + m_points[tf->nodes[1]] + m_points[tf->nodes[2]]) / Scalar(3);
            const std::array<int64_t, 3> key = wrapped_key(c, lb, ub, axes, tol, q);
            auto it = first.find(key);
            if (it == first.end()) {
                first.emplace(key, f);
                continue;
            }
            auto& keep = std::get<TriangleFace>(m_faces[it->second]);
            if (keep.neighbor_cells[1] != uInt(-1)) throw std::runtime_error("periodic face matched twice");
            keep.neighbor_cells[1] = tf->neighbor_cells[0];
            for (int k = 0; k < 3; ++k) {
                keep.nodes_r[k] = uInt(-1);
                for (int m = 0; m < 3; ++m)
                    if (cls[tf->nodes[m]] == cls[keep.nodes[k]]) keep.nodes_r[k] = tf->nodes[m];
                if (keep.nodes_r[k] == uInt(-1)) throw std::runtime_error("periodic node match failed");
            }
            drop[f] = 1;
            remap[f] = it->second;
            ++merged;
        }
        // compact faces and renumber cell -> face references
        std::vector<uInt> newid(m_faces.size());
        std::vector<GeneralFace> faces;
        faces.reserve(m_faces.size() - merged);
        for (uInt f = 0; f < m_faces.size(); ++f)
            if (!drop[f]) {
                newid[f] = faces.size();
                faces.push_back(m_faces[f]);
            }
        for (uInt f = 0; f < m_faces.size(); ++f)
            if (drop[f]) newid[f] = newid[remap[f]];
        for (auto& c : m_cells)
            if (auto* t = std::get_if<Tetrahedron>(&c))
                for (auto& f : t->faces)
                    if (f != uInt(-1)) f = newid[f];
        m_faces.swap(faces);
        return merged;
    }

    // Move every point by a seeded uniform offset in (-amp*h_d, amp*h_d) per
    // coordinate.  Points that are periodic images of each other (same wrapped
    // position) get the same offset; coordinates on a non-periodic boundary
    // plane are not moved along that axis.
    void jitter_points(Scalar amp, uint64_t seed, const vector3f& lb, const vector3f& ub, const vector3f& h,
                       std::array<bool, 3> periodic) {
        if (amp == 0) return;
        const std::vector<uInt> cls = periodic_classes(lb, ub, periodic);
        const Scalar tol = plane_tolerance(lb, ub);
        uInt ncls = 0;
        for (uInt c : cls) ncls = std::max(ncls, c + 1);
        std::vector<vector3f> off(ncls);
        std::mt19937_64 rng(seed);
        std::uniform_real_distribution<Scalar> U(-1.0, 1.0);
        for (auto& o : off)
            for (int d = 0; d < 3; ++d) o[d] = amp * h[d] * U(rng);
        for (uInt p = 0; p < m_points.size(); ++p)
            for (int d = 0; d < 3; ++d) {
                const bool on_plane = std::abs(m_points[p][d] - lb[d]) <= tol || std::abs(m_points[p][d] - ub[d]) <= tol;
                if (!periodic[d] && on_plane) continue;
                m_points[p][d] += off[cls[p]][d];
            }
    }

private:
    void reorder_from_faces() {
        std::vector<std::vector<uInt>> cell_faces(m_cells.size());
        for (uInt f = 0; f < m_faces.size(); ++f) {
            const vector2u nc = std::visit([](auto&& x) { return x.neighbor_cells; }, m_faces[f]);
            for (uInt c : nc)
                if (c != uInt(-1)) {
                    if (c >= m_cells.size()) throw std::runtime_error("face references a missing cell");
                    cell_faces[c].push_back(f);
                }
        }
        for (uInt c = 0; c < m_cells.size(); ++c) {
            auto& t = std::get<Tetrahedron>(m_cells[c]);
            if (cell_faces[c].size() != 4) throw std::runtime_error("Cell face count mismatch");
            std::unordered_set<uInt> nodes;
            for (uInt f : cell_faces[c]) {
                const auto* tf = std::get_if<TriangleFace>(&m_faces[f]);
                if (!tf) throw std::runtime_error("tetrahedron with a quadrilateral face");
                for (uInt v : tf->nodes) nodes.insert(v);
            }
            if (nodes.size() != 4) throw std::runtime_error("Cell node count mismatch");
            std::copy(nodes.begin(), nodes.end(), t.nodes.begin());
            std::copy(cell_faces[c].begin(), cell_faces[c].end(), t.faces.begin());
            t.reorder(m_points);
        }
    }

    // the reference's splitting of the hexahedra of an OrthHexMesh (HexSplitScan.h)
    void split_scan(bool five) {
        std::vector<fvdg_scan::Hex> hexes;
        hexes.reserve(m_cells.size());
        for (const auto& c : m_cells) {
            const auto* h = std::get_if<Hexahedron>(&c);
            if (!h || h->faces[0] == uInt(-1))
                throw std::runtime_error("split_hex*_scan: needs the hexahedra and faces of an OrthHexMesh");
            hexes.push_back({h->nodes, h->faces});
        }
        std::vector<fvdg_scan::Quad> quads(m_faces.size());
        for (uInt f = 0; f < m_faces.size(); ++f) {
            const auto* q = std::get_if<QuadFace>(&m_faces[f]);
            if (!q) throw std::runtime_error("split_hex*_scan: hexahedron faces must be quadrilaterals");
            quads[f].nodes = q->nodes;
            quads[f].cells = q->neighbor_cells;
            quads[f].diag = q->diagonal_nodes;
        }
        std::vector<vector4u> tets;
        std::vector<fvdg_scan::Face> tris;
        fvdg_scan::Splitter(m_points).run(hexes, quads, five, tets, tris);
        m_cells.assign(tets.size(), Tetrahedron{});
        for (uInt c = 0; c < tets.size(); ++c) std::get<Tetrahedron>(m_cells[c]).nodes = tets[c];
        m_faces.assign(tris.size(), TriangleFace{});
        for (uInt f = 0; f < tris.size(); ++f) {
            auto& tf = std::get<TriangleFace>(m_faces[f]);
            tf.nodes = {tris[f].nodes[0], tris[f].nodes[1], tris[f].nodes[2]};
            tf.neighbor_cells = tris[f].cells;
        }
    }

    static Scalar plane_tolerance(const vector3f& lb, const vector3f& ub) {
        Scalar L = 0;
        for (int d = 0; d < 3; ++d) L = std::max(L, ub[d] - lb[d]);
        return 1e-9 * L;
    }
    Scalar min_edge() const {
        Scalar h = INFINITY;
        auto upd = [&](uInt a, uInt b) { h = std::min(h, vec_length(m_points[a] - m_points[b])); };
        for (const auto& c : m_cells) {
            if (const auto* t = std::get_if<Tetrahedron>(&c)) {
                for (int a = 0; a < 4; ++a)
                    for (int b = a + 1; b < 4; ++b) upd(t->nodes[a], t->nodes[b]);
            } else if (const auto* hx = std::get_if<Hexahedron>(&c)) {
                upd(hx->nodes[0], hx->nodes[1]);
                upd(hx->nodes[0], hx->nodes[3]);
                upd(hx->nodes[0], hx->nodes[4]);
            }
        }
        return h;
    }
    // Periodic class of every point: its position wrapped onto [lb, ub) along the
    // periodic axes, quantised to 1/64 of the shortest cell edge (points of one
    // class are images of each other).  Classes are numbered by first appearance.
    std::vector<uInt> periodic_classes(const vector3f& lb, const vector3f& ub, std::array<bool, 3> axes) const {
        const Scalar tol = plane_tolerance(lb, ub);
        const Scalar q = min_edge() / 64;
        std::map<std::array<int64_t, 3>, uInt> cls_of;
        std::vector<uInt> cls(m_points.size());
        for (uInt p = 0; p < m_points.size(); ++p)
            cls[p] = cls_of.emplace(wrapped_key(m_points[p], lb, ub, axes, tol, q), (uInt)cls_of.size()).first->second;
        return cls;
    }
    // A position wrapped onto [lb, ub) along the periodic axes, quantised to q.
    static std::array<int64_t, 3> wrapped_key(const vector3f& p, const vector3f& lb, const vector3f& ub,
                                              std::array<bool, 3> axes, Scalar tol, Scalar q) {
        std::array<int64_t, 3> key;
        for (int d = 0; d < 3; ++d) {
            Scalar x = p[d];
            if (axes[d] && std::abs(x - ub[d]) <= tol) x = lb[d];
            key[d] = (int64_t)std::llround((x - lb[d]) / q);
        }
        return key;
    }

    using Key3 = std::array<uInt, 3>;
    struct Key3Hash {
        size_t operator()(const Key3& k) const {
            uint64_t h = k[0] * 0x9E3779B97F4A7C15ull;
            h ^= k[1] + 0x7F4A7C159E3779B9ull + (h << 6) + (h >> 2);
            h ^= k[2] + 0x94D049BB133111EBull + (h << 6) + (h >> 2);
            return (size_t)h;
        }
    };
    static Key3 sorted_key(vector3u k) {
        std::sort(k.begin(), k.end());
        return k;
    }

    void split_hex(bool five) {
        // 6 tets sharing the v0-v6 diagonal: one per monotone edge path v0 -> v6
        static const int six[6][4] = {{0, 1, 2, 6}, {0, 1, 5, 6}, {0, 3, 2, 6}, {0, 3, 7, 6}, {0, 4, 5, 6}, {0, 4, 7, 6}};
        // 5 tets: a central tet plus 4 corners, mirrored between the two parities
        static const int five_even[5][4] = {{0, 2, 5, 7}, {1, 0, 2, 5}, {3, 0, 2, 7}, {4, 0, 5, 7}, {6, 2, 5, 7}};
        static const int five_odd[5][4] = {{1, 3, 4, 6}, {0, 1, 3, 4}, {2, 1, 3, 6}, {5, 1, 4, 6}, {7, 3, 4, 6}};
        std::vector<GeneralCell> cells;
        cells.reserve(m_cells.size() * (five ? 5 : 6));
        for (const auto& c : m_cells) {
            if (const auto* h = std::get_if<Hexahedron>(&c)) {
                const int(*pat)[4] = five ? (h->parity ? five_odd : five_even) : six;
                const int nt = five ? 5 : 6;
                for (int t = 0; t < nt; ++t) {
                    Tetrahedron tet;
                    for (int a = 0; a < 4; ++a) tet.nodes[a] = h->nodes[pat[t][a]];
                    tet.reorder(m_points);
                    cells.push_back(tet);
                }
            } else {
                cells.push_back(c);
            }
        }
        m_cells.swap(cells);
        m_faces.clear();
        rebuild_cell_topology();
    }
};
