// Hexahedron -> tetrahedron splitting that reproduces the reference's
// GeneralMesh::split_hex5_scan / split_hex6_scan tet for tet
// (src/Mesh/GeneralMesh.cpp:78-300, src/Mesh/GeneralMeshTools.cpp:1-326):
// the same tetrahedra in the same order, the same faces with the same numbering,
// vertex order and face-cell links.  Together with the reference-form
// GeneralMesh::rebuild_cell_topology (vertex order of every tetrahedron) a caller
// that writes OrthHexMesh + split_hex*_scan + rebuild_cell_topology gets the
// reference's mesh, and hence its residual, from the same calls.
//
// The steps are the reference's, in its order, because its result depends on them:
//   * which diagonal a quadrilateral receives depends on the order in which
//     hexahedra and prisms are visited and on floating-point comparisons of the two
//     diagonal lengths (ties included);
//   * its triangle-existence test looks up one node hash (a hash_combine chain,
//     GeneralMeshTools.cpp:319-326) in a set filled with another (17 * 997^k + ...,
//     :288-298), so it almost never succeeds -- kept as is;
//   * faces are merged by the second hash alone (distinct faces with equal hashes
//     become one face, as there);
//   * the vertex order of the middle tetrahedron of a prism follows the iteration
//     order of a std::unordered_set<uint64_t> (GeneralMesh.cpp:228-247), which the
//     same standard library reproduces for the same insertions.
// Input: the hexahedra of an OrthHexMesh with their six quadrilateral faces in
// ascending face order (opposite pairs (0,1), (2,3), (4,5)), nodes v0..v7.
#pragma once
#include <algorithm>
#include <array>
#include <initializer_list>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include "base/Type.h"

namespace fvdg_scan {

constexpr uInt kNone = uInt(-1);

struct Quad {
    vector4u nodes;
    vector2u cells = {kNone, kNone};
    vector2u diag = {kNone, kNone};  // the two end nodes of the cutting diagonal, once decided
};
struct Hex {
    vector8u nodes;
    std::array<uInt, 6> faces;  // quadrilaterals; pairs (0,1), (2,3), (4,5) are opposite
};
// a face of the intermediate mesh (n = 3 triangle, n = 4 quadrilateral)
struct Face {
    uInt n = 3;
    vector4u nodes = {kNone, kNone, kNone, kNone};
    vector2u cells = {kNone, kNone};
    vector2u diag = {kNone, kNone};
};
// a cell of the intermediate mesh: a tetrahedron (nodes[0..3]) or a prism with
// triangles faces[0], faces[1] and quadrilaterals faces[2..4]
struct Cell {
    bool prism = false;
    vector6u nodes;
    std::array<uInt, 5> faces;
};

// node hash of a face for merging (GeneralMeshTools.cpp:288-298)
inline uInt merge_hash(const uInt* nodes, uInt n) {
    std::array<uInt, 4> s = {0, 0, 0, 0};
    std::copy(nodes, nodes + n, s.begin());
    std::sort(s.begin(), s.begin() + n);
    uInt h = 17;
    for (uInt i = 0; i < n; ++i) h = h * 997 + s[i];
    return h;
}
// node hash of the triangle-existence test (GeneralMeshTools.cpp:247-252, 319-326)
inline uInt lookup_hash(uInt a, uInt b, uInt c) {
    std::array<uInt, 3> s = {a, b, c};
    std::sort(s.begin(), s.end());
    uInt seed = 997;
    for (uInt v : s) seed ^= v + 0x9e3779b9u + (seed << 6) + (seed >> 2);
    return seed;
}

// Faces keyed by merge_hash; the first cell to add a face is its side 0, the
// second its side 1, later ones change nothing (GeneralMesh.cpp:81-91, 301-317).
struct FaceTable {
    std::vector<Face> faces;
    std::unordered_map<uInt, uInt> index;
    uInt add(std::initializer_list<uInt> nodes, uInt cell) {
        const uInt h = merge_hash(nodes.begin(), nodes.size());
        auto it = index.find(h);
        if (it != index.end()) {
            Face& f = faces[it->second];
            if (f.cells[1] == kNone) f.cells[1] = cell;
            return it->second;
        }
        Face f;
        f.n = nodes.size();
        std::copy(nodes.begin(), nodes.end(), f.nodes.begin());
        f.cells = {cell, kNone};
        faces.push_back(f);
        index.emplace(h, faces.size() - 1);
        return faces.size() - 1;
    }
};

inline bool contains(const uInt* set, uInt n, uInt v) { return std::find(set, set + n, v) != set + n; }
inline uInt distinct(std::array<uInt, 6> v) {
    std::sort(v.begin(), v.end());
    return std::unique(v.begin(), v.end()) - v.begin();
}

class Splitter {
public:
    explicit Splitter(const std::vector<vector3f>& points) : P(points) {}

    // Split every hexahedron (prefer5: 5 tets first, else 2 prisms first), settle
    // the prisms' side diagonals, split the prisms: tets + triangles out.
    void run(const std::vector<Hex>& hexes, std::vector<Quad>& quads, bool prefer5, std::vector<vector4u>& tets,
             std::vector<Face>& tris) {
        for (const Hex& h : hexes) {
            const bool ok = prefer5 ? (split5(h, quads) || split6(h, quads)) : (split6(h, quads) || split5(h, quads));
            if (!ok) throw std::runtime_error("Hexahedron splitting failed");
        }
        settle_prism_diagonals();
        split_cells(tets, tris);
    }

private:
    const std::vector<vector3f>& P;
    std::vector<Cell> cells;
    FaceTable ft;

    Scalar len(uInt a, uInt b) const { return vec_length(P[a] - P[b]); }
    vector2u shorter(const vector4u& q) const {  // GeneralMeshTools.cpp:256-268
        return len(q[0], q[2]) < len(q[1], q[3]) ? vector2u{q[0], q[2]} : vector2u{q[1], q[3]};
    }
    vector2u longer(const vector4u& q) const {  // GeneralMeshTools.cpp:269-280
        return len(q[0], q[2]) < len(q[1], q[3]) ? vector2u{q[1], q[3]} : vector2u{q[0], q[2]};
    }

    // 5 tets around the corner set {0,2,5,7} or {1,3,4,6}, allowed when every decided
    // face diagonal of the hexahedron lies in that set (GeneralMesh.cpp:138-157, 335-351, 483-500)
    bool split5(const Hex& h, std::vector<Quad>& quads) {
        const vector8u& n = h.nodes;
        const uInt even[4] = {n[0], n[2], n[5], n[7]}, odd[4] = {n[1], n[3], n[4], n[6]};
        bool use_even = true, use_odd = true;
        for (uInt f : h.faces) {
            const vector2u& d = quads[f].diag;
            if (d[0] == kNone) continue;
            use_even = use_even && contains(even, 4, d[0]) && contains(even, 4, d[1]);
            use_odd = use_odd && contains(odd, 4, d[0]) && contains(odd, 4, d[1]);
        }
        if (!use_even && !use_odd) return false;
        static const int tet_even[5][4] = {{0, 1, 2, 5}, {0, 2, 3, 7}, {7, 6, 5, 2}, {7, 5, 4, 0}, {0, 2, 5, 7}};
        static const int tet_odd[5][4] = {{1, 2, 3, 6}, {1, 3, 0, 4}, {6, 5, 4, 1}, {6, 4, 7, 3}, {1, 3, 4, 6}};
        const int(*pat)[4] = use_even ? tet_even : tet_odd;
        for (int t = 0; t < 5; ++t) {
            const uInt a = n[pat[t][0]], b = n[pat[t][1]], c = n[pat[t][2]], d = n[pat[t][3]];
            Cell cell;
            cell.nodes = {a, b, c, d, kNone, kNone};
            const uInt id = cells.size();  // faces opposite each vertex (GeneralMesh.cpp:404-418)
            cell.faces = {ft.add({d, b, c}, id), ft.add({a, d, c}, id), ft.add({a, b, d}, id), ft.add({a, b, c}, id),
                          kNone};
            cells.push_back(cell);
        }
        // the hexahedron's undecided faces take the diagonal through the chosen corners
        const uInt* S = use_even ? even : odd;
        for (uInt f : h.faces) {
            Quad& q = quads[f];
            if (q.diag[0] != kNone) continue;
            int k = 0;
            for (int s = 0; s < 4 && k < 2; ++s)
                if (contains(q.nodes.data(), 4, S[s])) q.diag[k++] = S[s];
        }
        return true;
    }

    // 2 prisms along the first axis whose face pair is not fully decided
    // (GeneralMesh.cpp:160-180, 354-381, 434-480, 502-525)
    bool split6(const Hex& h, std::vector<Quad>& quads) {
        int axis = -1;
        for (int a = 0; a < 3 && axis < 0; ++a) {
            const vector2u &d1 = quads[h.faces[2 * a]].diag, &d2 = quads[h.faces[2 * a + 1]].diag;
            if (d1[0] == kNone || d2[0] == kNone || d1[0] == d2[0] || d1[0] == d2[1]) axis = a;
        }
        if (axis < 0) return false;
        static const int prism[3][2][6] = {{{0, 1, 2, 4, 5, 6}, {2, 3, 0, 6, 7, 4}},
                                           {{0, 1, 5, 3, 2, 6}, {3, 0, 5, 7, 4, 6}},
                                           {{0, 2, 3, 4, 6, 7}, {1, 0, 3, 5, 4, 7}}};
        for (int p = 0; p < 2; ++p) {
            Cell cell;
            cell.prism = true;
            for (int k = 0; k < 6; ++k) cell.nodes[k] = h.nodes[prism[axis][p][k]];
            const vector6u& m = cell.nodes;
            const uInt id = cells.size();
            cell.faces = {ft.add({m[0], m[1], m[2]}, id), ft.add({m[3], m[4], m[5]}, id),
                          ft.add({m[0], m[1], m[4], m[3]}, id), ft.add({m[1], m[2], m[5], m[4]}, id),
                          ft.add({m[2], m[0], m[3], m[5]}, id)};
            cells.push_back(cell);
        }
        for (int s = 0; s < 2; ++s) {
            Quad& q = quads[h.faces[2 * axis + s]];
            if (q.diag[0] == kNone) q.diag = shorter(q.nodes);
        }
        return true;
    }

    // Diagonals of the prisms' side faces (GeneralMeshTools.cpp:4-127): forced ones
    // first, then the prism-consistency passes, then free faces by length.
    void settle_prism_diagonals() {
        for (const Face& f : ft.faces)
            if (f.n == 3) tri_hashes.insert(merge_hash(f.nodes.data(), 3));
        for (Face& f : ft.faces) {
            if (f.n != 4) continue;
            const int w = forced(f.nodes);
            if (w < 0) continue;
            const uInt a = f.nodes[w], b = f.nodes[w + 2];
            if (f.diag[0] != kNone && f.diag[0] != a && f.diag[0] != b)
                throw std::runtime_error("split_hex_scan: conflicting forced diagonal");
            f.diag = {a, b};
        }
        for (uInt flag : {0, 1, 0, 1, 0, 0, 0}) prism_pass(flag);
        free_faces(false);
        prism_pass(0);
        free_faces(true);
    }
    // 0: both triangles of the 0-2 cut exist, 1: those of the 1-3 cut, -1: neither
    int forced(const vector4u& q) const {
        auto exists = [&](uInt a, uInt b, uInt c) { return tri_hashes.count(lookup_hash(a, b, c)) != 0; };
        if (exists(q[0], q[1], q[2]) && exists(q[0], q[2], q[3])) return 0;
        if (exists(q[0], q[1], q[3]) && exists(q[1], q[2], q[3])) return 1;
        return -1;
    }
    // undecided quadrilaterals: forced cut, else the shorter (first sweep) or the
    // longer (second sweep) diagonal
    void free_faces(bool second) {
        for (Face& f : ft.faces) {
            if (f.n != 4 || f.diag[0] != kNone) continue;
            const int w = forced(f.nodes);
            if (w >= 0)
                f.diag = {f.nodes[w], f.nodes[w + 2]};
            else
                f.diag = second ? longer(f.nodes) : shorter(f.nodes);
        }
    }
    std::unordered_set<uInt> tri_hashes;  // merge_hash of every triangle (the reference's lookup set)

    // one consistency pass over the prisms (GeneralMeshTools.cpp:167-241)
    void prism_pass(uInt flag) {
        for (const Cell& c : cells) {
            if (!c.prism) continue;
            std::array<vector2u, 3> d;
            int count = 0;
            for (int i = 0; i < 3; ++i) {
                d[i] = ft.faces[c.faces[2 + i]].diag;
                count += d[i][0] != kNone && d[i][1] != kNone;
            }
            if (count == 3 && distinct({d[0][0], d[0][1], d[1][0], d[1][1], d[2][0], d[2][1]}) == 6)
                ft.faces[c.faces[2]].diag = {kNone, kNone};  // no common vertex: undecide the first side
            if (count == 2) {
                for (int idx = 0; idx < 3; ++idx) {
                    const int ii = idx, jj = (idx + 1) % 3, kk = (idx + 2) % 3;
                    // the reference tests the first two sides here, whatever idx is
                    if (d[0][0] == kNone || d[1][0] == kNone) continue;
                    Face& fk = ft.faces[c.faces[2 + kk]];
                    const vector4u q = fk.nodes;
                    const uInt n02 = distinct({d[ii][0], d[ii][1], d[jj][0], d[jj][1], q[0], q[2]});
                    const uInt n13 = distinct({d[ii][0], d[ii][1], d[jj][0], d[jj][1], q[1], q[3]});
                    if (n02 == 6 && n13 == 6) throw std::runtime_error("split_hex_scan: prism without a consistent split");
                    fk.diag = (n02 == 4 && n13 == 6) ? vector2u{q[0], q[2]} : vector2u{q[1], q[3]};
                }
            }
            if (count < 2) {
                for (int i = 0; i < 3; ++i) {
                    Face& f = ft.faces[c.faces[2 + i]];
                    if (f.diag[0] != kNone) continue;
                    f.diag = flag ? shorter(f.nodes) : longer(f.nodes);
                    break;
                }
            }
        }
    }

    // tets of the final mesh, faces renumbered from scratch (GeneralMesh.cpp:183-273, 387-401)
    void split_cells(std::vector<vector4u>& tets, std::vector<Face>& tris) {
        FaceTable out;
        tets.clear();
        auto tet = [&](uInt a, uInt b, uInt c, uInt d) {
            const uInt id = tets.size();
            out.add({a, b, c}, id);
            out.add({a, b, d}, id);
            out.add({b, c, d}, id);
            out.add({c, a, d}, id);
            tets.push_back({a, b, c, d});
        };
        for (const Cell& c : cells) {
            const vector6u& n = c.nodes;
            if (!c.prism) {
                tet(n[0], n[1], n[2], n[3]);
                continue;
            }
            std::array<uInt, 6> flat;
            for (int i = 0; i < 3; ++i) {
                const vector2u& d = ft.faces[c.faces[2 + i]].diag;
                flat[2 * i] = d[0];
                flat[2 * i + 1] = d[1];
            }
            std::unordered_set<uInt> uniq;
            for (uInt v : flat) uniq.insert(v);
            if (uniq.size() == 6) throw std::runtime_error("diags split for prism error 333.");
            // the vertex shared by two side diagonals on the first triangle -> s[0], on the
            // second -> s[1]; the other diagonal ends -> s[2], s[3] in set iteration order
            vector4u s = {kNone, kNone, kNone, kNone};
            for (uInt v : uniq) {
                if (std::count(flat.begin(), flat.end(), v) == 2)
                    (contains(n.data(), 3, v) ? s[0] : s[1]) = v;
                else
                    (s[2] == kNone ? s[2] : s[3]) = v;
            }
            tet(s[0], s[1], s[2], s[3]);
            tet(n[0], n[1], n[2], s[1]);
            tet(n[3], n[4], n[5], s[0]);
        }
        tris = std::move(out.faces);
    }
};

}  // namespace fvdg_scan
