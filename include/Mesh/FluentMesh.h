// Reader of Fluent ASCII mesh files (.msh, the format of the reference's data/case11.msh and
// data/case46.msh) into a conforming tetrahedral GeneralMesh.
//
// Sections read: (0 comment), (2 dimension), (10 nodes), (12 cells), (13 faces),
// (39 / 45 zone names and types); numbers in the section headers and in the node, cell and
// face lists are hexadecimal, node coordinates decimal.  A face line lists its nodes (after
// a node count when the zone's face type is 0, mixed) followed by its two cells, 0 for none.
// Fluent describes cells only through their faces, so every cell is rebuilt from its faces:
// a cell with four triangular faces is kept as a tetrahedron; any other cell (hexahedron,
// wedge, pyramid, polyhedron) is split into tetrahedra joining its centroid to its faces,
// every quadrilateral face being cut along the diagonal through its smallest node index --
// both neighbours of a face cut it the same way, so the result is conforming.
// Boundary faces keep their zone (id, Fluent type, name) through the same triangulation.
#pragma once
#include <cctype>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

#include "Mesh/ComputingMesh.h"

struct FluentZone {
    int id = 0;
    int bc_type = 0;       // Fluent boundary-condition code of the face zone (2 interior, 3 wall, ...)
    std::string kind;      // zone type word of the (39 ...) section, e.g. "wall", "velocity-inlet"
    std::string name;
};

class FluentMesh {
public:
    struct Face {
        std::vector<uInt> nodes;  // 0-based
        int64_t c0 = -1, c1 = -1;  // 0-based cells (Fluent's right / left cell), -1 none
        int zone = 0;
    };
    int dimension = 3;
    std::vector<vector3f> points;
    std::vector<Face> faces;
    int64_t n_cells = 0;
    std::map<int, FluentZone> zones;

    void read(const std::string& path) {
        std::ifstream in(path);
        if (!in) throw std::runtime_error("FluentMesh: cannot open " + path);
        std::stringstream ss;
        ss << in.rdbuf();
        text_ = ss.str();
        pos_ = 0;
        while (skip_to_open()) section();
        if (points.empty() || faces.empty() || n_cells <= 0) throw std::runtime_error("FluentMesh: no nodes, faces or cells");
        if (dimension != 3) throw std::runtime_error("FluentMesh: only 3D meshes are supported");
    }

    // Tetrahedral mesh (points, tets, faces rebuilt) and, per face of the result, the Fluent
    // zone of the original face it lies in (-1 for faces inside a split cell).
    GeneralMesh to_tets(std::vector<int32_t>& face_zone) const {
        std::vector<std::vector<uInt>> cell_faces(n_cells);
        for (uInt f = 0; f < faces.size(); ++f) {
            if (faces[f].c0 >= n_cells || faces[f].c1 >= n_cells) throw std::runtime_error("FluentMesh: face cell out of range");
            if (faces[f].c0 >= 0) cell_faces[faces[f].c0].push_back(f);
            if (faces[f].c1 >= 0) cell_faces[faces[f].c1].push_back(f);
        }
        GeneralMesh g;
        g.m_points = points;
        for (int64_t c = 0; c < n_cells; ++c) {
            const auto& cf = cell_faces[c];
            if (cf.empty()) throw std::runtime_error("FluentMesh: cell without faces");
            bool tet = cf.size() == 4;
            for (uInt f : cf) tet = tet && faces[f].nodes.size() == 3;
            if (tet) {
                std::vector<uInt> nodes;
                for (uInt f : cf)
                    for (uInt n : faces[f].nodes)
                        if (std::find(nodes.begin(), nodes.end(), n) == nodes.end()) nodes.push_back(n);
                if (nodes.size() != 4) throw std::runtime_error("FluentMesh: inconsistent tetrahedron");
                Tetrahedron t;
                t.nodes = {nodes[0], nodes[1], nodes[2], nodes[3]};
                g.m_cells.push_back(t);
                continue;
            }
            // centroid of the cell's distinct nodes
            std::vector<uInt> nodes;
            for (uInt f : cf)
                for (uInt n : faces[f].nodes)
                    if (std::find(nodes.begin(), nodes.end(), n) == nodes.end()) nodes.push_back(n);
            vector3f ctr = {0, 0, 0};
            for (uInt n : nodes) ctr = ctr + points[n];
            ctr = ctr / Scalar(nodes.size());
            const uInt ic = g.m_points.size();
            g.m_points.push_back(ctr);
            for (uInt f : cf)
                for (const auto& tri : triangulate(faces[f].nodes)) {
                    Tetrahedron t;
                    t.nodes = {tri[0], tri[1], tri[2], ic};
                    g.m_cells.push_back(t);
                }
        }
        g.rebuild_cell_topology();
        g.validate_mesh();
        // zones of the boundary (and zone-interface) triangles
        std::map<vector3u, int32_t> zone_of;
        for (const auto& F : faces)
            for (const auto& tri : triangulate(F.nodes)) zone_of[key(tri)] = F.zone;
        face_zone.assign(g.m_faces.size(), -1);
        for (uInt f = 0; f < g.m_faces.size(); ++f) {
            const auto& t = std::get<TriangleFace>(g.m_faces[f]);
            auto it = zone_of.find(key(t.nodes));
            if (it != zone_of.end()) face_zone[f] = it->second;
        }
        return g;
    }

    // the reference's BoundaryType for a Fluent face-zone code (interior zones: none)
    static BoundaryType boundary_type_of(int bc_type) {
        switch (bc_type) {
            case 4: case 10: case 20: return BoundaryType::Inflow;   // pressure / velocity / mass-flow inlet
            case 5: case 36: return BoundaryType::Outflow;           // pressure outlet, outflow
            case 7: return BoundaryType::Neumann;                    // symmetry
            case 9: return BoundaryType::Dirichlet;                  // pressure far field
            default: return BoundaryType::Wall;
        }
    }

private:
    std::string text_;
    size_t pos_ = 0;

    // the two triangles of a quadrilateral cut along the diagonal through its smallest node
    // (a triangle is returned as it is; larger polygons as a fan from the smallest node)
    static std::vector<vector3u> triangulate(const std::vector<uInt>& p) {
        const size_t n = p.size();
        if (n == 3) return {vector3u{p[0], p[1], p[2]}};
        size_t m = 0;
        for (size_t i = 1; i < n; ++i)
            if (p[i] < p[m]) m = i;
        std::vector<vector3u> out;
        for (size_t i = 1; i + 1 < n; ++i) out.push_back({p[m], p[(m + i) % n], p[(m + i + 1) % n]});
        return out;
    }
    static vector3u key(vector3u k) {
        std::sort(k.begin(), k.end());
        return k;
    }

    // ---- tokenizer over the S-expression text
    bool skip_to_open() {
        while (pos_ < text_.size() && text_[pos_] != '(') ++pos_;
        return pos_ < text_.size();
    }
    void skip_ws() {
        while (pos_ < text_.size() && std::isspace((unsigned char)text_[pos_])) ++pos_;
    }
    void expect(char c) {
        skip_ws();
        if (pos_ >= text_.size() || text_[pos_] != c)
            throw std::runtime_error(std::string("FluentMesh: expected '") + c + "' at offset " + std::to_string(pos_));
        ++pos_;
    }
    std::string token() {
        skip_ws();
        const size_t b = pos_;
        while (pos_ < text_.size() && !std::isspace((unsigned char)text_[pos_]) && text_[pos_] != '(' &&
               text_[pos_] != ')')
            ++pos_;
        if (b == pos_) throw std::runtime_error("FluentMesh: expected a token at offset " + std::to_string(pos_));
        return text_.substr(b, pos_ - b);
    }
    int64_t hex() { return (int64_t)std::stoll(token(), nullptr, 16); }
    bool at(char c) {
        skip_ws();
        return pos_ < text_.size() && text_[pos_] == c;
    }
    // skip a balanced parenthesised group starting at '(' (strings in quotes are opaque)
    void skip_group() {
        expect('(');
        int depth = 1;
        bool quoted = false;
        while (pos_ < text_.size() && depth > 0) {
            const char ch = text_[pos_++];
            if (ch == '"') quoted = !quoted;
            else if (!quoted && ch == '(') ++depth;
            else if (!quoted && ch == ')') --depth;
        }
    }
    std::vector<int64_t> hex_header() {
        expect('(');
        std::vector<int64_t> h;
        while (!at(')')) h.push_back(hex());
        expect(')');
        return h;
    }

    void section() {
        const size_t start = pos_;
        expect('(');
        const std::string id = token();
        if (id == "2") {
            dimension = (int)std::stol(token());
            expect(')');
        } else if (id == "10") {
            const auto h = hex_header();  // zone first last type [nd]
            if (h.size() < 4) throw std::runtime_error("FluentMesh: bad node section header");
            if (h.size() >= 5 && h[4] != 3) throw std::runtime_error("FluentMesh: only 3D nodes are supported");
            if (h[0] == 0) {
                points.resize(h[2]);
            } else {
                if ((int64_t)points.size() < h[2]) points.resize(h[2]);
                expect('(');
                for (int64_t i = h[1] - 1; i < h[2]; ++i)
                    for (int d = 0; d < 3; ++d) points[i][d] = std::stod(token());
                expect(')');
            }
            expect(')');
        } else if (id == "12") {
            const auto h = hex_header();  // zone first last type element-type
            if (h.size() < 4) throw std::runtime_error("FluentMesh: bad cell section header");
            n_cells = std::max<int64_t>(n_cells, h[2]);
            if (at('(')) skip_group();  // per-cell element types of a mixed zone
            expect(')');
        } else if (id == "13") {
            const auto h = hex_header();  // zone first last bc-type face-type
            if (h.size() < 5) throw std::runtime_error("FluentMesh: bad face section header");
            if (h[0] != 0) {
                const int zone = (int)h[0], ftype = (int)h[4];
                zones[zone].id = zone;
                zones[zone].bc_type = (int)h[3];
                if (ftype == 2) throw std::runtime_error("FluentMesh: 2D (line) faces are not supported");
                if ((int64_t)faces.size() < h[2]) faces.resize(h[2]);
                expect('(');
                for (int64_t i = h[1] - 1; i < h[2]; ++i) {
                    const int64_t nn = ftype == 0 ? hex() : ftype;
                    Face& F = faces[i];
                    F.nodes.resize(nn);
                    for (int64_t k = 0; k < nn; ++k) F.nodes[k] = (uInt)(hex() - 1);
                    F.c0 = hex() - 1;
                    F.c1 = hex() - 1;
                    F.zone = zone;
                }
                expect(')');
            }
            expect(')');
        } else if (id == "39" || id == "45") {
            expect('(');
            const int zone = (int)std::stol(token());
            const std::string kind = token();
            std::string name = at(')') ? std::string() : token();
            zones[zone].id = zone;
            zones[zone].kind = kind;
            zones[zone].name = name;
            pos_ = start;
            skip_group();
        } else {
            pos_ = start;
            skip_group();
        }
    }
};
