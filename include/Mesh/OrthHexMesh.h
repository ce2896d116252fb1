// Structured box of Nx*Ny*Nz hexahedra (reference: include/Mesh/OrthHexMesh.h:3-106).
// Point (i,j,k) has index i + (j + k*(Ny+1))*(Nx+1) and position
// lb + (i,j,k) * (ub-lb)/(Nx,Ny,Nz); the quadrilateral faces are numbered, noded
// and linked to their cells as in the reference (z-normal faces, then x-normal,
// then y-normal; boundary faces have one cell), because the reference's splitting
// (split_hex*_scan) depends on that numbering.  Each hexahedron gets the nodes v0..v7
// (v0=(i,j,k) v1=(i+1,j,k) v2=(i+1,j+1,k) v3=(i,j+1,k), v4..v7 the same at k+1) and
// its faces in ascending order (z low, z high, x low, x high, y low, y high): what the
// reference's rebuild_cell_topology + Hexahedron::reorder (HexahedronCell.h:10-37)
// give an axis-parallel box with lb < ub -- the bottom pair is the lowest in z, the
// polar sort about the centre starts at the (-,-) corner counter-clockwise, the top
// nodes align vertically, and the face list stays in ascending order.
#pragma once
#include "Mesh/GeneralMesh.h"

class OrthHexMesh : public GeneralMesh {
public:
    OrthHexMesh() = default;
    OrthHexMesh(vector3f lb, vector3f ub, vector3u Nxyz) {
        const uInt Nx = Nxyz[0], Ny = Nxyz[1], Nz = Nxyz[2];
        if (Nx == 0 || Ny == 0 || Nz == 0) throw std::runtime_error("OrthHexMesh: zero cells along an axis");
        for (int d = 0; d < 3; ++d)
            if (!(ub[d] > lb[d])) throw std::runtime_error("OrthHexMesh: ub must exceed lb");
        m_points.resize((Nx + 1) * (Ny + 1) * (Nz + 1));
        for (uInt k = 0; k <= Nz; ++k)
            for (uInt j = 0; j <= Ny; ++j)
                for (uInt i = 0; i <= Nx; ++i)
                    m_points[pid(i, j, k, Nx, Ny)] = {lb[0] + i * (ub[0] - lb[0]) / Nx,
                                                      lb[1] + j * (ub[1] - lb[1]) / Ny,
                                                      lb[2] + k * (ub[2] - lb[2]) / Nz};
        const uInt nzf = Nx * Ny * (Nz + 1), nxf = (Nx + 1) * Ny * Nz;
        auto cell = [&](uInt i, uInt j, uInt k) { return i + (j + k * Ny) * Nx; };
        auto zf = [&](uInt i, uInt j, uInt k) { return i + (j + k * Ny) * Nx; };
        auto xf = [&](uInt i, uInt j, uInt k) { return i + (j + k * Ny) * (Nx + 1) + nzf; };
        auto yf = [&](uInt i, uInt j, uInt k) { return i + (j + k * (Ny + 1)) * Nx + nzf + nxf; };
        m_faces.assign(nzf + nxf + Nx * (Ny + 1) * Nz, QuadFace{});
        auto quad = [&](uInt f, vector4u nodes, uInt c0, uInt c1) {
            auto& q = std::get<QuadFace>(m_faces[f]);
            q.nodes = nodes;
            q.neighbor_cells = {c0, c1};
        };
        for (uInt i = 0; i < Nx; ++i)
            for (uInt j = 0; j < Ny; ++j)
                for (uInt k = 0; k <= Nz; ++k)
                    quad(zf(i, j, k),
                         {pid(i, j, k, Nx, Ny), pid(i + 1, j, k, Nx, Ny), pid(i + 1, j + 1, k, Nx, Ny), pid(i, j + 1, k, Nx, Ny)},
                         k == 0 ? cell(i, j, 0) : cell(i, j, k - 1), (k == 0 || k == Nz) ? uInt(-1) : cell(i, j, k));
        for (uInt i = 0; i <= Nx; ++i)
            for (uInt j = 0; j < Ny; ++j)
                for (uInt k = 0; k < Nz; ++k)
                    quad(xf(i, j, k),
                         {pid(i, j, k, Nx, Ny), pid(i, j + 1, k, Nx, Ny), pid(i, j + 1, k + 1, Nx, Ny), pid(i, j, k + 1, Nx, Ny)},
                         i == 0 ? cell(0, j, k) : cell(i - 1, j, k), (i == 0 || i == Nx) ? uInt(-1) : cell(i, j, k));
        for (uInt i = 0; i < Nx; ++i)
            for (uInt j = 0; j <= Ny; ++j)
                for (uInt k = 0; k < Nz; ++k)
                    quad(yf(i, j, k),
                         {pid(i, j, k, Nx, Ny), pid(i, j, k + 1, Nx, Ny), pid(i + 1, j, k + 1, Nx, Ny), pid(i + 1, j, k, Nx, Ny)},
                         j == 0 ? cell(i, 0, k) : cell(i, j - 1, k), (j == 0 || j == Ny) ? uInt(-1) : cell(i, j, k));
        m_cells.reserve(Nx * Ny * Nz);
        for (uInt k = 0; k < Nz; ++k)
            for (uInt j = 0; j < Ny; ++j)
                for (uInt i = 0; i < Nx; ++i) {
                    Hexahedron h;
                    h.nodes = {pid(i, j, k, Nx, Ny),         pid(i + 1, j, k, Nx, Ny),
                               pid(i + 1, j + 1, k, Nx, Ny), pid(i, j + 1, k, Nx, Ny),
                               pid(i, j, k + 1, Nx, Ny),     pid(i + 1, j, k + 1, Nx, Ny),
                               pid(i + 1, j + 1, k + 1, Nx, Ny), pid(i, j + 1, k + 1, Nx, Ny)};
                    h.faces = {zf(i, j, k), zf(i, j, k + 1), xf(i, j, k), xf(i + 1, j, k), yf(i, j, k), yf(i, j + 1, k)};
                    h.parity = (i + j + k) & 1;
                    m_cells.push_back(h);
                }
        rebuild_cell_topology();
        validate_mesh();
    }

private:
    static uInt pid(uInt i, uInt j, uInt k, uInt Nx, uInt Ny) { return i + (j + k * (Ny + 1)) * (Nx + 1); }
};
