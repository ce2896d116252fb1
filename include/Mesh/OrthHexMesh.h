// Structured box of Nx*Ny*Nz hexahedra (reference: include/Mesh/OrthHexMesh.h:3-106).
// Point (i,j,k) has index i + (j + k*(Ny+1))*(Nx+1) and position
// lb + (i,j,k) * (ub-lb)/(Nx,Ny,Nz), as in the reference.  Cells are
// Hexahedron with corners in the v0..v7 order documented in GeneralMesh.h;
// faces are produced when the hexahedra are split into tetrahedra.
#pragma once
#include "Mesh/GeneralMesh.h"

class OrthHexMesh : public GeneralMesh {
public:
    OrthHexMesh() = default;
    OrthHexMesh(vector3f lb, vector3f ub, vector3u Nxyz) {
        const uInt Nx = Nxyz[0], Ny = Nxyz[1], Nz = Nxyz[2];
        if (Nx == 0 || Ny == 0 || Nz == 0) throw std::runtime_error("OrthHexMesh: zero cells along an axis");
        m_points.resize((Nx + 1) * (Ny + 1) * (Nz + 1));
        for (uInt k = 0; k <= Nz; ++k)
            for (uInt j = 0; j <= Ny; ++j)
                for (uInt i = 0; i <= Nx; ++i)
                    m_points[pid(i, j, k, Nx, Ny)] = {lb[0] + i * (ub[0] - lb[0]) / Nx,
                                                      lb[1] + j * (ub[1] - lb[1]) / Ny,
                                                      lb[2] + k * (ub[2] - lb[2]) / Nz};
        m_cells.reserve(Nx * Ny * Nz);
        for (uInt k = 0; k < Nz; ++k)
            for (uInt j = 0; j < Ny; ++j)
                for (uInt i = 0; i < Nx; ++i) {
                    Hexahedron h;
                    h.nodes = {pid(i, j, k, Nx, Ny),         pid(i + 1, j, k, Nx, Ny),
                               pid(i + 1, j + 1, k, Nx, Ny), pid(i, j + 1, k, Nx, Ny),
                               pid(i, j, k + 1, Nx, Ny),     pid(i + 1, j, k + 1, Nx, Ny),
                               pid(i + 1, j + 1, k + 1, Nx, Ny), pid(i, j + 1, k + 1, Nx, Ny)};
                    h.parity = (i + j + k) & 1;
                    m_cells.push_back(h);
                }
        rebuild_cell_topology();
        validate_mesh();
    }

private:
    static uInt pid(uInt i, uInt j, uInt k, uInt Nx, uInt Ny) { return i + (j + k * (Ny + 1)) * (Nx + 1); }
};
