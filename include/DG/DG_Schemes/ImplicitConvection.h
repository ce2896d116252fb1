// Picard-linearised implicit DG Euler operator, assembled on the GPU.
//
// Same class template, method names and signatures as the reference's
// include/DG/DG_Schemes/ImplicitConvection.h:10-54:
//
//     ImplicitConvection<3> convection;
//     BlockSparseMatrix<5*N, 5*N> sparse_mat;  LongVector<5*N> rhs(n_cells);
//     convection.assemble(cmesh, dx, t, sparse_mat, rhs);   // R(U) ~ A(dx) U - rhs
//     sparse_mat.add_block(e, e, DenseMatrix<5*N,5*N>::Diag(mass[e] / dt));  ... finalize(); solve
//
// What is assembled (src/DG/DG_Schemes/ImplicitConvection.cpp): per element the block
// -sum_q w_q |det| phi_i (sum_d J_d(U_q) d phi_j/dx_d) (assemble_cells, :266-325); per interior face
// the four blocks of 1/2 (F_n(U_L) + lambda I) and 1/2 (F_n(U_R) - lambda I) between the two
// elements' traces (assemble_internals, :329-420); per boundary face the linearisation around the
// reference's ghost state (Dirichlet, WallTD, WallTN from the application's rho_Xi .. rhoe_Xi hooks
// of base/exact.h at curr_time; Pseudo3DX/Y/Z mirrors) with its remainder subtracted from the
// right-hand side (assemble_boundarys, :20-167).  J_d = dF_d/dU (EulerPhysicalFlux.cpp:21-42).
// Blocks are added to the caller's matrix with add_block (summed per position, as the reference's
// BlockSparseMatrix does) and the remainders added to the caller's right-hand side.
//
// The blocks are computed on the GPU (fvdg_implicit_assemble) and copied to the host: 5 blocks of
// (5N)^2 doubles per element (400 KB at P3), so host assembly suits the meshes the reference's
// drivers use.  For large meshes keep the system on the device: fvdg_implicit_euler_step, or
// ExplicitSSPRK3-style use of the C ABI (fvdg_implicit_solve).
#pragma once
#include <memory>
#include <vector>

#include "DG/DG_Schemes/ExplicitConvection.h"
#include "Matrix/SparseMatrix.h"

template <uInt Order = 3, typename Flux = AirFluxC, typename GaussQuadCell = GaussLegendreTet::Auto,
          typename GaussQuadFace = GaussLegendreTri::Auto>
class ImplicitConvection {
private:
    using Basis = DGBasisEvaluator<Order>;
    static constexpr uInt N = Basis::NumBasis;
    static constexpr uInt NB = 5 * N;

public:
    void assemble(const ComputingMesh& mesh, const LongVector<NB>& old_solution, const Scalar curr_time,
                  BlockSparseMatrix<NB, NB>& sparse_mat, LongVector<NB>& sparse_rhs) {
        run(mesh, old_solution, curr_time, FVDG_PART_ALL, sparse_mat, sparse_rhs);
    }
    void assemble_cells(const ComputingMesh& mesh, const LongVector<NB>& old_solution, const Scalar curr_time,
                        BlockSparseMatrix<NB, NB>& sparse_mat, LongVector<NB>& sparse_rhs) {
        run(mesh, old_solution, curr_time, FVDG_PART_CELLS, sparse_mat, sparse_rhs);
    }
    void assemble_internals(const ComputingMesh& mesh, const LongVector<NB>& old_solution, const Scalar curr_time,
                            BlockSparseMatrix<NB, NB>& sparse_mat, LongVector<NB>& sparse_rhs) {
        run(mesh, old_solution, curr_time, FVDG_PART_INTERNALS, sparse_mat, sparse_rhs);
    }
    void assemble_boundarys(const ComputingMesh& mesh, const LongVector<NB>& old_solution, const Scalar curr_time,
                            BlockSparseMatrix<NB, NB>& sparse_mat, LongVector<NB>& sparse_rhs) {
        run(mesh, old_solution, curr_time, FVDG_PART_BOUNDARYS, sparse_mat, sparse_rhs);
    }

    // the explicit operator that owns the device context and the boundary hooks
    // (set_boundary_function replaces the link-time hooks)
    ExplicitConvection<Order, Flux, GaussQuadCell, GaussQuadFace>& explicit_operator() { return conv_; }

private:
    void run(const ComputingMesh& mesh, const LongVector<NB>& U, Scalar t, int parts,
             BlockSparseMatrix<NB, NB>& mat, LongVector<NB>& rhs) {
        const size_t E = mesh.m_cells.size();
        if (U.size() != E) throw std::runtime_error("ImplicitConvection: state size != cell count");
        if (rhs.size() != E) rhs.resize(E);
        fvdg::Session& s = conv_.session(mesh);
        if (parts & FVDG_PART_BOUNDARYS) conv_.update_boundary_states(mesh, t);
        s.upload(U.data());
        fvdg::check(fvdg_implicit_assemble(s.ctx(), parts), "fvdg_implicit_assemble");
        const size_t bb = (size_t)NB * NB;
        std::vector<double> diag(E * bb), off(E * 4 * bb), r(E * NB);
        std::vector<int32_t> cols(E * 4);
        fvdg::check(fvdg_implicit_blocks(s.ctx(), diag.data(), off.data(), cols.data(), r.data()),
                    "fvdg_implicit_blocks");
        auto blk = std::make_unique<DenseMatrix<NB, NB>>();
        for (size_t e = 0; e < E; ++e) {
            std::copy(diag.begin() + e * bb, diag.begin() + (e + 1) * bb, blk->data_ptr());
            mat.add_block((uInt)e, (uInt)e, *blk);
            for (int l = 0; l < 4; ++l) {
                if (cols[e * 4 + l] < 0) continue;
                std::copy(off.begin() + (e * 4 + l) * bb, off.begin() + (e * 4 + l + 1) * bb, blk->data_ptr());
                mat.add_block((uInt)e, (uInt)cols[e * 4 + l], *blk);
            }
            for (uInt k = 0; k < NB; ++k) rhs[e][k] += r[e * NB + k];
        }
    }

    ExplicitConvection<Order, Flux, GaussQuadCell, GaussQuadFace> conv_;
};
