// Linear solver for a BlockSparseMatrix with the reference's interface
// (include/EigenSolver/EigenSparseSolver.h, src/EigenSolver/EigenSparseSolver.cpp), solving on the GPU.
//
//     EigenSparseSolver<DoFs, DoFs> solver(sparse_mat, rhs);        // sparse_mat finalized
//     LongVector<DoFs> x = solver.DGMRES(x0);                      // Sedov.cpp:179-180
//
// Every method runs restarted GMRES (Saad & Schultz 1986) with a block-Jacobi preconditioner (the
// inverse or LU factors of each diagonal block) on the device (fvdg_bsr_solve) from the initial guess x0, to the relative
// residual |b - A x| / |b| < tol (set_tol, default 1e-12) within set_iterations iterations (default
// twice the number of unknowns) and restart length set_restart (default 30).  The reference's DGMRES
// is Eigen's deflated GMRES with a diagonal preconditioner; SparseLU and BiCGSTAB are answered with
// the same GMRES, so every method returns the solution to the tolerance rather than its own
// iterates.  Errors (singular diagonal block, non-finite values) throw std::runtime_error.
#pragma once
#include <algorithm>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "DG/DG_Schemes/DeviceSession.h"
#include "Matrix/SparseMatrix.h"

template <uInt BlockRows, uInt BlockCols>
class EigenSparseSolver {
    static_assert(BlockRows == BlockCols, "square blocks");

public:
    EigenSparseSolver() = default;
    explicit EigenSparseSolver(const BlockSparseMatrix<BlockRows, BlockCols>& mat) { set_matrix(mat); }
    EigenSparseSolver(const BlockSparseMatrix<BlockRows, BlockCols>& mat, const LongVector<BlockRows>& rhs) {
        set_matrix(mat);
        set_rhs(rhs);
    }

    LongVector<BlockCols> SparseLU(const LongVector<BlockCols>& x0) { return solve(x0, nullptr, nullptr); }
    LongVector<BlockCols> BiCGSTAB(const LongVector<BlockCols>& x0) { return solve(x0, nullptr, nullptr); }
    LongVector<BlockCols> DGMRES(const LongVector<BlockCols>& x0) { return solve(x0, nullptr, nullptr); }
    LongVector<BlockCols> DGMRES(const LongVector<BlockCols>& x0, uInt& iter, Scalar& residual) {
        int it = 0;
        double rr = 0;
        auto x = solve(x0, &it, &rr);
        iter = (uInt)it;
        residual = rr;
        return x;
    }
    std::tuple<uInt, Scalar> DGMRES(const LongVector<BlockCols>& x0, LongVector<BlockCols>& ret) {
        int it = 0;
        double rr = 0;
        ret = solve(x0, &it, &rr);
        return {(uInt)it, rr};
    }

    void set_tol(Scalar tol = 1e-12) { m_tol = tol; }
    void set_iterations(uInt it) { m_maxiters = it; }
    void set_restart(uInt it = 30) { m_restart = it; }
    void set_rhs(const LongVector<BlockRows>& rhs) {
        m_rhs.assign(rhs.data(), rhs.data() + (size_t)rhs.size() * BlockRows);
    }

private:
    void set_matrix(const BlockSparseMatrix<BlockRows, BlockCols>& mat) {
        if (!mat.finalized) throw std::runtime_error("EigenSparseSolver: the matrix must be finalized");
        m_rows = mat.num_block_rows;
        m_row_ptr.assign(mat.row_ptr().begin(), mat.row_ptr().end());
        m_cols.assign(mat.cols().begin(), mat.cols().end());
        m_values = mat.values();
    }
    LongVector<BlockCols> solve(const LongVector<BlockCols>& x0, int* it, double* rr) {
        if (m_rhs.size() != (size_t)m_rows * BlockRows) throw std::runtime_error("EigenSparseSolver: rhs size");
        LongVector<BlockCols> x = x0;
        if (x.size() != m_rows) x.resize(m_rows);
        const int64_t n = (int64_t)m_rows * BlockRows;
        const int maxit = m_maxiters == uInt(-1) ? (int)std::min<int64_t>(2 * n, 1 << 30) : (int)m_maxiters;
        const int restart = m_restart == uInt(-1) ? 30 : (int)std::min<uInt>(m_restart, 63);
        fvdg::check(fvdg_bsr_solve(fvdg::default_device(), m_rows, (int)BlockRows, m_row_ptr.data(), m_cols.data(),
                                   m_values.data(), m_rhs.data(), x.data(), m_tol, restart, maxit, it, rr),
                    "fvdg_bsr_solve");
        return x;
    }

    uInt m_rows = 0;
    std::vector<int64_t> m_row_ptr, m_cols;
    std::vector<Scalar> m_values, m_rhs;
    Scalar m_tol = 1e-12;
    uInt m_maxiters = uInt(-1);
    uInt m_restart = uInt(-1);
};
