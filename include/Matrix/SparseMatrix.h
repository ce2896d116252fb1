// Block-sparse matrix with the reference's interface (include/Matrix/SparseMatrix.h:7-125):
// blocks are added by (row, col) and summed when a position repeats (add_block), fixed by
// finalize(), and applied by multiply().  Storage: an ordered map while assembling, then block
// CSR (rows ascending, columns ascending within a row), which is also the form handed to the GPU
// solver (EigenSparseSolver.h).
#pragma once
#include <map>
#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

#include "Matrix/DenseMatrix.h"
#include "Matrix/LongVector.h"

template <uInt BlockRows, uInt BlockCols>
class BlockSparseMatrix {
public:
    using BlockType = DenseMatrix<BlockRows, BlockCols>;

    void add_block(uInt row, uInt col, const BlockType& blk) {
        if (finalized) throw std::runtime_error("BlockSparseMatrix: add_block after finalize");
        auto it = coo_.find({row, col});
        if (it == coo_.end()) coo_.emplace(std::make_pair(row, col), std::make_unique<BlockType>(blk));
        else *it->second += blk;
        if (row + 1 > num_block_rows) num_block_rows = row + 1;
    }

    void finalize() {
        if (finalized) throw std::runtime_error("BlockSparseMatrix: finalize called twice");
        row_ptr_.assign(num_block_rows + 1, 0);
        for (const auto& kv : coo_) ++row_ptr_[kv.first.first + 1];
        for (uInt r = 0; r < num_block_rows; ++r) row_ptr_[r + 1] += row_ptr_[r];
        cols_.clear();
        values_.clear();
        cols_.reserve(coo_.size());
        values_.reserve(coo_.size() * BlockRows * BlockCols);
        for (const auto& kv : coo_) {  // map order: rows, then columns, ascending
            cols_.push_back(kv.first.second);
            const Scalar* p = kv.second->data_ptr();
            values_.insert(values_.end(), p, p + BlockRows * BlockCols);
        }
        coo_.clear();
        finalized = true;
    }

    LongVector<BlockRows> multiply(const LongVector<BlockCols>& vec) const {
        if (!finalized) throw std::runtime_error("BlockSparseMatrix: multiply before finalize");
        LongVector<BlockRows> out(num_block_rows);
        for (uInt r = 0; r < num_block_rows; ++r)
            for (uInt k = row_ptr_[r]; k < row_ptr_[r + 1]; ++k) {
                const Scalar* B = values_.data() + (size_t)k * BlockRows * BlockCols;
                const auto& x = vec[cols_[k]];
                auto& y = out[r];
                for (uInt i = 0; i < BlockRows; ++i) {
                    Scalar s = 0;
                    for (uInt j = 0; j < BlockCols; ++j) s += B[i * BlockCols + j] * x[j];
                    y[i] += s;
                }
            }
        return out;
    }

    // block CSR after finalize(): row_ptr (num_block_rows + 1), column per block, row-major blocks
    const std::vector<uInt>& row_ptr() const { return row_ptr_; }
    const std::vector<uInt>& cols() const { return cols_; }
    const std::vector<Scalar>& values() const { return values_; }

    uInt num_block_rows = 0;
    bool finalized = false;

private:
    std::map<std::pair<uInt, uInt>, std::unique_ptr<BlockType>> coo_;
    std::vector<uInt> row_ptr_, cols_;
    std::vector<Scalar> values_;
};
