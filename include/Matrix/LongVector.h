// Block vector of DenseMatrix<B,1>: the DG state container, one block of 5*N
// coefficients per element with index 5*k + v (mode k, conserved variable v).
// API-compatible subset of the reference's include/Matrix/LongVector.h:7-61.
#pragma once
#include "Matrix/DenseMatrix.h"

template <uInt BlockSize>
class LongVector {
public:
    std::vector<DenseMatrix<BlockSize, 1>> blocks;

    LongVector() = default;
    explicit LongVector(uInt length) : blocks(length, DenseMatrix<BlockSize, 1>(0.0)) {}

    DenseMatrix<BlockSize, 1>& operator[](uInt i) { return blocks[i]; }
    const DenseMatrix<BlockSize, 1>& operator[](uInt i) const { return blocks[i]; }
    uInt size() const { return blocks.size(); }
    void resize(uInt n) { blocks.resize(n, DenseMatrix<BlockSize, 1>(0.0)); }
    void fill_scalar(Scalar v) {
        for (auto& b : blocks) b = v;
    }
    // Contiguous view of all coefficients: element e, entry 5k+v at [e*BlockSize + 5k+v].
    Scalar* data() { return blocks.empty() ? nullptr : blocks[0].data_ptr(); }
    const Scalar* data() const { return blocks.empty() ? nullptr : blocks[0].data_ptr(); }

#define FVDG_LV_OP(op)                                                        \
    LongVector& operator op##=(const LongVector& r) {                         \
        for (uInt i = 0; i < blocks.size(); ++i) blocks[i] op## = r.blocks[i]; \
        return *this;                                                         \
    }                                                                         \
    LongVector operator op(const LongVector& r) const {                       \
        LongVector t(*this);                                                  \
        return t op## = r;                                                    \
    }                                                                         \
    LongVector& operator op##=(Scalar v) {                                    \
        for (auto& b : blocks) b op## = v;                                    \
        return *this;                                                         \
    }                                                                         \
    LongVector operator op(Scalar v) const {                                  \
        LongVector t(*this);                                                  \
        return t op## = v;                                                    \
    }                                                                         \
    friend LongVector operator op(Scalar v, LongVector r) {                   \
        for (auto& b : r.blocks) b = v op b;                                  \
        return r;                                                             \
    }
    FVDG_LV_OP(+)
    FVDG_LV_OP(-)
    FVDG_LV_OP(*)
    FVDG_LV_OP(/)
#undef FVDG_LV_OP
    LongVector operator-() const {
        LongVector t(*this);
        for (auto& b : t.blocks) b = -b;
        return t;
    }

    Scalar dot(const LongVector& r) const {
        Scalar s = 0;
        for (uInt i = 0; i < blocks.size(); ++i) s += blocks[i].dot(r.blocks[i]);
        return s;
    }
    Scalar norm() const { return std::sqrt(dot(*this)); }
};

static_assert(sizeof(DenseMatrix<5, 1>) == 5 * sizeof(Scalar), "state blocks must be tightly packed");
