// Umbrella header (reference: include/Matrix/Matrix.h).
#pragma once
#include "base/Type.h"
#include "Matrix/DenseMatrix.h"
#include "Matrix/LongVector.h"
#include "Matrix/SparseMatrix.h"
