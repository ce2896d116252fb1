// Umbrella header (reference: include/Matrix/Matrix.h).  The sparse/Krylov
// part of the reference's Matrix library serves its implicit solver and is
// not provided here.
#pragma once
#include "base/Type.h"
#include "Matrix/DenseMatrix.h"
#include "Matrix/LongVector.h"
