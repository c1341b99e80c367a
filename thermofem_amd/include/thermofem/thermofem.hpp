// Umbrella header of the C++20 drop-in API: the reference's public headers
// (DataStructure/*, LinearAlgebra/{IterationContext,Diagonal,Operations,
// Operators}, Multigrid/*, Smoother/Jacobi, Preconditioner/{Amg,Identity},
// Solver/*, IO/MatrixMarket) in one include over libthermofem_hip.so.
#pragma once

#include "Memory.hpp"
#include "CsrMatrix.hpp"
#include "DenseMatrix.hpp"
#include "LinearAlgebra.hpp"
#include "Operations.hpp"
#include "Multigrid.hpp"
#include "Solver.hpp"
#include "MatrixMarket.hpp"
