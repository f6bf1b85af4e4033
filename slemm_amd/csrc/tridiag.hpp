// Symmetric tridiagonal eigenproblem for stochastic Lanczos quadrature.
#pragma once

namespace slemm {

// Eigenvalues D (ascending) and first components W of the unit eigenvectors of the J x J
// symmetric tridiagonal matrix with diagonal d[0..J) and off-diagonal e[0..J-1), in FP64.
// Implicit QL with Wilkinson shifts, accumulating only the first row of the eigenvector
// matrix (Golub-Welsch).  Throws slemm::Error if an eigenvalue fails to converge.
void tridiag_eig_first_row(int J, const double* d, const double* e, double* D, double* W);

}  // namespace slemm
