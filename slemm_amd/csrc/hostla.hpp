// Small dense FP64 linear algebra for the REML drivers (covariate QR, c x c solves).
// Sizes are n x c with c = number of covariates (a handful), so plain loops suffice.
#pragma once

#include <cmath>
#include <vector>

namespace slemm {

// Column-major n x c matrix in FP64.
struct MatD {
  int64_t n = 0;
  int c = 0;
  std::vector<double> a;
  MatD() = default;
  MatD(int64_t rows, int cols) : n(rows), c(cols), a((size_t)rows * cols, 0.0) {}
  double& operator()(int64_t i, int j) { return a[i + (size_t)j * n]; }
  double operator()(int64_t i, int j) const { return a[i + (size_t)j * n]; }
};

// Thin QR of A (n x c) by Householder reflections: Q (n x c, orthonormal columns) and R
// (c x c upper triangular) with A = Q R.  Returns false if A is rank deficient, judged by
// |R_kk| <= tol * max_k |R_kk|.
bool thin_qr(const MatD& A, MatD& Q, MatD& R, double tol);

// Solves S x = b for symmetric positive definite S (c x c, column-major) by Cholesky.
bool chol_solve(std::vector<double> S, int c, std::vector<double>& b);

// Inverse of a c x c matrix (Gauss-Jordan with partial pivoting).
bool invert(std::vector<double> S, int c, std::vector<double>& inv);

// log |det S| of a c x c matrix (LU with partial pivoting).
double log_abs_det(std::vector<double> S, int c);

}  // namespace slemm
