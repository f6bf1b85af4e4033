#include "hostla.hpp"

#include <algorithm>
#include <cmath>

namespace slemm {

bool thin_qr(const MatD& A, MatD& Q, MatD& R, double tol) {
  const int64_t n = A.n;
  const int c = A.c;
  MatD W = A;                           // reflected in place
  std::vector<std::vector<double>> vs;  // Householder vectors (length n - k)
  std::vector<double> betas;
  R = MatD(c, c);
  for (int k = 0; k < c; ++k) {
    double norm2 = 0.0;
    for (int64_t i = k; i < n; ++i) norm2 += W(i, k) * W(i, k);
    const double norm = std::sqrt(norm2);
    const double x0 = W(k, k);
    const double alpha = x0 > 0 ? -norm : norm;
    std::vector<double> v(n - k);
    for (int64_t i = k; i < n; ++i) v[i - k] = W(i, k);
    v[0] -= alpha;
    double vn2 = 0.0;
    for (double e : v) vn2 += e * e;
    const double beta = vn2 > 0 ? 2.0 / vn2 : 0.0;
    for (int j = k; j < c; ++j) {
      double s = 0.0;
      for (int64_t i = k; i < n; ++i) s += v[i - k] * W(i, j);
      s *= beta;
      for (int64_t i = k; i < n; ++i) W(i, j) -= s * v[i - k];
    }
    vs.push_back(std::move(v));
    betas.push_back(beta);
  }
  double rmax = 0.0;
  for (int k = 0; k < c; ++k) {
    for (int j = k; j < c; ++j) R(k, j) = W(k, j);
    rmax = std::max(rmax, std::fabs(R(k, k)));
  }
  for (int k = 0; k < c; ++k)
    if (!(std::fabs(R(k, k)) > tol * rmax)) return false;
  // Q = H_0 H_1 ... H_{c-1} [I_c; 0]
  Q = MatD(n, c);
  for (int j = 0; j < c; ++j) Q(j, j) = 1.0;
  for (int k = c - 1; k >= 0; --k) {
    const auto& v = vs[k];
    for (int j = 0; j < c; ++j) {
      double s = 0.0;
      for (int64_t i = k; i < n; ++i) s += v[i - k] * Q(i, j);
      s *= betas[k];
      for (int64_t i = k; i < n; ++i) Q(i, j) -= s * v[i - k];
    }
  }
  return true;
}

bool chol_solve(std::vector<double> S, int c, std::vector<double>& b) {
  for (int j = 0; j < c; ++j) {
    double d = S[j + j * c];
    for (int k = 0; k < j; ++k) d -= S[j + k * c] * S[j + k * c];
    if (!(d > 0)) return false;
    d = std::sqrt(d);
    S[j + j * c] = d;
    for (int i = j + 1; i < c; ++i) {
      double s = S[i + j * c];
      for (int k = 0; k < j; ++k) s -= S[i + k * c] * S[j + k * c];
      S[i + j * c] = s / d;
    }
  }
  for (int i = 0; i < c; ++i) {  // L y = b
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= S[i + k * c] * b[k];
    b[i] = s / S[i + i * c];
  }
  for (int i = c - 1; i >= 0; --i) {  // L' x = y
    double s = b[i];
    for (int k = i + 1; k < c; ++k) s -= S[k + i * c] * b[k];
    b[i] = s / S[i + i * c];
  }
  return true;
}

bool invert(std::vector<double> S, int c, std::vector<double>& inv) {
  inv.assign((size_t)c * c, 0.0);
  for (int i = 0; i < c; ++i) inv[i + i * c] = 1.0;
  for (int col = 0; col < c; ++col) {
    int piv = col;
    for (int r = col + 1; r < c; ++r)
      if (std::fabs(S[r + col * c]) > std::fabs(S[piv + col * c])) piv = r;
    if (S[piv + col * c] == 0.0) return false;
    if (piv != col)
      for (int k = 0; k < c; ++k) {
        std::swap(S[col + k * c], S[piv + k * c]);
        std::swap(inv[col + k * c], inv[piv + k * c]);
      }
    const double d = S[col + col * c];
    for (int k = 0; k < c; ++k) {
      S[col + k * c] /= d;
      inv[col + k * c] /= d;
    }
    for (int r = 0; r < c; ++r) {
      if (r == col) continue;
      const double f = S[r + col * c];
      if (f == 0.0) continue;
      for (int k = 0; k < c; ++k) {
        S[r + k * c] -= f * S[col + k * c];
        inv[r + k * c] -= f * inv[col + k * c];
      }
    }
  }
  return true;
}

double log_abs_det(std::vector<double> S, int c) {
  double acc = 0.0;
  for (int col = 0; col < c; ++col) {
    int piv = col;
    for (int r = col + 1; r < c; ++r)
      if (std::fabs(S[r + col * c]) > std::fabs(S[piv + col * c])) piv = r;
    if (S[piv + col * c] == 0.0) return -INFINITY;
    if (piv != col)
      for (int k = 0; k < c; ++k) std::swap(S[col + k * c], S[piv + k * c]);
    const double d = S[col + col * c];
    acc += std::log(std::fabs(d));
    for (int r = col + 1; r < c; ++r) {
      const double f = S[r + col * c] / d;
      for (int k = col; k < c; ++k) S[r + k * c] -= f * S[col + k * c];
    }
  }
  return acc;
}

}  // namespace slemm
