// Implicit-shift QL iteration on a symmetric tridiagonal matrix (first eigenvector row only).
#include "tridiag.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "common.hpp"

namespace slemm {

void tridiag_eig_first_row(int J, const double* d_in, const double* e_in, double* D, double* W) {
  if (J <= 0) return;
  std::vector<double> d(d_in, d_in + J), e(J, 0.0), z(J, 0.0);
  for (int i = 0; i + 1 < J; ++i) e[i] = e_in[i];
  z[0] = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < J; ++l) {
    int iter = 0;
    for (;;) {
      int m = l;
      for (; m < J - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (++iter > 100) fail(SLEMM_ERR_DATA, "tridiagonal eigenvalue iteration did not converge");
      // Wilkinson-type shift from the leading 2x2 block
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      int i = m - 1;
      bool early = false;
      for (; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {  // split: undo the pending shift and restart
          d[i + 1] -= p;
          e[m] = 0.0;
          early = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        const double zf = z[i + 1];
        z[i + 1] = s * z[i] + c * zf;
        z[i] = c * z[i] - s * zf;
      }
      if (early) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> idx(J);
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  for (int q = 0; q < J; ++q) {
    D[q] = d[idx[q]];
    W[q] = z[idx[q]];
  }
}

}  // namespace slemm
