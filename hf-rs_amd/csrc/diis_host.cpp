// diis_host.cpp -- the small dense solves of the DIIS extrapolations (the
// CDIIS bordered system, the ADIIS model's minimiser on the probability
// simplex), host code shared by the RHF drivers and the CPU check library.
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "scf_host.h"

namespace hfg {

// Projected-gradient minimisation of 1/2 c^T Q c + b^T c on the probability
// simplex, fixed step 1/L with L the Gershgorin bound of sym(Q)
// (src/diis.rs:195-266); projection by the sort-and-threshold rule.
void simplex_project(std::vector<double>& v) {
  std::vector<double> u = v;
  std::sort(u.begin(), u.end(), std::greater<double>());
  double css = 0, theta = 0;
  for (size_t k = 0; k < u.size(); ++k) {
    css += u[k];
    const double t = (css - 1.0) / (double)(k + 1);
    if (u[k] - t > 0) theta = t;
  }
  for (double& x : v) x = std::max(x - theta, 0.0);
}
std::vector<double> simplex_minimise(const std::vector<double>& q, const std::vector<double>& b) {
  const int m = (int)b.size();
  double lip = 1e-12;
  for (int i = 0; i < m; ++i) {
    double s = 0;
    for (int j = 0; j < m; ++j) s += std::fabs(0.5 * (q[(size_t)i * m + j] + q[(size_t)j * m + i]));
    lip = std::max(lip, s);
  }
  std::vector<double> c(m, 0.0), nx(m);
  c[m - 1] = 1.0;
  simplex_project(c);
  for (int it = 0; it < 5000; ++it) {
    for (int i = 0; i < m; ++i) {
      double g = b[i];
      for (int j = 0; j < m; ++j) g += 0.5 * (q[(size_t)i * m + j] + q[(size_t)j * m + i]) * c[j];
      nx[i] = c[i] - g / lip;
    }
    simplex_project(nx);
    double d = 0;
    for (int i = 0; i < m; ++i) d = std::max(d, std::fabs(nx[i] - c[i]));
    c = nx;
    if (d < 1e-14) break;
  }
  return c;
}

// Gaussian elimination with partial pivoting (src/diis.rs:532-589); false if singular.
bool solve_small(std::vector<double> a, std::vector<double> b, int n, std::vector<double>& x) {
  for (int i = 0; i < n; ++i) {
    int piv = i;
    for (int k = i + 1; k < n; ++k)
      if (std::fabs(a[(size_t)k * n + i]) > std::fabs(a[(size_t)piv * n + i])) piv = k;
    if (std::fabs(a[(size_t)piv * n + i]) < 1e-28) return false;
    if (piv != i) {
      std::swap(b[i], b[piv]);
      for (int c = 0; c < n; ++c) std::swap(a[(size_t)i * n + c], a[(size_t)piv * n + c]);
    }
    for (int j = i + 1; j < n; ++j) {
      const double f = a[(size_t)j * n + i] / a[(size_t)i * n + i];
      for (int k = i; k < n; ++k) a[(size_t)j * n + k] -= f * a[(size_t)i * n + k];
      b[j] -= f * b[i];
    }
  }
  x.assign(n, 0.0);
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < n; ++k) s -= a[(size_t)i * n + k] * x[k];
    x[i] = s / a[(size_t)i * n + i];
  }
  return true;
}

}  // namespace hfg
