// onee.cpp -- host one-electron matrices S, T, V for the atomic SAD solves
// (a few dozen functions each; the molecules' matrices are built on the
// device, onee_dev.hip) and the real spherical transform.
//
// Obara-Saika recurrences: 1-D overlaps S_ij with S_{i+1,j} = X_PA S_ij +
// (i S_{i-1,j} + j S_{i,j-1})/(2p); kinetic t_ij = b(2j+1) S_ij - 2b^2 S_{i,j+2}
// - j(j-1)/2 S_{i,j-2}; nuclear attraction by the vertical recurrence
// Th^m[a+1_i] = PA_i Th^m[a] - PC_i Th^{m+1}[a] + a_i/(2p)(Th^m[a-1_i] - Th^{m+1}[a-1_i])
// from Th^m[0] = 2 pi/p exp(-mu AB^2) F_m(p PC^2), contracted, then the
// horizontal recurrence (a, b+1_i) = (a+1_i, b) + AB_i (a, b).  These replace
// compute_one_electron_integrals (src/scf.rs:247-271), whose reference
// evaluation is McMurchie-Davidson (src/integrals.rs:407-537, 705-793).
#include <array>
#include <cmath>
#include <functional>
#include <map>
#include <vector>

#include "eri_device.h"
#include "host_prep.h"
#include "scf_host.h"

namespace hfg {

static const std::vector<double>& host_boys() {
  static std::vector<double> t = boys_table();
  return t;
}

static std::vector<std::array<int, 3>> comps(int L) {
  if (L == 2) return {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
  std::vector<std::array<int, 3>> o;
  for (int lx = L; lx >= 0; --lx)
    for (int ly = L - lx; ly >= 0; --ly) o.push_back({lx, ly, L - lx - ly});
  return o;
}

// Shell-pair block of S, T, V in (component a, component b) order, unscaled.
static void pair_block(const Shell& A, const Shell& B, const std::vector<std::array<double, 4>>& atoms,
                       std::vector<double>& s, std::vector<double>& t, std::vector<double>& v) {
  const int la = A.L, lb = B.L, Lt = la + lb;
  auto ca = comps(la), cb = comps(lb);
  const int na = (int)ca.size(), nb = (int)cb.size();
  s.assign(na * nb, 0.0);
  t.assign(na * nb, 0.0);
  v.assign(na * nb, 0.0);
  const double PI = 3.14159265358979323846;
  double AB[3], ab2 = 0;
  for (int k = 0; k < 3; ++k) {
    AB[k] = A.c[k] - B.c[k];
    ab2 += AB[k] * AB[k];
  }
  const int D = Lt + 1;  // index stride for [x][y][z]
  std::vector<double> ve((size_t)D * D * D, 0.0);  // contracted V_[e|0]
  std::vector<double> th((size_t)(Lt + 1) * D * D * D);
  double F[64];
  for (int ka = 0; ka < A.K; ++ka)
    for (int kb = 0; kb < B.K; ++kb) {
      const double a = A.e[ka], b = B.e[kb], p = a + b, mu = a * b / p, o2p = 0.5 / p;
      const double cc = A.r[ka] * B.r[kb];
      const double eab = std::exp(-mu * ab2);
      double P[3], PA[3], PB[3];
      for (int k = 0; k < 3; ++k) {
        P[k] = (a * A.c[k] + b * B.c[k]) / p;
        PA[k] = P[k] - A.c[k];
        PB[k] = P[k] - B.c[k];
      }
      // 1-D overlaps up to (la, lb+2)
      double S1[3][8][10];
      for (int ax = 0; ax < 3; ++ax) {
        S1[ax][0][0] = 1.0;
        for (int i = 0; i <= la; ++i)
          for (int j = 0; j <= lb + 2; ++j) {
            if (i == 0 && j == 0) continue;
            double val;
            if (i > 0) {
              val = PA[ax] * S1[ax][i - 1][j];
              if (i > 1) val += (i - 1) * o2p * S1[ax][i - 2][j];
              if (j > 0) val += j * o2p * S1[ax][i - 1][j - 1];
            } else {
              val = PB[ax] * S1[ax][i][j - 1];
              if (j > 1) val += (j - 1) * o2p * S1[ax][i][j - 2];
            }
            S1[ax][i][j] = val;
          }
      }
      const double s00 = std::pow(PI / p, 1.5) * eab * cc;
      for (int x = 0; x < na; ++x)
        for (int y = 0; y < nb; ++y) {
          double sx = S1[0][ca[x][0]][cb[y][0]], sy = S1[1][ca[x][1]][cb[y][1]], sz = S1[2][ca[x][2]][cb[y][2]];
          double tk[3];
          for (int ax = 0; ax < 3; ++ax) {
            const int i = ca[x][ax], j = cb[y][ax];
            double tt = b * (2 * j + 1) * S1[ax][i][j] - 2 * b * b * S1[ax][i][j + 2];
            if (j >= 2) tt -= 0.5 * j * (j - 1) * S1[ax][i][j - 2];
            tk[ax] = tt;
          }
          s[x * nb + y] += s00 * sx * sy * sz;
          t[x * nb + y] += s00 * (tk[0] * sy * sz + sx * tk[1] * sz + sx * sy * tk[2]);
        }
      // nuclear attraction: vertical recurrence on the bra index to level Lt
      for (size_t ia = 0; ia < atoms.size(); ++ia) {
        const auto& at = atoms[ia];
        double PC[3];
        for (int k = 0; k < 3; ++k) PC[k] = P[k] - at[1 + k];
        const double Tn = p * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
        boys_eval_rt(Lt, Tn, host_boys().data(), F);
        const double base = -at[0] * 2.0 * PI / p * eab * cc;
        auto I = [&](int m, int x, int y, int z) -> double& { return th[(((size_t)m * D + x) * D + y) * D + z]; };
        for (int m = 0; m <= Lt; ++m) I(m, 0, 0, 0) = base * F[m];
        for (int l = 1; l <= Lt; ++l)
          for (int x = l; x >= 0; --x)
            for (int y = l - x; y >= 0; --y) {
              const int z = l - x - y;
              int e[3] = {x, y, z};
              int i = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
              int e1[3] = {e[0], e[1], e[2]};
              e1[i]--;
              for (int m = 0; m <= Lt - l; ++m) {
                double val = PA[i] * I(m, e1[0], e1[1], e1[2]) - PC[i] * I(m + 1, e1[0], e1[1], e1[2]);
                if (e1[i] > 0) {
                  int e2[3] = {e1[0], e1[1], e1[2]};
                  e2[i]--;
                  val += e1[i] * o2p * (I(m, e2[0], e2[1], e2[2]) - I(m + 1, e2[0], e2[1], e2[2]));
                }
                I(m, x, y, z) = val;
              }
            }
        for (int l = la; l <= Lt; ++l)
          for (int x = l; x >= 0; --x)
            for (int y = l - x; y >= 0; --y) {
              ve[((size_t)x * D + y) * D + (l - x - y)] += I(0, x, y, l - x - y);
            }
      }
    }
  // horizontal recurrence on the contracted V
  const std::vector<double>* src = &ve;
  std::map<std::array<int, 6>, double> memo;
  std::function<double(std::array<int, 3>, std::array<int, 3>)> hrr = [&](std::array<int, 3> ea,
                                                                         std::array<int, 3> eb) -> double {
    if (eb[0] + eb[1] + eb[2] == 0) return (*src)[((size_t)ea[0] * D + ea[1]) * D + ea[2]];
    std::array<int, 6> key{ea[0], ea[1], ea[2], eb[0], eb[1], eb[2]};
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
    int i = eb[0] > 0 ? 0 : (eb[1] > 0 ? 1 : 2);
    auto a1 = ea, b1 = eb;
    a1[i]++;
    b1[i]--;
    double r = hrr(a1, b1) + AB[i] * hrr(ea, b1);
    memo[key] = r;
    return r;
  };
  for (int x = 0; x < na; ++x)
    for (int y = 0; y < nb; ++y) v[x * nb + y] = hrr(ca[x], cb[y]);
}

void one_electron(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms,
                  std::vector<double>& S, std::vector<double>& T, std::vector<double>& V) {
  S.assign((size_t)n * n, 0.0);
  T = S;
  V = S;
  const int ns = (int)sh.size();
  std::vector<std::pair<int, int>> pr;
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b <= a; ++b) pr.push_back({a, b});
#pragma omp parallel for schedule(dynamic)
  for (long q = 0; q < (long)pr.size(); ++q) {
    const Shell& A = sh[pr[q].first];
    const Shell& B = sh[pr[q].second];
    std::vector<double> s, t, v;
    pair_block(A, B, atoms, s, t, v);
    const int na = (int)A.scale.size(), nb = (int)B.scale.size();
    for (int x = 0; x < na; ++x)
      for (int y = 0; y < nb; ++y) {
        const double f = A.scale[x] * B.scale[y];
        const size_t i = A.ao + x, j = B.ao + y;
        S[i * n + j] = S[j * n + i] = f * s[x * nb + y];
        T[i * n + j] = T[j * n + i] = f * t[x * nb + y];
        V[i * n + j] = V[j * n + i] = f * v[x * nb + y];
      }
  }
}

// Real solid harmonics S_lm in Cartesian monomials (Helgaker, Jorgensen, Olsen,
// Molecular Electronic-Structure Theory, eqs. 6.4.47-6.4.50), m = -l..l; each
// spherical function is then normalised with the shell's Cartesian overlap.
// The orthogonaliser depends only on the span and norms, not on the sign
// convention of the harmonics.
static double fct(int k) {
  double r = 1;
  for (int i = 2; i <= k; ++i) r *= i;
  return r;
}
static double binom(int a, int b) { return (b < 0 || b > a) ? 0.0 : fct(a) / (fct(b) * fct(a - b)); }

static std::map<std::array<int, 3>, double> solid_harmonic(int l, int m) {
  std::map<std::array<int, 3>, double> out;
  const int am = std::abs(m);
  const double vm = m < 0 ? 0.5 : 0.0;
  const double N = 1.0 / (std::pow(2.0, am) * fct(l)) * std::sqrt(2.0 * fct(l + am) * fct(l - am) / (m == 0 ? 2.0 : 1.0));
  for (int t = 0; t <= (l - am) / 2; ++t)
    for (int u = 0; u <= t; ++u) {
      // v runs over vm, vm+1, ..., floor(am/2 - vm) + vm
      const int vmax2 = (int)std::floor(am / 2.0 - vm);
      for (int vi = 0; vi <= vmax2; ++vi) {
        const double v = vi + vm;
        const double sign = ((int)std::lround(t + v - vm) % 2) ? -1.0 : 1.0;
        const double c = sign * std::pow(0.25, t) * binom(l, t) * binom(l - t, am + t) * binom(t, u) *
                         binom(am, (int)std::lround(2 * v));
        const int px = 2 * t + am - (int)std::lround(2 * (u + v));
        const int py = (int)std::lround(2 * (u + v));
        const int pz = l - 2 * t - am;
        out[{px, py, pz}] += N * c;
      }
    }
  return out;
}

void spherical_transform(const std::vector<Shell>& sh, int n, std::vector<double>& Tm, int& ncol) {
  ncol = 0;
  for (auto& s : sh) ncol += s.L <= 1 ? (int)s.scale.size() : 2 * s.L + 1;
  Tm.assign((size_t)n * ncol, 0.0);
  int col = 0;
  const std::vector<std::array<double, 4>> none;
  for (auto& s : sh) {
    const int cnt = (int)s.scale.size();
    if (s.L <= 1) {
      for (int r = 0; r < cnt; ++r) Tm[(size_t)(s.ao + r) * ncol + col++] = 1.0;
      continue;
    }
    std::vector<double> so, to, vo;
    pair_block(s, s, none, so, to, vo);
    auto cc = comps(s.L);
    for (int m = -s.L; m <= s.L; ++m) {
      auto poly = solid_harmonic(s.L, m);
      // coefficients on individually normalised components: monomial x^a y^b z^c
      // equals (normalised component) / (norm factor); the overlap below is in
      // unscaled components, so work with unscaled coefficients and rescale.
      std::vector<double> c(cnt, 0.0);
      for (int i = 0; i < cnt; ++i) {
        auto it = poly.find(cc[i]);
        if (it != poly.end()) c[i] = it->second;
      }
      double nrm = 0;
      for (int i = 0; i < cnt; ++i)
        for (int j = 0; j < cnt; ++j) nrm += c[i] * so[i * cnt + j] * c[j];
      if (!(nrm > 0)) throw HfgError{HFG_INTEGRAL, "invalid spherical normalization"};
      // column in the basis of normalised components: unscaled = scale * normalised
      for (int i = 0; i < cnt; ++i) Tm[(size_t)(s.ao + i) * ncol + col] = c[i] / std::sqrt(nrm) / s.scale[i];
      ++col;
    }
  }
}

}  // namespace hfg
