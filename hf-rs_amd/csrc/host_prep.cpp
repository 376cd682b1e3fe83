// host_prep.cpp -- shell grouping, primitive pairs and the Boys table.
#include "host_prep.h"

#include <cmath>
#include <cstring>

#include "eri_device.h"

namespace hfg {

static std::vector<std::array<int, 3>> cart_order(int L) {
  std::vector<std::array<int, 3>> out;
  if (L == 2) return {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
  for (int lx = L; lx >= 0; --lx)
    for (int ly = L - lx; ly >= 0; --ly) out.push_back({lx, ly, L - lx - ly});
  return out;
}

std::vector<Shell> group_shells(const hfg_basis_view& b) {
  std::vector<Shell> out;
  const int n = b.n_functions;
  if (n <= 0) throw HfgError{HFG_INVALID_INPUT, "empty basis"};
  int start = 0;
  while (start < n) {
    const int* l0 = b.lxyz + 3 * start;
    const int L = l0[0] + l0[1] + l0[2];
    const int cnt = (L + 1) * (L + 2) / 2;
    if (L > 7 || start + cnt > n) throw HfgError{HFG_INVALID_INPUT, "incomplete Cartesian shell"};
    const int k0 = b.prim_offset[start], K = b.prim_offset[start + 1] - k0;
    if (K <= 0) throw HfgError{HFG_INVALID_INPUT, "function without primitives"};
    auto order = cart_order(L);
    Shell s;
    s.L = L;
    s.K = K;
    s.ao = start;
    s.atom = b.atom_index[start];
    for (int k = 0; k < 3; ++k) s.c[k] = b.center[3 * start + k];
    for (int c = 0; c < cnt; ++c) {
      const int f = start + c;
      const int* l = b.lxyz + 3 * f;
      if (l[0] != order[c][0] || l[1] != order[c][1] || l[2] != order[c][2])
        throw HfgError{HFG_INVALID_INPUT, "noncontiguous or incomplete Cartesian shell (components out of hf-rs order)"};
      if (b.atom_index[f] != s.atom || std::memcmp(b.center + 3 * f, s.c, sizeof s.c) != 0 ||
          b.prim_offset[f + 1] - b.prim_offset[f] != K)
        throw HfgError{HFG_INVALID_INPUT, "noncontiguous or incomplete Cartesian shell"};
      for (int k = 0; k < K; ++k)
        if (b.exponent[b.prim_offset[f] + k] != b.exponent[k0 + k])
          throw HfgError{HFG_INVALID_INPUT, "noncontiguous or incomplete Cartesian shell"};
    }
    // radial coefficients from component 0; factor of component c from the
    // largest primitive, then a consistency check on every primitive
    int kmax = 0;
    for (int k = 1; k < K; ++k)
      if (std::fabs(b.coefficient[k0 + k]) > std::fabs(b.coefficient[k0 + kmax])) kmax = k;
    for (int k = 0; k < K; ++k) {
      s.e.push_back(b.exponent[k0 + k]);
      s.r.push_back(b.coefficient[k0 + k]);
      if (!(b.exponent[k0 + k] > 0.0) || !std::isfinite(b.coefficient[k0 + k]))
        throw HfgError{HFG_INVALID_INPUT, "invalid primitive"};
    }
    for (int c = 0; c < cnt; ++c) {
      const int fk = b.prim_offset[start + c];
      const double sc = b.coefficient[fk + kmax] / b.coefficient[k0 + kmax];
      for (int k = 0; k < K; ++k) {
        const double v = b.coefficient[fk + k], pred = s.r[k] * sc;
        if (std::fabs(v - pred) > 1e-13 * std::fabs(v) + 1e-300)
          throw HfgError{HFG_INVALID_INPUT, "shell components do not share one radial contraction"};
      }
      s.scale.push_back(sc);
    }
    out.push_back(std::move(s));
    start += cnt;
  }
  return out;
}

void append_pair(const Shell& A, const Shell& B, PrimPairHost& pp) {
  const double ab2 = (A.c[0] - B.c[0]) * (A.c[0] - B.c[0]) + (A.c[1] - B.c[1]) * (A.c[1] - B.c[1]) +
                     (A.c[2] - B.c[2]) * (A.c[2] - B.c[2]);
  for (int ka = 0; ka < A.K; ++ka)
    for (int kb = 0; kb < B.K; ++kb) {
      const double a = A.e[ka], b = B.e[kb], p = a + b;
      double P[3];
      for (int k = 0; k < 3; ++k) P[k] = (a * A.c[k] + b * B.c[k]) / p;
      pp.p.push_back(p);
      pp.Px.push_back(P[0]);
      pp.Py.push_back(P[1]);
      pp.Pz.push_back(P[2]);
      pp.PAx.push_back(P[0] - A.c[0]);
      pp.PAy.push_back(P[1] - A.c[1]);
      pp.PAz.push_back(P[2] - A.c[2]);
      pp.K.push_back(A.r[ka] * B.r[kb] * std::exp(-a * b / p * ab2) / p);
      pp.H.push_back(0.5 / p);
    }
}

void append_aux_pair(const Shell& A, PrimPairHost& pp) {
  for (int k = 0; k < A.K; ++k) {
    const double p = A.e[k];
    pp.p.push_back(p);
    pp.Px.push_back(A.c[0]);
    pp.Py.push_back(A.c[1]);
    pp.Pz.push_back(A.c[2]);
    pp.PAx.push_back(0.0);
    pp.PAy.push_back(0.0);
    pp.PAz.push_back(0.0);
    pp.K.push_back(A.r[k] / p);
    pp.H.push_back(0.5 / p);
  }
}

void build_pairs(const std::vector<Shell>& sh, std::vector<PairInfo>& pairs, PrimPairHost& pp) {
  pairs.clear();
  const int ns = (int)sh.size();
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b <= a; ++b) {
      const bool swap = sh[b].L > sh[a].L;
      const Shell& A = swap ? sh[b] : sh[a];
      const Shell& B = swap ? sh[a] : sh[b];
      PairInfo pi;
      pi.p0 = (int)pp.size();
      append_pair(A, B, pp);
      pi.p1 = (int)pp.size();
      pi.ao_a = A.ao;
      pi.ao_b = B.ao;
      pi.sh_a = swap ? b : a;
      pi.sh_b = swap ? a : b;
      for (int k = 0; k < 3; ++k) pi.AB[k] = A.c[k] - B.c[k];
      pi.la = A.L;
      pi.lb = B.L;
      pi.hf_index = a * (a + 1) / 2 + b;
      pi.swapped = swap;
      pairs.push_back(pi);
    }
}

// F_m(T) for m = 0..MT by the convergent series at MT and downward recursion,
// in long double; F_MT = e^{-T} sum_i (2T)^i / ((2MT+1)(2MT+3)...(2MT+2i+1)).
std::vector<double> boys_table() {
  std::vector<double> tab((size_t)BOYS_ROWS * BOYS_ROW);
  const int MT = BOYS_ORDERS - 1;
  for (int r = 0; r < BOYS_ROWS; ++r) {
    const long double T = (long double)r / BOYS_INV_STEP;
    const long double et = expl(-T);
    long double term = 1.0L / (2 * MT + 1), sum = term;
    for (int i = 1; i < 100000; ++i) {
      term *= 2.0L * T / (2 * MT + 2 * i + 1);
      sum += term;
      if (term < sum * 1e-22L) break;
    }
    std::vector<long double> f(MT + 1);
    f[MT] = et * sum;
    for (int m = MT - 1; m >= 0; --m) f[m] = (2.0L * T * f[m + 1] + et) / (2 * m + 1);
    double* row = &tab[(size_t)r * BOYS_ROW];
    row[0] = (double)et;
    for (int m = 0; m <= MT; ++m) row[1 + m] = (double)f[m];
  }
  return tab;
}

int pair_class(int la, int lb) {
  if (lb > la) return -1;
  if (la <= 2) return la * (la + 1) / 2 + lb;  // ss0 ps1 pp2 ds3 dp4 dd5
  if (lb == 0 && la <= 4) return 3 + la;     // fs6 gs7 (auxiliary shell with the unit function)
  if (la == 3) return 7 + lb;                // fp8 fd9 ff10 (orbital pairs of the 3-index integrals)
  return -1;
}
int quartet_class(int pb, int pk) { return pb * (pb + 1) / 2 + pk; }

}  // namespace hfg
