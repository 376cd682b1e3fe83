// hfo.cpp -- CPU oracle for the hf-rs Fock-build path.  TEST INFRASTRUCTURE ONLY.
//
// An FP64 restatement of the reference algorithms, used as the checker for the
// HIP engine and as the CPU baseline that bench.py times.  The product library
// never links, loads or calls this code.  Each routine names the reference
// file:line it restates.  The arithmetic follows the reference's McMurchie-
// Davidson formulation (Hermite E coefficients, Hermite Coulomb R recursion,
// per-component 1-D convolutions), so that a mistake in the HIP engine's
// different (Obara-Saika / Head-Gordon-Pople) recursion cannot cancel here.
#include "hfo.h"

#include <omp.h>

#include <strings.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;
int fail(const std::string& m, int code = 1) {
  g_err = m;
  return code;
}

const double PI = 3.14159265358979323846264338327950288;
const double TWO_PI_POW_2_5 = 34.986836655249726164477;  // 2 pi^(5/2)
const double PI_POW_1_5 = 5.56832799683170784528;
const double HALF_SQRT_PI = 0.88622692545275801365;

// ---------------------------------------------------------------------------
// Boys function: src/integrals.rs:20-101.  Table of F_0..F_47 on T = 0..50.05
// (step 0.05, 1002 rows) from the convergent series at m = 47 followed by
// downward recursion; evaluation by a 7-term Taylor expansion around the
// nearest grid point; T > 50 from the erf closed form and upward recursion.
// ---------------------------------------------------------------------------
const int BOYS_MAX_M = 40;
const int BOYS_TAYLOR = 7;
const int BOYS_TABLE_M = BOYS_MAX_M + BOYS_TAYLOR;
const double BOYS_STEP = 0.05;
const double BOYS_SWITCH = 50.0;
const int BOYS_GRID = 1002;

void boys_downward(int max_n, double t, double* out) {  // src/integrals.rs:39-57
  double exp_t = std::exp(-t);
  double term = 1.0 / (2 * max_n + 1);
  double sum = term;
  int i = 1;
  for (;;) {
    term *= 2.0 * t / (2 * max_n + 2 * i + 1);
    sum += term;
    if (term < sum * 1e-17 || i > 10000) break;
    ++i;
  }
  out[max_n] = exp_t * sum;
  for (int m = max_n - 1; m >= 0; --m) out[m] = (2.0 * t * out[m + 1] + exp_t) / (2 * m + 1);
}

const std::vector<double>& boys_table() {
  static std::vector<double> tab = [] {
    std::vector<double> v((size_t)BOYS_GRID * (BOYS_TABLE_M + 1));
    for (int i = 0; i < BOYS_GRID; ++i) boys_downward(BOYS_TABLE_M, i * BOYS_STEP, &v[(size_t)i * (BOYS_TABLE_M + 1)]);
    return v;
  }();
  return tab;
}

void boys_array(int max_n, double t, double* out) {  // src/integrals.rs:70-101
  if (max_n > BOYS_MAX_M) throw std::runtime_error("Boys order exceeds supported maximum");
  if (t > BOYS_SWITCH) {
    double exp_t = std::exp(-t), inv_2t = 0.5 / t, st = std::sqrt(t);
    out[0] = HALF_SQRT_PI * std::erf(st) / st;
    for (int n = 1; n <= max_n; ++n) out[n] = ((2 * n - 1) * out[n - 1] - exp_t) * inv_2t;
    return;
  }
  int idx = (int)std::round(t * (1.0 / BOYS_STEP));
  double dt = t - idx * BOYS_STEP;
  const double* row = &boys_table()[(size_t)idx * (BOYS_TABLE_M + 1)];
  double coeff[BOYS_TAYLOR + 1];
  coeff[0] = 1.0;
  for (int k = 1; k <= BOYS_TAYLOR; ++k) coeff[k] = coeff[k - 1] * (-dt) / k;
  for (int n = 0; n <= max_n; ++n) {
    double acc = 0.0;
    for (int k = BOYS_TAYLOR; k >= 0; --k) acc += coeff[k] * row[n + k];
    out[n] = acc;
  }
}

// ---------------------------------------------------------------------------
// Hermite expansion coefficients E^{ij}_t: src/integrals.rs:135-250.
// Layout [i][j][t] with i<=la, j<=lb, t<=la+lb.
// ---------------------------------------------------------------------------
struct Herm {
  int la, lb, nt;
  std::vector<double> d;
  double get(int i, int j, int t) const { return d[((size_t)i * (lb + 1) + j) * nt + t]; }
  double& at(int i, int j, int t) { return d[((size_t)i * (lb + 1) + j) * nt + t]; }
};

void hermite(int la, int lb, double alpha, double beta, double a, double b, Herm& e) {
  e.la = la;
  e.lb = lb;
  e.nt = la + lb + 1;
  e.d.assign((size_t)(la + 1) * (lb + 1) * e.nt, 0.0);
  double p = alpha + beta, inv_p = 1.0 / p;
  double pc = (alpha * a + beta * b) * inv_p;
  double xpa = pc - a, xpb = pc - b, mu = alpha * beta * inv_p;
  int max_t = la + lb;
  double inv_2p = 0.5 * inv_p;
  double diff = a - b;
  e.at(0, 0, 0) = std::exp(-mu * diff * diff);
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j)
      for (int t = 0; t <= i + j; ++t) {
        if (i == 0 && j == 0) continue;
        if (i > 0) {
          double t1 = t > 0 ? e.get(i - 1, j, t - 1) * inv_2p : 0.0;
          double t2 = xpa * e.get(i - 1, j, t);
          double t3 = t < max_t ? (t + 1.0) * e.get(i - 1, j, t + 1) : 0.0;
          e.at(i, j, t) = t1 + t2 + t3;
        } else {
          double t1 = t > 0 ? e.get(i, j - 1, t - 1) * inv_2p : 0.0;
          double t2 = xpb * e.get(i, j - 1, t);
          double t3 = t < max_t ? (t + 1.0) * e.get(i, j - 1, t + 1) : 0.0;
          e.at(i, j, t) = t1 + t2 + t3;
        }
      }
}

// Hermite Coulomb integrals R^n_{tuv}: src/integrals.rs:259-346 (flat [t][u][v][n]).
void coulomb_r(int max_t, int max_u, int max_v, int max_n, double p, const double pmc[3], double* r) {
  double x = pmc[0], y = pmc[1], z = pmc[2];
  double r2 = x * x + y * y + z * z;
  size_t sv = max_n + 1, su = (max_v + 1) * sv, st = (max_u + 1) * su;
  double f[BOYS_MAX_M + 1];
  boys_array(max_n, p * r2, f);
  double neg_2p = -2.0 * p, pw = 1.0;
  for (int n = 0; n <= max_n; ++n) {
    r[n] = pw * f[n];
    pw *= neg_2p;
  }
  for (int t = 0; t < std::min(max_t, max_n); ++t) {
    size_t tc = t * st, tp = t > 0 ? (t - 1) * st : 0, tn = (t + 1) * st;
    int lim = max_n - t - 1;
    for (int n = 0; n <= lim; ++n)
      r[tn + n] = (t > 0 ? t * r[tp + n + 1] : 0.0) + x * r[tc + n + 1];
  }
  for (int t = 0; t <= std::min(max_t, max_n); ++t) {
    size_t to = t * st;
    for (int u = 0; u < std::min(max_u, max_n - t); ++u) {
      size_t uc = to + u * su, up = u > 0 ? to + (u - 1) * su : 0, un = to + (u + 1) * su;
      int lim = max_n - t - u - 1;
      for (int n = 0; n <= lim; ++n)
        r[un + n] = (u > 0 ? u * r[up + n + 1] : 0.0) + y * r[uc + n + 1];
    }
  }
  for (int t = 0; t <= std::min(max_t, max_n); ++t) {
    size_t to = t * st;
    for (int u = 0; u <= std::min(max_u, max_n - t); ++u) {
      size_t tuo = to + u * su;
      for (int v = 0; v < std::min(max_v, max_n - t - u); ++v) {
        size_t vc = tuo + v * sv, vp = v > 0 ? tuo + (v - 1) * sv : 0, vn = tuo + (v + 1) * sv;
        int lim = max_n - t - u - v - 1;
        for (int n = 0; n <= lim; ++n)
          r[vn + n] = (v > 0 ? v * r[vp + n + 1] : 0.0) + z * r[vc + n + 1];
      }
    }
  }
}

// Per-component contraction: src/integrals.rs:905-971.
double eri_contraction(int tab, int uab, int vab, int tcd, int ucd, int vcd, size_t st, size_t su,
                       size_t sv, const double* exab, const double* eyab, const double* ezab,
                       const double* excd, const double* eycd, const double* ezcd, const double* r,
                       double prefactor) {
  double cx[BOYS_MAX_M + 1] = {0}, cy[BOYS_MAX_M + 1] = {0}, cz[BOYS_MAX_M + 1] = {0};
  for (int t = 0; t <= tab; ++t) {
    double et = exab[t];
    if (et == 0.0) continue;
    for (int tp = 0; tp <= tcd; ++tp) cx[t + tp] += et * ((tp & 1) ? -excd[tp] : excd[tp]);
  }
  for (int u = 0; u <= uab; ++u) {
    double eu = eyab[u];
    if (eu == 0.0) continue;
    for (int up = 0; up <= ucd; ++up) cy[u + up] += eu * ((up & 1) ? -eycd[up] : eycd[up]);
  }
  for (int v = 0; v <= vab; ++v) {
    double ev = ezab[v];
    if (ev == 0.0) continue;
    for (int vp = 0; vp <= vcd; ++vp) cz[v + vp] += ev * ((vp & 1) ? -ezcd[vp] : ezcd[vp]);
  }
  double sum = 0.0;
  for (int t = 0; t <= tab + tcd; ++t) {
    double cxt = cx[t];
    if (cxt == 0.0) continue;
    for (int u = 0; u <= uab + ucd; ++u) {
      double cyu = cy[u];
      if (cyu == 0.0) continue;
      double cc = cxt * cyu;
      size_t ri = t * st + u * su;
      for (int v = 0; v <= vab + vcd; ++v) {
        double czv = cz[v];
        if (czv != 0.0) sum += cc * czv * r[ri];
        ri += sv;
      }
    }
  }
  return prefactor * sum;
}

// ---------------------------------------------------------------------------
// Basis: src/basis.rs:14-47 (atoms), 89-132 (types, normalization),
// 170-188 (Cartesian order), 296-357 (G94), 402-496 (build + renormalize).
// ---------------------------------------------------------------------------
const char* PERIODIC[] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F", "Ne", "Na", "Mg", "Al",
                          "Si", "P", "S", "Cl", "Ar"};
int symbol_z(const std::string& s) {
  for (int i = 0; i < 18; ++i) {
    std::string u = PERIODIC[i];
    if (u.size() == s.size() && strcasecmp(u.c_str(), s.c_str()) == 0) return i + 1;
  }
  return 0;
}

struct Prim {
  double e, c;
};
struct CG {
  int atom;
  double center[3];
  int l[3];
  std::vector<Prim> prims;
  int L() const { return l[0] + l[1] + l[2]; }
};
struct ShellDef {
  int L;  // -1 for SP
  std::vector<double> exps, coefs, pcoefs;
};

double dfact(long n) {
  if (n <= 0) return 1.0;
  double r = 1.0;
  for (long k = n; k > 0; k -= 2) r *= (double)k;
  return r;
}
double primitive_norm(double a, const int l[3]) {  // src/basis.rs:123-132
  int lt = l[0] + l[1] + l[2];
  double pre = std::pow(2.0 * a / PI, 0.75);
  double num = std::pow(4.0 * a, lt);
  double den = dfact(2 * l[0] - 1) * dfact(2 * l[1] - 1) * dfact(2 * l[2] - 1);
  return pre * std::sqrt(num / den);
}

std::vector<std::array<int, 3>> cart_components(int L) {
  std::vector<std::array<int, 3>> out;
  if (L == 2) {  // src/basis.rs:430-436
    out = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    return out;
  }
  for (int lx = L; lx >= 0; --lx)  // src/basis.rs:179-188
    for (int ly = L - lx; ly >= 0; --ly) out.push_back({lx, ly, L - lx - ly});
  return out;
}

CG make_cg(int atom, const double c[3], const int l[3], const std::vector<double>& exps,
           const std::vector<double>& coefs) {  // src/basis.rs:455-496
  CG g;
  g.atom = atom;
  for (int k = 0; k < 3; ++k) {
    g.center[k] = c[k];
    g.l[k] = l[k];
  }
  for (size_t i = 0; i < exps.size(); ++i) g.prims.push_back({exps[i], coefs[i] * primitive_norm(exps[i], l)});
  double so = 0.0;
  for (auto& p1 : g.prims)
    for (auto& p2 : g.prims) {
      double p = p1.e + p2.e;
      double s = std::pow(PI / p, 1.5) * (dfact(2 * l[0] - 1) / std::pow(2.0 * p, l[0])) *
                 (dfact(2 * l[1] - 1) / std::pow(2.0 * p, l[1])) *
                 (dfact(2 * l[2] - 1) / std::pow(2.0 * p, l[2]));
      so += p1.c * p2.c * s;
    }
  double nf = 1.0 / std::sqrt(so);
  for (auto& p : g.prims) p.c *= nf;
  return g;
}

double parse_num(std::string s) {
  for (auto& ch : s)
    if (ch == 'D' || ch == 'd') ch = 'E';
  size_t pos = 0;
  double v = std::stod(s, &pos);
  if (pos != s.size()) throw std::runtime_error("bad number " + s);
  return v;
}

int shell_letter_l(const std::string& t) {
  static const std::string L = "SPDFGHIK";
  if (t.size() == 1) {
    auto p = L.find((char)toupper(t[0]));
    if (p != std::string::npos) return (int)p;
  }
  return -2;
}

std::map<std::string, std::vector<ShellDef>> parse_g94(const std::string& text) {
  std::vector<std::string> lines;
  std::istringstream is(text);
  std::string raw;
  while (std::getline(is, raw)) {
    auto cut = raw.find_first_of("!#");
    if (cut != std::string::npos) raw = raw.substr(0, cut);
    size_t a = raw.find_first_not_of(" \t\r"), b = raw.find_last_not_of(" \t\r");
    if (a == std::string::npos) continue;
    lines.push_back(raw.substr(a, b - a + 1));
  }
  std::map<std::string, std::vector<ShellDef>> map;
  std::string elem;
  size_t i = 0;
  while (i < lines.size()) {
    if (lines[i] == "****") {
      elem.clear();
      ++i;
      continue;
    }
    std::istringstream ls(lines[i]);
    std::vector<std::string> tok;
    std::string t;
    while (ls >> t) tok.push_back(t);
    if (elem.empty()) {
      if (tok.size() != 2 || tok[1] != "0" || !symbol_z(tok[0])) throw std::runtime_error("bad G94 element header: " + lines[i]);
      elem = tok[0];
      for (auto& ch : elem) ch = (char)toupper(ch);
      map[elem];
      ++i;
      continue;
    }
    std::string st = tok[0];
    for (auto& ch : st) ch = (char)toupper(ch);
    if (tok.size() != 3) throw std::runtime_error("bad G94 shell header: " + lines[i]);
    int L = st == "SP" ? -1 : shell_letter_l(st);
    if (L == -2) throw std::runtime_error("bad shell type " + st);
    int n = std::stoi(tok[1]);
    double scale = parse_num(tok[2]);
    std::vector<double> exps;
    std::vector<std::vector<double>> cols;
    for (int r = 0; r < n; ++r) {
      ++i;
      if (i >= lines.size()) throw std::runtime_error("truncated G94 shell");
      std::istringstream ps(lines[i]);
      std::vector<double> vals;
      while (ps >> t) vals.push_back(parse_num(t));
      if (vals.size() < 2) throw std::runtime_error("bad primitive line");
      if (r == 0) cols.assign(vals.size() - 1, {});
      exps.push_back(vals[0] * scale * scale);
      for (size_t c = 1; c < vals.size(); ++c) cols[c - 1].push_back(vals[c]);
    }
    auto& shells = map[elem];
    if (L == -1) {
      for (size_t c = 0; c + 1 < cols.size(); c += 2) shells.push_back({-1, exps, cols[c], cols[c + 1]});
    } else {
      for (auto& c : cols) shells.push_back({L, exps, c, {}});
    }
    ++i;
  }
  return map;
}

}  // namespace

struct hfo_basis {
  std::vector<CG> f;
  std::vector<int> shell_start;  // complete shells, src/ao.rs:133-164
};

namespace {

int compute_shells(hfo_basis& b) {  // src/ao.rs:133-164
  b.shell_start.clear();
  size_t start = 0;
  while (start < b.f.size()) {
    const CG& a = b.f[start];
    int l = a.L();
    size_t cnt = (size_t)(l + 1) * (l + 2) / 2;
    if (l > 7 || start + cnt > b.f.size()) return fail("incomplete Cartesian shell");
    std::vector<std::array<int, 3>> seen;
    for (size_t k = start; k < start + cnt; ++k) {
      const CG& g = b.f[k];
      std::array<int, 3> key{g.l[0], g.l[1], g.l[2]};
      bool dup = std::find(seen.begin(), seen.end(), key) != seen.end();
      bool bad = g.atom != a.atom || g.center[0] != a.center[0] || g.center[1] != a.center[1] ||
                 g.center[2] != a.center[2] || g.L() != l || g.prims.size() != a.prims.size() || dup;
      for (size_t p = 0; !bad && p < g.prims.size(); ++p) bad = g.prims[p].e != a.prims[p].e;
      if (bad) return fail("noncontiguous or incomplete Cartesian shell");
      seen.push_back(key);
    }
    b.shell_start.push_back((int)start);
    start += cnt;
  }
  b.shell_start.push_back((int)b.f.size());
  return 0;
}

// ---------------------------------------------------------------------------
// Shell pairs and quartets: src/shell_integrals.rs:12-135.
// ---------------------------------------------------------------------------
struct PairPrim {
  double p, P[3];
  std::vector<double> herm[3];  // [i*(lb+1)*(la+lb+1) + j*(la+lb+1) + t]
  std::vector<double> coef;     // per component pair ia*nb+ib
};
struct ShellPair {
  int sa, sb, fa, fb, na, nb, la, lb;  // shell ids, first function index, counts, L
  std::vector<PairPrim> prims;
};

ShellPair make_pair(const hfo_basis& B, int sa, int sb) {
  ShellPair sp;
  sp.sa = sa;
  sp.sb = sb;
  sp.fa = B.shell_start[sa];
  sp.fb = B.shell_start[sb];
  sp.na = B.shell_start[sa + 1] - sp.fa;
  sp.nb = B.shell_start[sb + 1] - sp.fb;
  const CG& a = B.f[sp.fa];
  const CG& b = B.f[sp.fb];
  sp.la = a.L();
  sp.lb = b.L();
  int stride = (sp.lb + 1) * (sp.la + sp.lb + 1);
  for (size_t ia = 0; ia < a.prims.size(); ++ia)
    for (size_t ib = 0; ib < b.prims.size(); ++ib) {
      PairPrim pp;
      double al = a.prims[ia].e, be = b.prims[ib].e;
      pp.p = al + be;
      for (int k = 0; k < 3; ++k) pp.P[k] = (al * a.center[k] + be * b.center[k]) / pp.p;
      for (int ax = 0; ax < 3; ++ax) {
        Herm e;
        hermite(sp.la, sp.lb, al, be, a.center[ax], b.center[ax], e);
        pp.herm[ax].assign((size_t)(sp.la + 1) * stride, 0.0);
        for (int i = 0; i <= sp.la; ++i)
          for (int j = 0; j <= sp.lb; ++j)
            for (int t = 0; t <= i + j; ++t) pp.herm[ax][i * stride + j * (sp.la + sp.lb + 1) + t] = e.get(i, j, t);
      }
      for (int x = 0; x < sp.na; ++x)
        for (int y = 0; y < sp.nb; ++y)
          pp.coef.push_back(B.f[sp.fa + x].prims[ia].c * B.f[sp.fb + y].prims[ib].c);
      sp.prims.push_back(std::move(pp));
    }
  return sp;
}

const double* expansion(const hfo_basis& B, const ShellPair& sp, const PairPrim& p, int axis, int i, int j) {
  int stride = sp.la + sp.lb + 1;
  int off = (B.f[sp.fa + i].l[axis] * (sp.lb + 1) + B.f[sp.fb + j].l[axis]) * stride;
  return &p.herm[axis][off];
}

void quartet(const hfo_basis& B, const ShellPair& ab, const ShellPair& cd, bool diag, std::vector<double>& res) {
  int nab = ab.na * ab.nb, ncd = cd.na * cd.nb;
  res.assign(diag ? nab : (size_t)nab * ncd, 0.0);
  int order = ab.la + ab.lb + cd.la + cd.lb;
  size_t sv = order + 1, su = sv * sv, st = su * sv;
  std::vector<double> r(st * sv);
  for (const auto& p : ab.prims)
    for (const auto& q : cd.prims) {
      double rho = p.p * q.p / (p.p + q.p);
      double delta[3] = {p.P[0] - q.P[0], p.P[1] - q.P[1], p.P[2] - q.P[2]};
      coulomb_r(order, order, order, order, rho, delta, r.data());
      double pref = TWO_PI_POW_2_5 / (p.p * q.p * std::sqrt(p.p + q.p));
      for (int ij = 0; ij < nab; ++ij) {
        int i = ij / ab.nb, j = ij % ab.nb;
        const double* eab[3];
        int lab[3];
        for (int k = 0; k < 3; ++k) {
          eab[k] = expansion(B, ab, p, k, i, j);
          lab[k] = B.f[ab.fa + i].l[k] + B.f[ab.fb + j].l[k];
        }
        int s0 = diag ? ij : 0, s1 = diag ? ij + 1 : ncd;
        for (int kl = s0; kl < s1; ++kl) {
          int k = kl / cd.nb, l = kl % cd.nb;
          const double* ecd[3];
          int lcd[3];
          for (int ax = 0; ax < 3; ++ax) {
            ecd[ax] = expansion(B, cd, q, ax, k, l);
            lcd[ax] = B.f[cd.fa + k].l[ax] + B.f[cd.fb + l].l[ax];
          }
          double v = eri_contraction(lab[0], lab[1], lab[2], lcd[0], lcd[1], lcd[2], st, su, sv, eab[0], eab[1],
                                     eab[2], ecd[0], ecd[1], ecd[2], r.data(), pref);
          size_t idx = diag ? ij : (size_t)ij * ncd + kl;
          res[idx] += p.coef[ij] * q.coef[kl] * v;
        }
      }
    }
}

struct PairSet {
  std::vector<std::pair<int, int>> idx;
  std::vector<ShellPair> pairs;
  std::vector<double> bound;
};

void build_pairs(const hfo_basis& B, PairSet& ps) {  // src/shell_integrals.rs:158-173
  int ns = (int)B.shell_start.size() - 1;
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b <= a; ++b) ps.idx.push_back({a, b});
  ps.pairs.resize(ps.idx.size());
  ps.bound.resize(ps.idx.size());
#pragma omp parallel for schedule(dynamic)
  for (long i = 0; i < (long)ps.idx.size(); ++i) {
    ps.pairs[i] = make_pair(B, ps.idx[i].first, ps.idx[i].second);
    std::vector<double> d;
    quartet(B, ps.pairs[i], ps.pairs[i], true, d);
    double m = 0.0;
    for (double v : d) m = std::max(m, v);
    ps.bound[i] = std::sqrt(m);
  }
}

inline size_t compound(size_t i, size_t j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }
inline size_t eri_index(size_t i, size_t j, size_t k, size_t l) { return compound(compound(i, j), compound(k, l)); }

// The per-value digestion of src/scf.rs:345-357: all distinct permutations of
// the unique quartet (mu nu|la si) add D_tot[c,d] v to J[a,b] and D[b,d] v to K[a,c].
inline void digest(double* J, double* K, const double* P, int n, int mu, int nu, int la, int si, bool same_pair,
                   double v) {
  for (int pf = 0; pf < (same_pair ? 1 : 2); ++pf) {
    int aa = pf == 0 ? mu : la, bb = pf == 0 ? nu : si, cc = pf == 0 ? la : mu, dd = pf == 0 ? si : nu;
    for (int abf = 0; abf < (aa == bb ? 1 : 2); ++abf) {
      int a = abf == 0 ? aa : bb, b = abf == 0 ? bb : aa;
      for (int cdf = 0; cdf < (cc == dd ? 1 : 2); ++cdf) {
        int c = cdf == 0 ? cc : dd, d = cdf == 0 ? dd : cc;
        J[(size_t)a * n + b] += P[(size_t)c * n + d] * v;
        K[(size_t)a * n + c] += P[(size_t)b * n + d] * v;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Pair engine with primitive screening (src/integrals.rs:819-902, 1032-1072),
// used by contracted_eri and hence by the DF metric and 3-index integrals.
// ---------------------------------------------------------------------------
struct PPData {
  double p, P[3], coeff;
  int tm, um, vm;
  std::vector<double> ex, ey, ez;
};
std::vector<PPData> basis_pair_data(const CG& a, const CG& b) {
  std::vector<PPData> out;
  double dx = a.center[0] - b.center[0], dy = a.center[1] - b.center[1], dz = a.center[2] - b.center[2];
  double d2 = dx * dx + dy * dy + dz * dz;
  for (auto& pa : a.prims)
    for (auto& pb : b.prims) {
      double coeff = pa.c * pb.c;
      if (std::fabs(coeff) < 1e-15) continue;
      double p = pa.e + pb.e, inv_p = 1.0 / p, mu = pa.e * pb.e * inv_p;
      if (std::fabs(coeff * std::exp(-mu * d2)) < 1e-13) continue;
      PPData d;
      d.p = p;
      d.coeff = coeff;
      for (int k = 0; k < 3; ++k) d.P[k] = (pa.e * a.center[k] + pb.e * b.center[k]) * inv_p;
      Herm e[3];
      for (int ax = 0; ax < 3; ++ax) hermite(a.l[ax], b.l[ax], pa.e, pb.e, a.center[ax], b.center[ax], e[ax]);
      d.tm = a.l[0] + b.l[0];
      d.um = a.l[1] + b.l[1];
      d.vm = a.l[2] + b.l[2];
      for (int t = 0; t <= d.tm; ++t) d.ex.push_back(e[0].get(a.l[0], b.l[0], t));
      for (int t = 0; t <= d.um; ++t) d.ey.push_back(e[1].get(a.l[1], b.l[1], t));
      for (int t = 0; t <= d.vm; ++t) d.ez.push_back(e[2].get(a.l[2], b.l[2], t));
      out.push_back(std::move(d));
    }
  return out;
}

double prim_eri_pairs(const PPData& ab, const PPData& cd, std::vector<double>& r) {  // src/integrals.rs:973-1030
  double p = ab.p, q = cd.p, pq = p + q, ap = p * q / pq;
  double pmq[3] = {ab.P[0] - cd.P[0], ab.P[1] - cd.P[1], ab.P[2] - cd.P[2]};
  int mt = ab.tm + cd.tm, mu = ab.um + cd.um, mv = ab.vm + cd.vm, mn = mt + mu + mv;
  double r2 = pmq[0] * pmq[0] + pmq[1] * pmq[1] + pmq[2] * pmq[2];
  double pref = TWO_PI_POW_2_5 / (p * q * std::sqrt(pq));
  if (mn == 0) {
    double f0;
    boys_array(0, ap * r2, &f0);
    return pref * (ab.ex[0] * ab.ey[0] * ab.ez[0]) * (cd.ex[0] * cd.ey[0] * cd.ez[0]) * f0;
  }
  size_t sv = mn + 1, su = (mv + 1) * sv, st = (mu + 1) * su;
  r.resize((mt + 1) * st);
  coulomb_r(mt, mu, mv, mn, ap, pmq, r.data());
  return eri_contraction(ab.tm, ab.um, ab.vm, cd.tm, cd.um, cd.vm, st, su, sv, ab.ex.data(), ab.ey.data(),
                         ab.ez.data(), cd.ex.data(), cd.ey.data(), cd.ez.data(), r.data(), pref);
}

double pair_eri(const std::vector<PPData>& ab, const std::vector<PPData>& cd, std::vector<double>& r) {
  double v = 0.0;  // src/integrals.rs:1032-1059
  for (auto& x : ab)
    for (auto& y : cd) {
      double tc = x.coeff * y.coeff;
      if (std::fabs(tc) < 1e-13) continue;
      v += tc * prim_eri_pairs(x, y, r);
    }
  return v;
}

CG constant_fn(const double c[3]) {  // src/df.rs:30-40
  CG g;
  g.atom = 0;
  for (int k = 0; k < 3; ++k) {
    g.center[k] = c[k];
    g.l[k] = 0;
  }
  g.prims.push_back({0.0, 1.0});
  return g;
}

// ---------------------------------------------------------------------------
// Dense linear algebra (src/matrix.rs).
// ---------------------------------------------------------------------------
typedef std::vector<double> Mat;
Mat mm(const Mat& a, const Mat& b, int m, int k, int n) {  // a m x k, b k x n
  Mat c((size_t)m * n, 0.0);
#pragma omp parallel for if ((long)m * k * n > 100000)
  for (int i = 0; i < m; ++i)
    for (int p = 0; p < k; ++p) {
      double aip = a[(size_t)i * k + p];
      if (aip == 0.0) continue;
      const double* br = &b[(size_t)p * n];
      double* cr = &c[(size_t)i * n];
      for (int j = 0; j < n; ++j) cr[j] += aip * br[j];
    }
  return c;
}
Mat tr(const Mat& a, int m, int n) {
  Mat t((size_t)m * n);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) t[(size_t)j * m + i] = a[(size_t)i * n + j];
  return t;
}

void sort_sign(int n, std::vector<double>& w, Mat& v) {  // src/matrix.rs:150-190 sort + phase
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return w[a] < w[b]; });
  std::vector<double> w2(n);
  Mat v2((size_t)n * n);
  for (int c = 0; c < n; ++c) {
    int o = ord[c];
    w2[c] = w[o];
    double mx = 0.0, sg = 1.0;
    for (int r = 0; r < n; ++r) {
      double x = v[(size_t)r * n + o];
      if (std::fabs(x) > mx) {
        mx = std::fabs(x);
        sg = x >= 0.0 ? 1.0 : -1.0;
      }
    }
    for (int r = 0; r < n; ++r) v2[(size_t)r * n + c] = sg * v[(size_t)r * n + o];
  }
  w = w2;
  v = v2;
}

void eigh_jacobi(int n, const Mat& in, std::vector<double>& w, Mat& v) {  // src/matrix.rs:192-290
  Mat a = in;
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  double scale = 1.0;
  for (double x : a) scale = std::max(scale, std::fabs(x));
  double tol = 1e-15 * scale;
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += a[(size_t)i * n + j] * a[(size_t)i * n + j];
    if (off < tol * tol) break;
    double thr = sweep < 4 ? 0.2 * std::sqrt(off) / n : 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = a[(size_t)p * n + q];
        if (std::fabs(apq) <= thr) continue;
        double app = a[(size_t)p * n + p], aqq = a[(size_t)q * n + q];
        double theta = 0.5 * (aqq - app) / apq;
        double t = theta >= 0.0 ? 1.0 / (theta + std::sqrt(theta * theta + 1.0))
                                : -1.0 / (-theta + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
        for (int i = 0; i < n; ++i)
          if (i != p && i != q) {
            double aip = a[(size_t)i * n + p], aiq = a[(size_t)i * n + q];
            a[(size_t)i * n + p] = c * aip - s * aiq;
            a[(size_t)p * n + i] = a[(size_t)i * n + p];
            a[(size_t)i * n + q] = s * aip + c * aiq;
            a[(size_t)q * n + i] = a[(size_t)i * n + q];
          }
        a[(size_t)p * n + p] = c * c * app - 2.0 * s * c * apq + s * s * aqq;
        a[(size_t)q * n + q] = s * s * app + 2.0 * s * c * apq + c * c * aqq;
        a[(size_t)p * n + q] = 0.0;
        a[(size_t)q * n + p] = 0.0;
        for (int i = 0; i < n; ++i) {
          double vip = v[(size_t)i * n + p], viq = v[(size_t)i * n + q];
          v[(size_t)i * n + p] = c * vip - s * viq;
          v[(size_t)i * n + q] = s * vip + c * viq;
        }
      }
  }
  w.resize(n);
  for (int i = 0; i < n; ++i) w[i] = a[(size_t)i * n + i];
  sort_sign(n, w, v);
}

// Householder tridiagonalisation + implicit QL with shifts (the EISPACK
// tred2/tql2 method, Wilkinson & Reinsch, Handbook for Automatic Computation
// II, 1971) -- stands in for faer's self-adjoint solver (src/matrix.rs:148-170);
// eigenvalues agree to rounding, vectors after the sort/phase convention.
void eigh_ql(int n, const Mat& in, std::vector<double>& w, Mat& vout) {
  std::vector<std::vector<double>> V(n, std::vector<double>(n));
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) V[i][j] = 0.5 * (in[(size_t)i * n + j] + in[(size_t)j * n + i]);
  std::vector<double> d(n), e(n);
  for (int j = 0; j < n; ++j) d[j] = V[n - 1][j];
  for (int i = n - 1; i > 0; --i) {
    double scale = 0.0, h = 0.0;
    for (int k = 0; k < i; ++k) scale += std::fabs(d[k]);
    if (scale == 0.0) {
      e[i] = d[i - 1];
      for (int j = 0; j < i; ++j) {
        d[j] = V[i - 1][j];
        V[i][j] = 0.0;
        V[j][i] = 0.0;
      }
    } else {
      for (int k = 0; k < i; ++k) {
        d[k] /= scale;
        h += d[k] * d[k];
      }
      double f = d[i - 1], g = std::sqrt(h);
      if (f > 0) g = -g;
      e[i] = scale * g;
      h -= f * g;
      d[i - 1] = f - g;
      for (int j = 0; j < i; ++j) e[j] = 0.0;
      for (int j = 0; j < i; ++j) {
        f = d[j];
        V[j][i] = f;
        g = e[j] + V[j][j] * f;
        for (int k = j + 1; k <= i - 1; ++k) {
          g += V[k][j] * d[k];
          e[k] += V[k][j] * f;
        }
        e[j] = g;
      }
      f = 0.0;
      for (int j = 0; j < i; ++j) {
        e[j] /= h;
        f += e[j] * d[j];
      }
      double hh = f / (h + h);
      for (int j = 0; j < i; ++j) e[j] -= hh * d[j];
      for (int j = 0; j < i; ++j) {
        f = d[j];
        g = e[j];
        for (int k = j; k <= i - 1; ++k) V[k][j] -= (f * e[k] + g * d[k]);
        d[j] = V[i - 1][j];
        V[i][j] = 0.0;
      }
    }
    d[i] = h;
  }
  for (int i = 0; i < n - 1; ++i) {
    V[n - 1][i] = V[i][i];
    V[i][i] = 1.0;
    double h = d[i + 1];
    if (h != 0.0) {
      for (int k = 0; k <= i; ++k) d[k] = V[k][i + 1] / h;
      for (int j = 0; j <= i; ++j) {
        double g = 0.0;
        for (int k = 0; k <= i; ++k) g += V[k][i + 1] * V[k][j];
        for (int k = 0; k <= i; ++k) V[k][j] -= g * d[k];
      }
    }
    for (int k = 0; k <= i; ++k) V[k][i + 1] = 0.0;
  }
  for (int j = 0; j < n; ++j) {
    d[j] = V[n - 1][j];
    V[n - 1][j] = 0.0;
  }
  V[n - 1][n - 1] = 1.0;
  e[0] = 0.0;
  // tql2
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  double f = 0.0, tst1 = 0.0, eps = std::pow(2.0, -52.0);
  for (int l = 0; l < n; ++l) {
    tst1 = std::max(tst1, std::fabs(d[l]) + std::fabs(e[l]));
    int m = l;
    while (m < n) {
      if (std::fabs(e[m]) <= eps * tst1) break;
      ++m;
    }
    if (m > l) {
      int iter = 0;
      do {
        ++iter;
        double g = d[l];
        double p = (d[l + 1] - g) / (2.0 * e[l]);
        double r = std::hypot(p, 1.0);
        if (p < 0) r = -r;
        d[l] = e[l] / (p + r);
        d[l + 1] = e[l] * (p + r);
        double dl1 = d[l + 1], h = g - d[l];
        for (int i = l + 2; i < n; ++i) d[i] -= h;
        f += h;
        p = d[m];
        double c = 1.0, c2 = c, c3 = c, el1 = e[l + 1], s = 0.0, s2 = 0.0;
        for (int i = m - 1; i >= l; --i) {
          c3 = c2;
          c2 = c;
          s2 = s;
          g = c * e[i];
          h = c * p;
          r = std::hypot(p, e[i]);
          e[i + 1] = s * r;
          s = e[i] / r;
          c = p / r;
          p = c * d[i] - s * g;
          d[i + 1] = h + s * (c * g + s * d[i]);
          for (int k = 0; k < n; ++k) {
            h = V[k][i + 1];
            V[k][i + 1] = s * V[k][i] + c * h;
            V[k][i] = c * V[k][i] - s * h;
          }
        }
        p = -s * s2 * c3 * el1 * e[l] / dl1;
        e[l] = s * p;
        d[l] = c * p;
      } while (std::fabs(e[l]) > eps * tst1 && iter < 200);
    }
    d[l] += f;
    e[l] = 0.0;
  }
  w = d;
  vout.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) vout[(size_t)i * n + j] = V[i][j];
  sort_sign(n, w, vout);
}

void eigh(int n, const Mat& a, std::vector<double>& w, Mat& v) {  // src/matrix.rs:293-303
  if (n >= 8)
    eigh_ql(n, a, w, v);
  else
    eigh_jacobi(n, a, w, v);
}

// src/matrix.rs:329-346
int canonical_orthogonalizer(int n, const Mat& s, double cutoff, Mat& x, int& cols) {
  std::vector<double> w;
  Mat v;
  eigh(n, s, w, v);
  if (w[0] < -cutoff) return fail("overlap has negative eigenvalue", 2);
  std::vector<int> keep;
  for (int i = 0; i < n; ++i)
    if (w[i] >= cutoff) keep.push_back(i);
  if (keep.empty()) return fail("overlap has zero retained rank", 2);
  cols = (int)keep.size();
  x.assign((size_t)n * cols, 0.0);
  for (int c = 0; c < cols; ++c)
    for (int r = 0; r < n; ++r) x[(size_t)r * cols + c] = v[(size_t)r * n + keep[c]] / std::sqrt(w[keep[c]]);
  return 0;
}

// ---------------------------------------------------------------------------
// One-electron integrals: src/integrals.rs:407-537 and 705-793.
// ---------------------------------------------------------------------------
double prim_overlap(double al, const int* la, const double* A, double be, const int* lb, const double* Bc) {
  Herm e[3];
  for (int k = 0; k < 3; ++k) hermite(la[k], lb[k], al, be, A[k], Bc[k], e[k]);
  double inv_p = 1.0 / (al + be);
  return e[0].get(la[0], lb[0], 0) * e[1].get(la[1], lb[1], 0) * e[2].get(la[2], lb[2], 0) *
         (PI_POW_1_5 * inv_p * std::sqrt(inv_p));
}
double prim_kinetic(double al, const int* la, const double* A, double be, const int* lb, const double* Bc) {
  Herm e[3];
  for (int k = 0; k < 3; ++k) hermite(la[k], lb[k] + 2, al, be, A[k], Bc[k], e[k]);
  double inv_p = 1.0 / (al + be);
  auto t1d = [&](int k) {
    int li = la[k], lj = lb[k];
    double t1 = be * (2.0 * lj + 1.0) * e[k].get(li, lj, 0);
    double t2 = -2.0 * be * be * e[k].get(li, lj + 2, 0);
    double t3 = lj >= 2 ? -0.5 * (lj * (lj - 1)) * e[k].get(li, lj - 2, 0) : 0.0;
    return t1 + t2 + t3;
  };
  double sx = e[0].get(la[0], lb[0], 0), sy = e[1].get(la[1], lb[1], 0), sz = e[2].get(la[2], lb[2], 0);
  double pre = PI_POW_1_5 * inv_p * std::sqrt(inv_p);
  return pre * (t1d(0) * sy * sz + sx * t1d(1) * sz + sx * sy * t1d(2));
}
double contracted_nuclear(const CG& a, const CG& b, const std::vector<std::array<double, 4>>& atoms) {
  double total = 0.0;
  std::vector<double> r;
  for (auto& pa : a.prims)
    for (auto& pb : b.prims) {
      double coeff = pa.c * pb.c;
      if (std::fabs(coeff) < 1e-15) continue;
      double p = pa.e + pb.e, inv_p = 1.0 / p;
      double P[3];
      for (int k = 0; k < 3; ++k) P[k] = (pa.e * a.center[k] + pb.e * b.center[k]) * inv_p;
      Herm e[3];
      for (int k = 0; k < 3; ++k) hermite(a.l[k], b.l[k], pa.e, pb.e, a.center[k], b.center[k], e[k]);
      int mt = a.l[0] + b.l[0], mu = a.l[1] + b.l[1], mv = a.l[2] + b.l[2], mn = mt + mu + mv;
      size_t sv = mn + 1, su = (mv + 1) * sv, st = (mu + 1) * su;
      r.resize((mt + 1) * st);
      double pref = 2.0 * PI * inv_p;
      for (auto& at : atoms) {
        double pmc[3] = {P[0] - at[1], P[1] - at[2], P[2] - at[3]};
        coulomb_r(mt, mu, mv, mn, p, pmc, r.data());
        double sum = 0.0;
        for (int t = 0; t <= mt; ++t) {
          double et = e[0].get(a.l[0], b.l[0], t);
          if (et == 0.0) continue;
          for (int u = 0; u <= mu; ++u) {
            double eu = e[1].get(a.l[1], b.l[1], u);
            if (eu == 0.0) continue;
            for (int v = 0; v <= mv; ++v) {
              double ev = e[2].get(a.l[2], b.l[2], v);
              if (ev == 0.0) continue;
              sum += et * eu * ev * r[t * st + u * su + v * sv];
            }
          }
        }
        total += coeff * (-at[0] * pref * sum);
      }
    }
  return total;
}

void one_electron(const hfo_basis& B, const std::vector<std::array<double, 4>>& atoms, Mat& S, Mat& T, Mat& V) {
  int n = (int)B.f.size();
  S.assign((size_t)n * n, 0.0);
  T = S;
  V = S;
#pragma omp parallel for schedule(dynamic)
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) {
      const CG &a = B.f[i], &b = B.f[j];
      double s = 0, t = 0;
      for (auto& pa : a.prims)
        for (auto& pb : b.prims) {
          s += pa.c * pb.c * prim_overlap(pa.e, a.l, a.center, pb.e, b.l, b.center);
          t += pa.c * pb.c * prim_kinetic(pa.e, a.l, a.center, pb.e, b.l, b.center);
        }
      double v = contracted_nuclear(a, b, atoms);
      S[(size_t)i * n + j] = S[(size_t)j * n + i] = s;
      T[(size_t)i * n + j] = T[(size_t)j * n + i] = t;
      V[(size_t)i * n + j] = V[(size_t)j * n + i] = v;
    }
}

double contracted_overlap(const CG& a, const CG& b) {
  double s = 0;
  for (auto& pa : a.prims)
    for (auto& pb : b.prims) s += pa.c * pb.c * prim_overlap(pa.e, a.l, a.center, pb.e, b.l, b.center);
  return s;
}

// Real spherical transform: src/ao.rs:99-131 (polynomial) and 166-217.
double fact(int n) {
  double r = 1;
  for (int k = 2; k <= n; ++k) r *= k;
  return r;
}
std::map<std::array<int, 3>, double> solid_harmonic(int l, int m) {
  int am = std::abs(m);
  std::map<std::array<int, 3>, double> out;
  double norm = std::sqrt(fact(l - am) / fact(l + am)) * (am == 0 ? 1.0 : std::sqrt(2.0));
  for (int k = 0; k <= (l - am) / 2; ++k) {
    double coef = ((k & 1) ? -1.0 : 1.0) * fact(2 * l - 2 * k) / (std::pow(2.0, l) * fact(k) * fact(l - k) * fact(l - am - 2 * k));
    for (int y = 0; y <= am; ++y) {
      double phase = 0.0;
      int ym = y % 4;
      if ((m < 0 && ym == 1) || (m >= 0 && ym == 0)) phase = 1.0;
      if ((m < 0 && ym == 3) || (m >= 0 && ym == 2)) phase = -1.0;
      if (phase == 0.0) continue;
      double binom = fact(am) / (fact(y) * fact(am - y));
      for (int xx = 0; xx <= k; ++xx)
        for (int yy = 0; yy <= k - xx; ++yy) {
          int zz = k - xx - yy;
          std::array<int, 3> pw{am - y + 2 * xx, y + 2 * yy, l - am - 2 * k + 2 * zz};
          double multi = fact(k) / (fact(xx) * fact(yy) * fact(zz));
          out[pw] += norm * coef * binom * phase * multi;
        }
    }
  }
  return out;
}

int spherical_transform(const hfo_basis& B, Mat& T, int& ncol) {
  int n = (int)B.f.size(), ns = (int)B.shell_start.size() - 1;
  ncol = 0;
  for (int s = 0; s < ns; ++s) ncol += 2 * B.f[B.shell_start[s]].L() + 1;
  T.assign((size_t)n * ncol, 0.0);
  int col = 0;
  for (int s = 0; s < ns; ++s) {
    int st = B.shell_start[s], en = B.shell_start[s + 1], l = B.f[st].L(), cnt = en - st;
    if (l <= 1) {
      for (int r = st; r < en; ++r) T[(size_t)r * ncol + col++] = 1.0;
      continue;
    }
    Mat ov((size_t)cnt * cnt);
    for (int i = 0; i < cnt; ++i)
      for (int j = 0; j < cnt; ++j) ov[(size_t)i * cnt + j] = contracted_overlap(B.f[st + i], B.f[st + j]);
    for (int m = -l; m <= l; ++m) {
      auto p = solid_harmonic(l, m);
      std::vector<double> c(cnt);
      for (int i = 0; i < cnt; ++i) {
        const CG& g = B.f[st + i];
        double wgt = std::sqrt(dfact(2 * g.l[0] - 1) * dfact(2 * g.l[1] - 1) * dfact(2 * g.l[2] - 1));
        auto it = p.find({g.l[0], g.l[1], g.l[2]});
        c[i] = (it == p.end() ? 0.0 : it->second) * wgt;
      }
      double nrm = 0;
      for (int i = 0; i < cnt; ++i)
        for (int j = 0; j < cnt; ++j) nrm += c[i] * ov[(size_t)i * cnt + j] * c[j];
      if (nrm <= 0.0 || !std::isfinite(nrm)) return fail("invalid spherical normalization", 2);
      for (int i = 0; i < cnt; ++i) T[(size_t)(st + i) * ncol + col] = c[i] / std::sqrt(nrm);
      ++col;
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// J/K engines.
// ---------------------------------------------------------------------------
struct Engine {
  const hfo_basis* B;
  PairSet ps;
  bool incore = false;
  std::vector<double> eri;  // packed (in-core mode)
  int n;
};

void packed_tensor(const hfo_basis& B, const PairSet& ps, std::vector<double>& out) {  // src/shell_integrals.rs:174-240
  size_t n = B.f.size(), np = n * (n + 1) / 2;
  out.assign(np * (np + 1) / 2, 0.0);
  long npair = (long)ps.pairs.size();
#pragma omp parallel for schedule(dynamic)
  for (long ij = 0; ij < npair; ++ij) {
    std::vector<double> blk;
    const ShellPair& ab = ps.pairs[ij];
    for (long kl = 0; kl <= ij; ++kl) {
      if (!(ps.bound[ij] * ps.bound[kl] > 1e-12)) continue;
      const ShellPair& cd = ps.pairs[kl];
      quartet(B, ab, cd, false, blk);
      for (int i = 0; i < ab.na; ++i)
        for (int j = 0; j < ab.nb; ++j) {
          int mu = ab.fa + i, nu = ab.fb + j;
          if (mu < nu) continue;
          for (int k = 0; k < cd.na; ++k)
            for (int l = 0; l < cd.nb; ++l) {
              int la = cd.fa + k, si = cd.fb + l;
              if (la < si) continue;
              out[eri_index(mu, nu, la, si)] = blk[(size_t)(i * ab.nb + j) * cd.na * cd.nb + k * cd.nb + l];
            }
        }
    }
  }
}

void contract_jk(const double* P, const double* eri, int n, Mat& J, Mat& K) {  // src/scf.rs:337-364
  J.assign((size_t)n * n, 0.0);
  K = J;
  long np = (long)n * (n + 1) / 2;
  std::vector<std::pair<int, int>> pairs;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j <= i; ++j) pairs.push_back({i, j});
#pragma omp parallel
  {
    Mat j((size_t)n * n, 0.0), k((size_t)n * n, 0.0);
#pragma omp for schedule(dynamic, 16)
    for (long ij = 0; ij < np; ++ij) {
      int mu = pairs[ij].first, nu = pairs[ij].second;
      const double* row = eri + (size_t)ij * (ij + 1) / 2;
      for (long kl = 0; kl <= ij; ++kl) {
        double v = row[kl];
        if (v == 0.0) continue;
        digest(j.data(), k.data(), P, n, mu, nu, pairs[kl].first, pairs[kl].second, ij == kl, v);
      }
    }
#pragma omp critical
    for (size_t x = 0; x < J.size(); ++x) {
      J[x] += j[x];
      K[x] += k[x];
    }
  }
}

// Direct digestion of the same screened shell quartets (no stored tensor).
void direct_jk(const hfo_basis& B, const PairSet& ps, const double* P, Mat& J, Mat& K, int stride, int offset) {
  int n = (int)B.f.size();
  J.assign((size_t)n * n, 0.0);
  K = J;
  long npair = (long)ps.pairs.size();
#pragma omp parallel
  {
    Mat j((size_t)n * n, 0.0), k((size_t)n * n, 0.0);
    std::vector<double> blk;
#pragma omp for schedule(dynamic)
    for (long ij = npair - 1; ij >= 0; --ij) {
      if (stride > 1 && ij % stride != offset) continue;
      const ShellPair& ab = ps.pairs[ij];
      for (long kl = 0; kl <= ij; ++kl) {
        if (!(ps.bound[ij] * ps.bound[kl] > 1e-12)) continue;
        const ShellPair& cd = ps.pairs[kl];
        quartet(B, ab, cd, false, blk);
        int ncd = cd.na * cd.nb;
        for (int i = 0; i < ab.na; ++i)
          for (int jj = 0; jj < ab.nb; ++jj) {
            int mu = ab.fa + i, nu = ab.fb + jj;
            if (mu < nu) continue;
            size_t cmn = compound(mu, nu);
            for (int kk = 0; kk < cd.na; ++kk)
              for (int l = 0; l < cd.nb; ++l) {
                int la = cd.fa + kk, si = cd.fb + l;
                if (la < si) continue;
                size_t cls = compound(la, si);
                if (ij == kl && cls > cmn) continue;
                double v = blk[(size_t)(i * ab.nb + jj) * ncd + kk * cd.nb + l];
                if (v == 0.0) continue;
                digest(j.data(), k.data(), P, n, mu, nu, la, si, cmn == cls, v);
              }
          }
      }
    }
#pragma omp critical
    for (size_t x = 0; x < J.size(); ++x) {
      J[x] += j[x];
      K[x] += k[x];
    }
  }
}

void fock_from(const Engine& E, const Mat& H, const Mat& P, Mat& F) {  // src/scf.rs:366-369
  int n = E.n;
  Mat J, K;
  if (E.incore)
    contract_jk(P.data(), E.eri.data(), n, J, K);
  else
    direct_jk(*E.B, E.ps, P.data(), J, K, 1, 0);
  F.resize((size_t)n * n);
  for (size_t x = 0; x < F.size(); ++x) F[x] = H[x] + J[x] - 0.5 * K[x];
}

// ---------------------------------------------------------------------------
// DIIS family: src/diis.rs:53-137 (CDIIS), 195-266 (simplex), 289-359 (ADIIS),
// 532-589 (Gaussian elimination).
// ---------------------------------------------------------------------------
Mat diis_error(const Mat& F, const Mat& P, const Mat& S, int n) {
  Mat fps = mm(mm(F, P, n, n, n), S, n, n, n);
  Mat spf = mm(mm(S, P, n, n, n), F, n, n, n);
  for (size_t x = 0; x < fps.size(); ++x) fps[x] -= spf[x];
  return fps;
}
double max_abs(const Mat& m) {
  double r = 0;
  for (double x : m) r = std::max(r, std::fabs(x));
  return r;
}
bool solve_linear(Mat a, std::vector<double> b, int n, std::vector<double>& x) {
  for (int i = 0; i < n; ++i) {
    int mr = i;
    double mv = std::fabs(a[(size_t)i * n + i]);
    for (int k = i + 1; k < n; ++k)
      if (std::fabs(a[(size_t)k * n + i]) > mv) {
        mv = std::fabs(a[(size_t)k * n + i]);
        mr = k;
      }
    if (mv < 1e-28) return false;
    if (mr != i) {
      std::swap(b[i], b[mr]);
      for (int c = 0; c < n; ++c) std::swap(a[(size_t)i * n + c], a[(size_t)mr * n + c]);
    }
    double dg = a[(size_t)i * n + i];
    for (int j = i + 1; j < n; ++j) {
      double f = a[(size_t)j * n + i] / dg;
      a[(size_t)j * n + i] = f;
      for (int k = i + 1; k < n; ++k) a[(size_t)j * n + k] -= f * a[(size_t)i * n + k];
      b[j] -= f * b[i];
    }
  }
  x.assign(n, 0);
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < n; ++k) s -= a[(size_t)i * n + k] * x[k];
    x[i] = s / a[(size_t)i * n + i];
  }
  return true;
}
double frob(const Mat& a, const Mat& b) {
  double s = 0;
  for (size_t x = 0; x < a.size(); ++x) s += a[x] * b[x];
  return s;
}

struct Cdiis {
  size_t maxs = 8;
  std::vector<Mat> f, e;
  void push(const Mat& F, const Mat& E) {
    if (f.size() >= maxs) {
      f.erase(f.begin());
      e.erase(e.begin());
    }
    f.push_back(F);
    e.push_back(E);
  }
  // Pulay coefficients: the bordered system [B -1; -1 0] [c; l] = [0; -1]
  // with B_ij = <e_i, e_j> (src/diis.rs:72-137).
  bool coefficients(std::vector<double>& c) const {
    int m = (int)e.size();
    if (m < 2) return false;
    int dim = m + 1;
    Mat b((size_t)dim * dim, 0.0);
    for (int i = 0; i < m; ++i) {
      for (int j = 0; j <= i; ++j) {
        double d = frob(e[i], e[j]);
        b[(size_t)i * dim + j] = b[(size_t)j * dim + i] = d;
      }
      b[(size_t)i * dim + m] = -1.0;
      b[(size_t)m * dim + i] = -1.0;
    }
    std::vector<double> rhs(dim, 0.0);
    rhs[m] = -1.0;
    if (!solve_linear(b, rhs, dim, c)) return false;
    c.resize(m);
    return true;
  }
  bool extrapolate(Mat& out) const {
    std::vector<double> c;
    if (!coefficients(c)) return false;
    const int m = (int)c.size();
    out.assign(f[0].size(), 0.0);
    for (int i = 0; i < m; ++i)
      for (size_t k = 0; k < out.size(); ++k) out[k] += c[i] * f[i][k];
    return true;
  }
};

void project_simplex(std::vector<double>& v) {
  std::vector<double> u = v;
  std::stable_sort(u.begin(), u.end(), [](double a, double b) { return a > b; });
  double css = 0, theta = 0;
  for (size_t k = 0; k < u.size(); ++k) {
    css += u[k];
    double t = (css - 1.0) / (double)(k + 1);
    if (u[k] - t > 0.0) theta = t;
  }
  for (auto& x : v) x = std::max(x - theta, 0.0);
}
std::vector<double> min_quad_simplex(const std::vector<std::vector<double>>& q, const std::vector<double>& b,
                                     const std::vector<double>& start) {
  int m = (int)b.size();
  double lip = 1e-12;
  for (int i = 0; i < m; ++i) {
    double s = 0;
    for (int j = 0; j < m; ++j) s += std::fabs(0.5 * (q[i][j] + q[j][i]));
    lip = std::max(lip, s);
  }
  double step = 1.0 / lip;
  std::vector<double> c = start;
  project_simplex(c);
  for (int it = 0; it < 5000; ++it) {
    std::vector<double> nx(m);
    for (int i = 0; i < m; ++i) {
      double g = b[i];
      for (int j = 0; j < m; ++j) g += 0.5 * (q[i][j] + q[j][i]) * c[j];
      nx[i] = c[i] - step * g;
    }
    project_simplex(nx);
    double diff = 0;
    for (int i = 0; i < m; ++i) diff = std::max(diff, std::fabs(nx[i] - c[i]));
    c = nx;
    if (diff < 1e-14) break;
  }
  return c;
}
struct Adiis {
  size_t maxs = 8;
  std::vector<Mat> f, p;
  std::vector<double> en;
  void push(const Mat& F, const Mat& P, double E) {
    if (f.size() >= maxs) {
      f.erase(f.begin());
      p.erase(p.begin());
      en.erase(en.begin());
    }
    f.push_back(F);
    p.push_back(P);
    en.push_back(E);
  }
  // The quadratic model f(c) = E_n + b^T c + 1/2 c^T Q c of the energy of
  // sum_i c_i P_i about the newest entry n (src/diis.rs:325-333).
  void model(std::vector<std::vector<double>>& q, std::vector<double>& b) const {
    int m = (int)f.size();
    int nl = m - 1;
    std::vector<Mat> dp(m), df(m);
    for (int i = 0; i < m; ++i) {
      dp[i] = p[i];
      df[i] = f[i];
      for (size_t k = 0; k < dp[i].size(); ++k) {
        dp[i][k] -= p[nl][k];
        df[i][k] -= f[nl][k];
      }
    }
    b.assign(m, 0.0);
    q.assign(m, std::vector<double>(m));
    for (int i = 0; i < m; ++i) {
      b[i] = frob(dp[i], f[nl]);
      for (int j = 0; j < m; ++j) q[i][j] = frob(dp[i], df[j]);
    }
  }
  bool extrapolate(Mat& out) const {
    int m = (int)f.size();
    if (m < 2) return false;
    std::vector<double> b;
    std::vector<std::vector<double>> q;
    model(q, b);
    std::vector<double> st(m, 0.0);
    st[m - 1] = 1.0;
    auto c = min_quad_simplex(q, b, st);
    out.assign(f[0].size(), 0.0);
    for (int i = 0; i < m; ++i)
      for (size_t k = 0; k < out.size(); ++k) out[k] += c[i] * f[i][k];
    return true;
  }
};

Mat occupied_density(const Mat& C, int rows, int cols, int nocc, double scale) {  // src/matrix.rs:107-110
  Mat d((size_t)rows * rows, 0.0);
  for (int i = 0; i < rows; ++i)
    for (int j = 0; j < rows; ++j) {
      double s = 0;
      for (int k = 0; k < nocc; ++k) s += C[(size_t)i * cols + k] * C[(size_t)j * cols + k];
      d[(size_t)i * rows + j] = s * scale;
    }
  return d;
}
double rmsd(const Mat& a, const Mat& b) {
  double s = 0;
  for (size_t x = 0; x < a.size(); ++x) s += (a[x] - b[x]) * (a[x] - b[x]);
  return std::sqrt(s / (double)a.size());
}

// X^T F X eigh and back-transform: C = X C'.
void diag_in(const Mat& F, const Mat& X, int n, int m, std::vector<double>& eps, Mat& C) {
  Mat xt = tr(X, n, m);
  Mat fp = mm(mm(xt, F, m, n, n), X, m, n, m);
  Mat cp;
  eigh(m, fp, eps, cp);
  C = mm(X, cp, n, m, m);
}

// ---------------------------------------------------------------------------
// SAD: src/sad.rs:24-134 (atomic fractional HF), 217-259 (superposition).
// ---------------------------------------------------------------------------
std::vector<double> occupations(const std::vector<double>& e, double nel) {
  if (nel < 0.0 || nel > 2.0 * e.size() + 1e-9) throw std::runtime_error("Atomic electron count exceeds retained basis capacity");
  std::vector<double> occ(e.size(), 0.0);
  double left = nel;
  size_t i = 0;
  while (i < e.size() && left > 1e-12) {
    size_t end = i + 1;
    while (end < e.size() && std::fabs(e[end] - e[i]) < 1e-6) ++end;
    double pop = std::min(left, 2.0 * (end - i));
    for (size_t k = i; k < end; ++k) occ[k] = pop / (end - i);
    left -= pop;
    i = end;
  }
  return occ;
}
Mat fractional_density(const Mat& C, int rows, int cols, const std::vector<double>& occ) {
  Mat w = C;
  for (int r = 0; r < rows; ++r)
    for (int c = 0; c < cols; ++c) w[(size_t)r * cols + c] *= std::sqrt(occ[c]);
  return mm(w, tr(w, rows, cols), rows, cols, rows);
}

struct AtomicResult {
  Mat density;
  double energy;
  int iterations;
};

AtomicResult atomic_fractional_hf(const std::vector<CG>& fns, double z) {
  hfo_basis B;
  B.f = fns;
  for (auto& g : B.f) {
    g.center[0] = g.center[1] = g.center[2] = 0.0;
    g.atom = 0;
  }
  if (compute_shells(B)) throw std::runtime_error(g_err);
  int n = (int)B.f.size();
  std::vector<std::array<double, 4>> atoms{{z, 0, 0, 0}};
  Mat S, T, V;
  one_electron(B, atoms, S, T, V);
  Mat H(S.size());
  for (size_t x = 0; x < H.size(); ++x) H[x] = T[x] + V[x];
  Mat X;
  int m;
  if (canonical_orthogonalizer(n, S, 1e-8, X, m)) throw std::runtime_error(g_err);
  Engine E;
  E.B = &B;
  E.n = n;
  build_pairs(B, E.ps);
  packed_tensor(B, E.ps, E.eri);
  E.incore = true;
  auto afock = [&](const Mat& P) {
    if (std::fabs(z - 1.0) < 1e-12) return H;
    Mat F;
    fock_from(E, H, P, F);
    return F;
  };
  std::vector<double> eps;
  Mat C;
  diag_in(H, X, n, m, eps, C);
  auto occ = occupations(eps, z);
  Mat dens = fractional_density(C, n, m, occ);
  double prev = INFINITY;
  Cdiis diis;
  for (int it = 1; it <= 200; ++it) {
    Mat F = afock(dens);
    double en = 0;
    for (size_t x = 0; x < F.size(); ++x) en += dens[x] * (H[x] + F[x]);
    en *= 0.5;
    diag_in(F, X, n, m, eps, C);
    occ = occupations(eps, z);
    Mat next = fractional_density(C, n, m, occ);
    Mat err = diis_error(F, dens, S, n);
    if (std::fabs(en - prev) < 1e-11 && rmsd(dens, next) < 1e-9 && max_abs(err) < 1e-8) {
      Mat Fn = afock(next);
      double e2 = 0;
      for (size_t x = 0; x < Fn.size(); ++x) e2 += next[x] * (H[x] + Fn[x]);
      return {next, 0.5 * e2, it};
    }
    prev = en;
    diis.push(F, err);
    Mat fu;
    if (!diis.extrapolate(fu)) fu = F;
    diag_in(fu, X, n, m, eps, C);
    occ = occupations(eps, z);
    next = fractional_density(C, n, m, occ);
    if (it <= 5)
      for (size_t x = 0; x < dens.size(); ++x) dens[x] = 0.5 * dens[x] + 0.5 * next[x];
    else
      dens = next;
  }
  throw std::runtime_error("Fractional atomic HF did not converge");
}

std::string atom_key(const std::vector<CG>& fns, double z) {
  std::ostringstream os;
  os.precision(17);
  os << z << "|" << fns.size();
  for (auto& g : fns) {
    os << "|" << g.l[0] << g.l[1] << g.l[2] << ":" << g.prims.size();
    for (auto& p : g.prims) os << "," << p.e << "," << p.c;
  }
  return os.str();
}

int sad_density(const hfo_basis& B, const std::vector<std::array<double, 4>>& atoms, int charge, const Mat& S, Mat& P) {
  static std::mutex mu;
  static std::map<std::string, AtomicResult> cache;  // src/sad.rs:136-214
  int n = (int)B.f.size();
  double target = -charge;
  for (auto& a : atoms) target += a[0];
  if (target <= 0.0) return fail("Invalid SAD electron count or overlap");
  P.assign((size_t)n * n, 0.0);
  for (size_t a = 0; a < atoms.size(); ++a) {
    std::vector<int> idx;
    for (int i = 0; i < n; ++i)
      if (B.f[i].atom == (int)a) idx.push_back(i);
    if (idx.empty()) return fail("No basis functions for atom");
    std::vector<CG> loc;
    for (int i : idx) loc.push_back(B.f[i]);
    std::string key = atom_key(loc, atoms[a][0]);
    AtomicResult ar;
    {
      std::lock_guard<std::mutex> lk(mu);
      auto it = cache.find(key);
      if (it != cache.end()) ar = it->second;
    }
    if (ar.density.empty()) {
      ar = atomic_fractional_hf(loc, atoms[a][0]);
      std::lock_guard<std::mutex> lk(mu);
      cache[key] = ar;
    }
    int m = (int)idx.size();
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < m; ++j) P[(size_t)idx[i] * n + idx[j]] = ar.density[(size_t)i * m + j];
  }
  double el = 0;
  for (size_t x = 0; x < P.size(); ++x) el += P[x] * S[x];
  if (el <= 1e-12) return fail("Zero SAD population");
  for (auto& x : P) x *= target / el;
  return 0;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

std::vector<std::array<double, 4>> atoms_of(const double* zxyz, int natom) {
  std::vector<std::array<double, 4>> a(natom);
  for (int i = 0; i < natom; ++i)
    for (int k = 0; k < 4; ++k) a[i][k] = zxyz[4 * i + k];
  return a;
}

// ---------------------------------------------------------------------------
// DF: src/df.rs:16-174.
// ---------------------------------------------------------------------------
Mat coulomb_metric(const hfo_basis& aux) {
  int n = (int)aux.f.size();
  Mat m((size_t)n * n);
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < n; ++p) {
    std::vector<double> r;
    auto pp = basis_pair_data(aux.f[p], constant_fn(aux.f[p].center));
    for (int q = 0; q < n; ++q) {
      auto qq = basis_pair_data(aux.f[q], constant_fn(aux.f[q].center));
      m[(size_t)p * n + q] = pair_eri(pp, qq, r);
    }
  }
  return m;
}
Mat three_center(const hfo_basis& B, const hfo_basis& aux) {
  int n = (int)B.f.size(), na = (int)aux.f.size();
  size_t np = (size_t)n * (n + 1) / 2;
  Mat t((size_t)na * np);
  std::vector<std::vector<PPData>> mn(np);
  for (int mu = 0; mu < n; ++mu)
    for (int nu = 0; nu <= mu; ++nu) mn[compound(mu, nu)] = basis_pair_data(B.f[mu], B.f[nu]);
#pragma omp parallel for schedule(dynamic)
  for (int p = 0; p < na; ++p) {
    std::vector<double> r;
    auto pp = basis_pair_data(aux.f[p], constant_fn(aux.f[p].center));
    for (size_t k = 0; k < np; ++k) t[(size_t)p * np + k] = pair_eri(mn[k], pp, r);
  }
  return t;
}

int df_jk(const double* bt, int rank, int n, const double* co, int no, Mat& J, Mat& K) {  // src/df.rs:133-174
  for (int x = 0; x < n * no; ++x)
    if (!std::isfinite(co[x])) return fail("invalid occupied coefficients for DF J/K");
  size_t np = (size_t)n * (n + 1) / 2;
  J.assign((size_t)n * n, 0.0);
  K = J;
#pragma omp parallel
  {
    Mat j((size_t)n * n, 0.0), k((size_t)n * n, 0.0), b((size_t)n * n), act((size_t)n * no);
#pragma omp for schedule(dynamic)
    for (int q = 0; q < rank; ++q) {
      for (int p = 0; p < n; ++p)
        for (int s = 0; s < n; ++s) b[(size_t)p * n + s] = bt[(size_t)q * np + compound(p, s)];
      std::fill(act.begin(), act.end(), 0.0);
      for (int p = 0; p < n; ++p)
        for (int s = 0; s < n; ++s) {
          double bv = b[(size_t)p * n + s];
          for (int i = 0; i < no; ++i) act[(size_t)p * no + i] += bv * co[(size_t)s * no + i];
        }
      double charge = 0;
      for (size_t x = 0; x < act.size(); ++x) charge += act[x] * co[x];
      charge *= 2.0;
      for (size_t x = 0; x < b.size(); ++x) j[x] += charge * b[x];
      for (int p = 0; p < n; ++p)
        for (int s = 0; s < n; ++s) {
          double acc = 0;
          for (int i = 0; i < no; ++i) acc += act[(size_t)p * no + i] * act[(size_t)s * no + i];
          k[(size_t)p * n + s] += 2.0 * acc;
        }
    }
#pragma omp critical
    for (size_t x = 0; x < J.size(); ++x) {
      J[x] += j[x];
      K[x] += k[x];
    }
  }
  return 0;
}

int df_build(const hfo_basis& B, const hfo_basis& aux, double cutoff, int& rank, Mat& bt) {  // src/df.rs:89-127
  if (B.f.empty() || aux.f.empty()) return fail("DF requires orbital and auxiliary bases");
  if (!std::isfinite(cutoff) || cutoff <= 0) return fail("DF metric cutoff must be finite and positive");
  int na = (int)aux.f.size(), n = (int)B.f.size();
  size_t np = (size_t)n * (n + 1) / 2;
  Mat metric = coulomb_metric(aux), X;
  if (canonical_orthogonalizer(na, metric, cutoff, X, rank)) return 2;
  Mat three = three_center(B, aux);
  bt = mm(tr(X, na, rank), three, rank, na, (int)np);
  return 0;
}

}  // namespace

// ===========================================================================
// C interface
// ===========================================================================
extern "C" {

const char* hfo_last_error(void) { return g_err.c_str(); }

int hfo_boys(int m, double t, double* out) {
  try {
    boys_array(m, t, out);
    return 0;
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

int hfo_basis_build(const char* g94, const char* symbols, const double* xyz, int natom, hfo_basis** out) {
  try {
    auto map = parse_g94(g94);
    std::istringstream ss(symbols);
    std::vector<std::string> sym;
    std::string s;
    while (ss >> s) sym.push_back(s);
    if ((int)sym.size() != natom) return fail("symbol count does not match atom count");
    auto* b = new hfo_basis;
    for (int a = 0; a < natom; ++a) {
      std::string u = sym[a];
      for (auto& ch : u) ch = (char)toupper(ch);
      auto it = map.find(u);
      if (it == map.end()) {
        delete b;
        return fail("missing basis for " + sym[a]);
      }
      const double* c = xyz + 3 * a;
      for (auto& sh : it->second) {  // src/basis.rs:402-452
        if (sh.L == -1) {
          int l0[3] = {0, 0, 0};
          b->f.push_back(make_cg(a, c, l0, sh.exps, sh.coefs));
          for (auto& comp : cart_components(1)) b->f.push_back(make_cg(a, c, comp.data(), sh.exps, sh.pcoefs));
        } else {
          for (auto& comp : cart_components(sh.L)) b->f.push_back(make_cg(a, c, comp.data(), sh.exps, sh.coefs));
        }
      }
    }
    if (compute_shells(*b)) {
      delete b;
      return 1;
    }
    *out = b;
    return 0;
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

void hfo_basis_free(hfo_basis* b) { delete b; }
int hfo_basis_nfunc(const hfo_basis* b) { return (int)b->f.size(); }
int hfo_basis_nprim_total(const hfo_basis* b) {
  int n = 0;
  for (auto& g : b->f) n += (int)g.prims.size();
  return n;
}
int hfo_basis_export(const hfo_basis* b, int32_t* atom_index, double* center, int32_t* lxyz, int32_t* prim_offset,
                     double* exponent, double* coefficient) {
  int off = 0;
  for (size_t i = 0; i < b->f.size(); ++i) {
    const CG& g = b->f[i];
    atom_index[i] = g.atom;
    for (int k = 0; k < 3; ++k) {
      center[3 * i + k] = g.center[k];
      lxyz[3 * i + k] = g.l[k];
    }
    prim_offset[i] = off;
    for (auto& p : g.prims) {
      exponent[off] = p.e;
      coefficient[off] = p.c;
      ++off;
    }
  }
  prim_offset[b->f.size()] = off;
  return 0;
}
int hfo_basis_from_arrays(int nfunc, const int32_t* atom_index, const double* center, const int32_t* lxyz,
                          const int32_t* prim_offset, const double* exponent, const double* coefficient,
                          hfo_basis** out) {
  auto* b = new hfo_basis;
  for (int i = 0; i < nfunc; ++i) {
    CG g;
    g.atom = atom_index[i];
    for (int k = 0; k < 3; ++k) {
      g.center[k] = center[3 * i + k];
      g.l[k] = lxyz[3 * i + k];
    }
    for (int p = prim_offset[i]; p < prim_offset[i + 1]; ++p) g.prims.push_back({exponent[p], coefficient[p]});
    b->f.push_back(g);
  }
  if (compute_shells(*b)) {
    delete b;
    return 1;
  }
  *out = b;
  return 0;
}
int hfo_nshells(const hfo_basis* b, int32_t* ss) {
  int ns = (int)b->shell_start.size() - 1;
  if (ss)
    for (int i = 0; i <= ns; ++i) ss[i] = b->shell_start[i];
  return ns;
}

int hfo_shell_quartet(const hfo_basis* b, int sa, int sb, int sc, int sd, double* out) {
  int ns = (int)b->shell_start.size() - 1;
  for (int s : {sa, sb, sc, sd})
    if (s < 0 || s >= ns) return fail("shell index out of range");
  ShellPair ab = make_pair(*b, sa, sb), cd = make_pair(*b, sc, sd);
  std::vector<double> r;
  quartet(*b, ab, cd, false, r);
  std::memcpy(out, r.data(), r.size() * sizeof(double));
  return 0;
}

double hfo_contracted_eri(const hfo_basis* b, int i, int j, int k, int l) {
  std::vector<double> r;
  return pair_eri(basis_pair_data(b->f[i], b->f[j]), basis_pair_data(b->f[k], b->f[l]), r);
}

int hfo_eri_packed(const hfo_basis* b, double* out, size_t len) {
  size_t n = b->f.size(), np = n * (n + 1) / 2;
  if (len != np * (np + 1) / 2) return fail("packed length mismatch");
  PairSet ps;
  build_pairs(*b, ps);
  std::vector<double> t;
  packed_tensor(*b, ps, t);
  std::memcpy(out, t.data(), len * sizeof(double));
  return 0;
}

int hfo_schwarz(const hfo_basis* b, double* q) {
  PairSet ps;
  build_pairs(*b, ps);
  std::memcpy(q, ps.bound.data(), ps.bound.size() * sizeof(double));
  return 0;
}

int hfo_contract_jk(const double* p, const double* eri, int n, double* j, double* k) {
  Mat J, K;
  contract_jk(p, eri, n, J, K);
  std::memcpy(j, J.data(), J.size() * 8);
  std::memcpy(k, K.data(), K.size() * 8);
  return 0;
}

int hfo_direct_jk(const hfo_basis* b, const double* p, double* j, double* k, int stride, int offset,
                  double* seconds) {
  auto t0 = std::chrono::steady_clock::now();
  PairSet ps;
  build_pairs(*b, ps);
  Mat J, K;
  direct_jk(*b, ps, p, J, K, std::max(1, stride), offset);
  std::memcpy(j, J.data(), J.size() * 8);
  std::memcpy(k, K.data(), K.size() * 8);
  if (seconds) *seconds = seconds_since(t0);
  return 0;
}

int hfo_one_electron(const hfo_basis* b, const double* zxyz, int natom, double* s, double* t, double* v) {
  Mat S, T, V;
  one_electron(*b, atoms_of(zxyz, natom), S, T, V);
  std::memcpy(s, S.data(), S.size() * 8);
  std::memcpy(t, T.data(), T.size() * 8);
  std::memcpy(v, V.data(), V.size() * 8);
  return 0;
}

int hfo_eigh(const double* a, int n, double* w, double* v) {
  Mat A(a, a + (size_t)n * n), V;
  std::vector<double> W;
  eigh(n, A, W, V);
  std::memcpy(w, W.data(), n * 8);
  std::memcpy(v, V.data(), V.size() * 8);
  return 0;
}

int hfo_spherical_transform(const hfo_basis* b, double* t, int* ncols) {
  Mat T;
  int nc;
  if (spherical_transform(*b, T, nc)) return 2;
  if (t) std::memcpy(t, T.data(), T.size() * 8);
  *ncols = nc;
  return 0;
}

int hfo_sad_density(const hfo_basis* b, const double* zxyz, int natom, int charge, const double* s, double* p_out) {
  try {
    int n = (int)b->f.size();
    Mat S(s, s + (size_t)n * n), P;
    if (sad_density(*b, atoms_of(zxyz, natom), charge, S, P)) return 1;
    std::memcpy(p_out, P.data(), P.size() * 8);
    return 0;
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

int hfo_atomic_hf(const hfo_basis* b, double z, double* energy, double* density) {
  try {
    auto r = atomic_fractional_hf(b->f, z);
    *energy = r.energy;
    if (density) std::memcpy(density, r.density.data(), r.density.size() * 8);
    return 0;
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

// RHF driver: src/scf.rs:410-627.
int hfo_scf_rhf(const hfo_basis* b, const double* zxyz, int natom, int charge, const hfo_scf_options* opt,
                hfo_scf_result* res, double* p_out, double* f_out, double* eps_out) {
  try {
    auto t_all = std::chrono::steady_clock::now();
    std::memset(res, 0, sizeof(*res));
    int n = (int)b->f.size();
    double tol = opt->energy_tol;
    if (opt->max_iter == 0 || n == 0 || !(tol > 0) || !(opt->residual_tol > 0) || !(opt->density_tol > 0))
      return fail("SCF requires a nonempty basis, positive tolerance and iteration limit");
    auto atoms = atoms_of(zxyz, natom);
    double zt = 0;
    for (auto& a : atoms) zt += a[0];
    int nel = (int)std::lround(zt) - charge;
    if (nel <= 0) return fail("Invalid electron count");
    if (nel % 2) return fail("Molecule has an odd number of electrons. RHF requires an even number of electrons.");
    int nocc = nel / 2;
    if (nocc > n) return fail("More occupied orbitals than basis functions");
    double enuc = 0;  // src/scf.rs:206-219
    for (int i = 0; i < natom; ++i)
      for (int j = i + 1; j < natom; ++j) {
        double dx = atoms[i][1] - atoms[j][1], dy = atoms[i][2] - atoms[j][2], dz = atoms[i][3] - atoms[j][3];
        enuc += atoms[i][0] * atoms[j][0] / std::sqrt(dx * dx + dy * dy + dz * dz);
      }
    Mat S, T, V;
    one_electron(*b, atoms, S, T, V);
    Mat H(S.size());
    for (size_t x = 0; x < H.size(); ++x) H[x] = T[x] + V[x];
    // AoSpace::orthogonalizer, src/ao.rs:58-67
    Mat Tr;
    int m0 = n;
    if (opt->spherical) {
      if (spherical_transform(*b, Tr, m0)) return 2;
    } else {
      Tr.assign((size_t)n * n, 0.0);
      for (int i = 0; i < n; ++i) Tr[(size_t)i * n + i] = 1.0;
    }
    Mat Sp = mm(mm(tr(Tr, n, m0), S, m0, n, n), Tr, m0, n, m0), Xp;
    int m;
    if (canonical_orthogonalizer(m0, Sp, 1e-8, Xp, m)) return 2;
    Mat X = mm(Tr, Xp, n, m0, m);
    if (nocc > m) return fail("More occupied orbitals than retained overlap rank");
    Engine E;
    E.B = b;
    E.n = n;
    auto t0 = std::chrono::steady_clock::now();
    build_pairs(*b, E.ps);
    E.incore = !opt->direct;
    if (E.incore) packed_tensor(*b, E.ps, E.eri);
    res->seconds_eri = seconds_since(t0);
    double t_fock = 0;
    int builds = 0;
    auto fock = [&](const Mat& P) {
      auto tf = std::chrono::steady_clock::now();
      Mat F;
      fock_from(E, H, P, F);
      t_fock += seconds_since(tf);
      ++builds;
      return F;
    };
    Mat P, C;
    std::vector<double> eps;
    Mat sad;
    if (sad_density(*b, atoms, charge, S, sad) == 0) {
      Mat fi = fock(sad);
      diag_in(fi, X, n, m, eps, C);
      P = sad;
    } else {
      diag_in(H, X, n, m, eps, C);
      P = occupied_density(C, n, m, nocc, 2.0);
    }
    double prev_e = 0;
    bool conv = false;
    Mat F = H;
    int iters = 0;
    Cdiis cd;
    Adiis ad;
    Mat prevP = P;
    for (int it = 1; it <= opt->max_iter; ++it) {
      iters = it;
      F = fock(P);
      double ee = 0;
      for (size_t x = 0; x < F.size(); ++x) ee += 0.5 * P[x] * (H[x] + F[x]);
      double tot = ee + enuc, de = std::fabs(tot - prev_e);
      prev_e = tot;
      Mat err = diis_error(F, P, S, n);
      if (m < n) err = mm(mm(tr(X, n, m), err, m, n, n), X, m, n, m);
      double en = max_abs(err);
      if (opt->verbose) fprintf(stderr, "  oracle iter %3d  E=%.12f  dE=%.3e  |e|=%.3e\n", it, tot, de, en);
      if (it > 1 && de < tol && en < opt->residual_tol && rmsd(P, prevP) < opt->density_tol) {
        conv = true;
        break;
      }
      prevP = P;
      Mat ex;
      bool ok;
      if (opt->conv_algo == 1) {
        cd.push(F, err);
        ok = cd.extrapolate(ex);
      } else if (opt->conv_algo == 2) {
        ad.push(F, P, tot);
        ok = ad.extrapolate(ex);
      } else {
        double sw = opt->conv_algo == 3 ? 0.08 : 0.05;
        if (en > sw || it <= 2) {
          ad.push(F, P, tot);
          ok = ad.extrapolate(ex);
        } else {
          cd.push(F, err);
          ok = cd.extrapolate(ex);
        }
      }
      diag_in(ok ? ex : F, X, n, m, eps, C);
      P = occupied_density(C, n, m, nocc, 2.0);
    }
    if (!conv) F = fock(P);
    diag_in(F, X, n, m, eps, C);
    P = occupied_density(C, n, m, nocc, 2.0);
    F = fock(P);
    double fe = 0;
    for (size_t x = 0; x < F.size(); ++x) fe += P[x] * (H[x] + F[x]);
    fe *= 0.5;
    res->total_energy = fe + enuc;
    res->electronic_energy = fe;
    res->nuclear_repulsion = enuc;
    res->iterations = iters;
    res->converged = conv;
    res->n_occ = nocc;
    res->fock_builds = builds;
    res->seconds_fock = t_fock;
    res->seconds_total = seconds_since(t_all);
    if (p_out) std::memcpy(p_out, P.data(), P.size() * 8);
    if (f_out) std::memcpy(f_out, F.data(), F.size() * 8);
    if (eps_out) std::memcpy(eps_out, eps.data(), eps.size() * 8);
    return 0;
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

int hfo_df_metric(const hfo_basis* aux, double* metric) {
  Mat m = coulomb_metric(*aux);
  std::memcpy(metric, m.data(), m.size() * 8);
  return 0;
}
int hfo_df_three_center(const hfo_basis* b, const hfo_basis* aux, double* three) {
  Mat t = three_center(*b, *aux);
  std::memcpy(three, t.data(), t.size() * 8);
  return 0;
}
int hfo_df_build(const hfo_basis* b, const hfo_basis* aux, double cutoff, int* rank, double** b_out) {
  Mat bt;
  int r;
  if (int st = df_build(*b, *aux, cutoff, r, bt)) return st;
  *rank = r;
  *b_out = (double*)malloc(bt.size() * 8);
  std::memcpy(*b_out, bt.data(), bt.size() * 8);
  return 0;
}
void hfo_free(void* p) { free(p); }
int hfo_df_jk(const double* bt, int rank, int n, const double* cocc, int nocc, double* j, double* k) {
  Mat J, K;
  if (df_jk(bt, rank, n, cocc, nocc, J, K)) return 1;
  std::memcpy(j, J.data(), J.size() * 8);
  std::memcpy(k, K.data(), K.size() * 8);
  return 0;
}

// Pivoted Cholesky of the AO two-electron supermatrix: src/cholesky.rs:44-192.
// The diagonal (pq|pq) comes from the (AB|AB) shell quartets; each step takes
// the largest residual diagonal as the pivot (values within 1e-12 relative of
// it are ties, resolved by the smallest packed pair index), the pivot's
// column (pq|rs) from the complete shell quartets of the pivot's shell pair
// (cached while consecutive pivots share it), and forms the new factor row
// L(pq) = ((pq|rs) - sum_K L_K(pq) L_K(rs)) / sqrt(D_rs), L(rs) = sqrt(D_rs);
// the residual diagonal drops by L^2.  It stops when the largest residual is
// at most the threshold (or every pair is a pivot).  bt: rank x np, packed.
int cholesky_eri(const hfo_basis& B, double threshold, double budget, int& rank, Mat& bt, double& max_residual,
                 std::vector<int>& pivots) {
  const int n = (int)B.f.size();
  if (n == 0 || !std::isfinite(threshold) || threshold < 0)
    return fail("Cholesky requires a basis and a finite, nonnegative threshold");
  const int ns = (int)B.shell_start.size() - 1;
  const size_t np = (size_t)n * (n + 1) / 2;
  std::vector<int> shell_of(n);
  for (int s = 0; s < ns; ++s)
    for (int p = B.shell_start[s]; p < B.shell_start[s + 1]; ++p) shell_of[p] = s;
  const size_t nsp = (size_t)ns * (ns + 1) / 2;
  std::vector<std::pair<int, int>> sp_list(nsp);
  for (int a = 0; a < ns; ++a)
    for (int b = 0; b <= a; ++b) sp_list[(size_t)a * (a + 1) / 2 + b] = {a, b};
  std::vector<size_t> sp_of(np);
  for (int p = 0; p < n; ++p)
    for (int q = 0; q <= p; ++q) {
      const size_t a = shell_of[p], b = shell_of[q];
      sp_of[compound(p, q)] = a * (a + 1) / 2 + b;
    }
  auto start = [&](int s) { return B.shell_start[s]; };
  auto len = [&](int s) { return B.shell_start[s + 1] - B.shell_start[s]; };
  auto block = [&](int a, int b, int c, int d, std::vector<double>& r) {
    ShellPair ab = make_pair(B, a, b), cd = make_pair(B, c, d);
    quartet(B, ab, cd, false, r);
  };
  Mat diag(np, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (long x = 0; x < (long)nsp; ++x) {
    const int a = sp_list[x].first, b = sp_list[x].second;
    std::vector<double> r;
    block(a, b, a, b, r);
    const int na = len(a), nb = len(b);
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < nb; ++j) {
        const int p = start(a) + i, q = start(b) + j;
        if (p < q) continue;
        const int ij = i * nb + j;
        diag[compound(p, q)] = r[(size_t)ij * na * nb + ij];
      }
  }
  double scale = 1.0;
  for (double v : diag) scale = std::max(scale, v);
  for (double v : diag)
    if (v < -1e-12 * scale) return fail("negative ERI diagonal; the Coulomb metric is not positive semidefinite");
  std::vector<double> qsp(nsp, 0.0);
  for (size_t x = 0; x < np; ++x) qsp[sp_of[x]] = std::max(qsp[sp_of[x]], std::sqrt(std::max(diag[x], 0.0)));
  // columns of every AO pair of shell pair sp (those with r >= s), over all packed rows
  auto columns_of = [&](size_t sp, std::vector<size_t>& idx, std::vector<Mat>& cols) {
    const int r = sp_list[sp].first, s2 = sp_list[sp].second, nr = len(r), nq = len(s2);
    std::vector<std::array<int, 2>> tg;
    idx.clear();
    for (int k = 0; k < nr; ++k)
      for (int l = 0; l < nq; ++l)
        if (start(r) + k >= start(s2) + l) {
          tg.push_back({k, l});
          idx.push_back(compound(start(r) + k, start(s2) + l));
        }
    cols.assign(tg.size(), Mat(np, 0.0));
#pragma omp parallel for schedule(dynamic)
    for (long x = 0; x < (long)nsp; ++x) {
      if (!(qsp[x] * qsp[sp] > 1e-15)) continue;
      const int a = sp_list[x].first, b = sp_list[x].second, na = len(a), nb = len(b);
      std::vector<double> blk;
      block(a, b, r, s2, blk);
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < nb; ++j) {
          const int p = start(a) + i, q = start(b) + j;
          if (p < q) continue;
          const size_t row = compound(p, q);
          for (size_t t = 0; t < tg.size(); ++t) cols[t][row] = blk[((size_t)(i * nb + j) * nr + tg[t][0]) * nq + tg[t][1]];
        }
    }
  };
  std::vector<Mat> factors;
  pivots.clear();
  size_t cached_sp = (size_t)-1;
  std::vector<size_t> cidx;
  std::vector<Mat> cols;
  for (;;) {
    double largest = -INFINITY;
    for (double v : diag) largest = std::max(largest, v);
    size_t piv = 0;
    while (!(diag[piv] >= largest - 1e-12 * std::fabs(largest))) ++piv;
    const double maximum = diag[piv];
    if (maximum <= threshold || factors.size() == np) {
      max_residual = std::max(maximum, 0.0);
      break;
    }
    if (8.0 * (double)np * (double)(factors.size() + 2) > budget) return fail("factor storage exceeds the memory budget");
    const size_t sp = sp_of[piv];
    if (sp != cached_sp) {
      columns_of(sp, cidx, cols);
      cached_sp = sp;
    }
    const Mat& col = cols[std::find(cidx.begin(), cidx.end(), piv) - cidx.begin()];
    const double root = std::sqrt(maximum);
    Mat row(np);
#pragma omp parallel for schedule(static)
    for (long k = 0; k < (long)np; ++k) {
      double acc = 0.0;
      for (const Mat& L : factors) acc += L[k] * L[piv];
      row[k] = (col[k] - acc) / root;
    }
    row[piv] = root;
    for (size_t k = 0; k < np; ++k) {
      diag[k] -= row[k] * row[k];
      if (diag[k] < -1e-10 * scale) return fail("nonpositive residual Coulomb diagonal");
      diag[k] = std::max(diag[k], 0.0);
    }
    factors.push_back(std::move(row));
    pivots.push_back((int)piv);
  }
  rank = (int)factors.size();
  if (rank == 0) return fail("threshold exceeds every ERI diagonal; no Cholesky vectors were generated");
  bt.assign((size_t)rank * np, 0.0);
  for (int k = 0; k < rank; ++k) std::copy(factors[k].begin(), factors[k].end(), bt.begin() + (size_t)k * np);
  return 0;
}

int hfo_cholesky_eri(const hfo_basis* b, double threshold, double budget, int* rank, double* max_residual,
                     int* pivots, int max_pivots, double** b_out) {
  try {
    Mat bt;
    std::vector<int> piv;
    int r = 0;
    double mr = 0;
    if (int st = cholesky_eri(*b, threshold, budget, r, bt, mr, piv)) return st;
    *rank = r;
    *max_residual = mr;
    for (int k = 0; k < r && k < max_pivots; ++k) pivots[k] = piv[k];
    if (b_out) {
      *b_out = (double*)std::malloc(bt.size() * sizeof(double));
      std::memcpy(*b_out, bt.data(), bt.size() * sizeof(double));
    }
    return 0;
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

// Factorized RHF over packed factors bt (rank x np): src/hf_variants/ri_df_hf.rs:160-263.
static int factorized_rhf(const hfo_basis* b, const Mat& bt, int rank, const std::vector<std::array<double, 4>>& atoms,
                          double el, int max_iter, double e_tol, double r_tol, double d_tol, hfo_dfhf_result* res,
                          double* p_out, double* f_out, double* j_out, double* k_out);

// DF-RHF: src/hf_variants/ri_df_hf.rs:122-263.
int hfo_df_rhf(const hfo_basis* b, const hfo_basis* aux, const double* zxyz, int natom, int charge, int max_iter,
               double e_tol, double r_tol, double d_tol, double cutoff, hfo_dfhf_result* res, double* p_out,
               double* f_out, double* j_out, double* k_out) {
  try {
    std::memset(res, 0, sizeof(*res));
    auto atoms = atoms_of(zxyz, natom);
    if (natom == 0 || b->f.empty() || max_iter == 0 || !(e_tol > 0) || !(r_tol > 0) || !(d_tol > 0) ||
        !(cutoff > 0))
      return fail("DF-HF requires finite geometry and positive convergence controls");
    double el = -charge;
    for (auto& a : atoms) el += a[0];
    if (el <= 0 || std::fabs(el - std::round(el)) > 1e-8 || ((long)std::round(el)) % 2)
      return fail("DF-RHF requires a positive even electron count");
    int rank;
    Mat bt;
    if (int st = df_build(*b, *aux, cutoff, rank, bt)) return st;
    return factorized_rhf(b, bt, rank, atoms, el, max_iter, e_tol, r_tol, d_tol, res, p_out, f_out, j_out, k_out);
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

// CD-HF: run_cd_hf (src/hf_variants/cholesky_df.rs:43-95): the Cholesky
// factors to threshold delta (budget 512 MiB) in the factorized RHF driver,
// residual tolerance max(sqrt(tol) 1e-2, 1e-11), density tolerance
// sqrt(tol) 0.1 clamped to [1e-12, 1e-10].
int hfo_cd_rhf(const hfo_basis* b, const double* zxyz, int natom, int charge, double delta, int max_iter, double tol,
               hfo_dfhf_result* res, double* p_out) {
  try {
    std::memset(res, 0, sizeof(*res));
    auto atoms = atoms_of(zxyz, natom);
    double el = -charge;
    for (auto& a : atoms) el += a[0];
    if (natom == 0 || b->f.empty() || max_iter == 0 || !(tol > 0))
      return fail("CD-HF requires finite geometry and positive convergence controls");
    if (el <= 0 || std::fabs(el - std::round(el)) > 1e-8 || ((long)std::round(el)) % 2)
      return fail("DF-RHF requires a positive even electron count");
    int rank;
    Mat bt;
    double mr;
    std::vector<int> piv;
    if (int st = cholesky_eri(*b, delta, 512.0 * 1024 * 1024, rank, bt, mr, piv)) return st;
    const double r_tol = std::max(std::sqrt(tol) * 1e-2, 1e-11);
    const double d_tol = std::min(std::max(std::sqrt(tol) * 0.1, 1e-12), 1e-10);
    return factorized_rhf(b, bt, rank, atoms, el, max_iter, tol, r_tol, d_tol, res, p_out, nullptr, nullptr, nullptr);
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

static int factorized_rhf(const hfo_basis* b, const Mat& bt, int rank, const std::vector<std::array<double, 4>>& atoms,
                          double el, int max_iter, double e_tol, double r_tol, double d_tol, hfo_dfhf_result* res,
                          double* p_out, double* f_out, double* j_out, double* k_out) {
  try {
    const int natom = (int)atoms.size();
    int n = (int)b->f.size(), no = (int)std::lround(el) / 2;
    Mat S, T, V;
    one_electron(*b, atoms, S, T, V);
    Mat H(S.size());
    for (size_t x = 0; x < H.size(); ++x) H[x] = T[x] + V[x];
    Mat X;
    int m;
    if (canonical_orthogonalizer(n, S, 1e-8, X, m)) return 2;
    if (no > m) return fail("factorized RHF occupied count exceeds overlap rank");
    double enuc = 0;
    for (int i = 0; i < natom; ++i)
      for (int j = i + 1; j < natom; ++j) {
        double dx = atoms[i][1] - atoms[j][1], dy = atoms[i][2] - atoms[j][2], dz = atoms[i][3] - atoms[j][3];
        enuc += atoms[i][0] * atoms[j][0] / std::sqrt(dx * dx + dy * dy + dz * dz);
      }
    auto occ = [&](const Mat& C) {
      Mat co((size_t)n * no);
      for (int p = 0; p < n; ++p)
        for (int i = 0; i < no; ++i) co[(size_t)p * no + i] = C[(size_t)p * m + i];
      return co;
    };
    auto fockdf = [&](const Mat& co, Mat& F, Mat* Jo, Mat* Ko) {
      Mat J, K;
      if (df_jk(bt.data(), rank, n, co.data(), no, J, K)) throw std::runtime_error(g_err);
      F.resize(J.size());
      for (size_t x = 0; x < F.size(); ++x) F[x] = H[x] + J[x] - 0.5 * K[x];
      if (Jo) *Jo = J;
      if (Ko) *Ko = K;
    };
    auto energy = [&](const Mat& P, const Mat& F) {
      double s = 0;
      for (size_t x = 0; x < P.size(); ++x) s += P[x] * (H[x] + F[x]);
      return 0.5 * s;
    };
    std::vector<double> eps;
    Mat C;
    diag_in(H, X, n, m, eps, C);
    Mat co = occ(C), P = mm(co, tr(co, n, no), n, no, n);
    for (auto& x : P) x *= 2.0;
    Cdiis diis;
    double last = INFINITY;
    for (int it = 1; it <= max_iter; ++it) {
      Mat F;
      fockdf(co, F, nullptr, nullptr);
      double ce = energy(P, F) + enuc;
      Mat err = mm(mm(tr(X, n, m), diis_error(F, P, S, n), m, n, n), X, m, n, m);
      double resid = max_abs(err);
      Mat nc;
      diag_in(F, X, n, m, eps, nc);
      Mat nco = occ(nc), np2 = mm(nco, tr(nco, n, no), n, no, n);
      for (auto& x : np2) x *= 2.0;
      Mat dd = np2;
      for (size_t x = 0; x < dd.size(); ++x) dd[x] -= P[x];
      double dch = max_abs(dd);
      if (!std::isfinite(ce) || !std::isfinite(resid)) return fail("DF-HF energy/residual became nonfinite");
      if (std::fabs(ce - last) <= e_tol && resid <= r_tol && dch <= d_tol) {
        Mat Ff, Jf, Kf;
        fockdf(nco, Ff, &Jf, &Kf);
        res->total_energy = energy(np2, Ff) + enuc;
        res->iterations = it;
        res->converged = 1;
        res->n_aux = rank;
        res->residual_norm = max_abs(mm(mm(tr(X, n, m), diis_error(Ff, np2, S, n), m, n, n), X, m, n, m));
        if (p_out) std::memcpy(p_out, np2.data(), np2.size() * 8);
        if (f_out) std::memcpy(f_out, Ff.data(), Ff.size() * 8);
        if (j_out) std::memcpy(j_out, Jf.data(), Jf.size() * 8);
        if (k_out) std::memcpy(k_out, Kf.data(), Kf.size() * 8);
        return 0;
      }
      last = ce;
      diis.push(F, err);
      Mat ex;
      bool ok = it >= 2 && diis.extrapolate(ex);
      diag_in(ok ? ex : F, X, n, m, eps, C);
      co = occ(C);
      P = mm(co, tr(co, n, no), n, no, n);
      for (auto& x : P) x *= 2.0;
    }
    return fail("factorized RHF did not converge", 4);
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

// ---------------------------------------------------------------------------
// Analytic RHF gradient for a density P and energy-weighted density W:
// compute_gradient_from_density (src/gradient.rs:287-512).  Every term is
// evaluated without the reference's 1e-11 / 1e-15 screening of small
// products (exact up to rounding); the sums are the reference's:
//   nuclear repulsion (:220-243);
//   2 sum_nu [P (dT + dV) - W dS] over the derivative of each basis function
//   (:411-447);
//   -2 P_mu,nu (<dmu|V_C|nu> + <mu|V_C|dnu>) for the operator derivative at
//   each nucleus C (:451-490);
//   sum_{mu, nu, lambda >= sigma} gamma (dmu nu|lambda sigma) with gamma of
//   compute_gradient_from_density (:293-303) (:493-509).
// ---------------------------------------------------------------------------
void differentiate_cg(const CG& cg, int tau, std::vector<CG>& terms) {  // src/gradient.rs:71-113
  terms.clear();
  CG p = cg;
  p.l[tau] += 1;
  for (auto& pr : p.prims) pr.c *= 2.0 * pr.e;
  terms.push_back(p);
  if (cg.l[tau] > 0) {
    CG m = cg;
    m.l[tau] -= 1;
    const double lf = cg.l[tau];
    for (auto& pr : m.prims) pr.c *= -lf;
    terms.push_back(m);
  }
}

double contracted_kinetic(const CG& a, const CG& b) {
  double t = 0;
  for (auto& pa : a.prims)
    for (auto& pb : b.prims) t += pa.c * pb.c * prim_kinetic(pa.e, a.l, a.center, pb.e, b.l, b.center);
  return t;
}

int rhf_gradient(const hfo_basis& B, const std::vector<std::array<double, 4>>& atoms, const double* P, const double* W,
                 std::vector<double>& grad) {
  const int n = (int)B.f.size(), na = (int)atoms.size();
  grad.assign((size_t)3 * na, 0.0);
  // nuclear repulsion
  for (int a = 0; a < na; ++a)
    for (int b = 0; b < na; ++b) {
      if (a == b) continue;
      double d[3], r2 = 0;
      for (int k = 0; k < 3; ++k) {
        d[k] = atoms[a][1 + k] - atoms[b][1 + k];
        r2 += d[k] * d[k];
      }
      const double f = atoms[a][0] * atoms[b][0] / (r2 * std::sqrt(r2));
      for (int k = 0; k < 3; ++k) grad[3 * a + k] -= f * d[k];
    }
  std::vector<std::array<std::vector<CG>, 3>> dfn(n);
  for (int mu = 0; mu < n; ++mu)
    for (int t = 0; t < 3; ++t) differentiate_cg(B.f[mu], t, dfn[mu][t]);
  // one-electron basis-function derivatives
  std::vector<double> g1((size_t)3 * n, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int task = 0; task < 3 * n; ++task) {
    const int mu = task / 3, t = task % 3;
    double ds = 0, dt = 0, dv = 0;
    for (int nu = 0; nu < n; ++nu) {
      const double p = P[(size_t)mu * n + nu], w = W[(size_t)mu * n + nu];
      for (const CG& term : dfn[mu][t]) {
        ds += w * contracted_overlap(term, B.f[nu]);
        dt += p * contracted_kinetic(term, B.f[nu]);
        dv += p * contracted_nuclear(term, B.f[nu], atoms);
      }
    }
    g1[task] = 2.0 * (dt + dv - ds);
  }
  for (int mu = 0; mu < n; ++mu)
    for (int t = 0; t < 3; ++t) grad[3 * B.f[mu].atom + t] += g1[3 * mu + t];
  // operator derivative of the nuclear attraction, per nucleus
  std::vector<double> g2((size_t)n * 3 * na, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int mu = 0; mu < n; ++mu) {
    double* out = &g2[(size_t)mu * 3 * na];
    for (int nu = 0; nu <= mu; ++nu) {
      const double p = P[(size_t)mu * n + nu];
      for (int c = 0; c < na; ++c) {
        const std::vector<std::array<double, 4>> one{atoms[c]};
        for (int t = 0; t < 3; ++t) {
          double v = 0;
          for (const CG& term : dfn[mu][t]) v += contracted_nuclear(term, B.f[nu], one);
          if (mu != nu)
            for (const CG& term : dfn[nu][t]) v += contracted_nuclear(term, B.f[mu], one);
          out[3 * c + t] -= 2.0 * p * v;
        }
      }
    }
  }
  for (int mu = 0; mu < n; ++mu)
    for (size_t x = 0; x < (size_t)3 * na; ++x) grad[x] += g2[(size_t)mu * 3 * na + x];
  // two-electron: sum over mu, nu, lambda >= sigma of gamma (dmu nu|lambda sigma)
  std::vector<std::pair<int, int>> ls;
  std::vector<std::vector<PPData>> lsd;
  for (int l = 0; l < n; ++l)
    for (int s2 = 0; s2 <= l; ++s2) {
      ls.push_back({l, s2});
      lsd.push_back(basis_pair_data(B.f[l], B.f[s2]));
    }
  auto Pm = [&](int i, int j) { return P[(size_t)i * n + j]; };
  std::vector<double> g3((size_t)3 * n, 0.0);
#pragma omp parallel for schedule(dynamic)
  for (int task = 0; task < 3 * n; ++task) {
    const int mu = task / 3, t = task % 3;
    std::vector<double> r;
    double acc = 0;
    for (int nu = 0; nu < n; ++nu)
      for (const CG& term : dfn[mu][t]) {
        const std::vector<PPData> ab = basis_pair_data(term, B.f[nu]);
        for (size_t x = 0; x < ls.size(); ++x) {
          const int l = ls[x].first, sg = ls[x].second;
          const double gam = l == sg ? 2.0 * Pm(mu, nu) * Pm(l, l) - Pm(mu, l) * Pm(nu, l)
                                     : 4.0 * Pm(mu, nu) * Pm(l, sg) - Pm(mu, l) * Pm(nu, sg) - Pm(mu, sg) * Pm(nu, l);
          if (gam == 0.0) continue;
          acc += gam * pair_eri(ab, lsd[x], r);
        }
      }
    g3[task] = acc;
  }
  for (int mu = 0; mu < n; ++mu)
    for (int t = 0; t < 3; ++t) grad[3 * B.f[mu].atom + t] += g3[3 * mu + t];
  return 0;
}

int hfo_rhf_gradient(const hfo_basis* b, const double* zxyz, int natom, const double* p, const double* w,
                     double* grad) {
  try {
    std::vector<double> g;
    if (int st = rhf_gradient(*b, atoms_of(zxyz, natom), p, w, g)) return st;
    std::memcpy(grad, g.data(), g.size() * sizeof(double));
    return 0;
  } catch (std::exception& e) {
    return fail(e.what());
  }
}

// McMurchie-Davidson flop model per primitive quartet (mirrors the loop
// structure of src/shell_integrals.rs:101-133 and src/integrals.rs:70-101,
// 259-346, 905-971): setup 20, Boys 8+16(m+1), R scaling 2(L+1), R recursion
// 3(C(L+4,4)-(L+1)), and per Cartesian component quartet
// 2*sum_axis(lab+1)(lcd+1) + 3*prod_axis(Lax+1) + 3.
double hfo_md_flops(int la, int lb, int lc, int ld) {
  int L = la + lb + lc + ld;
  auto binom = [](int n, int k) {
    double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
  };
  double f = 20 + 8 + 16.0 * (L + 1) + 2.0 * (L + 1) + 3.0 * (binom(L + 4, 4) - (L + 1));
  auto ca = cart_components(la), cb = cart_components(lb), cc = cart_components(lc), cdd = cart_components(ld);
  for (auto& a : ca)
    for (auto& b : cb)
      for (auto& c : cc)
        for (auto& d : cdd) {
          double s = 0, pr = 1;
          for (int k = 0; k < 3; ++k) {
            int lab = a[k] + b[k], lcd = c[k] + d[k];
            s += (lab + 1.0) * (lcd + 1.0);
            pr *= (lab + lcd + 1.0);
          }
          f += 2 * s + 3 * pr + 3;
        }
  return f;
}

int hfo_fock_flops_sample(const hfo_basis* b, int stride, int offset, double* flops, double* quartets,
                          double* prim_quartets) {
  PairSet ps;
  build_pairs(*b, ps);
  long np = (long)ps.pairs.size();
  double fl = 0, nq = 0, npq = 0;
  std::map<std::array<int, 4>, double> memo;
  for (int la = 0; la <= 4; ++la)
    for (int lb = 0; lb <= 4; ++lb)
      for (int lc = 0; lc <= 4; ++lc)
        for (int ld = 0; ld <= 4; ++ld)
          if (la + lb + lc + ld <= 8) memo[{la, lb, lc, ld}] = hfo_md_flops(la, lb, lc, ld);
#pragma omp parallel for reduction(+ : fl, nq, npq) schedule(dynamic)
  for (long ij = 0; ij < np; ++ij) {
    if (stride > 1 && ij % stride != offset) continue;
    const ShellPair& ab = ps.pairs[ij];
    for (long kl = 0; kl <= ij; ++kl) {
      if (!(ps.bound[ij] * ps.bound[kl] > 1e-12)) continue;
      const ShellPair& cd = ps.pairs[kl];
      double k4 = (double)ab.prims.size() * cd.prims.size();
      auto it = memo.find({ab.la, ab.lb, cd.la, cd.lb});
      double f = it != memo.end() ? it->second : hfo_md_flops(ab.la, ab.lb, cd.la, cd.lb);
      fl += k4 * f;
      nq += 1;
      npq += k4;
    }
  }
  *flops = fl;
  if (quartets) *quartets = nq;
  if (prim_quartets) *prim_quartets = npq;
  return 0;
}

int hfo_fock_flops(const hfo_basis* b, double* flops, double* quartets, double* prim_quartets) {
  return hfo_fock_flops_sample(b, 1, 0, flops, quartets, prim_quartets);
}


// DIIS pieces on their own, for the tests that pin them.
void hfo_project_simplex(double* v, int m) {
  std::vector<double> x(v, v + m);
  project_simplex(x);
  std::copy(x.begin(), x.end(), v);
}
void hfo_min_quad_simplex(const double* q, const double* b, int m, double* c) {
  std::vector<std::vector<double>> qq(m, std::vector<double>(m));
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) qq[i][j] = q[(size_t)i * m + j];
  std::vector<double> st(m, 0.0);
  st[m - 1] = 1.0;
  auto r = min_quad_simplex(qq, std::vector<double>(b, b + m), st);
  std::copy(r.begin(), r.end(), c);
}
// errors: m entries of len doubles each; c: m coefficients.
int hfo_cdiis_coefficients(const double* errors, int m, int len, double* c) {
  Cdiis d;
  d.maxs = (size_t)std::max(m, 1);
  for (int i = 0; i < m; ++i) d.push(Mat(len, 0.0), Mat(errors + (size_t)i * len, errors + (size_t)(i + 1) * len));
  std::vector<double> cc;
  if (!d.coefficients(cc)) return fail("CDIIS system is singular or has fewer than two entries");
  std::copy(cc.begin(), cc.end(), c);
  return 0;
}
// dens, focks: m entries of len doubles each (oldest first); q (m x m), b (m).
int hfo_adiis_model(const double* dens, const double* focks, int m, int len, double* q, double* b) {
  if (m < 1) return fail("empty ADIIS history");
  Adiis a;
  a.maxs = (size_t)m;
  for (int i = 0; i < m; ++i)
    a.push(Mat(focks + (size_t)i * len, focks + (size_t)(i + 1) * len),
           Mat(dens + (size_t)i * len, dens + (size_t)(i + 1) * len), 0.0);
  std::vector<std::vector<double>> qq;
  std::vector<double> bb;
  a.model(qq, bb);
  for (int i = 0; i < m; ++i) {
    b[i] = bb[i];
    for (int j = 0; j < m; ++j) q[(size_t)i * m + j] = qq[i][j];
  }
  return 0;
}
}  // extern "C"
