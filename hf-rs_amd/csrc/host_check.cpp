// host_check.cpp -- CPU build of the generated Head-Gordon-Pople recurrences
// and of the Boys evaluation, for the CPU test-suite only (compared there with
// the oracle).  Not part of libhfg.so and never used as a fallback.
#define HFG_HOST_CHECK 1
#include <algorithm>
#include <cstring>
#include <string>

#include "classes.h"
#include "eri_device.h"
#include "host_prep.h"
#include "orient.h"

using namespace hfg;

static thread_local std::string g_err;

extern "C" const char* hfgc_last_error() { return g_err.c_str(); }

extern "C" int hfgc_boys(int m, double t, double* out) {
  static std::vector<double> tab = boys_table();
  if (m < 0 || m > BOYS_MAXM) return 1;
  boys_eval_rt(m, t, tab.data(), out);
  return 0;
}

template <int M, bool C>
static void boys_class(double t, double* out) {
  static std::vector<double> tab = boys_table();
  using B = BoysSlice<C>;
  std::vector<double> slice(B::size);
  for (int r = 0; r < B::rows; ++r) boys_slice_row<M, C>(tab.data(), r, &slice[r * B::cols]);
  boys_eval<M, false, C>(t, slice.data(), out);
}

// Templated per-class evaluation (the kernels' path) for M = 0..10, from the
// full (compact = 0) or the compact (compact = 1) LDS slice.
extern "C" int hfgc_boys_class(int m, int compact, double t, double* out) {
  switch (m) {
#define B(M)                            \
  case M:                               \
    if (compact)                        \
      boys_class<M, true>(t, out);      \
    else                                \
      boys_class<M, false>(t, out);     \
    return 0;
    B(0) B(1) B(2) B(3) B(4) B(5) B(6) B(7) B(8) B(9) B(10)
#undef B
  }
  return 1;
}

static int run_class(int cls, const PrimPairs& bp, int nb, const PrimPairs& kp, int nk, const double* AB,
                     const double* CD, double* blk) {
  static std::vector<double> tab = boys_table();
  std::vector<double> slice(BoysSlice<false>::size);
  switch (cls) {
#define X(a, b_, c, d)                                                                              \
  case (a * (a + 1) / 2 + b_ > 5 ? 3 + a : a * (a + 1) / 2 + b_) *                                 \
               ((a * (a + 1) / 2 + b_ > 5 ? 3 + a : a * (a + 1) / 2 + b_) + 1) / 2 +              \
           (c * (c + 1) / 2 + d > 5 ? 3 + c : c * (c + 1) / 2 + d):                                \
    for (int r = 0; r < BoysSlice<false>::rows; ++r) boys_slice_row<a + b_ + c + d>(tab.data(), r, &slice[r * BoysSlice<false>::cols]); \
    eri_block<a, b_, c, d>(bp, 0, nb, kp, 0, nk, 1, AB, CD, slice.data(), blk);                   \
    return 0;
    HFG_FOR_EACH_CLASS_ALL(X)
#undef X
  }
  return 1;
}

static PrimPairs view_of(const PrimPairHost& h) {
  return PrimPairs{h.p.data(), h.Px.data(), h.Py.data(), h.Pz.data(), h.PAx.data(),
                   h.PAy.data(), h.PAz.data(), h.K.data(), h.H.data()};
}

// (ab|P) with the unit function in the fourth slot: out[a][b][P], scaled.
extern "C" int hfgc_three_center(const hfg_basis_view* orb, const hfg_basis_view* aux, int sa, int sb, int sp,
                                 double* out) {
  try {
    auto so = group_shells(*orb);
    auto sx = group_shells(*aux);
    const Shell* A = &so[sa];
    const Shell* B = &so[sb];
    const bool swab = B->L > A->L;
    if (swab) std::swap(A, B);
    const Shell& P = sx[sp];
    PrimPairHost ab, px;
    append_pair(*A, *B, ab);
    append_aux_pair(P, px);
    const int pcab = pair_class(A->L, B->L), pcp = pair_class(P.L, 0);
    if (pcab < 0 || pcp < 0) return HFG_UNSUPPORTED;
    double AB[3], Z[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) AB[k] = A->c[k] - B->c[k];
    const int na = (int)A->scale.size(), nb = (int)B->scale.size(), np = (int)P.scale.size();
    std::vector<double> blk(2048);
    const bool aux_bra = pcp > pcab;
    if (aux_bra)
      run_class(quartet_class(pcp, pcab), view_of(px), (int)px.size(), view_of(ab), (int)ab.size(), Z, AB, blk.data());
    else
      run_class(quartet_class(pcab, pcp), view_of(ab), (int)ab.size(), view_of(px), (int)px.size(), AB, Z, blk.data());
    for (int i = 0; i < na; ++i)
      for (int j = 0; j < nb; ++j)
        for (int k = 0; k < np; ++k) {
          const double v = aux_bra ? blk[(k * na + i) * nb + j] : blk[(i * nb + j) * np + k];
          const int ii = swab ? j : i, jj = swab ? i : j;
          const int nbo = swab ? na : nb;
          out[(ii * nbo + jj) * np + k] = v * A->scale[i] * B->scale[j] * P.scale[k];
        }
    return 0;
  } catch (HfgError& e) {
    g_err = e.msg;
    return e.code;
  }
}

// (P|Q) metric block out[P][Q], scaled.
extern "C" int hfgc_metric(const hfg_basis_view* aux, int sp, int sq, double* out) {
  try {
    auto sx = group_shells(*aux);
    const Shell* P = &sx[sp];
    const Shell* Q = &sx[sq];
    const bool sw = Q->L > P->L;
    if (sw) std::swap(P, Q);
    PrimPairHost pp, qq;
    append_aux_pair(*P, pp);
    append_aux_pair(*Q, qq);
    double Z[3] = {0, 0, 0};
    std::vector<double> blk(2048);
    if (run_class(quartet_class(pair_class(P->L, 0), pair_class(Q->L, 0)), view_of(pp), (int)pp.size(), view_of(qq),
                  (int)qq.size(), Z, Z, blk.data()))
      return HFG_UNSUPPORTED;
    const int np = (int)P->scale.size(), nq = (int)Q->scale.size();
    for (int i = 0; i < np; ++i)
      for (int j = 0; j < nq; ++j) {
        const double v = blk[i * nq + j] * P->scale[i] * Q->scale[j];
        if (sw)
          out[j * np + i] = v;
        else
          out[i * nq + j] = v;
      }
    return 0;
  } catch (HfgError& e) {
    g_err = e.msg;
    return e.code;
  }
}

extern "C" int hfgc_shell_quartet(const hfg_basis_view* b, int sa, int sb, int sc, int sd, double* out) {
  static std::vector<double> tab = boys_table();
  try {
    auto sh = group_shells(*b);
    Orientation o = orient(sh, sa, sb, sc, sd);
    if (o.cls < 0) {
      g_err = "angular momentum beyond the compiled classes";
      return HFG_UNSUPPORTED;
    }
    PrimPairHost bra, ket;
    append_pair(sh[o.k[0]], sh[o.k[1]], bra);
    append_pair(sh[o.k[2]], sh[o.k[3]], ket);
    auto view = [](const PrimPairHost& h) {
      return PrimPairs{h.p.data(), h.Px.data(), h.Py.data(), h.Pz.data(),
                       h.PAx.data(), h.PAy.data(), h.PAz.data(), h.K.data(), h.H.data()};
    };
    PrimPairs bp = view(bra), kp = view(ket);
    double AB[3], CD[3];
    for (int k = 0; k < 3; ++k) {
      AB[k] = sh[o.k[0]].c[k] - sh[o.k[1]].c[k];
      CD[k] = sh[o.k[2]].c[k] - sh[o.k[3]].c[k];
    }
    std::vector<double> blk(2048), slice(BoysSlice<false>::size);
    switch (o.cls) {
#define X(a, b_, c, d)                                                                              \
  case (a * (a + 1) / 2 + b_) * ((a * (a + 1) / 2 + b_) + 1) / 2 + (c * (c + 1) / 2 + d):          \
    for (int r = 0; r < BoysSlice<false>::rows; ++r) boys_slice_row<a + b_ + c + d>(tab.data(), r, &slice[r * BoysSlice<false>::cols]); \
    eri_block<a, b_, c, d>(bp, 0, (int)bra.size(), kp, 0, (int)ket.size(), 1, AB, CD, slice.data(), blk.data()); \
    break;
      HFG_FOR_EACH_CLASS_ALL(X)
#undef X
    }
    unorient(sh, o, blk.data(), out);
    return 0;
  } catch (HfgError& e) {
    g_err = e.msg;
    return e.code;
  }
}

#include "scf_host.h"

// The RHF drivers' ADIIS minimiser (diis_host.cpp) and simplex projection, for
// the CPU tests that check them against brute force.
extern "C" void hfgc_simplex_minimise(const double* q, const double* b, int m, double* c) {
  const auto r = hfg::simplex_minimise(std::vector<double>(q, q + (size_t)m * m), std::vector<double>(b, b + m));
  std::copy(r.begin(), r.end(), c);
}
extern "C" void hfgc_simplex_project(double* v, int m) {
  std::vector<double> x(v, v + m);
  hfg::simplex_project(x);
  std::copy(x.begin(), x.end(), v);
}
extern "C" int hfgc_solve_small(const double* a, const double* b, int n, double* x) {
  std::vector<double> xx;
  if (!hfg::solve_small(std::vector<double>(a, a + (size_t)n * n), std::vector<double>(b, b + n), n, xx)) return 0;
  std::copy(xx.begin(), xx.end(), x);
  return 1;
}

#include "bscf_layout.h"

// Offsets of a server slot's matrices (bscf_layout.h) and the buffer size,
// for the CPU test that checks the carving stays inside the buffer.
extern "C" size_t hfgc_slot_layout(size_t n, size_t m, size_t* off) { return hfg::slot_layout(n, m, off); }
