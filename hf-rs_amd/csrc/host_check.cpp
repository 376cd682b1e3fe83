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

extern "C" int hfgc_shell_quartet(const hfg_basis_view* b, int sa, int sb, int sc, int sd, double* out) {
  static std::vector<double> tab = boys_table();
  try {
    auto sh = group_shells(*b);
    Orientation o = orient(sh, sa, sb, sc, sd);
    if (o.cls < 0) {
      g_err = "angular momentum beyond d";
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
    std::vector<double> blk(1296), slice(BOYS_ROWS * BOYS_SLICE);
    switch (o.cls) {
#define X(a, b_, c, d)                                                                              \
  case (a * (a + 1) / 2 + b_) * ((a * (a + 1) / 2 + b_) + 1) / 2 + (c * (c + 1) / 2 + d):          \
    for (int r = 0; r < BOYS_ROWS; ++r) boys_slice_row<a + b_ + c + d>(tab.data(), r, &slice[r * BOYS_SLICE]); \
    eri_block<a, b_, c, d>(bp, 0, (int)bra.size(), kp, 0, (int)ket.size(), 1, AB, CD, slice.data(), blk.data()); \
    break;
      HFG_FOR_EACH_CLASS(X)
#undef X
    }
    unorient(sh, o, blk.data(), out);
    return 0;
  } catch (HfgError& e) {
    g_err = e.msg;
    return e.code;
  }
}
