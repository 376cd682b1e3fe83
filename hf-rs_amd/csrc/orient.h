// orient.h -- map a requested shell quartet (a b|c d) onto kernel orientation
// and back.  (ab|cd) = (ba|cd) = (ab|dc) = (cd|ab) for real Gaussians.
#pragma once
#include "host_prep.h"

namespace hfg {

struct Orientation {
  int k[4];    // shells in kernel order
  int pos[4];  // kernel slot j holds requested slot pos[j]
  int cls;     // quartet class index
};

inline Orientation orient(const std::vector<Shell>& sh, int sa, int sb, int sc, int sd) {
  int s[4] = {sa, sb, sc, sd}, p[4] = {0, 1, 2, 3};
  if (sh[s[1]].L > sh[s[0]].L) { std::swap(s[0], s[1]); std::swap(p[0], p[1]); }
  if (sh[s[3]].L > sh[s[2]].L) { std::swap(s[2], s[3]); std::swap(p[2], p[3]); }
  int pb = pair_class(sh[s[0]].L, sh[s[1]].L), pk = pair_class(sh[s[2]].L, sh[s[3]].L);
  if (pk > pb) {
    std::swap(s[0], s[2]); std::swap(s[1], s[3]); std::swap(p[0], p[2]); std::swap(p[1], p[3]);
    std::swap(pb, pk);
  }
  Orientation o;
  for (int j = 0; j < 4; ++j) { o.k[j] = s[j]; o.pos[j] = p[j]; }
  o.cls = (pb < 0 || pk < 0) ? -1 : quartet_class(pb, pk);
  return o;
}

// Scatter a kernel-orientation block (unscaled) into requested order,
// applying the per-component normalisation factors.
inline void unorient(const std::vector<Shell>& sh, const Orientation& o, const double* blk, double* out) {
  int n[4], nr[4];
  for (int j = 0; j < 4; ++j) n[j] = (int)sh[o.k[j]].scale.size();
  for (int j = 0; j < 4; ++j) nr[o.pos[j]] = n[j];
  int r[4];
  size_t idx = 0;
  for (int i0 = 0; i0 < n[0]; ++i0)
    for (int i1 = 0; i1 < n[1]; ++i1)
      for (int i2 = 0; i2 < n[2]; ++i2)
        for (int i3 = 0; i3 < n[3]; ++i3, ++idx) {
          int ii[4] = {i0, i1, i2, i3};
          double f = 1.0;
          for (int j = 0; j < 4; ++j) {
            r[o.pos[j]] = ii[j];
            f *= sh[o.k[j]].scale[ii[j]];
          }
          out[((size_t)(r[0] * nr[1] + r[1]) * nr[2] + r[2]) * nr[3] + r[3]] = f * blk[idx];
        }
}

}  // namespace hfg
