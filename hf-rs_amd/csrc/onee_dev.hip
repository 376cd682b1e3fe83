// onee_dev.hip -- the one-electron matrices S, T, V and the SAD density on
// the device: compute_one_electron_integrals (src/scf.rs:246-271, whose
// contracted_overlap / contracted_kinetic / contracted_nuclear_attraction are
// McMurchie-Davidson, src/integrals.rs:407-537, 705-793) and the
// superposition step of compute_sad_density (src/sad.rs:217-259).
//
// One 64-lane workgroup per unique shell pair (A, B), oriented so that
// L_A >= L_B, one launch per (L_A, L_B) class (L <= 3).  Lanes take the
// primitive pairs for S and T and the (primitive pair, nucleus) products for
// V; the per-lane sums are reduced through LDS in a fixed order, so a result
// does not depend on the schedule.  Obara-Saika recurrences:
//   1-D overlaps  S_{i+1,j} = PA S_ij + (i S_{i-1,j} + j S_{i,j-1})/(2p),
//                 S_{i,j+1} = PB S_ij + (i S_{i-1,j} + j S_{i,j-1})/(2p);
//   kinetic       t_ij = b(2j+1) S_ij - 2b^2 S_{i,j+2} - j(j-1)/2 S_{i,j-2};
//   nuclear       Th^m[e+1_i] = PA_i Th^m[e] - PC_i Th^{m+1}[e]
//                   + e_i/(2p) (Th^m[e-1_i] - Th^{m+1}[e-1_i]),
//                 Th^m[0] = -Z 2pi/p exp(-mu AB^2) F_m(p PC^2), to L_A + L_B
//                 on the bra, summed over primitives and nuclei, then
//                 (a|b) = sum_k prod_x C(b_x,k_x) AB_x^(b_x-k_x) (a+k|0).
// Each workgroup reads its two shells' primitives and the nuclei and writes
// its block and the transposed block: about 3 n^2 doubles of HBM traffic per
// call.  The work is FP64 VALU, K_A K_B N_atom recurrences per pair.
#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "engine.h"
#include "eri_device.h"
#include "host_prep.h"
#include "la_device.h"
#include "rhf.h"
#include "scf_host.h"

namespace hfg {
namespace {

constexpr int ONEE_LMAX = 3;  // f: cc-pVTZ in the DF drivers

constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }
// number of exponent triples (x, y, z) with x + y + z <= L
constexpr int ncum(int L) { return L < 0 ? 0 : (L + 1) * (L + 2) * (L + 3) / 6; }
// position of (x, y, z) among all triples, ordered by total, then x
// descending, then y descending
constexpr int hpos(int x, int y, int z) { return ncum(x + y + z - 1) + (y + z) * (y + z + 1) / 2 + z; }
// the k-th triple of total l in that order
constexpr int trip(int l, int k, int ax) {
  int i = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y, ++i)
      if (i == k) return ax == 0 ? x : ax == 1 ? y : l - x - y;
  return 0;
}
// the k-th Cartesian component of a shell of angular momentum L in hf-rs
// order: d = xx, yy, zz, xy, xz, yz (src/basis.rs:430-436), otherwise x
// descending then y descending
constexpr int cexp(int L, int k, int ax) {
  if (L == 2) {
    const int d[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    return d[k][ax];
  }
  return trip(L, k, ax);
}

template <class F, int... I>
HFG_DEV void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, i>) for i = 0..N-1: compile-time indices
template <int N, class F>
HFG_DEV void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// F_0..F_M(T) from the global Boys table (rows T_r = r/8): the Taylor
// expansion of F_M about the nearest row, |dT| <= 1/16, nine terms, then the
// downward recursion; the asymptotic form with the upward recursion for
// T >= 50 (as boys_eval, with the table instead of an LDS slice).
template <int M>
HFG_DEV void boys_rows(double T, const double* __restrict__ tab, double* __restrict__ F) {
  if (T < BOYS_TMAX) {
    const int r = (int)(T * BOYS_INV_STEP + 0.5);
    const double x = r * (1.0 / BOYS_INV_STEP) - T;
    const double* row = tab + r * BOYS_ROW;
    double acc = row[1 + M + BOYS_TAYLOR - 1] * BOYS_EXP_C[BOYS_TAYLOR - 1];
#pragma unroll
    for (int k = BOYS_TAYLOR - 2; k >= 0; --k) acc = fma(acc, x, row[1 + M + k] * BOYS_EXP_C[k]);
    F[M] = acc;
    if (M > 0) {
      double e = BOYS_EXP_C[12];
#pragma unroll
      for (int k = 11; k >= 0; --k) e = fma(e, x, BOYS_EXP_C[k]);
      const double emt = row[0] * e, t2 = 2.0 * T;
#pragma unroll
      for (int m = M - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + emt) * (1.0 / (2 * m + 1));
    }
  } else {
    const double rt = rsqrt_nr(T);
    F[0] = HALF_SQRT_PI * rt;
    if (M > 0) {
      const double emt = exp(-T), h = 0.5 * rt * rt;
#pragma unroll
      for (int m = 1; m <= M; ++m) F[m] = ((2 * m - 1) * F[m - 1] - emt) * h;
    }
  }
}

struct OneShell {
  double c[3];
  int L, K, p0, ao, atom;
};

struct OneArgs {
  const int2* pairs;  // (A, B) shell ids, L_A >= L_B
  int npair;
  const OneShell* sh;
  const double* pe;  // primitive exponents
  const double* pr;  // primitive coefficients (unscaled components)
  const double* fscale;  // per-function component normalisation
  const double4* atoms;  // (Z, x, y, z)
  int natom;
  const double* boys;
  int n;
  double *S, *T, *V, *H;  // any may be null; H = T + V
  const double *P, *W;     // gradient: density and energy-weighted density
  double* grad;            // gradient: natom x 3, accumulated
};

constexpr int RED_ROWS = 32;

// out[r] = sum over the 64 lanes of v[r], r < N, in lane order.
template <int N>
HFG_DEV void wave_sum(const double (&v)[N], double* __restrict__ red, double* __restrict__ out) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int c0 = 0; c0 < N; c0 += RED_ROWS) {
#pragma unroll
    for (int r = 0; r < RED_ROWS; ++r)
      if (c0 + r < N) red[lane * (RED_ROWS + 1) + r] = v[c0 + r];
    __syncthreads();
    if (lane < RED_ROWS && c0 + lane < N) {
      double s = 0.0;
      for (int i = 0; i < 64; ++i) s += red[i * (RED_ROWS + 1) + lane];
      out[c0 + lane] = s;
    }
    __syncthreads();
  }
}

HFG_DEV void comp_rt(int L, int k, int* e) {
  if (L == 2) {
    const int d[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    e[0] = d[k][0];
    e[1] = d[k][1];
    e[2] = d[k][2];
    return;
  }
  int i = 0;
  for (int x = L; x >= 0; --x)
    for (int y = L - x; y >= 0; --y, ++i)
      if (i == k) {
        e[0] = x;
        e[1] = y;
        e[2] = L - x - y;
      }
}

// (x, y, z) at position pos of the order of hpos
HFG_DEV void trip_rt(int pos, int* e) {
  int l = 0;
  while (ncum(l) <= pos) ++l;
  const int k = pos - ncum(l - 1);
  int i = 0;
  for (int x = l; x >= 0; --x)
    for (int y = l - x; y >= 0; --y, ++i)
      if (i == k) {
        e[0] = x;
        e[1] = y;
        e[2] = l - x - y;
      }
}

// Coefficient of [e|0] in (a|b) by the horizontal transfer
// (a|b) = sum_k prod_x C(b_x,k_x) AB_x^(b_x-k_x) [a+k|0].
HFG_DEV double hrr_coef(const int* a, const int* b, const int* e, const double* AB);

HFG_DEV double ipow(double x, int k) {
  double r = 1.0;
  for (int i = 0; i < k; ++i) r *= x;
  return r;
}
HFG_DEV double binom_small(int a, int b) {  // small a
  double r = 1.0;
  for (int i = 0; i < b; ++i) r = r * (a - i) / (i + 1);
  return r;
}

HFG_DEV double hrr_coef(const int* a, const int* b, const int* e, const double* AB) {
  double c = 1.0;
  for (int x = 0; x < 3; ++x) {
    const int k = e[x] - a[x];
    if (k < 0 || k > b[x]) return 0.0;
    c *= binom_small(b[x], k) * ipow(AB[x], b[x] - k);
  }
  return c;
}

template <int LA, int LB>
__global__ __launch_bounds__(64) void k_onee(OneArgs g) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NO = NA * NB, LT = LA + LB;
  constexpr int E0 = ncum(LA - 1), NE = ncum(LT) - E0;
  __shared__ double red[64 * (RED_ROWS + 1)];
  __shared__ double st[2 * NO];
  __shared__ double ve[NE];
  const int lane = threadIdx.x;
  const int2 pq = g.pairs[blockIdx.x];
  const OneShell A = g.sh[pq.x], B = g.sh[pq.y];
  double AB[3];
  double ab2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    AB[k] = A.c[k] - B.c[k];
    ab2 += AB[k] * AB[k];
  }
  const int KK = A.K * B.K;
  const double PI = 3.14159265358979323846;

  // S and T: lanes over the primitive pairs
  {
    double acc[2 * NO];
#pragma unroll
    for (int k = 0; k < 2 * NO; ++k) acc[k] = 0.0;
    for (int kp = lane; kp < KK; kp += 64) {
      const int ka = kp / B.K, kb = kp - ka * B.K;
      const double a = g.pe[A.p0 + ka], b = g.pe[B.p0 + kb], p = a + b, o2p = 0.5 / p;
      const double mu = a * b / p;
      const double s00 = (PI / p) * sqrt(PI / p) * exp(-mu * ab2) * g.pr[A.p0 + ka] * g.pr[B.p0 + kb];
      double s1[3][LA + 1][LB + 3];
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const double P = (a * A.c[ax] + b * B.c[ax]) / p;
        const double PA = P - A.c[ax], PB = P - B.c[ax];
        s1[ax][0][0] = 1.0;
#pragma unroll
        for (int i = 0; i <= LA; ++i)
#pragma unroll
          for (int j = 0; j <= LB + 2; ++j) {
            if (i == 0 && j == 0) continue;
            double v;
            if (i > 0) {
              v = PA * s1[ax][i - 1][j];
              if (i > 1) v += (i - 1) * o2p * s1[ax][i - 2][j];
              if (j > 0) v += j * o2p * s1[ax][i - 1][j - 1];
            } else {
              v = PB * s1[ax][i][j - 1];
              if (j > 1) v += (j - 1) * o2p * s1[ax][i][j - 2];
            }
            s1[ax][i][j] = v;
          }
      }
      sfor<NA>([&](auto X) {
        constexpr int x = decltype(X)::value;
        constexpr int a0 = cexp(LA, x, 0), a1 = cexp(LA, x, 1), a2 = cexp(LA, x, 2);
        sfor<NB>([&](auto Y) {
          constexpr int y = decltype(Y)::value;
          constexpr int b0 = cexp(LB, y, 0), b1 = cexp(LB, y, 1), b2 = cexp(LB, y, 2);
          const double sx = s1[0][a0][b0], sy = s1[1][a1][b1], sz = s1[2][a2][b2];
          auto kin = [&](int ax, auto I, auto J) {
            constexpr int i = decltype(I)::value, j = decltype(J)::value;
            double t = b * (2 * j + 1) * s1[ax][i][j] - 2.0 * b * b * s1[ax][i][j + 2];
            if constexpr (j >= 2) t -= 0.5 * j * (j - 1) * s1[ax][i][j - 2];
            return t;
          };
          const double tx = kin(0, std::integral_constant<int, a0>{}, std::integral_constant<int, b0>{});
          const double ty = kin(1, std::integral_constant<int, a1>{}, std::integral_constant<int, b1>{});
          const double tz = kin(2, std::integral_constant<int, a2>{}, std::integral_constant<int, b2>{});
          acc[x * NB + y] += s00 * sx * sy * sz;
          acc[NO + x * NB + y] += s00 * (tx * sy * sz + sx * ty * sz + sx * sy * tz);
        });
      });
    }
    wave_sum<2 * NO>(acc, red, st);
  }

  // V: lanes over (primitive pair, nucleus)
  {
    double acc[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) acc[k] = 0.0;
    const int ntask = KK * g.natom;
    for (int t = lane; t < ntask; t += 64) {
      const int kp = t / g.natom, c = t - kp * g.natom;
      const int ka = kp / B.K, kb = kp - ka * B.K;
      const double a = g.pe[A.p0 + ka], b = g.pe[B.p0 + kb], p = a + b, o2p = 0.5 / p;
      const double mu = a * b / p;
      const double4 C = g.atoms[c];
      double PA[3], PC[3];
      double pc2 = 0.0;
#pragma unroll
      for (int ax = 0; ax < 3; ++ax) {
        const double P = (a * A.c[ax] + b * B.c[ax]) / p;
        PA[ax] = P - A.c[ax];
        PC[ax] = P - (ax == 0 ? C.y : ax == 1 ? C.z : C.w);
        pc2 += PC[ax] * PC[ax];
      }
      double F[LT + 1];
      boys_rows<LT>(p * pc2, g.boys, F);
      const double base = -C.x * 2.0 * PI / p * exp(-mu * ab2) * g.pr[A.p0 + ka] * g.pr[B.p0 + kb];
      double th[LT + 1][ncum(LT)];
#pragma unroll
      for (int m = 0; m <= LT; ++m) th[m][0] = base * F[m];
      sfor<LT>([&](auto Lm) {
        constexpr int l = decltype(Lm)::value + 1;
        sfor<ncart(l)>([&](auto Kc) {
          constexpr int k = decltype(Kc)::value;
          constexpr int x = trip(l, k, 0), y = trip(l, k, 1), z = trip(l, k, 2);
          constexpr int i = x > 0 ? 0 : (y > 0 ? 1 : 2);
          constexpr int x1 = x - (i == 0), y1 = y - (i == 1), z1 = z - (i == 2);
          constexpr int e1 = i == 0 ? x1 : (i == 1 ? y1 : z1);
          constexpr int p1 = hpos(x1, y1, z1);
          constexpr int p2 = e1 > 0 ? hpos(x1 - (i == 0), y1 - (i == 1), z1 - (i == 2)) : 0;
          constexpr int pt = hpos(x, y, z);
#pragma unroll
          for (int m = 0; m <= LT - l; ++m) {
            double v = PA[i] * th[m][p1] - PC[i] * th[m + 1][p1];
            if constexpr (e1 > 0) v += (e1 * o2p) * (th[m][p2] - th[m + 1][p2]);
            th[m][pt] = v;
          }
        });
      });
#pragma unroll
      for (int k = 0; k < NE; ++k) acc[k] += th[0][E0 + k];
    }
    wave_sum<NE>(acc, red, ve);
  }

  // horizontal transfer to the ket, scale and store
  const bool diag = pq.x == pq.y;
  for (int k = lane; k < NO; k += 64) {
    const int x = k / NB, y = k - x * NB;
    if (diag && x < y) continue;
    int ea[3], eb[3];
    comp_rt(LA, x, ea);
    comp_rt(LB, y, eb);
    double v = 0.0;
    for (int kx = 0; kx <= eb[0]; ++kx)
      for (int ky = 0; ky <= eb[1]; ++ky)
        for (int kz = 0; kz <= eb[2]; ++kz) {
          const double cf = binom_small(eb[0], kx) * binom_small(eb[1], ky) * binom_small(eb[2], kz) *
                            ipow(AB[0], eb[0] - kx) * ipow(AB[1], eb[1] - ky) * ipow(AB[2], eb[2] - kz);
          const int xx = ea[0] + kx, yy = ea[1] + ky, zz = ea[2] + kz, l = xx + yy + zz;
          const int pos = (l * (l + 1) * (l + 2)) / 6 + ((yy + zz) * (yy + zz + 1)) / 2 + zz - E0;
          v += cf * ve[pos];
        }
    const int i = A.ao + x, j = B.ao + y;
    const double f = g.fscale[i] * g.fscale[j];
    const double s = f * st[k], t = f * st[NO + k];
    v *= f;
    const size_t ij = (size_t)i * g.n + j, ji = (size_t)j * g.n + i;
    if (g.S) g.S[ij] = g.S[ji] = s;
    if (g.T) g.T[ij] = g.T[ji] = t;
    if (g.V) g.V[ij] = g.V[ji] = v;
    if (g.H) g.H[ij] = g.H[ji] = t + v;
  }
}

// One-electron terms of the nuclear gradient for the ordered shell pair
// (A, B): derivatives of the A functions, d/dA_t phi_a = 2 alpha phi_{a+1_t}
// - a_t phi_{a-1_t} (src/gradient.rs:411-490).  With w_ab = 2 f_ab P_ab and
// u_ab = -2 f_ab W_ab (f: component norms),
//   F_A,t += sum_ab w_ab (dT + dV)_ab,t + u_ab dS_ab,t,
//   F_C,t -= sum_ab w_ab dV^(C)_ab,t  for every nucleus C (the operator term),
// dV = sum_C dV^(C).  dV^(C) is linear in the bra-only [e|0] of each
// (primitive pair, nucleus), so a task's contribution is th . (2 alpha R_t - L_t),
// with R_t[e] = sum_ab w_ab h(a+1_t, b; e), L_t[e] = sum_ab w_ab a_t h(a-1_t, b; e)
// and h the transfer coefficients, formed once per pair in LDS.  Forces of the
// nuclei are summed in LDS (3 natom doubles, dynamic) and added to the
// output once per workgroup.
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_grad_onee(OneArgs g) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NO = NA * NB, LT = LA + LB + 1;
  constexpr int E0 = ncum(LA - 2), NEG = ncum(LT) - E0;
  extern __shared__ double force[];  // natom x 3
  __shared__ double red[64 * (RED_ROWS + 1)];
  __shared__ double wv[NO], uv[NO];
  __shared__ double rl[6 * NEG];
  __shared__ double fa[3];
  const int lane = threadIdx.x;
  const int2 pq = g.pairs[blockIdx.x];
  const OneShell A = g.sh[pq.x], B = g.sh[pq.y];
  double AB[3];
  double ab2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    AB[k] = A.c[k] - B.c[k];
    ab2 += AB[k] * AB[k];
  }
  for (int i = lane; i < 3 * g.natom; i += 64) force[i] = 0.0;
  for (int k = lane; k < NO; k += 64) {
    const int x = k / NB, y = k - x * NB;
    const int i = A.ao + x, j = B.ao + y;
    const double f = 2.0 * g.fscale[i] * g.fscale[j];
    wv[k] = f * g.P[(size_t)i * g.n + j];
    uv[k] = -f * g.W[(size_t)i * g.n + j];
  }
  __syncthreads();
  for (int q = lane; q < 6 * NEG; q += 64) {
    const int kt = q / NEG, ei = q - kt * NEG;
    const int t = kt % 3;
    const bool up = kt < 3;
    int e[3];
    trip_rt(E0 + ei, e);
    double acc = 0.0;
    for (int k = 0; k < NO; ++k) {
      const int x = k / NB, y = k - x * NB;
      int a[3], b[3];
      comp_rt(LA, x, a);
      comp_rt(LB, y, b);
      double f = wv[k];
      if (up) {
        ++a[t];
      } else {
        if (a[t] == 0) continue;
        f *= a[t];
        --a[t];
      }
      acc += f * hrr_coef(a, b, e, AB);
    }
    rl[q] = acc;
  }
  __syncthreads();
  const int KK = A.K * B.K;
  const double PI = 3.14159265358979323846;
  double fA[3] = {0.0, 0.0, 0.0};

  // overlap and kinetic derivatives: lanes over the primitive pairs
  for (int kp = lane; kp < KK; kp += 64) {
    const int ka = kp / B.K, kb = kp - ka * B.K;
    const double a = g.pe[A.p0 + ka], b = g.pe[B.p0 + kb], p = a + b, o2p = 0.5 / p;
    const double mu = a * b / p;
    const double s00 = (PI / p) * sqrt(PI / p) * exp(-mu * ab2) * g.pr[A.p0 + ka] * g.pr[B.p0 + kb];
    double s1[3][LA + 2][LB + 3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const double P = (a * A.c[ax] + b * B.c[ax]) / p;
      const double PA = P - A.c[ax], PB = P - B.c[ax];
      s1[ax][0][0] = 1.0;
#pragma unroll
      for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
        for (int j = 0; j <= LB + 2; ++j) {
          if (i == 0 && j == 0) continue;
          double v;
          if (i > 0) {
            v = PA * s1[ax][i - 1][j];
            if (i > 1) v += (i - 1) * o2p * s1[ax][i - 2][j];
            if (j > 0) v += j * o2p * s1[ax][i - 1][j - 1];
          } else {
            v = PB * s1[ax][i][j - 1];
            if (j > 1) v += (j - 1) * o2p * s1[ax][i][j - 2];
          }
          s1[ax][i][j] = v;
        }
    }
    // 1-D kinetic factor of (i | j) along ax
    auto kin = [&](int ax, int i, int j) {
      double t = b * (2 * j + 1) * s1[ax][i][j] - 2.0 * b * b * s1[ax][i][j + 2];
      if (j >= 2) t -= 0.5 * j * (j - 1) * s1[ax][i][j - 2];
      return t;
    };
    sfor<NA>([&](auto X) {
      constexpr int x = decltype(X)::value;
      constexpr int ea[3] = {cexp(LA, x, 0), cexp(LA, x, 1), cexp(LA, x, 2)};
      sfor<NB>([&](auto Y) {
        constexpr int y = decltype(Y)::value;
        constexpr int eb[3] = {cexp(LB, y, 0), cexp(LB, y, 1), cexp(LB, y, 2)};
        const double w = wv[x * NB + y], u = uv[x * NB + y];
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          // (a +- 1_t | b): overlap and kinetic products with axis t shifted
          double sx[3], tx[3], sl[3], tl[3];
#pragma unroll
          for (int ax = 0; ax < 3; ++ax) {
            const int i = ea[ax] + (ax == t ? 1 : 0);
            sx[ax] = s1[ax][i][eb[ax]];
            tx[ax] = kin(ax, i, eb[ax]);
            const int il = ea[ax] - (ax == t ? 1 : 0);
            sl[ax] = il >= 0 ? s1[ax][il][eb[ax]] : 0.0;
            tl[ax] = il >= 0 ? kin(ax, il, eb[ax]) : 0.0;
          }
          const double sr = sx[0] * sx[1] * sx[2];
          const double tr = tx[0] * sx[1] * sx[2] + sx[0] * tx[1] * sx[2] + sx[0] * sx[1] * tx[2];
          double ds = 2.0 * a * sr, dt = 2.0 * a * tr;
          if (ea[t] > 0) {
            ds -= ea[t] * (sl[0] * sl[1] * sl[2]);
            dt -= ea[t] * (tl[0] * sl[1] * sl[2] + sl[0] * tl[1] * sl[2] + sl[0] * sl[1] * tl[2]);
          }
          fA[t] += s00 * (w * dt + u * ds);
        }
      });
    });
  }

  // nuclear attraction: lanes over (primitive pair, nucleus)
  const int ntask = KK * g.natom;
  for (int q = lane; q < ntask; q += 64) {
    const int kp = q / g.natom, c = q - kp * g.natom;
    const int ka = kp / B.K, kb = kp - ka * B.K;
    const double a = g.pe[A.p0 + ka], b = g.pe[B.p0 + kb], p = a + b, o2p = 0.5 / p;
    const double mu = a * b / p;
    const double4 C = g.atoms[c];
    double PA[3], PC[3];
    double pc2 = 0.0;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
      const double P = (a * A.c[ax] + b * B.c[ax]) / p;
      PA[ax] = P - A.c[ax];
      PC[ax] = P - (ax == 0 ? C.y : ax == 1 ? C.z : C.w);
      pc2 += PC[ax] * PC[ax];
    }
    double F[LT + 1];
    boys_rows<LT>(p * pc2, g.boys, F);
    const double base = -C.x * 2.0 * PI / p * exp(-mu * ab2) * g.pr[A.p0 + ka] * g.pr[B.p0 + kb];
    double th[LT + 1][ncum(LT)];
#pragma unroll
    for (int m = 0; m <= LT; ++m) th[m][0] = base * F[m];
    sfor<LT>([&](auto Lm) {
      constexpr int l = decltype(Lm)::value + 1;
      sfor<ncart(l)>([&](auto Kc) {
        constexpr int k = decltype(Kc)::value;
        constexpr int x = trip(l, k, 0), y = trip(l, k, 1), z = trip(l, k, 2);
        constexpr int i = x > 0 ? 0 : (y > 0 ? 1 : 2);
        constexpr int x1 = x - (i == 0), y1 = y - (i == 1), z1 = z - (i == 2);
        constexpr int e1 = i == 0 ? x1 : (i == 1 ? y1 : z1);
        constexpr int p1 = hpos(x1, y1, z1);
        constexpr int p2 = e1 > 0 ? hpos(x1 - (i == 0), y1 - (i == 1), z1 - (i == 2)) : 0;
        constexpr int pt = hpos(x, y, z);
#pragma unroll
        for (int m = 0; m <= LT - l; ++m) {
          double v = PA[i] * th[m][p1] - PC[i] * th[m + 1][p1];
          if constexpr (e1 > 0) v += (e1 * o2p) * (th[m][p2] - th[m + 1][p2]);
          th[m][pt] = v;
        }
      });
    });
    double gt[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < NEG; ++k) {
      const double v = th[0][E0 + k];
#pragma unroll
      for (int t = 0; t < 3; ++t) gt[t] += v * (2.0 * a * rl[t * NEG + k] - rl[(3 + t) * NEG + k]);
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      fA[t] += gt[t];
      atomicAdd(&force[3 * c + t], -gt[t]);
    }
  }
  wave_sum<3>(fA, red, fa);
  if (lane < 3) atomicAdd(&force[3 * A.atom + lane], fa[lane]);
  __syncthreads();
  for (int i = lane; i < 3 * g.natom; i += 64)
    if (force[i] != 0.0) atomicAdd(&g.grad[i], force[i]);
}

template <int LA, int LB>
void launch_grad(const OneArgs& g, hipStream_t st) {
  const size_t lds = (size_t)3 * g.natom * sizeof(double);
  hipLaunchKernelGGL((k_grad_onee<LA, LB>), dim3((unsigned)g.npair), dim3(64), lds, st, g);
  HIP_CHECK(hipGetLastError());
}

void launch_grad_class(int la, int lb, const OneArgs& g, hipStream_t st) {
  switch (la * 4 + lb) {
    case 0: return launch_grad<0, 0>(g, st);
    case 1: return launch_grad<0, 1>(g, st);
    case 2: return launch_grad<0, 2>(g, st);
    case 4: return launch_grad<1, 0>(g, st);
    case 5: return launch_grad<1, 1>(g, st);
    case 6: return launch_grad<1, 2>(g, st);
    case 8: return launch_grad<2, 0>(g, st);
    case 9: return launch_grad<2, 1>(g, st);
    case 10: return launch_grad<2, 2>(g, st);
  }
  throw HfgError{HFG_UNSUPPORTED, "one-electron gradient class beyond d"};
}

template <int LA, int LB>
void launch_onee(const OneArgs& g, hipStream_t st) {
  hipLaunchKernelGGL((k_onee<LA, LB>), dim3((unsigned)g.npair), dim3(64), 0, st, g);
  HIP_CHECK(hipGetLastError());
}

void launch_class(int la, int lb, const OneArgs& g, hipStream_t st) {
  switch (la * 4 + lb) {
    case 0: return launch_onee<0, 0>(g, st);
    case 4: return launch_onee<1, 0>(g, st);
    case 5: return launch_onee<1, 1>(g, st);
    case 8: return launch_onee<2, 0>(g, st);
    case 9: return launch_onee<2, 1>(g, st);
    case 10: return launch_onee<2, 2>(g, st);
    case 12: return launch_onee<3, 0>(g, st);
    case 13: return launch_onee<3, 1>(g, st);
    case 14: return launch_onee<3, 2>(g, st);
    case 15: return launch_onee<3, 3>(g, st);
  }
  throw HfgError{HFG_UNSUPPORTED, "one-electron class beyond f"};
}

__global__ void k_sad_scatter(const double* __restrict__ blk, const int* __restrict__ boff,
                              const int* __restrict__ idx, const int* __restrict__ ioff, int n, double* __restrict__ P) {
  const int a = blockIdx.x;
  const int m = ioff[a + 1] - ioff[a];
  const double* b = blk + boff[a];
  const int* id = idx + ioff[a];
  for (int q = threadIdx.x; q < m * m; q += blockDim.x) {
    const int i = q / m, j = q - i * m;
    P[(size_t)id[i] * n + id[j]] = b[q];
  }
}

__global__ void k_scale(double* __restrict__ x, size_t cnt, double f) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) x[i] *= f;
}

__global__ void k_clear(double* __restrict__ x, size_t cnt) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) x[i] = 0.0;
}

}  // namespace

namespace {

// Device copy of what the one-electron kernels read: shells, primitives,
// component norms, nuclei, the Boys table and the shell-pair list by class
// (la * 4 + lb).  Unique pairs oriented L_A >= L_B, or every ordered pair.
struct OneStage {
  DevBuf d;
  OneArgs g{};
  std::vector<int> start;  // 17 entries: first pair of each class
  const int2* pairs = nullptr;

  OneStage(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms, bool ordered,
           hipStream_t st) {
    const int ns = (int)sh.size();
    if (ns == 0 || atoms.empty()) throw HfgError{HFG_INVALID_INPUT, "one-electron matrices need shells and nuclei"};
    std::vector<OneShell> hs(ns);
    std::vector<double> pe, pr, fscale(n, 0.0);
    for (int s = 0; s < ns; ++s) {
      const Shell& x = sh[s];
      if (x.L > ONEE_LMAX) throw HfgError{HFG_UNSUPPORTED, "one-electron integrals support shells up to f"};
      if (x.ao < 0 || x.ao + (int)x.scale.size() > n) throw HfgError{HFG_INVALID_INPUT, "shell beyond the basis"};
      hs[s] = OneShell{{x.c[0], x.c[1], x.c[2]}, x.L, x.K, (int)pe.size(), x.ao, x.atom};
      pe.insert(pe.end(), x.e.begin(), x.e.end());
      pr.insert(pr.end(), x.r.begin(), x.r.end());
      for (size_t c = 0; c < x.scale.size(); ++c) fscale[x.ao + c] = x.scale[c];
    }
    std::vector<std::vector<int2>> cls(16);
    for (int a = 0; a < ns; ++a)
      for (int b = 0; b < (ordered ? ns : a + 1); ++b) {
        int2 q = make_int2(a, b);
        if (!ordered && sh[a].L < sh[b].L) q = make_int2(b, a);
        cls[sh[q.x].L * 4 + sh[q.y].L].push_back(q);
      }
    std::vector<int2> pl;
    start.assign(17, 0);
    for (int c = 0; c < 16; ++c) {
      start[c] = (int)pl.size();
      pl.insert(pl.end(), cls[c].begin(), cls[c].end());
    }
    start[16] = (int)pl.size();
    std::vector<double4> at(atoms.size());
    for (size_t i = 0; i < atoms.size(); ++i)
      at[i] = make_double4(atoms[i][0], atoms[i][1], atoms[i][2], atoms[i][3]);
    const std::vector<double> tab = boys_table();

    auto bytes = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_sh = 0, o_pe = o_sh + bytes(hs.size() * sizeof(OneShell)), o_pr = o_pe + bytes(pe.size() * 8),
                 o_fs = o_pr + bytes(pr.size() * 8), o_at = o_fs + bytes(fscale.size() * 8),
                 o_pq = o_at + bytes(at.size() * sizeof(double4)), o_bt = o_pq + bytes(pl.size() * sizeof(int2)),
                 total = o_bt + bytes(tab.size() * 8);
    std::vector<char> host(total);
    auto put = [&](size_t off, const void* src, size_t len) {
      if (len) std::memcpy(host.data() + off, src, len);
    };
    put(o_sh, hs.data(), hs.size() * sizeof(OneShell));
    put(o_pe, pe.data(), pe.size() * 8);
    put(o_pr, pr.data(), pr.size() * 8);
    put(o_fs, fscale.data(), fscale.size() * 8);
    put(o_at, at.data(), at.size() * sizeof(double4));
    put(o_pq, pl.data(), pl.size() * sizeof(int2));
    put(o_bt, tab.data(), tab.size() * 8);
    d.alloc(total);
    char* base = d.as<char>();
    HIP_CHECK(hipMemcpyAsync(base, host.data(), total, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));  // the host staging copy goes out of scope
    g.sh = (const OneShell*)(base + o_sh);
    g.pe = (const double*)(base + o_pe);
    g.pr = (const double*)(base + o_pr);
    g.fscale = (const double*)(base + o_fs);
    g.atoms = (const double4*)(base + o_at);
    g.natom = (int)atoms.size();
    g.boys = (const double*)(base + o_bt);
    g.n = n;
    pairs = (const int2*)(base + o_pq);
  }
};

}  // namespace

void one_electron_device(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms,
                         double* dS, double* dT, double* dV, double* dH, hipStream_t st) {
  OneStage o(sh, n, atoms, false, st);
  OneArgs g = o.g;
  g.S = dS;
  g.T = dT;
  g.V = dV;
  g.H = dH;
  // every entry of the outputs is written by exactly one pair's block
  for (int c = 0; c < 16; ++c) {
    const int cnt = o.start[c + 1] - o.start[c];
    if (!cnt) continue;
    g.pairs = o.pairs + o.start[c];
    g.npair = cnt;
    launch_class(c / 4, c % 4, g, st);
  }
  HIP_CHECK(hipStreamSynchronize(st));  // the staging buffer goes out of scope
}

void onee_gradient_device(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms,
                          const double* dP, const double* dW, double* dgrad, hipStream_t st) {
  if ((size_t)3 * atoms.size() * sizeof(double) > 48 * 1024)
    throw HfgError{HFG_UNSUPPORTED, "gradient: too many atoms for the per-workgroup force sums"};
  for (const Shell& s : sh)
    if (s.atom < 0 || s.atom >= (int)atoms.size())
      throw HfgError{HFG_INVALID_INPUT, "basis atom index outside the geometry"};
  OneStage o(sh, n, atoms, true, st);
  OneArgs g = o.g;
  g.P = dP;
  g.W = dW;
  g.grad = dgrad;
  for (int c = 0; c < 16; ++c) {
    const int cnt = o.start[c + 1] - o.start[c];
    if (!cnt) continue;
    g.pairs = o.pairs + o.start[c];
    g.npair = cnt;
    launch_grad_class(c / 4, c % 4, g, st);
  }
  HIP_CHECK(hipStreamSynchronize(st));
}

void sad_guess_device(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms,
                      int charge, const double* dS, double* dP, Blas& bl, hipStream_t st) {
  double target = -charge;
  for (auto& a : atoms) target += a[0];
  for (auto& s : sh)
    if (s.atom < 0 || s.atom >= (int)atoms.size())
      throw HfgError{HFG_INVALID_INPUT, "the basis names an atom beyond the geometry"};
  if (!(target > 0)) throw HfgError{HFG_INVALID_INPUT, "SAD needs a positive electron count"};
  std::vector<double> blk;
  std::vector<int> boff, idx, ioff{0};
  for (size_t a = 0; a < atoms.size(); ++a) {
    std::vector<int> id;
    const std::vector<double>& p = atomic_density(sh, (int)a, atoms[a][0], id);
    boff.push_back((int)blk.size());
    blk.insert(blk.end(), p.begin(), p.end());
    idx.insert(idx.end(), id.begin(), id.end());
    ioff.push_back((int)idx.size());
  }
  for (int i : idx)
    if (i < 0 || i >= n) throw HfgError{HFG_INVALID_INPUT, "atomic density beyond the basis"};
  const size_t nn = (size_t)n * n;
  DevBuf d;
  auto bytes = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_b = 0, o_bo = bytes(blk.size() * 8), o_i = o_bo + bytes(boff.size() * 4),
               o_io = o_i + bytes(idx.size() * 4), total = o_io + bytes(ioff.size() * 4);
  std::vector<char> host(total);
  std::memcpy(host.data() + o_b, blk.data(), blk.size() * 8);
  std::memcpy(host.data() + o_bo, boff.data(), boff.size() * 4);
  std::memcpy(host.data() + o_i, idx.data(), idx.size() * 4);
  std::memcpy(host.data() + o_io, ioff.data(), ioff.size() * 4);
  d.alloc(total);
  char* base = d.as<char>();
  HIP_CHECK(hipMemcpyAsync(base, host.data(), total, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_clear, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, dP, nn);
  hipLaunchKernelGGL(k_sad_scatter, dim3((unsigned)atoms.size()), dim3(256), 0, st, (const double*)(base + o_b),
                     (const int*)(base + o_bo), (const int*)(base + o_i), (const int*)(base + o_io), n, dP);
  HIP_CHECK(hipGetLastError());
  // N_e = tr(P S) scaled to the electron count (src/sad.rs:240-256)
  const double el = bl.dot(nn, dP, dS);
  if (!(el > 1e-12)) throw HfgError{HFG_INTEGRAL, "zero SAD population"};
  hipLaunchKernelGGL(k_scale, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, dP, nn, target / el);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace hfg
