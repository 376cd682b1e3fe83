// eri_device.h -- integral arithmetic shared by the HIP kernels.
//
// Compiled by hipcc for gfx950.  The same text also compiles as plain C++
// (HFG_HOST_CHECK defined) so that the CPU test-suite can compare the
// generated recurrences with the oracle before any GPU run; that host build is
// a test artefact, never a fallback of the engine.
#pragma once
#include <cmath>
#include <cstdint>

#ifdef HFG_HOST_CHECK
#define HFG_DEV inline
#else
#include <hip/hip_runtime.h>
#define HFG_DEV __host__ __device__ __forceinline__
#endif

namespace hfg {

// Read-only data whose address is uniform across the wave: loads through the
// constant address space become scalar loads (the compiler cannot otherwise
// prove that the kernel's atomics do not alias them).
#if defined(__HIP_DEVICE_COMPILE__)
typedef const double __attribute__((address_space(4)))* cdptr;
#else
typedef const double* cdptr;
#endif

constexpr double TWO_PI_2_5 = 34.986836655249726164477;  // 2 pi^(5/2)
constexpr double HALF_SQRT_PI = 0.88622692545275801365;

// ---------------------------------------------------------------------------
// Boys function F_m(T), m = 0..M.
//
// Table: rows T_r = r/8 (r = 0..BOYS_ROWS-1, T up to 50), each row holding
// exp(-T_r) followed by F_0(T_r) .. F_{BOYS_ORDERS-1}(T_r).  F_M is a 9-term
// Taylor expansion about the nearest row, d^k F_m/dT^k = (-1)^k F_{m+k}, with
// |dT| <= 1/16 so the truncation is below F_M * (1/16)^9/9! ~ 5e-17 relative;
// lower orders follow by the stable downward recursion
// F_{m-1} = (2T F_m + e^{-T})/(2m-1), e^{-T} = e^{-T_r} e^{-dT}.  For T >= 50,
// F_0 = sqrt(pi/T)/2 (erf(sqrt T) = 1 to 3e-24) and the upward recursion
// F_m = ((2m-1) F_{m-1} - e^{-T})/(2T) is stable.
// ---------------------------------------------------------------------------
// 1/sqrt(x) for x > 0: the hardware estimate refined by two Newton steps
// y <- y (3/2 - x/2 y^2), each squaring the relative error.
HFG_DEV double rsqrt_nr(double x) {
#ifdef HFG_HOST_CHECK
  double y = 1.0 / std::sqrt(x);
#elif defined(__HIP_DEVICE_COMPILE__)
  double y = __builtin_amdgcn_rsq(x);
#else
  double y = 1.0 / std::sqrt(x);
#endif
  const double h = 0.5 * x;
  y = y * (1.5 - h * y * y);
  y = y * (1.5 - h * y * y);
  return y;
}

constexpr int BOYS_TAYLOR = 9;
constexpr int BOYS_MAXM = 24;                       // highest order served
constexpr int BOYS_ORDERS = BOYS_MAXM + BOYS_TAYLOR;  // F_0 .. F_32 stored
constexpr int BOYS_ROW = BOYS_ORDERS + 1;           // + exp(-T_r)
constexpr double BOYS_INV_STEP = 8.0;
constexpr double BOYS_TMAX = 50.0;
constexpr int BOYS_ROWS = 402;

// Per-class slice of the table, as the kernels keep it in LDS: rows of
// exp(-T_r) followed by the Taylor coefficients F_{M+k}(T_r)/k!, so that each
// Horner step is one FMA.  Two layouts, picked per kernel at compile time:
//   full     every table row (T_r = r/8, 402 rows), nine terms, |dT| <= 1/16,
//            remainder below (1/16)^9/9! = 4e-17 of F_M; 10 doubles a row, 32 KB;
//   compact  every other row (T_r = r/4, 201 rows), eleven terms, |dT| <= 1/8,
//            remainder below (1/8)^11/11! = 3e-18; 12 doubles a row, 19 KB.
// The slice is the largest LDS item of a J/K workgroup: the compact one leaves
// room for more workgroups per CU at two more FMAs per evaluation.
template <bool C>
struct BoysSlice {
  static constexpr int taylor = C ? 11 : BOYS_TAYLOR;
  static constexpr int cols = taylor + 1;
  static constexpr int step = C ? 2 : 1;  // table rows per slice row
  static constexpr int rows = C ? (BOYS_ROWS + 1) / 2 : BOYS_ROWS;
  static constexpr int size = rows * cols;
  static constexpr double inv_step = BOYS_INV_STEP / step;
  static constexpr int nexp = C ? 12 : 10;  // terms of e^{-dT}: remainder < 2e-19
};
// Measured on the def2-SVP batch and on (H2O)16/cc-pVDZ: the one-member
// (segmented-basis) variants of the classes above (ss|ss) run faster with the
// compact layout (they are LDS-bound); the general-contraction kernels, which
// evaluate more primitive quartets per digested block, with the full one.
constexpr bool boys_compact(int V, int L) { return V == 1 && L > 0; }

template <int M, bool C = false>
HFG_DEV void boys_slice_row(const double* __restrict__ tab, int r, double* __restrict__ row) {
  using B = BoysSlice<C>;
  static_assert(M + B::taylor <= BOYS_ORDERS, "Boys table holds too few orders for this slice");
  const int g = B::step * r;  // table row of T = r / inv_step
  row[0] = tab[g * BOYS_ROW];
  double inv_fact = 1.0;
  for (int k = 0; k < B::taylor; ++k) {
    if (k > 0) inv_fact /= k;
    row[1 + k] = tab[g * BOYS_ROW + 1 + M + k] * inv_fact;
  }
}

// 1/k!, k = 0..12: the series of e^{-dT} around a slice row.
constexpr double BOYS_EXP_C[13] = {1.0,
                                   1.0,
                                   1.0 / 2,
                                   1.0 / 6,
                                   1.0 / 24,
                                   1.0 / 120,
                                   1.0 / 720,
                                   1.0 / 5040,
                                   1.0 / 40320,
                                   1.0 / 362880,
                                   1.0 / 3628800,
                                   1.0 / 39916800,
                                   1.0 / 479001600};

template <int M, bool SELECT = false, bool C = false>
HFG_DEV void boys_eval(double T, const double* __restrict__ slice, double* __restrict__ F) {
  using B = BoysSlice<C>;
  if constexpr (SELECT && M <= 4) {
    // Both branches below, evaluated without control flow and selected per
    // lane: the loop body stays one basic block, so that the scheduler can
    // overlap the table reads of consecutive primitives (used where the bra
    // loop is unrolled; it costs registers elsewhere).
    const bool near = T < BOYS_TMAX;
    const double Tn = near ? T : 0.0;
    const int r = (int)(Tn * B::inv_step + 0.5);
    const double x = r * (1.0 / B::inv_step) - Tn;
    const double* row = slice + r * B::cols;
    double acc = row[B::taylor];
#pragma unroll
    for (int k = B::taylor - 1; k >= 1; --k) acc = fma(acc, x, row[k]);
    double Fn[M + 1];
    Fn[M] = acc;
    if (M > 0) {
      double e = BOYS_EXP_C[B::nexp - 1];
#pragma unroll
      for (int k = B::nexp - 2; k >= 0; --k) e = fma(e, x, BOYS_EXP_C[k]);
      const double emt = row[0] * e;
      const double t2 = 2.0 * Tn;
#pragma unroll
      for (int m = M - 1; m >= 0; --m) Fn[m] = (t2 * Fn[m + 1] + emt) * (1.0 / (2 * m + 1));
    }
    const double rt = rsqrt_nr(near ? 1.0 : T);
    double Ff = HALF_SQRT_PI * rt;
    const double h = 0.5 * rt * rt;
    F[0] = near ? Fn[0] : Ff;
#pragma unroll
    for (int m = 1; m <= M; ++m) {
      Ff = (2 * m - 1) * Ff * h;
      F[m] = near ? Fn[m] : Ff;
    }
    return;
  }
  if (T < BOYS_TMAX) {
    const int r = (int)(T * B::inv_step + 0.5);
    const double x = r * (1.0 / B::inv_step) - T;  // -dT
    const double* row = slice + r * B::cols;
    // F_M(T) = sum_k [F_{M+k}(T_r)/k!] (-dT)^k
    double acc = row[B::taylor];
#pragma unroll
    for (int k = B::taylor - 1; k >= 1; --k) acc = fma(acc, x, row[k]);
    F[M] = acc;
    if (M > 0) {
      // e^{-dT} = sum_k x^k/k!
      double e = BOYS_EXP_C[B::nexp - 1];
#pragma unroll
      for (int k = B::nexp - 2; k >= 0; --k) e = fma(e, x, BOYS_EXP_C[k]);
      const double emt = row[0] * e;
      const double t2 = 2.0 * T;
#pragma unroll
      for (int m = M - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + emt) * (1.0 / (2 * m + 1));
    }
  } else {
    // F_0 = sqrt(pi/T)/2 (erf(sqrt T) = 1 to 3e-24), upward recursion
    // F_m = ((2m-1) F_{m-1} - e^{-T})/(2T).  For M <= 4 the e^{-T} terms
    // (< 2e-22/T) change F_m by less than 2e-17 relative and are left out.
    const double rt = rsqrt_nr(T);
    F[0] = HALF_SQRT_PI * rt;
    if (M > 0) {
      const double emt = M > 4 ? exp(-T) : 0.0;
      const double h = 0.5 * rt * rt;
#pragma unroll
      for (int m = 1; m <= M; ++m) F[m] = ((2 * m - 1) * F[m - 1] - emt) * h;
    }
  }
}

// Runtime-order variant for the general path.
HFG_DEV void boys_eval_rt(int M, double T, const double* __restrict__ tab, double* __restrict__ F) {
  if (T < BOYS_TMAX) {
    const int r = (int)(T * BOYS_INV_STEP + 0.5);
    const double dT = T - r * (1.0 / BOYS_INV_STEP);
    const double* row = tab + r * BOYS_ROW;
    double acc = row[1 + M + BOYS_TAYLOR - 1];
    for (int k = BOYS_TAYLOR - 1; k >= 1; --k) acc = row[1 + M + k - 1] + acc * (-dT) * (1.0 / k);
    F[M] = acc;
    double e = 1.0;
    for (int k = 9; k >= 1; --k) e = 1.0 + e * (-dT) * (1.0 / k);
    const double emt = row[0] * e, t2 = 2.0 * T;
    for (int m = M - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + emt) / (2 * m + 1);
  } else {
    const double inv_t = 1.0 / T;
    F[0] = HALF_SQRT_PI * sqrt(inv_t);
    const double emt = exp(-T), h = 0.5 * inv_t;
    for (int m = 1; m <= M; ++m) F[m] = ((2 * m - 1) * F[m - 1] - emt) * h;
  }
}

// Primitive-pair data, structure of arrays.  For a pair of shells (A, B) in
// kernel orientation (L_A >= L_B): p = alpha+beta, P = (alpha A + beta B)/p,
// PA = P - A, K = c_A c_B exp(-alpha beta |AB|^2 / p) / p and H = 1/(2p).
struct PrimPairs {
  const double* __restrict__ p;
  const double* __restrict__ Px;
  const double* __restrict__ Py;
  const double* __restrict__ Pz;
  const double* __restrict__ PAx;
  const double* __restrict__ PAy;
  const double* __restrict__ PAz;
  const double* __restrict__ K;
  const double* __restrict__ H;
};

// Bra primitives b0..b1-1 (contiguous); ket primitives k0 + j*ks, j < nk.
template <int LA, int LB, int LC, int LD>
HFG_DEV void eri_block(const PrimPairs& bp, int b0, int b1, const PrimPairs& kp, int k0, int nk, int ks,
                       const double AB[3], const double CD[3], const double* __restrict__ boys, double* __restrict__ I);

constexpr int ncart(int L) { return (L + 1) * (L + 2) / 2; }

#ifndef HFG_HOST_CHECK
// Boys slice of order M into LDS by the whole workgroup (layout above).
template <int M, bool C = false>
__device__ __forceinline__ void load_boys_slice(const double* __restrict__ g, double* __restrict__ lds) {
  using B = BoysSlice<C>;
  for (int r = threadIdx.x; r < B::rows; r += blockDim.x) boys_slice_row<M, C>(g, r, lds + r * B::cols);
  __syncthreads();
}
#endif

}  // namespace hfg

namespace hfg {
#include "gen/eri_classes.inc"
}
