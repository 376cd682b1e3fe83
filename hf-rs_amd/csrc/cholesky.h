// cholesky.h -- pivoted Cholesky factors of the AO ERIs on the device
// (cholesky.hip), the factor source of CD-HF (src/hf_variants/cholesky_df.rs).
#pragma once
#include <vector>

#include "engine.h"

namespace hfg {

struct Blas;

struct CholeskyInfo {
  int rank = 0;
  double max_residual = 0.0;  // largest remaining residual diagonal
  std::vector<int> pivots;    // packed AO-pair index of each pivot, in order
};

// Factors of the AO ERIs of shells orb (n functions) to residual diagonal
// threshold, with at most budget bytes of packed factors (the reference's
// rule: 8 np (K + 2) <= budget before row K is added), on stream st; the
// result is d_B = [rank][n][n] (each factor unpacked, symmetric).
void cholesky_factors(const std::vector<Shell>& orb, int n, hipStream_t st, Blas& bl, double threshold, double budget,
                      DevBuf& d_B, int& rank, CholeskyInfo& info);

}  // namespace hfg
