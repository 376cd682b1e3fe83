// scf_host.h -- host-side helpers of the RHF driver.
#pragma once
#include <array>
#include <vector>

#include "host_prep.h"

namespace hfg {

void one_electron(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms,
                  std::vector<double>& S, std::vector<double>& T, std::vector<double>& V);
void spherical_transform(const std::vector<Shell>& sh, int n, std::vector<double>& Tm, int& ncol);

// Euclidean projection onto the probability simplex, and the minimiser of
// 1/2 c^T Q c + b^T c on it (Q row-major m x m): the ADIIS model
// (src/diis.rs:195-266), diis_host.cpp.
void simplex_project(std::vector<double>& v);
std::vector<double> simplex_minimise(const std::vector<double>& q, const std::vector<double>& b);
// Solution of a x = b (a row-major n x n) by Gaussian elimination with
// partial pivoting (src/diis.rs:532-589); false if a pivot vanishes.
bool solve_small(std::vector<double> a, std::vector<double> b, int n, std::vector<double>& x);

}  // namespace hfg
