// rhf.h -- pieces of the RHF set-up shared by the single-molecule driver
// (scf.cpp) and the batched one (bscf.hip).
#pragma once
#include <array>
#include <vector>

#include "engine.h"
#include "la_device.h"
#include "scf_host.h"

namespace hfg {

typedef std::vector<double> HMat;

// Atoms (Z, x, y, z in bohr), electron and occupied-orbital counts and the
// nuclear repulsion of one RHF problem, with the input checks of
// run_scf_in_space_with_convergence (src/scf.rs:415-447).
struct RhfInputs {
  std::vector<std::array<double, 4>> atoms;
  int nel = 0, nocc = 0;
  double enuc = 0;
};
RhfInputs rhf_inputs(const hfg_basis_view& basis, const double* zxyz, int natom, int charge);

// Canonical orthogonaliser of the overlap dS (device, n x n), spherical or
// Cartesian (src/ao.rs:58-67); returns m and X (host, n x m).
int orthogonalizer(Blas& bl, hipStream_t st, const std::vector<Shell>& shells, int n, const double* dS, bool spherical,
                   HMat& X);

// Superposition of atomic densities (src/sad.rs:56-134), scaled to the
// electron count; throws when an atom has no basis functions.
HMat sad_guess(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms, int charge,
               const HMat& S);

}  // namespace hfg
