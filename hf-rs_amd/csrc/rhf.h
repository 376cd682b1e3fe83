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

// The fractional-occupation atomic density of atom `atom` (nuclear charge z)
// in its own functions, row-major m x m, solved once per element and basis
// and cached (src/sad.rs:136-214); idx receives the m basis-function indices.
HMat atomic_density(const std::vector<Shell>& sh, int atom, double z, std::vector<int>& idx);
// The SAD guess on the device (src/sad.rs:217-259): the cached atomic blocks
// are scattered into dP (n x n), tr(P S) with the device overlap dS is taken
// and P is scaled to the electron count; throws when an atom has no basis
// functions.  onee_dev.hip.
void sad_guess_device(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms,
                      int charge, const double* dS, double* dP, Blas& bl, hipStream_t st);
// S, T, V and H = T + V (n x n each, device; any may be null) of the
// Cartesian basis sh for the nuclei (Z, x, y, z), compute_one_electron_integrals
// (src/scf.rs:246-271) on the device; onee_dev.hip.  Shells up to f.
void one_electron_device(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms,
                         double* dS, double* dT, double* dV, double* dH, hipStream_t st);
// Adds the one-electron terms of the RHF nuclear gradient (overlap, kinetic
// and attraction derivatives of every basis function and the operator
// derivative at every nucleus, src/gradient.rs:411-490) for the device
// density dP and energy-weighted density dW (n x n) to dgrad (natom x 3,
// device); onee_dev.hip.
// Nuclear gradient (natom x 3, host) for the device density dP and
// energy-weighted density dW (n x n) of the engine's molecule, on stream st;
// grad.hip.  Reads the engine's tables only.
void rhf_gradient_device(Engine& e, const std::vector<std::array<double, 4>>& atoms, const double* dP,
                         const double* dW, double* grad, hipStream_t st);
void onee_gradient_device(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms,
                          const double* dP, const double* dW, double* dgrad, hipStream_t st);

}  // namespace hfg
