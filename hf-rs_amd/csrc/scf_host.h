// scf_host.h -- host-side helpers of the RHF driver.
#pragma once
#include <array>
#include <vector>

#include "host_prep.h"

namespace hfg {

void one_electron(const std::vector<Shell>& sh, int n, const std::vector<std::array<double, 4>>& atoms,
                  std::vector<double>& S, std::vector<double>& T, std::vector<double>& V);
void spherical_transform(const std::vector<Shell>& sh, int n, std::vector<double>& Tm, int& ncol);

}  // namespace hfg
