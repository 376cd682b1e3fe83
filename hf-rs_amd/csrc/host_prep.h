// host_prep.h -- host-side preparation of the engine's device data:
// shell grouping, primitive-pair tables, Boys table, quartet work lists.
#pragma once
#include <array>
#include <string>
#include <vector>

#include "../../include/hfg.h"

namespace hfg {

struct HfgError {
  int code;
  std::string msg;
};

// One complete Cartesian shell (src/ao.rs:133-164): L, centre, exponents,
// radial coefficients r_k and per-component factors s_c such that the hf-rs
// normalised coefficient of component c, primitive k is r_k * s_c.
struct Shell {
  int L = 0, K = 0, ao = 0, atom = 0;
  double c[3] = {0, 0, 0};
  std::vector<double> e, r, scale;
};

std::vector<Shell> group_shells(const hfg_basis_view& b);  // throws HfgError

// A unique shell pair a >= b (hf-rs pair order), stored in kernel orientation
// first = the shell of higher L.
struct PairInfo {
  int p0, p1;        // primitive-pair range
  int ao_a, ao_b;    // AO offsets of first / second shell (kernel orientation)
  int sh_a, sh_b;    // shell ids (kernel orientation)
  double AB[3];      // A - B
  int la, lb;        // kernel orientation
  int hf_index;      // position in hf-rs pair order (a*(a+1)/2 + b)
  int swapped;       // 1 if first = hf-rs b
};

struct PrimPairHost {
  std::vector<double> p, Px, Py, Pz, PAx, PAy, PAz, K, H;
  size_t size() const { return p.size(); }
};

void build_pairs(const std::vector<Shell>& sh, std::vector<PairInfo>& pairs, PrimPairHost& pp);
void append_pair(const Shell& A, const Shell& B, PrimPairHost& pp);  // kernel orientation A, B
// Auxiliary shell P paired with the exponent-0 unit function at its centre
// (src/df.rs:30-40): p = alpha, P = C, PA = 0, K = c / alpha.
void append_aux_pair(const Shell& P, PrimPairHost& pp);

// Boys table (layout documented in eri_device.h), computed in long double.
std::vector<double> boys_table();

int pair_class(int la, int lb);  // (ss,ps,pp,ds,dp,dd,fs,gs,fp,fd,ff) index, -1 beyond
int quartet_class(int pc_bra, int pc_ket);  // pc_bra >= pc_ket, index 0..20

}  // namespace hfg
