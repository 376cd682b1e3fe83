// jk_plan.h -- tables of the direct J/K kernel and their construction
// (jk_plan.cpp).  Plain C++: the CPU check library builds the same plan.
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

#include "classes.h"
#include "host_prep.h"

namespace hfg {

// Shell pair as the kernels read it (kernel orientation, see PairInfo).
struct PairDev {
  int p0, p1, ao_a, ao_b, sh_a, sh_b, pad0, pad1;
  double AB[3], pad2;
};

// A view of a pair of generally contracted shells as the J/K kernel reads it:
// na (<= RA) members of the first shell times nb (<= RB) members of the
// second, kernel orientation (higher L first), with the primitive pairs that
// matter for them.  Primitive records [p0, p1) in the view-contiguous bra
// table, each holding the contraction coefficients of the RA * RB
// combinations (ga * RB + gb, zero for missing members; only when
// RA * RB > 1, otherwise the single combination's coefficient is folded into
// K); coef0 is unused; RA * RB combinations at c0 in
// the combination table (missing members: Schwarz bound 0); same = 1 when
// both sides are one shell.
struct ViewDev {
  int p0, p1, na, nb, c0, coef0, same, mol;  // mol: molecule of a batch (0 alone)
  double AB[3], pad2;
};

// One contraction combination = one hf-rs shell pair: AO offsets of its two
// shells (kernel orientation) and its Schwarz bound.
struct ComboDev {
  int ao_a, ao_b;
  double q;
};

// One wave of work: bra view, first ket view and number of ket views (<= 64);
// all ket views share one primitive count.
struct JKTask {
  int bra, ket0, nket, grp;
};

// A group of ket views of one pair class, cap and primitive count, contiguous
// in the view order; its primitive data is stored transposed,
// X[base + prim * size + (view - start)], and its coefficients as
// C[cbase + (prim * cap + g) * size + (view - start)], so that the lanes of a
// wave read consecutive addresses.
struct KetGroup {
  int start, size, base, npp, cbase, pad0, pad1, pad2;
};

struct BlockTask {
  int bra, ket;
  long long off;
};

// Contraction-member caps (RA, RB, RC, RD) of the J/K classes.
constexpr int GC_CAPS[HFG_NCLASS][4] = {HFG_GC_CAPS};
// Classes whose kernels are also compiled with one contraction member per
// shell (bit c for class c): a molecule whose shells of the class's angular
// momenta are all single -- every shell of a segmented basis -- runs that
// variant, which does not carry the RA x RB x RC x RD combinations the capped
// kernels evaluate for one nonzero.
constexpr unsigned UNIT_VARIANT_MASK = HFG_UNIT_VARIANT_MASK;
// Caps of class cls in kernel variant v (0: GC_CAPS, 1: one member per shell).
constexpr int gc_cap(int cls, int v, int k) { return v ? 1 : GC_CAPS[cls][k]; }
// Launch slots of a build: class c in variant v is slot c + v * HFG_NCLASS.
constexpr int HFG_NSLOT = 2 * HFG_NCLASS;

// Output-slice mode of the J/K classes: the Cartesian block of a quartet is
// evaluated whole (0) or in slices of one a component (1), one c component
// (2) or one (a, c) pair (3); slice_count is the number of slices of a class
// with bra/ket first shells of angular momentum la and lc.
constexpr int SLICE_MODE[HFG_NCLASS] = {HFG_SLICE_MODES};
constexpr int slice_count(int cls, int la, int lc) {
  const int na = (la + 1) * (la + 2) / 2, nc = (lc + 1) * (lc + 2) / 2;
  return SLICE_MODE[cls] == 0 ? 1 : SLICE_MODE[cls] == 1 ? na : SLICE_MODE[cls] == 2 ? nc : na * nc;
}

// Everything the J/K kernels read besides the density, built on the host.
struct JKPlan {
  std::vector<ViewDev> views;
  std::vector<double> view_q;  // per view: the largest Schwarz bound of its combinations
  std::vector<ComboDev> combos;
  std::vector<KetGroup> groups;
  std::vector<double> vb;     // bra primitive records, HFG_BRA_REC doubles each (vb_size records)
  std::vector<double> kt;     // ket primitive pairs, 9 columns of kt_size
  std::vector<double> kcoef;  // ket coefficients
  std::vector<JKTask> tasks;
  std::vector<size_t> task_begin;  // HFG_NCLASS + 1 entries
  // Jobs: tasks of one bra view, [jobs[j], jobs[j + 1]) (absolute task
  // indices, one sentinel at the end); class c owns jobs job_begin[c] ..
  // job_begin[c + 1] - 1.
  std::vector<int> jobs;
  std::vector<double> job_cost;   // estimated cost per job (same order)
  std::vector<size_t> job_begin;  // HFG_NCLASS + 1 entries
  size_t vb_size = 0, kt_size = 0, prim_pairs_kept = 0;
  int n_gen_shells = 0;
  // Classes planned with one member per shell (kernel variant 1), bit c.
  unsigned unit_mask = 0;
};

// Unscaled blocks of the listed quartets (one task per quartet, grouped by
// class, offsets into out) of primitive pair table `prims` through pair table
// `tab`: supplied by the caller (GPU block kernels, or the CPU check).
using BlockFn = std::function<void(const PrimPairHost& prims, const std::vector<PairDev>& tab,
                                   const std::vector<BlockTask>& tasks, const std::vector<int>& cls,
                                   std::vector<double>& out)>;

// Q: Schwarz bound per hf-rs shell pair (index a(a+1)/2 + b, a >= b);
// ao_scale: component factor per AO.
// caps: contraction-member caps per class.  When null, GC_CAPS, except that
// a class of UNIT_VARIANT_MASK whose angular momenta have only single-member
// shells in this basis gets unit caps and its unit_mask bit (HFG_UNIT_VARIANT=0
// turns that off); other values only for work-count studies on the CPU, the
// kernels are compiled for these two variants.
JKPlan make_jk_plan(const std::vector<Shell>& shells, const std::vector<double>& Q, const std::vector<double>& ao_scale,
                    double screen, double prim_screen, const BlockFn& blocks, const int (*caps)[4] = nullptr);

}  // namespace hfg
