// jk_plan.cpp -- the work plan of the direct J/K kernel.
//
// hf-rs stores a generally contracted basis (cc-pVXZ: several contractions
// over one exponent set) as separate segmented shells, and its shell engine
// evaluates the primitive integrals of each contracted quartet on their own
// (src/shell_integrals.rs:101-133, src/basis.rs:89-98).  Here the shells of
// one atom and angular momentum whose exponent sets nest are merged into one
// generally contracted shell; the primitive integrals of a pair of such shells
// are evaluated once and contracted with the coefficients of every
// combination of member shells.  Each combination is one hf-rs shell pair and
// is digested with its own Schwarz bound, so the set of quartets that enter J
// and K is exactly the reference's (Q_ab * Q_cd > 1e-12,
// src/shell_integrals.rs:184,226).
//
// Weights: a view of a pair of one merged shell with itself lists both
// orders of every member pair, a task whose bra and ket are one view lists
// both orders of every combination pair; each such duplicate carries weight
// 1/2, which reproduces 1/((1+d_AB)(1+d_CD)(1+d_AB,CD)) of the unique hf-rs
// quartets (the digestion sums the eight permutations).
#include <algorithm>
#include <cstdlib>
#include <cstdio>
#include <chrono>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>

#include "jk_plan.h"

namespace hfg {

namespace {

struct GenShell {
  int L = 0, atom = 0;
  double c[3] = {0, 0, 0};
  std::vector<double> e;               // exponents
  std::vector<int> mem;                // member hf-rs shells
  std::vector<std::vector<double>> r;  // per member: coefficient of each exponent (0 if absent)
};

// Merge the shells of one atom and L whose exponents are all found (exactly)
// in an earlier shell's set; larger sets are visited first.
std::vector<GenShell> merge_shells(const std::vector<Shell>& sh) {
  std::vector<int> order(sh.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    if (sh[x].atom != sh[y].atom) return sh[x].atom < sh[y].atom;
    if (sh[x].L != sh[y].L) return sh[x].L < sh[y].L;
    return sh[x].K > sh[y].K;
  });
  std::vector<GenShell> gs;
  for (int s : order) {
    const Shell& S = sh[s];
    int found = -1;
    std::vector<int> map(S.K);
    for (int g = (int)gs.size() - 1; g >= 0 && found < 0; --g) {
      const GenShell& G = gs[g];
      if (G.atom != S.atom || G.L != S.L) break;  // sorted: earlier entries differ too
      if (std::memcmp(G.c, S.c, sizeof G.c) != 0) continue;
      bool ok = true;
      for (int k = 0; k < S.K && ok; ++k) {
        auto it = std::find(G.e.begin(), G.e.end(), S.e[k]);
        if (it == G.e.end())
          ok = false;
        else
          map[k] = (int)(it - G.e.begin());
      }
      if (ok) found = g;
    }
    if (found < 0) {
      GenShell G;
      G.L = S.L;
      G.atom = S.atom;
      std::memcpy(G.c, S.c, sizeof G.c);
      G.e = S.e;
      gs.push_back(G);
      found = (int)gs.size() - 1;
      for (int k = 0; k < S.K; ++k) map[k] = k;
    }
    GenShell& G = gs[found];
    std::vector<double> r(G.e.size(), 0.0);
    for (int k = 0; k < S.K; ++k) r[map[k]] += S.r[k];
    G.mem.push_back(s);
    G.r.push_back(r);
  }
  return gs;
}

// A pair of merged shells, kernel orientation, with its primitive pairs
// (unit contraction coefficients) and the coefficient of every combination.
struct GenPair {
  int A, B, la, lb, pc;
  bool same;
  double AB[3];
  int q0, q1;                            // primitive pairs in the unit-coefficient table
  std::vector<std::array<int, 2>> combos;  // member indices (a, b)
  std::vector<std::vector<double>> coef;   // [combo][prim]
};

const int CLASS_L4[HFG_NCLASS][4] = {
#define X(a, b, c, d) {a, b, c, d},
    HFG_FOR_EACH_CLASS(X)
#undef X
};

int ncomp(int L) { return (L + 1) * (L + 2) / 2; }

}  // namespace

JKPlan make_jk_plan(const std::vector<Shell>& shells, const std::vector<double>& Q, const std::vector<double>& ao_scale,
                    double screen, double prim_screen, const BlockFn& blocks, const int (*caps)[4]) {
  JKPlan plan;
  const bool trace = std::getenv("HFG_TRACE") != nullptr;
  auto t0 = std::chrono::steady_clock::now();
  auto lap = [&](const char* what) {
    if (!trace) return;
    const auto t1 = std::chrono::steady_clock::now();
    std::fprintf(stderr, "hfg plan   %-10s %8.2f ms\n", what, std::chrono::duration<double>(t1 - t0).count() * 1e3);
    t0 = t1;
  };
  const std::vector<GenShell> gs = merge_shells(shells);
  plan.n_gen_shells = (int)gs.size();
  int auto_caps[HFG_NCLASS][4];
  if (!caps) {
    static const bool unit_on = [] {
      const char* e = std::getenv("HFG_UNIT_VARIANT");
      return !(e && e[0] == '0');
    }();
    int members[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // largest member count per L
    for (const GenShell& g : gs) members[g.L & 7] = std::max(members[g.L & 7], (int)g.mem.size());
    for (int c = 0; c < HFG_NCLASS; ++c) {
      bool single = unit_on && ((UNIT_VARIANT_MASK >> c) & 1u);
      for (int k = 0; k < 4; ++k) single = single && members[CLASS_L4[c][k]] <= 1;
      for (int k = 0; k < 4; ++k) auto_caps[c][k] = gc_cap(c, single ? 1 : 0, k);
      if (single) plan.unit_mask |= 1u << c;
    }
    caps = auto_caps;
  }
  auto seg_pair = [](int a, int b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; };

  // ---- merged shell pairs and their unit-coefficient primitive pairs
  std::vector<GenPair> gp;
  PrimPairHost up;  // unit-coefficient primitive pairs
  std::vector<std::array<int, 2>> uk;  // (ka, kb) of each
  for (int i = 0; i < (int)gs.size(); ++i)
    for (int j = 0; j <= i; ++j) {
      const bool sw = gs[j].L > gs[i].L;
      const GenShell& A = sw ? gs[j] : gs[i];
      const GenShell& B = sw ? gs[i] : gs[j];
      GenPair P;
      P.A = sw ? j : i;
      P.B = sw ? i : j;
      P.la = A.L;
      P.lb = B.L;
      P.pc = pair_class(A.L, B.L);
      P.same = i == j;
      for (int k = 0; k < 3; ++k) P.AB[k] = A.c[k] - B.c[k];
      const double ab2 = P.AB[0] * P.AB[0] + P.AB[1] * P.AB[1] + P.AB[2] * P.AB[2];
      P.q0 = (int)up.size();
      for (int ka = 0; ka < (int)A.e.size(); ++ka)
        for (int kb = 0; kb < (int)B.e.size(); ++kb) {
          if (P.same && kb > ka) continue;  // (ka,kb) and (kb,ka) share one integral: folded
          const double a = A.e[ka], b = B.e[kb], p = a + b;
          double Pc[3];
          for (int k = 0; k < 3; ++k) Pc[k] = (a * A.c[k] + b * B.c[k]) / p;
          up.p.push_back(p);
          up.Px.push_back(Pc[0]);
          up.Py.push_back(Pc[1]);
          up.Pz.push_back(Pc[2]);
          up.PAx.push_back(Pc[0] - A.c[0]);
          up.PAy.push_back(Pc[1] - A.c[1]);
          up.PAz.push_back(Pc[2] - A.c[2]);
          up.K.push_back(std::exp(-a * b / p * ab2) / p);
          up.H.push_back(0.5 / p);
          uk.push_back({ka, kb});
        }
      P.q1 = (int)up.size();
      for (int ma = 0; ma < (int)A.mem.size(); ++ma)
        for (int mb = 0; mb < (int)B.mem.size(); ++mb) {
          P.combos.push_back({ma, mb});
          std::vector<double> c(P.q1 - P.q0);
          for (int q = P.q0; q < P.q1; ++q) {
            const int ka = uk[q][0], kb = uk[q][1];
            double v = A.r[ma][ka] * B.r[mb][kb];
            if (P.same && ka != kb) v += A.r[ma][kb] * B.r[mb][ka];
            c[q - P.q0] = v;
          }
          P.coef.push_back(std::move(c));
        }
      gp.push_back(std::move(P));
    }

  lap("pairs");
  // ---- bound of every unit primitive pair: B_q = sqrt(max_ij |(ij|ij)|) with
  // the largest component factors of the two shells, from one block each.
  std::vector<double> Bu(up.size(), 0.0);
  if (up.size() > 0) {
    const size_t nu = up.size();
    std::vector<int> order(gp.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return gp[x].pc < gp[y].pc; });
    std::vector<PairDev> tab(nu);
    std::vector<BlockTask> tasks;
    std::vector<int> cls;
    std::vector<long long> off(nu);
    long long o = 0;
    for (int i : order) {
      const GenPair& P = gp[i];
      const int nab = ncomp(P.la) * ncomp(P.lb);
      for (int q = P.q0; q < P.q1; ++q) {
        PairDev& d = tab[q];
        std::memset(&d, 0, sizeof d);
        d.p0 = q;
        d.p1 = q + 1;
        for (int k = 0; k < 3; ++k) d.AB[k] = P.AB[k];
        tasks.push_back({q, q, o});
        cls.push_back(quartet_class(P.pc, P.pc));
        off[q] = o;
        o += (long long)nab * nab;
      }
    }
    std::vector<double> out(o);
    blocks(up, tab, tasks, cls, out);
    for (const GenPair& P : gp) {
      const GenShell& A = gs[P.A];
      const GenShell& B = gs[P.B];
      const int na = ncomp(P.la), nb = ncomp(P.lb), nab = na * nb;
      std::vector<double> fa(na, 0.0), fb(nb, 0.0);
      for (int s : A.mem)
        for (int x = 0; x < na; ++x) fa[x] = std::max(fa[x], std::fabs(ao_scale[shells[s].ao + x]));
      for (int s : B.mem)
        for (int y = 0; y < nb; ++y) fb[y] = std::max(fb[y], std::fabs(ao_scale[shells[s].ao + y]));
      for (int q = P.q0; q < P.q1; ++q) {
        double m = 0.0;
        for (int x = 0; x < na; ++x)
          for (int y = 0; y < nb; ++y) {
            const int ij = x * nb + y;
            m = std::max(m, std::fabs(out[off[q] + (size_t)ij * nab + ij]) * fa[x] * fa[x] * fb[y] * fb[y]);
          }
        Bu[q] = std::sqrt(m);
      }
    }
  }
  double bmax = 0.0;
  for (const GenPair& P : gp)
    for (const auto& c : P.coef)
      for (int q = P.q0; q < P.q1; ++q) bmax = std::max(bmax, std::fabs(c[q - P.q0]) * Bu[q]);

  lap("bounds");
  // ---- views per (pair class, member caps RA x RB) that some class uses:
  // the members of each shell are cut into chunks of at most RA (RB); a view
  // is one chunk of each, its combinations padded to RA x RB (index -1).
  struct View {
    int pair, na, nb;
    std::vector<int> combos;  // RA * RB entries: index into GenPair::combos, or -1
    std::vector<int> prims;   // kept primitive pairs (unit table ids)
    double qmax;
  };
  struct ViewSet {
    int pc, ra, rb;
    std::vector<View> views;
  };
  std::vector<ViewSet> sets;
  auto set_of = [&](int pc, int ra, int rb) {
    for (size_t s = 0; s < sets.size(); ++s)
      if (sets[s].pc == pc && sets[s].ra == ra && sets[s].rb == rb) return (int)s;
    sets.push_back({pc, ra, rb, {}});
    return (int)sets.size() - 1;
  };
  std::vector<std::array<int, 2>> class_sets(HFG_NCLASS);
  for (int c = 0; c < HFG_NCLASS; ++c) {
    const int pcb = pair_class(CLASS_L4[c][0], CLASS_L4[c][1]), pck = pair_class(CLASS_L4[c][2], CLASS_L4[c][3]);
    class_sets[c] = {set_of(pcb, caps[c][0], caps[c][1]), set_of(pck, caps[c][2], caps[c][3])};
  }
  for (ViewSet& S : sets)
    for (int ip = 0; ip < (int)gp.size(); ++ip) {
      const GenPair& P = gp[ip];
      if (P.pc != S.pc) continue;
      const int Ga = (int)gs[P.A].mem.size(), Gb = (int)gs[P.B].mem.size();
      for (int a0 = 0; a0 < Ga; a0 += S.ra)
        for (int b0 = 0; b0 < Gb; b0 += S.rb) {
          View v;
          v.pair = ip;
          v.na = std::min(S.ra, Ga - a0);
          v.nb = std::min(S.rb, Gb - b0);
          v.combos.assign(S.ra * S.rb, -1);
          v.qmax = 0.0;
          for (int x = 0; x < v.na; ++x)
            for (int y = 0; y < v.nb; ++y) {
              const int ci = (a0 + x) * Gb + (b0 + y);
              v.combos[x * S.rb + y] = ci;
              v.qmax = std::max(v.qmax, Q[seg_pair(gs[P.A].mem[a0 + x], gs[P.B].mem[b0 + y])]);
            }
          for (int q = P.q0; q < P.q1; ++q) {
            double cm = 0.0;
            for (int ci : v.combos)
              if (ci >= 0) cm = std::max(cm, std::fabs(P.coef[ci][q - P.q0]));
            if (cm > 0.0 && cm * Bu[q] * bmax >= prim_screen) v.prims.push_back(q);
          }
          if (!v.prims.empty()) S.views.push_back(std::move(v));
        }
    }

  lap("views");
  // ---- ket groups per set: views whose primitive counts fall in one bucket
  // (steps of at most 25%), padded to the largest count of the group with
  // zero-weight primitives, sorted by qmax descending; heavy groups first.
  // Few, large groups keep the lanes of a task full (exact counts left most
  // groups small and the surviving prefix of a bra short).  Global view ids:
  // groups are contiguous.
  struct Group {
    int set, npp, bucket;
    std::vector<int> v;  // view indices in the set
  };
  auto bucket_of = [](int npp) {
    static const int B[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 56, 64, 72, 81};
    for (int b : B)
      if (npp <= b) return b;
    return npp;
  };
  std::vector<Group> groups;
  std::vector<std::vector<int>> set_groups(sets.size());
  for (size_t s = 0; s < sets.size(); ++s) {
    std::vector<Group> gl;
    for (int i = 0; i < (int)sets[s].views.size(); ++i) {
      const int npp = (int)sets[s].views[i].prims.size(), bk = bucket_of(npp);
      auto it = std::find_if(gl.begin(), gl.end(), [&](const Group& g) { return g.bucket == bk; });
      if (it == gl.end()) {
        gl.push_back({(int)s, 0, bk, {}});
        it = gl.end() - 1;
      }
      it->npp = std::max(it->npp, npp);
      it->v.push_back(i);
    }
    for (auto& g : gl)
      std::stable_sort(g.v.begin(), g.v.end(),
                       [&](int x, int y) { return sets[s].views[x].qmax > sets[s].views[y].qmax; });
    std::stable_sort(gl.begin(), gl.end(), [](const Group& x, const Group& y) { return x.npp > y.npp; });
    for (auto& g : gl) {
      set_groups[s].push_back((int)groups.size());
      groups.push_back(std::move(g));
    }
  }

  lap("groups");
  // ---- device tables.  Each group's views, combinations and bra records
  // start at offsets fixed by the groups before it, so the groups are filled
  // in parallel.
  std::vector<KetGroup> kg(groups.size());
  std::vector<size_t> v_off(groups.size() + 1, 0), c_off(groups.size() + 1, 0), r_off(groups.size() + 1, 0),
      k_off(groups.size() + 1, 0), f_off(groups.size() + 1, 0);
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const Group& g = groups[gi];
    const ViewSet& S = sets[g.set];
    const int cap = S.ra * S.rb;
    size_t nc_ = 0, nr = 0;
    for (int vi : g.v) {
      nc_ += S.views[vi].combos.size();
      nr += S.views[vi].prims.size();
    }
    v_off[gi + 1] = v_off[gi] + g.v.size();
    c_off[gi + 1] = c_off[gi] + nc_;
    r_off[gi + 1] = r_off[gi] + nr;
    k_off[gi + 1] = k_off[gi] + (size_t)g.npp * g.v.size();
    f_off[gi + 1] = f_off[gi] + (size_t)g.npp * cap * g.v.size();
  }
  const size_t ktot = k_off[groups.size()], ctot = f_off[groups.size()], nbrec = r_off[groups.size()];
  std::vector<ViewDev> vd(v_off[groups.size()]);
  std::vector<ComboDev> cd(c_off[groups.size()]);
  std::vector<double> brec(nbrec * HFG_BRA_REC, 0.0);  // bra records, HFG_BRA_REC doubles per primitive pair
  std::vector<double> kt_cols[9];
  std::vector<double> kcoef;
  const std::vector<double>* ucol[9] = {&up.p, &up.Px, &up.Py, &up.Pz, &up.PAx, &up.PAy, &up.PAz, &up.K, &up.H};
  for (auto& c : kt_cols) c.assign(std::max<size_t>(1, ktot), 0.0);
  kcoef.assign(std::max<size_t>(1, ctot), 0.0);
  plan.view_q.assign(vd.size(), 0.0);
  plan.prim_pairs_kept = nbrec;
  for (const Group& g : groups) {
    const ViewSet& S = sets[g.set];
    const int cap = S.ra * S.rb, coff = S.pc == 0 ? HFG_BRA_COEF_SS : HFG_BRA_COEF;
    if (cap > 1 && coff + cap > HFG_BRA_REC) throw HfgError{HFG_INVALID_INPUT, "contraction cap beyond the bra record"};
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (long gi_ = 0; gi_ < (long)groups.size(); ++gi_) {
    const size_t gi = (size_t)gi_;
    const Group& g = groups[gi];
    const ViewSet& S = sets[g.set];
    const int cap = S.ra * S.rb, size = (int)g.v.size();
    const size_t kb = k_off[gi], cb = f_off[gi];
    size_t ci_next = c_off[gi], rec = r_off[gi];
    kg[gi] = KetGroup{(int)v_off[gi], size, (int)kb, g.npp, (int)cb, 0, 0, 0};
    const int coff = S.pc == 0 ? HFG_BRA_COEF_SS : HFG_BRA_COEF;
    for (int j = 0; j < size; ++j) {
      const View& v = S.views[g.v[j]];
      const GenPair& P = gp[v.pair];
      ViewDev d;
      std::memset(&d, 0, sizeof d);
      const int npv = (int)v.prims.size();  // this view's primitives; the group pads to g.npp
      d.p0 = (int)rec;
      d.p1 = d.p0 + npv;
      d.na = v.na;
      d.nb = v.nb;
      d.c0 = (int)ci_next;
      d.coef0 = 0;
      d.same = P.same ? 1 : 0;
      for (int k = 0; k < 3; ++k) d.AB[k] = P.AB[k];
      vd[v_off[gi] + j] = d;
      plan.view_q[v_off[gi] + j] = v.qmax;
      for (int ci : v.combos) {
        // a missing member repeats the view's first combination with bound 0
        const auto& mm = P.combos[ci >= 0 ? ci : v.combos[0]];
        const int sa = gs[P.A].mem[mm[0]], sb = gs[P.B].mem[mm[1]];
        cd[ci_next++] = ComboDev{shells[sa].ao, shells[sb].ao, ci >= 0 ? Q[seg_pair(sa, sb)] : 0.0};
      }
      for (int ip = npv; ip < g.npp; ++ip) {  // padding: p = 1, K = 0 (contributes exactly 0)
        kt_cols[0][kb + (size_t)ip * size + j] = 1.0;
        kt_cols[8][kb + (size_t)ip * size + j] = 0.5;
      }
      for (int ip = 0; ip < npv; ++ip) {
        const int q = v.prims[ip];
        double x[9];
        for (int c = 0; c < 9; ++c) {
          x[c] = (*ucol[c])[q];
          if (c == 7 && cap == 1) x[c] *= P.coef[v.combos[0]][q - P.q0];
          kt_cols[c][kb + (size_t)ip * size + j] = x[c];
        }
        // record: p, K, Px, Py, Pz, [PAx, PAy, PAz, 1/(2p) beyond ss], coefficients
        double* r = &brec[rec * HFG_BRA_REC];
        r[0] = x[0];
        r[1] = x[7];
        r[2] = x[1];
        r[3] = x[2];
        r[4] = x[3];
        if (S.pc != 0) {
          r[5] = x[4];
          r[6] = x[5];
          r[7] = x[6];
          r[8] = x[8];
        }
        if (cap > 1)
          for (int gg = 0; gg < cap; ++gg) {
            const double cf = v.combos[gg] >= 0 ? P.coef[v.combos[gg]][q - P.q0] : 0.0;
            r[coff + gg] = cf;
            kcoef[cb + ((size_t)ip * cap + gg) * size + j] = cf;
          }
        ++rec;
      }
    }
  }

  lap("tables");
  // ---- tasks per class: bra views of the class's bra set, ket groups of its
  // ket set; kets with qmax_bra * qmax_ket > screen form a prefix.  When bra
  // and ket sets coincide each unordered pair of views is listed once.  The
  // tasks of one job share their bra view and are contiguous; jobs are ordered
  // by estimated cost, heaviest first.
  struct JobRange {
    double cost;
    size_t a, b;  // tasks [a, b) of the class list
  };
  std::vector<std::vector<JKTask>> tl_c(HFG_NCLASS);
  std::vector<std::vector<JobRange>> jobs_cv(HFG_NCLASS);
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < HFG_NCLASS; ++c) {
    const int sb = class_sets[c][0], sk = class_sets[c][1];
    const bool same_set = sb == sk;
    const double digest = 40.0 * caps[c][0] * caps[c][1] * caps[c][2] * caps[c][3];
    std::vector<JKTask>& tl = tl_c[c];  // grouped by bra view
    std::vector<JobRange> per_bra;
    for (size_t ib = 0; ib < set_groups[sb].size(); ++ib) {
      const int gb = set_groups[sb][ib];
      const Group& B = groups[gb];
      for (size_t i = 0; i < B.v.size(); ++i) {
        const double qb = sets[sb].views[B.v[i]].qmax;
        const size_t a = tl.size();
        double cost = 0.0;
        for (size_t ik = 0; ik < set_groups[sk].size(); ++ik) {
          if (same_set && ik > ib) continue;
          const int gk = set_groups[sk][ik];
          const Group& K = groups[gk];
          size_t lo = 0, hi = K.v.size();
          while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (qb * sets[sk].views[K.v[mid]].qmax > screen)
              lo = mid + 1;
            else
              hi = mid;
          }
          size_t nk = lo;
          if (gb == gk) nk = std::min(nk, i + 1);
          for (size_t k0 = 0; k0 < nk; k0 += 64) {
            const int m = (int)std::min<size_t>(64, nk - k0);
            tl.push_back(JKTask{kg[gb].start + (int)i, kg[gk].start + (int)k0, m, gk});
            cost += (double)m * ((double)B.npp * K.npp + digest);
          }
        }
        if (tl.size() > a) per_bra.push_back(JobRange{cost, a, tl.size()});
      }
    }
    // A bra with many tasks is cut into jobs of at most jt tasks, jt chosen so
    // that the class has about 1024 jobs (four per CU and more) when it has the
    // tasks for it: enough workgroups for the whole chip, and as many tasks per
    // job as that allows (each job ends with one flush of its LDS rows).
    const size_t jt = std::max<size_t>(4, tl.size() / 1024);
    std::vector<JobRange>& jobs = jobs_cv[c];
    for (const JobRange& pb : per_bra) {
      const size_t nt = pb.b - pb.a, nj = (nt + jt - 1) / jt;
      for (size_t j = 0; j < nj; ++j) {
        const size_t a0 = j * nt / nj, a1 = (j + 1) * nt / nj;
        jobs.push_back(JobRange{pb.cost * (double)(a1 - a0) / nt, pb.a + a0, pb.a + a1});
      }
    }
    std::stable_sort(jobs.begin(), jobs.end(), [](const JobRange& x, const JobRange& y) { return x.cost > y.cost; });
  }
  std::vector<JKTask> all;
  plan.task_begin.assign(HFG_NCLASS + 1, 0);
  plan.job_begin.assign(HFG_NCLASS + 1, 0);
  plan.jobs.clear();
  size_t ntot = 0;
  for (auto& t : tl_c) ntot += t.size();
  all.reserve(ntot);
  for (int c = 0; c < HFG_NCLASS; ++c) {
    plan.task_begin[c] = all.size();
    plan.job_begin[c] = plan.jobs.size();
    for (const JobRange& jb : jobs_cv[c]) {
      plan.jobs.push_back((int)all.size());
      plan.job_cost.push_back(jb.cost);
      all.insert(all.end(), tl_c[c].begin() + jb.a, tl_c[c].begin() + jb.b);
    }
  }
  plan.job_begin[HFG_NCLASS] = plan.jobs.size();
  plan.jobs.push_back((int)all.size());
  plan.task_begin[HFG_NCLASS] = all.size();

  plan.vb_size = nbrec;
  plan.kt_size = ktot;
  if (brec.empty()) brec.assign(HFG_BRA_REC, 0.0);
  plan.vb = std::move(brec);
  plan.kt.assign(9 * std::max<size_t>(1, ktot), 0.0);
  for (int c = 0; c < 9; ++c) std::copy(kt_cols[c].begin(), kt_cols[c].begin() + ktot, plan.kt.begin() + c * ktot);
  if (vd.empty()) {
    vd.emplace_back();
    plan.view_q.push_back(0.0);
  }
  if (cd.empty()) cd.push_back(ComboDev{0, 0, 0.0});
  plan.kcoef = std::move(kcoef);
  plan.views = std::move(vd);
  plan.combos = std::move(cd);
  plan.groups = std::move(kg);
  plan.tasks = std::move(all);
  lap("tasks");
  return plan;
}

}  // namespace hfg
