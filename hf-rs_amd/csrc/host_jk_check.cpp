// host_jk_check.cpp -- CPU replay of the J/K kernels' work plan, for the CPU
// test-suite only (compared there with the oracle's contract_jk).  It builds
// the same plan (jk_plan.cpp) with primitive bounds from the CPU build of the
// block recurrences, runs the generated combination blocks (eri_gc) of every
// task and lane on the host, and digests them with the kernels' weights and
// per-combination Schwarz test.  Not part of libhfg.so; never a fallback.
#define HFG_HOST_CHECK 1
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "classes.h"
#include "eri_device.h"
#include "host_prep.h"
#include "jk_plan.h"

namespace hfg {
template <int LA, int LB, int LC, int LD, int S, int V>
HFG_DEV void eri_gc(const double* __restrict__ br, int b0, int b1, const PrimPairs& kp,
                    const double* __restrict__ kc, int k0, int nk, int ks, const double AB[3], const double CD[3],
                    const double* __restrict__ boys, double* __restrict__ I);
#include "gen/eri_gc.inc"
}  // namespace hfg

using namespace hfg;

namespace {

thread_local std::string g_err;

const std::vector<double>& boys_tab() {
  static std::vector<double> tab = boys_table();
  return tab;
}

template <int M, bool C = false>
const double* slice() {
  static std::vector<double> s = [] {
    using B = BoysSlice<C>;
    std::vector<double> v(B::size);
    for (int r = 0; r < B::rows; ++r) boys_slice_row<M, C>(boys_tab().data(), r, &v[r * B::cols]);
    return v;
  }();
  return s.data();
}

typedef void (*BlockHost)(const PrimPairs&, int, int, const PrimPairs&, int, int, const double*, const double*,
                          double*);
template <int A, int B, int C, int D>
void block_host(const PrimPairs& bp, int b0, int b1, const PrimPairs& kp, int k0, int k1, const double* AB,
                const double* CD, double* I) {
  eri_block<A, B, C, D>(bp, b0, b1, kp, k0, k1 - k0, 1, AB, CD, slice<A + B + C + D>(), I);
}
const BlockHost BLOCK_HOST[HFG_NCLASS_ALL] = {
#define X(a, b, c, d) &block_host<a, b, c, d>,
    HFG_FOR_EACH_CLASS_ALL(X)
#undef X
};

PrimPairs view_of(const PrimPairHost& h) {
  return PrimPairs{h.p.data(), h.Px.data(), h.Py.data(), h.Pz.data(), h.PAx.data(),
                   h.PAy.data(), h.PAz.data(), h.K.data(), h.H.data()};
}
PrimPairs view_cols(const std::vector<double>& flat, size_t m) {
  const double* b = flat.data();
  return PrimPairs{b, b + m, b + 2 * m, b + 3 * m, b + 4 * m, b + 5 * m, b + 6 * m, b + 7 * m, b + 8 * m};
}

struct Replay {
  const JKPlan& P;
  const double* D;
  int n;
  const std::vector<double>& scale;
  double screen;
  std::vector<double>& Jacc;
  std::vector<double>& Kacc;
};

// Every output slice of a block, so that I holds the whole block.
template <int LA, int LB, int LC, int LD, int CLS, int V, int S = 0>
void eri_gc_all(const double* br, int b0, int b1, const PrimPairs& kp, const double* kc, int k0, int nk, int ks,
                const double AB[3], const double CD[3], const double* boys, double* I) {
  eri_gc<LA, LB, LC, LD, S, V>(br, b0, b1, kp, kc, k0, nk, ks, AB, CD, boys, I);
  if constexpr (S + 1 < slice_count(CLS, LA, LC))
    eri_gc_all<LA, LB, LC, LD, CLS, V, S + 1>(br, b0, b1, kp, kc, k0, nk, ks, AB, CD, boys, I);
}

// The class's tasks evaluated with the kernels' generated recurrences of
// variant V and digested element by element.
template <int LA, int LB, int LC, int LD, int CLS, int V>
void replay_class_v(Replay& r) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NCD = NC * ND, NBLK = NA * NB * NCD;
  constexpr int CA = gc_cap(CLS, V, 0) * gc_cap(CLS, V, 1), CC = gc_cap(CLS, V, 2) * gc_cap(CLS, V, 3);
  const JKPlan& P = r.P;
  const PrimPairs kp = view_cols(P.kt, P.kt_size);
  const int n = r.n;
  std::vector<double> I(CA * CC * NBLK);
  for (size_t t = P.task_begin[CLS]; t < P.task_begin[CLS + 1]; ++t) {
    const JKTask& task = P.tasks[t];
    const ViewDev& vb = P.views[task.bra];
    const KetGroup& g = P.groups[task.grp];
    for (int lane = 0; lane < task.nket; ++lane) {
      const int ket = task.ket0 + lane;
      const ViewDev& vk = P.views[ket];
      const int j = ket - g.start;
      eri_gc_all<LA, LB, LC, LD, CLS, V>(P.vb.data(), vb.p0, vb.p1, kp, P.kcoef.data() + g.cbase + j, g.base + j,
                                      g.npp, g.size, vb.AB, vk.AB,
                                      slice<LA + LB + LC + LD, boys_compact(V, LA + LB + LC + LD)>(), I.data());
      double w = 1.0;
      if (vb.same) w *= 0.5;
      if (vk.same) w *= 0.5;
      if (task.bra == ket) w *= 0.5;
      for (int x = 0; x < CA; ++x)
        for (int y = 0; y < CC; ++y) {  // padded combinations carry bound 0
          const ComboDev& cb = P.combos[vb.c0 + x];
          const ComboDev& ck = P.combos[vk.c0 + y];
          if (!(cb.q * ck.q > r.screen)) continue;
          const int A0 = cb.ao_a, B0 = cb.ao_b, C0 = ck.ao_a, D0 = ck.ao_b;
          const double* blk = &I[(x * CC + y) * NBLK];
          for (int i = 0; i < NA; ++i)
            for (int jj = 0; jj < NB; ++jj)
              for (int k = 0; k < NC; ++k)
                for (int l = 0; l < ND; ++l) {
                  const double v = w * blk[((i * NB + jj) * NC + k) * ND + l] * r.scale[A0 + i] * r.scale[B0 + jj] *
                                   r.scale[C0 + k] * r.scale[D0 + l];
                  const int a = A0 + i, b = B0 + jj, c = C0 + k, d = D0 + l;
                  r.Jacc[(size_t)a * n + b] += 2.0 * r.D[(size_t)c * n + d] * v;
                  r.Jacc[(size_t)c * n + d] += 2.0 * r.D[(size_t)a * n + b] * v;
                  r.Kacc[(size_t)a * n + c] += r.D[(size_t)b * n + d] * v;
                  r.Kacc[(size_t)b * n + c] += r.D[(size_t)a * n + d] * v;
                  r.Kacc[(size_t)a * n + d] += r.D[(size_t)b * n + c] * v;
                  r.Kacc[(size_t)b * n + d] += r.D[(size_t)a * n + c] * v;
                }
        }
    }
  }
}

// The class in the variant its plan chose.
template <int LA, int LB, int LC, int LD, int CLS>
void replay_class(Replay& r) {
  if constexpr (((UNIT_VARIANT_MASK >> CLS) & 1u) != 0)
    if ((r.P.unit_mask >> CLS) & 1u) return replay_class_v<LA, LB, LC, LD, CLS, 1>(r);
  if ((r.P.unit_mask >> CLS) & 1u) throw HfgError{HFG_UNSUPPORTED, "no one-member variant of this class"};
  replay_class_v<LA, LB, LC, LD, CLS, 0>(r);
}

}  // namespace

// J and K of density D (n x n, row-major) through the J/K plan, on the CPU.
// Outputs the plan's size figures in stats (if not null): generally
// contracted shells, views, primitive pairs kept, tasks, and the mask of the
// classes planned with one member per shell.
extern "C" int hfgc_jk(const hfg_basis_view* b, const double* D, double* J, double* K, long long* stats) {
  try {
    std::vector<Shell> sh = group_shells(*b);
    for (auto& s : sh)
      if (s.L > 2) throw HfgError{HFG_UNSUPPORTED, "shell-quartet classes are compiled up to d functions (L<=2)"};
    const int n = b->n_functions;
    std::vector<PairInfo> pairs;
    PrimPairHost pp;
    build_pairs(sh, pairs, pp);
    std::vector<double> scale(n, 1.0);
    for (auto& s : sh)
      for (size_t c = 0; c < s.scale.size(); ++c) scale[s.ao + c] = s.scale[c];
    // Schwarz bounds (src/shell_integrals.rs:165-173)
    const PrimPairs pv = view_of(pp);
    std::vector<double> Q(pairs.size()), blk(1296);
    for (size_t i = 0; i < pairs.size(); ++i) {
      const PairInfo& p = pairs[i];
      const int pc = pair_class(p.la, p.lb);
      BLOCK_HOST[quartet_class(pc, pc)](pv, p.p0, p.p1, pv, p.p0, p.p1, p.AB, p.AB, blk.data());
      const int na = ncart(p.la), nb = ncart(p.lb), nab = na * nb;
      double m = 0.0;
      for (int x = 0; x < na; ++x)
        for (int y = 0; y < nb; ++y) {
          const int ij = x * nb + y;
          const double f = scale[p.ao_a + x] * scale[p.ao_b + y];
          m = std::max(m, blk[(size_t)ij * nab + ij] * f * f);
        }
      Q[p.hf_index] = std::sqrt(m);
    }
    BlockFn host_blocks = [](const PrimPairHost& prims, const std::vector<PairDev>& tab,
                             const std::vector<BlockTask>& tasks, const std::vector<int>& cls,
                             std::vector<double>& out) {
      const PrimPairs v = view_of(prims);
      for (size_t t = 0; t < tasks.size(); ++t) {
        const PairDev& pb = tab[tasks[t].bra];
        const PairDev& pk = tab[tasks[t].ket];
        BLOCK_HOST[cls[t]](v, pb.p0, pb.p1, v, pk.p0, pk.p1, pb.AB, pk.AB, &out[tasks[t].off]);
      }
    };
    const double screen = 1e-12;
    JKPlan plan = make_jk_plan(sh, Q, scale, screen, 1e-20, host_blocks);
    std::vector<double> Jacc((size_t)n * n, 0.0), Kacc((size_t)n * n, 0.0);
    Replay r{plan, D, n, scale, screen, Jacc, Kacc};
    {
      // class indices in HFG_FOR_EACH_CLASS order
      replay_class<0, 0, 0, 0, 0>(r);
      replay_class<1, 0, 0, 0, 1>(r);
      replay_class<1, 0, 1, 0, 2>(r);
      replay_class<1, 1, 0, 0, 3>(r);
      replay_class<1, 1, 1, 0, 4>(r);
      replay_class<1, 1, 1, 1, 5>(r);
      replay_class<2, 0, 0, 0, 6>(r);
      replay_class<2, 0, 1, 0, 7>(r);
      replay_class<2, 0, 1, 1, 8>(r);
      replay_class<2, 0, 2, 0, 9>(r);
      replay_class<2, 1, 0, 0, 10>(r);
      replay_class<2, 1, 1, 0, 11>(r);
      replay_class<2, 1, 1, 1, 12>(r);
      replay_class<2, 1, 2, 0, 13>(r);
      replay_class<2, 1, 2, 1, 14>(r);
      replay_class<2, 2, 0, 0, 15>(r);
      replay_class<2, 2, 1, 0, 16>(r);
      replay_class<2, 2, 1, 1, 17>(r);
      replay_class<2, 2, 2, 0, 18>(r);
      replay_class<2, 2, 2, 1, 19>(r);
      replay_class<2, 2, 2, 2, 20>(r);
    }
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        J[(size_t)i * n + j] = Jacc[(size_t)i * n + j] + Jacc[(size_t)j * n + i];
        K[(size_t)i * n + j] = Kacc[(size_t)i * n + j] + Kacc[(size_t)j * n + i];
      }
    if (stats) {
      stats[0] = plan.n_gen_shells;
      stats[1] = (long long)plan.views.size();
      stats[2] = (long long)plan.prim_pairs_kept;
      stats[3] = (long long)plan.tasks.size();
      stats[4] = (long long)plan.unit_mask;
    }
    return 0;
  } catch (HfgError& e) {
    g_err = e.msg;
    return e.code;
  }
}

extern "C" const char* hfgc_jk_last_error() { return g_err.c_str(); }

// Work of the plan built with the given caps (21 x 4 ints, or null for the
// compiled ones), without evaluating it: per class, out[4c] = tasks,
// out[4c+1] = primitive quartets of the active lanes, out[4c+2] = combination
// blocks of the active lanes, out[4c+3] = J/K atomics.
extern "C" int hfgc_jk_plan_work2(const hfg_basis_view* b, const int* caps, double* out, double* wave_out);
extern "C" int hfgc_jk_plan_work(const hfg_basis_view* b, const int* caps, double* out) {
  return hfgc_jk_plan_work2(b, caps, out, nullptr);
}
// As hfgc_jk_plan_work; wave_out[c] (if not null): primitive-quartet
// iterations of the waves (every lane of a task runs its kets' count).
extern "C" int hfgc_jk_plan_work2(const hfg_basis_view* b, const int* caps, double* out, double* wave_out) {
  try {
    std::vector<Shell> sh = group_shells(*b);
    const int n = b->n_functions;
    std::vector<PairInfo> pairs;
    PrimPairHost pp;
    build_pairs(sh, pairs, pp);
    std::vector<double> scale(n, 1.0);
    for (auto& s : sh)
      for (size_t c = 0; c < s.scale.size(); ++c) scale[s.ao + c] = s.scale[c];
    const PrimPairs pv = view_of(pp);
    std::vector<double> Q(pairs.size()), blk(1296);
    for (size_t i = 0; i < pairs.size(); ++i) {
      const PairInfo& p = pairs[i];
      const int pc = pair_class(p.la, p.lb);
      BLOCK_HOST[quartet_class(pc, pc)](pv, p.p0, p.p1, pv, p.p0, p.p1, p.AB, p.AB, blk.data());
      const int na = ncart(p.la), nb = ncart(p.lb), nab = na * nb;
      double m = 0.0;
      for (int x = 0; x < na; ++x)
        for (int y = 0; y < nb; ++y) {
          const int ij = x * nb + y;
          const double f = scale[p.ao_a + x] * scale[p.ao_b + y];
          m = std::max(m, blk[(size_t)ij * nab + ij] * f * f);
        }
      Q[p.hf_index] = std::sqrt(m);
    }
    BlockFn host_blocks = [](const PrimPairHost& prims, const std::vector<PairDev>& tab,
                             const std::vector<BlockTask>& tasks, const std::vector<int>& cls,
                             std::vector<double>& o) {
      const PrimPairs v = view_of(prims);
      for (size_t t = 0; t < tasks.size(); ++t) {
        const PairDev& pb = tab[tasks[t].bra];
        const PairDev& pk = tab[tasks[t].ket];
        BLOCK_HOST[cls[t]](v, pb.p0, pb.p1, v, pk.p0, pk.p1, pb.AB, pk.AB, &o[tasks[t].off]);
      }
    };
    JKPlan plan = make_jk_plan(sh, Q, scale, 1e-12, 1e-20, host_blocks, (const int(*)[4])caps);
    for (int c = 0; c < HFG_NCLASS; ++c) {
      static const int LC[HFG_NCLASS][4] = {
#define X(a_, b_, c_, d_) {a_, b_, c_, d_},
          HFG_FOR_EACH_CLASS(X)
#undef X
      };
      const int NA = ncart(LC[c][0]), NB = ncart(LC[c][1]), NC = ncart(LC[c][2]), ND = ncart(LC[c][3]);
      double nt = 0, pq = 0, nb = 0, at = 0, wave = 0;
      for (size_t t = plan.task_begin[c]; t < plan.task_begin[c + 1]; ++t) {
        const JKTask& task = plan.tasks[t];
        const ViewDev& vb = plan.views[task.bra];
        const KetGroup& g = plan.groups[task.grp];
        nt += 1;
        pq += (double)(vb.p1 - vb.p0) * g.npp * task.nket;
        wave += (double)(vb.p1 - vb.p0) * g.npp;
        at += (double)vb.na * vb.nb * NA * NB;  // bra Coulomb: one per wave
        for (int l = 0; l < task.nket; ++l) {
          const ViewDev& vk = plan.views[task.ket0 + l];
          nb += (double)vb.na * vb.nb * vk.na * vk.nb;
          at += (double)vk.na * vk.nb * NC * ND + (double)vb.na * vk.na * NA * NC + (double)vb.nb * vk.na * NB * NC +
                (double)vb.na * vk.nb * NA * ND + (double)vb.nb * vk.nb * NB * ND;
        }
      }
      std::vector<int> bras;
      for (size_t t = plan.task_begin[c]; t < plan.task_begin[c + 1]; ++t) bras.push_back(plan.tasks[t].bra);
      std::sort(bras.begin(), bras.end());
      const double njob = (double)(std::unique(bras.begin(), bras.end()) - bras.begin());
      out[4 * c] = nt + 1e-6 * njob;  // tasks, and distinct bra views in the fraction
      out[4 * c + 1] = pq;
      out[4 * c + 2] = nb;
      out[4 * c + 3] = at;
      if (wave_out) wave_out[c] = wave;
    }
    return 0;
  } catch (HfgError& e) {
    g_err = e.msg;
    return e.code;
  }
}

// Tasks, jobs and active lanes (ket views summed over tasks) of the plan per
// class (out[3c], out[3c + 1], out[3c + 2]): the work granularity of the J/K
// kernels (a task is one wave of 64 lanes; a job, one workgroup's unit).
extern "C" int hfgc_jk_plan_jobs(const hfg_basis_view* b, double* out) {
  try {
    std::vector<Shell> sh = group_shells(*b);
    const int n = b->n_functions;
    std::vector<PairInfo> pairs;
    PrimPairHost pp;
    build_pairs(sh, pairs, pp);
    std::vector<double> scale(n, 1.0);
    for (auto& s : sh)
      for (size_t c = 0; c < s.scale.size(); ++c) scale[s.ao + c] = s.scale[c];
    const PrimPairs pv = view_of(pp);
    std::vector<double> Q(pairs.size()), blk(1296);
    for (size_t i = 0; i < pairs.size(); ++i) {
      const PairInfo& p = pairs[i];
      const int pc = pair_class(p.la, p.lb);
      BLOCK_HOST[quartet_class(pc, pc)](pv, p.p0, p.p1, pv, p.p0, p.p1, p.AB, p.AB, blk.data());
      const int na = ncart(p.la), nb = ncart(p.lb), nab = na * nb;
      double m = 0.0;
      for (int x = 0; x < na; ++x)
        for (int y = 0; y < nb; ++y) {
          const int ij = x * nb + y;
          const double f = scale[p.ao_a + x] * scale[p.ao_b + y];
          m = std::max(m, blk[(size_t)ij * nab + ij] * f * f);
        }
      Q[p.hf_index] = std::sqrt(m);
    }
    BlockFn host_blocks = [](const PrimPairHost& prims, const std::vector<PairDev>& tab,
                             const std::vector<BlockTask>& tasks, const std::vector<int>& cls,
                             std::vector<double>& o) {
      const PrimPairs v = view_of(prims);
      for (size_t t = 0; t < tasks.size(); ++t) {
        const PairDev& pb = tab[tasks[t].bra];
        const PairDev& pk = tab[tasks[t].ket];
        BLOCK_HOST[cls[t]](v, pb.p0, pb.p1, v, pk.p0, pk.p1, pb.AB, pk.AB, &o[tasks[t].off]);
      }
    };
    JKPlan plan = make_jk_plan(sh, Q, scale, 1e-12, 1e-20, host_blocks);
    for (int c = 0; c < HFG_NCLASS; ++c) {
      out[3 * c] = (double)(plan.task_begin[c + 1] - plan.task_begin[c]);
      out[3 * c + 1] = (double)(plan.job_begin[c + 1] - plan.job_begin[c]);
      double lanes = 0;
      for (size_t t = plan.task_begin[c]; t < plan.task_begin[c + 1]; ++t) lanes += plan.tasks[t].nket;
      out[3 * c + 2] = lanes;
    }
    return 0;
  } catch (std::exception& e) {
    g_err = e.what();
    return 1;
  }
}
