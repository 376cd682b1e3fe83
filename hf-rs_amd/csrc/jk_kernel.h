// jk_kernel.h -- direct J/K digestion of screened shell quartets on gfx950.
//
// Work decomposition: one wavefront per task = one bra view (uniform across
// the wave, so its primitive data and contraction coefficients are read
// through scalar loads) and up to 64 ket views (one per lane) of one ket
// group, drawn from a Schwarz-sorted list so that the surviving kets of a bra
// are a prefix.  A view is a pair of generally contracted shells restricted to
// at most RA x RB (bra) / RC x RD (ket) contraction members that share their
// primitive pairs (see jk_plan.cpp).  Each lane evaluates the CA x CC blocks
// of its quartet with the generated Head-Gordon-Pople recurrences (eri_gc),
// then digests every combination whose hf-rs quartet passes the Schwarz test
// into J and K:
//   Jacc[ij] += 2w sum_kl D_kl (ij|kl)   (wave-reduced, one atomic per ij)
//   Jacc[kl] += 2w sum_ij D_ij (ij|kl)
//   Kacc[ik] += w sum_jl D_jl v, Kacc[jk] += w sum_il D_il v,
//   Kacc[il] += w sum_jk D_jk v, Kacc[jl] += w sum_ik D_ik v
// then J = Jacc + Jacc^T and K = Kacc + Kacc^T.  This equals contract_jk
// (src/scf.rs:337-364) applied to the packed tensor of
// compute_shell_eri_tensor (src/shell_integrals.rs:156-240) without ever
// storing it.
#pragma once
#include <algorithm>

#include "engine.h"
#include "eri_device.h"

namespace hfg {

// Slice S of the blocks of CA x CC contraction combinations, caps of variant
// V (generated, see gen_eri.py).
template <int LA, int LB, int LC, int LD, int S, int V>
HFG_DEV void eri_gc(const double* __restrict__ br, int b0, int b1, const PrimPairs& kp,
                    const double* __restrict__ kc, int k0, int nk, int ks, const double AB[3], const double CD[3],
                    const double* __restrict__ boys, double* __restrict__ I);
#include "gen/eri_gc.inc"

template <int LA, int LB, int LC, int LD>
constexpr int class_index() {
  int c = 0, found = -1;
#define X(a, b, cc, d) \
  if (a == LA && b == LB && cc == LC && d == LD) found = c; \
  ++c;
  HFG_FOR_EACH_CLASS(X)
#undef X
  return found;
}

// Output slices of a class: the Cartesian block of a contracted quartet is
// split by its a component (mode 1), its c component (mode 2) or both (mode
// 3), and each slice is evaluated and digested by its own wave with only the
// recurrence terms its outputs need (high-L classes, whose whole blocks do
// not fit the registers of a lane).  A task of such a class is run once per
// slice; the partial sums of the slices add up to the quartet's J and K.
template <int LA, int LB, int LC, int LD>
constexpr int slice_count() {
  return hfg::slice_count(class_index<LA, LB, LC, LD>(), LA, LC);
}
// Whether slice S holds the outputs with a component i and c component k.
template <int LA, int LB, int LC, int LD, int S>
constexpr bool in_slice(int i, int k) {
  constexpr int m = SLICE_MODE[class_index<LA, LB, LC, LD>()];
  return m == 0 || (m == 1 && i == S) || (m == 2 && k == S) || (m == 3 && i * ncart(LC) + k == S);
}
template <int LA, int LB, int LC, int LD, int S>
constexpr bool slice_has_a(int i) {
  for (int k = 0; k < ncart(LC); ++k)
    if (in_slice<LA, LB, LC, LD, S>(i, k)) return true;
  return false;
}
template <int LA, int LB, int LC, int LD, int S>
constexpr bool slice_has_c(int k) {
  for (int i = 0; i < ncart(LA); ++i)
    if (in_slice<LA, LB, LC, LD, S>(i, k)) return true;
  return false;
}

// Waves per SIMD the register allocator must leave room for in k_jk_rows:
// two for (ss|ss), (ps|ss), (ps|ps) and (pp|ss), whose few spills cost less
// than the latency a second wave hides; one (no constraint) elsewhere.
// Measured on (H2O)16/cc-pVDZ: three waves for (ss|ss) and (ps|ss) spill
// far more and run 1.5-3.7x slower; two for (pp|ps) cost 15%.
template <int LA, int LB, int LC, int LD>
constexpr int jk_min_waves() {
  constexpr int c = class_index<LA, LB, LC, LD>();
  return (c == 0 || c == 1 || c == 2 || c == 3) ? 2 : 1;
}

// J/K partial sums live in one copy per XCC (accelerator complex die, each
// with its own L2).  A wave never migrates, so every update to copy x comes
// from XCC x and a workgroup-scope atomic -- performed in that XCC's L2 --
// is enough; no update has to travel to another die's cache.  The copies are
// summed by k_symmetrize after the kernel boundary.  The XCC id is the 4-bit
// hardware register HIP's __smid also reads; NXCC_SLOTS covers all 16 values.
__device__ __forceinline__ unsigned xcc_id() {
  return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (NXCC_SLOTS - 1);
}
__device__ __forceinline__ void add_local(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_add(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// One molecule's addresses for a job or task, read from its table once into
// registers (wave-uniform): the plan tables, the density, this XCC's
// accumulator copies, AO factors, size and the build's Schwarz threshold.
// Held by value, the loads are not repeated after the kernel's atomics.
struct MolView {
  const JKTask* tasks;
  const ViewDev* views;
  const ComboDev* combos;
  const KetGroup* groups;
  const double* brec;
  PrimPairs kt;
  const double* kcoef;
  const double* vq;
  const double* scale;
  const double* D;
  double* Jx;
  double* Kx;
  const int* jobs;
  double screen;
  int n;
};
// The table is read through the constant address space: its address is
// wave-uniform and it does not change during a launch, so every field comes
// in by scalar loads and stays in scalar registers.  (Through the global
// address space the kernel's atomics would keep the compiler from proving
// that, and the bra records and factors that follow from these pointers
// would be read by vector loads.)
typedef const MolTab __attribute__((address_space(4)))* ctab;
__device__ __forceinline__ MolView mol_view(const MolTab* T_) {
  const ctab T = (ctab)T_;
  MolView m;
  m.tasks = T->tasks;
  m.views = T->views;
  m.combos = T->combos;
  m.groups = T->groups;
  m.brec = T->brec;
  const double* kb = T->kt;
  const size_t ks = T->kt_size;
  m.kt = PrimPairs{kb, kb + ks, kb + 2 * ks, kb + 3 * ks, kb + 4 * ks, kb + 5 * ks, kb + 6 * ks, kb + 7 * ks, kb + 8 * ks};
  m.kcoef = T->kcoef;
  m.vq = T->vq;
  m.scale = T->scale;
  m.D = T->D;
  const size_t slot = (size_t)xcc_id() * T->nn;
  m.Jx = T->Jacc + slot;
  m.Kx = T->Kacc + slot;
  m.jobs = T->jobs;
  m.screen = T->screen;
  m.n = T->n;
  return m;
}
// A field of a molecule's table (constant address space, as above).
__device__ __forceinline__ int tab_task_begin(const MolTab* T, int c) { return ((ctab)T)->task_begin[c]; }
__device__ __forceinline__ int tab_job_begin(const MolTab* T, int c) { return ((ctab)T)->job_begin[c]; }
// Directory arrays and job offsets do not change during a launch either.
typedef const int __attribute__((address_space(4)))* ciptr;

// Entry of a directory prefix array (nd + 1 ascending offsets) that holds
// item x: the last e with pre[e] <= x.  Wave-uniform.
__device__ __forceinline__ int dir_entry(const int* __restrict__ pre_, int nd, int x) {
  const ciptr pre = (ciptr)pre_;
  int lo = 0, hi = nd - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= x)
      lo = mid;
    else
      hi = mid - 1;
  }
  return __builtin_amdgcn_readfirstlane(lo);
}

// One task: the bra view (wave-uniform) against up to 64 ket views.  The
// CA x CC blocks of a lane are scaled (weight, component factors, and 0 for a
// combination whose hf-rs quartet fails the Schwarz test or whose member is
// missing), then digested with every sum over the members that do not address
// the output taken in registers first: K[a_i, c_k] sums over the b and d
// members, K[b_j, c_k] over a and d, K[a_i, d_l] over b and c, K[b_j, d_l]
// over a and c, J[c_k, d_l] over a and b, and J[a_i, b_j] over c, d and the
// 64 lanes (one wave reduction).  One atomic per output element and member
// tuple remains.
// ROWS: rows is the workgroup's LDS accumulator of the bra's K rows and bra
// Coulomb block (see k_jk_rows); otherwise J and K are updated in global
// memory only.
// V: the kernel variant (caps gc_cap(CLS, V, .)).
template <int LA, int LB, int LC, int LD, bool ROWS, int S, int V>
__device__ __forceinline__ void jk_task(const MolView& a, double screen0, const double* __restrict__ boys, int task,
                                        int lane, double* __restrict__ rows) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND, NBLK = NAB * NCD;
  constexpr int CLS = class_index<LA, LB, LC, LD>();
  constexpr int RA = gc_cap(CLS, V, 0), RB = gc_cap(CLS, V, 1), RC = gc_cap(CLS, V, 2), RD = gc_cap(CLS, V, 3);
  constexpr int CA = RA * RB, CC = RC * RD;
  constexpr int NR = RA * NA + RB * NB;  // K rows of the bra
  const JKTask t = a.tasks[task];
  const int bra = __builtin_amdgcn_readfirstlane(t.bra);
  const int ket0 = __builtin_amdgcn_readfirstlane(t.ket0);
  int nket = __builtin_amdgcn_readfirstlane(t.nket);
  const int grp = __builtin_amdgcn_readfirstlane(t.grp);
  if (a.screen > screen0) {
    // A raised threshold (a density change): the task's kets, sorted by
    // bound, keep the prefix with qmax_bra * qmax_ket > screen.
    const cdptr vq = (cdptr)a.vq;
    const double qb = vq[bra];
    int lo = 0, hi = nket;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (qb * vq[ket0 + mid] > a.screen)
        lo = mid + 1;
      else
        hi = mid;
    }
    nket = lo;
    if (nket == 0) return;  // wave-uniform
  }
  const bool act = lane < nket;
  const int ket = ket0 + (act ? lane : 0);
  const ViewDev& vb = a.views[bra];
  const ViewDev& vk = a.views[ket];
  const KetGroup g = a.groups[grp];
  const int j0 = ket - g.start;
  double I[CA * CC * NBLK];
  eri_gc<LA, LB, LC, LD, S, V>(a.brec, vb.p0, vb.p1, a.kt, a.kcoef + g.cbase + j0, g.base + j0,
                         g.npp, g.size, vb.AB, vk.AB, boys, I);  // idle lanes run a valid pair, masked below

  double w = 1.0;
  if (vb.same) w *= 0.5;
  if (vk.same) w *= 0.5;
  if (bra == ket) w *= 0.5;
  const int na = __builtin_amdgcn_readfirstlane(vb.na), nb = __builtin_amdgcn_readfirstlane(vb.nb);
  const int nc = vk.na, nd = vk.nb;
  const ComboDev* __restrict__ cbra = a.combos + vb.c0;
  const ComboDev* __restrict__ cket = a.combos + vk.c0;
  int Ao[RA], Bo[RB], Co[RC], Do[RD];
#pragma unroll
  for (int x = 0; x < RA; ++x) Ao[x] = __builtin_amdgcn_readfirstlane(cbra[x * RB].ao_a);
#pragma unroll
  for (int y = 0; y < RB; ++y) Bo[y] = __builtin_amdgcn_readfirstlane(cbra[y].ao_b);
#pragma unroll
  for (int z = 0; z < RC; ++z) Co[z] = cket[z * RD].ao_a;
#pragma unroll
  for (int u = 0; u < RD; ++u) Do[u] = cket[u].ao_b;
  const cdptr sc_ = (cdptr)a.scale;  // read-only in the kernel: loads survive the atomics
  // scaling
#pragma unroll
  for (int xy = 0; xy < CA; ++xy) {
    const double qb = cbra[xy].q;
    const int x = xy / RB, y = xy % RB;
#pragma unroll
    for (int zu = 0; zu < CC; ++zu) {
      const int z = zu / RD, u = zu % RD;
      const double f = (act && qb * cket[zu].q > a.screen) ? w : 0.0;
      double* blk = I + (xy * CC + zu) * NBLK;
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int jj = 0; jj < NB; ++jj) {
          const double fab = f * sc_[Ao[x] + i] * sc_[Bo[y] + jj];
#pragma unroll
          for (int k = 0; k < NC; ++k) {
            if (!in_slice<LA, LB, LC, LD, S>(i, k)) continue;
#pragma unroll
            for (int l = 0; l < ND; ++l) blk[((i * NB + jj) * NC + k) * ND + l] *= fab * sc_[Co[z] + k] * sc_[Do[u] + l];
          }
        }
    }
  }

  const cdptr Dm = (cdptr)a.D;  // read-only in the kernel: loads survive the atomics
  const int n = a.n;
  double* __restrict__ Jx = a.Jx;
  double* __restrict__ Kx = a.Kx;
  // Each term is its own pass over the blocks, one output element at a time,
  // so that only the blocks, one density block and one sum are live.
  // Coulomb, bra side: J[a_i, b_j] over c, d and the wave (all lanes share the bra).
#pragma unroll
  for (int x = 0; x < RA; ++x)
#pragma unroll
    for (int y = 0; y < RB; ++y) {
      if (x >= na || y >= nb) continue;  // wave-uniform
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int jj = 0; jj < NB; ++jj) {
          if (!slice_has_a<LA, LB, LC, LD, S>(i)) continue;
          double s = 0.0;
#pragma unroll
          for (int z = 0; z < RC; ++z)
#pragma unroll
            for (int u = 0; u < RD; ++u) {
              const double* blk = I + ((x * RB + y) * CC + z * RD + u) * NBLK + (i * NB + jj) * NCD;
#pragma unroll
              for (int k = 0; k < NC; ++k) {
                if (!in_slice<LA, LB, LC, LD, S>(i, k)) continue;
#pragma unroll
                for (int l = 0; l < ND; ++l) s += Dm[(size_t)(Co[z] + k) * n + Do[u] + l] * blk[k * ND + l];
              }
            }
          s = wave_sum(s);
          if (lane == 0) {
            if constexpr (ROWS)
              lds_add(&rows[NR * n + (x * RB + y) * NAB + i * NB + jj], 2.0 * s);
            else
              add_local(&Jx[(size_t)(Ao[x] + i) * n + Bo[y] + jj], 2.0 * s);
          }
        }
    }
  if (!act) return;
  // Coulomb, ket side: J[c_k, d_l] over a and b.
#pragma unroll
  for (int z = 0; z < RC; ++z)
#pragma unroll
    for (int u = 0; u < RD; ++u) {
      if (z >= nc || u >= nd) continue;
#pragma unroll
      for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int l = 0; l < ND; ++l) {
          if (!slice_has_c<LA, LB, LC, LD, S>(k)) continue;
          double s = 0.0;
#pragma unroll
          for (int x = 0; x < RA; ++x)
#pragma unroll
            for (int y = 0; y < RB; ++y) {
              const double* blk = I + ((x * RB + y) * CC + z * RD + u) * NBLK + k * ND + l;
#pragma unroll
              for (int i = 0; i < NA; ++i) {
                if (!in_slice<LA, LB, LC, LD, S>(i, k)) continue;
#pragma unroll
                for (int jj = 0; jj < NB; ++jj) s += Dm[(size_t)(Ao[x] + i) * n + Bo[y] + jj] * blk[(i * NB + jj) * NCD];
              }
            }
          add_local(&Jx[(size_t)(Co[z] + k) * n + Do[u] + l], 2.0 * s);
        }
    }
  // Exchange: K[a_i, c_k] over b, d.
#pragma unroll
  for (int x = 0; x < RA; ++x)
#pragma unroll
    for (int z = 0; z < RC; ++z) {
      if (x >= na || z >= nc) continue;
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          if (!in_slice<LA, LB, LC, LD, S>(i, k)) continue;
          double s = 0.0;
#pragma unroll
          for (int y = 0; y < RB; ++y)
#pragma unroll
            for (int u = 0; u < RD; ++u) {
              const double* blk = I + ((x * RB + y) * CC + z * RD + u) * NBLK;
#pragma unroll
              for (int jj = 0; jj < NB; ++jj)
#pragma unroll
                for (int l = 0; l < ND; ++l)
                  s += Dm[(size_t)(Bo[y] + jj) * n + Do[u] + l] * blk[((i * NB + jj) * NC + k) * ND + l];
            }
          if constexpr (ROWS)
            lds_add(&rows[(size_t)(x * NA + i) * n + Co[z] + k], s);
          else
            add_local(&Kx[(size_t)(Ao[x] + i) * n + Co[z] + k], s);
        }
    }
  // K[b_j, c_k] over a, d.
#pragma unroll
  for (int y = 0; y < RB; ++y)
#pragma unroll
    for (int z = 0; z < RC; ++z) {
      if (y >= nb || z >= nc) continue;
#pragma unroll
      for (int jj = 0; jj < NB; ++jj)
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          if (!slice_has_c<LA, LB, LC, LD, S>(k)) continue;
          double s = 0.0;
#pragma unroll
          for (int x = 0; x < RA; ++x)
#pragma unroll
            for (int u = 0; u < RD; ++u) {
              const double* blk = I + ((x * RB + y) * CC + z * RD + u) * NBLK;
#pragma unroll
              for (int i = 0; i < NA; ++i) {
                if (!in_slice<LA, LB, LC, LD, S>(i, k)) continue;
#pragma unroll
                for (int l = 0; l < ND; ++l)
                  s += Dm[(size_t)(Ao[x] + i) * n + Do[u] + l] * blk[((i * NB + jj) * NC + k) * ND + l];
              }
            }
          if constexpr (ROWS)
            lds_add(&rows[(size_t)(RA * NA + y * NB + jj) * n + Co[z] + k], s);
          else
            add_local(&Kx[(size_t)(Bo[y] + jj) * n + Co[z] + k], s);
        }
    }
  // K[a_i, d_l] over b, c.
#pragma unroll
  for (int x = 0; x < RA; ++x)
#pragma unroll
    for (int u = 0; u < RD; ++u) {
      if (x >= na || u >= nd) continue;
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int l = 0; l < ND; ++l) {
          if (!slice_has_a<LA, LB, LC, LD, S>(i)) continue;
          double s = 0.0;
#pragma unroll
          for (int y = 0; y < RB; ++y)
#pragma unroll
            for (int z = 0; z < RC; ++z) {
              const double* blk = I + ((x * RB + y) * CC + z * RD + u) * NBLK;
#pragma unroll
              for (int jj = 0; jj < NB; ++jj)
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                  if (!in_slice<LA, LB, LC, LD, S>(i, k)) continue;
                  s += Dm[(size_t)(Bo[y] + jj) * n + Co[z] + k] * blk[((i * NB + jj) * NC + k) * ND + l];
                }
            }
          if constexpr (ROWS)
            lds_add(&rows[(size_t)(x * NA + i) * n + Do[u] + l], s);
          else
            add_local(&Kx[(size_t)(Ao[x] + i) * n + Do[u] + l], s);
        }
    }
  // K[b_j, d_l] over a, c.
#pragma unroll
  for (int y = 0; y < RB; ++y)
#pragma unroll
    for (int u = 0; u < RD; ++u) {
      if (y >= nb || u >= nd) continue;
#pragma unroll
      for (int jj = 0; jj < NB; ++jj)
#pragma unroll
        for (int l = 0; l < ND; ++l) {
          double s = 0.0;
#pragma unroll
          for (int x = 0; x < RA; ++x)
#pragma unroll
            for (int z = 0; z < RC; ++z) {
              const double* blk = I + ((x * RB + y) * CC + z * RD + u) * NBLK;
#pragma unroll
              for (int i = 0; i < NA; ++i)
#pragma unroll
                for (int k = 0; k < NC; ++k) {
                  if (!in_slice<LA, LB, LC, LD, S>(i, k)) continue;
                  s += Dm[(size_t)(Ao[x] + i) * n + Co[z] + k] * blk[((i * NB + jj) * NC + k) * ND + l];
                }
            }
          if constexpr (ROWS)
            lds_add(&rows[(size_t)(RA * NA + y * NB + jj) * n + Do[u] + l], s);
          else
            add_local(&Kx[(size_t)(Bo[y] + jj) * n + Do[u] + l], s);
        }
    }
}

// Slice s (wave-uniform) of a task.
template <int LA, int LB, int LC, int LD, bool ROWS, int V, int S = 0>
__device__ __forceinline__ void jk_task_slice(const MolView& a, double screen0, const double* __restrict__ boys,
                                              int task, int s, int lane, double* __restrict__ rows) {
  if constexpr (S + 1 < slice_count<LA, LB, LC, LD>()) {
    if (s != S) {
      jk_task_slice<LA, LB, LC, LD, ROWS, V, S + 1>(a, screen0, boys, task, s, lane, rows);
      return;
    }
  }
  jk_task<LA, LB, LC, LD, ROWS, S, V>(a, screen0, boys, task, lane, rows);
}

// Persistent waves: every wave walks the launch's tasks (over all directory
// entries, in order; each molecule's tasks ordered by job, heaviest first)
// with a stride of the total wave count, so the LDS copy of the Boys slice is
// made once per workgroup.  J and K are updated in global memory (per-XCC
// copies).
template <int LA, int LB, int LC, int LD, int V>
__global__ __launch_bounds__(256) void k_jk(JKArgs a) {
  constexpr bool BC = boys_compact(V, LA + LB + LC + LD);
  __shared__ double boys[BoysSlice<BC>::size];
  load_boys_slice<LA + LB + LC + LD, BC>(a.boys, boys);
  constexpr int NS = slice_count<LA, LB, LC, LD>();
  constexpr int CLS = class_index<LA, LB, LC, LD>();
  const int lane = threadIdx.x & 63;
  const int waves = (int)(gridDim.x * (blockDim.x >> 6));
  for (int v = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); v < a.ntask * NS; v += waves) {
    const int gt = __builtin_amdgcn_readfirstlane(v / NS);
    const int e = dir_entry(a.dir_task, a.ndir, gt);
    const MolTab* T = a.tabs + ((ciptr)a.dir_mol)[e];
    const MolView m = mol_view(T);
    const int task = tab_task_begin(T, CLS) + gt - ((ciptr)a.dir_task)[e];
    jk_task_slice<LA, LB, LC, LD, false, V>(m, a.screen0, boys, task, __builtin_amdgcn_readfirstlane(v - gt * NS),
                                         lane, nullptr);
  }
}

// Persistent workgroups over jobs (tasks of one bra view of one molecule),
// taken from a counter in order (each molecule's heaviest first).  The four
// waves share the job's tasks and accumulate every K update -- all of them
// have a bra index as their row -- and the bra Coulomb block in LDS, NR x n
// rows plus CA x NAB; after the job the workgroup adds the nonzero entries to
// the molecule's global copies and clears them.  Only the ket Coulomb updates
// go to global memory directly.
template <int LA, int LB, int LC, int LD, int V>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(jk_min_waves<LA, LB, LC, LD>()))) void k_jk_rows(JKArgs a) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NAB = NA * NB;
  constexpr int CLS = class_index<LA, LB, LC, LD>();
  constexpr int RA = gc_cap(CLS, V, 0), RB = gc_cap(CLS, V, 1), CA = RA * RB;
  constexpr int NR = RA * NA + RB * NB;
  constexpr int NS = slice_count<LA, LB, LC, LD>();
  constexpr bool BC = boys_compact(V, LA + LB + LC + LD);
  __shared__ double boys[BoysSlice<BC>::size];
  __shared__ int job_sh;
  extern __shared__ double rows[];
  // a.nmax is the largest molecule's size: it bounds every job's rows
  for (int e = threadIdx.x; e < NR * a.nmax + CA * NAB; e += blockDim.x) rows[e] = 0.0;
  load_boys_slice<LA + LB + LC + LD, BC>(a.boys, boys);  // ends with a barrier
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  for (;;) {
    if (threadIdx.x == 0) job_sh = atomicAdd(a.job_counter, 1);
    __syncthreads();
    const int job = __builtin_amdgcn_readfirstlane(job_sh);
    __syncthreads();
    if (job >= a.njob) break;
    const int de = dir_entry(a.dir_job, a.ndir, job);
    const MolTab* T = a.tabs + ((ciptr)a.dir_mol)[de];
    const MolView m = mol_view(T);
    const int lj = tab_job_begin(T, CLS) + job - ((ciptr)a.dir_job)[de];
    const int t0 = ((ciptr)m.jobs)[lj], t1 = ((ciptr)m.jobs)[lj + 1];
    for (int v = t0 * NS + wave; v < t1 * NS; v += nwave) {
      const int task = __builtin_amdgcn_readfirstlane(v / NS);
      jk_task_slice<LA, LB, LC, LD, true, V>(m, a.screen0, boys, task, __builtin_amdgcn_readfirstlane(v - task * NS),
                                          lane, rows);
    }
    __syncthreads();
    const ViewDev& vb = m.views[m.tasks[t0].bra];
    const ComboDev* __restrict__ cbra = m.combos + vb.c0;
    const int n = m.n;
    const int nrow = NR * n + CA * NAB;
    double* __restrict__ Jx = m.Jx;
    double* __restrict__ Kx = m.Kx;
    for (int e = threadIdx.x; e < nrow; e += blockDim.x) {
      const double v = rows[e];
      if (v == 0.0) continue;
      rows[e] = 0.0;
      if (e < NR * n) {
        const int r = e / n, col = e - r * n;
        const int ao = r < RA * NA ? cbra[(r / NA) * RB].ao_a + r % NA : cbra[(r - RA * NA) / NB].ao_b + (r - RA * NA) % NB;
        add_local(&Kx[(size_t)ao * n + col], v);
      } else {
        const int f = e - NR * n, xy = f / NAB, ij = f % NAB;
        const int ao_a = cbra[(xy / RB) * RB].ao_a + ij / NB, ao_b = cbra[xy % RB].ao_b + ij % NB;
        add_local(&Jx[(size_t)ao_a * n + ao_b], v);
      }
    }
    __syncthreads();
  }
}

// LDS bytes of k_jk_rows for n basis functions, beyond the static Boys slice.
template <int LA, int LB, int LC, int LD, int V>
constexpr size_t jk_rows_bytes(int n) {
  constexpr int CLS = class_index<LA, LB, LC, LD>();
  constexpr int RA = gc_cap(CLS, V, 0), RB = gc_cap(CLS, V, 1);
  return sizeof(double) * ((size_t)(RA * ncart(LA) + RB * ncart(LB)) * n + RA * RB * ncart(LA) * ncart(LB));
}

template <int LA, int LB, int LC, int LD, int V>
void launch_jk(const JKArgs& a, int max_blocks, hipStream_t s) {
  const int waves_per_block = 4;
  // Device limits and the resident-workgroup count of this kernel for the
  // current LDS size, queried once per (device, size): the queries cost more
  // than a launch.
  static thread_local int cached_dev = -1, lds_limit = 0, cus = 0, resident = 0;
  static thread_local size_t cached_lds = 0;
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  const size_t lds = jk_rows_bytes<LA, LB, LC, LD, V>(a.nmax);
  if (dev != cached_dev) {
    HIP_CHECK(hipDeviceGetAttribute(&lds_limit, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    cached_dev = dev;
    cached_lds = 0;
  }
  if (a.use_rows && a.njob > 0 && lds + sizeof(double) * BoysSlice<boys_compact(V, LA + LB + LC + LD)>::size + 64 <= (size_t)lds_limit) {
    if (lds != cached_lds) {
      int per_cu = 0;
      HIP_CHECK(hipFuncSetAttribute((const void*)k_jk_rows<LA, LB, LC, LD, V>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_jk_rows<LA, LB, LC, LD, V>,
                                                             64 * waves_per_block, lds));
      resident = std::max(1, per_cu) * cus;
      cached_lds = lds;
    }
    const int blocks = std::max(1, std::min(std::min(a.njob, max_blocks), resident));
    hipLaunchKernelGGL((k_jk_rows<LA, LB, LC, LD, V>), dim3(blocks), dim3(64 * waves_per_block), lds, s, a);
    return;
  }
  if (a.ntask <= 0) return;
  const int blocks = std::min(max_blocks, (a.ntask + waves_per_block - 1) / waves_per_block);
  hipLaunchKernelGGL((k_jk<LA, LB, LC, LD, V>), dim3(blocks), dim3(64 * waves_per_block), 0, s, a);
}

// Launch of the class in variant v; v = 1 exists for the classes of
// UNIT_VARIANT_MASK only (the plan never asks for another).
template <int LA, int LB, int LC, int LD>
void launch_jk_v(int v, const JKArgs& a, int max_blocks, hipStream_t s) {
  constexpr int CLS = class_index<LA, LB, LC, LD>();
  if constexpr (((UNIT_VARIANT_MASK >> CLS) & 1u) != 0) {
    if (v) {
      launch_jk<LA, LB, LC, LD, 1>(a, max_blocks, s);
      return;
    }
  } else {
    if (v) throw HfgError{HFG_UNSUPPORTED, "no one-member kernel variant for J/K class " + std::to_string(CLS)};
  }
  launch_jk<LA, LB, LC, LD, 0>(a, max_blocks, s);
}

}  // namespace hfg
