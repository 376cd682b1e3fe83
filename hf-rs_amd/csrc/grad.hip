// grad.hip -- analytic RHF nuclear gradient: compute_gradient_from_density
// (src/gradient.rs:287-512) for a density P and an energy-weighted density W.
//
// Two-electron part on the device.  dE2/dX = 1/2 sum over ordered AO
// quadruples of d(mu nu|lam sig)/dX * G, G = P_mn P_ls - 1/4 (P_ml P_ns +
// P_ms P_nl); over unique shell quartets with their degeneracy.  A
// basis-function derivative is d/dA_t phi_a = 2 alpha phi_{a+1_t} - a_t
// phi_{a-1_t} (src/gradient.rs:71-113), so the derivatives of a quartet with
// respect to the centres of three of its shells come from ordinary ERI blocks
// of "raised" shells (L+1, radial coefficients times 2 alpha) and "lowered"
// shells (L-1); the fourth centre follows from translational invariance.  The
// block kernels evaluate those blocks (the fp and fd raised pairs are the extra
// classes of gen/eri_grad.inc); one kernel then contracts every block with G
// and adds the quartet's forces to per-workgroup LDS sums of the atoms' forces.
//
// One-electron part on the host (the size of a Fock-matrix set-up): the
// overlap, kinetic and nuclear-attraction derivatives of each basis function
// (src/gradient.rs:411-447), the operator derivative of the attraction at each
// nucleus (:451-490), and the nuclear repulsion (:220-243).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/hfg.h"
#include "engine.h"
#include "eri_device.h"
#include "la_device.h"
#include "rhf.h"
#include "scf_host.h"

namespace hfg {
#include "gen/eri_grad.inc"

namespace {

// One thread per block, as the engine's block kernels.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_gblock(BlockArgs a) {
  constexpr int NT = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
  __shared__ double boys[BoysSlice<false>::size];
  load_boys_slice<LA + LB + LC + LD>(a.boys, boys);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.ntask) return;
  const BlockTask bt = a.tasks[t];
  const PairDev& pb = a.pairs[bt.bra];
  const PairDev& pk = a.pairs[bt.ket];
  double I[NT];
  eri_block<LA, LB, LC, LD>(a.pp, pb.p0, pb.p1, a.pp, pk.p0, pk.p1 - pk.p0, 1, pb.AB, pk.AB, boys, I);
  double* o = a.out + bt.off;
#pragma unroll
  for (int i = 0; i < NT; ++i) o[i] = I[i];
}

template <int LA, int LB, int LC, int LD>
void launch_gblock(const BlockArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_gblock<LA, LB, LC, LD>), dim3((a.ntask + 63) / 64), dim3(64), 0, s, a);
}

// Rank of a pair class (la >= lb) in the block kernels' order: the orbital
// pairs, the auxiliary f-s and g-s pairs, then the gradient's f-p and f-d.
int pair_rank(int la, int lb) {
  static const int R[5][3] = {{0, -1, -1}, {1, 2, -1}, {3, 4, 5}, {6, 8, 9}, {7, -1, -1}};
  return (la < 0 || la > 4 || lb < 0 || lb > 2) ? -1 : R[la][lb];
}

// Launch of block class (l1 l2 | l3 l4) (kernel orientation, first pair of
// rank >= second).
void launch_any_block(int l1, int l2, int l3, int l4, const BlockArgs& a, hipStream_t s) {
  int idx = 0;
#define X(p, q, r, u)                                              \
  if (p == l1 && q == l2 && r == l3 && u == l4) {                    \
    launch_block_class(idx, a, s);                                   \
    return;                                                          \
  }                                                                  \
  ++idx;
  HFG_FOR_EACH_CLASS_ALL(X)
#undef X
#define X(p, q, r, u)                                          \
  if (p == l1 && q == l2 && r == l3 && u == l4) {              \
    launch_gblock<p, q, r, u>(a, s);                           \
    HIP_CHECK(hipGetLastError());                              \
    return;                                                    \
  }
  HFG_FOR_EACH_GRAD_CLASS(X)
#undef X
  throw HfgError{HFG_UNSUPPORTED, "no gradient block kernel for class (" + std::to_string(l1) + std::to_string(l2) +
                                      "|" + std::to_string(l3) + std::to_string(l4) + ")"};
}

// Cartesian components in hf-rs order (as the generator's cart()): exponents
// per (L, index) and index per (L, lx, ly).
struct CartTab {
  signed char e[4][10][3];
  signed char idx[4][4][4];
};
constexpr CartTab make_cart() {
  CartTab t{};
  for (int L = 0; L < 4; ++L) {
    int k = 0;
    if (L == 2) {
      const int d[6][3] = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
      for (; k < 6; ++k)
        for (int c = 0; c < 3; ++c) t.e[L][k][c] = (signed char)d[k][c];
    } else {
      for (int lx = L; lx >= 0; --lx)
        for (int ly = L - lx; ly >= 0; --ly, ++k) {
          t.e[L][k][0] = (signed char)lx;
          t.e[L][k][1] = (signed char)ly;
          t.e[L][k][2] = (signed char)(L - lx - ly);
        }
    }
    for (int i = 0; i < k; ++i) t.idx[L][t.e[L][i][0]][t.e[L][i][1]] = (signed char)i;
  }
  return t;
}
__constant__ CartTab g_cart = make_cart();

// One derivative block of a quartet: storage position p (0..3: first pair's
// first and second shell, second pair's first and second shell) holds
// original shell sh[p] (0..3 = A, B, C, D of the quartet) with angular
// momentum L[p]; position mod holds the raised (raised = 1) or lowered shell
// whose centre is differentiated.
struct GBlock {
  long long off;
  signed char sh[4], L[4];
  signed char mod, raised, pad[2];
};
// A quartet: its shells' AO offsets, angular momenta and atoms; its blocks;
// w = 1/2 x degeneracy.
struct GQuartet {
  int ao[4];
  int atom[4];
  signed char L[4];
  int nblk;
  GBlock blk[6];
  double w;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Contraction of every quartet's derivative blocks with G: one wave per
// quartet (waves stride over the list), forces summed per atom in LDS, added
// to grad (natom x 3) at the end.
__global__ __launch_bounds__(256) void k_grad_contract(const GQuartet* __restrict__ qs, int nq,
                                                       const double* __restrict__ blk, const double* __restrict__ P,
                                                       const double* __restrict__ scale, int n, int natom,
                                                       double* __restrict__ grad) {
  extern __shared__ double acc[];  // natom x 3
  for (int i = threadIdx.x; i < 3 * natom; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (blockDim.x >> 6);
  for (int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < nq; q += waves) {
    const GQuartet& Q = qs[q];
    double gA[3] = {0, 0, 0}, gB[3] = {0, 0, 0}, gC[3] = {0, 0, 0};  // d/dA, d/dB, d/dC (x, y, z)
    for (int b = 0; b < Q.nblk; ++b) {
      const GBlock B = Q.blk[b];
      const int n0 = ncart(B.L[0]), n1 = ncart(B.L[1]), n2 = ncart(B.L[2]), n3 = ncart(B.L[3]);
      const int ne = n0 * n1 * n2 * n3;
      const int X = B.sh[B.mod];                 // differentiated shell (0, 1 or 2)
      const int Lo = Q.L[X];                     // its own angular momentum
      for (int e = lane; e < ne; e += 64) {
        int c[4];
        c[3] = e % n3;
        c[2] = (e / n3) % n2;
        c[1] = (e / (n3 * n2)) % n1;
        c[0] = e / (n3 * n2 * n1);
        const double v = blk[B.off + e];
        int o[4];
        for (int p = 0; p < 4; ++p) o[B.sh[p]] = c[p];
        const signed char* ex = g_cart.e[B.L[B.mod]][c[B.mod]];
        for (int t = 0; t < 3; ++t) {
          int ee[3] = {ex[0], ex[1], ex[2]};
          double coef;
          if (B.raised) {
            if (ee[t] == 0) continue;
            --ee[t];
            coef = 1.0;
          } else {
            coef = -(double)(ee[t] + 1);
            ++ee[t];
          }
          o[X] = g_cart.idx[Lo][ee[0]][ee[1]];
          const int mu = Q.ao[0] + o[0], nu = Q.ao[1] + o[1], la = Q.ao[2] + o[2], si = Q.ao[3] + o[3];
          const double G = P[(size_t)mu * n + nu] * P[(size_t)la * n + si] -
                           0.25 * (P[(size_t)mu * n + la] * P[(size_t)nu * n + si] +
                                   P[(size_t)mu * n + si] * P[(size_t)nu * n + la]);
          const double c = coef * v * G * scale[mu] * scale[nu] * scale[la] * scale[si];
          if (X == 0)
            gA[t] += c;
          else if (X == 1)
            gB[t] += c;
          else
            gC[t] += c;
        }
      }
    }
    double s[9];
    for (int t = 0; t < 3; ++t) {
      s[t] = wave_sum(gA[t]) * Q.w;
      s[3 + t] = wave_sum(gB[t]) * Q.w;
      s[6 + t] = wave_sum(gC[t]) * Q.w;
    }
    if (lane == 0)
      for (int t = 0; t < 3; ++t) {
        atomicAdd(&acc[3 * Q.atom[0] + t], s[t]);
        atomicAdd(&acc[3 * Q.atom[1] + t], s[3 + t]);
        atomicAdd(&acc[3 * Q.atom[2] + t], s[6 + t]);
        atomicAdd(&acc[3 * Q.atom[3] + t], -(s[t] + s[3 + t] + s[6 + t]));
      }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * natom; i += blockDim.x)
    if (acc[i] != 0.0) atomicAdd(&grad[i], acc[i]);
}

Shell raised(const Shell& s) {
  Shell r = s;
  r.L = s.L + 1;
  for (int k = 0; k < s.K; ++k) r.r[k] = s.r[k] * 2.0 * s.e[k];
  return r;
}
Shell lowered(const Shell& s) {
  Shell r = s;
  r.L = s.L - 1;
  return r;
}

// The derivative pairs of one shell pair (kernel orientation A, B): raised
// and lowered A, raised and lowered B, each stored with the higher L first.
struct DPair {
  int idx[4] = {-1, -1, -1, -1};   // pair-table entry: RA, LA, RB, LB
  bool swap[4] = {false, false, false, false};  // stored as (B', A) / (B, A')
};

}  // namespace

// Nuclear gradient (natom x 3) of the engine's molecule for density P and
// energy-weighted density W (host, row-major n x n).
void rhf_gradient(Engine& e, const std::vector<std::array<double, 4>>& atoms, const double* P, const double* W,
                  double* grad) {
  const int n = e.n, na = (int)atoms.size();
  std::vector<double> g((size_t)3 * na, 0.0);
  for (const Shell& s : e.shells)
    if (s.atom < 0 || s.atom >= na) throw HfgError{HFG_INVALID_INPUT, "basis atom index outside the geometry"};
  // nuclear repulsion (src/gradient.rs:220-243)
  for (int a = 0; a < na; ++a)
    for (int b = 0; b < na; ++b) {
      if (a == b) continue;
      double d[3], r2 = 0;
      for (int k = 0; k < 3; ++k) {
        d[k] = atoms[a][1 + k] - atoms[b][1 + k];
        r2 += d[k] * d[k];
      }
      const double f = atoms[a][0] * atoms[b][0] / (r2 * std::sqrt(r2));
      for (int k = 0; k < 3; ++k) g[3 * a + k] -= f * d[k];
    }
  const std::vector<Shell>& sh = e.shells;
  const std::vector<double>& sc = e.ao_scale;
  // ---- two-electron part on the device
  hipStream_t st = e.stream;
  const size_t npair = e.pairs.size();
  // pair table: the engine's pairs, then the derivative pairs
  PrimPairHost pp = e.pp;
  std::vector<PairDev> pd(npair);
  std::vector<std::array<int, 2>> pl(npair);  // (L first, L second) per table entry
  for (size_t i = 0; i < npair; ++i) {
    const PairInfo& p = e.pairs[i];
    std::memset(&pd[i], 0, sizeof(PairDev));
    pd[i].p0 = p.p0;
    pd[i].p1 = p.p1;
    for (int k = 0; k < 3; ++k) pd[i].AB[k] = p.AB[k];
    pl[i] = {p.la, p.lb};
  }
  auto add_pair = [&](const Shell& X, const Shell& Y) {
    PairDev d;
    std::memset(&d, 0, sizeof d);
    d.p0 = (int)pp.size();
    append_pair(X, Y, pp);
    d.p1 = (int)pp.size();
    for (int k = 0; k < 3; ++k) d.AB[k] = X.c[k] - Y.c[k];
    pd.push_back(d);
    pl.push_back({X.L, Y.L});
    return (int)pd.size() - 1;
  };
  std::vector<DPair> dp(npair);
  for (size_t i = 0; i < npair; ++i) {
    const PairInfo& p = e.pairs[i];
    const Shell &A = sh[p.sh_a], &B = sh[p.sh_b];
    DPair& d = dp[i];
    d.idx[0] = add_pair(raised(A), B);  // L_A + 1 > L_B
    if (A.L > 0) {
      const Shell la = lowered(A);
      d.swap[1] = la.L < B.L;
      d.idx[1] = d.swap[1] ? add_pair(B, la) : add_pair(la, B);
    }
    const Shell rb = raised(B);
    d.swap[2] = rb.L > A.L;
    d.idx[2] = d.swap[2] ? add_pair(rb, A) : add_pair(A, rb);
    if (B.L > 0) d.idx[3] = add_pair(A, lowered(B));  // L_A >= L_B - 1
  }
  const size_t npp = pp.size();
  DevBuf d_pp, d_pd, d_P, d_scale, d_grad;
  d_pp.alloc(9 * npp * sizeof(double));
  const std::vector<double>* cols[9] = {&pp.p, &pp.Px, &pp.Py, &pp.Pz, &pp.PAx, &pp.PAy, &pp.PAz, &pp.K, &pp.H};
  for (int c = 0; c < 9; ++c)
    HIP_CHECK(hipMemcpyAsync(d_pp.as<double>() + c * npp, cols[c]->data(), npp * 8, hipMemcpyHostToDevice, st));
  d_pd.alloc(pd.size() * sizeof(PairDev));
  HIP_CHECK(hipMemcpyAsync(d_pd.p, pd.data(), pd.size() * sizeof(PairDev), hipMemcpyHostToDevice, st));
  d_P.alloc((size_t)n * n * 8);
  HIP_CHECK(hipMemcpyAsync(d_P.p, P, (size_t)n * n * 8, hipMemcpyHostToDevice, st));
  d_scale.alloc((size_t)n * 8);
  HIP_CHECK(hipMemcpyAsync(d_scale.p, sc.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
  d_grad.alloc((size_t)3 * na * 8);
  hipLaunchKernelGGL(k_zero, dim3((3 * na + 255) / 256), dim3(256), 0, st, d_grad.as<double>(), (size_t)3 * na);
  // ---- one-electron terms on the device (onee_dev.hip)
  {
    DevBuf d_W;
    d_W.alloc((size_t)n * n * 8);
    HIP_CHECK(hipMemcpyAsync(d_W.p, W, (size_t)n * n * 8, hipMemcpyHostToDevice, st));
    onee_gradient_device(sh, n, atoms, d_P.as<double>(), d_W.as<double>(), d_grad.as<double>(), st);
  }
  const double* b = d_pp.as<double>();
  const PrimPairs view{b, b + npp, b + 2 * npp, b + 3 * npp, b + 4 * npp, b + 5 * npp, b + 6 * npp, b + 7 * npp,
                       b + 8 * npp};
  // Blocks of a batch of quartets: (class, task) lists, then the contraction.
  struct Item {
    int l[4];
    BlockTask t;
  };
  std::vector<Item> items;
  std::vector<GQuartet> quartets;
  long long off = 0;
  DevBuf d_tasks, d_blk, d_q;
  size_t cap_t = 0, cap_b = 0, cap_q = 0;
  const size_t lds = (size_t)3 * na * sizeof(double);
  if (lds > 60000) throw HfgError{HFG_UNSUPPORTED, "gradient: too many atoms for the per-workgroup force sums"};
  auto flush = [&]() {
    if (quartets.empty()) return;
    std::stable_sort(items.begin(), items.end(), [](const Item& x, const Item& y) {
      return std::lexicographical_compare(x.l, x.l + 4, y.l, y.l + 4);
    });
    std::vector<BlockTask> tasks(items.size());
    for (size_t i = 0; i < items.size(); ++i) tasks[i] = items[i].t;
    if (tasks.size() > cap_t) {
      d_tasks.alloc(tasks.size() * sizeof(BlockTask));
      cap_t = tasks.size();
    }
    if ((size_t)off > cap_b) {
      d_blk.alloc((size_t)off * 8);
      cap_b = off;
    }
    if (quartets.size() > cap_q) {
      d_q.alloc(quartets.size() * sizeof(GQuartet));
      cap_q = quartets.size();
    }
    HIP_CHECK(hipMemcpyAsync(d_tasks.p, tasks.data(), tasks.size() * sizeof(BlockTask), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(d_q.p, quartets.data(), quartets.size() * sizeof(GQuartet), hipMemcpyHostToDevice, st));
    size_t a = 0;
    while (a < items.size()) {
      size_t z = a;
      while (z < items.size() && std::equal(items[z].l, items[z].l + 4, items[a].l)) ++z;
      BlockArgs args{d_tasks.as<BlockTask>() + a, (int)(z - a), d_pd.as<PairDev>(), view, e.d_boys.as<double>(),
                     d_blk.as<double>()};
      launch_any_block(items[a].l[0], items[a].l[1], items[a].l[2], items[a].l[3], args, st);
      a = z;
    }
    const int blocks = (int)std::min<size_t>(2048, (quartets.size() + 3) / 4);
    hipLaunchKernelGGL(k_grad_contract, dim3(blocks), dim3(256), lds, st, d_q.as<GQuartet>(), (int)quartets.size(),
                       d_blk.as<double>(), d_P.as<double>(), d_scale.as<double>(), n, na, d_grad.as<double>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));  // the host lists are reused
    items.clear();
    quartets.clear();
    off = 0;
  };
  const long long budget = 1ll << 24;  // block doubles per batch
  // one block of a quartet: derivative pair dpi (lengths dl) against pair other
  auto add_block = [&](GQuartet& Q, int kind, int dpi, bool dswap, int other, const int sh_d[2],
                       const int sh_o[2], bool raised_) {
    const std::array<int, 2> ld = pl[dpi], lo = pl[other];
    // storage: the pair of higher rank first
    const bool dfirst = pair_rank(ld[0], ld[1]) >= pair_rank(lo[0], lo[1]);
    GBlock& B = Q.blk[Q.nblk++];
    B.off = off;
    B.raised = raised_ ? 1 : 0;
    const int dpos0 = dfirst ? 0 : 2, opos0 = dfirst ? 2 : 0;
    // derivative pair storage order: (mod, partner) unless swapped
    B.sh[dpos0] = (signed char)(dswap ? sh_d[1] : sh_d[0]);
    B.sh[dpos0 + 1] = (signed char)(dswap ? sh_d[0] : sh_d[1]);
    B.L[dpos0] = (signed char)ld[0];
    B.L[dpos0 + 1] = (signed char)ld[1];
    B.mod = (signed char)(dswap ? dpos0 + 1 : dpos0);
    B.sh[opos0] = (signed char)sh_o[0];
    B.sh[opos0 + 1] = (signed char)sh_o[1];
    B.L[opos0] = (signed char)lo[0];
    B.L[opos0 + 1] = (signed char)lo[1];
    Item it;
    if (dfirst) {
      it.l[0] = ld[0];
      it.l[1] = ld[1];
      it.l[2] = lo[0];
      it.l[3] = lo[1];
      it.t = BlockTask{dpi, other, off};
    } else {
      it.l[0] = lo[0];
      it.l[1] = lo[1];
      it.l[2] = ld[0];
      it.l[3] = ld[1];
      it.t = BlockTask{other, dpi, off};
    }
    items.push_back(it);
    off += (long long)ncart(ld[0]) * ncart(ld[1]) * ncart(lo[0]) * ncart(lo[1]);
    (void)kind;
  };
  for (size_t ij = 0; ij < npair; ++ij)
    for (size_t kl = 0; kl <= ij; ++kl) {
      if (!(e.Q[ij] * e.Q[kl] > e.screen)) continue;
      const PairInfo& pb = e.pairs[ij];
      const PairInfo& pk = e.pairs[kl];
      GQuartet Q;
      std::memset(&Q, 0, sizeof Q);
      const int shells4[4] = {pb.sh_a, pb.sh_b, pk.sh_a, pk.sh_b};
      for (int k = 0; k < 4; ++k) {
        const Shell& s = sh[shells4[k]];
        Q.ao[k] = s.ao;
        Q.atom[k] = s.atom;
        Q.L[k] = (signed char)s.L;
      }
      double deg = 1.0;
      if (pb.sh_a != pb.sh_b) deg *= 2.0;
      if (pk.sh_a != pk.sh_b) deg *= 2.0;
      if (ij != kl) deg *= 2.0;
      Q.w = 0.5 * deg;
      const int AB[2] = {0, 1}, BA[2] = {1, 0}, CD[2] = {2, 3};
      const DPair& db = dp[ij];
      const DPair& dk = dp[kl];
      // d/dA: raised and lowered A against (C D)
      add_block(Q, 0, db.idx[0], false, (int)kl, AB, CD, true);
      if (db.idx[1] >= 0) add_block(Q, 1, db.idx[1], db.swap[1], (int)kl, AB, CD, false);
      // d/dB: the pair (A, B') or (B', A)
      add_block(Q, 2, db.idx[2], !db.swap[2], (int)kl, BA, CD, true);
      if (db.idx[3] >= 0) add_block(Q, 3, db.idx[3], true, (int)kl, BA, CD, false);
      // d/dC: raised and lowered C against (A B)
      add_block(Q, 4, dk.idx[0], false, (int)ij, CD, AB, true);
      if (dk.idx[1] >= 0) add_block(Q, 5, dk.idx[1], dk.swap[1], (int)ij, CD, AB, false);
      quartets.push_back(Q);
      if (off >= budget) flush();
    }
  flush();
  std::vector<double> g2((size_t)3 * na);
  download(g2.data(), d_grad.as<double>(), g2.size(), st);
  for (size_t x = 0; x < g.size(); ++x) grad[x] = g[x] + g2[x];
}

}  // namespace hfg
