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
#include <chrono>
#include <cstdio>
#include <cstdlib>
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

// Nuclear gradient (natom x 3, host) of the engine's molecule for the device
// density dP and energy-weighted density dW (row-major n x n), on stream st.
// Reads the engine's tables only, so it may run beside builds of other
// molecules.
void rhf_gradient_device(Engine& e, const std::vector<std::array<double, 4>>& atoms, const double* dP,
                         const double* dW, double* grad, hipStream_t st) {
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
  DevBuf d_pp, d_pd, d_scale, d_grad;
  d_pp.alloc(9 * npp * sizeof(double));
  const std::vector<double>* cols[9] = {&pp.p, &pp.Px, &pp.Py, &pp.Pz, &pp.PAx, &pp.PAy, &pp.PAz, &pp.K, &pp.H};
  for (int c = 0; c < 9; ++c)
    HIP_CHECK(hipMemcpyAsync(d_pp.as<double>() + c * npp, cols[c]->data(), npp * 8, hipMemcpyHostToDevice, st));
  d_pd.alloc(pd.size() * sizeof(PairDev));
  HIP_CHECK(hipMemcpyAsync(d_pd.p, pd.data(), pd.size() * sizeof(PairDev), hipMemcpyHostToDevice, st));
  d_scale.alloc((size_t)n * 8);
  HIP_CHECK(hipMemcpyAsync(d_scale.p, sc.data(), (size_t)n * 8, hipMemcpyHostToDevice, st));
  d_grad.alloc((size_t)3 * na * 8);
  hipLaunchKernelGGL(k_zero, dim3((3 * na + 255) / 256), dim3(256), 0, st, d_grad.as<double>(), (size_t)3 * na);
  // ---- one-electron terms on the device (onee_dev.hip)
  onee_gradient_device(sh, n, atoms, dP, dW, d_grad.as<double>(), st);
  const double* b = d_pp.as<double>();
  const PrimPairs view{b, b + npp, b + 2 * npp, b + 3 * npp, b + 4 * npp, b + 5 * npp, b + 6 * npp, b + 7 * npp,
                       b + 8 * npp};
  // The quartets of pair rows ij (kl <= ij, passing the Schwarz test) and
  // their derivative blocks.  A counting pass (parallel over rows) gives each
  // row's quartet, block and block-double counts; consecutive rows form
  // batches of at most `budget` block doubles; a second parallel pass fills
  // each row's segment of the batch, and the block tasks are bucketed by
  // class (row order kept within a class) for one launch per class.
  const size_t lds = (size_t)3 * na * sizeof(double);
  if (lds > 60000) throw HfgError{HFG_UNSUPPORTED, "gradient: too many atoms for the per-workgroup force sums"};
  const bool trace = std::getenv("HFG_TRACE") != nullptr;
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto secs = [](std::chrono::steady_clock::time_point t0, std::chrono::steady_clock::time_point t1) {
    return std::chrono::duration<double>(t1 - t0).count();
  };
  const auto t_start = now();
  double t_count = 0, t_fill = 0, t_gpu = 0;
  auto cls_key = [](int l0, int l1, int l2, int l3) { return ((l0 * 5 + l1) * 5 + l2) * 5 + l3; };
  constexpr int NKEY = 625;
  // One block of a quartet: derivative pair dpi against pair other.  With Q
  // null only the block's size is returned.
  auto add_block = [&](GQuartet* Q, long long& off, BlockTask* t_out, int* k_out, int dpi, bool dswap, int other,
                       const int sh_d[2], const int sh_o[2], bool raised_) -> long long {
    const std::array<int, 2> ld = pl[dpi], lo = pl[other];
    const long long size = (long long)ncart(ld[0]) * ncart(ld[1]) * ncart(lo[0]) * ncart(lo[1]);
    if (!Q) return size;
    // storage: the pair of higher rank first
    const bool dfirst = pair_rank(ld[0], ld[1]) >= pair_rank(lo[0], lo[1]);
    GBlock& B = Q->blk[Q->nblk++];
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
    if (dfirst) {
      *k_out = cls_key(ld[0], ld[1], lo[0], lo[1]);
      *t_out = BlockTask{dpi, other, off};
    } else {
      *k_out = cls_key(lo[0], lo[1], ld[0], ld[1]);
      *t_out = BlockTask{other, dpi, off};
    }
    off += size;
    return size;
  };
  // Quartet (ij, kl): its blocks (d/dA raised and lowered, d/dB raised and
  // lowered, d/dC raised and lowered); returns the block count and adds the
  // block doubles to off.  Count-only with Q null.
  auto make_quartet = [&](size_t ij, size_t kl, GQuartet* Q, BlockTask* t_out, int* k_out, long long& off) {
    const PairInfo& pb = e.pairs[ij];
    const PairInfo& pk = e.pairs[kl];
    if (Q) {
      std::memset(Q, 0, sizeof *Q);
      const int shells4[4] = {pb.sh_a, pb.sh_b, pk.sh_a, pk.sh_b};
      for (int k = 0; k < 4; ++k) {
        const Shell& s = sh[shells4[k]];
        Q->ao[k] = s.ao;
        Q->atom[k] = s.atom;
        Q->L[k] = (signed char)s.L;
      }
      double deg = 1.0;
      if (pb.sh_a != pb.sh_b) deg *= 2.0;
      if (pk.sh_a != pk.sh_b) deg *= 2.0;
      if (ij != kl) deg *= 2.0;
      Q->w = 0.5 * deg;
    }
    const int AB[2] = {0, 1}, BA[2] = {1, 0}, CD[2] = {2, 3};
    const DPair& db = dp[ij];
    const DPair& dk = dp[kl];
    int nb = 0;
    long long count_off = 0;
    long long& o = Q ? off : count_off;
    auto add = [&](int dpi, bool dswap, int other, const int* sd, const int* so, bool r) {
      const long long sz = add_block(Q, o, t_out + nb, k_out + nb, dpi, dswap, other, sd, so, r);
      if (!Q) o += sz;
      ++nb;
    };
    add(db.idx[0], false, (int)kl, AB, CD, true);
    if (db.idx[1] >= 0) add(db.idx[1], db.swap[1], (int)kl, AB, CD, false);
    add(db.idx[2], !db.swap[2], (int)kl, BA, CD, true);
    if (db.idx[3] >= 0) add(db.idx[3], true, (int)kl, BA, CD, false);
    add(dk.idx[0], false, (int)ij, CD, AB, true);
    if (dk.idx[1] >= 0) add(dk.idx[1], dk.swap[1], (int)ij, CD, AB, false);
    if (!Q) off += count_off;
    return nb;
  };
  // counting pass
  std::vector<long long> row_q(npair), row_b(npair), row_d(npair);
  {
    const auto t0 = now();
#pragma omp parallel for schedule(dynamic, 16)
    for (long long ij = 0; ij < (long long)npair; ++ij) {
      long long q = 0, nb = 0, d = 0;
      for (size_t kl = 0; kl <= (size_t)ij; ++kl) {
        if (!(e.Q[ij] * e.Q[kl] > e.screen)) continue;
        ++q;
        nb += make_quartet((size_t)ij, kl, nullptr, nullptr, nullptr, d);
      }
      row_q[ij] = q;
      row_b[ij] = nb;
      row_d[ij] = d;
    }
    t_count = secs(t0, now());
  }
  const long long budget = 1ll << 26;  // block doubles per batch (512 MB)
  DevBuf d_tasks, d_blk, d_q;
  size_t cap_t = 0, cap_b = 0, cap_q = 0;
  std::vector<GQuartet> quartets;
  std::vector<BlockTask> rtask, tasks;
  std::vector<int> rkey;
  size_t r0 = 0;
  int batches = 0;
  while (r0 < npair) {
    // rows [r0, r1): at least one row, then as many as fit the budget
    size_t r1 = r0 + 1;
    long long dsum = row_d[r0];
    while (r1 < npair && dsum + row_d[r1] <= budget) dsum += row_d[r1++];
    std::vector<long long> qo(r1 - r0 + 1, 0), bo(r1 - r0 + 1, 0), dof(r1 - r0 + 1, 0);
    for (size_t r = r0; r < r1; ++r) {
      qo[r - r0 + 1] = qo[r - r0] + row_q[r];
      bo[r - r0 + 1] = bo[r - r0] + row_b[r];
      dof[r - r0 + 1] = dof[r - r0] + row_d[r];
    }
    const size_t nq = (size_t)qo.back(), nbk = (size_t)bo.back();
    const long long ndbl = dof.back();
    r0 = r1;
    if (nq == 0) continue;
    const auto t0 = now();
    quartets.resize(nq);
    rtask.resize(nbk);
    rkey.resize(nbk);
    const size_t rbase = r1 - (qo.size() - 1);
#pragma omp parallel for schedule(dynamic, 16)
    for (long long r = 0; r < (long long)(qo.size() - 1); ++r) {
      const size_t ij = rbase + (size_t)r;
      long long qi = qo[r], bi = bo[r], off = dof[r];
      for (size_t kl = 0; kl <= ij; ++kl) {
        if (!(e.Q[ij] * e.Q[kl] > e.screen)) continue;
        bi += make_quartet(ij, kl, &quartets[qi], &rtask[bi], &rkey[bi], off);
        ++qi;
      }
    }
    // bucket the block tasks by class, row order kept within a class
    std::vector<size_t> cnt(NKEY + 1, 0);
    for (int k : rkey) ++cnt[k + 1];
    for (int k = 0; k < NKEY; ++k) cnt[k + 1] += cnt[k];
    std::vector<size_t> start(cnt.begin(), cnt.end() - 1);
    tasks.resize(nbk);
    for (size_t i = 0; i < nbk; ++i) tasks[cnt[rkey[i]]++] = rtask[i];
    t_fill += secs(t0, now());
    const auto t1 = now();
    if (nbk > cap_t) {
      d_tasks.alloc(nbk * sizeof(BlockTask));
      cap_t = nbk;
    }
    if ((size_t)ndbl > cap_b) {
      d_blk.alloc((size_t)ndbl * 8);
      cap_b = ndbl;
    }
    if (nq > cap_q) {
      d_q.alloc(nq * sizeof(GQuartet));
      cap_q = nq;
    }
    HIP_CHECK(hipMemcpyAsync(d_tasks.p, tasks.data(), nbk * sizeof(BlockTask), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(d_q.p, quartets.data(), nq * sizeof(GQuartet), hipMemcpyHostToDevice, st));
    for (int k = 0; k < NKEY; ++k) {
      const size_t a0 = start[k], z = cnt[k];  // cnt[k] is now the end of bucket k
      if (z == a0) continue;
      BlockArgs args{d_tasks.as<BlockTask>() + a0, (int)(z - a0), d_pd.as<PairDev>(), view, e.d_boys.as<double>(),
                     d_blk.as<double>()};
      launch_any_block(k / 125, (k / 25) % 5, (k / 5) % 5, k % 5, args, st);
    }
    const int blocks = (int)std::min<size_t>(2048, (nq + 3) / 4);
    hipLaunchKernelGGL(k_grad_contract, dim3(blocks), dim3(256), lds, st, d_q.as<GQuartet>(), (int)nq,
                       d_blk.as<double>(), dP, d_scale.as<double>(), n, na, d_grad.as<double>());
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));  // the host lists are reused
    t_gpu += secs(t1, now());
    ++batches;
  }
  if (trace)
    fprintf(stderr, "hfg gradient: %zu pairs, %d batches; count %.3f s, fill %.3f s, device %.3f s, total %.3f s\n",
            npair, batches, t_count, t_fill, t_gpu, secs(t_start, now()));
  std::vector<double> g2((size_t)3 * na);
  download(g2.data(), d_grad.as<double>(), g2.size(), st);
  for (size_t x = 0; x < g.size(); ++x) grad[x] = g[x] + g2[x];
}

// The same for host P and W, on the engine's stream.
void rhf_gradient(Engine& e, const std::vector<std::array<double, 4>>& atoms, const double* P, const double* W,
                  double* grad) {
  const size_t nn = (size_t)e.n * e.n;
  DevBuf d;
  d.alloc(2 * nn * sizeof(double));
  HIP_CHECK(hipMemcpyAsync(d.as<double>(), P, nn * 8, hipMemcpyHostToDevice, e.stream));
  HIP_CHECK(hipMemcpyAsync(d.as<double>() + nn, W, nn * 8, hipMemcpyHostToDevice, e.stream));
  rhf_gradient_device(e, atoms, d.as<double>(), d.as<double>() + nn, grad, e.stream);
}

}  // namespace hfg
