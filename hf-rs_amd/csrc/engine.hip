// engine.hip -- direct J/K digestion of screened shell quartets on gfx950.
//
// Work decomposition: one wavefront per task = one bra shell pair (uniform
// across the wave, so its primitive data is read through scalar loads) and up
// to 64 ket shell pairs (one per lane) of one primitive count, drawn from a
// Schwarz-sorted list so that the surviving kets of a bra are a prefix.  Each
// lane contracts its quartet in registers with the generated Head-Gordon-Pople
// recurrences, then digests the block into J and K:
//   Jacc[ij] += 2w sum_kl D_kl (ij|kl)   (wave-reduced, one atomic per ij)
//   Jacc[kl] += 2w sum_ij D_ij (ij|kl)
//   Kacc[ik] += w sum_jl D_jl v, Kacc[jk] += w sum_il D_il v,
//   Kacc[il] += w sum_jk D_jk v, Kacc[jl] += w sum_ik D_ik v
// with w = 1/((1+d_AB)(1+d_CD)(1+d_AB,CD)); then J = Jacc + Jacc^T and
// K = Kacc + Kacc^T.  This equals contract_jk (src/scf.rs:337-364) applied to
// the packed tensor of compute_shell_eri_tensor (src/shell_integrals.rs:156-240)
// without ever storing it.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "engine.h"
#include "eri_device.h"
#include "orient.h"

namespace hfg {

// J/K partial sums live in one copy per XCC (accelerator complex die, each
// with its own L2).  A wave never migrates, so every update to copy x comes
// from XCC x and a workgroup-scope atomic -- performed in that XCC's L2 --
// is enough; no update has to travel to another die's cache.  The copies are
// summed by k_symmetrize after the kernel boundary.  The XCC id is the 4-bit
// hardware register HIP's __smid also reads; NXCC_SLOTS covers all 16 values.
constexpr int NXCC_SLOTS = 16;
__device__ __forceinline__ unsigned xcc_id() {
  return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & (NXCC_SLOTS - 1);
}
__device__ __forceinline__ void add_local(double* p, double v) {
  __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

struct JKArgs {
  const JKTask* tasks;
  int ntask, n;
  const PairDev* pairs;
  const KetGroup* groups;
  PrimPairs pp;  // bra side, contiguous per pair
  PrimPairs kt;  // ket side, transposed per group
  size_t nn;      // n*n, the stride between per-XCC accumulator copies
  const double* scale;
  const double* boys;
  const double* D;
  double* Jacc;
  double* Kacc;
};

// Boys slice of the class's order M into LDS (see eri_device.h).
template <int M>
__device__ __forceinline__ void load_boys_slice(const double* __restrict__ g, double* __restrict__ lds) {
  for (int r = threadIdx.x; r < BOYS_ROWS; r += blockDim.x) boys_slice_row<M>(g, r, lds + r * BOYS_SLICE);
  __syncthreads();
}

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void jk_task(const JKArgs& a, const double* __restrict__ boys, int task, int lane);

// Persistent waves: every wave walks the class's task list with a stride of
// the total wave count (tasks are ordered heaviest first), so the LDS copy of
// the Boys slice is made once per workgroup.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(256) void k_jk(JKArgs a) {
  __shared__ double boys[BOYS_ROWS * BOYS_SLICE];
  load_boys_slice<LA + LB + LC + LD>(a.boys, boys);
  const int lane = threadIdx.x & 63;
  const int waves = (int)(gridDim.x * (blockDim.x >> 6));
  for (int task = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); task < a.ntask; task += waves)
    jk_task<LA, LB, LC, LD>(a, boys, task, lane);
}

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void jk_task(const JKArgs& a, const double* __restrict__ boys, int task, int lane) {
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
  constexpr int NAB = NA * NB, NCD = NC * ND;
  const JKTask t = a.tasks[task];
  const int bra = __builtin_amdgcn_readfirstlane(t.bra);
  const int ket0 = __builtin_amdgcn_readfirstlane(t.ket0);
  const int nket = __builtin_amdgcn_readfirstlane(t.nket);
  const int grp = __builtin_amdgcn_readfirstlane(t.grp);
  const bool act = lane < nket;
  const int ket = ket0 + (act ? lane : 0);
  const PairDev& pb = a.pairs[bra];
  const PairDev& pk = a.pairs[ket];
  const KetGroup g = a.groups[grp];
  double I[NAB * NCD];
  eri_block<LA, LB, LC, LD>(a.pp, pb.p0, pb.p1, a.kt, g.base + (ket - g.start), act ? g.npp : 0, g.size, pb.AB,
                            pk.AB, boys, I);

  double w = 1.0;
  if (pb.sh_a == pb.sh_b) w *= 0.5;
  if (pk.sh_a == pk.sh_b) w *= 0.5;
  if (bra == ket) w *= 0.5;
  const int A0 = pb.ao_a, B0 = pb.ao_b, C0 = pk.ao_a, D0 = pk.ao_b, n = a.n;
  double sa[NA], sb[NB], sc[NC], sd[ND];
#pragma unroll
  for (int i = 0; i < NA; ++i) sa[i] = a.scale[A0 + i] * w;
#pragma unroll
  for (int i = 0; i < NB; ++i) sb[i] = a.scale[B0 + i];
#pragma unroll
  for (int i = 0; i < NC; ++i) sc[i] = a.scale[C0 + i];
#pragma unroll
  for (int i = 0; i < ND; ++i) sd[i] = a.scale[D0 + i];
#pragma unroll
  for (int i = 0; i < NA; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int l = 0; l < ND; ++l) I[((i * NB + j) * NC + k) * ND + l] *= sa[i] * sb[j] * sc[k] * sd[l];

  const double* __restrict__ Dm = a.D;
  const size_t slot = (size_t)xcc_id() * a.nn;
  double* __restrict__ Jx = a.Jacc + slot;
  double* __restrict__ Kx = a.Kacc + slot;
  // Coulomb, bra side: reduce over the wave (all lanes share the bra pair).
  {
    double Dcd[NCD];
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
      for (int l = 0; l < ND; ++l) Dcd[k * ND + l] = Dm[(size_t)(C0 + k) * n + D0 + l];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        double s = 0.0;
#pragma unroll
        for (int kl = 0; kl < NCD; ++kl) s += Dcd[kl] * I[(i * NB + j) * NCD + kl];
        s = wave_sum(s);
        if (lane == 0) add_local(&Jx[(size_t)(A0 + i) * n + B0 + j], 2.0 * s);
      }
  }
  if (!act) return;
  {
    double Dab[NAB];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) Dab[i * NB + j] = Dm[(size_t)(A0 + i) * n + B0 + j];
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
      for (int l = 0; l < ND; ++l) {
        double s = 0.0;
#pragma unroll
        for (int ij = 0; ij < NAB; ++ij) s += Dab[ij] * I[ij * NCD + k * ND + l];
        add_local(&Jx[(size_t)(C0 + k) * n + D0 + l], 2.0 * s);
      }
  }
  // Exchange.
  {
    double Dx[NB * ND];  // D[b_j, d_l] -> K[a_i, c_k]
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int l = 0; l < ND; ++l) Dx[j * ND + l] = Dm[(size_t)(B0 + j) * n + D0 + l];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int l = 0; l < ND; ++l) s += Dx[j * ND + l] * I[((i * NB + j) * NC + k) * ND + l];
        add_local(&Kx[(size_t)(A0 + i) * n + C0 + k], s);
      }
  }
  {
    double Dx[NA * ND];  // D[a_i, d_l] -> K[b_j, c_k]
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int l = 0; l < ND; ++l) Dx[i * ND + l] = Dm[(size_t)(A0 + i) * n + D0 + l];
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
          for (int l = 0; l < ND; ++l) s += Dx[i * ND + l] * I[((i * NB + j) * NC + k) * ND + l];
        add_local(&Kx[(size_t)(B0 + j) * n + C0 + k], s);
      }
  }
  {
    double Dx[NB * NC];  // D[b_j, c_k] -> K[a_i, d_l]
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int k = 0; k < NC; ++k) Dx[j * NC + k] = Dm[(size_t)(B0 + j) * n + C0 + k];
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int l = 0; l < ND; ++l) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NB; ++j)
#pragma unroll
          for (int k = 0; k < NC; ++k) s += Dx[j * NC + k] * I[((i * NB + j) * NC + k) * ND + l];
        add_local(&Kx[(size_t)(A0 + i) * n + D0 + l], s);
      }
  }
  {
    double Dx[NA * NC];  // D[a_i, c_k] -> K[b_j, d_l]
#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int k = 0; k < NC; ++k) Dx[i * NC + k] = Dm[(size_t)(A0 + i) * n + C0 + k];
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int l = 0; l < ND; ++l) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NA; ++i)
#pragma unroll
          for (int k = 0; k < NC; ++k) s += Dx[i * NC + k] * I[((i * NB + j) * NC + k) * ND + l];
        add_local(&Kx[(size_t)(B0 + j) * n + D0 + l], s);
      }
  }
}

// One thread per quartet: the unscaled block in kernel orientation.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_block(BlockArgs a) {
  constexpr int NT = ncart(LA) * ncart(LB) * ncart(LC) * ncart(LD);
  __shared__ double boys[BOYS_ROWS * BOYS_SLICE];
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

__global__ void k_fill(double* p, size_t n, double v);

__global__ void k_copy_calibration(const double4* __restrict__ src, double4* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

double calibration_copy(size_t bytes) {
  const size_t n = bytes / sizeof(double4);
  DevBuf a, b;
  a.alloc(n * sizeof(double4));
  b.alloc(n * sizeof(double4));
  hipStream_t s;
  HIP_CHECK(hipStreamCreate(&s));
  hipLaunchKernelGGL(k_fill, dim3((4 * n + 255) / 256), dim3(256), 0, s, a.as<double>(), 4 * n, 1.0);
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, s));
  hipLaunchKernelGGL(k_copy_calibration, dim3(4096), dim3(256), 0, s, a.as<double4>(), b.as<double4>(), n);
  HIP_CHECK(hipEventRecord(e1, s));
  HIP_CHECK(hipStreamSynchronize(s));
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  return ms;
}

__global__ void k_fill(double* p, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_symmetrize(const double* Jacc, const double* Kacc, double* J, double* K, int n, const double* H,
                             double* F) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nn = (size_t)n * n;
  if (idx >= nn) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  const size_t tr = (size_t)c * n + r;
  double j = 0.0, k = 0.0;
  for (int x = 0; x < NXCC_SLOTS; ++x) {
    j += Jacc[x * nn + idx] + Jacc[x * nn + tr];
    k += Kacc[x * nn + idx] + Kacc[x * nn + tr];
  }
  if (J) J[idx] = j;
  if (K) K[idx] = k;
  if (F) F[idx] = H[idx] + j - 0.5 * k;
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
typedef void (*JKLaunch)(const JKArgs&, int, hipStream_t);
typedef void (*BlockLaunch)(const BlockArgs&, hipStream_t);

static int g_max_blocks = 1024;  // persistent grid: workgroups per launch (set from the CU count)

template <int LA, int LB, int LC, int LD>
void launch_jk(const JKArgs& a, int ntask, hipStream_t s) {
  JKArgs b = a;
  b.ntask = ntask;
  const int waves_per_block = 4;
  const int blocks = std::min(g_max_blocks, (ntask + waves_per_block - 1) / waves_per_block);
  hipLaunchKernelGGL((k_jk<LA, LB, LC, LD>), dim3(blocks), dim3(64 * waves_per_block), 0, s, b);
}
template <int LA, int LB, int LC, int LD>
void launch_block(const BlockArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_block<LA, LB, LC, LD>), dim3((a.ntask + 63) / 64), dim3(64), 0, s, a);
}

static const JKLaunch JK_LAUNCH[HFG_NCLASS] = {
#define X(a, b, c, d) &launch_jk<a, b, c, d>,
    HFG_FOR_EACH_CLASS(X)
#undef X
};
static const BlockLaunch BLOCK_LAUNCH[HFG_NCLASS_ALL] = {
#define X(a, b, c, d) &launch_block<a, b, c, d>,
    HFG_FOR_EACH_CLASS_ALL(X)
#undef X
};
static const int CLASS_L[HFG_NCLASS][4] = {
#define X(a, b, c, d) {a, b, c, d},
    HFG_FOR_EACH_CLASS(X)
#undef X
};

static int nc(int L) { return (L + 1) * (L + 2) / 2; }

void launch_block_class(int cls, const BlockArgs& a, hipStream_t s) {
  BLOCK_LAUNCH[cls](a, s);
  HIP_CHECK(hipGetLastError());
}

Engine::Engine(const hfg_basis_view& b) {
  shells = group_shells(b);
  n = b.n_functions;
  for (auto& s : shells)
    if (s.L > 2) throw HfgError{HFG_UNSUPPORTED, "shell-quartet classes are compiled up to d functions (L<=2)"};
  int ndev = 0;
  HIP_CHECK(hipGetDeviceCount(&ndev));
  if (ndev <= 0) throw HfgError{HFG_DEVICE, "no GPU visible"};
  HIP_CHECK(hipGetDevice(&device));
  HIP_CHECK(hipStreamCreate(&stream));  // default priority
  {
    int cus = 0;
    HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    g_max_blocks = std::max(1, cus) * 4;  // 4 workgroups (16 waves) per CU; LDS 4 x 32 KB
    if (const char* env = std::getenv("HFG_PERSISTENT_BLOCKS")) g_max_blocks = std::max(1, std::atoi(env));
  }
  build_pairs(shells, pairs, pp);
  ao_scale.assign(n, 1.0);
  for (auto& s : shells)
    for (size_t c = 0; c < s.scale.size(); ++c) ao_scale[s.ao + c] = s.scale[c];

  // primitive pairs: 8 SoA arrays in one allocation
  const size_t np = pp.size();
  d_pp.alloc(9 * np * sizeof(double));
  const std::vector<double>* cols[9] = {&pp.p, &pp.Px, &pp.Py, &pp.Pz, &pp.PAx, &pp.PAy, &pp.PAz, &pp.K, &pp.H};
  for (int c = 0; c < 9; ++c)
    HIP_CHECK(hipMemcpyAsync(d_pp.as<double>() + c * np, cols[c]->data(), np * sizeof(double), hipMemcpyHostToDevice,
                             stream));
  std::vector<PairDev> pd(pairs.size());
  for (size_t i = 0; i < pairs.size(); ++i) {
    const PairInfo& p = pairs[i];
    PairDev& d = pd[i];
    std::memset(&d, 0, sizeof d);
    d.p0 = p.p0;
    d.p1 = p.p1;
    d.ao_a = p.ao_a;
    d.ao_b = p.ao_b;
    d.sh_a = p.sh_a;
    d.sh_b = p.sh_b;
    for (int k = 0; k < 3; ++k) d.AB[k] = p.AB[k];
  }
  d_pairs.alloc(pd.size() * sizeof(PairDev));
  HIP_CHECK(hipMemcpyAsync(d_pairs.p, pd.data(), pd.size() * sizeof(PairDev), hipMemcpyHostToDevice, stream));
  auto tab = boys_table();
  d_boys.alloc(tab.size() * sizeof(double));
  HIP_CHECK(hipMemcpyAsync(d_boys.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  d_scale.alloc(n * sizeof(double));
  HIP_CHECK(hipMemcpyAsync(d_scale.p, ao_scale.data(), n * sizeof(double), hipMemcpyHostToDevice, stream));
  const size_t nn = (size_t)n * n * sizeof(double);
  d_D.alloc(nn);
  d_Jacc.alloc(nn * NXCC_SLOTS);
  d_Kacc.alloc(nn * NXCC_SLOTS);
  d_J.alloc(nn);
  d_K.alloc(nn);
  d_H.alloc(nn);
  d_F.alloc(nn);
  sync();
  compute_schwarz();
  build_tasks();
}

Engine::~Engine() {
  for (auto e : events) (void)hipEventDestroy(e);
  if (stream) (void)hipStreamDestroy(stream);
}

// McMurchie-Davidson loop-count model of one primitive quartet (see hfg.h).
double md_flops(int la, int lb, int lc, int ld) {
  auto comps = [](int L) {
    std::vector<std::array<int, 3>> o;
    for (int x = L; x >= 0; --x)
      for (int y = L - x; y >= 0; --y) o.push_back({x, y, L - x - y});
    return o;
  };
  const int L = la + lb + lc + ld;
  double c4 = 1;
  for (int i = 1; i <= 4; ++i) c4 = c4 * (L + i) / i;  // C(L+4, 4)
  double f = 20 + 8 + 16.0 * (L + 1) + 2.0 * (L + 1) + 3.0 * (c4 - (L + 1));
  for (auto& a : comps(la))
    for (auto& b : comps(lb))
      for (auto& c : comps(lc))
        for (auto& d : comps(ld)) {
          double s = 0, pr = 1;
          for (int k = 0; k < 3; ++k) {
            const int lab = a[k] + b[k], lcd = c[k] + d[k];
            s += (lab + 1.0) * (lcd + 1.0);
            pr *= lab + lcd + 1.0;
          }
          f += 2 * s + 3 * pr + 3;
        }
  return f;
}

void Engine::sync() { HIP_CHECK(hipStreamSynchronize(stream)); }

PrimPairs Engine::pps_view() const {
  const double* b = d_pps.as<double>();
  const size_t m = pps_size;
  return PrimPairs{b, b + m, b + 2 * m, b + 3 * m, b + 4 * m, b + 5 * m, b + 6 * m, b + 7 * m, b + 8 * m};
}

// Primitive-pair screening for the J/K work.  B_q = sqrt(max_ij |(ij|ij)|) of
// the single primitive pair q (component factors applied) bounds every
// primitive integral it enters: |(q|r)| <= B_q B_r.  A primitive pair is left
// out of the J/K kernels when B_q * max_r B_r < prim_screen (1e-20), so an
// integral changes by at most K_ab K_cd * 1e-20 (<= 7e-17 for 81 x 81
// primitive pairs).  The reference's shell engine evaluates every primitive
// (src/shell_integrals.rs:101-133); shell blocks, Schwarz bounds and the
// packed tensor here still do.
void Engine::screen_primitives(std::vector<int>& s0, std::vector<int>& s1, PrimPairHost& pps) {
  const size_t np = pp.size();
  std::vector<PairDev> tmp(np);
  std::vector<BlockTask> tasks;
  std::vector<int> cls;
  std::vector<long long> off(np);
  std::vector<int> owner(np);
  std::vector<size_t> order(pairs.size());
  std::iota(order.begin(), order.end(), 0);
  auto pcl = [&](size_t i) { return pair_class(pairs[i].la, pairs[i].lb); };
  std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return pcl(x) < pcl(y); });
  long long o = 0;
  for (size_t i : order) {
    const PairInfo& p = pairs[i];
    const int nab = nc(p.la) * nc(p.lb);
    for (int q = p.p0; q < p.p1; ++q) {
      PairDev& d = tmp[q];
      std::memset(&d, 0, sizeof d);
      d.p0 = q;
      d.p1 = q + 1;
      d.ao_a = p.ao_a;
      d.ao_b = p.ao_b;
      d.sh_a = p.sh_a;
      d.sh_b = p.sh_b;
      for (int k = 0; k < 3; ++k) d.AB[k] = p.AB[k];
      tasks.push_back({q, q, o});
      cls.push_back(quartet_class(pcl(i), pcl(i)));
      off[q] = o;
      owner[q] = (int)i;
      o += (long long)nab * nab;
    }
  }
  DevBuf dtmp;
  dtmp.alloc(std::max<size_t>(1, np) * sizeof(PairDev));
  HIP_CHECK(hipMemcpyAsync(dtmp.p, tmp.data(), np * sizeof(PairDev), hipMemcpyHostToDevice, stream));
  std::vector<double> out(o);
  blocks(tasks, cls, out, dtmp.as<PairDev>());
  std::vector<double> B(np, 0.0);
  double bmax = 0.0;
  for (size_t q = 0; q < np; ++q) {
    const PairInfo& p = pairs[owner[q]];
    const int na = nc(p.la), nb = nc(p.lb), nab = na * nb;
    double m = 0.0;
    for (int x = 0; x < na; ++x)
      for (int y = 0; y < nb; ++y) {
        const int ij = x * nb + y;
        const double f = ao_scale[p.ao_a + x] * ao_scale[p.ao_b + y];
        m = std::max(m, std::fabs(out[off[q] + (size_t)ij * nab + ij]) * f * f);
      }
    B[q] = std::sqrt(m);
    bmax = std::max(bmax, B[q]);
  }
  const std::vector<double>* src[9] = {&pp.p, &pp.Px, &pp.Py, &pp.Pz, &pp.PAx, &pp.PAy, &pp.PAz, &pp.K, &pp.H};
  std::vector<double>* dst[9] = {&pps.p, &pps.Px, &pps.Py, &pps.Pz, &pps.PAx, &pps.PAy, &pps.PAz, &pps.K, &pps.H};
  s0.assign(pairs.size(), 0);
  s1.assign(pairs.size(), 0);
  for (size_t i = 0; i < pairs.size(); ++i) {
    s0[i] = (int)pps.size();
    for (int q = pairs[i].p0; q < pairs[i].p1; ++q)
      if (B[q] * bmax >= prim_screen)
        for (int c = 0; c < 9; ++c) dst[c]->push_back((*src[c])[q]);
    s1[i] = (int)pps.size();
  }
  prim_pairs_kept = pps.size();
}

PrimPairs Engine::kt_view() const {
  const double* b = d_kt.as<double>();
  const size_t m = kt_size;
  return PrimPairs{b, b + m, b + 2 * m, b + 3 * m, b + 4 * m, b + 5 * m, b + 6 * m, b + 7 * m, b + 8 * m};
}

PrimPairs Engine::pp_view() const {
  const double* b = d_pp.as<double>();
  const size_t np = pp.size();
  return PrimPairs{b, b + np, b + 2 * np, b + 3 * np, b + 4 * np, b + 5 * np, b + 6 * np, b + 7 * np, b + 8 * np};
}

void Engine::blocks(const std::vector<BlockTask>& tasks, const std::vector<int>& cls, std::vector<double>& out,
                    const PairDev* pair_table) {
  // tasks are grouped by class: cls[i] is the class of tasks[i]; offsets are
  // into `out`, whose size the caller has set.
  DevBuf dt, dout;
  dt.alloc(tasks.size() * sizeof(BlockTask));
  dout.alloc(out.size() * sizeof(double));
  HIP_CHECK(hipMemcpyAsync(dt.p, tasks.data(), tasks.size() * sizeof(BlockTask), hipMemcpyHostToDevice, stream));
  size_t i = 0;
  while (i < tasks.size()) {
    size_t j = i;
    while (j < tasks.size() && cls[j] == cls[i]) ++j;
    BlockArgs a{dt.as<BlockTask>() + i, (int)(j - i), pair_table ? pair_table : d_pairs.as<PairDev>(), pp_view(),
                d_boys.as<double>(), dout.as<double>()};
    BLOCK_LAUNCH[cls[i]](a, stream);
    HIP_CHECK(hipGetLastError());
    i = j;
  }
  HIP_CHECK(hipMemcpyAsync(out.data(), dout.p, out.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
  sync();
}

void Engine::compute_schwarz() {
  // (ab|ab) blocks of every pair, Q_ab = sqrt(max_ij (ij|ij)) with the
  // component factors applied (src/shell_integrals.rs:165-173).
  std::vector<int> order(pairs.size());
  std::iota(order.begin(), order.end(), 0);
  auto pcl = [&](int i) { return pair_class(pairs[i].la, pairs[i].lb); };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pcl(x) < pcl(y); });
  std::vector<BlockTask> tasks;
  std::vector<int> cls;
  long long off = 0;
  std::vector<long long> offs(pairs.size());
  for (int i : order) {
    const int nab = nc(pairs[i].la) * nc(pairs[i].lb);
    tasks.push_back({i, i, off});
    cls.push_back(quartet_class(pcl(i), pcl(i)));
    offs[i] = off;
    off += (long long)nab * nab;
  }
  std::vector<double> out(off);
  blocks(tasks, cls, out);
  Q.assign(pairs.size(), 0.0);
  for (size_t i = 0; i < pairs.size(); ++i) {
    const PairInfo& p = pairs[i];
    const int na = nc(p.la), nb = nc(p.lb), nab = na * nb;
    double m = 0.0;
    for (int x = 0; x < na; ++x)
      for (int y = 0; y < nb; ++y) {
        const int ij = x * nb + y;
        const double f = ao_scale[p.ao_a + x] * ao_scale[p.ao_b + y];
        m = std::max(m, out[offs[i] + (size_t)ij * nab + ij] * f * f);
      }
    Q[i] = std::sqrt(m);
  }
}

void Engine::build_tasks() {
  // Pair groups: (pair class, primitive count), each sorted by Q descending.
  struct Group {
    int pc, npp;
    std::vector<int> idx;
    std::vector<double> q;
    std::vector<double> full;
  };
  std::vector<int> s0, s1;
  PrimPairHost pps;
  screen_primitives(s0, s1, pps);
  std::vector<Group> groups;
  for (size_t i = 0; i < pairs.size(); ++i) {
    const int pc = pair_class(pairs[i].la, pairs[i].lb), npp = s1[i] - s0[i];
    if (npp == 0) continue;  // every primitive pair negligible: no J/K work
    auto it = std::find_if(groups.begin(), groups.end(), [&](const Group& g) { return g.pc == pc && g.npp == npp; });
    if (it == groups.end()) {
      groups.push_back({pc, npp, {}, {}, {}});
      it = groups.end() - 1;
    }
    it->idx.push_back((int)i);
  }
  for (auto& g : groups) {
    std::stable_sort(g.idx.begin(), g.idx.end(), [&](int x, int y) { return Q[x] > Q[y]; });
    for (int i : g.idx) g.q.push_back(Q[i]);
    // unscreened primitive counts, prefix-summed, for the algorithmic FLOP count
    g.full.assign(1, 0.0);
    for (int i : g.idx) g.full.push_back(g.full.back() + (pairs[i].p1 - pairs[i].p0));
  }
  // Heavy groups first inside each class.
  std::sort(groups.begin(), groups.end(), [](const Group& x, const Group& y) {
    return x.pc != y.pc ? x.pc < y.pc : x.npp > y.npp;
  });
  std::vector<std::vector<JKTask>> per_class(HFG_NCLASS);
  n_quartets = 0;
  prim_quartets = 0;
  class_flops.assign(HFG_NCLASS, 0.0);
  class_quartets.assign(HFG_NCLASS, 0);
  class_ms.assign(HFG_NCLASS, 0.0);
  class_launches.assign(HFG_NCLASS, 0);
  // ket lists are stored as contiguous index arrays: the task's ket0 indexes
  // into a flattened ket order, translated to pair ids through `ketmap`.
  for (size_t gb = 0; gb < groups.size(); ++gb)
    for (size_t gk = 0; gk < groups.size(); ++gk) {
      const Group& B = groups[gb];
      const Group& K = groups[gk];
      if (K.pc > B.pc) continue;
      if (K.pc == B.pc && gk > gb) continue;  // unordered pairs of groups once
      const bool tri = gb == gk;
      const int cls = quartet_class(B.pc, K.pc);
      for (size_t i = 0; i < B.idx.size(); ++i) {
        const double qb = B.q[i];
        // kets with qb * qk > screen form a prefix (sorted descending)
        size_t lo = 0, hi = K.q.size();
        while (lo < hi) {
          size_t mid = (lo + hi) / 2;
          if (qb * K.q[mid] > screen)
            lo = mid + 1;
          else
            hi = mid;
        }
        size_t nk = lo;
        if (tri) nk = std::min(nk, i + 1);
        n_quartets += (int64_t)nk;
        class_quartets[cls] += (int64_t)nk;
        const double pq = (double)(B.full[i + 1] - B.full[i]) * K.full[nk];
        prim_quartets += pq;
        class_flops[cls] += pq * md_flops(CLASS_L[cls][0], CLASS_L[cls][1], CLASS_L[cls][2], CLASS_L[cls][3]);
        for (size_t k0 = 0; k0 < nk; k0 += 64) {
          JKTask t;
          t.bra = B.idx[i];
          t.ket0 = (int)k0;  // position in K.idx; fixed up below
          t.nket = (int)std::min<size_t>(64, nk - k0);
          t.grp = (int)gk;
          per_class[cls].push_back(t);
        }
      }
    }
  // Ket pairs of one task must be consecutive pair ids: renumber pairs so that
  // every group is a contiguous, Q-sorted range.
  std::vector<int> newid(pairs.size(), -1), oldid;
  std::vector<int> gstart(groups.size());
  for (size_t g = 0; g < groups.size(); ++g) {
    gstart[g] = (int)oldid.size();
    for (int i : groups[g].idx) {
      newid[i] = (int)oldid.size();
      oldid.push_back(i);
    }
  }
  std::vector<PairDev> pd(std::max<size_t>(1, oldid.size()));
  for (size_t k = 0; k < oldid.size(); ++k) {
    const PairInfo& p = pairs[oldid[k]];
    PairDev& d = pd[k];
    std::memset(&d, 0, sizeof d);
    d.p0 = s0[oldid[k]];
    d.p1 = s1[oldid[k]];
    d.ao_a = p.ao_a;
    d.ao_b = p.ao_b;
    d.sh_a = p.sh_a;
    d.sh_b = p.sh_b;
    for (int c = 0; c < 3; ++c) d.AB[c] = p.AB[c];
  }
  std::vector<JKTask> all;
  task_begin.assign(HFG_NCLASS + 1, 0);
  for (int c = 0; c < HFG_NCLASS; ++c) {
    task_begin[c] = all.size();
    for (JKTask t : per_class[c]) {
      t.bra = newid[t.bra];
      t.ket0 = gstart[t.grp] + t.ket0;
      all.push_back(t);
    }
  }
  task_begin[HFG_NCLASS] = all.size();
  // transposed ket primitive data per group
  std::vector<KetGroup> kg(groups.size());
  size_t tot = 0;
  for (size_t g = 0; g < groups.size(); ++g) {
    kg[g] = KetGroup{gstart[g], (int)groups[g].idx.size(), (int)tot, groups[g].npp};
    tot += (size_t)groups[g].npp * groups[g].idx.size();
  }
  kt_size = tot;
  const std::vector<double>* cols[9] = {&pps.p, &pps.Px, &pps.Py, &pps.Pz, &pps.PAx, &pps.PAy, &pps.PAz, &pps.K, &pps.H};
  std::vector<double> kt(9 * std::max<size_t>(1, tot));
  for (size_t g = 0; g < groups.size(); ++g)
    for (size_t j = 0; j < groups[g].idx.size(); ++j) {
      const int first = s0[groups[g].idx[j]];
      for (int ik = 0; ik < groups[g].npp; ++ik)
        for (int c = 0; c < 9; ++c)
          kt[c * tot + kg[g].base + (size_t)ik * kg[g].size + j] = (*cols[c])[first + ik];
    }
  pps_size = pps.size();
  d_pps.alloc(9 * std::max<size_t>(1, pps_size) * sizeof(double));
  for (int c = 0; c < 9; ++c)
    if (pps_size)
      HIP_CHECK(hipMemcpyAsync(d_pps.as<double>() + c * pps_size, cols[c]->data(), pps_size * sizeof(double),
                               hipMemcpyHostToDevice, stream));
  d_kt.alloc(kt.size() * sizeof(double));
  HIP_CHECK(hipMemcpyAsync(d_kt.p, kt.data(), kt.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  d_groups.alloc(std::max<size_t>(1, kg.size()) * sizeof(KetGroup));
  HIP_CHECK(hipMemcpyAsync(d_groups.p, kg.data(), kg.size() * sizeof(KetGroup), hipMemcpyHostToDevice, stream));
  d_jkpairs.alloc(pd.size() * sizeof(PairDev));
  HIP_CHECK(hipMemcpyAsync(d_jkpairs.p, pd.data(), pd.size() * sizeof(PairDev), hipMemcpyHostToDevice, stream));
  d_tasks.alloc(std::max<size_t>(1, all.size()) * sizeof(JKTask));
  if (!all.empty())
    HIP_CHECK(hipMemcpyAsync(d_tasks.p, all.data(), all.size() * sizeof(JKTask), hipMemcpyHostToDevice, stream));
  sync();
}

void Engine::jk_device(const double* dP, double* dJ, double* dK, const double* dH, double* dF) {
  const size_t nn = (size_t)n * n;
  const int tb = 256;
  const size_t nacc = nn * NXCC_SLOTS;
  hipLaunchKernelGGL(k_fill, dim3((nacc + tb - 1) / tb), dim3(tb), 0, stream, d_Jacc.as<double>(), nacc, 0.0);
  hipLaunchKernelGGL(k_fill, dim3((nacc + tb - 1) / tb), dim3(tb), 0, stream, d_Kacc.as<double>(), nacc, 0.0);
  JKArgs a;
  a.n = n;
  a.nn = nn;
  a.pairs = d_jkpairs.as<PairDev>();
  a.groups = d_groups.as<KetGroup>();
  a.pp = pps_view();
  a.kt = kt_view();
  a.scale = d_scale.as<double>();
  a.boys = d_boys.as<double>();
  a.D = dP;
  a.Jacc = d_Jacc.as<double>();
  a.Kacc = d_Kacc.as<double>();
  for (int c = 0; c < HFG_NCLASS; ++c) {
    const size_t nt = task_begin[c + 1] - task_begin[c];
    if (!nt) continue;
    a.tasks = d_tasks.as<JKTask>() + task_begin[c];
    if (profile) HIP_CHECK(hipEventRecord(events[2 * c], stream));
    JK_LAUNCH[c](a, (int)nt, stream);
    HIP_CHECK(hipGetLastError());
    if (profile) HIP_CHECK(hipEventRecord(events[2 * c + 1], stream));
  }
  hipLaunchKernelGGL(k_symmetrize, dim3((nn + tb - 1) / tb), dim3(tb), 0, stream, d_Jacc.as<double>(),
                     d_Kacc.as<double>(), dJ, dK, n, dH, dF);
  HIP_CHECK(hipGetLastError());
  if (profile) {
    sync();
    for (int c = 0; c < HFG_NCLASS; ++c) {
      if (task_begin[c + 1] == task_begin[c]) continue;
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, events[2 * c], events[2 * c + 1]));
      class_ms[c] += ms;
      class_launches[c] += 1;
    }
  }
}

void Engine::set_profile(bool on) {
  if (on && events.empty()) {
    events.resize(2 * HFG_NCLASS);
    for (auto& e : events) HIP_CHECK(hipEventCreate(&e));
  }
  if (on) {  // a new measurement; disabling keeps the statistics for reading
    class_ms.assign(HFG_NCLASS, 0.0);
    class_launches.assign(HFG_NCLASS, 0);
  }
  profile = on;
}

int persistent_blocks() { return g_max_blocks; }

void class_angular(int c, int* l4) {
  for (int k = 0; k < 4; ++k) l4[k] = CLASS_L[c][k];
}

}  // namespace hfg
