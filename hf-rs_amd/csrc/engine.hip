// engine.hip -- per-molecule engine state: shell-pair tables, Schwarz bounds,
// the block kernels (one thread per shell quartet, used for Schwarz bounds,
// shell blocks, the packed tensor and the density-fitting integrals) and the
// Fock-build driver over the J/K kernels (jk_kernel.h, jk_plan.cpp).
#include <algorithm>
#include <array>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "engine.h"
#include "eri_device.h"
#include "orient.h"

namespace hfg {

// One thread per quartet: the unscaled block in kernel orientation.
template <int LA, int LB, int LC, int LD>
__global__ __launch_bounds__(64) void k_block(BlockArgs a) {
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

__global__ void k_zero_int(int* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
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

// The density pointer and Schwarz threshold of the next build in a
// molecule's table (a kernel, so that it orders with the launches around it).
__global__ void k_tab_set(MolTab* tab, const double* D, double screen) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    tab->D = D;
    tab->screen = screen;
  }
}

// ---------------------------------------------------------------------------
// Host side.

JKDirectory make_directory(const std::vector<int>& mol, const std::vector<const MolTab*>& tabs_host) {
  JKDirectory d;
  d.off.assign(HFG_NSLOT, 0);
  d.count.assign(HFG_NSLOT, 0);
  d.njob.assign(HFG_NSLOT, 0);
  d.ntask.assign(HFG_NSLOT, 0);
  for (size_t i = 0; i < mol.size(); ++i) d.nmax = std::max(d.nmax, tabs_host[i]->n);
  for (int slot = 0; slot < HFG_NSLOT; ++slot) {
    const int c = slot % HFG_NCLASS, v = slot / HFG_NCLASS;
    std::vector<int> ids, jp{0}, tp{0};
    for (size_t i = 0; i < mol.size(); ++i) {
      const MolTab& t = *tabs_host[i];
      const int nj = t.job_begin[c + 1] - t.job_begin[c], nt = t.task_begin[c + 1] - t.task_begin[c];
      if (!nt || (int)((t.unit_mask >> c) & 1u) != v) continue;
      ids.push_back(mol[i]);
      jp.push_back(jp.back() + nj);
      tp.push_back(tp.back() + nt);
    }
    d.off[slot] = (int)d.host.size();
    d.count[slot] = (int)ids.size();
    d.njob[slot] = jp.back();
    d.ntask[slot] = tp.back();
    d.host.insert(d.host.end(), ids.begin(), ids.end());
    d.host.insert(d.host.end(), jp.begin(), jp.end());
    d.host.insert(d.host.end(), tp.begin(), tp.end());
  }
  if (d.host.empty()) d.host.push_back(0);
  return d;
}

unsigned default_side_mask() {
  static const unsigned mask = [] {
    unsigned m = 0;
    for (int c = 9; c < HFG_NCLASS; ++c) m |= 1u << c;  // (ds|ds), (dp|ss) .. (dd|dd)
    if (const char* env = std::getenv("HFG_SIDE_CLASSES")) {
      m = 0;
      for (const char* p = env; *p;) {
        char* end = nullptr;
        const long c = std::strtol(p, &end, 10);
        if (end == p) break;
        if (c >= 0 && c < HFG_NCLASS) m |= 1u << c;
        p = *end ? end + 1 : end;
      }
    }
    return m;
  }();
  return mask;
}

bool side_class(int cls) { return cls >= 0 && cls < HFG_NCLASS && ((default_side_mask() >> cls) & 1u); }
// ---------------------------------------------------------------------------
typedef void (*BlockLaunch)(const BlockArgs&, hipStream_t);

static int g_max_blocks = 1024;  // persistent grid: workgroups per launch (set from the CU count)

template <int LA, int LB, int LC, int LD>
void launch_block(const BlockArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_block<LA, LB, LC, LD>), dim3((a.ntask + 63) / 64), dim3(64), 0, s, a);
}

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

Engine::Engine(const hfg_basis_view& b, hipStream_t shared) {
  shells = group_shells(b);
  n = b.n_functions;
  for (auto& s : shells)
    if (s.L > 2) throw HfgError{HFG_UNSUPPORTED, "shell-quartet classes are compiled up to d functions (L<=2)"};
  int ndev = 0;
  HIP_CHECK(hipGetDeviceCount(&ndev));
  if (ndev <= 0) throw HfgError{HFG_DEVICE, "no GPU visible"};
  HIP_CHECK(hipGetDevice(&device));
  if (shared) {
    // A batch member: no streams of its own (every stream of a process maps
    // onto one of a few hardware queues, each of which keeps its own scratch
    // reservation), single-stream builds on the batch's stream.
    stream = shared;
    own_streams = false;
    side_blocks = 0;
  } else {
    HIP_CHECK(hipStreamCreate(&stream));  // default priority
    HIP_CHECK(hipStreamCreate(&side));    // default priority
    HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    if (const char* env = std::getenv("HFG_SIDE_BLOCKS")) side_blocks = std::max(0, std::atoi(env));
  }
  if (const char* env = std::getenv("HFG_GRAPH")) use_graph = std::atoi(env) != 0;
  use_graph = use_graph && own_streams;
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
  const bool trace = std::getenv("HFG_TRACE") != nullptr;
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  auto lap = [&](const char* what) {
    if (!trace) return;
    const auto t1 = clk::now();
    fprintf(stderr, "hfg engine %-10s %8.2f ms\n", what, std::chrono::duration<double>(t1 - t0).count() * 1e3);
    t0 = t1;
  };
  lap("upload");
  compute_schwarz();
  lap("schwarz");
  build_tasks();
  lap("plan");
}

Engine::~Engine() {
  for (auto& g : graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
  }
  for (auto e : events) (void)hipEventDestroy(e);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (own_streams) {
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }
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

PrimPairs Engine::pp_view() const {
  const double* b = d_pp.as<double>();
  const size_t np = pp.size();
  return PrimPairs{b, b + np, b + 2 * np, b + 3 * np, b + 4 * np, b + 5 * np, b + 6 * np, b + 7 * np, b + 8 * np};
}

void Engine::blocks(const std::vector<BlockTask>& tasks, const std::vector<int>& cls, std::vector<double>& out,
                    const PairDev* pair_table, const PrimPairs* prims) {
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
    BlockArgs a{dt.as<BlockTask>() + i, (int)(j - i), pair_table ? pair_table : d_pairs.as<PairDev>(), prims ? *prims : pp_view(),
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

PrimPairs Engine::kt_view() const {
  const double* b = d_kt.as<double>();
  const size_t m = kt_size;
  return PrimPairs{b, b + m, b + 2 * m, b + 3 * m, b + 4 * m, b + 5 * m, b + 6 * m, b + 7 * m, b + 8 * m};
}

// The J/K plan (jk_plan.cpp) with primitive-pair bounds from the GPU block
// kernels, uploaded once per molecule.
void Engine::build_tasks() {
  BlockFn gpu_blocks = [&](const PrimPairHost& prims, const std::vector<PairDev>& tab,
                           const std::vector<BlockTask>& tasks, const std::vector<int>& cls, std::vector<double>& out) {
    const size_t np = prims.size();
    std::vector<double> flat(9 * std::max<size_t>(1, np));
    const std::vector<double>* cols[9] = {&prims.p,   &prims.Px,  &prims.Py, &prims.Pz, &prims.PAx,
                                          &prims.PAy, &prims.PAz, &prims.K,  &prims.H};
    for (int c = 0; c < 9; ++c) std::copy(cols[c]->begin(), cols[c]->end(), flat.begin() + c * np);
    DevBuf dp, dtab;
    dp.alloc(flat.size() * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(dp.p, flat.data(), flat.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    dtab.alloc(std::max<size_t>(1, tab.size()) * sizeof(PairDev));
    HIP_CHECK(hipMemcpyAsync(dtab.p, tab.data(), tab.size() * sizeof(PairDev), hipMemcpyHostToDevice, stream));
    const double* b = dp.as<double>();
    PrimPairs v{b, b + np, b + 2 * np, b + 3 * np, b + 4 * np, b + 5 * np, b + 6 * np, b + 7 * np, b + 8 * np};
    blocks(tasks, cls, out, dtab.as<PairDev>(), &v);
  };
  std::vector<double> Qhf(pairs.size());
  for (size_t i = 0; i < pairs.size(); ++i) Qhf[pairs[i].hf_index] = Q[i];
  plan = make_jk_plan(shells, Qhf, ao_scale, screen, prim_screen, gpu_blocks);
  auto up_load = [&](DevBuf& d, const void* src, size_t bytes) {
    d.alloc(std::max<size_t>(bytes, 16));
    if (bytes) HIP_CHECK(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, stream));
  };
  up_load(d_vb, plan.vb.data(), plan.vb.size() * sizeof(double));
  up_load(d_kt, plan.kt.data(), plan.kt.size() * sizeof(double));
  up_load(d_kcoef, plan.kcoef.data(), plan.kcoef.size() * sizeof(double));
  up_load(d_groups, plan.groups.data(), plan.groups.size() * sizeof(KetGroup));
  up_load(d_views, plan.views.data(), plan.views.size() * sizeof(ViewDev));
  up_load(d_vq, plan.view_q.data(), plan.view_q.size() * sizeof(double));
  up_load(d_combos, plan.combos.data(), plan.combos.size() * sizeof(ComboDev));
  up_load(d_tasks, plan.tasks.data(), plan.tasks.size() * sizeof(JKTask));
  up_load(d_jobs, plan.jobs.data(), plan.jobs.size() * sizeof(int));
  d_counters.alloc(HFG_NSLOT * sizeof(int));
  job_begin = plan.job_begin;
  if (const char* env = std::getenv("HFG_JK_ROWS")) use_rows = std::atoi(env) != 0;
  vb_size = plan.vb_size;
  kt_size = plan.kt_size;
  task_begin = plan.task_begin;
  prim_pairs_kept = plan.prim_pairs_kept;
  n_gen_shells = plan.n_gen_shells;
  n_views = (int)plan.views.size();
  std::memset(&tab_host, 0, sizeof tab_host);
  tab_host.tasks = d_tasks.as<JKTask>();
  tab_host.views = d_views.as<ViewDev>();
  tab_host.combos = d_combos.as<ComboDev>();
  tab_host.groups = d_groups.as<KetGroup>();
  tab_host.brec = d_vb.as<double>();
  tab_host.kt = d_kt.as<double>();
  tab_host.kcoef = d_kcoef.as<double>();
  tab_host.vq = d_vq.as<double>();
  tab_host.scale = d_scale.as<double>();
  tab_host.D = d_D.as<double>();
  tab_host.Jacc = d_Jacc.as<double>();
  tab_host.Kacc = d_Kacc.as<double>();
  tab_host.jobs = d_jobs.as<int>();
  tab_host.kt_size = kt_size;
  tab_host.nn = (size_t)n * n;
  tab_host.screen = screen;
  tab_host.n = n;
  tab_host.unit_mask = plan.unit_mask;
  for (int c = 0; c <= HFG_NCLASS; ++c) {
    tab_host.task_begin[c] = (int)task_begin[c];
    tab_host.job_begin[c] = (int)job_begin[c];
  }
  d_tab.alloc(sizeof(MolTab));
  HIP_CHECK(hipMemcpyAsync(d_tab.p, &tab_host, sizeof(MolTab), hipMemcpyHostToDevice, stream));
  dir = make_directory({0}, {&tab_host});
  d_dir.alloc(dir.host.size() * sizeof(int));
  HIP_CHECK(hipMemcpyAsync(d_dir.p, dir.host.data(), dir.host.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  sync();
  count_work();
}

// Algorithmic work of one Fock build as the reference's shell engine performs
// it: every unique segmented quartet with Q_ab * Q_cd > screen, all of its
// primitive quartets, priced by md_flops.
void Engine::count_work() {
  std::vector<std::vector<int>> bypc(6);
  for (size_t i = 0; i < pairs.size(); ++i) bypc[pair_class(pairs[i].la, pairs[i].lb)].push_back((int)i);
  std::vector<std::vector<double>> qs(6), kpre(6);
  for (int pc = 0; pc < 6; ++pc) {
    auto& v = bypc[pc];
    std::stable_sort(v.begin(), v.end(), [&](int x, int y) { return Q[x] > Q[y]; });
    kpre[pc].assign(1, 0.0);
    for (int i : v) {
      qs[pc].push_back(Q[i]);
      kpre[pc].push_back(kpre[pc].back() + (pairs[i].p1 - pairs[i].p0));
    }
  }
  n_quartets = 0;
  prim_quartets = 0;
  class_flops.assign(HFG_NCLASS, 0.0);
  class_quartets.assign(HFG_NCLASS, 0);
  class_ms.assign(HFG_NCLASS, 0.0);
  class_launches.assign(HFG_NCLASS, 0);
  for (int c = 0; c < HFG_NCLASS; ++c) {
    const int* l = CLASS_L[c];
    const int pb = pair_class(l[0], l[1]), pk = pair_class(l[2], l[3]);
    const double f = md_flops(l[0], l[1], l[2], l[3]);
    for (size_t i = 0; i < bypc[pb].size(); ++i) {
      const double qb = qs[pb][i];
      size_t lo = 0, hi = qs[pk].size();
      while (lo < hi) {
        const size_t mid = (lo + hi) / 2;
        if (qb * qs[pk][mid] > screen)
          lo = mid + 1;
        else
          hi = mid;
      }
      size_t nk = lo;
      if (pb == pk) nk = std::min(nk, i + 1);
      const double pq = (kpre[pb][i + 1] - kpre[pb][i]) * kpre[pk][nk];
      n_quartets += (int64_t)nk;
      class_quartets[c] += (int64_t)nk;
      prim_quartets += pq;
      class_flops[c] += pq * f;
    }
  }
}

void launch_jk_set(const JKLaunch& L, const MolTab* d_tabs, const JKDirectory& dr, const int* d_dr, hipStream_t main,
                   hipStream_t side_stream, int side_cap) {
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, main, L.counters, HFG_NSLOT);
  // The high-L classes, few and long quartets, run on the side stream with a
  // capped grid and fill the room the low-L classes leave.
  auto on_side = [&](int c) { return side_cap > 0 && side_stream && ((L.side_mask >> c) & 1u); };
  const bool two = side_cap > 0 && side_stream;
  if (two) {
    HIP_CHECK(hipEventRecord(L.fork, main));
    HIP_CHECK(hipStreamWaitEvent(side_stream, L.fork, 0));
  }
  JKArgs a;
  std::memset(&a, 0, sizeof a);
  a.tabs = d_tabs;
  a.nmax = dr.nmax;
  a.screen0 = L.screen0;
  a.boys = L.boys;
  a.use_rows = L.use_rows ? 1 : 0;
  // A class's two variants run back to back on one stream; its event pair
  // brackets both.
  for (int pass = 0; pass < 2; ++pass)
    for (int c = 0; c < HFG_NCLASS; ++c) {
      if ((!dr.count[c] && !dr.count[c + HFG_NCLASS]) || on_side(c) != (pass == 0)) continue;
      hipStream_t s = on_side(c) ? side_stream : main;
      if (L.events) HIP_CHECK(hipEventRecord(L.events[2 * c], s));
      for (int v = 0; v < 2; ++v) {
        const int slot = c + v * HFG_NCLASS;
        if (!dr.count[slot]) continue;
        const int* base = d_dr + dr.off[slot];
        a.ndir = dr.count[slot];
        a.dir_mol = base;
        a.dir_job = base + dr.count[slot];
        a.dir_task = base + 2 * dr.count[slot] + 1;
        a.njob = dr.njob[slot];
        a.ntask = dr.ntask[slot];
        a.job_counter = L.counters + slot;
        launch_jk_class(c, v, a, on_side(c) ? side_cap : (L.max_blocks > 0 ? L.max_blocks : g_max_blocks), s);
        HIP_CHECK(hipGetLastError());
      }
      if (L.events) HIP_CHECK(hipEventRecord(L.events[2 * c + 1], s));
    }
  if (two) {
    HIP_CHECK(hipEventRecord(L.join, side_stream));
    HIP_CHECK(hipStreamWaitEvent(main, L.join, 0));
  }
}

JKLaunch Engine::launch_ctx() {
  return JKLaunch{d_counters.as<int>(), ev_fork, ev_join, d_boys.as<double>(), screen, use_rows,
                  profile ? events.data() : nullptr, default_side_mask(), 0};
}

void Engine::jk_device(const double* dP, double* dJ, double* dK, const double* dH, double* dF, double dmax) {
  const size_t nn = (size_t)n * n;
  const int tb = 256;
  const size_t nacc = nn * NXCC_SLOTS;
  hipLaunchKernelGGL(k_tab_set, dim3(1), dim3(64), 0, stream, d_tab.as<MolTab>(), dP,
                     dmax > 0.0 && dmax < 1.0 ? screen / dmax : screen);
  auto launches = [&] {
    hipLaunchKernelGGL(k_fill, dim3((nacc + tb - 1) / tb), dim3(tb), 0, stream, d_Jacc.as<double>(), nacc, 0.0);
    hipLaunchKernelGGL(k_fill, dim3((nacc + tb - 1) / tb), dim3(tb), 0, stream, d_Kacc.as<double>(), nacc, 0.0);
    launch_jk_set(launch_ctx(), d_tab.as<MolTab>(), dir, d_dir.as<int>(), stream, side, side_blocks);
    hipLaunchKernelGGL(k_symmetrize, dim3((nn + tb - 1) / tb), dim3(tb), 0, stream, d_Jacc.as<double>(),
                       d_Kacc.as<double>(), dJ, dK, n, dH, dF);
  };
  JKGraph* g = nullptr;
  if (use_graph && !profile) {
    const void* key[4] = {dJ, dK, dH, dF};
    for (auto& x : graphs)
      if (std::equal(key, key + 4, x.key)) g = &x;
    if (!g && graphs.size() < 8) {
      graphs.emplace_back();
      g = &graphs.back();
      std::copy(key, key + 4, g->key);
    }
  }
  if (g && g->exec) {
    HIP_CHECK(hipGraphLaunch(g->exec, stream));
    ++graph_launches;
  } else if (g && g->seen >= 1) {
    // second build with these outputs: capture it (the first ran eagerly, so
    // the launch helpers' one-time attribute and occupancy queries are done)
    HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      launches();
    } catch (...) {
      hipGraph_t partial = nullptr;
      (void)hipStreamEndCapture(stream, &partial);  // leave the stream usable
      if (partial) (void)hipGraphDestroy(partial);
      use_graph = false;
      throw;
    }
    HIP_CHECK(hipStreamEndCapture(stream, &g->graph));
    HIP_CHECK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    HIP_CHECK(hipGraphLaunch(g->exec, stream));
    ++graph_launches;
  } else {
    launches();
    if (g) ++g->seen;
  }
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
