// batch.hip -- Fock builds of many molecules in fused launches.
//
// Each molecule keeps its own engine (shell pairs, Schwarz bounds, J/K plan,
// density and accumulators in HBM).  The batch concatenates the molecules'
// plans into one view table, one ket-group table and, per quartet class, one
// task and job list sorted by cost across molecules; every view carries its
// molecule's index, through which the J/K kernels reach that molecule's
// density, accumulators, AO factors and size (MolDev).  One build of the whole
// batch is 21 class launches (two streams), one clearing and one
// symmetrisation launch, whatever the number of molecules: quartets of uneven
// classes from many small molecules fill the chip together, where molecule by
// molecule each launch would be nearly empty.  The reference builds each
// molecule's Fock matrix on its own (src/scf.rs:366-369, via
// src/shell_integrals.rs:156-240 and src/scf.rs:337-364); the results per
// molecule are the same.
#include <algorithm>
#include <cstring>
#include <memory>

#include "../../include/hfg.h"
#include "engine.h"

namespace hfg {

__global__ void k_batch_clear(const MolDev* __restrict__ mols, int count) {
  const int m = blockIdx.y;
  if (m >= count) return;
  const MolDev md = mols[m];
  const size_t len = md.nn * NXCC_SLOTS;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
    md.Jacc[i] = 0.0;
    md.Kacc[i] = 0.0;
  }
}

// J = sum over the XCC copies of Jacc + Jacc^T, K likewise, per molecule.
__global__ void k_batch_symmetrize(const MolDev* __restrict__ mols, double* const* __restrict__ Jout,
                                   double* const* __restrict__ Kout, int count) {
  const int m = blockIdx.y;
  if (m >= count) return;
  const MolDev md = mols[m];
  const int n = md.n;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < md.nn;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / n), c = (int)(idx % n);
    const size_t tr = (size_t)c * n + r;
    double j = 0.0, k = 0.0;
    for (int x = 0; x < NXCC_SLOTS; ++x) {
      j += md.Jacc[x * md.nn + idx] + md.Jacc[x * md.nn + tr];
      k += md.Kacc[x * md.nn + idx] + md.Kacc[x * md.nn + tr];
    }
    Jout[m][idx] = j;
    Kout[m][idx] = k;
  }
}

struct Batch {
  std::vector<std::unique_ptr<Engine>> eng;
  hipStream_t stream = nullptr, side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  DevBuf d_views, d_combos, d_groups, d_vb, d_kt, d_kcoef, d_vq, d_tasks, d_jobs, d_counters, d_mols, d_jout, d_kout,
      d_boys;
  std::vector<size_t> task_begin, job_begin;
  size_t vb_size = 0, kt_size = 0;
  int nmax = 0;
  int side_blocks = 512;
  double flops = 0;
  int64_t quartets = 0;

  explicit Batch(const hfg_basis_view* bases, int count);
  ~Batch();
  void jk();
};

Batch::Batch(const hfg_basis_view* bases, int count) {
  if (count <= 0) throw HfgError{HFG_INVALID_INPUT, "empty batch"};
  for (int m = 0; m < count; ++m) {
    eng.emplace_back(new Engine(bases[m]));
    nmax = std::max(nmax, eng.back()->n);
    for (double f : eng.back()->class_flops) flops += f;
    quartets += eng.back()->n_quartets;
  }
  HIP_CHECK(hipStreamCreate(&stream));  // default priority
  HIP_CHECK(hipStreamCreate(&side));    // default priority
  HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  if (const char* env = std::getenv("HFG_SIDE_BLOCKS")) side_blocks = std::max(0, std::atoi(env));

  // concatenated tables
  std::vector<ViewDev> views;
  std::vector<ComboDev> combos;
  std::vector<KetGroup> groups;
  std::vector<double> brec, kcoef, vq;
  std::vector<std::vector<double>> kcols(9);
  std::vector<int> vo(count), go(count);
  for (int m = 0; m < count; ++m) {
    const JKPlan& P = eng[m]->plan;
    vo[m] = (int)views.size();
    go[m] = (int)groups.size();
    const int co = (int)combos.size(), bo = (int)(brec.size() / HFG_BRA_REC), ko = (int)kcols[0].size(),
              cfo = (int)kcoef.size();
    for (size_t i = 0; i < P.views.size(); ++i) {
      ViewDev v = P.views[i];
      v.p0 += bo;
      v.p1 += bo;
      v.c0 += co;
      v.mol = m;
      views.push_back(v);
      vq.push_back(P.view_q[i]);
    }
    combos.insert(combos.end(), P.combos.begin(), P.combos.end());
    for (KetGroup g : P.groups) {
      g.start += vo[m];
      g.base += ko;
      g.cbase += cfo;
      groups.push_back(g);
    }
    brec.insert(brec.end(), P.vb.begin(), P.vb.begin() + P.vb_size * HFG_BRA_REC);
    for (int c = 0; c < 9; ++c)
      kcols[c].insert(kcols[c].end(), P.kt.begin() + c * P.kt_size, P.kt.begin() + (c + 1) * P.kt_size);
    kcoef.insert(kcoef.end(), P.kcoef.begin(), P.kcoef.end());
  }
  vb_size = brec.size() / HFG_BRA_REC;
  kt_size = kcols[0].size();
  std::vector<double> kt(9 * std::max<size_t>(1, kt_size), 0.0);
  for (int c = 0; c < 9; ++c) std::copy(kcols[c].begin(), kcols[c].end(), kt.begin() + c * kt_size);
  if (brec.empty()) brec.assign(HFG_BRA_REC, 0.0);
  if (kcoef.empty()) kcoef.push_back(0.0);

  // per class: the molecules' jobs merged by cost, their tasks renumbered
  std::vector<JKTask> tasks;
  std::vector<int> jobs;
  task_begin.assign(HFG_NCLASS + 1, 0);
  job_begin.assign(HFG_NCLASS + 1, 0);
  for (int c = 0; c < HFG_NCLASS; ++c) {
    task_begin[c] = tasks.size();
    job_begin[c] = jobs.size();
    struct J {
      double cost;
      int m, j;
    };
    std::vector<J> js;
    for (int m = 0; m < count; ++m) {
      const JKPlan& P = eng[m]->plan;
      for (size_t j = P.job_begin[c]; j < P.job_begin[c + 1]; ++j) js.push_back({P.job_cost[j], m, (int)j});
    }
    std::stable_sort(js.begin(), js.end(), [](const J& x, const J& y) { return x.cost > y.cost; });
    for (const J& x : js) {
      const JKPlan& P = eng[x.m]->plan;
      jobs.push_back((int)tasks.size());
      for (int t = P.jobs[x.j]; t < P.jobs[x.j + 1]; ++t) {
        JKTask k = P.tasks[t];
        k.bra += vo[x.m];
        k.ket0 += vo[x.m];
        k.grp += go[x.m];
        tasks.push_back(k);
      }
    }
  }
  task_begin[HFG_NCLASS] = tasks.size();
  job_begin[HFG_NCLASS] = jobs.size();
  jobs.push_back((int)tasks.size());

  std::vector<MolDev> mols(count);
  std::vector<double*> jout(count), kout(count);
  for (int m = 0; m < count; ++m) {
    Engine& e = *eng[m];
    mols[m] = MolDev{e.d_D.as<double>(), e.d_Jacc.as<double>(), e.d_Kacc.as<double>(), e.d_scale.as<double>(),
                     (size_t)e.n * e.n, e.n, 0};
    jout[m] = e.d_J.as<double>();
    kout[m] = e.d_K.as<double>();
  }
  auto up = [&](DevBuf& d, const void* src, size_t bytes) {
    d.alloc(std::max<size_t>(bytes, 16));
    if (bytes) HIP_CHECK(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, stream));
  };
  up(d_views, views.data(), views.size() * sizeof(ViewDev));
  up(d_combos, combos.data(), combos.size() * sizeof(ComboDev));
  up(d_groups, groups.data(), groups.size() * sizeof(KetGroup));
  up(d_vb, brec.data(), brec.size() * sizeof(double));
  up(d_kt, kt.data(), kt.size() * sizeof(double));
  up(d_kcoef, kcoef.data(), kcoef.size() * sizeof(double));
  up(d_vq, vq.data(), vq.size() * sizeof(double));
  up(d_tasks, tasks.data(), tasks.size() * sizeof(JKTask));
  up(d_jobs, jobs.data(), jobs.size() * sizeof(int));
  up(d_mols, mols.data(), mols.size() * sizeof(MolDev));
  up(d_jout, jout.data(), jout.size() * sizeof(double*));
  up(d_kout, kout.data(), kout.size() * sizeof(double*));
  d_counters.alloc(HFG_NCLASS * sizeof(int));
  HIP_CHECK(hipStreamSynchronize(stream));
}

Batch::~Batch() {
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
  if (stream) (void)hipStreamDestroy(stream);
}

__global__ void k_zero_int(int* p, int n);

void Batch::jk() {
  const int count = (int)eng.size();
  // every molecule's density is staged by its engine's stream; order it first
  for (auto& e : eng) HIP_CHECK(hipStreamSynchronize(e->stream));
  hipLaunchKernelGGL(k_batch_clear, dim3(64, count), dim3(256), 0, stream, d_mols.as<MolDev>(), count);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, stream, d_counters.as<int>(), HFG_NCLASS);
  JKArgs a;
  std::memset(&a, 0, sizeof a);
  a.n = nmax;  // sizes the LDS rows; everything else per molecule
  a.views = d_views.as<ViewDev>();
  a.combos = d_combos.as<ComboDev>();
  a.groups = d_groups.as<KetGroup>();
  a.brec = d_vb.as<double>();
  const double* kb = d_kt.as<double>();
  const size_t m = kt_size;
  a.kt = PrimPairs{kb, kb + m, kb + 2 * m, kb + 3 * m, kb + 4 * m, kb + 5 * m, kb + 6 * m, kb + 7 * m, kb + 8 * m};
  a.kcoef = d_kcoef.as<double>();
  a.screen = a.screen0 = eng[0]->screen;
  a.vq = d_vq.as<double>();
  a.boys = eng[0]->d_boys.as<double>();
  a.mols = d_mols.as<MolDev>();
  a.tasks = d_tasks.as<JKTask>();
  auto on_side = [&](int c) { return side_blocks > 0 && side_class(c); };
  const bool two = side_blocks > 0;
  if (two) {
    HIP_CHECK(hipEventRecord(ev_fork, stream));
    HIP_CHECK(hipStreamWaitEvent(side, ev_fork, 0));
  }
  for (int pass = 0; pass < 2; ++pass)
    for (int c = 0; c < HFG_NCLASS; ++c) {
      const size_t nt = task_begin[c + 1] - task_begin[c];
      if (!nt || on_side(c) != (pass == 0)) continue;
      a.task0 = (int)task_begin[c];
      a.jobs = d_jobs.as<int>() + job_begin[c];
      a.njob = (int)(job_begin[c + 1] - job_begin[c]);
      a.job_counter = d_counters.as<int>() + c;
      launch_jk_class(c, a, (int)nt, on_side(c) ? side_blocks : 4 * 256, on_side(c) ? side : stream);
      HIP_CHECK(hipGetLastError());
    }
  if (two) {
    HIP_CHECK(hipEventRecord(ev_join, side));
    HIP_CHECK(hipStreamWaitEvent(stream, ev_join, 0));
  }
  hipLaunchKernelGGL(k_batch_symmetrize, dim3(64, count), dim3(256), 0, stream, d_mols.as<MolDev>(),
                     d_jout.as<double*>(), d_kout.as<double*>(), count);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(stream));
}

}  // namespace hfg

using namespace hfg;

struct hfg_batch {
  std::unique_ptr<Batch> b;
};

#define BGUARD(...)                                        \
  try {                                                    \
    __VA_ARGS__                                            \
  } catch (HfgError & ex) {                                \
    return set_error(ex.code, ex.msg);                     \
  } catch (std::bad_alloc&) {                              \
    return set_error(HFG_DEVICE, "host allocation failed"); \
  } catch (std::exception & ex) {                          \
    return set_error(HFG_INTEGRAL, ex.what());             \
  }

int hfg_batch_create(const hfg_basis_view* bases, int count, hfg_batch** out) {
  if (!bases || !out || count <= 0) return set_error(HFG_INVALID_INPUT, "null argument or empty batch");
  BGUARD({
    auto* h = new hfg_batch;
    try {
      h->b.reset(new Batch(bases, count));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return HFG_OK;
  })
}

void hfg_batch_destroy(hfg_batch* b) { delete b; }

int hfg_batch_info(const hfg_batch* b, int64_t* count, int64_t* n_max, int64_t* quartets, double* flops) {
  if (!b) return set_error(HFG_INVALID_INPUT, "null batch");
  if (count) *count = (int64_t)b->b->eng.size();
  if (n_max) *n_max = b->b->nmax;
  if (quartets) *quartets = b->b->quartets;
  if (flops) *flops = b->b->flops;
  return HFG_OK;
}

int hfg_batch_set_density(hfg_batch* b, int mol, const double* p) {
  if (!b || !p) return set_error(HFG_INVALID_INPUT, "null argument");
  if (mol < 0 || mol >= (int)b->b->eng.size()) return set_error(HFG_INVALID_INPUT, "molecule index out of range");
  BGUARD({
    Engine& e = *b->b->eng[mol];
    HIP_CHECK(hipMemcpyAsync(e.d_D.p, p, (size_t)e.n * e.n * 8, hipMemcpyHostToDevice, e.stream));
    return HFG_OK;
  })
}

int hfg_batch_jk(hfg_batch* b) {
  if (!b) return set_error(HFG_INVALID_INPUT, "null batch");
  BGUARD({
    b->b->jk();
    return HFG_OK;
  })
}

int hfg_batch_jk_each(hfg_batch* b) {
  if (!b) return set_error(HFG_INVALID_INPUT, "null batch");
  BGUARD({
    for (auto& e : b->b->eng) e->jk_device(e->d_D.as<double>(), e->d_J.as<double>(), e->d_K.as<double>());
    for (auto& e : b->b->eng) e->sync();
    return HFG_OK;
  })
}

int hfg_batch_get_jk(hfg_batch* b, int mol, double* j, double* k) {
  if (!b) return set_error(HFG_INVALID_INPUT, "null batch");
  if (mol < 0 || mol >= (int)b->b->eng.size()) return set_error(HFG_INVALID_INPUT, "molecule index out of range");
  BGUARD({
    Engine& e = *b->b->eng[mol];
    const size_t bytes = (size_t)e.n * e.n * 8;
    if (j) HIP_CHECK(hipMemcpyAsync(j, e.d_J.p, bytes, hipMemcpyDeviceToHost, b->b->stream));
    if (k) HIP_CHECK(hipMemcpyAsync(k, e.d_K.p, bytes, hipMemcpyDeviceToHost, b->b->stream));
    HIP_CHECK(hipStreamSynchronize(b->b->stream));
    return HFG_OK;
  })
}
