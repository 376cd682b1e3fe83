// batch.hip -- the resident molecule set (fleet.h) and the hfg_batch_* entry
// points: Fock builds of many molecules in fused launches.
#include <algorithm>
#include <cstring>
#include <cstdio>
#include <memory>

#include "../../include/hfg.h"
#include "fleet.h"

namespace hfg {

__global__ void k_zero_int(int* p, int n);

// Clears the per-XCC accumulator copies of the active slots (grid (x, count)).
__global__ void k_fleet_clear(const MolTab* __restrict__ tabs, const int* __restrict__ slots, int count) {
  const int m = blockIdx.y;
  if (m >= count) return;
  const MolTab& t = tabs[slots[m]];
  double* J = t.Jacc;
  double* K = t.Kacc;
  const size_t len = t.nn * NXCC_SLOTS;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (size_t)gridDim.x * blockDim.x) {
    J[i] = 0.0;
    K[i] = 0.0;
  }
}

__global__ void k_fleet_jk_out(const MolTab* __restrict__ tabs, const int* __restrict__ slots,
                               double* const* __restrict__ jout, double* const* __restrict__ kout, int count) {
  const int m = blockIdx.y;
  if (m >= count) return;
  const MolTab& t = tabs[slots[m]];
  const int n = t.n;
  const size_t nn = t.nn;
  double* J = jout ? jout[m] : nullptr;
  double* K = kout ? kout[m] : nullptr;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(idx / n), c = (int)(idx % n);
    const size_t tr = (size_t)c * n + r;
    double j = 0.0, k = 0.0;
    for (int x = 0; x < NXCC_SLOTS; ++x) {
      j += t.Jacc[x * nn + idx] + t.Jacc[x * nn + tr];
      k += t.Kacc[x * nn + idx] + t.Kacc[x * nn + tr];
    }
    if (J) J[idx] = j;
    if (K) K[idx] = k;
  }
}

Fleet::Fleet(int cap) : capacity(cap) {
  if (cap <= 0) throw HfgError{HFG_INVALID_INPUT, "a fleet needs at least one slot"};
  HIP_CHECK(hipStreamCreate(&stream));  // default priority
  HIP_CHECK(hipStreamCreate(&side));    // default priority
  HIP_CHECK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
  HIP_CHECK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
  if (const char* env = std::getenv("HFG_SIDE_BLOCKS")) side_blocks = std::max(0, std::atoi(env));
  if (const char* env = std::getenv("HFG_JK_ROWS")) use_rows = std::atoi(env) != 0;
  tabs.assign(cap, MolTab{});
  occupied.assign(cap, 0);
  d_tabs.alloc(cap * sizeof(MolTab));
  d_active.alloc(cap * sizeof(int));
  d_counters.alloc(HFG_NSLOT * sizeof(int));
  auto tab = boys_table();
  d_boys.alloc(tab.size() * sizeof(double));
  HIP_CHECK(hipMemcpyAsync(d_boys.p, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, stream));
  slot_quartets.assign(cap, 0.0);
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // Candidates: the default split (the one measured best on (H2O)16/cc-pVDZ),
  // one stream, only the dp/dd-bra classes aside, the default split with a
  // larger side grid, only the dd-bra classes aside.
  const unsigned def = default_side_mask();
  unsigned dpdd = 0, dd = 0;
  for (int c = 10; c < HFG_NCLASS; ++c) dpdd |= 1u << c;
  for (int c = 15; c < HFG_NCLASS; ++c) dd |= 1u << c;
  candidates = {{def, side_blocks, 4}, {def, 0, 4}, {dpdd, side_blocks, 4}, {def, 2 * side_blocks, 4}, {dd, side_blocks / 2, 4}};
  score.assign(candidates.size(), -1.0);
  const char* at = std::getenv("HFG_AUTOTUNE");
  tuned = side_blocks == 0 || (at && std::atoi(at) == 0);
  HIP_CHECK(hipEventCreate(&tb));
  HIP_CHECK(hipEventCreate(&te));
  sync();
}

Fleet::~Fleet() {
  for (auto e : events) (void)hipEventDestroy(e);
  if (tb) (void)hipEventDestroy(tb);
  if (te) (void)hipEventDestroy(te);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (side) (void)hipStreamDestroy(side);
  if (stream) (void)hipStreamDestroy(stream);
}

std::unique_ptr<Engine> Fleet::make_engine(const hfg_basis_view& b) {
  return std::unique_ptr<Engine>(new Engine(b, stream));
}

void Fleet::attach(int s, const Engine& e, hipStream_t st) {
  if (s < 0 || s >= capacity) throw HfgError{HFG_INVALID_INPUT, "fleet slot out of range"};
  if (!st) st = stream;
  tabs[s] = e.tab_host;
  occupied[s] = 1;
  slot_quartets[s] = (double)e.n_quartets;
  HIP_CHECK(hipMemcpyAsync(d_tabs.as<MolTab>() + s, &tabs[s], sizeof(MolTab), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipStreamSynchronize(st));  // the host copy may change before the next build
}

void Fleet::detach(int s) {
  if (s >= 0 && s < capacity) occupied[s] = 0;
}

void Fleet::set_active(const std::vector<int>& slots) {
  std::vector<const MolTab*> th;
  for (int s : slots) {
    if (s < 0 || s >= capacity || !occupied[s]) throw HfgError{HFG_INVALID_INPUT, "inactive fleet slot"};
    th.push_back(&tabs[s]);
  }
  active = slots;
  dir = make_directory(slots, th);
  if (d_dir.bytes < dir.host.size() * sizeof(int)) d_dir.alloc(std::max<size_t>(dir.host.size(), 1024) * sizeof(int));
  HIP_CHECK(hipMemcpyAsync(d_dir.p, dir.host.data(), dir.host.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  if (!slots.empty())
    HIP_CHECK(hipMemcpyAsync(d_active.p, slots.data(), slots.size() * sizeof(int), hipMemcpyHostToDevice, stream));
  HIP_CHECK(hipStreamSynchronize(stream));  // host vectors are reused by the next call
}

void Fleet::tune_step(bool full_build) {
  if (pending >= 0) {
    // the previous build's time (its events are complete once the stream
    // has run past them)
    HIP_CHECK(hipEventSynchronize(te));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, tb, te));
    if (pending_work > 0) score[pending] = ms / pending_work;
    pending = -1;
    size_t next = 0;
    while (next < score.size() && score[next] >= 0) ++next;
    if (next == score.size()) {
      current = (int)(std::min_element(score.begin(), score.end()) - score.begin());
      tuned = true;
      if (std::getenv("HFG_TRACE"))
        for (size_t i = 0; i < score.size(); ++i)
          fprintf(stderr, "hfg autotune: candidate %zu (side mask %#x, cap %d): %.3e ms/quartet%s\n", i,
                  candidates[i].side_mask, candidates[i].side_cap, score[i], (int)i == current ? "  <- chosen" : "");
    } else {
      current = (int)next;
    }
  }
  if (!tuned && full_build) {
    pending = current;
    pending_work = 0;
    for (int s : active) pending_work += slot_quartets[s];
  }
}

void Fleet::build(bool full_build) {
  const int count = (int)active.size();
  if (!count) return;
  if (!tuned) tune_step(full_build);
  hipLaunchKernelGGL(k_fleet_clear, dim3(64, count), dim3(256), 0, stream, d_tabs.as<MolTab>(), d_active.as<int>(),
                     count);
  const JKConfig& cfg = candidates[pending >= 0 ? pending : current];
  JKLaunch L{d_counters.as<int>(), ev_fork, ev_join, d_boys.as<double>(), screen0, use_rows,
             profile ? events.data() : nullptr, cfg.side_mask, cfg.blocks_per_cu * cus};
  if (pending >= 0) HIP_CHECK(hipEventRecord(tb, stream));
  launch_jk_set(L, d_tabs.as<MolTab>(), dir, d_dir.as<int>(), stream, side, profile && exclusive ? 0 : cfg.side_cap);
  if (pending >= 0) HIP_CHECK(hipEventRecord(te, stream));
  if (profile) {
    sync();
    for (int c = 0; c < HFG_NCLASS; ++c) {
      if (!dir.count[c] && !dir.count[c + HFG_NCLASS]) continue;
      float ms = 0;
      HIP_CHECK(hipEventElapsedTime(&ms, events[2 * c], events[2 * c + 1]));
      class_ms[c] += ms;
      class_launches[c] += 1;
    }
  }
}

void Fleet::set_profile(bool on, bool exclusive_) {
  exclusive = on && exclusive_;
  if (on && events.empty()) {
    events.resize(2 * HFG_NCLASS);
    for (auto& e : events) HIP_CHECK(hipEventCreate(&e));
  }
  if (on) {
    class_ms.assign(HFG_NCLASS, 0.0);
    class_launches.assign(HFG_NCLASS, 0);
  }
  profile = on;
}

void Fleet::build_jk(double* const* d_jout, double* const* d_kout) {
  build();
  const int count = (int)active.size();
  if (!count) return;
  hipLaunchKernelGGL(k_fleet_jk_out, dim3(64, count), dim3(256), 0, stream, d_tabs.as<MolTab>(), d_active.as<int>(),
                     d_jout, d_kout, count);
  HIP_CHECK(hipGetLastError());
}

void Fleet::sync() { HIP_CHECK(hipStreamSynchronize(stream)); }

// The hfg_batch_* entry points: a fleet holding every molecule, all active.
struct Batch {
  std::unique_ptr<Fleet> fleet;
  std::vector<std::unique_ptr<Engine>> eng;
  DevBuf d_jout, d_kout;
  int nmax = 0;
  double flops = 0;
  int64_t quartets = 0;

  Batch(const hfg_basis_view* bases, int count) {
    if (count <= 0) throw HfgError{HFG_INVALID_INPUT, "empty batch"};
    fleet.reset(new Fleet(count));
    std::vector<int> slots;
    std::vector<double*> jout, kout;
    for (int m = 0; m < count; ++m) {
      eng.push_back(fleet->make_engine(bases[m]));
      Engine& e = *eng.back();
      fleet->attach(m, e);
      slots.push_back(m);
      jout.push_back(e.d_J.as<double>());
      kout.push_back(e.d_K.as<double>());
      nmax = std::max(nmax, e.n);
      for (double f : e.class_flops) flops += f;
      quartets += e.n_quartets;
    }
    fleet->set_active(slots);
    d_jout.alloc(count * sizeof(double*));
    d_kout.alloc(count * sizeof(double*));
    HIP_CHECK(hipMemcpyAsync(d_jout.p, jout.data(), count * sizeof(double*), hipMemcpyHostToDevice, fleet->stream));
    HIP_CHECK(hipMemcpyAsync(d_kout.p, kout.data(), count * sizeof(double*), hipMemcpyHostToDevice, fleet->stream));
    fleet->sync();
  }
  ~Batch() {
    eng.clear();  // engines use the fleet's stream: release them first
    fleet.reset();
  }
  void jk() {
    fleet->build_jk(d_jout.as<double*>(), d_kout.as<double*>());
    fleet->sync();
  }
};

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

int hfg_batch_profile(hfg_batch* b, int enable) {
  if (!b) return set_error(HFG_INVALID_INPUT, "null batch");
  BGUARD({
    b->b->fleet->set_profile(enable != 0, enable == 2);
    return HFG_OK;
  })
}

int hfg_batch_class_stats(hfg_batch* b, int cls, int32_t* l4, double* flops, int64_t* quartets, double* ms,
                          int64_t* launches) {
  if (!b || cls < 0 || cls >= HFG_NCLASS) return set_error(HFG_INVALID_INPUT, "class index out of range");
  Batch& B = *b->b;
  int l[4];
  class_angular(cls, l);
  if (l4)
    for (int k = 0; k < 4; ++k) l4[k] = l[k];
  double f = 0;
  int64_t q = 0;
  for (auto& e : B.eng) {
    f += e->class_flops[cls];
    q += e->class_quartets[cls];
  }
  if (flops) *flops = f;
  if (quartets) *quartets = q;
  const Fleet& F = *B.fleet;
  if (ms) *ms = F.class_ms.empty() ? 0.0 : F.class_ms[cls];
  if (launches) *launches = F.class_launches.empty() ? 0 : F.class_launches[cls];
  return HFG_OK;
}

int hfg_batch_set_density(hfg_batch* b, int mol, const double* p) {
  if (!b || !p) return set_error(HFG_INVALID_INPUT, "null argument");
  if (mol < 0 || mol >= (int)b->b->eng.size()) return set_error(HFG_INVALID_INPUT, "molecule index out of range");
  BGUARD({
    Engine& e = *b->b->eng[mol];
    HIP_CHECK(hipMemcpyAsync(e.d_D.p, p, (size_t)e.n * e.n * 8, hipMemcpyHostToDevice, e.stream));
    HIP_CHECK(hipStreamSynchronize(e.stream));
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
    b->b->fleet->sync();
    return HFG_OK;
  })
}

int hfg_batch_get_jk(hfg_batch* b, int mol, double* j, double* k) {
  if (!b) return set_error(HFG_INVALID_INPUT, "null batch");
  if (mol < 0 || mol >= (int)b->b->eng.size()) return set_error(HFG_INVALID_INPUT, "molecule index out of range");
  BGUARD({
    Engine& e = *b->b->eng[mol];
    const size_t bytes = (size_t)e.n * e.n * 8;
    if (j) HIP_CHECK(hipMemcpyAsync(j, e.d_J.p, bytes, hipMemcpyDeviceToHost, e.stream));
    if (k) HIP_CHECK(hipMemcpyAsync(k, e.d_K.p, bytes, hipMemcpyDeviceToHost, e.stream));
    HIP_CHECK(hipStreamSynchronize(e.stream));
    return HFG_OK;
  })
}
