// bscf.hip -- RHF solves of many molecules at once, with continuous batching.
//
// A server holds a fixed number of slots.  Each slot carries one molecule's
// engine (its J/K plan in the fleet, fleet.h), its RHF state in HBM (S, H, X,
// F, P, the previous density, the DIIS histories, the purification buffers)
// and its host-side iteration state.  One server iteration advances every
// occupied slot by one RHF iteration of run_scf_in_space_with_convergence
// (src/scf.rs:410-627):
//   Fock build (one fused launch per shell-quartet class over all slots),
//   energy, DIIS error FPS - SPF, convergence test, ADIIS/CDIIS extrapolation
//   (src/diis.rs), density by SP2 purification of X^T F X,
// each step one batched launch over all slots, whatever their sizes (la_batch.h
// and the kernels below); the host reads the per-slot scalars once per
// iteration.  A slot whose solve has finished is refilled at once from the
// queue of requests that have arrived, if the resident footprint stays within
// the HBM budget, so molecules join and leave while others are mid-solve.
// A request is a trajectory of geometry steps: each step's solve starts from
// the previous step's density (the reference's run_scf_with_initial_density,
// src/scf.rs:630-639), the slot's buffers stay resident and the DIIS history
// restarts.  The per-molecule arithmetic is that of scf.cpp, so each result
// is the one the molecule gets when solved alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <future>
#include <map>
#include <memory>
#include <numeric>
#include <thread>
#include <atomic>
#include <array>
#include <condition_variable>
#include <mutex>

#include "../../include/hfg.h"
#include "bscf_layout.h"
#include "fleet.h"
#include "la_batch.h"
#include "rhf.h"

namespace hfg {

constexpr int BMAXS = SLOT_DIIS;  // DIIS subspace (src/diis.rs:53-70)

// One slot's matrices as the batched kernels address them (row-major).
struct BMol {
  int n, m, elen, pad;
  double *S, *H, *X, *F, *P, *Pprev, *G, *Pb, *Dd, *T1, *T2, *E, *err, *Fu, *Fo, *Xs, *X2;
  double *cF, *cE, *aF, *aP;  // BMAXS slots each, strides n*n (cF, aF, aP) and elen (cE)
  MolTab* tab;
};

// Per-slot scalars the kernels write and the host reads once per iteration.
struct BStat {
  double e_el;        // 1/2 sum P (H + F)
  double emax;        // max |DIIS error|
  double rss;         // sum (P - Pprev)^2
  double dmax;        // max |P - Pb| (incremental build)
  double cdot[BMAXS + 1];      // error . cE[t], error . error
  double adot[2 * BMAXS + 1];  // P . aF[t], aP[t] . F, P . F
  double sb[2];       // Gershgorin bounds of X^T F X
  double st[2];       // tr X, tr X^2 of the current purification step
  int extra, fin, done, pad;
};

// Per-slot control of one iteration, written by the host.
struct BCtrl {
  int build;  // 0 full, 1 incremental (J/K of P - Pb)
  int keep;   // full build that starts an incremental run: keep G = F - H and Pb = P
  int store;  // 0 none, 1 CDIIS (F, error) into slot, 2 ADIIS (F, P) into slot
  int slot;
  int src;    // density from: 0 F, 1 CDIIS combination, 2 ADIIS combination, 3 H
  int nocc;
  double coef[BMAXS];
};

__device__ __forceinline__ void atomic_max_pos(double* p, double v) {
  // non-negative doubles order like their bit patterns
  atomicMax((unsigned long long*)p, (unsigned long long)__double_as_longlong(v));
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < nw; ++i) r += sh[i];
  return r;  // valid in thread 0
}

__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < nw; ++i) r = fmax(r, sh[i]);
  return r;
}

#define BSLOT()                      \
  const int slot = list[blockIdx.y]; \
  const BMol& M = mols[slot];        \
  const size_t nn = (size_t)M.n * M.n;
#define BLOOP(i, len) \
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < (len); i += (size_t)gridDim.x * blockDim.x)

// Clears the iteration's accumulated scalars.
__global__ void k_b_begin(const int* __restrict__ list, BStat* __restrict__ st) {
  BStat& s = st[list[blockIdx.y]];
  if (threadIdx.x == 0) {
    s.e_el = s.emax = s.rss = s.dmax = 0.0;
    for (int t = 0; t <= BMAXS; ++t) s.cdot[t] = 0.0;
    for (int t = 0; t <= 2 * BMAXS; ++t) s.adot[t] = 0.0;
  }
}

// Density of the build: P, or P - Pb for an incremental one (with max |.|).
__global__ void k_b_prep(const int* __restrict__ list, const BMol* __restrict__ mols, const BCtrl* __restrict__ ctrl,
                         BStat* __restrict__ st) {
  BSLOT();
  const bool inc = ctrl[slot].build == 1;
  double mx = 0.0;
  BLOOP(i, nn) {
    const double v = inc ? M.P[i] - M.Pb[i] : M.P[i];
    M.Dd[i] = v;
    mx = fmax(mx, fabs(v));
  }
  if (inc) {
    __shared__ double sh[16];
    mx = block_max(mx, sh);
    if (threadIdx.x == 0) atomic_max_pos(&st[slot].dmax, mx);
  }
}

// Schwarz threshold of the build: screen0, raised to screen0 / max|dP| for an
// incremental build (a vanishing change screens every quartet out).
__global__ void k_b_screen(const int* __restrict__ list, const BMol* __restrict__ mols,
                           const BCtrl* __restrict__ ctrl, const BStat* __restrict__ st, double screen0) {
  const int slot = list[blockIdx.y];
  if (threadIdx.x) return;
  double s = screen0;
  if (ctrl[slot].build == 1) {
    const double d = st[slot].dmax;
    s = d > 0.0 ? (d < 1.0 ? screen0 / d : screen0) : 1e300;
  }
  mols[slot].tab->screen = s;
}

// F from the per-XCC J/K copies: full builds F = H + J - K/2 (keeping
// G = J - K/2 and Pb = P when an incremental run starts), incremental builds
// G += J(dP) - K(dP)/2, F = H + G, Pb = P.
__global__ void k_b_fock(const int* __restrict__ list, const BMol* __restrict__ mols, const BCtrl* __restrict__ ctrl) {
  BSLOT();
  const MolTab& t = *M.tab;
  const int n = M.n;
  const BCtrl& c = ctrl[slot];
  BLOOP(idx, nn) {
    const size_t r = idx / n, col = idx % n, tr = col * n + r;
    double j = 0.0, k = 0.0;
    for (int x = 0; x < NXCC_SLOTS; ++x) {
      j += t.Jacc[x * nn + idx] + t.Jacc[x * nn + tr];
      k += t.Kacc[x * nn + idx] + t.Kacc[x * nn + tr];
    }
    const double g = j - 0.5 * k;
    if (c.build == 1) {
      const double gn = M.G[idx] + g;
      M.G[idx] = gn;
      M.F[idx] = M.H[idx] + gn;
      M.Pb[idx] = M.P[idx];
    } else {
      M.F[idx] = M.H[idx] + g;
      if (c.keep) {
        M.G[idx] = g;
        M.Pb[idx] = M.P[idx];
      }
    }
  }
}

// e_el = 1/2 sum P (H + F).
__global__ void k_b_energy(const int* __restrict__ list, const BMol* __restrict__ mols, BStat* __restrict__ st) {
  BSLOT();
  double s = 0.0;
  BLOOP(i, nn) s += M.P[i] * (M.H[i] + M.F[i]);
  __shared__ double sh[16];
  s = block_sum(s, sh);
  if (threadIdx.x == 0) atomicAdd(&st[slot].e_el, 0.5 * s);
}

// E = T2 - T2^T (T2 = F P S), the DIIS error FPS - SPF.
__global__ void k_b_antisym(const int* __restrict__ list, const BMol* __restrict__ mols) {
  BSLOT();
  const int n = M.n;
  BLOOP(idx, nn) {
    const size_t r = idx / n, c = idx % n;
    M.E[idx] = M.T2[idx] - M.T2[c * n + r];
  }
}

// max |error| and sum (P - Pprev)^2.
__global__ void k_b_norms(const int* __restrict__ list, const BMol* __restrict__ mols, BStat* __restrict__ st) {
  BSLOT();
  double mx = 0.0, ss = 0.0;
  BLOOP(i, (size_t)M.elen) mx = fmax(mx, fabs(M.err[i]));
  BLOOP(i, nn) {
    const double d = M.P[i] - M.Pprev[i];
    ss += d * d;
  }
  __shared__ double sh[16];
  mx = block_max(mx, sh);
  if (threadIdx.x == 0) atomic_max_pos(&st[slot].emax, mx);
  ss = block_sum(ss, sh);
  if (threadIdx.x == 0) atomicAdd(&st[slot].rss, ss);
}

// The inner products either DIIS variant needs for this iteration's entry,
// against the histories as they stand (the host then picks the variant and
// the slot the entry replaces): error . cE[t] and error . error; P . aF[t],
// aP[t] . F and P . F.
__global__ void k_b_dots(const int* __restrict__ list, const BMol* __restrict__ mols, BStat* __restrict__ st,
                         const int* __restrict__ want) {
  BSLOT();
  const int w = want[slot];  // bit 0: CDIIS products, bit 1: ADIIS products
  __shared__ double sh[16];
  if (w & 1) {
    double acc[BMAXS + 1];
    for (int t = 0; t <= BMAXS; ++t) acc[t] = 0.0;
    BLOOP(i, (size_t)M.elen) {
      const double e = M.err[i];
      for (int t = 0; t < BMAXS; ++t) acc[t] += e * M.cE[(size_t)t * M.elen + i];
      acc[BMAXS] += e * e;
    }
    for (int t = 0; t <= BMAXS; ++t) {
      const double v = block_sum(acc[t], sh);
      if (threadIdx.x == 0) atomicAdd(&st[slot].cdot[t], v);
    }
  }
  if (w & 2) {
    double acc[2 * BMAXS + 1];
    for (int t = 0; t <= 2 * BMAXS; ++t) acc[t] = 0.0;
    BLOOP(i, nn) {
      const double p = M.P[i], f = M.F[i];
      for (int t = 0; t < BMAXS; ++t) {
        acc[t] += p * M.aF[(size_t)t * nn + i];
        acc[BMAXS + t] += M.aP[(size_t)t * nn + i] * f;
      }
      acc[2 * BMAXS] += p * f;
    }
    for (int t = 0; t <= 2 * BMAXS; ++t) {
      const double v = block_sum(acc[t], sh);
      if (threadIdx.x == 0) atomicAdd(&st[slot].adot[t], v);
    }
  }
}

// The DIIS store (this iteration's entry into its history slot), Pprev = P,
// and the Fock matrix the density comes from (Fu).
__global__ void k_b_store(const int* __restrict__ list, const BMol* __restrict__ mols, const BCtrl* __restrict__ ctrl) {
  BSLOT();
  const BCtrl& c = ctrl[slot];
  BLOOP(i, nn) {
    const double f = M.F[i];
    if (c.store == 1) {
      M.cF[(size_t)c.slot * nn + i] = f;
      if (i < (size_t)M.elen) M.cE[(size_t)c.slot * M.elen + i] = M.err[i];
    } else if (c.store == 2) {
      M.aF[(size_t)c.slot * nn + i] = f;
      M.aP[(size_t)c.slot * nn + i] = M.P[i];
    }
    M.Pprev[i] = M.P[i];
    double u;
    if (c.src == 1 || c.src == 2) {
      const double* h = c.src == 1 ? M.cF : M.aF;
      u = 0.0;
      for (int t = 0; t < BMAXS; ++t)
        if (c.coef[t] != 0.0) u += c.coef[t] * h[(size_t)t * nn + i];
    } else {
      u = c.src == 3 ? M.H[i] : f;
    }
    M.Fu[i] = u;
  }
}

// SP2 purification of Fo = X^T Fu X (m x m) onto its nocc lowest
// eigenvectors (Niklasson, Phys. Rev. B 66, 155115 (2002)); the same steps
// and stopping rule as la_device.h's Sp2, per slot.
__global__ void k_b_gersh(const int* __restrict__ list, const BMol* __restrict__ mols, BStat* __restrict__ st) {
  const int slot = list[blockIdx.y];
  const BMol& M = mols[slot];
  const int m = M.m;
  double lo = 1e300, hi = -1e300;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    double r = 0.0;
    for (int j = 0; j < m; ++j)
      if (j != i) r += fabs(M.Fo[(size_t)i * m + j]);
    lo = fmin(lo, M.Fo[(size_t)i * m + i] - r);
    hi = fmax(hi, M.Fo[(size_t)i * m + i] + r);
  }
  __shared__ double sl[256], su[256];
  sl[threadIdx.x] = lo;
  su[threadIdx.x] = hi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sl[threadIdx.x] = fmin(sl[threadIdx.x], sl[threadIdx.x + s]);
      su[threadIdx.x] = fmax(su[threadIdx.x], su[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    BStat& b = st[slot];
    b.sb[0] = sl[0];
    b.sb[1] = su[0];
    b.extra = -1;
    b.fin = 0;
    b.done = 0;
  }
}

__global__ void k_b_sp2_init(const int* __restrict__ list, const BMol* __restrict__ mols,
                             const BStat* __restrict__ st) {
  const int slot = list[blockIdx.y];
  const BMol& M = mols[slot];
  const size_t mm = (size_t)M.m * M.m;
  const double emin = st[slot].sb[0], emax = st[slot].sb[1];
  const double s = 1.0 / fmax(emax - emin, 1e-300);
  BLOOP(i, mm) {
    const size_t r = i / M.m, c = i % M.m;
    M.Xs[i] = ((r == c ? emax : 0.0) - M.Fo[i]) * s;
  }
}

// Traces of X and X^2 of step k, then the stopping rule: a slot whose last
// step this is sets fin (its following products are skipped); one that had
// finished before sets done (its updates are skipped).
__global__ void k_b_sp2_traces(const int* __restrict__ list, const BMol* __restrict__ mols, BStat* __restrict__ st,
                               int k) {
  const int slot = list[blockIdx.y];
  const BMol& M = mols[slot];
  BStat& b = st[slot];
  if (b.done) return;
  if (b.fin) {
    __syncthreads();
    if (threadIdx.x == 0) b.done = 1;
    return;
  }
  const int m = M.m;
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    s0 += M.Xs[(size_t)i * m + i];
    s1 += M.X2[(size_t)i * m + i];
  }
  __shared__ double sh[16];
  s0 = block_sum(s0, sh);
  s1 = block_sum(s1, sh);
  if (threadIdx.x == 0) {
    b.st[0] = s0;
    b.st[1] = s1;
    const double err = s0 - s1;
    if (b.extra > 0) {
      if (--b.extra == 0) b.fin = 1;
    } else if (b.extra < 0 && k >= 11 && (k + 1) % 4 == 0 && fabs(err) < 1e-9) {
      b.extra = 3;
    }
  }
}

__global__ void k_b_sp2_update(const int* __restrict__ list, const BMol* __restrict__ mols,
                               const BStat* __restrict__ st, const BCtrl* __restrict__ ctrl) {
  const int slot = list[blockIdx.y];
  const BStat& b = st[slot];
  if (b.done) return;
  const BMol& M = mols[slot];
  const size_t mm = (size_t)M.m * M.m;
  const double t0 = b.st[0], t1 = b.st[1], nocc = ctrl[slot].nocc;
  const bool square = fabs(t1 - nocc) < fabs(2.0 * t0 - t1 - nocc);
  BLOOP(i, mm) M.Xs[i] = square ? M.X2[i] : 2.0 * M.Xs[i] - M.X2[i];
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------

struct Request {
  int id = 0;
  double arrival = 0;  // seconds after the start of the run
  int natom = 0, charge = 0;
  hfg_scf_options opt{};
  // per geometry step: the basis (deep copy) and (Z, x, y, z) per atom
  struct Step {
    std::vector<int32_t> atom, lxyz, poff;
    std::vector<double> center, expo, coef, zxyz;
    hfg_basis_view view() const {
      return hfg_basis_view{(int32_t)atom.size(), atom.data(), center.data(), lxyz.data(), poff.data(),
                            expo.data(), coef.data()};
    }
  };
  std::vector<Step> steps;
  // results
  std::vector<hfg_step_result> res;
  double admitted = -1, finished = -1;
  std::vector<double> last_p;  // final density of the last step
  int n = 0;
  bool keep_density = false;
  bool want_gradient = false;  // nuclear gradient of every step
  std::vector<double> grad;    // nsteps x natom x 3
  int status = HFG_OK;
  std::string error;
};

// FREE: empty; ITER: iterating; PRE: one more build after max_iter; FINAL:
// the final build; NEXT: finished a geometry step, the next one's set-up is
// due; SETUP: a worker is setting the slot up (the main loop leaves it alone
// until it collects the set-up).
enum Phase { FREE = 0, ITER, PRE, FINAL, NEXT, SETUP };

struct Slot {
  Phase phase = FREE;
  int req = -1, step = 0;
  std::unique_ptr<Engine> eng;
  int n = 0, m = 0, nocc = 0, elen = 0;
  double enuc = 0;
  DevBuf mem;        // every matrix of the slot, carved by layout()
  size_t cap = 0;    // doubles held by mem
  size_t footprint = 0;
  BMol mol{};
  // iteration state (scf.cpp's loop variables)
  int it = 0, since_full = -1, builds = 0;
  double prev_e = 0;
  std::vector<int> cslot, aslot;
  std::map<std::pair<int, int>, double> edot;
  double G[BMAXS][BMAXS];
  bool converged = false;
  double e_final = 0;
  std::vector<double> guess;  // density handed to the next geometry step (NEXT)
};

// A set-up worker: its own stream and BLAS handle, so that several slots are
// set up at once (engine, one-electron matrices, orthogonaliser, guess).
struct Worker {
  hipStream_t stream = nullptr;
  std::unique_ptr<Blas> blas;
};

static size_t slot_doubles(int n, int m) { return slot_layout(n, m, nullptr); }

// Carves the slot's buffer by slot_layout (every sub-buffer 256-byte aligned).
static void layout(Slot& s) {
  size_t off[SM_COUNT];
  const size_t total = slot_layout(s.n, s.m, off);
  if (total > s.cap) throw HfgError{HFG_DEVICE, "slot buffer smaller than its layout"};
  double* p = s.mem.as<double>();
  BMol& M = s.mol;
  M.n = s.n;
  M.m = s.m;
  M.elen = s.elen;
  M.S = p + off[SM_S];
  M.H = p + off[SM_H];
  M.X = p + off[SM_X];
  M.F = p + off[SM_F];
  M.P = p + off[SM_P];
  M.Pprev = p + off[SM_PPREV];
  M.G = p + off[SM_G];
  M.Pb = p + off[SM_PB];
  M.T1 = p + off[SM_T1];
  M.T2 = p + off[SM_T2];
  M.E = p + off[SM_E];
  M.Fu = p + off[SM_FU];
  M.Fo = p + off[SM_FO];
  M.Xs = p + off[SM_XS];
  M.X2 = p + off[SM_X2];
  M.cF = p + off[SM_CF];
  M.aF = p + off[SM_AF];
  M.aP = p + off[SM_AP];
  M.cE = p + off[SM_CE];
  M.err = s.m < s.n ? M.T2 : M.E;
  M.Dd = s.eng->d_D.as<double>();
}

struct Server {
  hfg_server_options opt{};
  std::unique_ptr<Fleet> fleet;
  std::vector<Slot> slots;
  std::vector<Request> reqs;
  std::deque<int> queue;  // request ids in arrival order, not yet admitted (or mid-trajectory)
  DevBuf d_mols, d_stat, d_ctrl, d_lists, d_want;
  std::vector<BMol> h_mols;
  std::vector<BStat> h_stat;
  std::vector<BCtrl> h_ctrl;
  std::vector<int> h_want;
  std::vector<int> h_lists;  // staging of the slot lists (4 x capacity)
  bool dirty = true;         // slot contents changed: the J/K directory must be rebuilt
  std::unique_ptr<Blas> blas;
  std::vector<Worker> workers;
  size_t budget = 0, resident = 0;
  // Admission pool: slot set-ups run on the worker threads while the slots
  // already set up iterate; the main loop collects the finished ones at the
  // start of each iteration.
  // A worker job for slot si: the gradient of step grad_step (if >= 0) of
  // request ri, then the set-up of step `step` (none when step is past the
  // request's last).  result: 0 failed, 1 set up, 2 set up without SAD
  // (core guess), 3 gradient only (the request has finished).
  struct Setup {
    int si, ri, step, result, grad_step;
  };
  std::mutex pool_mu;
  std::condition_variable pool_cv, done_cv;
  std::deque<Setup> pool_jobs;
  std::vector<Setup> pool_done;
  int pool_pending = 0;  // submitted and not yet collected (main thread only)
  bool pool_stop = false;
  void pool_work(Worker& w);
  void submit(const std::vector<std::array<int, 3>>& todo);
  // Moves finished set-ups into the iteration (or frees their slots); with
  // wait, blocks until at least one has finished.
  void collect(bool wait);
  // statistics
  int64_t fused_builds = 0, iterations = 0;
  double seconds = 0, seconds_setup = 0, seconds_sync = 0;
  std::chrono::steady_clock::time_point t0;

  explicit Server(const hfg_server_options& o);
  ~Server();
  double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
  void run();
  // 0: failed (recorded in the request), 1: ready, 2: ready but the density
  // must come from the core Hamiltonian (no SAD guess)
  int admit(int slot, int req, int step, const double* guess, Worker& w);
  void step_gradient(int slot, int req, int step, Worker& w);

  void finish_step(int slot);
  void density(const std::vector<int>& L);
  void upload_lists(const std::vector<int>& L, int which);
  const int* list_ptr(int which) const { return d_lists.as<int>() + (size_t)which * opt.capacity; }
};

Server::Server(const hfg_server_options& o) : opt(o) {
  if (opt.capacity <= 0) throw HfgError{HFG_INVALID_INPUT, "server capacity must be positive"};
  fleet.reset(new Fleet(opt.capacity));
  slots.resize(opt.capacity);
  const int cap = opt.capacity;
  d_mols.alloc(cap * sizeof(BMol));
  d_stat.alloc(cap * sizeof(BStat));
  d_ctrl.alloc(cap * sizeof(BCtrl));
  d_lists.alloc(4 * (size_t)cap * sizeof(int));
  d_want.alloc(cap * sizeof(int));
  h_mols.assign(cap, BMol{});
  h_stat.assign(cap, BStat{});
  h_ctrl.assign(cap, BCtrl{});
  h_want.assign(cap, 0);
  h_lists.assign(4 * (size_t)cap, 0);
  blas.reset(new Blas(fleet->stream));
  int nw = std::max(1, std::min(8, (int)std::thread::hardware_concurrency() / 2));
  if (const char* env = std::getenv("HFG_SERVER_WORKERS")) nw = std::max(1, std::atoi(env));
  nw = std::min(nw, opt.capacity);
  workers.resize(nw);
  for (Worker& w : workers) {
    HIP_CHECK(hipStreamCreate(&w.stream));  // default priority
    w.blas.reset(new Blas(w.stream));
  }
  size_t fr = 0, tot = 0;
  HIP_CHECK(hipMemGetInfo(&fr, &tot));
  budget = opt.hbm_budget_gb > 0 ? (size_t)(opt.hbm_budget_gb * 1e9) : (size_t)(0.8 * fr);
}

Server::~Server() {
  for (Slot& sl : slots) sl.eng.reset();  // engines were set up on the workers' streams
  for (Worker& w : workers) {
    w.blas.reset();
    if (w.stream) (void)hipStreamDestroy(w.stream);
  }
}

void Server::upload_lists(const std::vector<int>& L, int which) {
  if (L.empty()) return;
  int* h = h_lists.data() + (size_t)which * opt.capacity;
  std::copy(L.begin(), L.end(), h);
  HIP_CHECK(hipMemcpyAsync(d_lists.as<int>() + (size_t)which * opt.capacity, h, L.size() * sizeof(int),
                           hipMemcpyHostToDevice, fleet->stream));
}

// Footprint of a molecule of n functions in a slot: the engine's buffers
// (plan tables are small next to these) and the RHF state.
static size_t footprint(int n) {
  const size_t nn = (size_t)n * n;
  return (2 * NXCC_SLOTS + 8) * nn * 8 + slot_doubles(n, n) * 8;
}

// Set-up of one geometry step in a slot: engine (J/K plan), one-electron
// matrices, orthogonaliser, guess (SAD, or the given density), as scf.cpp.
int Server::admit(int si, int ri, int step, const double* guess, Worker& w) {
  Request& R = reqs[ri];
  Slot& s = slots[si];
  const Request::Step& st = R.steps[step];
  const hfg_basis_view bv = st.view();
  const double ts = now();
  try {
    const RhfInputs in = rhf_inputs(bv, st.zxyz.data(), R.natom, R.charge);
    const int n = bv.n_functions;
    if (s.eng) fleet->detach(si);
    s.eng.reset();
    hipStream_t stq = w.stream;
    s.eng.reset(new Engine(bv, stq));
    s.n = n;
    s.nocc = in.nocc;
    s.enuc = in.enuc;
    const size_t nn = (size_t)n * n;
    // S and H = T + V on the device, copied into the slot's buffer below once
    // the orthogonaliser has fixed its size
    Dev dS, dH;
    dS.alloc(nn);
    dH.alloc(nn);
    one_electron_device(s.eng->shells, n, in.atoms, dS.p, nullptr, nullptr, dH.p, stq);
    HMat X;
    s.m = orthogonalizer(*w.blas, stq, s.eng->shells, n, dS.p, R.opt.spherical != 0, X);
    if (s.nocc > s.m) throw HfgError{HFG_INVALID_INPUT, "more occupied orbitals than retained overlap rank"};
    s.elen = s.m < n ? s.m * s.m : (int)nn;
    const size_t need = slot_doubles(n, s.m);
    if (need > s.cap) {
      s.mem.alloc(need * sizeof(double));
      s.cap = need;
    }
    layout(s);
    HIP_CHECK(hipMemcpyAsync(s.mol.S, dS.p, nn * 8, hipMemcpyDeviceToDevice, stq));
    HIP_CHECK(hipMemcpyAsync(s.mol.H, dH.p, nn * 8, hipMemcpyDeviceToDevice, stq));
    upload(s.mol.X, X.data(), (size_t)n * s.m, stq);
    // DIIS histories start at zero (scf.cpp clears them the same way)
    for (double* h : {s.mol.cF, s.mol.aF, s.mol.aP})
      hipLaunchKernelGGL(k_zero, dim3((unsigned)((BMAXS * nn + 255) / 256)), dim3(256), 0, stq, h, BMAXS * nn);
    hipLaunchKernelGGL(k_zero, dim3((unsigned)(((size_t)BMAXS * s.elen + 255) / 256)), dim3(256), 0, stq, s.mol.cE,
                       (size_t)BMAXS * s.elen);
    bool core_guess = false;
    if (guess) {
      upload(s.mol.P, guess, nn, stq);
    } else {
      try {
        sad_guess_device(s.eng->shells, n, in.atoms, R.charge, s.mol.S, s.mol.P, *w.blas, stq);
      } catch (HfgError&) {
        core_guess = true;  // src/scf.rs:452-471: core Hamiltonian when SAD is unavailable
      }
    }
    fleet->attach(si, *s.eng, stq);
    s.mol.tab = fleet->d_tabs.as<MolTab>() + si;
    h_mols[si] = s.mol;
    HIP_CHECK(hipMemcpyAsync(d_mols.as<BMol>() + si, &h_mols[si], sizeof(BMol), hipMemcpyHostToDevice, stq));
    s.req = ri;
    s.step = step;
    s.it = 0;
    s.since_full = -1;
    s.builds = 0;
    s.prev_e = 0;
    s.cslot.clear();
    s.aslot.clear();
    s.edot.clear();
    std::memset(s.G, 0, sizeof s.G);
    s.converged = false;
    s.footprint = footprint(n);
    HIP_CHECK(hipMemcpyAsync(s.mol.Pprev, s.mol.P, nn * 8, hipMemcpyDeviceToDevice, stq));
    HIP_CHECK(hipStreamSynchronize(stq));
    if (R.admitted < 0) R.admitted = ts;
    R.n = n;
    return core_guess ? 2 : 1;
  } catch (HfgError& e) {
    R.status = e.code;
    R.error = e.msg;
  } catch (std::exception& e) {
    R.status = HFG_INTEGRAL;
    R.error = e.what();
  }
  R.finished = now();
  if (s.eng) fleet->detach(si);
  s.eng.reset();
  s.req = -1;
  return 0;
}

// Nuclear gradient of a finished geometry step from the slot's final density
// P and Fock matrix F (W = 1/2 P F P, the energy-weighted density of a
// converged RHF), on the worker's stream; the slot stays reserved (SETUP)
// meanwhile, and the gradient reads only its engine's tables.
void Server::step_gradient(int si, int ri, int step, Worker& w) {
  Slot& s = slots[si];
  Request& R = reqs[ri];
  const int n = s.n;
  const size_t nn = (size_t)n * n;
  Dev t;
  t.alloc(2 * nn);
  w.blas->gemm(false, false, n, n, n, 1.0, s.mol.F, n, s.mol.P, n, 0.0, t.p, n);     // F P
  w.blas->gemm(false, false, n, n, n, 0.5, s.mol.P, n, t.p, n, 0.0, t.p + nn, n);  // 1/2 P F P
  std::vector<std::array<double, 4>> atoms(R.natom);
  const std::vector<double>& z = R.steps[step].zxyz;
  for (int a = 0; a < R.natom; ++a)
    for (int k = 0; k < 4; ++k) atoms[a][k] = z[4 * a + k];
  rhf_gradient_device(*s.eng, atoms, s.mol.P, t.p + nn, R.grad.data() + (size_t)step * 3 * R.natom, w.stream);
}

// Set-ups run on the worker threads, each with its own stream and BLAS
// handle; a finished one waits in pool_done for the main loop.
void Server::pool_work(Worker& w) {
  for (;;) {
    Setup job;
    {
      std::unique_lock<std::mutex> lk(pool_mu);
      pool_cv.wait(lk, [&] { return pool_stop || !pool_jobs.empty(); });
      if (pool_jobs.empty()) return;  // stopping
      job = pool_jobs.front();
      pool_jobs.pop_front();
    }
    Slot& s = slots[job.si];
    Request& R = reqs[job.ri];
    bool ok = true;
    if (job.grad_step >= 0) {
      try {
        step_gradient(job.si, job.ri, job.grad_step, w);
      } catch (HfgError& e) {
        R.status = e.code;
        R.error = e.msg;
        ok = false;
      } catch (std::exception& e) {
        R.status = HFG_INTEGRAL;
        R.error = e.what();
        ok = false;
      }
    }
    if (job.step >= (int)R.steps.size()) {
      job.result = ok ? 3 : 0;
    } else if (!ok) {
      job.result = 0;
    } else {
      const bool warm = !s.guess.empty() &&
                        s.guess.size() == R.steps[job.step].atom.size() * R.steps[job.step].atom.size();
      job.result = admit(job.si, job.ri, job.step, warm ? s.guess.data() : nullptr, w);
    }
    {
      std::lock_guard<std::mutex> lk(pool_mu);
      pool_done.push_back(job);
    }
    done_cv.notify_all();
  }
}

void Server::submit(const std::vector<std::array<int, 3>>& todo) {
  if (todo.empty()) return;
  {
    std::lock_guard<std::mutex> lk(pool_mu);
    for (const auto& t : todo) {
      slots[t[0]].phase = SETUP;
      // a step after the first follows a finished step of the same request
      const int gs = t[2] > 0 && reqs[t[1]].want_gradient ? t[2] - 1 : -1;
      pool_jobs.push_back(Setup{t[0], t[1], t[2], 0, gs});
    }
  }
  pool_pending += (int)todo.size();
  pool_cv.notify_all();
}

void Server::collect(bool wait) {
  std::vector<Setup> got;
  {
    std::unique_lock<std::mutex> lk(pool_mu);
    if (wait && pool_done.empty() && pool_pending > 0) {
      const double ts = now();
      done_cv.wait(lk, [&] { return !pool_done.empty(); });
      seconds_setup += now() - ts;
    }
    got.swap(pool_done);
  }
  if (got.empty()) return;
  pool_pending -= (int)got.size();
  dirty = true;
  std::vector<int> core;
  for (const Setup& j : got) {
    Slot& s = slots[j.si];
    s.guess.clear();
    s.guess.shrink_to_fit();
    if (j.result == 3 || (j.result == 0 && s.eng)) {
      // the request's last gradient is done (or a gradient failed): free
      // the slot as finish_step does for a request without gradients
      Request& R = reqs[j.ri];
      R.finished = now();
      fleet->detach(j.si);
      s.eng.reset();
      resident -= std::min(resident, s.footprint);
      s.phase = FREE;
      s.req = -1;
      continue;
    }
    if (j.result == 0) {
      resident -= std::min(resident, footprint((int)reqs[j.ri].steps[j.step].atom.size()));
      s.phase = FREE;
      continue;
    }
    s.phase = ITER;
    if (j.result == 2) core.push_back(j.si);
  }
  if (!core.empty()) {
    // density from H before the first build (src/scf.rs:452-471)
    for (int si : core) {
      h_ctrl[si] = BCtrl{};
      h_ctrl[si].src = 3;
      h_ctrl[si].nocc = slots[si].nocc;
    }
    HIP_CHECK(hipMemcpyAsync(d_ctrl.p, h_ctrl.data(), h_ctrl.size() * sizeof(BCtrl), hipMemcpyHostToDevice,
                             fleet->stream));
    density(core);
    for (int si : core) {
      const Slot& s = slots[si];
      HIP_CHECK(hipMemcpyAsync(s.mol.Pprev, s.mol.P, (size_t)s.n * s.n * 8, hipMemcpyDeviceToDevice, fleet->stream));
    }
    HIP_CHECK(hipStreamSynchronize(fleet->stream));
  }
}

// Density of the listed slots from their Fu (k_b_store wrote it): Fo = X^T Fu
// X, SP2, P = 2 X D X^T; slots whose purification does not converge in 100
// steps fall back to the eigensolver, as scf.cpp.
void Server::density(const std::vector<int>& L) {
  if (L.empty()) return;
  hipStream_t stq = fleet->stream;
  const int k = (int)L.size();
  upload_lists(L, 1);
  const int* dl = list_ptr(1);
  int nmax = 0, mmax = 0;
  for (int si : L) {
    nmax = std::max(nmax, slots[si].n);
    mmax = std::max(mmax, slots[si].m);
  }
  hipLaunchKernelGGL(k_b_store, dim3(std::max(1, std::min(64, (nmax * nmax + 255) / 256)), k), dim3(256), 0, stq, dl,
                     d_mols.as<BMol>(), d_ctrl.as<BCtrl>());
  // Purification needs 0 < nocc < m; otherwise (and with HFG_SCF_EIGH=1)
  // the eigensolver gives the density, as in scf.cpp.
  static const bool eigh_only = std::getenv("HFG_SCF_EIGH") != nullptr;
  std::vector<int> Ls, failed;
  for (int si : L) {
    const Slot& s = slots[si];
    if (!eigh_only && s.nocc > 0 && s.nocc < s.m)
      Ls.push_back(si);
    else
      failed.push_back(si);
  }
  GemmList g1, g2, sq, g3, g4;
  BStat* dst = d_stat.as<BStat>();
  for (int si : L) {
    const Slot& s = slots[si];
    const BMol& M = s.mol;
    g1.add(false, false, s.n, s.m, s.n, 1.0, M.Fu, s.n, M.X, s.m, 0.0, M.T1, s.m);   // Fu X
    g2.add(true, false, s.m, s.m, s.n, 1.0, M.X, s.m, M.T1, s.m, 0.0, M.Fo, s.m);    // X^T Fu X
  }
  for (int si : Ls) {
    const Slot& s = slots[si];
    const BMol& M = s.mol;
    sq.add(false, false, s.m, s.m, s.m, 1.0, M.Xs, s.m, M.Xs, s.m, 0.0, M.X2, s.m, &dst[si].fin);
    g3.add(false, false, s.n, s.m, s.m, 1.0, M.X, s.m, M.Xs, s.m, 0.0, M.T1, s.m);   // X D
    g4.add(false, true, s.n, s.n, s.m, 2.0, M.T1, s.m, M.X, s.m, 0.0, M.P, s.n);     // 2 X D X^T
  }
  for (GemmList* g : {&g1, &g2, &sq, &g3, &g4}) g->upload(stq);
  g1.launch(stq);
  g2.launch(stq);
  if (!Ls.empty()) {
    const int ks = (int)Ls.size();
    upload_lists(Ls, 3);
    const int* ds = list_ptr(3);
    const dim3 gm(std::max(1, std::min(64, (mmax * mmax + 255) / 256)), ks);
    hipLaunchKernelGGL(k_b_gersh, dim3(1, ks), dim3(256), 0, stq, ds, d_mols.as<BMol>(), dst);
    hipLaunchKernelGGL(k_b_sp2_init, gm, dim3(256), 0, stq, ds, d_mols.as<BMol>(), dst);
    int step = 0;
    for (; step < Sp2::MAX_STEPS; ++step) {
      sq.launch(stq);
      hipLaunchKernelGGL(k_b_sp2_traces, dim3(1, ks), dim3(256), 0, stq, ds, d_mols.as<BMol>(), dst, step);
      hipLaunchKernelGGL(k_b_sp2_update, gm, dim3(256), 0, stq, ds, d_mols.as<BMol>(), dst, d_ctrl.as<BCtrl>());
      if (step >= 14 && (step + 1) % 4 == 0) {
        const auto ts = now();
        HIP_CHECK(hipMemcpyAsync(h_stat.data(), dst, h_stat.size() * sizeof(BStat), hipMemcpyDeviceToHost, stq));
        HIP_CHECK(hipStreamSynchronize(stq));
        seconds_sync += now() - ts;
        bool all = true;
        for (int si : Ls) all = all && h_stat[si].fin;
        if (all) break;
      }
    }
    if (step == Sp2::MAX_STEPS) {
      HIP_CHECK(hipMemcpyAsync(h_stat.data(), dst, h_stat.size() * sizeof(BStat), hipMemcpyDeviceToHost, stq));
      HIP_CHECK(hipStreamSynchronize(stq));
      for (int si : Ls)
        if (!h_stat[si].fin) failed.push_back(si);
    }
    g3.launch(stq);
    g4.launch(stq);
  }
  // eigensolver (no purification, or a vanishing HOMO-LUMO gap), one slot at
  // a time; Fo = X^T Fu X is in place for every listed slot
  for (int si : failed) {
    const Slot& s = slots[si];
    const BMol& M = s.mol;
    blas->gemm(false, false, s.n, s.m, s.n, 1.0, M.Fu, s.n, M.X, s.m, 0.0, M.T1, s.m);
    blas->gemm(true, false, s.m, s.m, s.n, 1.0, M.X, s.m, M.T1, s.m, 0.0, M.Fo, s.m);
    Eigh eig;
    Dev w;
    w.alloc(s.m);
    eig.run(*blas, M.Fo, w.p, s.m);
    // C = X V (the row-major view of Fo holds V^T), P = 2 C_occ C_occ^T
    blas->gemm(false, true, s.n, s.m, s.m, 1.0, M.X, s.m, M.Fo, s.m, 0.0, M.T1, s.m);
    blas->gemm(false, true, s.n, s.n, s.nocc, 2.0, M.T1, s.m, M.T1, s.m, 0.0, M.P, s.n);
    HIP_CHECK(hipStreamSynchronize(stq));
  }
}

// Records the finished step of slot si and moves the slot on: the next
// geometry step of the same request (the density stays as its guess), or
// free.
void Server::finish_step(int si) {
  Slot& s = slots[si];
  Request& R = reqs[s.req];
  hfg_step_result& r = R.res[s.step];
  r.energy = s.e_final;
  r.iterations = s.it;
  r.converged = s.converged;
  r.fock_builds = s.builds;
  r.n = s.n;
  if (s.step + 1 < (int)R.steps.size()) {
    // warm start of the next geometry (run_scf_with_initial_density,
    // src/scf.rs:630-639): the final density is kept for the next step's
    // set-up in the coming admission round; the slot stays reserved
    s.guess.resize((size_t)s.n * s.n);
    download(s.guess.data(), s.mol.P, s.guess.size(), fleet->stream);
    s.phase = NEXT;
    return;
  }
  if (R.keep_density) {
    R.last_p.resize((size_t)s.n * s.n);
    download(R.last_p.data(), s.mol.P, R.last_p.size(), fleet->stream);
  }
  if (R.want_gradient) {
    // the last step's gradient runs on a worker; the slot stays reserved
    // until it is collected (a job with no set-up after it)
    s.phase = NEXT;
    return;
  }
  R.finished = now();
  fleet->detach(si);
  s.eng.reset();
  resident -= s.footprint;
  s.phase = FREE;
  s.req = -1;
  dirty = true;
}

void Server::run() {
  hipStream_t stq = fleet->stream;
  t0 = std::chrono::steady_clock::now();
  const double screen0 = 1e-12;
  std::vector<int> order(reqs.size());
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival < reqs[b].arrival; });
  queue.assign(order.begin(), order.end());
  std::vector<int> prev_active;
  // the admission pool for this run; stopped and joined on every way out
  pool_stop = false;
  pool_pending = 0;
  std::vector<std::thread> pool;
  for (Worker& w : workers) pool.emplace_back([this, &w] { pool_work(w); });
  struct PoolJoin {
    Server* S;
    std::vector<std::thread>& th;
    ~PoolJoin() {
      {
        std::lock_guard<std::mutex> lk(S->pool_mu);
        S->pool_stop = true;
        S->pool_jobs.clear();
      }
      S->pool_cv.notify_all();
      for (auto& t : th) t.join();
      S->pool_done.clear();
    }
  } pool_join{this, pool};
  for (;;) {
    // admission: the next geometry steps of requests in flight, then arrived
    // requests into free slots while the footprint fits; all set up at once
    std::vector<std::array<int, 3>> todo;
    for (int si = 0; si < opt.capacity; ++si)
      if (slots[si].phase == NEXT) todo.push_back({si, slots[si].req, slots[si].step + 1});
    const double tnow = now();
    // Requests that have arrived are admitted largest first: a large molecule
    // admitted late would finish last (longest-processing-time order).
    size_t arrived = 0;
    while (arrived < queue.size() && reqs[queue[arrived]].arrival <= tnow) ++arrived;
    std::stable_sort(queue.begin(), queue.begin() + arrived, [&](int a, int b) {
      return reqs[a].steps[0].atom.size() > reqs[b].steps[0].atom.size();
    });
    for (int si = 0; si < opt.capacity && arrived > 0; ++si) {
      if (slots[si].phase != FREE) continue;
      const int ri = queue.front();
      const size_t fp = footprint((int)reqs[ri].steps[0].atom.size());
      if (resident > 0 && resident + fp > budget) break;
      queue.pop_front();
      --arrived;
      resident += fp;
      slots[si].phase = SETUP;  // reserved until its set-up is collected
      todo.push_back({si, ri, 0});
    }
    submit(todo);
    collect(false);
    std::vector<int> act;
    for (int si = 0; si < opt.capacity; ++si)
      if (slots[si].phase != FREE && slots[si].phase != SETUP) act.push_back(si);
    if (act.empty()) {
      if (pool_pending > 0) {
        collect(true);  // nothing can iterate until a set-up finishes
        continue;
      }
      if (queue.empty()) break;
      // idle until the next arrival
      const double wait = reqs[queue.front()].arrival - now();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
      continue;
    }
    const int k = (int)act.size();
    int nmax = 0;
    for (int si : act) nmax = std::max(nmax, slots[si].n);
    // build type per slot (scf.cpp: fock_iter / fock)
    for (int si : act) {
      Slot& s = slots[si];
      BCtrl& c = h_ctrl[si];
      c = BCtrl{};
      c.nocc = s.nocc;
      const bool incremental = reqs[s.req].opt.incremental_fock != 0;
      if (s.phase == ITER && incremental && !(s.since_full < 0 || s.since_full >= 12)) {
        c.build = 1;
      } else {
        c.build = 0;
        c.keep = s.phase == ITER && incremental;
      }
    }
    HIP_CHECK(hipMemcpyAsync(d_ctrl.p, h_ctrl.data(), h_ctrl.size() * sizeof(BCtrl), hipMemcpyHostToDevice, stq));
    upload_lists(act, 0);
    const int* L0 = list_ptr(0);
    const dim3 ge(std::max(1, std::min(64, (nmax * nmax + 255) / 256)), k);
    hipLaunchKernelGGL(k_b_begin, dim3(1, k), dim3(64), 0, stq, L0, d_stat.as<BStat>());
    hipLaunchKernelGGL(k_b_prep, ge, dim3(256), 0, stq, L0, d_mols.as<BMol>(), d_ctrl.as<BCtrl>(),
                       d_stat.as<BStat>());
    hipLaunchKernelGGL(k_b_screen, dim3(1, k), dim3(64), 0, stq, L0, d_mols.as<BMol>(), d_ctrl.as<BCtrl>(),
                       d_stat.as<BStat>(), screen0);
    if (dirty || act != prev_active) {
      fleet->set_active(act);
      prev_active = act;
      dirty = false;
    }
    // the tuner times full builds of a settled set only (a growing set
    // would favour the later candidates)
    bool all_full = pool_pending == 0;
    for (int si : act) all_full = all_full && h_ctrl[si].build == 0;
    fleet->build(all_full);
    ++fused_builds;
    hipLaunchKernelGGL(k_b_fock, ge, dim3(256), 0, stq, L0, d_mols.as<BMol>(), d_ctrl.as<BCtrl>());
    hipLaunchKernelGGL(k_b_energy, ge, dim3(256), 0, stq, L0, d_mols.as<BMol>(), d_stat.as<BStat>());
    for (int si : act) slots[si].builds++;
    // error, norms and DIIS products of the iterating slots
    std::vector<int> Li;
    for (int si : act)
      if (slots[si].phase == ITER) Li.push_back(si);
    if (!Li.empty()) {
      upload_lists(Li, 2);
      const int* L2 = list_ptr(2);
      const int ki = (int)Li.size();
      GemmList e1, e2, p1, p2;
      for (int si : Li) {
        const Slot& s = slots[si];
        const BMol& M = s.mol;
        e1.add(false, false, s.n, s.n, s.n, 1.0, M.F, s.n, M.P, s.n, 0.0, M.T1, s.n);   // F P
        e2.add(false, false, s.n, s.n, s.n, 1.0, M.T1, s.n, M.S, s.n, 0.0, M.T2, s.n);  // F P S
        if (s.m < s.n) {
          p1.add(false, false, s.n, s.m, s.n, 1.0, M.E, s.n, M.X, s.m, 0.0, M.T1, s.m);  // E X
          p2.add(true, false, s.m, s.m, s.n, 1.0, M.X, s.m, M.T1, s.m, 0.0, M.T2, s.m);  // X^T E X
        }
        const int algo = reqs[s.req].opt.conv_algo;
        h_want[si] = algo == 1 ? 1 : algo == 2 ? 2 : 3;
      }
      for (GemmList* g : {&e1, &e2, &p1, &p2}) g->upload(stq);
      HIP_CHECK(hipMemcpyAsync(d_want.p, h_want.data(), h_want.size() * sizeof(int), hipMemcpyHostToDevice, stq));
      e1.launch(stq);
      e2.launch(stq);
      const dim3 gi(std::max(1, std::min(64, (nmax * nmax + 255) / 256)), ki);
      hipLaunchKernelGGL(k_b_antisym, gi, dim3(256), 0, stq, L2, d_mols.as<BMol>());
      p1.launch(stq);
      p2.launch(stq);
      hipLaunchKernelGGL(k_b_norms, gi, dim3(256), 0, stq, L2, d_mols.as<BMol>(), d_stat.as<BStat>());
      hipLaunchKernelGGL(k_b_dots, gi, dim3(256), 0, stq, L2, d_mols.as<BMol>(), d_stat.as<BStat>(),
                         d_want.as<int>());
    }
    {
      const auto ts = now();
      HIP_CHECK(hipMemcpyAsync(h_stat.data(), d_stat.p, h_stat.size() * sizeof(BStat), hipMemcpyDeviceToHost, stq));
      HIP_CHECK(hipStreamSynchronize(stq));
      seconds_sync += now() - ts;
    }
    ++iterations;
    // host decisions (scf.cpp's loop body after the Fock build)
    std::vector<int> Ld, done;
    for (int si : act) {
      Slot& s = slots[si];
      const Request& R = reqs[s.req];
      const BStat& b = h_stat[si];
      BCtrl& c = h_ctrl[si];
      if (s.phase == ITER) {
        if (c.build == 1)
          ++s.since_full;
        else if (R.opt.incremental_fock)
          s.since_full = 0;
      }
      if (s.phase == FINAL) {
        s.e_final = b.e_el + s.enuc;
        done.push_back(si);
        continue;
      }
      if (s.phase == PRE) {
        c.store = 0;
        c.src = 0;
        s.phase = FINAL;
        Ld.push_back(si);
        continue;
      }
      // ITER
      const int it = ++s.it;
      const double tot = b.e_el + s.enuc;
      const double de = std::fabs(tot - s.prev_e);
      s.prev_e = tot;
      const double en = b.emax;
      const double rms = std::sqrt(b.rss / ((double)s.n * s.n));
      if (R.opt.verbose)
        fprintf(stderr, "  hfg slot %d iter %3d  E=%.12f  dE=%.3e  |e|=%.3e\n", si, it, tot, de, en);
      if (it > 1 && de < R.opt.energy_tol && en < R.opt.residual_tol && rms < R.opt.density_tol) {
        s.converged = true;
        c.store = 0;
        c.src = 0;
        s.phase = FINAL;
        Ld.push_back(si);
        continue;
      }
      bool use_cdiis;
      if (R.opt.conv_algo == 1)
        use_cdiis = true;
      else if (R.opt.conv_algo == 2)
        use_cdiis = false;
      else
        use_cdiis = !(en > (R.opt.conv_algo == 3 ? 0.08 : 0.05) || it <= 2);
      auto take_slot = [&](std::vector<int>& order) {
        int sl;
        if ((int)order.size() >= BMAXS) {
          sl = order.front();
          order.erase(order.begin());
        } else {
          sl = (int)order.size();
          for (int q = 0; q < BMAXS; ++q)
            if (std::find(order.begin(), order.end(), q) == order.end()) {
              sl = q;
              break;
            }
        }
        order.push_back(sl);
        return sl;
      };
      c.src = 0;
      for (int t = 0; t < BMAXS; ++t) c.coef[t] = 0.0;
      if (use_cdiis) {
        const int sl = take_slot(s.cslot);
        c.store = 1;
        c.slot = sl;
        for (auto e = s.edot.begin(); e != s.edot.end();)
          if (e->first.first == sl || e->first.second == sl)
            e = s.edot.erase(e);
          else
            ++e;
        for (int o : s.cslot) s.edot[{std::min(sl, o), std::max(sl, o)}] = o == sl ? b.cdot[BMAXS] : b.cdot[o];
        const int kk = (int)s.cslot.size();
        if (kk >= 2) {
          HMat B((size_t)(kk + 1) * (kk + 1), 0.0);
          for (int i = 0; i < kk; ++i) {
            for (int j = 0; j < kk; ++j)
              B[(size_t)i * (kk + 1) + j] =
                  s.edot[{std::min(s.cslot[i], s.cslot[j]), std::max(s.cslot[i], s.cslot[j])}];
            B[(size_t)i * (kk + 1) + kk] = B[(size_t)kk * (kk + 1) + i] = -1.0;
          }
          std::vector<double> rhs(kk + 1, 0.0), cc;
          rhs[kk] = -1.0;
          if (solve_small(B, rhs, kk + 1, cc)) {
            for (int i = 0; i < kk; ++i) c.coef[s.cslot[i]] = cc[i];
            c.src = 1;
          }
        }
      } else {
        const int sl = take_slot(s.aslot);
        c.store = 2;
        c.slot = sl;
        for (int t = 0; t < BMAXS; ++t) {
          if (t == sl) continue;
          s.G[sl][t] = b.adot[t];          // P_new . F_t
          s.G[t][sl] = b.adot[BMAXS + t];  // P_t . F_new
        }
        s.G[sl][sl] = b.adot[2 * BMAXS];
        const int kk = (int)s.aslot.size();
        if (kk >= 2) {
          auto g = [&](int i, int j) { return s.G[s.aslot[i]][s.aslot[j]]; };  // P_i . F_j
          const int last = kk - 1;
          HMat Q((size_t)kk * kk);
          std::vector<double> bb(kk);
          for (int i = 0; i < kk; ++i) {
            bb[i] = g(i, last) - g(last, last);
            for (int j = 0; j < kk; ++j) Q[(size_t)i * kk + j] = g(i, j) - g(i, last) - g(last, j) + g(last, last);
          }
          const auto cc = simplex_minimise(Q, bb);
          for (int i = 0; i < kk; ++i) c.coef[s.aslot[i]] = cc[i];
          c.src = 2;
        }
      }
      if (it >= R.opt.max_iter) s.phase = PRE;  // after this density: a build, a density from F, the final build
      Ld.push_back(si);
    }
    if (!Ld.empty()) {
      HIP_CHECK(hipMemcpyAsync(d_ctrl.p, h_ctrl.data(), h_ctrl.size() * sizeof(BCtrl), hipMemcpyHostToDevice, stq));
      density(Ld);
    }
    for (int si : done) finish_step(si);
  }
  HIP_CHECK(hipStreamSynchronize(stq));
  seconds = now();
}

}  // namespace hfg

using namespace hfg;

struct hfg_server {
  std::unique_ptr<Server> s;
};

#define SGUARD(...)                                        \
  try {                                                    \
    __VA_ARGS__                                            \
  } catch (HfgError & ex) {                                \
    return set_error(ex.code, ex.msg);                     \
  } catch (std::bad_alloc&) {                              \
    return set_error(HFG_DEVICE, "host allocation failed"); \
  } catch (std::exception & ex) {                          \
    return set_error(HFG_INTEGRAL, ex.what());             \
  }

int hfg_server_create(const hfg_server_options* opt, hfg_server** out) {
  if (!opt || !out) return set_error(HFG_INVALID_INPUT, "null argument");
  SGUARD({
    auto* h = new hfg_server;
    try {
      h->s.reset(new Server(*opt));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return HFG_OK;
  })
}

void hfg_server_destroy(hfg_server* s) {
  if (!s) return;
  for (auto& sl : s->s->slots) sl.eng.reset();  // engines use the fleet's stream
  delete s;
}

int hfg_server_submit(hfg_server* srv, double arrival_s, int nsteps, const hfg_basis_view* bases,
                      const double* const* zxyz, int natom, int charge, const hfg_scf_options* opt, int flags,
                      int* request_id) {
  if (!srv || !bases || !zxyz || !opt || nsteps <= 0 || natom <= 0)
    return set_error(HFG_INVALID_INPUT, "null argument or empty request");
  if (opt->max_iter <= 0 || !(opt->energy_tol > 0) || !(opt->residual_tol > 0) || !(opt->density_tol > 0))
    return set_error(HFG_INVALID_INPUT, "SCF requires positive tolerances and iteration limit");
  SGUARD({
    Server& S = *srv->s;
    Request R;
    R.id = (int)S.reqs.size();
    R.arrival = arrival_s;
    R.natom = natom;
    R.charge = charge;
    R.opt = *opt;
    R.keep_density = (flags & HFG_SERVE_KEEP_DENSITY) != 0;
    // Switched off: a run of the served-gradient test after the rest of the
    // GPU suite stopped making progress (the same test passed alone); the
    // cause is not established, so the path stays closed.
    if (flags & HFG_SERVE_GRADIENT)
      throw HfgError{HFG_UNSUPPORTED, "gradients of served trajectories are switched off in this build"};
    R.want_gradient = false;
    if (R.want_gradient) R.grad.assign((size_t)nsteps * 3 * natom, 0.0);
    for (int k = 0; k < nsteps; ++k) {
      const hfg_basis_view& b = bases[k];
      if (b.n_functions <= 0 || !zxyz[k]) throw HfgError{HFG_INVALID_INPUT, "empty basis or geometry"};
      Request::Step st;
      const int n = b.n_functions, np = b.prim_offset[n];
      st.atom.assign(b.atom_index, b.atom_index + n);
      st.lxyz.assign(b.lxyz, b.lxyz + 3 * n);
      st.poff.assign(b.prim_offset, b.prim_offset + n + 1);
      st.center.assign(b.center, b.center + 3 * n);
      st.expo.assign(b.exponent, b.exponent + np);
      st.coef.assign(b.coefficient, b.coefficient + np);
      st.zxyz.assign(zxyz[k], zxyz[k] + 4 * natom);
      R.steps.push_back(std::move(st));
    }
    R.res.assign(nsteps, hfg_step_result{});
    S.reqs.push_back(std::move(R));
    if (request_id) *request_id = S.reqs.back().id;
    return HFG_OK;
  })
}

int hfg_server_run(hfg_server* srv) {
  if (!srv) return set_error(HFG_INVALID_INPUT, "null server");
  SGUARD({
    srv->s->run();
    return HFG_OK;
  })
}

int hfg_server_result(hfg_server* srv, int id, hfg_request_result* r, hfg_step_result* steps, double* p_out) {
  if (!srv || !r) return set_error(HFG_INVALID_INPUT, "null argument");
  Server& S = *srv->s;
  if (id < 0 || id >= (int)S.reqs.size()) return set_error(HFG_INVALID_INPUT, "request id out of range");
  const Request& R = S.reqs[id];
  r->arrival_s = R.arrival;
  r->admitted_s = R.admitted;
  r->finished_s = R.finished;
  r->steps = (int32_t)R.steps.size();
  int conv = R.status == HFG_OK;
  for (auto& x : R.res) conv = conv && x.converged;
  r->converged = conv;
  r->status = R.status;
  r->n = R.n;
  if (steps) std::memcpy(steps, R.res.data(), R.res.size() * sizeof(hfg_step_result));
  if (p_out) {
    if (R.last_p.empty()) return set_error(HFG_INVALID_INPUT, "density not kept for this request");
    std::memcpy(p_out, R.last_p.data(), R.last_p.size() * sizeof(double));
  }
  if (R.status != HFG_OK) return set_error(R.status, R.error);
  return HFG_OK;
}

int hfg_server_gradient(hfg_server* srv, int id, double* grad) {
  if (!srv || !grad) return set_error(HFG_INVALID_INPUT, "null argument");
  Server& S = *srv->s;
  if (id < 0 || id >= (int)S.reqs.size()) return set_error(HFG_INVALID_INPUT, "request id out of range");
  const Request& R = S.reqs[id];
  if (!R.want_gradient) return set_error(HFG_INVALID_INPUT, "gradients not requested for this request");
  if (R.status != HFG_OK) return set_error(R.status, R.error);
  if (R.finished < 0) return set_error(HFG_INVALID_INPUT, "request has not finished");
  std::memcpy(grad, R.grad.data(), R.grad.size() * sizeof(double));
  return HFG_OK;
}

int hfg_server_stats(const hfg_server* srv, hfg_server_stats_t* st) {
  if (!srv || !st) return set_error(HFG_INVALID_INPUT, "null argument");
  const Server& S = *srv->s;
  st->fused_builds = S.fused_builds;
  st->iterations = S.iterations;
  st->seconds = S.seconds;
  st->seconds_setup = S.seconds_setup;
  st->seconds_sync = S.seconds_sync;
  st->hbm_budget_bytes = (double)S.budget;
  st->jk_config = S.fleet->tuned ? S.fleet->current : -1;
  st->jk_candidates = (int32_t)S.fleet->candidates.size();
  return HFG_OK;
}
