// engine.h -- per-molecule device state of the Fock-build engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "classes.h"
#include "eri_device.h"
#include "host_prep.h"
#include "jk_plan.h"

namespace hfg {

// Number of per-XCC copies of the J/K accumulators (see jk_kernel.h).
constexpr int NXCC_SLOTS = 16;

// One molecule as the J/K kernels address it: its plan tables (jk_plan.cpp),
// the density of the current build, its per-XCC accumulator copies, AO
// factors and size, and the Schwarz threshold of the current build.  It lives
// in device memory and a launch reaches it through a directory (JKArgs), so
// one launch per class serves any set of resident molecules; a single
// molecule is a directory of one entry.
struct MolTab {
  const JKTask* tasks;
  const ViewDev* views;
  const ComboDev* combos;
  const KetGroup* groups;
  const double* brec;   // bra primitive records, contiguous per view
  const double* kt;     // ket primitive pairs, 9 columns of kt_size, transposed per group
  const double* kcoef;  // ket coefficients, transposed per group
  const double* vq;     // per view: largest Schwarz bound of its combinations
  const double* scale;  // per AO component factor
  const double* D;      // density of the build
  double* Jacc;         // NXCC_SLOTS copies of n*n
  double* Kacc;
  const int* jobs;      // absolute task offsets of the molecule's jobs, plus a sentinel
  size_t kt_size;
  size_t nn;            // n*n, the stride between per-XCC accumulator copies
  double screen;        // Schwarz threshold of a combination pair in this build
  int n;
  unsigned unit_mask;   // classes planned with one member per shell (kernel variant 1), bit c
  int task_begin[HFG_NCLASS + 1];
  int job_begin[HFG_NCLASS + 1];
};

// Arguments of one class launch over a directory: entry e is molecule
// tabs[dir_mol[e]]; the class's jobs of the entries are numbered
// consecutively, entry e owning [dir_job[e], dir_job[e + 1]) (its tasks
// likewise [dir_task[e], dir_task[e + 1])).
struct JKArgs {
  const MolTab* tabs;
  const int* dir_mol;
  const int* dir_job;
  const int* dir_task;
  int ndir, njob, ntask;
  int nmax;            // largest n of the entries: sizes the LDS rows
  double screen0;      // the threshold the task lists were built with
  const double* boys;
  int* job_counter;    // zero at launch (k_jk_rows)
  int use_rows;        // 0: the global-memory kernel k_jk
};
// Per-slot directory arrays of one launch set, in one int buffer: slot
// c + v * HFG_NCLASS holds the molecules that run class c in kernel variant v
// (MolTab::unit_mask); its entries start at off[slot] (dir_mol, then dir_job
// and dir_task with one extra entry each).
struct JKDirectory {
  std::vector<int> host;
  std::vector<int> off;     // HFG_NSLOT entries into host
  std::vector<int> count;   // entries per slot
  std::vector<int> njob, ntask;
  int nmax = 0;
};
// Directory of the molecules (tab indices) for every slot, skipping entries
// without work in a class; tabs_host holds the molecules' tables.
JKDirectory make_directory(const std::vector<int>& mol, const std::vector<const MolTab*>& tabs_host);
// Launch of class cls's J/K kernel in variant v (1: one contraction member per
// shell) with at most max_blocks workgroups; the kernels are split over five
// translation units.
void launch_jk_class(int cls, int v, const JKArgs& a, int max_blocks, hipStream_t s);
// Whether class cls runs on the side stream: (ds|ds) and the classes with a dp
// or dd bra, (dp|ss) .. (dd|dd), unless HFG_SIDE_CLASSES lists class indices
// (comma-separated; measured on (H2O)16/cc-pVDZ).
bool side_class(int cls);
// The classes side_class() names, as a bit mask.
unsigned default_side_mask();

// Arguments of the block kernels: one thread per task computes the unscaled
// block of pair `bra` with pair `ket` (kernel orientation) at out + off.
struct BlockArgs {
  const BlockTask* tasks;
  int ntask;
  const PairDev* pairs;
  PrimPairs pp;
  const double* boys;
  double* out;
};
void launch_block_class(int cls, const BlockArgs& a, hipStream_t s);
// The 3-index classes with an orbital f-p, f-d or f-f pair (df_f.hip).
void launch_df_f_block(int l1, int l2, int l3, int l4, const BlockArgs& a, hipStream_t s);

#define HIP_CHECK(expr)                                                                           \
  do {                                                                                            \
    hipError_t _e = (expr);                                                                       \
    if (_e != hipSuccess)                                                                         \
      throw ::hfg::HfgError{HFG_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)};     \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void alloc(size_t b) {
    free();
    if (b) HIP_CHECK(hipMalloc(&p, b));
    bytes = b;
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return (T*)p; }
  ~DevBuf() { free(); }
};

double md_flops(int la, int lb, int lc, int ld);
// Records the message returned by hfg_last_error(); returns code.
int set_error(int code, const std::string& msg);

// What one set of class launches needs besides the directory: job counters
// (HFG_NSLOT ints), fork/join events of the side stream, the Boys table,
// the threshold the task lists were built with, the LDS-row choice and
// optional per-class event pairs.
struct JKLaunch {
  int* counters;
  hipEvent_t fork, join;
  const double* boys;
  double screen0;
  bool use_rows;
  hipEvent_t* events;
  unsigned side_mask;  // classes on the side stream (bit c)
  int max_blocks;      // workgroup cap of a main-stream class; 0: the persistent default
};
// The J/K class launches of one build over directory dir (device copy d_dir)
// of the tables at d_tabs; accumulators must be cleared beforehand.  Classes
// side_class() names run on side_stream (at most side_cap workgroups each)
// when side_cap > 0; the streams are joined at the end.
void launch_jk_set(const JKLaunch& L, const MolTab* d_tabs, const JKDirectory& dir, const int* d_dir, hipStream_t main,
                   hipStream_t side_stream, int side_cap);

struct Engine {
  int device = 0;
  // The engine's stream: its own, or the one a batch shares among its members
  // (then the engine creates no stream and runs single-stream builds).
  hipStream_t stream = nullptr;
  bool own_streams = true;
  // Second stream (default priority) for the latency-bound high-L classes,
  // which then overlap the large low-L classes; joined before symmetrisation.
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int side_blocks = 512;  // workgroup cap of a side-stream class (HFG_SIDE_BLOCKS; 0: one stream)
  int n = 0;
  std::vector<Shell> shells;
  std::vector<PairInfo> pairs;  // hf-rs pair order
  PrimPairHost pp;
  std::vector<double> Q;        // Schwarz bound per pair
  std::vector<double> ao_scale; // per AO component factor
  int64_t n_quartets = 0;
  double screen = 1e-12;        // Q_ab*Q_cd > screen (src/shell_integrals.rs:184,226)

  DevBuf d_vq, d_pp, d_vb, d_kt, d_kcoef, d_groups, d_pairs, d_views, d_combos, d_boys, d_scale, d_tasks, d_jobs,
      d_counters, d_D, d_Jacc, d_Kacc, d_J, d_K, d_H, d_F;
  std::vector<size_t> task_begin;  // per class, into d_tasks (HFG_NCLASS+1 entries)
  std::vector<size_t> job_begin;   // per class, into d_jobs (HFG_NCLASS+1 entries)
  bool use_rows = true;            // LDS-row J/K kernels where they fit (HFG_JK_ROWS=0: global only)
  std::vector<double> class_flops;      // algorithmic FLOPs per launch
  std::vector<int64_t> class_quartets;  // quartets per launch
  double prim_quartets = 0;
  bool profile = false;
  std::vector<double> class_ms;
  std::vector<int64_t> class_launches;
  std::vector<hipEvent_t> events;  // 2 per class

  // shared: a stream owned by the caller (batch members); null creates the
  // engine's own two streams.
  explicit Engine(const hfg_basis_view& b, hipStream_t shared = nullptr);
  ~Engine();
  PrimPairs pp_view() const;
  PrimPairs kt_view() const;
  size_t vb_size = 0;
  size_t kt_size = 0;

  void compute_schwarz();
  // The J/K work plan (jk_plan.cpp): generally contracted shells, views,
  // ket groups and per-class task lists; then the algorithmic work count.
  void build_tasks();
  void count_work();
  // Blocks of the listed quartets (kernel orientation, unscaled) into host out;
  // pair_table / prims replace the engine's shell-pair tables when given.
  void blocks(const std::vector<BlockTask>& tasks, const std::vector<int>& cls, std::vector<double>& out,
              const PairDev* pair_table = nullptr, const PrimPairs* prims = nullptr);
  double prim_screen = 1e-20;
  size_t prim_pairs_kept = 0;  // primitive pairs over all views
  JKPlan plan;                 // host copy of the J/K plan (a batch concatenates them)
  int n_gen_shells = 0;
  int n_views = 0;
  // J/K from the device density dP into dJ, dK (and F = H + J - K/2 when H
  // given; F may alias H).  dmax < 1 (the largest |element| of a density
  // change) raises the Schwarz threshold to screen / dmax for this build.
  void jk_device(const double* dP, double* dJ, double* dK, const double* dH = nullptr, double* dF = nullptr,
                 double dmax = 1.0);
  // The molecule's table (host copy: D = d_D, screen = the plan's threshold)
  // and its one-entry directory in HBM.
  MolTab tab_host;
  DevBuf d_tab, d_dir;
  JKDirectory dir;
  // The engine's launch resources (job counters, fork/join events, Boys table,
  // threshold, profiling events) for launch_jk_set.
  JKLaunch launch_ctx();
  // HIP graphs of jk_device's launches after its table update (accumulator
  // fills, job-counter clear, the class launches on both streams and their
  // fork/join, the symmetrisation), one per set of output pointers (J, K,
  // H, F), captured on the second build with those pointers and replayed from
  // then on: every launch argument of those kernels is fixed per engine and
  // output set (the density and threshold of a build reach the kernels
  // through the table).  Engines with their own streams only; HFG_GRAPH=0
  // turns them off.
  bool use_graph = true;
  struct JKGraph {
    const void* key[4];
    int seen = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  };
  std::vector<JKGraph> graphs;
  int64_t graph_launches = 0;
  void sync();
  void set_profile(bool on);
};

void class_angular(int c, int* l4);
// The packed 8-fold screened ERI tensor of the engine's molecule into host
// out (np(np+1)/2 doubles, np = n(n+1)/2).
void eri_packed_host(Engine& e, double* out);
int persistent_blocks();
double calibration_copy(size_t bytes);

}  // namespace hfg
