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

// One molecule of a batch as the J/K kernels address it: its density, the
// first of its per-XCC accumulator copies, its AO component factors.
struct MolDev {
  const double* D;
  double* Jacc;
  double* Kacc;
  const double* scale;
  size_t nn;
  int n, pad;
};

// Arguments of the J/K kernels (jk_kernel.h).
struct JKArgs {
  const JKTask* tasks;   // all classes' tasks; this class's start at task0
  int task0, ntask, n;
  const ViewDev* views;
  const ComboDev* combos;
  const KetGroup* groups;
  const double* brec;    // bra primitive records, contiguous per view
  PrimPairs kt;          // ket side, transposed per group
  const double* kcoef;   // ket coefficients, transposed per group
  size_t nn;             // n*n, the stride between per-XCC accumulator copies
  double screen;         // Schwarz threshold of a combination pair in this build
  double screen0;        // the threshold the task lists were built with
  const double* vq;      // per view: largest Schwarz bound of its combinations
  const double* scale;
  const double* boys;
  const double* D;
  double* Jacc;
  double* Kacc;
  // LDS-row path (k_jk_rows): the class's jobs (absolute task offsets, njob + 1
  // entries) and its job counter, zero at launch; njob = 0 selects k_jk.
  const int* jobs;
  int njob;
  int* job_counter;
  // Batch: per-molecule addressing (ViewDev::mol); null for one molecule,
  // which then uses n, nn, D, Jacc, Kacc and scale above (n is the largest
  // molecule's for a batch: it sizes the LDS rows).
  const MolDev* mols;
};
// Launch of class cls's J/K kernel over ntask tasks with at most max_blocks
// workgroups; the kernels are split over three translation units.
void launch_jk_class(int cls, const JKArgs& a, int ntask, int max_blocks, hipStream_t s);
// Whether class cls runs on the side stream: (ds|ds) and the classes with a dp
// or dd bra, (dp|ss) .. (dd|dd), unless HFG_SIDE_CLASSES lists class indices
// (comma-separated; measured on (H2O)16/cc-pVDZ).
bool side_class(int cls);

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

struct Engine {
  int device = 0;
  hipStream_t stream = nullptr;
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

  explicit Engine(const hfg_basis_view& b);
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
  void sync();
  void set_profile(bool on);
};

void class_angular(int c, int* l4);
int persistent_blocks();
double calibration_copy(size_t bytes);

}  // namespace hfg
