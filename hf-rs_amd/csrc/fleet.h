// fleet.h -- a resident set of molecules whose J/K builds run in fused
// launches.
//
// Each molecule keeps its own engine (shell pairs, Schwarz bounds, J/K plan,
// density and accumulators in HBM) and occupies a slot: its table (MolTab)
// sits in the fleet's device array.  A build of the active slots is one
// launch per shell-quartet class over a directory of those slots (JKArgs),
// whatever their number and sizes: quartets of uneven classes from many small
// molecules fill the chip together, where molecule by molecule each launch
// would be nearly empty.  The active set can change between builds (a new
// directory of a few hundred ints), so molecules join and leave while others
// are mid-solve.  The reference builds each molecule's Fock matrix on its own
// (src/scf.rs:366-369, via src/shell_integrals.rs:156-240 and
// src/scf.rs:337-364); the results per molecule are the same.
#pragma once
#include <memory>
#include <vector>

#include "engine.h"

namespace hfg {

// One way to run a build's class launches: which classes go to the side
// stream (bit mask), that stream's workgroup cap (0: one stream) and the
// main-stream classes' workgroups per CU.
struct JKConfig {
  unsigned side_mask;
  int side_cap;
  int blocks_per_cu;
};

struct Fleet {
  hipStream_t stream = nullptr, side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  int side_blocks = 512;
  int capacity = 0;
  std::vector<MolTab> tabs;       // host copies, one per slot
  std::vector<char> occupied;
  DevBuf d_tabs, d_dir, d_active, d_counters, d_boys;
  JKDirectory dir;
  std::vector<int> active;        // slots of the current directory
  double screen0 = 1e-12;
  bool use_rows = true;
  // per-class launch times (HIP events on the stream each class runs on;
  // the side-stream classes' times include their overlap with the others,
  // unless exclusive: then the profiled builds run every class on the main
  // stream, one after the other, and each time is the class's alone)
  bool profile = false, exclusive = false;
  std::vector<hipEvent_t> events;
  std::vector<double> class_ms;
  std::vector<int64_t> class_launches;
  void set_profile(bool on, bool exclusive_ = false);
  // Autotuner: the first full builds of the fleet each run one candidate
  // configuration, timed by events on the stream; the score is milliseconds
  // per surviving shell quartet of the build's molecules, and the best
  // candidate runs every later build.  (The results do not depend on the
  // configuration, only the time does.)  HFG_AUTOTUNE=0 keeps the default.
  std::vector<JKConfig> candidates;
  std::vector<double> score;       // per candidate (< 0: not measured)
  int current = 0;                 // candidate of the next build
  int pending = -1;                // candidate whose timing is in flight
  double pending_work = 0;
  bool tuned = false;
  hipEvent_t tb = nullptr, te = nullptr;
  std::vector<double> slot_quartets;  // per slot (tuner's work measure)
  int cus = 256;
  void tune_step(bool full_build);
  const JKConfig& config() const { return candidates[current]; }

  explicit Fleet(int capacity);
  ~Fleet();
  Fleet(const Fleet&) = delete;
  Fleet& operator=(const Fleet&) = delete;
  // An engine created on this fleet's stream (no streams of its own).
  std::unique_ptr<Engine> make_engine(const hfg_basis_view& b);
  // Puts engine e's table into slot s (its density buffer d_D), copied on
  // stream st (the fleet's when null) before returning.  Slots set up by
  // different threads at once must differ.
  void attach(int s, const Engine& e, hipStream_t st = nullptr);
  void detach(int s);
  // The slots of the next builds; uploads their directory.
  void set_active(const std::vector<int>& slots);
  // Clears the accumulators of the active slots and runs the class launches;
  // the caller reduces the per-XCC copies (k_fleet_jk_out or its own kernel).
  // full_build: every active slot builds the whole quartet set (the tuner
  // only times such builds).
  void build(bool full_build = true);
  // build() followed by J and K of every active slot into jout[i], kout[i]
  // (device arrays of device pointers, one per active slot).
  void build_jk(double* const* d_jout, double* const* d_kout);
  void sync();
};

// Symmetrised J = sum_xcc Jacc + Jacc^T (K likewise) of active slot i into
// jout[i] / kout[i] (either may be null).  grid (x, count).
__global__ void k_fleet_jk_out(const MolTab* __restrict__ tabs, const int* __restrict__ slots,
                               double* const* __restrict__ jout, double* const* __restrict__ kout, int count);

}  // namespace hfg
