// engine.h -- per-molecule device state of the Fock-build engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "classes.h"
#include "eri_device.h"
#include "host_prep.h"

namespace hfg {

// Shell pair as the kernels read it (kernel orientation, see PairInfo).
struct PairDev {
  int p0, p1, ao_a, ao_b, sh_a, sh_b, pad0, pad1;
  double AB[3], pad2;
};

// One wave of work: bra pair, first ket pair and number of ket pairs (<= 64);
// all ket pairs share one primitive count.
struct JKTask {
  int bra, ket0, nket, grp;
};

// A group of ket pairs of one class and primitive count, contiguous in the
// J/K pair order; its primitive data is stored transposed,
// X[base + prim * size + (pair - start)], so that the lanes of a wave read
// consecutive addresses.
struct KetGroup {
  int start, size, base, npp;
};

struct BlockTask {
  int bra, ket;
  long long off;
};

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
  int n = 0;
  std::vector<Shell> shells;
  std::vector<PairInfo> pairs;  // hf-rs pair order
  PrimPairHost pp;
  std::vector<double> Q;        // Schwarz bound per pair
  std::vector<double> ao_scale; // per AO component factor
  int64_t n_quartets = 0;
  double screen = 1e-12;        // Q_ab*Q_cd > screen (src/shell_integrals.rs:184,226)

  DevBuf d_pp, d_pps, d_kt, d_groups, d_pairs, d_jkpairs, d_boys, d_scale, d_tasks, d_D, d_Jacc, d_Kacc, d_J, d_K, d_H, d_F;
  std::vector<size_t> task_begin;  // per class, into d_tasks (HFG_NCLASS+1 entries)
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
  PrimPairs pps_view() const;
  size_t pps_size = 0;
  size_t kt_size = 0;

  void compute_schwarz();
  void build_tasks();
  // Blocks of the listed quartets (kernel orientation, unscaled) into host out.
  void blocks(const std::vector<BlockTask>& tasks, const std::vector<int>& cls, std::vector<double>& out,
              const PairDev* pair_table = nullptr);
  // Primitive pairs kept for the J/K work (see engine.hip): per pair [s0, s1) into pps.
  void screen_primitives(std::vector<int>& s0, std::vector<int>& s1, PrimPairHost& pps);
  double prim_screen = 1e-20;
  size_t prim_pairs_kept = 0;
  // J/K from the device density d_D into d_J, d_K (and F = H + J - K/2 when H given).
  void jk_device(const double* dP, double* dJ, double* dK, const double* dH = nullptr, double* dF = nullptr);
  void sync();
  void set_profile(bool on);
};

void class_angular(int c, int* l4);
int persistent_blocks();
double calibration_copy(size_t bytes);

}  // namespace hfg
