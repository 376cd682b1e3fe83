// noalloc.hip -- J and K of a device-resident density with no device
// allocation in the call: the scratch memory of the build (the molecule's
// table, the job counters and the per-XCC J/K accumulators) lives in one
// file-scope device array instead of the context's buffers.
//
// The array's host-side name is not a device address: its address on each
// device comes from hipGetSymbolAddress, taken once per device on the first
// call there (with that device current) and kept; only that address is used
// afterwards.  Every call on one device shares the array, so calls that
// overlap on one device race; each call clears what must start at zero.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <mutex>

#include "../../include/hfg.h"
#include "engine.h"

namespace hfg {

__global__ void k_fill(double* p, size_t n, double v);
__global__ void k_symmetrize(const double* Jacc, const double* Kacc, double* J, double* K, int n, const double* H,
                             double* F);

// Largest basis the array serves, and its layout in doubles: a head holding
// the table (MolTab) and the job counters, then the two accumulators of
// NXCC_SLOTS copies of n x n each.
constexpr int NOALLOC_MAX_N = 256;
constexpr size_t NOALLOC_HEAD = 1024;
constexpr size_t NOALLOC_ACC = (size_t)NXCC_SLOTS * NOALLOC_MAX_N * NOALLOC_MAX_N;
static_assert(sizeof(MolTab) + HFG_NSLOT * sizeof(int) + 256 <= NOALLOC_HEAD * sizeof(double), "head too small");

static __device__ __align__(256) double scratch[NOALLOC_HEAD + 2 * NOALLOC_ACC];

// The array's address on device dev (that device current when first asked).
static double* scratch_on(int dev) {
  static std::mutex mu;
  static std::map<int, double*> addr;
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(dev);
  if (it != addr.end()) return it->second;
  int cur = 0;
  HIP_CHECK(hipGetDevice(&cur));
  if (cur != dev) HIP_CHECK(hipSetDevice(dev));
  double* p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
  if (cur != dev) (void)hipSetDevice(cur);
  if (e != hipSuccess || !p)
    throw HfgError{HFG_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)};
  addr[dev] = p;
  return p;
}

// Writes the call's table (passed by value) into the scratch head.
__global__ void k_tab_put(MolTab* dst, MolTab t) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = t;
}

void jk_device_noalloc(Engine& e, const double* dP, double* dJ, double* dK) {
  if (e.n > NOALLOC_MAX_N)
    throw HfgError{HFG_INVALID_INPUT, "hfg_contract_jk_device_noalloc serves at most " +
                                          std::to_string(NOALLOC_MAX_N) + " basis functions (got " +
                                          std::to_string(e.n) + ")"};
  double* base = scratch_on(e.device);
  MolTab* tab = (MolTab*)base;
  int* counters = (int*)(base + (sizeof(MolTab) + 255) / 256 * 32);
  double* Jacc = base + NOALLOC_HEAD;
  double* Kacc = Jacc + NOALLOC_ACC;
  const size_t nn = (size_t)e.n * e.n;
  const size_t nacc = nn * NXCC_SLOTS;
  const int tb = 256;
  MolTab t = e.tab_host;
  t.D = dP;
  t.Jacc = Jacc;
  t.Kacc = Kacc;
  t.screen = e.screen;
  hipLaunchKernelGGL(k_tab_put, dim3(1), dim3(64), 0, e.stream, tab, t);
  hipLaunchKernelGGL(k_fill, dim3((nacc + tb - 1) / tb), dim3(tb), 0, e.stream, Jacc, nacc, 0.0);
  hipLaunchKernelGGL(k_fill, dim3((nacc + tb - 1) / tb), dim3(tb), 0, e.stream, Kacc, nacc, 0.0);
  JKLaunch L = e.launch_ctx();
  L.counters = counters;  // launch_jk_set clears them
  L.events = nullptr;
  launch_jk_set(L, tab, e.dir, e.d_dir.as<int>(), e.stream, e.side, e.side_blocks);
  hipLaunchKernelGGL(k_symmetrize, dim3((nn + tb - 1) / tb), dim3(tb), 0, e.stream, Jacc, Kacc, dJ, dK, e.n, nullptr,
                     nullptr);
  HIP_CHECK(hipGetLastError());
}

}  // namespace hfg
