// df_mfma.hip -- the exchange product of DF-K, K = alpha X X^T (X: n x L,
// row-major, L = rank * n_occ), on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64): coulomb_exchange_rhf's K += 2 A_Q A_Q^T summed
// over Q (src/df.rs:143-173) as one symmetric rank-L update.
//
// Only the tiles on and above the diagonal are computed (half the FLOPs of a
// GEMM).  Each 256-thread workgroup takes one 128 x 128 output tile and one
// slice of the contraction index: the tiles alone are far too few for 256
// CUs (28 at n = 800), so L is cut into S slices whose partial tiles go to a
// buffer [S][n][n] and are summed in a fixed order by a second kernel, which
// also writes the lower triangle.  Per 16-wide step of the contraction the
// two 128 x 16 panels of X are staged in LDS (rows padded to 17 doubles: the
// 16 rows one MFMA operand reads fall in distinct banks), the next step's
// panels are loaded into registers meanwhile; each wave owns a 64 x 64
// quarter of the tile, sixteen 16 x 16 accumulators.
//
// MFMA f64 16x16x4 operand layout (lane l): A[i][k] with i = l % 16,
// k = l / 16; B[k][j] with j = l % 16, k = l / 16; D[i][j] in register r of
// lane l with j = l % 16, i = l / 16 + 4 r (rows interleaved by register,
// unlike the 4-row groups of the FP32 16x16x4 form).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "engine.h"

namespace hfg {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int ST = 128;     // output tile
constexpr int KB = 16;      // contraction step staged in LDS
constexpr int PAD = KB + 1;
constexpr int WT = ST / 2;  // a wave's quarter of the tile
constexpr int NB = WT / 16; // 16 x 16 MFMA blocks per wave side

// The thread's share of the two ST x KB panels at contraction offset k0:
// row t / 2, columns 8 (t % 2) .. +7 of each (zero beyond n rows or kend).
struct Panel {
  double a[8], b[8];
};
__device__ __forceinline__ void load_panel(const double* __restrict__ X, int n, long long ldx, int i0, int j0,
                                           long long k0, long long kend, bool vec, Panel& p) {
  const int t = threadIdx.x, r = t >> 1, c0 = (t & 1) * 8;
  const int ra = i0 + r, rb = j0 + r;
  const long long k = k0 + c0;
  if (vec && k + 8 <= kend) {
    // 16-byte loads: rows start 16-byte aligned when ldx is even, k0 is a multiple of 16
#pragma unroll
    for (int c = 0; c < 8; c += 2) {
      const double2 va = ra < n ? *(const double2*)(X + (size_t)ra * ldx + k + c) : double2{0.0, 0.0};
      const double2 vb = rb < n ? *(const double2*)(X + (size_t)rb * ldx + k + c) : double2{0.0, 0.0};
      p.a[c] = va.x;
      p.a[c + 1] = va.y;
      p.b[c] = vb.x;
      p.b[c + 1] = vb.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const bool kin = k + c < kend;
      p.a[c] = (ra < n && kin) ? X[(size_t)ra * ldx + k + c] : 0.0;
      p.b[c] = (rb < n && kin) ? X[(size_t)rb * ldx + k + c] : 0.0;
    }
  }
}

__global__ __launch_bounds__(256) void k_syrk_f64(const double* __restrict__ X, int n, long long L, long long ldx,
                                                  const int2* __restrict__ tiles, int ntile, long long kslice,
                                                  double* __restrict__ part) {
  __shared__ double As[ST * PAD], Bs[ST * PAD];
  const int2 tile = tiles[blockIdx.x % ntile];
  const int s = blockIdx.x / ntile;
  const int i0 = tile.x * ST, j0 = tile.y * ST;
  const long long kbeg = (long long)s * kslice, kend = L < kbeg + kslice ? L : kbeg + kslice;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = w >> 1, wj = w & 1;
  const bool vec = (ldx & 1) == 0;
  d4 acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  const int r = t >> 1, c0 = (t & 1) * 8;
  Panel cur;
  if (kbeg < kend) load_panel(X, n, ldx, i0, j0, kbeg, kend, vec, cur);
  for (long long k0 = kbeg; k0 < kend; k0 += KB) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      As[r * PAD + c0 + c] = cur.a[c];
      Bs[r * PAD + c0 + c] = cur.b[c];
    }
    __syncthreads();
    // the next panels' global loads are in flight during this step's MFMAs
    if (k0 + KB < kend) load_panel(X, n, ldx, i0, j0, k0 + KB, kend, vec, cur);
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      const int kq = kk + (lane >> 4), rq = lane & 15;
      double av[NB], bv[NB];
#pragma unroll
      for (int a = 0; a < NB; ++a) av[a] = As[(WT * wi + 16 * a + rq) * PAD + kq];
#pragma unroll
      for (int b = 0; b < NB; ++b) bv[b] = Bs[(WT * wj + 16 * b + rq) * PAD + kq];
#pragma unroll
      for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  double* out = part + (size_t)s * n * n;
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + WT * wi + 16 * a + (lane >> 4) + 4 * q;
        const int j = j0 + WT * wj + 16 * b + (lane & 15);
        if (i < n && j < n) out[(size_t)i * n + j] = acc[a][b][q];
      }
}

// K[i][j] = alpha sum_s part[s][tile-upper position of (i, j)], both triangles.
__global__ void k_syrk_reduce(const double* __restrict__ part, int nslice, int n, double alpha,
                              double* __restrict__ K) {
  const size_t nn = (size_t)n * n;
  for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nn; idx += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(idx / n), j = (int)(idx - (size_t)i * n);
    const size_t src = (i / ST <= j / ST) ? idx : (size_t)j * n + i;
    double v = 0.0;
    for (int s = 0; s < nslice; ++s) v += part[(size_t)s * nn + src];
    K[idx] = alpha * v;
  }
}

}  // namespace

// Contraction slices for an n x n output and contraction length L: enough
// workgroups (about 2048) for the chip, at least 64 of L each, at most 64.
size_t syrk_f64_slices(int n, long long L) {
  const long long nt = (n + ST - 1) / ST, upper = nt * (nt + 1) / 2;
  long long s = std::max<long long>(1, (2048 + upper - 1) / upper);
  s = std::min<long long>(s, std::max<long long>(1, L / (4 * KB)));  // at least 64 of L per slice
  return (size_t)std::min<long long>(s, 64);
}

// K = alpha X X^T, X row-major n x L (leading dimension ldx), K row-major
// n x n (both triangles), on stream st; scratch grows to the partial tiles
// (slices x n x n doubles) and the tile list.
void syrk_f64(const double* X, int n, long long L, long long ldx, double alpha, double* K, DevBuf& scratch,
              hipStream_t st) {
  const int nt = (n + ST - 1) / ST;
  const int upper = nt * (nt + 1) / 2;
  const int S = (int)syrk_f64_slices(n, L);
  const size_t need = (size_t)S * n * n * sizeof(double) + (size_t)upper * sizeof(int2);
  if (scratch.bytes < need) scratch.alloc(need);
  double* part = scratch.as<double>();
  int2* tiles = (int2*)(scratch.as<char>() + (size_t)S * n * n * sizeof(double));
  std::vector<int2> h;
  h.reserve(upper);
  for (int a = 0; a < nt; ++a)
    for (int b = a; b < nt; ++b) h.push_back(make_int2(a, b));
  HIP_CHECK(hipMemcpyAsync(tiles, h.data(), h.size() * sizeof(int2), hipMemcpyHostToDevice, st));
  const long long kslice = ((L + S - 1) / S + KB - 1) / KB * KB;
  hipLaunchKernelGGL(k_syrk_f64, dim3((unsigned)(upper * S)), dim3(256), 0, st, X, n, L, ldx, tiles, upper, kslice,
                     part);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_syrk_reduce, dim3(2048), dim3(256), 0, st, part, S, n, alpha, K);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));  // the tile list is a host vector
}

}  // namespace hfg
