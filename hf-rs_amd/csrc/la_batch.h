// la_batch.h -- FP64 linear algebra over many matrices of different sizes in
// one launch: the batched RHF iterations run every molecule's products,
// reductions and purification steps together instead of one library call per
// molecule.  Row-major matrices throughout.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "engine.h"

namespace hfg {

// One product C (M x N, leading dimension ldc) = alpha op(A) op(B) + beta C,
// op(A) M x K, op(B) K x N; skip (optional, device) nonzero skips it.
struct GemmOp {
  const double* A;
  const double* B;
  double* C;
  int M, N, K, lda, ldb, ldc;
  double alpha, beta;
  int ta, tb;
  const int* skip;
};

// 64 x 64 tile of one product per workgroup (grid: N tiles, M tiles, ops):
// 16 x 16 threads with 4 x 4 outputs each, k panels of 16 staged through LDS
// in [k][m] / [k][n] order (one padding column against bank conflicts).
// Out-of-range tiles of the smaller products of a launch exit at once.
__global__ __launch_bounds__(256) inline void k_gemm_batched(const GemmOp* __restrict__ ops) {
  const GemmOp& op = ops[blockIdx.z];
  if (op.skip && *op.skip) return;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  if (m0 >= op.M || n0 >= op.N) return;
  __shared__ double As[16][65];
  __shared__ double Bs[16][65];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  const double* __restrict__ A = op.A;
  const double* __restrict__ B = op.B;
  for (int k0 = 0; k0 < op.K; k0 += 16) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = threadIdx.x + 256 * r;
      int mm, kk;
      if (op.ta) {
        kk = i >> 6;
        mm = i & 63;
      } else {
        mm = i >> 4;
        kk = i & 15;
      }
      const int gm = m0 + mm, gk = k0 + kk;
      double v = 0.0;
      if (gm < op.M && gk < op.K) v = op.ta ? A[(size_t)gk * op.lda + gm] : A[(size_t)gm * op.lda + gk];
      As[kk][mm] = v;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = threadIdx.x + 256 * r;
      int kk, nn;
      if (op.tb) {
        nn = i >> 4;
        kk = i & 15;
      } else {
        kk = i >> 6;
        nn = i & 63;
      }
      const int gk = k0 + kk, gn = n0 + nn;
      double v = 0.0;
      if (gk < op.K && gn < op.N) v = op.tb ? B[(size_t)gn * op.ldb + gk] : B[(size_t)gk * op.ldb + gn];
      Bs[kk][nn] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int gm = m0 + ty + 16 * i;
    if (gm >= op.M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gn = n0 + tx + 16 * j;
      if (gn >= op.N) continue;
      double* c = op.C + (size_t)gm * op.ldc + gn;
      *c = op.beta == 0.0 ? op.alpha * acc[i][j] : op.alpha * acc[i][j] + op.beta * *c;
    }
  }
}

// A fixed list of products run together: built on the host and uploaded once
// (when the set of molecules changes), then launched any number of times.
struct GemmList {
  std::vector<GemmOp> ops;
  DevBuf d_ops;
  int mt = 0, nt = 0, count = 0;
  void clear() { ops.clear(); }
  void add(bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int lda, const double* B, int ldb,
           double beta, double* C, int ldc, const int* skip = nullptr) {
    ops.push_back(GemmOp{A, B, C, M, N, K, lda, ldb, ldc, alpha, beta, ta ? 1 : 0, tb ? 1 : 0, skip});
  }
  // Copies the list to the device (synchronously: the host list may change next).
  void upload(hipStream_t s) {
    count = (int)ops.size();
    mt = nt = 0;
    for (const GemmOp& o : ops) {
      mt = std::max(mt, (o.M + 63) / 64);
      nt = std::max(nt, (o.N + 63) / 64);
    }
    if (!count) return;
    if (d_ops.bytes < ops.size() * sizeof(GemmOp)) {
      HIP_CHECK(hipStreamSynchronize(s));  // a previous launch may still read the old array
      d_ops.alloc(std::max<size_t>(ops.size(), 64) * sizeof(GemmOp));
    }
    HIP_CHECK(hipMemcpyAsync(d_ops.p, ops.data(), ops.size() * sizeof(GemmOp), hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
  void launch(hipStream_t s) const {
    for (int z0 = 0; z0 < count; z0 += 65535) {
      const unsigned nz = (unsigned)std::min(65535, count - z0);
      hipLaunchKernelGGL(k_gemm_batched, dim3(nt, mt, nz), dim3(256), 0, s, d_ops.as<GemmOp>() + z0);
    }
    HIP_CHECK(hipGetLastError());
  }
};

}  // namespace hfg
