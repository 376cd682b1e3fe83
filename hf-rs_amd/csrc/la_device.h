// la_device.h -- FP64 device linear algebra shared by the SCF drivers:
// row-major wrappers over rocBLAS (column-major) and rocSOLVER's dsyevd, plus
// small elementwise kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <string>
#include <vector>

#include "engine.h"
#include "scf_host.h"

namespace hfg {

#define RB_CHECK(expr)                                                                   \
  do {                                                                                   \
    rocblas_status _s = (expr);                                                          \
    if (_s != rocblas_status_success)                                                    \
      throw HfgError{HFG_DEVICE, std::string(#expr) + ": " + rocblas_status_to_string(_s)}; \
  } while (0)


__global__ inline void k_antisym(const double* a, double* out, int n) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * n) return;
  const int r = (int)(idx / n), c = (int)(idx % n);
  out[idx] = a[idx] - a[(size_t)c * n + r];
}
__global__ inline void k_sub(const double* a, const double* b, double* out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] - b[i];
}
__global__ inline void k_add(const double* a, const double* b, double* out, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}
// Second-order spectral projection (SP2, trace-correcting purification;
// A. M. N. Niklasson, Phys. Rev. B 66, 155115 (2002)) of a symmetric m x m
// matrix H onto its nocc lowest eigenvectors, in three kernels per step
// around one GEMM X2 = X X.
//
// Gershgorin bounds of H: b[0] = min_i (H_ii - sum_{j!=i} |H_ij|),
// b[1] = max_i (H_ii + sum_{j!=i} |H_ij|).  One workgroup of 256 threads.
__global__ inline void k_gershgorin(const double* __restrict__ h, int m, double* __restrict__ b) {
  __shared__ double lo[256], hi[256];
  double l = 1e300, u = -1e300;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    double r = 0.0;
    for (int j = 0; j < m; ++j)
      if (j != i) r += fabs(h[(size_t)i * m + j]);
    l = fmin(l, h[(size_t)i * m + i] - r);
    u = fmax(u, h[(size_t)i * m + i] + r);
  }
  lo[threadIdx.x] = l;
  hi[threadIdx.x] = u;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      lo[threadIdx.x] = fmin(lo[threadIdx.x], lo[threadIdx.x + s]);
      hi[threadIdx.x] = fmax(hi[threadIdx.x], hi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    b[0] = lo[0];
    b[1] = hi[0];
  }
}
// X = (emax I - H) / (emax - emin): the spectrum mapped into [0, 1], lowest
// energies nearest 1.
__global__ inline void k_sp2_init(const double* __restrict__ h, const double* __restrict__ b, double* __restrict__ x,
                                  int m) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)m * m) return;
  const double emin = b[0], emax = b[1];
  const double s = 1.0 / fmax(emax - emin, 1e-300);
  const size_t r = i / m, c = i % m;
  x[i] = ((r == c ? emax : 0.0) - h[i]) * s;
}
// t[0] = tr X, t[1] = tr X2; err[k] = tr X - tr X2 (= sum lambda (1 - lambda)).
// One workgroup of 256 threads.
__global__ inline void k_sp2_traces(const double* __restrict__ x, const double* __restrict__ x2, int m,
                                    double* __restrict__ t, double* __restrict__ err, int k) {
  __shared__ double a[256], q[256];
  double s0 = 0.0, s1 = 0.0;
  for (int i = threadIdx.x; i < m; i += blockDim.x) {
    s0 += x[(size_t)i * m + i];
    s1 += x2[(size_t)i * m + i];
  }
  a[threadIdx.x] = s0;
  q[threadIdx.x] = s1;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      a[threadIdx.x] += a[threadIdx.x + s];
      q[threadIdx.x] += q[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    t[0] = a[0];
    t[1] = q[0];
    err[k] = a[0] - q[0];
  }
}
// X <- X2 when that brings the trace nearer nocc, else X <- 2X - X2.
__global__ inline void k_sp2_update(double* __restrict__ x, const double* __restrict__ x2,
                                    const double* __restrict__ t, int m, double nocc) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)m * m) return;
  const double t0 = t[0], t1 = t[1];
  const bool square = fabs(t1 - nocc) < fabs(2.0 * t0 - t1 - nocc);
  x[i] = square ? x2[i] : 2.0 * x[i] - x2[i];
}

__global__ inline void k_zero(double* a, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Device linear algebra, row-major matrices on top of column-major rocBLAS.
// ---------------------------------------------------------------------------
struct Blas {
  rocblas_handle h = nullptr;
  hipStream_t s;
  explicit Blas(hipStream_t st) : s(st) {
    RB_CHECK(rocblas_create_handle(&h));
    RB_CHECK(rocblas_set_stream(h, st));
    RB_CHECK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
  }
  ~Blas() {
    if (h) rocblas_destroy_handle(h);
  }
  // C (MxN) = alpha op(A) op(B) + beta C, all row-major.
  void gemm(bool ta, bool tb, int M, int N, int K, double alpha, const double* A, int lda, const double* B, int ldb,
            double beta, double* C, int ldc) {
    RB_CHECK(rocblas_dgemm(h, tb ? rocblas_operation_transpose : rocblas_operation_none,
                           ta ? rocblas_operation_transpose : rocblas_operation_none, N, M, K, &alpha, B, ldb, A, lda,
                           &beta, C, ldc));
  }
  double dot(size_t n, const double* x, const double* y) {
    double r = 0;
    RB_CHECK(rocblas_ddot(h, (rocblas_int)n, x, 1, y, 1, &r));
    return r;
  }
  double nrm2(size_t n, const double* x) {
    double r = 0;
    RB_CHECK(rocblas_dnrm2(h, (rocblas_int)n, x, 1, &r));
    return r;
  }
  double amax(size_t n, const double* x) {
    rocblas_int i = 0;
    RB_CHECK(rocblas_idamax(h, (rocblas_int)n, x, 1, &i));
    double v = 0;
    HIP_CHECK(hipMemcpyAsync(&v, x + (i - 1), 8, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return std::fabs(v);
  }
};

struct Dev {
  DevBuf b;
  double* p = nullptr;
  void alloc(size_t n) {
    b.alloc(n * sizeof(double));
    p = b.as<double>();
  }
};

static void upload(double* d, const double* h, size_t n, hipStream_t s) {
  HIP_CHECK(hipMemcpyAsync(d, h, n * 8, hipMemcpyHostToDevice, s));
}
static void download(double* h, const double* d, size_t n, hipStream_t s) {
  HIP_CHECK(hipMemcpyAsync(h, d, n * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

// Projector onto the nocc lowest eigenvectors of a symmetric m x m matrix by
// SP2 purification (the k_sp2_* kernels): h holds the matrix and is
// overwritten (it serves as the X^2 buffer); the projector is left in x.
// The steps stop three after tr(X - X^2) < 1e-9 (checked every fourth step
// from the twelfth); false when 100 steps have not converged.
struct Sp2 {
  Dev x, b, t, err;
  size_t cap = 0;
  static constexpr int MAX_STEPS = 100;
  bool run(Blas& bl, hipStream_t st, double* h, int m, int nocc) {
    if (nocc <= 0 || nocc >= m) return false;
    const size_t mm = (size_t)m * m;
    if (mm > cap) {
      x.alloc(mm);
      cap = mm;
    }
    if (!b.p) {
      b.alloc(2);
      t.alloc(2);
      err.alloc(MAX_STEPS);
    }
    const dim3 g((unsigned)((mm + 255) / 256));
    hipLaunchKernelGGL(k_gershgorin, dim3(1), dim3(256), 0, st, h, m, b.p);
    hipLaunchKernelGGL(k_sp2_init, g, dim3(256), 0, st, h, b.p, x.p, m);
    int extra = -1;
    for (int k = 0; k < MAX_STEPS; ++k) {
      bl.gemm(false, false, m, m, m, 1.0, x.p, m, x.p, m, 0.0, h, m);  // X^2
      hipLaunchKernelGGL(k_sp2_traces, dim3(1), dim3(256), 0, st, x.p, h, m, t.p, err.p, k);
      hipLaunchKernelGGL(k_sp2_update, g, dim3(256), 0, st, x.p, h, t.p, m, (double)nocc);
      if (extra > 0 && --extra == 0) return true;
      if (extra < 0 && k >= 11 && (k + 1) % 4 == 0) {
        double e = 0.0;
        download(&e, err.p + k, 1, st);
        if (std::fabs(e) < 1e-9) extra = 3;
      }
    }
    return false;
  }
};

// Symmetric eigensolver: eigenvalues ascending into w, eigenvectors as
// columns of the row-major result V (computed as V^T by rocSOLVER's
// column-major output and transposed in the following GEMM by the caller).
struct Eigh {
  Dev work, e;
  DevBuf info;
  int n = 0;
  void run(Blas& bl, double* a_colmajor_vecs, double* w, int m) {
    if (n != m) {
      e.alloc(m);
      info.alloc(sizeof(rocblas_int));
      n = m;
    }
    RB_CHECK(rocsolver_dsyevd(bl.h, rocblas_evect_original, rocblas_fill_lower, m, a_colmajor_vecs, m, w, e.p,
                              info.as<rocblas_int>()));
  }
};


}  // namespace hfg
