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


// Gaussian elimination with partial pivoting (src/diis.rs:532-589); false if singular.
inline bool solve_small(std::vector<double> a, std::vector<double> b, int n, std::vector<double>& x) {
  for (int i = 0; i < n; ++i) {
    int piv = i;
    for (int k = i + 1; k < n; ++k)
      if (std::fabs(a[(size_t)k * n + i]) > std::fabs(a[(size_t)piv * n + i])) piv = k;
    if (std::fabs(a[(size_t)piv * n + i]) < 1e-28) return false;
    if (piv != i) {
      std::swap(b[i], b[piv]);
      for (int c = 0; c < n; ++c) std::swap(a[(size_t)i * n + c], a[(size_t)piv * n + c]);
    }
    for (int j = i + 1; j < n; ++j) {
      const double f = a[(size_t)j * n + i] / a[(size_t)i * n + i];
      for (int k = i; k < n; ++k) a[(size_t)j * n + k] -= f * a[(size_t)i * n + k];
      b[j] -= f * b[i];
    }
  }
  x.assign(n, 0.0);
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < n; ++k) s -= a[(size_t)i * n + k] * x[k];
    x[i] = s / a[(size_t)i * n + i];
  }
  return true;
}

}  // namespace hfg
