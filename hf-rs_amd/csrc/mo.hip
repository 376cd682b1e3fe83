// mo.hip -- AO -> MO transformation of the two-electron integrals on the GPU.
//
// transform_eri_ao_to_mo (src/mo_integrals.rs:156-265): (pq|rs) =
// sum C_mp C_nq C_ls C_sr (mn|ls), in two half transformations, each an
// unpacking of one index pair followed by two products: T = M C over the
// flattened (pair, lambda) rows, then C^T T_pair per pair (strided batches),
// packed back to r >= s.  The AO tensor is the screened packed tensor of
// compute_eri_tensor; every product runs on the device in FP64 (rocBLAS), so
// the O(N^5) work is GEMM-shaped and the host only builds the AO tensor.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstring>
#include <vector>

#include "../../include/hfg.h"
#include "engine.h"
#include "la_device.h"

namespace hfg {

__device__ __forceinline__ size_t pidx(size_t i, size_t j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// M[a][l][s] = packed[cidx(a, cidx(l, s))] for every pair a (np rows) and
// every (l, s) in n x n: the first index pair kept packed, the second unpacked.
__global__ void k_mo_unpack_ao(const double* __restrict__ packed, size_t np, int n, double* __restrict__ M) {
  const size_t nn = (size_t)n * n, total = np * nn;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t a = i / nn, ls = i % nn, l = ls / n, s = ls % n;
    M[i] = packed[pidx(a, pidx(l, s))];
  }
}

// H[a][rs] = U[a][r][s] (r >= s) for the half-transformed blocks U (np x m x m).
__global__ void k_mo_pack_half(const double* __restrict__ U, size_t np, int m, double* __restrict__ H) {
  const size_t mp = (size_t)m * (m + 1) / 2, total = np * mp;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t a = i / mp, rs = i % mp;
    size_t r = (size_t)((sqrt(8.0 * (double)rs + 1.0) - 1.0) / 2.0);
    while (r * (r + 1) / 2 > rs) --r;
    while ((r + 1) * (r + 2) / 2 <= rs) ++r;
    const size_t s = rs - r * (r + 1) / 2;
    H[i] = U[(a * m + r) * m + s];
  }
}

// D[rs][m][n] = H[cidx(m, n)][rs]: the second half's input, AO pair unpacked.
__global__ void k_mo_unpack_half(const double* __restrict__ H, size_t np, size_t mp, int n, double* __restrict__ D) {
  const size_t nn = (size_t)n * n, total = mp * nn;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t rs = i / nn, mn = i % nn, mu = mn / n, nu = mn % n;
    D[i] = H[pidx(mu, nu) * mp + rs];
  }
}

// Largest AO basis of the transformation (the unpacked intermediate holds
// np * n^2 doubles: 6.4 GB at n = 200).
constexpr int MO_MAX_N = 200;

void eri_mo(Engine& e, const double* C, int nmo, double* out) {
  const int n = e.n;
  if (n > MO_MAX_N) throw HfgError{HFG_INVALID_INPUT, "the AO->MO transformation is limited to 200 basis functions"};
  if (nmo <= 0 || nmo > n) throw HfgError{HFG_INVALID_INPUT, "number of MOs must be in 1..n"};
  const size_t np = (size_t)n * (n + 1) / 2, mp = (size_t)nmo * (nmo + 1) / 2, nn = (size_t)n * n;
  std::vector<double> packed(np * (np + 1) / 2);
  eri_packed_host(e, packed.data());
  hipStream_t st = e.stream;
  Blas bl(st);
  Dev dPk, dC, dM, dT, dU, dH;
  dPk.alloc(packed.size());
  upload(dPk.p, packed.data(), packed.size(), st);
  dC.alloc((size_t)n * nmo);
  upload(dC.p, C, (size_t)n * nmo, st);
  const size_t big = std::max(np, mp) * nn;
  dM.alloc(big);
  dT.alloc(std::max(np, mp) * (size_t)n * nmo);
  dU.alloc(std::max(np, mp) * (size_t)nmo * nmo);
  dH.alloc(np * mp);
  const dim3 g(4096), b(256);
  // first half: (a | r s) for every AO pair a
  hipLaunchKernelGGL(k_mo_unpack_ao, g, b, 0, st, dPk.p, np, n, dM.p);
  dPk.b.free();
  auto half = [&](size_t rows) {
    // T[(x, l)][s'] = sum_s M[(x, l)][s] C[s][s']  (one product over all rows)
    bl.gemm(false, false, (int)(rows * n), nmo, n, 1.0, dM.p, n, dC.p, nmo, 0.0, dT.p, nmo);
    // U[x] = C^T T[x] (nmo x nmo) for every x (strided batches, column-major view)
    const double one = 1.0, zero = 0.0;
    RB_CHECK(rocblas_dgemm_strided_batched(bl.h, rocblas_operation_none, rocblas_operation_transpose, nmo, nmo, n,
                                           &one, dT.p, nmo, (rocblas_stride)n * nmo, dC.p, nmo, 0, &zero, dU.p, nmo,
                                           (rocblas_stride)nmo * nmo, (rocblas_int)rows));
  };
  half(np);
  hipLaunchKernelGGL(k_mo_pack_half, g, b, 0, st, dU.p, np, nmo, dH.p);
  // second half: (p q | r s) for every MO pair rs
  hipLaunchKernelGGL(k_mo_unpack_half, g, b, 0, st, dH.p, np, mp, n, dM.p);
  half(mp);
  // out[rs][pq]
  Dev dO;
  dO.alloc(mp * mp);
  hipLaunchKernelGGL(k_mo_pack_half, g, b, 0, st, dU.p, mp, nmo, dO.p);
  HIP_CHECK(hipGetLastError());
  download(out, dO.p, mp * mp, st);
}

}  // namespace hfg
