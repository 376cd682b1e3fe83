// df.hip -- Coulomb-metric density fitting on the GPU: metric (P|Q), the
// 3-index tensor (P|mu nu), the whitened factor B = X^T (P|mu nu) with
// X = U_kept lambda^{-1/2} of the metric (canonical orthogonaliser, cutoff
// 1e-10), DF-J/K, and the DF-RHF driver.
//
// Follows src/df.rs:16-174 (metric_orthogonalizer, coulomb_metric,
// three_center_integrals, DfTensor::build, coulomb_exchange_rhf) and
// run_factorized_rhf (src/hf_variants/ri_df_hf.rs:160-263).  The integrals are
// the generated Head-Gordon-Pople blocks with the auxiliary shell paired with
// the exponent-0 unit function (src/df.rs:30-40); B is kept unpacked,
// [rank][n][n], so that DF-J is two FP64 GEMVs over B (HBM-bound) and DF-K is
// one batched GEMM A_Q = B_Q C plus one GEMM K = 2 A' A'^T over the
// rank*n_occ index (FP64 matrix cores through rocBLAS).
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <future>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <type_traits>

#include "../../include/hfg.h"
#include "cholesky.h"
#include "engine.h"
#include "la_device.h"
#include "rhf.h"
#include "scf_host.h"

namespace hfg {

// Block element -> 3-index tensor T[P][mu][nu] (both mu nu orders), scaled.
struct Scatter3 {
  long long off;
  int aux_bra;        // 1: block is (P 1|a b), else (a b|P 1)
  int na, nb, np;     // component counts (kernel orientation of the orbital pair)
  int mu0, nu0, p0;   // AO offsets of first / second orbital shell and aux shell
};

__global__ void k_scatter3(const Scatter3* __restrict__ d, int ntask, const double* __restrict__ blk,
                           const double* __restrict__ so, const double* __restrict__ sx, double* __restrict__ T,
                           int n) {
  const int t = blockIdx.x;
  if (t >= ntask) return;
  const Scatter3 s = d[t];
  const int tot = s.na * s.nb * s.np;
  const size_t nn = (size_t)n * n;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int a, b, p;
    if (s.aux_bra) {
      p = e / (s.na * s.nb);
      const int r = e - p * s.na * s.nb;
      a = r / s.nb;
      b = r - a * s.nb;
    } else {
      a = e / (s.nb * s.np);
      const int r = e - a * s.nb * s.np;
      b = r / s.np;
      p = r - b * s.np;
    }
    const int mu = s.mu0 + a, nu = s.nu0 + b, P = s.p0 + p;
    const double v = blk[s.off + e] * so[mu] * so[nu] * sx[P];
    T[(size_t)P * nn + (size_t)mu * n + nu] = v;
    T[(size_t)P * nn + (size_t)nu * n + mu] = v;
  }
}

// Packed layout: column ij = i(i+1)/2 + j (i >= j) of the pair (mu, nu),
// the src/scf.rs:221-235 compound index.  Block element -> T[P][ij - c0] of a
// column chunk (cols wide); a block of one shell with itself writes only its
// a >= b elements, so every entry is written once.
__global__ void k_scatter3p(const Scatter3* __restrict__ d, int ntask, const double* __restrict__ blk,
                            const double* __restrict__ so, const double* __restrict__ sx, double* __restrict__ T,
                            long long c0, long long cols) {
  const int t = blockIdx.x;
  if (t >= ntask) return;
  const Scatter3 s = d[t];
  const int tot = s.na * s.nb * s.np;
  const bool diag = s.mu0 == s.nu0;
  for (int e = threadIdx.x; e < tot; e += blockDim.x) {
    int a, b, p;
    if (s.aux_bra) {
      p = e / (s.na * s.nb);
      const int r = e - p * s.na * s.nb;
      a = r / s.nb;
      b = r - a * s.nb;
    } else {
      a = e / (s.nb * s.np);
      const int r = e - a * s.nb * s.np;
      b = r / s.np;
      p = r - b * s.np;
    }
    if (diag && a < b) continue;
    const int mu = s.mu0 + a, nu = s.nu0 + b, P = s.p0 + p;
    const long long i = mu > nu ? mu : nu, j = mu > nu ? nu : mu;
    T[(size_t)P * cols + (size_t)(i * (i + 1) / 2 + j - c0)] = blk[s.off + e] * so[mu] * so[nu] * sx[P];
  }
}

// Symmetric n x n -> packed with the off-diagonal doubled: sum_ij X_ij Y_ij
// = sum_(i>=j) packed(X)_ij Y_ij for symmetric X, Y.
__global__ void k_pack_sym2(const double* __restrict__ X, int n, double* __restrict__ Xp) {
  const size_t ij = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t np = (size_t)n * (n + 1) / 2;
  if (ij >= np) return;
  int ii = (int)((sqrt(8.0 * (double)ij + 1.0) - 1.0) * 0.5);
  while ((size_t)ii * (ii + 1) / 2 > ij) --ii;
  while ((size_t)(ii + 1) * (ii + 2) / 2 <= ij) ++ii;
  const int j = (int)(ij - (size_t)ii * (ii + 1) / 2);
  Xp[ij] = (ii == j ? 1.0 : 2.0) * X[(size_t)ii * n + j];
}

// d[Q] = sum_j B[Q][j] v[j] over the nc columns of rows Q (row stride ld), four
// rows a workgroup so that each load of v serves four rows; 16-byte loads
// where rows and v are 16-byte aligned; per row a fixed-order sum (strided
// lanes, then a tree over the workgroup).  B P of DF-J (the transposed GEMV).
constexpr int RD_ROWS = 4, RD_NT = 512;
__global__ __launch_bounds__(RD_NT) void k_rowdot(const double* __restrict__ B, size_t ld, size_t nc, int rows,
                                                  const double* __restrict__ v, double* __restrict__ d) {
  __shared__ double red[RD_ROWS][RD_NT];
  const int q0 = blockIdx.x * RD_ROWS, t = threadIdx.x;
  double s[RD_ROWS];
  const double* row[RD_ROWS];
#pragma unroll
  for (int r = 0; r < RD_ROWS; ++r) {
    s[r] = 0.0;
    row[r] = B + (size_t)(q0 + r < rows ? q0 + r : q0) * ld;  // a missing row repeats the first, unused
  }
  const bool vec = (ld % 2 == 0) && ((uintptr_t)B % 16 == 0) && ((uintptr_t)v % 16 == 0);
  if (vec) {
    const size_t n2 = nc / 2;
    const double2* v2 = (const double2*)v;
    for (size_t x = t; x < n2; x += RD_NT) {
      const double2 b = v2[x];
#pragma unroll
      for (int r = 0; r < RD_ROWS; ++r) {
        const double2 a = ((const double2*)row[r])[x];
        s[r] = fma(a.x, b.x, s[r]);
        s[r] = fma(a.y, b.y, s[r]);
      }
    }
    if ((nc & 1) && t == 0)
#pragma unroll
      for (int r = 0; r < RD_ROWS; ++r) s[r] = fma(row[r][nc - 1], v[nc - 1], s[r]);
  } else {
    for (size_t x = t; x < nc; x += RD_NT) {
      const double b = v[x];
#pragma unroll
      for (int r = 0; r < RD_ROWS; ++r) s[r] = fma(row[r][x], b, s[r]);
    }
  }
#pragma unroll
  for (int r = 0; r < RD_ROWS; ++r) red[r][t] = s[r];
  __syncthreads();
  for (int h = RD_NT / 2; h > 0; h >>= 1) {
    if (t < h)
#pragma unroll
      for (int r = 0; r < RD_ROWS; ++r) red[r][t] += red[r][t + h];
    __syncthreads();
  }
  if (t < RD_ROWS && q0 + t < rows) d[q0 + t] = red[t][0];
}

// Packed -> symmetric n x n.
__global__ void k_unpack_sym(const double* __restrict__ Xp, int n, double* __restrict__ X) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n * n) return;
  const int i = (int)(idx / n), j = (int)(idx - (size_t)i * n);
  const int a = i > j ? i : j, b = i > j ? j : i;
  X[idx] = Xp[(size_t)a * (a + 1) / 2 + b];
}

// Rows q < nq of packed B (row stride np) -> nq symmetric n x n matrices of
// type T (double, float, or the two BF16 parts of a split, hi and lo), by
// 64 x 64 tiles through LDS, so that both the packed reads (along j of a row
// i >= j) and the two writes (row i of the lower tile, row j of its
// transpose) are coalesced.  grid (lower tiles, nq), 256 threads.
template <class T>
__global__ __launch_bounds__(256) void k_unpack_tiles(const double* __restrict__ Bp, int n, size_t np, int nt,
                                                      T* __restrict__ out, __hip_bfloat16* __restrict__ lo) {
  __shared__ double tile[64 * 65];
  int ti = 0, rem = blockIdx.x;  // lower tile (ti >= tj), row-major over the triangle
  while (rem > ti) {
    rem -= ti + 1;
    ++ti;
  }
  const int tj = rem;
  const int q = blockIdx.y;
  const double* B = Bp + (size_t)q * np;
  const size_t nn = (size_t)n * n;
  T* O = out + (size_t)q * nn;
  __hip_bfloat16* Lo = lo ? lo + (size_t)q * nn : nullptr;
  const int i0 = ti * 64, j0 = tj * 64, t = threadIdx.x;
  const bool diag = ti == tj;
  auto put = [&](size_t at, double v) {
    if constexpr (std::is_same<T, __hip_bfloat16>::value) {  // as k_split_bf16
      const float f = (float)v;
      const __hip_bfloat16 h = __float2bfloat16(f);
      O[at] = h;
      Lo[at] = __float2bfloat16(f - __bfloat162float(h));
    } else {
      O[at] = (T)v;
    }
  };
  for (int r = 0; r < 16; ++r) {
    const int li = (t >> 6) + 4 * r, lj = t & 63;
    const int i = i0 + li, j = j0 + lj;
    double v = 0.0;
    if (i < n && j < n && (!diag || j <= i)) {
      v = B[(size_t)i * (i + 1) / 2 + j];
      put((size_t)i * n + j, v);
    }
    tile[li * 65 + lj] = v;
  }
  __syncthreads();
  for (int r = 0; r < 16; ++r) {
    const int lj = (t >> 6) + 4 * r, li = t & 63;  // out row j, along i
    const int i = i0 + li, j = j0 + lj;
    if (i < n && j < n && (!diag || i > j)) put((size_t)j * n + i, tile[li * 65 + lj]);
  }
}

// Tasks i < cnt of one (orbital pair class, auxiliary class) product,
// t = t0 + i = (orbital pair io, auxiliary shell ix) with ix fastest: the
// block task (kernel orientation) and its scatter record, at output offset
// i * NT.  Generated on the device: the product has ~1e8 entries for
// (H2O)32/cc-pVDZ with def2-universal-JKFIT, far too many to build and sort
// on the host.
__global__ void k_expand3(const int* __restrict__ opair, const int* __restrict__ oao_a, const int* __restrict__ oao_b,
                          const int* __restrict__ xsh, const int* __restrict__ xao, int nx, long long t0, int cnt,
                          int NT, int aux_bra, int norb_pairs, int na, int nb, int np, BlockTask* __restrict__ tasks,
                          Scatter3* __restrict__ sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cnt) return;
  const long long t = t0 + i;
  const int io = (int)(t / nx), ix = (int)(t - (long long)io * nx);
  const int op = opair[io], xp = norb_pairs + xsh[ix];
  const long long off = (long long)i * NT;
  tasks[i] = BlockTask{aux_bra ? xp : op, aux_bra ? op : xp, off};
  sc[i] = Scatter3{off, aux_bra, na, nb, np, oao_a[io], oao_b[io], xao[ix]};
}

// Metric block (P 1|Q 1) -> V[P][Q] and V[Q][P].
struct ScatterM {
  long long off;
  int np, nq, p0, q0;
};

__global__ void k_scatter_metric(const ScatterM* __restrict__ d, int ntask, const double* __restrict__ blk,
                                 const double* __restrict__ sx, double* __restrict__ V, int naux) {
  const int t = blockIdx.x;
  if (t >= ntask) return;
  const ScatterM s = d[t];
  for (int e = threadIdx.x; e < s.np * s.nq; e += blockDim.x) {
    const int i = e / s.nq, j = e - i * s.nq;
    const int P = s.p0 + i, Q = s.q0 + j;
    const double v = blk[s.off + e] * sx[P] * sx[Q];
    V[(size_t)P * naux + Q] = v;
    V[(size_t)Q * naux + P] = v;
  }
}

// A[Q][m][i] -> At[m][Q][i] (rank x n x o to n x rank x o).
__global__ void k_permute_qmi(const double* __restrict__ A, double* __restrict__ At, int rank, int n, int o) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t tot = (size_t)rank * n * o;
  if (idx >= tot) return;
  const int i = (int)(idx % o);
  const size_t r = idx / o;
  const int m = (int)(r % n);
  const int q = (int)(r / n);
  At[((size_t)m * rank + q) * o + i] = A[idx];
}

__global__ void k_axpby_fock(const double* H, const double* J, const double* K, double* F, size_t nn) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nn) F[i] = H[i] + J[i] - 0.5 * K[i];
}

// Precision conversions for DF-K on the matrix cores.
__global__ void k_to_f32(const double* __restrict__ x, float* __restrict__ y, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = (float)x[i];
}
// x = hi + lo with hi = bf16(x) and lo = bf16(x - hi) (round to nearest): the
// pair carries about 16 significant bits.
template <class T>
__global__ void k_split_bf16(const T* __restrict__ x, __hip_bfloat16* __restrict__ hi, __hip_bfloat16* __restrict__ lo,
                             size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float v = (float)x[i];
    const __hip_bfloat16 h = __float2bfloat16(v);
    hi[i] = h;
    lo[i] = __float2bfloat16(v - __bfloat162float(h));
  }
}
__global__ void k_f32_to_f64_scaled(const float* __restrict__ x, double* __restrict__ y, size_t n, double s) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = s * (double)x[i];
}
// A[Q][m][i] -> At[m][Q][i] for float data.
__global__ void k_permute_qmi_f32(const float* __restrict__ A, float* __restrict__ At, int rank, int n, int o) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t tot = (size_t)rank * n * o;
  if (idx >= tot) return;
  const int i = (int)(idx % o);
  const int m = (int)((idx / o) % n);
  const int q = (int)(idx / ((size_t)o * n));
  At[((size_t)m * rank + q) * o + i] = A[idx];
}

// K = 2 sum_b part[b] over the split-K partial products.
template <class T>
__global__ void k_sum_parts(const T* __restrict__ part, int nb, size_t nn, double* __restrict__ K) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += (double)part[(size_t)b * nn + i];
    K[i] = 2.0 * s;
  }
}

// df_mfma.hip: K = alpha X X^T, and A = B C from the packed factor, on the
// FP64 matrix cores.
void syrk_f64(const double* X, int n, long long L, long long ldx, double alpha, double* K, DevBuf& scratch,
              hipStream_t st);
void bc_packed_f64(const double* Bp, size_t np, int rank, int n, const double* C, int o, double* At,
                   long long ldt, double alpha, double* gamma, DevBuf& scratch, hipStream_t st);
void bc_packed_bf16x3(const double* Bp, size_t np, int rank, int n, const double* C, int o, unsigned short* Ah,
                      unsigned short* Al, long long ldt, DevBuf& ct, hipStream_t st);

struct DfEngine {
  hipStream_t stream = nullptr;
  std::unique_ptr<Blas> bl;
  int n = 0, naux = 0, rank = 0;
  std::vector<Shell> orb, aux;
  // B: rank x n x n, or rank x n(n+1)/2 packed (columns ij = i(i+1)/2 + j,
  // i >= j) when the unpacked 3-index tensor and B would not fit in HBM
  bool packed = false;
  size_t npk = 0;
  DevBuf d_B;
  DevBuf d_pk, d_unp;  // packed path: packed P and J; unpacked rows of B for DF-K
  DevBuf d_syrk;       // partial tiles of the hand-written exchange product
  DevBuf d_bcp;        // tile partials of B_Q . P from the packed A = B C
  DevBuf d_ct;         // split BF16 C^T of the packed split-BF16 A = B C
  // FP64 exchange product by the hand-written MFMA kernel (HFG_DF_SYRK=0: rocBLAS dgemm)
  const bool use_syrk = [] {
    const char* e = std::getenv("HFG_DF_SYRK");
    return !(e && std::atoi(e) == 0);
  }();
  // packed factor, FP64 and split BF16: A = B C by the hand-written kernels
  // straight from the packed rows (HFG_DF_BCPACK=0: unpacked chunks and
  // rocBLAS GEMMs)
  // d = B P by k_rowdot (HFG_DF_ROWDOT=0: rocBLAS dgemv)
  const bool use_rowdot = [] {
    const char* e = std::getenv("HFG_DF_ROWDOT");
    return !(e && std::atoi(e) == 0);
  }();
  const bool use_bcpack = [] {
    const char* e = std::getenv("HFG_DF_BCPACK");
    return !(e && std::atoi(e) == 0);
  }();
  DevBuf d_work, d_A, d_At;
  size_t work_o = 0;
  // reduced-precision copies of B (made on first use) and work arrays
  DevBuf d_B32, d_Bh, d_Bl, d_C32, d_Ch, d_Cl, d_A32, d_At32, d_Ah, d_Al, d_K32, d_parts;
  size_t work_o32 = 0;
  // Split-K of the reduced-precision K = 2 At At^T (n x n output, inner
  // dimension rank*o): the output alone gives too few tiles for the chip, so
  // the auxiliary index is cut into nsplit blocks computed as one strided
  // batch, then summed in FP64.
  static constexpr size_t T_CHUNK = 8ull << 30;  // bytes of one column chunk of the packed 3-index tensor
  int nsplit() const {
    const int tiles = std::max(1, ((n + 127) / 128) * ((n + 127) / 128));
    return std::max(1, std::min(rank, (1024 + tiles - 1) / tiles));
  }

  DfEngine(const hfg_basis_view& basis, const hfg_basis_view& auxv, double cutoff) {
    if (!(cutoff > 0) || !std::isfinite(cutoff))
      throw HfgError{HFG_INVALID_INPUT, "DF metric cutoff must be finite and positive"};
    orb = group_shells(basis);
    aux = group_shells(auxv);
    n = basis.n_functions;
    naux = auxv.n_functions;
    npk = (size_t)n * (n + 1) / 2;
    for (auto& s : orb)
      if (s.L > 3) throw HfgError{HFG_UNSUPPORTED, "density fitting compiled for orbital shells up to f"};
    for (auto& s : aux)
      if (s.L > 4) throw HfgError{HFG_UNSUPPORTED, "density fitting compiled for auxiliary shells up to g"};
    // HBM footprint: unpacked, the 3-index tensor and B at naux n^2 doubles
    // each; packed, B at naux n(n+1)/2 and the tensor one column chunk at a
    // time (T_CHUNK bytes), so that config-size bases fit.
    size_t free_b = 0, total_b = 0;
    HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    const double metric = 8.0 * 2.0 * (double)naux * naux;
    const double unpacked = 8.0 * 2.0 * (double)naux * n * n + metric;
    const char* force = std::getenv("HFG_DF_PACKED");
    packed = force ? std::atoi(force) != 0 : unpacked > 0.4 * (double)free_b;
    if (packed && 8.0 * (double)naux * npk + (double)T_CHUNK + metric > 0.8 * (double)free_b)
      throw HfgError{HFG_DEVICE, "DF factor B does not fit in the free device memory"};
    HIP_CHECK(hipStreamCreate(&stream));
    bl.reset(new Blas(stream));
    build(cutoff);
  }
  // Factors of the pivoted Cholesky decomposition of the AO ERIs
  // (src/cholesky.rs:44-192, cholesky.hip) in place of the fitted B, for
  // CD-HF: n_aux = rank = the number of Cholesky vectors.
  DfEngine(const hfg_basis_view& basis, double threshold, double budget, CholeskyInfo& info) {
    orb = group_shells(basis);
    n = basis.n_functions;
    for (auto& s : orb)
      if (s.L > 2) throw HfgError{HFG_UNSUPPORTED, "Cholesky factors compiled for orbital shells up to d"};
    HIP_CHECK(hipStreamCreate(&stream));
    try {
      bl.reset(new Blas(stream));
      cholesky_factors(orb, n, stream, *bl, threshold, budget, d_B, rank, info);
    } catch (...) {
      bl.reset();
      (void)hipStreamDestroy(stream);
      stream = nullptr;
      throw;
    }
    naux = rank;
  }
  ~DfEngine() {
    bl.reset();
    if (stream) (void)hipStreamDestroy(stream);
  }

  void build(double cutoff) {
    // pair table: orbital pairs (a >= b) then auxiliary pairs (P, unit)
    std::vector<PairInfo> pairs;
    PrimPairHost pp;
    build_pairs(orb, pairs, pp);
    const int norb_pairs = (int)pairs.size();
    for (size_t p = 0; p < aux.size(); ++p) {
      PairInfo pi;
      std::memset(&pi, 0, sizeof pi);
      pi.p0 = (int)pp.size();
      append_aux_pair(aux[p], pp);
      pi.p1 = (int)pp.size();
      pi.ao_a = aux[p].ao;
      pi.ao_b = 0;
      pi.sh_a = (int)p;
      pi.sh_b = -1;
      pi.la = aux[p].L;
      pi.lb = 0;
      pairs.push_back(pi);
    }
    const size_t np = pp.size();
    DevBuf d_pp, d_pairs, d_boys, d_so, d_sx;
    d_pp.alloc(9 * np * sizeof(double));
    const std::vector<double>* cols[9] = {&pp.p, &pp.Px, &pp.Py, &pp.Pz, &pp.PAx, &pp.PAy, &pp.PAz, &pp.K, &pp.H};
    for (int c = 0; c < 9; ++c)
      HIP_CHECK(hipMemcpyAsync(d_pp.as<double>() + c * np, cols[c]->data(), np * 8, hipMemcpyHostToDevice, stream));
    std::vector<PairDev> pd(pairs.size());
    for (size_t i = 0; i < pairs.size(); ++i) {
      std::memset(&pd[i], 0, sizeof(PairDev));
      pd[i].p0 = pairs[i].p0;
      pd[i].p1 = pairs[i].p1;
      for (int k = 0; k < 3; ++k) pd[i].AB[k] = pairs[i].AB[k];
    }
    d_pairs.alloc(pd.size() * sizeof(PairDev));
    HIP_CHECK(hipMemcpyAsync(d_pairs.p, pd.data(), pd.size() * sizeof(PairDev), hipMemcpyHostToDevice, stream));
    auto tab = boys_table();
    d_boys.alloc(tab.size() * 8);
    HIP_CHECK(hipMemcpyAsync(d_boys.p, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, stream));
    std::vector<double> so(n), sx(naux);
    for (auto& s : orb)
      for (size_t c = 0; c < s.scale.size(); ++c) so[s.ao + c] = s.scale[c];
    for (auto& s : aux)
      for (size_t c = 0; c < s.scale.size(); ++c) sx[s.ao + c] = s.scale[c];
    d_so.alloc(n * 8);
    d_sx.alloc(naux * 8);
    HIP_CHECK(hipMemcpyAsync(d_so.p, so.data(), n * 8, hipMemcpyHostToDevice, stream));
    HIP_CHECK(hipMemcpyAsync(d_sx.p, sx.data(), naux * 8, hipMemcpyHostToDevice, stream));
    const PrimPairs view{d_pp.as<double>(),          d_pp.as<double>() + np,     d_pp.as<double>() + 2 * np,
                         d_pp.as<double>() + 3 * np, d_pp.as<double>() + 4 * np, d_pp.as<double>() + 5 * np,
                         d_pp.as<double>() + 6 * np, d_pp.as<double>() + 7 * np, d_pp.as<double>() + 8 * np};
    auto nc = [](int L) { return (L + 1) * (L + 2) / 2; };

    // Run block tasks grouped by class in bounded batches, then scatter.
    auto run_blocks = [&](std::vector<BlockTask>& tasks, std::vector<int>& cls, auto&& scatter) {
      std::vector<size_t> ord(tasks.size());
      std::iota(ord.begin(), ord.end(), 0);
      std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return cls[x] < cls[y]; });
      const long long budget = 1ll << 26;  // doubles per batch (512 MB)
      size_t i = 0;
      DevBuf dt, dout;
      while (i < ord.size()) {
        std::vector<BlockTask> bt;
        std::vector<int> bc;
        std::vector<size_t> which;
        long long off = 0;
        while (i < ord.size() && off < budget) {
          BlockTask t = tasks[ord[i]];
          const long long sz = t.off;  // size stashed by the caller
          t.off = off;
          bt.push_back(t);
          bc.push_back(cls[ord[i]]);
          which.push_back(ord[i]);
          off += sz;
          ++i;
        }
        dt.alloc(bt.size() * sizeof(BlockTask));
        dout.alloc(std::max<long long>(1, off) * 8);
        HIP_CHECK(hipMemcpyAsync(dt.p, bt.data(), bt.size() * sizeof(BlockTask), hipMemcpyHostToDevice, stream));
        size_t a = 0;
        while (a < bt.size()) {
          size_t b = a;
          while (b < bt.size() && bc[b] == bc[a]) ++b;
          BlockArgs args{dt.as<BlockTask>() + a, (int)(b - a), d_pairs.as<PairDev>(), view, d_boys.as<double>(),
                         dout.as<double>()};
          launch_block_class(bc[a], args, stream);
          a = b;
        }
        scatter(bt, which, dout.as<double>());
        HIP_CHECK(hipStreamSynchronize(stream));
      }
    };

    // Metric (P|Q), Q <= P.
    DevBuf d_V;
    d_V.alloc((size_t)naux * naux * 8);
    {
      std::vector<BlockTask> tasks;
      std::vector<int> cls;
      std::vector<ScatterM> sm;
      for (int p = 0; p < (int)aux.size(); ++p)
        for (int q = 0; q <= p; ++q) {
          int b = norb_pairs + p, k = norb_pairs + q;
          int pb = pair_class(aux[p].L, 0), pk = pair_class(aux[q].L, 0);
          if (pk > pb) {
            std::swap(b, k);
            std::swap(pb, pk);
          }
          const int s0 = b - norb_pairs, s1 = k - norb_pairs;
          const int n0 = nc(aux[s0].L), n1 = nc(aux[s1].L);
          tasks.push_back({b, k, (long long)n0 * n1});
          cls.push_back(quartet_class(pb, pk));
          sm.push_back({0, n0, n1, aux[s0].ao, aux[s1].ao});
        }
      DevBuf dsm;
      run_blocks(tasks, cls, [&](std::vector<BlockTask>& bt, std::vector<size_t>& which, double* blk) {
        std::vector<ScatterM> h(bt.size());
        for (size_t i = 0; i < bt.size(); ++i) {
          h[i] = sm[which[i]];
          h[i].off = bt[i].off;
        }
        dsm.alloc(h.size() * sizeof(ScatterM));
        HIP_CHECK(hipMemcpyAsync(dsm.p, h.data(), h.size() * sizeof(ScatterM), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_scatter_metric, dim3(h.size()), dim3(64), 0, stream, dsm.as<ScatterM>(), (int)h.size(),
                           blk, d_sx.as<double>(), d_V.as<double>(), naux);
        HIP_CHECK(hipGetLastError());
      });
    }
    // X = U_kept lambda^{-1/2} (src/matrix.rs:329-346 with cutoff)
    Dev dw;
    dw.alloc(naux);
    Eigh eig;
    eig.run(*bl, d_V.as<double>(), dw.p, naux);
    std::vector<double> w(naux);
    download(w.data(), dw.p, naux, stream);
    if (w[0] < -cutoff) throw HfgError{HFG_INTEGRAL, "metric has a negative eigenvalue"};
    int first = 0;
    while (first < naux && w[first] < cutoff) ++first;
    rank = naux - first;
    if (rank <= 0) throw HfgError{HFG_INTEGRAL, "metric has zero retained rank"};
    // rows of d_V's row-major view are eigenvectors; X^T rows = kept vectors / sqrt(lambda)
    std::vector<double> U((size_t)naux * naux);
    download(U.data(), d_V.as<double>(), (size_t)naux * naux, stream);
    std::vector<double> Xt((size_t)rank * naux);
    for (int r = 0; r < rank; ++r)
      for (int c = 0; c < naux; ++c) Xt[(size_t)r * naux + c] = U[(size_t)(first + r) * naux + c] / std::sqrt(w[first + r]);
    DevBuf d_Xt;
    d_Xt.alloc(Xt.size() * 8);
    HIP_CHECK(hipMemcpyAsync(d_Xt.p, Xt.data(), Xt.size() * 8, hipMemcpyHostToDevice, stream));
    d_V.free();

    // 3-index tensor T[P][mu][nu]: every orbital pair against every
    // auxiliary shell (src/df.rs:60-78), one (pair class, auxiliary class)
    // product at a time, in batches of at most `budget` block doubles, then
    // B = X^T T.  Packed, the tensor is formed one chunk of function rows
    // [f0, f1) at a time -- its packed columns f0(f0+1)/2 .. f1(f1+1)/2 - 1,
    // the pairs whose later shell starts in the chunk -- and each chunk's
    // columns of B follow from one GEMM, so the tensor never exists whole.
    const size_t nn = (size_t)n * n;
    std::vector<std::pair<int, int>> rows;  // function-row chunks
    if (!packed) {
      rows.push_back({0, n});
    } else {
      std::vector<int> starts;
      for (auto& sh : orb) starts.push_back(sh.ao);
      std::sort(starts.begin(), starts.end());
      starts.push_back(n);
      auto cols_of = [](long long f0, long long f1) { return f1 * (f1 + 1) / 2 - f0 * (f0 + 1) / 2; };
      size_t k = 0;
      while (k + 1 < starts.size()) {
        size_t e = k + 1;
        while (e + 1 < starts.size() && 8.0 * naux * cols_of(starts[k], starts[e + 1]) <= (double)T_CHUNK) ++e;
        rows.push_back({starts[k], starts[e]});
        k = e;
      }
    }
    d_B.alloc((size_t)rank * (packed ? npk : nn) * 8);
    DevBuf d_T;
    std::map<int, std::vector<int>> by_xc;  // auxiliary class -> shells
    for (int p = 0; p < (int)aux.size(); ++p) by_xc[pair_class(aux[p].L, 0)].push_back(p);
    const long long budget = 1ll << 26;
    DevBuf d_opair, d_oa, d_ob, d_xsh, d_xao, dt, ds3, dout;
    size_t cap_tasks = 0, cap_out = 0;
    for (const auto& rc : rows) {
      const long long c0 = (long long)rc.first * (rc.first + 1) / 2;
      const long long cols = packed ? (long long)rc.second * (rc.second + 1) / 2 - c0 : (long long)nn;
      const size_t tsz = (size_t)naux * cols;
      if (d_T.bytes < tsz * 8) d_T.alloc(tsz * 8);
      std::map<int, std::vector<int>> by_oc;  // orbital pair class -> pairs of this chunk
      for (int i = 0; i < norb_pairs; ++i) {
        const int row = std::max(pairs[i].ao_a, pairs[i].ao_b);
        if (row >= rc.first && row < rc.second) by_oc[pair_class(pairs[i].la, pairs[i].lb)].push_back(i);
      }
      for (auto& oc : by_oc) {
        const std::vector<int>& ol = oc.second;
        std::vector<int> oa(ol.size()), ob(ol.size());
        for (size_t k = 0; k < ol.size(); ++k) {
          oa[k] = pairs[ol[k]].ao_a;
          ob[k] = pairs[ol[k]].ao_b;
        }
        d_opair.alloc(ol.size() * sizeof(int));
        d_oa.alloc(ol.size() * sizeof(int));
        d_ob.alloc(ol.size() * sizeof(int));
        HIP_CHECK(hipMemcpyAsync(d_opair.p, ol.data(), ol.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d_oa.p, oa.data(), ol.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(d_ob.p, ob.data(), ol.size() * sizeof(int), hipMemcpyHostToDevice, stream));
        const PairInfo& o0 = pairs[ol[0]];
        for (auto& xc : by_xc) {
          const std::vector<int>& xl = xc.second;
          std::vector<int> xao(xl.size());
          for (size_t k = 0; k < xl.size(); ++k) xao[k] = aux[xl[k]].ao;
          d_xsh.alloc(xl.size() * sizeof(int));
          d_xao.alloc(xl.size() * sizeof(int));
          HIP_CHECK(hipMemcpyAsync(d_xsh.p, xl.data(), xl.size() * sizeof(int), hipMemcpyHostToDevice, stream));
          HIP_CHECK(hipMemcpyAsync(d_xao.p, xao.data(), xl.size() * sizeof(int), hipMemcpyHostToDevice, stream));
          const int pco = oc.first, pcx = xc.first;
          const bool aux_bra = pcx > pco;
          const int cls = aux_bra ? quartet_class(pcx, pco) : quartet_class(pco, pcx);
          const int na = nc(o0.la), nb = nc(o0.lb), np_ = nc(aux[xl[0]].L), NT = na * nb * np_;
          const long long total = (long long)ol.size() * (long long)xl.size();
          const long long chunk = std::max<long long>(1, std::min<long long>(budget / NT, 1ll << 24));
          for (long long t0 = 0; t0 < total; t0 += chunk) {
            const int cnt = (int)std::min(chunk, total - t0);
            if ((size_t)cnt > cap_tasks) {
              dt.alloc((size_t)cnt * sizeof(BlockTask));
              ds3.alloc((size_t)cnt * sizeof(Scatter3));
              cap_tasks = cnt;
            }
            if ((size_t)cnt * NT > cap_out) {
              dout.alloc((size_t)cnt * NT * 8);
              cap_out = (size_t)cnt * NT;
            }
            hipLaunchKernelGGL(k_expand3, dim3((cnt + 255) / 256), dim3(256), 0, stream, d_opair.as<int>(),
                               d_oa.as<int>(), d_ob.as<int>(), d_xsh.as<int>(), d_xao.as<int>(), (int)xl.size(), t0,
                               cnt, NT, aux_bra ? 1 : 0, norb_pairs, na, nb, np_, dt.as<BlockTask>(),
                               ds3.as<Scatter3>());
            HIP_CHECK(hipGetLastError());
            BlockArgs args{dt.as<BlockTask>(), cnt, d_pairs.as<PairDev>(), view, d_boys.as<double>(), dout.as<double>()};
            if (cls < HFG_NCLASS_ALL)
              launch_block_class(cls, args, stream);
            else  // an orbital f-p, f-d or f-f pair (never on the auxiliary side)
              launch_df_f_block(o0.la, o0.lb, aux[xl[0]].L, 0, args, stream);
            if (packed)
              hipLaunchKernelGGL(k_scatter3p, dim3(cnt), dim3(64), 0, stream, ds3.as<Scatter3>(), cnt,
                                 dout.as<double>(), d_so.as<double>(), d_sx.as<double>(), d_T.as<double>(), c0, cols);
            else
              hipLaunchKernelGGL(k_scatter3, dim3(cnt), dim3(64), 0, stream, ds3.as<Scatter3>(), cnt,
                                 dout.as<double>(), d_so.as<double>(), d_sx.as<double>(), d_T.as<double>(), n);
            HIP_CHECK(hipGetLastError());
          }
          // the lists are rewritten for the next product
          HIP_CHECK(hipStreamSynchronize(stream));
        }
      }
      // B[:, chunk] = X^T T_chunk  (rank x naux) (naux x cols)
      const int ldb = (int)(packed ? npk : nn);
      bl->gemm(false, false, rank, (int)cols, naux, 1.0, d_Xt.as<double>(), naux, d_T.as<double>(), (int)cols, 0.0,
               d_B.as<double>() + (packed ? c0 : 0), ldb);
      HIP_CHECK(hipStreamSynchronize(stream));
    }
  }

  // Packed B: f(q0, qn, W, lo) for chunks of rows q0 .. q0+qn-1 unpacked to
  // qn symmetric n x n matrices W of type T (for BF16, W the high and lo the
  // low part of the split), about 2 GB of scratch at a time.
  template <class T, class F>
  void for_unpacked_rows(F&& f) {
    const size_t nn = (size_t)n * n;
    const int qc = (int)std::max<size_t>(1, std::min<size_t>(rank, (2ull << 30) / (nn * 8)));
    if (d_unp.bytes < (size_t)qc * nn * 8) d_unp.alloc((size_t)qc * nn * 8);
    T* W = d_unp.as<T>();
    __hip_bfloat16* lo = std::is_same<T, __hip_bfloat16>::value ? d_unp.as<__hip_bfloat16>() + (size_t)qc * nn
                                                                 : nullptr;
    for (int q0 = 0; q0 < rank; q0 += qc) {
      const int qn = std::min(qc, rank - q0);
      const int nt = (n + 63) / 64;
      hipLaunchKernelGGL(k_unpack_tiles<T>, dim3((unsigned)(nt * (nt + 1) / 2), (unsigned)qn), dim3(256), 0, stream,
                         d_B.as<double>() + (size_t)q0 * npk, n, npk, nt, W, lo);
      HIP_CHECK(hipGetLastError());
      f(q0, qn, (const T*)W, (const __hip_bfloat16*)lo);
    }
  }

  // J = sum_Q B_Q (B_Q . P), P = 2 C C^T; K = 2 sum_Q (B_Q C)(B_Q C)^T
  // (src/df.rs:143-173).  dC: n x o device; outputs device n x n.
  void jk(const double* dC, int o, double* dJ, double* dK) {
    const size_t nn = (size_t)n * n;
    // packed factor with K: A = B C by the packed kernel straight into the
    // exchange layout, B P taken from its tiles (B_Q . P = 2 sum A_Q . C)
    const bool fused = packed && use_bcpack && dK;
    if ((size_t)o != work_o) {
      d_work.alloc(std::max(nn, (size_t)rank) * 8 + 8);
      d_A.free();
      d_At.alloc((size_t)rank * n * o * 8);
      work_o = o;
    }
    if (dK && !fused && d_A.bytes < (size_t)rank * n * o * 8) d_A.alloc((size_t)rank * n * o * 8);
    double* P = d_work.as<double>();
    Dev dq;
    dq.alloc(rank);
    bl->gemm(false, true, n, n, o, 2.0, dC, o, dC, o, 0.0, P, n);
    const double one = 1.0, zero = 0.0;
    const size_t bc = packed ? npk : nn;  // columns of B
    const double* pv = P;
    double* jv = dJ;
    if (packed) {
      // packed P with the off-diagonal doubled, so that B P over the packed
      // columns is the full contraction; J comes back packed
      if (d_pk.bytes < 2 * npk * 8) d_pk.alloc(2 * npk * 8);
      hipLaunchKernelGGL(k_pack_sym2, dim3((unsigned)((npk + 255) / 256)), dim3(256), 0, stream, P, n,
                         d_pk.as<double>());
      HIP_CHECK(hipGetLastError());
      pv = d_pk.as<double>();
      jv = d_pk.as<double>() + npk;
    }
    // column-major view of B: bc x rank (B^T); d = B P  ->  gemv(T)
    if (fused)
      bc_packed_f64(d_B.as<double>(), npk, rank, n, dC, o, d_At.as<double>(), (long long)rank * o, 2.0, dq.p, d_bcp,
                    stream);
    else if (use_rowdot) {
      hipLaunchKernelGGL(k_rowdot, dim3((unsigned)((rank + RD_ROWS - 1) / RD_ROWS)), dim3(RD_NT), 0, stream,
                         d_B.as<double>(), bc, bc, rank, pv, dq.p);
      HIP_CHECK(hipGetLastError());
    } else
      RB_CHECK(rocblas_dgemv(bl->h, rocblas_operation_transpose, (rocblas_int)bc, rank, &one, d_B.as<double>(),
                             (rocblas_int)bc, pv, 1, &zero, dq.p, 1));
    RB_CHECK(rocblas_dgemv(bl->h, rocblas_operation_none, (rocblas_int)bc, rank, &one, d_B.as<double>(),
                           (rocblas_int)bc, dq.p, 1, &zero, jv, 1));
    if (packed) {
      hipLaunchKernelGGL(k_unpack_sym, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, stream, jv, n, dJ);
      HIP_CHECK(hipGetLastError());
    }
    if (!dK) return;  // J only
    if (!fused) {
      // A = B C: rows (Q, m), (rank*n) x o, then permuted to At[m][Q o + i]
      if (!packed)
        bl->gemm(false, false, rank * n, o, n, 1.0, d_B.as<double>(), n, dC, o, 0.0, d_A.as<double>(), o);
      else
        for_unpacked_rows<double>([&](int q0, int qn, const double* W, const __hip_bfloat16*) {
          bl->gemm(false, false, qn * n, o, n, 1.0, W, n, dC, o, 0.0, d_A.as<double>() + (size_t)q0 * n * o, o);
        });
      const size_t tot = (size_t)rank * n * o;
      hipLaunchKernelGGL(k_permute_qmi, dim3((tot + 255) / 256), dim3(256), 0, stream, d_A.as<double>(),
                         d_At.as<double>(), rank, n, o);
      HIP_CHECK(hipGetLastError());
    }
    // FP64: the symmetric product on the matrix cores (upper tiles only,
    // split over the contraction index); or one rocBLAS GEMM (its dsyrk,
    // with half the flops, ran 2.1x slower than the GEMM on (H2O)32/cc-pVDZ)
    if (use_syrk)
      syrk_f64(d_At.as<double>(), n, (long long)rank * o, (long long)rank * o, 2.0, dK, d_syrk, stream);
    else
      bl->gemm(false, true, n, n, rank * o, 2.0, d_At.as<double>(), rank * o, d_At.as<double>(), rank * o, 0.0, dK,
               n);
  }

  // DF-K in reduced precision on the matrix cores (J stays FP64, HBM-bound):
  // prec 1: FP32 operands, FP32 accumulation; prec 2: each operand split into
  // two BF16 parts, three BF16 products (hi*hi + hi*lo + lo*hi) with FP32
  // accumulation, for both A = B C and K = 2 At At^T.  No reference
  // counterpart: the reference's DF-K is FP64 (src/df.rs:143-173).
  void gemm_ex(bool ta, bool tb, int m, int nn_, int k, const void* A, int lda, const void* B, int ldb, float* C,
               int ldc, float alpha, float beta, rocblas_datatype in) {
    // row-major C = op(A) op(B) through column-major C^T = op(B)^T op(A)^T
    RB_CHECK(rocblas_gemm_ex(bl->h, tb ? rocblas_operation_transpose : rocblas_operation_none,
                             ta ? rocblas_operation_transpose : rocblas_operation_none, nn_, m, k, &alpha, B, in, ldb,
                             A, in, lda, &beta, C, rocblas_datatype_f32_r, ldc, C, rocblas_datatype_f32_r, ldc,
                             rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0));
  }
  void k_reduced(const double* dC, int o, double* dK, int prec) {
    const size_t nn = (size_t)n * n, bsz = (size_t)rank * nn, csz = (size_t)n * o, asz = (size_t)rank * n * o;
    const dim3 g(2048), b(256);
    if (prec == 1 && !packed && !d_B32.p) {
      d_B32.alloc(bsz * 4);
      hipLaunchKernelGGL(k_to_f32, g, b, 0, stream, d_B.as<double>(), d_B32.as<float>(), bsz);
    }
    if (prec == 2 && !packed && !d_Bh.p) {
      d_Bh.alloc(bsz * 2);
      d_Bl.alloc(bsz * 2);
      hipLaunchKernelGGL(k_split_bf16<double>, g, b, 0, stream, d_B.as<double>(), d_Bh.as<__hip_bfloat16>(),
                         d_Bl.as<__hip_bfloat16>(), bsz);
    }
    // packed factor, split BF16: the hi and lo operands of the symmetric
    // product straight from the packed rows (no unpack, FP32 A or permute)
    const bool direct = prec == 2 && packed && use_bcpack;
    if ((size_t)o != work_o32) {
      d_C32.alloc(csz * 4);
      d_Ch.alloc(csz * 2);
      d_Cl.alloc(csz * 2);
      d_A32.free();
      d_At32.free();
      d_Ah.alloc(asz * 2);
      d_Al.alloc(asz * 2);
      d_K32.alloc(nn * 4);
      work_o32 = o;
    }
    if (!direct && !d_A32.p) {
      d_A32.alloc(asz * 4);
      d_At32.alloc(asz * 4);
    }
    float* A = d_A32.as<float>();
    float* At = d_At32.as<float>();
    float* K = d_K32.as<float>();
    if (direct) {
      bc_packed_bf16x3(d_B.as<double>(), npk, rank, n, dC, o, (unsigned short*)d_Ah.p, (unsigned short*)d_Al.p,
                       (long long)rank * o, d_ct, stream);
    } else if (prec == 1) {
      hipLaunchKernelGGL(k_to_f32, g, b, 0, stream, dC, d_C32.as<float>(), csz);
      if (!packed)
        gemm_ex(false, false, rank * n, o, n, d_B32.p, n, d_C32.p, o, A, o, 1.f, 0.f, rocblas_datatype_f32_r);
      else
        for_unpacked_rows<float>([&](int q0, int qn, const float* W, const __hip_bfloat16*) {
          gemm_ex(false, false, qn * n, o, n, W, n, d_C32.p, o, A + (size_t)q0 * n * o, o, 1.f, 0.f,
                  rocblas_datatype_f32_r);
        });
    } else {
      hipLaunchKernelGGL(k_split_bf16<double>, g, b, 0, stream, dC, d_Ch.as<__hip_bfloat16>(),
                         d_Cl.as<__hip_bfloat16>(), csz);
      const auto bf = rocblas_datatype_bf16_r;
      auto three = [&](const void* Bh, const void* Bl, int rows, float* Ao) {
        gemm_ex(false, false, rows, o, n, Bh, n, d_Ch.p, o, Ao, o, 1.f, 0.f, bf);
        gemm_ex(false, false, rows, o, n, Bh, n, d_Cl.p, o, Ao, o, 1.f, 1.f, bf);
        gemm_ex(false, false, rows, o, n, Bl, n, d_Ch.p, o, Ao, o, 1.f, 1.f, bf);
      };
      if (!packed)
        three(d_Bh.p, d_Bl.p, rank * n, A);
      else
        for_unpacked_rows<__hip_bfloat16>([&](int q0, int qn, const __hip_bfloat16* Wh, const __hip_bfloat16* Wl) {
          three(Wh, Wl, qn * n, A + (size_t)q0 * n * o);
        });
    }
    if (!direct)
      hipLaunchKernelGGL(k_permute_qmi_f32, dim3((asz + 255) / 256), dim3(256), 0, stream, A, At, rank, n, o);
    (void)K;
    const int nb = nsplit(), qb = (rank + nb - 1) / nb, nbe = (rank + qb - 1) / qb, full = rank / qb;
    const int kc = qb * o, ld = rank * o;
    d_parts.alloc((size_t)nbe * nn * 4);
    float* parts = d_parts.as<float>();
    // partial[b] (+)= X_b Y_b^T over aux block b, X, Y the row-major n x (rank*o)
    // operands; column-major: partial^T = Y_b^T-ish, same as gemm_ex above.
    auto split_gemm = [&](const void* X, const void* Y, float beta, rocblas_datatype in, size_t esz) {
      const float one = 1.f;
      if (full > 0)
        RB_CHECK(rocblas_gemm_strided_batched_ex(bl->h, rocblas_operation_transpose, rocblas_operation_none, n, n, kc,
                                                 &one, Y, in, ld, kc, X, in, ld, kc, &beta, parts,
                                                 rocblas_datatype_f32_r, n, (rocblas_stride)nn, parts,
                                                 rocblas_datatype_f32_r, n, (rocblas_stride)nn, full,
                                                 rocblas_datatype_f32_r, rocblas_gemm_algo_standard, 0, 0));
      if (full < nbe) {
        const int kt = (rank - full * qb) * o;
        const char* Xt = (const char*)X + (size_t)full * kc * esz;
        const char* Yt = (const char*)Y + (size_t)full * kc * esz;
        RB_CHECK(rocblas_gemm_ex(bl->h, rocblas_operation_transpose, rocblas_operation_none, n, n, kt, &one, Yt, in,
                                 ld, Xt, in, ld, &beta, parts + (size_t)full * nn, rocblas_datatype_f32_r, n,
                                 parts + (size_t)full * nn, rocblas_datatype_f32_r, n, rocblas_datatype_f32_r,
                                 rocblas_gemm_algo_standard, 0, 0));
      }
    };
    if (prec == 1) {
      split_gemm(At, At, 0.f, rocblas_datatype_f32_r, 4);
    } else {
      if (!direct)
        hipLaunchKernelGGL(k_split_bf16<float>, g, b, 0, stream, At, d_Ah.as<__hip_bfloat16>(),
                           d_Al.as<__hip_bfloat16>(), asz);
      const auto bf = rocblas_datatype_bf16_r;
      split_gemm(d_Ah.p, d_Ah.p, 0.f, bf, 2);
      split_gemm(d_Ah.p, d_Al.p, 1.f, bf, 2);
      split_gemm(d_Al.p, d_Ah.p, 1.f, bf, 2);
    }
    hipLaunchKernelGGL(k_sum_parts<float>, dim3(1024), dim3(256), 0, stream, parts, nbe, nn, dK);
    HIP_CHECK(hipGetLastError());
  }
};

// DF-RHF (src/hf_variants/ri_df_hf.rs:160-263): core-Hamiltonian guess, CDIIS
// from the second iteration, convergence on |dE|, the projected commutator
// residual and max|dP|, then one final Fock build with the new orbitals.
// make_engine builds the factors (DF: metric, 3-index and B; CD-HF: the
// Cholesky vectors) while the one-electron matrices are formed on a second
// host thread.
static int df_rhf(const hfg_basis_view* basis, const std::function<DfEngine*()>& make_engine, const double* zxyz,
                  int natom, int charge, const hfg_dfhf_options* opt, hfg_dfhf_result* res, double* p_out,
                  double* f_out, double* j_out, double* k_out) {
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  std::memset(res, 0, sizeof *res);
  if (natom <= 0 || opt->max_iter <= 0 || !(opt->energy_tol > 0) || !(opt->residual_tol > 0) ||
      !(opt->density_tol > 0))
    throw HfgError{HFG_INVALID_INPUT, "DF-HF requires finite geometry and positive convergence controls"};
  if (opt->k_precision < 0 || opt->k_precision > 2 || !(opt->switch_residual > 0))
    throw HfgError{HFG_INVALID_INPUT, "DF-HF k_precision must be 0, 1 or 2 and switch_residual positive"};
  std::vector<std::array<double, 4>> atoms(natom);
  double el = -charge;
  for (int a = 0; a < natom; ++a) {
    for (int k = 0; k < 4; ++k) atoms[a][k] = zxyz[4 * a + k];
    el += atoms[a][0];
  }
  if (!(el > 0) || std::fabs(el - std::round(el)) > 1e-8 || ((long)std::lround(el)) % 2)
    throw HfgError{HFG_INVALID_INPUT, "DF-RHF requires a positive even electron count"};
  const std::vector<Shell> sh1 = group_shells(*basis);
  std::unique_ptr<DfEngine> df_p(make_engine());
  DfEngine& df = *df_p;
  const int n = df.n, no = (int)std::lround(el) / 2;
  if (n != basis->n_functions) throw HfgError{HFG_INVALID_INPUT, "DF tensor and basis differ in size"};
  hipStream_t st = df.stream;
  Blas& bl = *df.bl;
  const size_t nn = (size_t)n * n;
  double enuc = 0;
  for (int i = 0; i < natom; ++i)
    for (int j = i + 1; j < natom; ++j) {
      const double dx = atoms[i][1] - atoms[j][1], dy = atoms[i][2] - atoms[j][2], dz = atoms[i][3] - atoms[j][3];
      enuc += atoms[i][0] * atoms[j][0] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  Dev dS, dH, dX, dTmp, dTmp2, dF, dP, dPn, dJ, dK, dC, dCo, dCn, dW, dErr, dFu;
  for (Dev* d : {&dS, &dH, &dTmp, &dTmp2, &dF, &dP, &dPn, &dJ, &dK, &dErr, &dFu}) d->alloc(nn);
  one_electron_device(sh1, n, atoms, dS.p, nullptr, nullptr, dH.p, st);  // S, H = T + V
  // canonical orthogonaliser of S (cutoff 1e-8, Cartesian)
  Eigh eig;
  int m;
  {
    Dev dw;
    dw.alloc(n);
    HIP_CHECK(hipMemcpyAsync(dTmp.p, dS.p, nn * 8, hipMemcpyDeviceToDevice, st));
    eig.run(bl, dTmp.p, dw.p, n);
    std::vector<double> w(n), U(nn);
    download(w.data(), dw.p, n, st);
    download(U.data(), dTmp.p, nn, st);
    if (w[0] < -1e-8) throw HfgError{HFG_INTEGRAL, "overlap has negative eigenvalue"};
    int first = 0;
    while (first < n && w[first] < 1e-8) ++first;
    m = n - first;
    if (no > m) throw HfgError{HFG_INVALID_INPUT, "factorized RHF occupied count exceeds overlap rank"};
    std::vector<double> X((size_t)n * m);
    for (int c = 0; c < m; ++c)
      for (int r = 0; r < n; ++r) X[(size_t)r * m + c] = U[(size_t)(first + c) * n + r] / std::sqrt(w[first + c]);
    dX.alloc((size_t)n * m);
    upload(dX.p, X.data(), (size_t)n * m, st);
  }
  dC.alloc((size_t)n * m);
  dCo.alloc((size_t)n * no);
  dCn.alloc((size_t)n * m);
  dW.alloc(m);
  const int tb = 256;
  auto grid = [&](size_t c) { return dim3((unsigned)((c + tb - 1) / tb)); };
  // C = X eigvecs(X^T F X); C_occ; P = 2 C_occ C_occ^T
  auto diag = [&](const double* F, double* C, double* P) {
    bl.gemm(false, false, n, m, n, 1.0, F, n, dX.p, m, 0.0, dTmp.p, m);
    bl.gemm(true, false, m, m, n, 1.0, dX.p, m, dTmp.p, m, 0.0, dTmp2.p, m);
    eig.run(bl, dTmp2.p, dW.p, m);
    bl.gemm(false, true, n, m, m, 1.0, dX.p, m, dTmp2.p, m, 0.0, C, m);
    bl.gemm(false, true, n, n, no, 2.0, C, m, C, m, 0.0, P, n);
  };
  auto occ_copy = [&](const double* C) {
    // C_occ (n x no): the first no columns of C (n x m), a strided device copy
    HIP_CHECK(hipMemcpy2DAsync(dCo.p, (size_t)no * 8, C, (size_t)m * 8, (size_t)no * 8, n, hipMemcpyDeviceToDevice, st));
  };
  double t_fock = 0;
  int builds = 0;
  int kprec = opt->k_precision;  // nonzero until the switch to FP64 builds
  int fp64_builds = 0;
  auto fock = [&](double* F) {
    const auto tf = clk::now();
    if (kprec != 0) {
      df.jk(dCo.p, no, dJ.p, nullptr);  // J in FP64
      df.k_reduced(dCo.p, no, dK.p, kprec);
    } else {
      df.jk(dCo.p, no, dJ.p, dK.p);
      ++fp64_builds;
    }
    hipLaunchKernelGGL(k_axpby_fock, grid(nn), dim3(tb), 0, st, dH.p, dJ.p, dK.p, F, nn);
    HIP_CHECK(hipStreamSynchronize(st));
    t_fock += std::chrono::duration<double>(clk::now() - tf).count();
    ++builds;
  };
  auto energy = [&](const double* P, const double* F) { return 0.5 * (bl.dot(nn, P, dH.p) + bl.dot(nn, P, F)); };
  diag(dH.p, dC.p, dP.p);
  occ_copy(dC.p);
  // CDIIS histories on the device (8 slots of n x n Focks and m x m
  // errors); the error inner products are kept per slot pair, so each
  // iteration adds at most 8 dot products; the (k+1)^2 system is solved on
  // the host.
  const int MAXS = 8;
  Dev hF, hE, dCoef;
  hF.alloc(MAXS * nn);
  hE.alloc((size_t)MAXS * (size_t)m * m);
  dCoef.alloc(MAXS);
  hipLaunchKernelGGL(k_zero, grid(MAXS * nn), dim3(tb), 0, st, hF.p, MAXS * nn);
  std::vector<int> order;  // logical order (oldest first) -> slot
  std::map<std::pair<int, int>, double> edot;
  double last = INFINITY;
  const size_t mm2 = (size_t)m * m;
  for (int it = 1; it <= opt->max_iter; ++it) {
    fock(dF.p);
    const double ce = energy(dP.p, dF.p) + enuc;
    // projected commutator X^T (FPS - SPF) X
    bl.gemm(false, false, n, n, n, 1.0, dF.p, n, dP.p, n, 0.0, dTmp.p, n);
    bl.gemm(false, false, n, n, n, 1.0, dTmp.p, n, dS.p, n, 0.0, dTmp2.p, n);
    hipLaunchKernelGGL(k_antisym, grid(nn), dim3(tb), 0, st, dTmp2.p, dErr.p, n);
    bl.gemm(false, false, n, m, n, 1.0, dErr.p, n, dX.p, m, 0.0, dTmp.p, m);
    bl.gemm(true, false, m, m, n, 1.0, dX.p, m, dTmp.p, m, 0.0, dTmp2.p, m);
    const double resid = bl.amax(mm2, dTmp2.p);
    // this iteration's error, kept before diag() reuses dTmp2
    int slot;
    if ((int)order.size() >= MAXS) {
      slot = order.front();
      order.erase(order.begin());
    } else {
      slot = (int)order.size();
    }
    HIP_CHECK(hipMemcpyAsync(hE.p + (size_t)slot * mm2, dTmp2.p, mm2 * 8, hipMemcpyDeviceToDevice, st));
    diag(dF.p, dCn.p, dPn.p);
    hipLaunchKernelGGL(k_sub, grid(nn), dim3(tb), 0, st, dPn.p, dP.p, dTmp.p, nn);
    const double dch = bl.amax(nn, dTmp.p);
    if (!std::isfinite(ce) || !std::isfinite(resid)) throw HfgError{HFG_INTEGRAL, "DF-HF energy/residual became nonfinite"};
    if (kprec == 0 && fp64_builds >= 2 && std::fabs(ce - last) <= opt->energy_tol && resid <= opt->residual_tol &&
        dch <= opt->density_tol) {
      occ_copy(dCn.p);
      fock(dF.p);
      res->total_energy = energy(dPn.p, dF.p) + enuc;
      res->iterations = it;
      res->converged = 1;
      res->n_aux = df.rank;
      res->fock_builds = builds;
      res->seconds_fock = t_fock;
      if (p_out) download(p_out, dPn.p, nn, st);
      if (f_out) download(f_out, dF.p, nn, st);
      if (j_out) download(j_out, dJ.p, nn, st);
      if (k_out) download(k_out, dK.p, nn, st);
      res->seconds_total = std::chrono::duration<double>(clk::now() - t0).count();
      return HFG_OK;
    }
    last = ce;
    if (kprec != 0 && resid < opt->switch_residual) {
      // from here on FP64 builds only; the reduced-precision Focks leave the
      // extrapolation history, this one included
      kprec = 0;
      order.clear();
      edot.clear();
      last = INFINITY;
      diag(dF.p, dC.p, dP.p);
      occ_copy(dC.p);
      continue;
    }
    HIP_CHECK(hipMemcpyAsync(hF.p + (size_t)slot * nn, dF.p, nn * 8, hipMemcpyDeviceToDevice, st));
    for (auto e = edot.begin(); e != edot.end();)
      if (e->first.first == slot || e->first.second == slot)
        e = edot.erase(e);
      else
        ++e;
    order.push_back(slot);
    for (int o : order)
      edot[{std::min(slot, o), std::max(slot, o)}] = bl.dot(mm2, hE.p + (size_t)slot * mm2, hE.p + (size_t)o * mm2);
    const double* fuse = dF.p;
    if (it >= 2 && order.size() >= 2) {
      const int k = (int)order.size();
      std::vector<double> B((size_t)(k + 1) * (k + 1), 0.0), rhs(k + 1, 0.0), c;
      for (int i = 0; i < k; ++i) {
        for (int j = 0; j < k; ++j)
          B[(size_t)i * (k + 1) + j] = edot[{std::min(order[i], order[j]), std::max(order[i], order[j])}];
        B[(size_t)i * (k + 1) + k] = B[(size_t)k * (k + 1) + i] = -1.0;
      }
      rhs[k] = -1.0;
      if (solve_small(B, rhs, k + 1, c)) {
        std::vector<double> cs(MAXS, 0.0);
        for (int i = 0; i < k; ++i) cs[order[i]] = c[i];
        upload(dCoef.p, cs.data(), MAXS, st);
        const double one = 1.0, zero = 0.0;
        RB_CHECK(rocblas_dgemv(bl.h, rocblas_operation_none, (rocblas_int)nn, MAXS, &one, hF.p, (rocblas_int)nn,
                               dCoef.p, 1, &zero, dFu.p, 1));
        fuse = dFu.p;
      }
    }
    diag(fuse, dC.p, dP.p);
    occ_copy(dC.p);
  }
  res->iterations = opt->max_iter;
  res->seconds_total = std::chrono::duration<double>(clk::now() - t0).count();
  throw HfgError{HFG_NOT_CONVERGED, "factorized RHF did not converge"};
}

}  // namespace hfg

using namespace hfg;

static int dferr(int code, const std::string& m) { return set_error(code, m); }

struct hfg_df {
  std::unique_ptr<DfEngine> e;
};

extern "C" {

int hfg_df_build(const hfg_basis_view* basis, const hfg_basis_view* aux, double metric_cutoff, hfg_df** out) {
  if (!basis || !aux || !out) return dferr(HFG_INVALID_INPUT, "null argument");
  try {
    auto* h = new hfg_df;
    try {
      h->e.reset(new DfEngine(*basis, *aux, metric_cutoff));
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
    return HFG_OK;
  } catch (HfgError& e) {
    return dferr(e.code, e.msg);
  } catch (std::exception& e) {
    return dferr(HFG_DEVICE, e.what());
  }
}

void hfg_df_destroy(hfg_df* df) { delete df; }

int hfg_df_info(const hfg_df* df, int64_t* n, int64_t* naux, int64_t* rank) {
  if (!df) return dferr(HFG_INVALID_INPUT, "null handle");
  if (n) *n = df->e->n;
  if (naux) *naux = df->e->naux;
  if (rank) *rank = df->e->rank;
  return HFG_OK;
}

int hfg_df_jk(hfg_df* df, const double* cocc, int nocc, double* j, double* k) {
  if (!df || !cocc || !j || !k || nocc <= 0) return dferr(HFG_INVALID_INPUT, "invalid occupied coefficients for DF J/K");
  try {
    DfEngine& e = *df->e;
    const size_t nn = (size_t)e.n * e.n;
    for (size_t x = 0; x < (size_t)e.n * nocc; ++x)
      if (!std::isfinite(cocc[x])) return dferr(HFG_INVALID_INPUT, "invalid occupied coefficients for DF J/K");
    Dev dc, dj, dk;
    dc.alloc((size_t)e.n * nocc);
    dj.alloc(nn);
    dk.alloc(nn);
    upload(dc.p, cocc, (size_t)e.n * nocc, e.stream);
    e.jk(dc.p, nocc, dj.p, dk.p);
    download(j, dj.p, nn, e.stream);
    download(k, dk.p, nn, e.stream);
    return HFG_OK;
  } catch (HfgError& ex) {
    return dferr(ex.code, ex.msg);
  }
}

int hfg_df_jk_ex(hfg_df* df, const double* cocc, int nocc, double* j, double* k, int precision,
                 double* device_ms) {
  if (!df || !cocc || !j || !k || nocc <= 0) return dferr(HFG_INVALID_INPUT, "invalid occupied coefficients for DF J/K");
  if (precision < 0 || precision > 2) return dferr(HFG_INVALID_INPUT, "precision must be 0 (FP64), 1 (FP32) or 2 (BF16x3)");
  try {
    DfEngine& e = *df->e;
    const size_t nn = (size_t)e.n * e.n;
    for (size_t x = 0; x < (size_t)e.n * nocc; ++x)
      if (!std::isfinite(cocc[x])) return dferr(HFG_INVALID_INPUT, "invalid occupied coefficients for DF J/K");
    Dev dc, dj, dk;
    dc.alloc((size_t)e.n * nocc);
    dj.alloc(nn);
    dk.alloc(nn);
    upload(dc.p, cocc, (size_t)e.n * nocc, e.stream);
    if (precision != 0) {  // first use: make the reduced copies of B outside the timed region
      e.k_reduced(dc.p, nocc, dk.p, precision);
      HIP_CHECK(hipStreamSynchronize(e.stream));
    }
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, e.stream));
    if (precision == 0) {
      e.jk(dc.p, nocc, dj.p, dk.p);
    } else {
      e.jk(dc.p, nocc, dj.p, nullptr);  // J in FP64
      e.k_reduced(dc.p, nocc, dk.p, precision);
    }
    HIP_CHECK(hipEventRecord(e1, e.stream));
    HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (device_ms) *device_ms = ms;
    download(j, dj.p, nn, e.stream);
    download(k, dk.p, nn, e.stream);
    return HFG_OK;
  } catch (HfgError& ex) {
    return dferr(ex.code, ex.msg);
  }
}

int hfg_df_rhf(const hfg_basis_view* basis, const hfg_basis_view* aux, const double* zxyz, int natom, int charge,
               const hfg_dfhf_options* opt, hfg_dfhf_result* res, double* p_out, double* f_out, double* j_out,
               double* k_out) {
  if (!basis || !aux || !zxyz || !opt || !res) return dferr(HFG_INVALID_INPUT, "null argument");
  try {
    if (!(opt->metric_cutoff > 0)) throw HfgError{HFG_INVALID_INPUT, "DF metric cutoff must be finite and positive"};
    return df_rhf(
        basis, [&] { return new DfEngine(*basis, *aux, opt->metric_cutoff); }, zxyz, natom, charge, opt, res, p_out,
        f_out, j_out, k_out);
  } catch (HfgError& e) {
    return dferr(e.code, e.msg);
  }
}

int hfg_cd_build(const hfg_basis_view* basis, double threshold, double budget_bytes, hfg_df** out, hfg_cd_info* info,
                 int32_t* pivots, int32_t max_pivots) {
  if (!basis || !out) return dferr(HFG_INVALID_INPUT, "null argument");
  try {
    CholeskyInfo ci;
    auto* d = new hfg_df;
    try {
      d->e.reset(new DfEngine(*basis, threshold, budget_bytes, ci));
    } catch (...) {
      delete d;
      throw;
    }
    if (info) {
      info->rank = ci.rank;
      info->max_residual = ci.max_residual;
    }
    if (pivots)
      for (int k = 0; k < ci.rank && k < max_pivots; ++k) pivots[k] = ci.pivots[k];
    *out = d;
    return HFG_OK;
  } catch (HfgError& e) {
    return dferr(e.code, e.msg);
  }
}

int hfg_cd_rhf(const hfg_basis_view* basis, const double* zxyz, int natom, int charge, double delta,
               double budget_bytes, const hfg_dfhf_options* opt, hfg_dfhf_result* res, hfg_cd_info* info,
               double* p_out, double* f_out) {
  if (!basis || !zxyz || !opt || !res) return dferr(HFG_INVALID_INPUT, "null argument");
  try {
    CholeskyInfo ci;
    const int st = df_rhf(
        basis, [&] { return new DfEngine(*basis, delta, budget_bytes, ci); }, zxyz, natom, charge, opt, res, p_out,
        f_out, nullptr, nullptr);
    if (info) {
      info->rank = ci.rank;
      info->max_residual = ci.max_residual;
    }
    return st;
  } catch (HfgError& e) {
    return dferr(e.code, e.msg);
  }
}

}  // extern "C"
