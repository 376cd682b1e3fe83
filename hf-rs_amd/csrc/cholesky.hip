// cholesky.hip -- pivoted Cholesky factors of the AO two-electron supermatrix
// on the device: pivoted_cholesky_eri (src/cholesky.rs:44-192).
//
// (pq|rs) ~ sum_K L_K(pq) L_K(rs).  The diagonal (pq|pq) comes from the
// (AB|AB) shell quartets; each step takes the largest residual diagonal as
// the pivot (values within 1e-12 relative of it are ties, resolved by the
// smallest packed pair index), the pivot's column (pq|rs) from complete shell
// quartets of its shell pair -- every column of that shell pair at once, kept
// while the following pivots share the shell pair -- and appends the factor
// row L(pq) = ((pq|rs) - sum_K L_K(pq) L_K(rs)) / sqrt(D_rs), L(rs) =
// sqrt(D_rs); the residual diagonal drops by L^2.  It stops when the largest
// residual is at most the threshold or every pair is a pivot.
//
// Device side: the shell-quartet blocks come from the block kernels; a
// scatter kernel writes them, scaled by the component factors, into the
// columns of the pivot shell pair; the factor row is one GEMV over the rows
// so far (L as an np x K column-major matrix, the pivot's entries read with
// stride np); one workgroup updates the residual diagonal, checks it and
// picks the next pivot, so each step returns one (pivot, maximum, flag)
// triple to the host.  The factors end as the unpacked [rank][n][n] tensor
// the density-fitting J/K kernels read (src/df.rs:121 keeps them packed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "cholesky.h"
#include "la_device.h"

namespace hfg {

namespace {

__device__ __forceinline__ size_t cpair(size_t i, size_t j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; }

// One shell-quartet block (kernel orientation) and where its elements go:
// the row pair (AO offsets ra, rb, component counts nra, nrb) and the pivot
// pair (pa, pb, npa, npb) are the block's bra and ket (pivot_ket = 1) or ket
// and bra; pswap = 1 when the pivot pair's kernel orientation is the reverse
// of its hf-rs orientation (r, s).
struct ColScatter {
  long long off;
  int ra, rb, nra, nrb;
  int pa, pb, npa, npb;
  int pivot_ket, pswap, same_row, pad;
};

// Diagonal of one (AB|AB) block: (pq|pq) with its component factors, p >= q
// for a pair of one shell with itself.
struct DiagScatter {
  long long off;
  int a, b, na, nb;
};

__global__ void k_chol_diag(const DiagScatter* __restrict__ ds, int count, const double* __restrict__ blk,
                            const double* __restrict__ scale, double* __restrict__ diag) {
  const int t = blockIdx.x;
  if (t >= count) return;
  const DiagScatter d = ds[t];
  const int nab = d.na * d.nb;
  for (int ij = threadIdx.x; ij < nab; ij += blockDim.x) {
    const int i = ij / d.nb, j = ij % d.nb;
    const int p = d.a + i, q = d.b + j;
    if (d.a == d.b && p < q) continue;
    const double f = scale[p] * scale[q];
    diag[cpair(p, q)] = blk[d.off + (size_t)ij * nab + ij] * f * f;
  }
}

// Columns of the pivot shell pair (r, s): cols[(k * ns + l) * np + pq] =
// (pq|rs) for AO r_k, s_l and row pair pq (rows of one shell with itself:
// p >= q only, as the reference fills them).
__global__ void k_chol_columns(const ColScatter* __restrict__ cs, int count, const double* __restrict__ blk,
                               const double* __restrict__ scale, double* __restrict__ cols, size_t np, int ns) {
  const int t = blockIdx.x;
  if (t >= count) return;
  const ColScatter c = cs[t];
  const int nrow = c.nra * c.nrb, npiv = c.npa * c.npb;
  for (int e = threadIdx.x; e < nrow * npiv; e += blockDim.x) {
    const int ri = e / npiv, pi = e % npiv;
    const int i = ri / c.nrb, j = ri % c.nrb, x = pi / c.npb, y = pi % c.npb;
    const int p = c.ra + i, q = c.rb + j;
    if (c.same_row && p < q) continue;
    const size_t idx = c.pivot_ket ? (size_t)ri * npiv + pi : (size_t)pi * nrow + ri;
    const double v = blk[c.off + idx] * scale[p] * scale[q] * scale[c.pa + x] * scale[c.pb + y];
    const int k = c.pswap ? y : x, l = c.pswap ? x : y;  // hf-rs orientation of the pivot pair
    cols[(size_t)(k * ns + l) * np + cpair(p, q)] = v;
  }
}

__global__ void k_chol_first_row(const double* __restrict__ col, double inv_root, double* __restrict__ row, size_t np) {
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < np; k += (size_t)gridDim.x * blockDim.x)
    row[k] = col[k] * inv_root;
}

// Pivot choice over the residual diagonal (one workgroup of 1024): the
// largest value M, then the smallest index with d >= M - 1e-12 |M|.
// With row given, first sets row[piv] = root and lowers the diagonal by
// row^2 (flag 1 when an entry falls below -1e-10 scale; entries are clamped
// at 0).  out: [0] pivot, [1] its value, [2] flag (doubles).
__global__ __launch_bounds__(1024) void k_chol_step(double* __restrict__ diag, size_t np, double* __restrict__ row,
                                                    long long piv, double root, double scale, double* __restrict__ out) {
  __shared__ double smax[1024];
  __shared__ unsigned long long smin[1024];
  __shared__ int sflag;
  const int t = threadIdx.x;
  if (t == 0) sflag = 0;
  __syncthreads();
  if (row && t == 0) row[piv] = root;
  __syncthreads();
  double m = -INFINITY;
  int bad = 0;
  for (size_t k = t; k < np; k += blockDim.x) {
    double d = diag[k];
    if (row) {
      const double l = row[k];
      d -= l * l;
      if (d < -1e-10 * scale) bad = 1;
      d = fmax(d, 0.0);
      diag[k] = d;
    }
    m = fmax(m, d);
  }
  if (bad) atomicOr(&sflag, 1);
  smax[t] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) smax[t] = fmax(smax[t], smax[t + s]);
    __syncthreads();
  }
  const double M = smax[0];
  const double cut = M - 1e-12 * fabs(M);
  unsigned long long first = ~0ull;
  for (size_t k = t; k < np; k += blockDim.x)
    if (diag[k] >= cut) {
      first = k;
      break;
    }
  smin[t] = first;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) smin[t] = smin[t] < smin[t + s] ? smin[t] : smin[t + s];
    __syncthreads();
  }
  if (t == 0) {
    const unsigned long long p = smin[0] == ~0ull ? 0 : smin[0];
    out[0] = (double)p;
    out[1] = diag[p];
    out[2] = sflag ? 1.0 : 0.0;
  }
}

// B[K][p][q] = B[K][q][p] = L[K][cpair(p, q)].
__global__ void k_chol_unpack(const double* __restrict__ L, int rank, int n, double* __restrict__ B) {
  const size_t nn = (size_t)n * n, np = (size_t)n * (n + 1) / 2, total = (size_t)rank * nn;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t K = i / nn, pq = i % nn;
    B[i] = L[K * np + cpair(pq / n, pq % n)];
  }
}

}  // namespace

void cholesky_factors(const std::vector<Shell>& orb, int n, hipStream_t st, Blas& bl, double threshold, double budget,
                      DevBuf& d_B, int& rank, CholeskyInfo& info) {
  if (n <= 0 || !std::isfinite(threshold) || threshold < 0)
    throw HfgError{HFG_INVALID_INPUT, "Cholesky requires a basis and a finite, nonnegative threshold"};
  if (!(budget > 0)) throw HfgError{HFG_INVALID_INPUT, "Cholesky factor budget must be positive"};
  const size_t np = (size_t)n * (n + 1) / 2;
  const int ns = (int)orb.size();
  std::vector<PairInfo> pairs;
  PrimPairHost pp;
  build_pairs(orb, pairs, pp);
  const size_t nsp = pairs.size();
  std::vector<int> pair_at(nsp);  // hf-rs shell-pair index -> position in pairs
  for (size_t i = 0; i < nsp; ++i) pair_at[pairs[i].hf_index] = (int)i;
  std::vector<int> shell_of(n);
  for (int s = 0; s < ns; ++s)
    for (size_t c = 0; c < orb[s].scale.size(); ++c) shell_of[orb[s].ao + c] = s;
  auto ncomp = [&](int s) { return (int)orb[s].scale.size(); };
  // device pair tables, Boys table, component factors
  const size_t npp = pp.size();
  DevBuf d_pp, d_pairs, d_boys, d_scale;
  d_pp.alloc(9 * std::max<size_t>(1, npp) * sizeof(double));
  const std::vector<double>* pcols[9] = {&pp.p, &pp.Px, &pp.Py, &pp.Pz, &pp.PAx, &pp.PAy, &pp.PAz, &pp.K, &pp.H};
  for (int c = 0; c < 9; ++c)
    HIP_CHECK(hipMemcpyAsync(d_pp.as<double>() + c * npp, pcols[c]->data(), npp * 8, hipMemcpyHostToDevice, st));
  std::vector<PairDev> pd(nsp);
  for (size_t i = 0; i < nsp; ++i) {
    std::memset(&pd[i], 0, sizeof(PairDev));
    pd[i].p0 = pairs[i].p0;
    pd[i].p1 = pairs[i].p1;
    for (int k = 0; k < 3; ++k) pd[i].AB[k] = pairs[i].AB[k];
  }
  d_pairs.alloc(nsp * sizeof(PairDev));
  HIP_CHECK(hipMemcpyAsync(d_pairs.p, pd.data(), nsp * sizeof(PairDev), hipMemcpyHostToDevice, st));
  const std::vector<double> tab = boys_table();
  d_boys.alloc(tab.size() * 8);
  HIP_CHECK(hipMemcpyAsync(d_boys.p, tab.data(), tab.size() * 8, hipMemcpyHostToDevice, st));
  std::vector<double> so(n);
  for (auto& s : orb)
    for (size_t c = 0; c < s.scale.size(); ++c) so[s.ao + c] = s.scale[c];
  d_scale.alloc(n * 8);
  HIP_CHECK(hipMemcpyAsync(d_scale.p, so.data(), n * 8, hipMemcpyHostToDevice, st));
  const double* ppb = d_pp.as<double>();
  const PrimPairs view{ppb, ppb + npp, ppb + 2 * npp, ppb + 3 * npp, ppb + 4 * npp,
                       ppb + 5 * npp, ppb + 6 * npp, ppb + 7 * npp, ppb + 8 * npp};
  auto nc = [](int L) { return (L + 1) * (L + 2) / 2; };

  // Block tasks grouped by class (the caller sorts), launched into one buffer.
  DevBuf d_tasks, d_blk, d_sc;
  size_t cap_tasks = 0, cap_blk = 0, cap_sc = 0;
  auto run = [&](const std::vector<BlockTask>& tasks, const std::vector<int>& cls, long long total) {
    if (tasks.size() > cap_tasks) {
      d_tasks.alloc(tasks.size() * sizeof(BlockTask));
      cap_tasks = tasks.size();
    }
    if ((size_t)total > cap_blk) {
      d_blk.alloc((size_t)total * 8);
      cap_blk = total;
    }
    HIP_CHECK(hipMemcpyAsync(d_tasks.p, tasks.data(), tasks.size() * sizeof(BlockTask), hipMemcpyHostToDevice, st));
    size_t a = 0;
    while (a < tasks.size()) {
      size_t b = a;
      while (b < tasks.size() && cls[b] == cls[a]) ++b;
      BlockArgs args{d_tasks.as<BlockTask>() + a, (int)(b - a), d_pairs.as<PairDev>(), view, d_boys.as<double>(),
                     d_blk.as<double>()};
      launch_block_class(cls[a], args, st);
      HIP_CHECK(hipGetLastError());
      a = b;
    }
  };
  auto upload_scatter = [&](const void* h, size_t bytes) {
    if (bytes > cap_sc) {
      d_sc.alloc(bytes);
      cap_sc = bytes;
    }
    HIP_CHECK(hipMemcpyAsync(d_sc.p, h, bytes, hipMemcpyHostToDevice, st));
  };

  // ---- diagonal (pq|pq)
  Dev d_diag;
  d_diag.alloc(np);
  {
    std::vector<int> order(nsp);
    std::iota(order.begin(), order.end(), 0);
    auto pcl = [&](int i) { return pair_class(pairs[i].la, pairs[i].lb); };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return pcl(x) < pcl(y); });
    std::vector<BlockTask> tasks;
    std::vector<int> cls;
    std::vector<DiagScatter> ds;
    long long off = 0;
    for (int i : order) {
      const PairInfo& p = pairs[i];
      const int nab = nc(p.la) * nc(p.lb);
      tasks.push_back({i, i, off});
      cls.push_back(quartet_class(pcl(i), pcl(i)));
      ds.push_back({off, p.ao_a, p.ao_b, nc(p.la), nc(p.lb)});
      off += (long long)nab * nab;
    }
    run(tasks, cls, off);
    upload_scatter(ds.data(), ds.size() * sizeof(DiagScatter));
    hipLaunchKernelGGL(k_chol_diag, dim3((unsigned)ds.size()), dim3(64), 0, st, (const DiagScatter*)d_sc.p,
                       (int)ds.size(), d_blk.as<double>(), d_scale.as<double>(), d_diag.p);
    HIP_CHECK(hipGetLastError());
  }
  std::vector<double> diag(np);
  download(diag.data(), d_diag.p, np, st);
  double scale = 1.0;
  for (double v : diag) scale = std::max(scale, v);
  for (double v : diag)
    if (v < -1e-12 * scale)
      throw HfgError{HFG_INTEGRAL, "pivoted Cholesky ERI decomposition: negative ERI diagonal; the Coulomb metric is "
                                   "not positive semidefinite"};
  // Schwarz bound per shell pair (hf-rs pair index), from the diagonal
  std::vector<double> qsp(nsp, 0.0);
  for (int p = 0; p < n; ++p)
    for (int q = 0; q <= p; ++q) {
      const int a = shell_of[p], b = shell_of[q];
      const size_t sp = (size_t)a * (a + 1) / 2 + b;
      qsp[sp] = std::max(qsp[sp], std::sqrt(std::max(diag[(size_t)p * (p + 1) / 2 + q], 0.0)));
    }

  // ---- factor rows: L[K][np] on the device, as many rows as the budget allows
  const double rows_allowed = std::floor(budget / (8.0 * (double)np)) - 1.0;  // K + 2 <= budget / (8 np)
  const size_t cap = (size_t)std::max(0.0, std::min((double)np, rows_allowed));
  Dev d_L, d_out, d_cols;
  d_L.alloc(std::max<size_t>(1, cap) * np);
  d_out.alloc(4);
  size_t cols_cap = 0;
  int cached = -1;  // hf-rs shell pair whose columns d_cols holds
  int cached_ns = 1;
  std::vector<double> h_out(3);
  hipLaunchKernelGGL(k_chol_step, dim3(1), dim3(1024), 0, st, d_diag.p, np, (double*)nullptr, 0ll, 0.0, scale, d_out.p);
  HIP_CHECK(hipGetLastError());
  download(h_out.data(), d_out.p, 3, st);
  info.pivots.clear();
  size_t K = 0;
  for (;;) {
    const size_t piv = (size_t)h_out[0];
    const double maximum = h_out[1];
    if (maximum <= threshold || K == np) {
      info.max_residual = std::max(maximum, 0.0);
      break;
    }
    if (8.0 * (double)np * (double)(K + 2) > budget || K >= cap)
      throw HfgError{HFG_INVALID_INPUT, "pivoted Cholesky ERI decomposition: factor storage exceeds the memory budget"};
    // pivot pair (p >= q) and its shell pair (r >= s)
    size_t p = (size_t)((std::sqrt(8.0 * (double)piv + 1.0) - 1.0) / 2.0);
    while (p * (p + 1) / 2 > piv) --p;
    while ((p + 1) * (p + 2) / 2 <= piv) ++p;
    const size_t q = piv - p * (p + 1) / 2;
    const int r = shell_of[p], s = shell_of[q];
    const int sp = r * (r + 1) / 2 + s;
    if (sp != cached) {
      // every column of shell pair (r, s): blocks of each shell pair x with
      // Q_x Q_sp > 1e-15 against it
      const int nr = ncomp(r), nss = ncomp(s);
      const size_t need = (size_t)nr * nss * np;
      if (need > cols_cap) {
        d_cols.alloc(need);
        cols_cap = need;
      }
      hipLaunchKernelGGL(k_zero, dim3((unsigned)((need + 255) / 256)), dim3(256), 0, st, d_cols.p, need);
      const PairInfo& P = pairs[pair_at[sp]];
      const int pcp = pair_class(P.la, P.lb);
      struct Item {
        int cls;
        BlockTask t;
        ColScatter c;
      };
      std::vector<Item> items;
      for (size_t xi = 0; xi < nsp; ++xi) {
        const PairInfo& X = pairs[xi];
        if (!(qsp[X.hf_index] * qsp[sp] > 1e-15)) continue;
        const int pcx = pair_class(X.la, X.lb);
        Item it;
        std::memset(&it, 0, sizeof it);
        it.c.ra = X.ao_a;
        it.c.rb = X.ao_b;
        it.c.nra = nc(X.la);
        it.c.nrb = nc(X.lb);
        it.c.pa = P.ao_a;
        it.c.pb = P.ao_b;
        it.c.npa = nc(P.la);
        it.c.npb = nc(P.lb);
        it.c.pswap = P.swapped;
        it.c.same_row = X.sh_a == X.sh_b;
        if (pcp > pcx) {  // the pivot pair is the bra (higher pair class first)
          it.t = BlockTask{pair_at[sp], (int)xi, 0};
          it.cls = quartet_class(pcp, pcx);
          it.c.pivot_ket = 0;
        } else {
          it.t = BlockTask{(int)xi, pair_at[sp], 0};
          it.cls = quartet_class(pcx, pcp);
          it.c.pivot_ket = 1;
        }
        items.push_back(it);
      }
      std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) { return a.cls < b.cls; });
      std::vector<BlockTask> tasks(items.size());
      std::vector<int> cls(items.size());
      std::vector<ColScatter> sc(items.size());
      long long off = 0;
      for (size_t i = 0; i < items.size(); ++i) {
        tasks[i] = items[i].t;
        tasks[i].off = off;
        cls[i] = items[i].cls;
        sc[i] = items[i].c;
        sc[i].off = off;
        off += (long long)sc[i].nra * sc[i].nrb * sc[i].npa * sc[i].npb;
      }
      if (!items.empty()) {
        run(tasks, cls, off);
        upload_scatter(sc.data(), sc.size() * sizeof(ColScatter));
        hipLaunchKernelGGL(k_chol_columns, dim3((unsigned)sc.size()), dim3(256), 0, st, (const ColScatter*)d_sc.p,
                           (int)sc.size(), d_blk.as<double>(), d_scale.as<double>(), d_cols.p, np, nss);
        HIP_CHECK(hipGetLastError());
      }
      cached = sp;
      cached_ns = nss;
    }
    const int k = (int)(p - orb[r].ao), l = (int)(q - orb[s].ao);
    const double* col = d_cols.p + (size_t)(k * cached_ns + l) * np;
    const double root = std::sqrt(maximum);
    double* row = d_L.p + K * np;
    hipLaunchKernelGGL(k_chol_first_row, dim3((unsigned)std::min<size_t>(2048, (np + 255) / 256)), dim3(256), 0, st,
                       col, 1.0 / root, row, np);
    if (K > 0) {
      // row -= (1/root) L^T L[:, piv]: L as an np x K column-major matrix
      const double alpha = -1.0 / root, beta = 1.0;
      RB_CHECK(rocblas_dgemv(bl.h, rocblas_operation_none, (rocblas_int)np, (rocblas_int)K, &alpha, d_L.p,
                             (rocblas_int)np, d_L.p + piv, (rocblas_int)np, &beta, row, 1));
    }
    hipLaunchKernelGGL(k_chol_step, dim3(1), dim3(1024), 0, st, d_diag.p, np, row, (long long)piv, root, scale,
                       d_out.p);
    HIP_CHECK(hipGetLastError());
    download(h_out.data(), d_out.p, 3, st);
    if (h_out[2] != 0.0)
      throw HfgError{HFG_INTEGRAL, "pivoted Cholesky ERI decomposition: nonpositive residual Coulomb diagonal"};
    info.pivots.push_back((int)piv);
    ++K;
  }
  rank = (int)K;
  if (rank == 0)
    throw HfgError{HFG_INTEGRAL, "pivoted Cholesky ERI decomposition: threshold exceeds every ERI diagonal; no "
                                 "Cholesky vectors were generated"};
  info.rank = rank;
  const size_t nn = (size_t)n * n;
  d_B.alloc((size_t)rank * nn * 8);
  hipLaunchKernelGGL(k_chol_unpack, dim3(4096), dim3(256), 0, st, d_L.p, rank, n, d_B.as<double>());
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
}

}  // namespace hfg
