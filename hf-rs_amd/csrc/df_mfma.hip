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
#include <cstdint>
#include <vector>

#include "engine.h"

namespace hfg {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int ST = 128;     // output tile
constexpr int KB = 32;      // contraction step staged in LDS
constexpr int PAD = KB + 1;
constexpr int SYRK_NT = 512;  // 8 waves, 4 x 2 over the tile
constexpr int WTI = ST / 4, WTJ = ST / 2;  // a wave's rows and columns of the tile
constexpr int NBI = WTI / 16, NBJ = WTJ / 16;  // 16 x 16 MFMA blocks per wave side

// The thread's share of the two ST x KB panels at contraction offset k0:
// row t / 4, columns 8 (t % 4) .. +7 of each (zero beyond n rows or kend).
struct Panel {
  double a[8], b[8];
};
__device__ __forceinline__ void load_panel(const double* __restrict__ X, int n, long long ldx, int i0, int j0,
                                           long long k0, long long kend, bool vec, Panel& p) {
  const int t = threadIdx.x, r = t >> 2, c0 = (t & 3) * 8;
  const int ra = i0 + r, rb = j0 + r;
  const long long k = k0 + c0;
  if (vec && k + 8 <= kend) {
    // 16-byte loads: rows start 16-byte aligned when ldx is even, k0 is a multiple of KB
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

__global__ __launch_bounds__(SYRK_NT) void k_syrk_f64(const double* __restrict__ X, int n, long long L,
                                                      long long ldx, const int2* __restrict__ tiles, int ntile,
                                                      long long kslice, double* __restrict__ part) {
  __shared__ double As[2][ST * PAD], Bs[2][ST * PAD];
  const int2 tile = tiles[blockIdx.x % ntile];
  const int s = blockIdx.x / ntile;
  const int i0 = tile.x * ST, j0 = tile.y * ST;
  const long long kbeg = (long long)s * kslice, kend = L < kbeg + kslice ? L : kbeg + kslice;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = w >> 1, wj = w & 1;
  const bool vec = (ldx & 1) == 0;
  d4 acc[NBI][NBJ];
#pragma unroll
  for (int a = 0; a < NBI; ++a)
#pragma unroll
    for (int b = 0; b < NBJ; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  const int r = t >> 2, c0 = (t & 3) * 8;
  auto store = [&](const Panel& p, int buf) {
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      As[buf][r * PAD + c0 + c] = p.a[c];
      Bs[buf][r * PAD + c0 + c] = p.b[c];
    }
  };
  // buffer cur holds the step's panels, the registers the next step's; a step
  // stores them into the other buffer (whose readers all passed the previous
  // barrier), loads the step after, multiplies, and meets the other waves at
  // one barrier
  Panel nxt;
  if (kbeg < kend) {
    load_panel(X, n, ldx, i0, j0, kbeg, kend, vec, nxt);
    store(nxt, 0);
    if (kbeg + KB < kend) load_panel(X, n, ldx, i0, j0, kbeg + KB, kend, vec, nxt);
  }
  __syncthreads();
  int cur = 0;
  for (long long k0 = kbeg; k0 < kend; k0 += KB, cur ^= 1) {
    if (k0 + KB < kend) {
      store(nxt, cur ^ 1);
      if (k0 + 2 * KB < kend) load_panel(X, n, ldx, i0, j0, k0 + 2 * KB, kend, vec, nxt);
    }
    const double* as = As[cur];
    const double* bs = Bs[cur];
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      const int kq = kk + (lane >> 4), rq = lane & 15;
      double av[NBI], bv[NBJ];
#pragma unroll
      for (int a = 0; a < NBI; ++a) av[a] = as[(WTI * wi + 16 * a + rq) * PAD + kq];
#pragma unroll
      for (int b = 0; b < NBJ; ++b) bv[b] = bs[(WTJ * wj + 16 * b + rq) * PAD + kq];
#pragma unroll
      for (int a = 0; a < NBI; ++a)
#pragma unroll
        for (int b = 0; b < NBJ; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  double* out = part + (size_t)s * n * n;
#pragma unroll
  for (int a = 0; a < NBI; ++a)
#pragma unroll
    for (int b = 0; b < NBJ; ++b)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int i = i0 + WTI * wi + 16 * a + (lane >> 4) + 4 * q;
        const int j = j0 + WTJ * wj + 16 * b + (lane & 15);
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

// ---- Work order of the packed A = B C kernels.  Each packed element of B_Q
// is read by two workgroups: the one of its row's 128-row tile and, as the
// transposed element, the one of its column's tile.  Two choices make the
// second read a cache hit:
// - XCD-aware mapping: workgroups are dealt round-robin to the 8 XCDs by
//   linear id, so linear id L = 8 j + x is given work item off(x) + j: each
//   XCD takes a contiguous run of items, the tiles of one auxiliary row run
//   together on one XCD and share its L2;
// - tournament order over the k-tiles (the circle method on M = T or T + 1
//   players, M odd, plus a fixed player F = M): in round r tile i multiplies
//   k-tile (2 r - i) mod M and the tile named there multiplies k-tile i, so
//   both reads of a 128 x 128 packed block fall in the same round; tile i
//   meets F in round i and takes its diagonal block then, a round against
//   the padding player (T even) is skipped.
// Measured on (H2O)32 / cc-pVTZ (ms a launch, neither / XCD map / tournament
// / both): FP64 111.8 / 111.3 / 119.6 / 119.2, split BF16 52.9 / 47.3 / 44.3
// / 43.3.  The split-BF16 kernel, bound by the traffic of B, takes both; the
// FP64 kernel, bound by the matrix cores, only the XCD map.
constexpr int BC_XCD_MAP = 1, BC_TOURNAMENT = 2;
struct WorkItem {
  int mt, it, q;
};
__device__ __forceinline__ WorkItem work_item(int gx, int gy, long long total, int flags) {
  const long long L = blockIdx.x, per = total / 8, rem = total % 8;
  const int x = (int)(L % 8);
  const long long w = (flags & BC_XCD_MAP) ? x * per + (x < rem ? x : rem) + L / 8 : L;
  WorkItem r;
  r.mt = (int)(w % gx);
  const long long rest = w / gx;
  r.it = (int)(rest % gy);
  r.q = (int)(rest / gy);
  return r;
}

template <int KS>
struct PanelOrder {
  static constexpr int TS = 128, PPT = TS / KS;  // tile width, panels a tile
  int mt, T, M, n, tour;
  __device__ PanelOrder(int mt_, int n_, int flags)
      : mt(mt_), T((n_ + TS - 1) / TS), M(T | 1), n(n_), tour(flags & BC_TOURNAMENT) {}
  __device__ int count() const { return tour ? M * PPT : (n + KS - 1) / KS; }
  // k0 of panel s (tournament: s = r PPT + j), or -1 where there is none
  __device__ int k0(int s) const {
    if (!tour) return s * KS;
    const int r = s / PPT, j = s - r * PPT;
    int t = r == mt ? mt : ((2 * r - mt) % M + M) % M;
    if (t >= T) return -1;
    const int k = t * TS + j * KS;
    return k < n ? k : -1;
  }
  __device__ int next(int s) const {
    for (++s; s < count(); ++s)
      if (k0(s) >= 0) break;
    return s;
  }
};

// A_Q = B_Q C for every row Q of the packed factor
// (B_Q symmetric n x n, lower triangle row-major, packed row stride np; C
// n x o row-major), written as At[m][Q o + i] (the exchange product's
// operand, row stride ldt), and with part, the tile's sum of A_Q . C
// (B_Q . P for P = C C^T without another pass over B).  One workgroup of NT
// threads per 128 x 160 tile of A_Q, in the work order above (all 160 columns of a 160-orbital C in one tile,
// so that B is read once).  Per 16-wide contraction step the B panel comes
// from the packed rows -- along packed row k where the panel lies above the
// diagonal (k > every m of the tile), along packed row m below it, by
// (max, min) where the diagonal crosses it -- consecutive lanes reading
// consecutive doubles; the next panel is loaded into registers while the
// current one is multiplied out of the other LDS buffer, one barrier a step.
// Waves (NT / 128) x 2, each 128 / (NT / 128) rows x 80 columns of 16 x 16
// MFMA blocks.
constexpr int AM = 128, AN = 160, AK = 32;
static_assert(AM == PanelOrder<AK>::TS, "tiles of the panel order are the row tiles");
constexpr int APB = AK + 1, APC = AN + 1;

template <int NT>
struct BcRegs {
  static constexpr int LR = NT / AK;         // lanes along a panel row
  static constexpr int NB = AM * AK / NT;    // B values per thread
  static constexpr int NC = AN * AK / NT;    // C values per thread
  double b[NB], c[NC];
};

template <int NT>
__device__ __forceinline__ void bc_load(const double* __restrict__ B, int n, const double* __restrict__ C, int o,
                                        int m0, int i0, int k0, int t, BcRegs<NT>& r) {
  constexpr int LR = BcRegs<NT>::LR, NB = BcRegs<NT>::NB, NC = BcRegs<NT>::NC;
  if (k0 >= m0 + AM) {
    const int k = k0 + t / LR;
    const size_t base = (size_t)k * (k + 1) / 2;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int m = m0 + t % LR + LR * c;
      r.b[c] = (k < n && m < n) ? B[base + m] : 0.0;
    }
  } else if (k0 + AK <= m0) {
    const int k = k0 + t % AK;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int m = m0 + t / AK + LR * c;
      r.b[c] = (k < n && m < n) ? B[(size_t)m * (m + 1) / 2 + k] : 0.0;
    }
  } else {
    const int k = k0 + t % AK;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const int m = m0 + t / AK + LR * c;
      const size_t hi = m > k ? m : k, lo = m > k ? k : m;
      r.b[c] = (k < n && m < n) ? B[hi * (hi + 1) / 2 + lo] : 0.0;
    }
  }
  const int k = k0 + t / LR;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int i = i0 + t % LR + LR * c;
    r.c[c] = (k < n && i < o) ? C[(size_t)k * o + i] : 0.0;
  }
}

template <int NT>
__device__ __forceinline__ void bc_store(int m0, int k0, int t, const BcRegs<NT>& r, double* Bs, double* Cs) {
  constexpr int LR = BcRegs<NT>::LR, NB = BcRegs<NT>::NB, NC = BcRegs<NT>::NC;
  if (k0 >= m0 + AM) {
#pragma unroll
    for (int c = 0; c < NB; ++c) Bs[(t % LR + LR * c) * APB + t / LR] = r.b[c];
  } else {
#pragma unroll
    for (int c = 0; c < NB; ++c) Bs[(t / AK + LR * c) * APB + t % AK] = r.b[c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) Cs[(t / LR) * APC + t % LR + LR * c] = r.c[c];
}

template <int NT>
__global__ __launch_bounds__(NT) void k_bc_packed_f64(const double* __restrict__ Bp, size_t np, int n,
                                                      const double* __restrict__ C, int o,
                                                      double* __restrict__ At, long long ldt,
                                                      double* __restrict__ part, int gx, int gy, long long total, int flags) {
  constexpr int WR = AM / (NT / 128), BA = WR / 16;  // rows per wave, MFMA blocks along them
  __shared__ double Bs[2][AM * APB], Cs[2][AK * APC];
  const WorkItem wi = work_item(gx, gy, total, flags);
  const int m0 = wi.mt * AM, i0 = wi.it * AN, q = wi.q;
  const double* B = Bp + (size_t)q * np;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  d4 acc[BA][5];
#pragma unroll
  for (int a = 0; a < BA; ++a)
#pragma unroll
    for (int b = 0; b < 5; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  // Panels in tournament order: panel s in buffer cur, the registers hold
  // panel s1.  A step stores the registers into the other buffer (whose
  // readers all passed the previous barrier), loads panel s2, multiplies, and
  // meets the other waves at one barrier.
  const PanelOrder<AK> po(wi.mt, n, flags);
  const int S = po.count();
  int s = po.next(-1), s1 = po.next(s);
  BcRegs<NT> r;
  bc_load<NT>(B, n, C, o, m0, i0, po.k0(s), t, r);
  bc_store<NT>(m0, po.k0(s), t, r, Bs[0], Cs[0]);
  if (s1 < S) bc_load<NT>(B, n, C, o, m0, i0, po.k0(s1), t, r);
  __syncthreads();
  for (int cur = 0; s < S; cur ^= 1) {
    const int s2 = s1 < S ? po.next(s1) : S;
    if (s1 < S) {
      bc_store<NT>(m0, po.k0(s1), t, r, Bs[cur ^ 1], Cs[cur ^ 1]);
      if (s2 < S) bc_load<NT>(B, n, C, o, m0, i0, po.k0(s2), t, r);
    }
    const double* bs = Bs[cur];
    const double* cs = Cs[cur];
#pragma unroll
    for (int kk = 0; kk < AK; kk += 4) {
      const int kq = kk + (lane >> 4), rq = lane & 15;
      double av[BA], bv[5];
#pragma unroll
      for (int a = 0; a < BA; ++a) av[a] = bs[(WR * wm + 16 * a + rq) * APB + kq];
#pragma unroll
      for (int b = 0; b < 5; ++b) bv[b] = cs[kq * APC + 80 * wn + 16 * b + rq];
#pragma unroll
      for (int a = 0; a < BA; ++a)
#pragma unroll
        for (int b = 0; b < 5; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
    s = s1;
    s1 = s2;
  }
  // A_Q[m][i] -> At[m][Q o + i]; with part, this tile's sum of A_Q . C
  double* Aq = At + (size_t)q * o;
  double sum = 0.0;
#pragma unroll
  for (int a = 0; a < BA; ++a)
#pragma unroll
    for (int b = 0; b < 5; ++b)
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const int m = m0 + WR * wm + 16 * a + (lane >> 4) + 4 * x;
        const int i = i0 + 80 * wn + 16 * b + (lane & 15);
        if (m < n && i < o) {
          Aq[(size_t)m * ldt + i] = acc[a][b][x];
          if (part) sum += acc[a][b][x] * C[(size_t)m * o + i];
        }
      }
  if (part) {  // fixed-order block sum
    __shared__ double red[NT];
    red[t] = sum;
    __syncthreads();
    for (int h = NT / 2; h > 0; h >>= 1) {
      if (t < h) red[t] += red[t + h];
      __syncthreads();
    }
    if (t == 0) part[(size_t)q * gx * gy + wi.it * gx + wi.mt] = red[0];
  }
}

// ---- split-BF16 form: A = B C with each operand x = hi + lo (hi = bf16(x),
// lo = bf16(x - hi), from the FP32 value) and three BF16 products hi hi +
// hi lo + lo hi accumulated in FP32, then A itself split into the hi and lo
// BF16 operands of the symmetric product.
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// FP32 -> BF16 bits, round to nearest even (finite values), as __float2bfloat16
__device__ __forceinline__ unsigned bf16_bits(float f) {
  const unsigned u = __float_as_uint(f);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_value(unsigned h) { return __uint_as_float(h << 16); }
__device__ __forceinline__ void split_bf16(float f, unsigned& h, unsigned& l) {
  h = bf16_bits(f);
  l = bf16_bits(f - bf16_value(h));
}

constexpr int SM = 128, SN = 160, SK = 32, SP = 40;  // tile rows, columns, contraction step, LDS row (BF16)
static_assert(SM == PanelOrder<SK>::TS, "tiles of the panel order are the row tiles");
constexpr int SBUF = (2 * SM + 2 * SN) * SP;         // BF16 per LDS buffer: B hi, B lo, C hi, C lo

// C (n x o, FP64) -> hi and lo parts of C^T, [oc][nk] BF16, zero past (o, n)
__global__ void k_split_ct(const double* __restrict__ C, int n, int o, int nk, int oc, unsigned short* __restrict__ Th,
                           unsigned short* __restrict__ Tl) {
  const size_t tot = (size_t)oc * nk;
  for (size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x; x < tot; x += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(x / nk), k = (int)(x - (size_t)i * nk);
    unsigned h = 0, l = 0;
    if (i < o && k < n) split_bf16((float)C[(size_t)k * o + i], h, l);
    Th[x] = (unsigned short)h;
    Tl[x] = (unsigned short)l;
  }
}

struct SplitRegs {
  double b[8];       // B[m][k0 + 8 g + c], c < 8
  uint4 ch[2], cl[2];  // 8 BF16 of C^T rows (chunks t and t + 512 of the 640)
};

__device__ __forceinline__ bool split_above(int m0, int k0) { return k0 >= m0 + SM; }

__device__ __forceinline__ void split_load(const double* __restrict__ B, int n, const unsigned short* __restrict__ Th,
                                           const unsigned short* __restrict__ Tl, int nk, int m0, int i0, int k0,
                                           int t, SplitRegs& r) {
  if (split_above(m0, k0)) {  // packed row k, consecutive m across lanes
    const int m = m0 + (t & 127), g = t >> 7;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int k = k0 + 8 * g + c;
      r.b[c] = (k < n && m < n) ? B[(size_t)k * (k + 1) / 2 + m] : 0.0;
    }
  } else if (k0 + SK <= m0) {  // packed row m, consecutive k
    const int m = m0 + (t >> 2), g = t & 3;
    const size_t base = (size_t)m * (m + 1) / 2;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int k = k0 + 8 * g + c;
      r.b[c] = (k < n && m < n) ? B[base + k] : 0.0;
    }
  } else {
    const int m = m0 + (t >> 2), g = t & 3;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int k = k0 + 8 * g + c;
      const size_t hi = m > k ? m : k, lo = m > k ? k : m;
      r.b[c] = (k < n && m < n) ? B[hi * (hi + 1) / 2 + lo] : 0.0;
    }
  }
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int cid = t + 512 * h;
    if (h == 0 || t < 128) {
      const size_t at = (size_t)(i0 + (cid >> 2)) * nk + k0 + 8 * (cid & 3);
      r.ch[h] = *(const uint4*)(Th + at);
      r.cl[h] = *(const uint4*)(Tl + at);
    }
  }
}

__device__ __forceinline__ void split_store(int m0, int k0, int t, const SplitRegs& r, unsigned short* S) {
  unsigned short *Bh = S, *Bl = S + SM * SP, *Ch = S + 2 * SM * SP, *Cl = Ch + SN * SP;
  const bool above = split_above(m0, k0);
  const int mm = above ? (t & 127) : (t >> 2), g = above ? (t >> 7) : (t & 3);
  unsigned h[8], l[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) split_bf16((float)r.b[c], h[c], l[c]);
  *(uint4*)(Bh + mm * SP + 8 * g) = uint4{h[0] | h[1] << 16, h[2] | h[3] << 16, h[4] | h[5] << 16, h[6] | h[7] << 16};
  *(uint4*)(Bl + mm * SP + 8 * g) = uint4{l[0] | l[1] << 16, l[2] | l[3] << 16, l[4] | l[5] << 16, l[6] | l[7] << 16};
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    const int cid = t + 512 * x;
    if (x == 0 || t < 128) {
      *(uint4*)(Ch + (cid >> 2) * SP + 8 * (cid & 3)) = r.ch[x];
      *(uint4*)(Cl + (cid >> 2) * SP + 8 * (cid & 3)) = r.cl[x];
    }
  }
}

// The split-BF16 A = B C from the packed factor: tiles, panel cases and
// pipeline as k_bc_packed_f64 (128 x 160 per 512-thread workgroup, waves
// 4 x 2 of 32 x 80), a 32-wide contraction step per v_mfma_f32_16x16x32_bf16;
// B split while it is staged, C^T pre-split ([oc][nk], k contiguous, so each
// lane's 8 contraction values are one 16-byte LDS read for either operand).
// Output: the hi and lo BF16 parts of A at Ah/Al[m][Q o + i] (row stride ldt).
__global__ __launch_bounds__(512) void k_bc_packed_bf16x3(const double* __restrict__ Bp, size_t np, int n,
                                                          const unsigned short* __restrict__ Th,
                                                          const unsigned short* __restrict__ Tl, int nk, int o,
                                                          unsigned short* __restrict__ Ah,
                                                          unsigned short* __restrict__ Al, long long ldt, int gx,
                                                          int gy, long long total, int flags) {
  __shared__ uint4 lds[2][SBUF / 8];
  const WorkItem wi = work_item(gx, gy, total, flags);
  const int m0 = wi.mt * SM, i0 = wi.it * SN, q = wi.q;
  const double* B = Bp + (size_t)q * np;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 1, wn = w & 1, rq = lane & 15, g = lane >> 4;
  f4 acc[2][5];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 5; ++b) acc[a][b] = f4{0.f, 0.f, 0.f, 0.f};
  const PanelOrder<SK> po(wi.mt, n, flags);
  const int S = po.count();
  int s = po.next(-1), s1 = po.next(s);
  SplitRegs r;
  split_load(B, n, Th, Tl, nk, m0, i0, po.k0(s), t, r);
  split_store(m0, po.k0(s), t, r, (unsigned short*)lds[0]);
  if (s1 < S) split_load(B, n, Th, Tl, nk, m0, i0, po.k0(s1), t, r);
  __syncthreads();
  for (int cur = 0; s < S; cur ^= 1) {
    const int s2 = s1 < S ? po.next(s1) : S;
    if (s1 < S) {
      split_store(m0, po.k0(s1), t, r, (unsigned short*)lds[cur ^ 1]);
      if (s2 < S) split_load(B, n, Th, Tl, nk, m0, i0, po.k0(s2), t, r);
    }
    const unsigned short* S = (const unsigned short*)lds[cur];
    const unsigned short *Bh = S, *Bl = S + SM * SP, *Ch = S + 2 * SM * SP, *Cl = Ch + SN * SP;
    bf8 ah[2], al[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      ah[a] = *(const bf8*)(Bh + (32 * wm + 16 * a + rq) * SP + 8 * g);
      al[a] = *(const bf8*)(Bl + (32 * wm + 16 * a + rq) * SP + 8 * g);
    }
#pragma unroll
    for (int b = 0; b < 5; ++b) {
      const bf8 ch = *(const bf8*)(Ch + (80 * wn + 16 * b + rq) * SP + 8 * g);
      const bf8 cl = *(const bf8*)(Cl + (80 * wn + 16 * b + rq) * SP + 8 * g);
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], ch, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[a], cl, acc[a][b], 0, 0, 0);
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[a], ch, acc[a][b], 0, 0, 0);
      }
    }
    __syncthreads();
    s = s1;
    s1 = s2;
  }
  // D of the FP32 16 x 16 forms: lane l, register x holds row 4 (l / 16) + x, column l % 16
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 5; ++b)
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const int m = m0 + 32 * wm + 16 * a + 4 * g + x;
        const int i = i0 + 80 * wn + 16 * b + rq;
        if (m < n && i < o) {
          unsigned h, l;
          split_bf16(acc[a][b][x], h, l);
          const size_t at = (size_t)m * ldt + (size_t)q * o + i;
          Ah[at] = (unsigned short)h;
          Al[at] = (unsigned short)l;
        }
      }
}

// gamma[Q] = alpha * (sum of the np tile partials of row Q, in order)
__global__ void k_part_reduce(const double* __restrict__ part, int np, int rank, double alpha,
                              double* __restrict__ gamma) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rank) return;
  double s = 0.0;
  for (int j = 0; j < np; ++j) s += part[(size_t)q * np + j];
  gamma[q] = alpha * s;
}

}  // namespace

// A = B C for every row of the packed factor (rank rows of stride np), into
// the layout of the exchange product: At[m][Q o + i] (row stride ldt >=
// rank o).  With gamma, also gamma[Q] = alpha sum_{m,i} A_Q[m][i] C[m][i]
// (= tr(B_Q P) for P = alpha C C^T), reduced in a fixed order through
// scratch.  On stream st.
void bc_packed_f64(const double* Bp, size_t np, int rank, int n, const double* C, int o, double* At,
                   long long ldt, double alpha, double* gamma, DevBuf& scratch, hipStream_t st) {
  const unsigned gx = (unsigned)((n + AM - 1) / AM), gy = (unsigned)((o + AN - 1) / AN);
  double* part = nullptr;
  if (gamma) {
    const size_t need = (size_t)rank * gx * gy * 8;
    if (scratch.bytes < need) scratch.alloc(need);
    part = scratch.as<double>();
  }
  const long long total = (long long)gx * gy * rank;
  if (total > INT32_MAX) throw HfgError{HFG_UNSUPPORTED, "packed A = B C: too many tiles for one launch"};
  // 512 threads (2 waves a SIMD) measured 5 % faster than 256 (1 wave a SIMD,
  // 160 accumulator registers) on (H2O)32 / cc-pVTZ
  hipLaunchKernelGGL(k_bc_packed_f64<512>, dim3((unsigned)total), dim3(512), 0, st, Bp, np, n, C, o, At, ldt, part,
                     (int)gx, (int)gy, total, BC_XCD_MAP);
  HIP_CHECK(hipGetLastError());
  if (gamma) {
    hipLaunchKernelGGL(k_part_reduce, dim3((unsigned)((rank + 255) / 256)), dim3(256), 0, st, part, (int)(gx * gy),
                       rank, alpha, gamma);
    HIP_CHECK(hipGetLastError());
  }
}

// The split-BF16 A = B C from the packed factor (rank rows of stride np),
// written as the hi and lo BF16 parts Ah/Al[m][Q o + i] (row stride ldt >=
// rank o); ct holds the split C^T.  On stream st.
void bc_packed_bf16x3(const double* Bp, size_t np, int rank, int n, const double* C, int o, unsigned short* Ah,
                      unsigned short* Al, long long ldt, DevBuf& ct, hipStream_t st) {
  const int nk = (n + SK - 1) / SK * SK, oc = (o + SN - 1) / SN * SN;
  const size_t tot = (size_t)oc * nk;
  if (ct.bytes < 2 * tot * 2) ct.alloc(2 * tot * 2);
  unsigned short* Th = ct.as<unsigned short>();
  unsigned short* Tl = Th + tot;
  hipLaunchKernelGGL(k_split_ct, dim3((unsigned)std::min<size_t>(2048, (tot + 255) / 256)), dim3(256), 0, st, C, n, o,
                     nk, oc, Th, Tl);
  HIP_CHECK(hipGetLastError());
  const unsigned gx = (unsigned)((n + SM - 1) / SM), gy = (unsigned)(oc / SN);
  const long long total = (long long)gx * gy * rank;
  if (total > INT32_MAX) throw HfgError{HFG_UNSUPPORTED, "packed A = B C: too many tiles for one launch"};
  hipLaunchKernelGGL(k_bc_packed_bf16x3, dim3((unsigned)total), dim3(512), 0, st, Bp, np, n, Th, Tl, nk, o, Ah, Al,
                     ldt, (int)gx, (int)gy, total, BC_XCD_MAP | BC_TOURNAMENT);
  HIP_CHECK(hipGetLastError());
}

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
  hipLaunchKernelGGL(k_syrk_f64, dim3((unsigned)(upper * S)), dim3(SYRK_NT), 0, st, X, n, L, ldx, tiles, upper,
                     kslice, part);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_syrk_reduce, dim3(2048), dim3(256), 0, st, part, S, n, alpha, K);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));  // the tile list is a host vector
}

}  // namespace hfg
