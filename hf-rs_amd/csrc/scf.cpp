// scf.cpp -- restricted Hartree-Fock driver with device-resident state.
//
// Follows run_scf_in_space_with_convergence (src/scf.rs:410-627): SAD guess
// (src/sad.rs), loop {F = H + J - K/2 on the GPU, E = 1/2 sum P(H+F),
// e = FPS - SPF (projected by X when the orthogonaliser is rectangular),
// converged when dE < tol, max|e| < residual and rms(dP) < density,
// ADIIS/CDIIS extrapolation (src/diis.rs), eigh(X^T F X), P = 2 C_occ C_occ^T},
// then a final diagonalisation of the true Fock matrix and a last Fock build.
// F, P, S, H, X, the DIIS histories and the eigensolver live in HBM; FP64
// GEMMs go to rocBLAS and the symmetric eigensolver to rocSOLVER (dsyevd).
// Only scalars (energies, norms, DIIS inner products) come back to the host.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>

#include "engine.h"
#include "la_device.h"
#include "rhf.h"

namespace hfg {

// ---------------------------------------------------------------------------
// Small dense host algebra for the atomic (SAD) problems, n <= ~40.
// ---------------------------------------------------------------------------
typedef std::vector<double> HMat;
static HMat hmm(const HMat& a, const HMat& b, int m, int k, int n, bool ta = false, bool tb = false) {
  HMat c((size_t)m * n, 0.0);
  for (int i = 0; i < m; ++i)
    for (int p = 0; p < k; ++p) {
      const double x = ta ? a[(size_t)p * m + i] : a[(size_t)i * k + p];
      if (x == 0.0) continue;
      for (int j = 0; j < n; ++j) c[(size_t)i * n + j] += x * (tb ? b[(size_t)j * k + p] : b[(size_t)p * n + j]);
    }
  return c;
}
// Cyclic Jacobi with ascending eigenvalues (vectors in columns).
static void jacobi(int n, HMat a, std::vector<double>& w, HMat& v) {
  v.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, tot = 0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) (i == j ? tot : off) += a[(size_t)i * n + j] * a[(size_t)i * n + j];
    if (off <= 1e-30 * (tot + off) || off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a[(size_t)p * n + q];
        if (apq == 0.0) continue;
        const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
          a[(size_t)k * n + p] = c * akp - s * akq;
          a[(size_t)k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
          a[(size_t)p * n + k] = c * apk - s * aqk;
          a[(size_t)q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = v[(size_t)k * n + p], vkq = v[(size_t)k * n + q];
          v[(size_t)k * n + p] = c * vkp - s * vkq;
          v[(size_t)k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> ord(n);
  for (int i = 0; i < n; ++i) ord[i] = i;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return a[(size_t)x * n + x] < a[(size_t)y * n + y]; });
  w.resize(n);
  HMat v2((size_t)n * n);
  for (int c = 0; c < n; ++c) {
    w[c] = a[(size_t)ord[c] * n + ord[c]];
    for (int r = 0; r < n; ++r) v2[(size_t)r * n + c] = v[(size_t)r * n + ord[c]];
  }
  v = v2;
}
static bool host_orthogonalizer(int n, const HMat& s, double cut, HMat& x, int& m) {
  std::vector<double> w;
  HMat v;
  jacobi(n, s, w, v);
  if (w[0] < -cut) return false;
  std::vector<int> keep;
  for (int i = 0; i < n; ++i)
    if (w[i] >= cut) keep.push_back(i);
  m = (int)keep.size();
  if (!m) return false;
  x.assign((size_t)n * m, 0.0);
  for (int c = 0; c < m; ++c)
    for (int r = 0; r < n; ++r) x[(size_t)r * m + c] = v[(size_t)r * n + keep[c]] / std::sqrt(w[keep[c]]);
  return true;
}
// Fractional occupations with equal shares inside degenerate groups (|de| < 1e-6).
static std::vector<double> frac_occ(const std::vector<double>& e, double nel) {
  if (nel < 0 || nel > 2.0 * e.size() + 1e-9) throw HfgError{HFG_INVALID_INPUT, "atomic electron count exceeds basis capacity"};
  std::vector<double> occ(e.size(), 0.0);
  double left = nel;
  size_t i = 0;
  while (i < e.size() && left > 1e-12) {
    size_t j = i + 1;
    while (j < e.size() && std::fabs(e[j] - e[i]) < 1e-6) ++j;
    const double pop = std::min(left, 2.0 * (j - i));
    for (size_t k = i; k < j; ++k) occ[k] = pop / (j - i);
    left -= pop;
    i = j;
  }
  return occ;
}

// ---------------------------------------------------------------------------
// SAD: fractional-occupation atomic RHF per element (src/sad.rs:56-134), with
// the two-electron part from the GPU engine on the atom's own basis.
// ---------------------------------------------------------------------------
struct AtomicDensity {
  HMat p;
  double energy = 0;
};

static AtomicDensity atomic_hf(const std::vector<Shell>& sh_in, double z) {
  std::vector<Shell> sh = sh_in;
  int n = 0;
  for (auto& s : sh) {
    s.c[0] = s.c[1] = s.c[2] = 0.0;
    s.atom = 0;
    s.ao = n;
    n += (int)s.scale.size();
  }
  std::vector<std::array<double, 4>> atoms{{z, 0, 0, 0}};
  HMat S, T, V;
  one_electron(sh, n, atoms, S, T, V);
  HMat H(S.size());
  for (size_t i = 0; i < H.size(); ++i) H[i] = T[i] + V[i];
  HMat X;
  int m;
  if (!host_orthogonalizer(n, S, 1e-8, X, m)) throw HfgError{HFG_INTEGRAL, "atomic overlap is singular"};
  // device engine over the atomic basis
  std::vector<int32_t> atom(n, 0), lx(3 * n), po(n + 1);
  std::vector<double> cen(3 * n, 0.0), ex, co;
  int f = 0;
  for (auto& s : sh) {
    std::vector<std::array<int, 3>> cmp;
    if (s.L == 2)
      cmp = {{2, 0, 0}, {0, 2, 0}, {0, 0, 2}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
    else
      for (int a = s.L; a >= 0; --a)
        for (int b = s.L - a; b >= 0; --b) cmp.push_back({a, b, s.L - a - b});
    for (size_t c = 0; c < cmp.size(); ++c, ++f) {
      for (int k = 0; k < 3; ++k) lx[3 * f + k] = cmp[c][k];
      po[f] = (int32_t)ex.size();
      for (int k = 0; k < s.K; ++k) {
        ex.push_back(s.e[k]);
        co.push_back(s.r[k] * s.scale[c]);
      }
    }
  }
  po[n] = (int32_t)ex.size();
  hfg_basis_view view{n, atom.data(), cen.data(), lx.data(), po.data(), ex.data(), co.data()};
  std::unique_ptr<Engine> eng;
  const bool hydrogenic = std::fabs(z - 1.0) < 1e-12;
  const bool trace = std::getenv("HFG_TRACE") != nullptr;
  const auto t_start = std::chrono::steady_clock::now();
  auto since = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count(); };
  if (!hydrogenic) eng.reset(new Engine(view));
  if (trace) std::fprintf(stderr, "[hfg] atomic HF Z=%g n=%d: engine %.4f s\n", z, n, since());
  const size_t nn = (size_t)n * n;
  auto fock = [&](const HMat& P) {
    if (hydrogenic) return H;
    HMat F(nn);
    HIP_CHECK(hipMemcpyAsync(eng->d_D.p, P.data(), nn * 8, hipMemcpyHostToDevice, eng->stream));
    HIP_CHECK(hipMemcpyAsync(eng->d_H.p, H.data(), nn * 8, hipMemcpyHostToDevice, eng->stream));
    eng->jk_device(eng->d_D.as<double>(), nullptr, nullptr, eng->d_H.as<double>(), eng->d_F.as<double>());
    HIP_CHECK(hipMemcpyAsync(F.data(), eng->d_F.p, nn * 8, hipMemcpyDeviceToHost, eng->stream));
    eng->sync();
    return F;
  };
  auto density = [&](const HMat& F, std::vector<double>& eps) {
    HMat fp = hmm(hmm(X, F, m, n, n, true), X, m, n, m);
    HMat cp;
    jacobi(m, fp, eps, cp);
    HMat C = hmm(X, cp, n, m, m);
    auto occ = frac_occ(eps, z);
    HMat w = C;
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < m; ++c) w[(size_t)r * m + c] *= std::sqrt(occ[c]);
    return hmm(w, w, n, m, n, false, true);
  };
  auto commutator = [&](const HMat& F, const HMat& P) {
    HMat a = hmm(hmm(F, P, n, n, n), S, n, n, n), b = hmm(hmm(S, P, n, n, n), F, n, n, n);
    for (size_t i = 0; i < a.size(); ++i) a[i] -= b[i];
    return a;
  };
  auto energy = [&](const HMat& P, const HMat& F) {
    double e = 0;
    for (size_t i = 0; i < nn; ++i) e += P[i] * (H[i] + F[i]);
    return 0.5 * e;
  };
  std::vector<double> eps;
  HMat P = density(H, eps);
  double prev = INFINITY;
  std::vector<HMat> fh, eh;
  for (int it = 1; it <= 200; ++it) {
    HMat F = fock(P);
    const double e = energy(P, F);
    HMat next = density(F, eps);
    HMat err = commutator(F, P);
    double emax = 0, rms = 0;
    for (double v : err) emax = std::max(emax, std::fabs(v));
    for (size_t i = 0; i < nn; ++i) rms += (P[i] - next[i]) * (P[i] - next[i]);
    rms = std::sqrt(rms / nn);
    if (std::fabs(e - prev) < 1e-11 && rms < 1e-9 && emax < 1e-8) {
      if (trace) std::fprintf(stderr, "[hfg] atomic HF Z=%g: %d iterations, %.4f s\n", z, it, since());
      HMat Fn = fock(next);
      return {next, energy(next, Fn)};
    }
    prev = e;
    if (fh.size() >= 8) {
      fh.erase(fh.begin());
      eh.erase(eh.begin());
    }
    fh.push_back(F);
    eh.push_back(err);
    HMat fu = F;
    const int k = (int)eh.size();
    if (k >= 2) {
      HMat B((size_t)(k + 1) * (k + 1), 0.0);
      for (int i = 0; i < k; ++i) {
        for (int j = 0; j < k; ++j) {
          double d = 0;
          for (size_t x = 0; x < nn; ++x) d += eh[i][x] * eh[j][x];
          B[(size_t)i * (k + 1) + j] = d;
        }
        B[(size_t)i * (k + 1) + k] = B[(size_t)k * (k + 1) + i] = -1.0;
      }
      std::vector<double> rhs(k + 1, 0.0), c;
      rhs[k] = -1.0;
      if (solve_small(B, rhs, k + 1, c)) {
        fu.assign(nn, 0.0);
        for (int i = 0; i < k; ++i)
          for (size_t x = 0; x < nn; ++x) fu[x] += c[i] * fh[i][x];
      }
    }
    next = density(fu, eps);
    if (it <= 5)
      for (size_t i = 0; i < nn; ++i) P[i] = 0.5 * P[i] + 0.5 * next[i];
    else
      P = next;
  }
  throw HfgError{HFG_NOT_CONVERGED, "fractional atomic HF did not converge"};
}

static std::string shell_key(const std::vector<Shell>& sh, double z) {
  std::ostringstream os;
  os.precision(17);
  os << z;
  for (auto& s : sh) {
    os << "|" << s.L << ":" << s.K;
    for (int k = 0; k < s.K; ++k) os << "," << s.e[k] << "," << s.r[k];
    for (double c : s.scale) os << ";" << c;
  }
  return os.str();
}

HMat atomic_density(const std::vector<Shell>& sh, int atom, double z, std::vector<int>& idx) {
  static std::mutex mu;
  static std::map<std::string, AtomicDensity> cache;  // src/sad.rs:136-214
  std::vector<Shell> loc;
  for (auto& s : sh)
    if (s.atom == atom) loc.push_back(s);
  if (loc.empty()) throw HfgError{HFG_INVALID_INPUT, "atom without basis functions"};
  const std::string key = shell_key(loc, z);
  AtomicDensity ad;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) ad = it->second;
  }
  if (ad.p.empty()) {
    ad = atomic_hf(loc, z);
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = ad;
  }
  idx.clear();
  for (auto& s : loc)
    for (size_t c = 0; c < s.scale.size(); ++c) idx.push_back(s.ao + (int)c);
  return ad.p;
}

RhfInputs rhf_inputs(const hfg_basis_view& basis, const double* zxyz, int natom, int charge) {
  RhfInputs in;
  if (natom <= 0 || !zxyz) throw HfgError{HFG_INVALID_INPUT, "no atoms"};
  in.atoms.resize(natom);
  double zsum = 0;
  for (int a = 0; a < natom; ++a) {
    for (int k = 0; k < 4; ++k) in.atoms[a][k] = zxyz[4 * a + k];
    zsum += in.atoms[a][0];
  }
  in.nel = (int)std::lround(zsum) - charge;
  if (in.nel <= 0) throw HfgError{HFG_INVALID_INPUT, "invalid electron count"};
  if (in.nel % 2) throw HfgError{HFG_INVALID_INPUT, "RHF requires an even number of electrons; use UHF for open shells"};
  in.nocc = in.nel / 2;
  if (in.nocc > basis.n_functions) throw HfgError{HFG_INVALID_INPUT, "more occupied orbitals than basis functions"};
  for (int i = 0; i < natom; ++i)
    for (int j = i + 1; j < natom; ++j) {
      const double dx = in.atoms[i][1] - in.atoms[j][1], dy = in.atoms[i][2] - in.atoms[j][2],
                   dz = in.atoms[i][3] - in.atoms[j][3];
      in.enuc += in.atoms[i][0] * in.atoms[j][0] / std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  return in;
}

// Canonical orthogonaliser X = Tr U_kept s^{-1/2} of S' = Tr^T S Tr (src/ao.rs:58-67,
// Matrix::canonical_orthogonalizer src/matrix.rs:329-346, cut 1e-8), the
// eigensolver on the device; X (n x m) returned on the host.
int orthogonalizer(Blas& bl, hipStream_t st, const std::vector<Shell>& shells, int n, const double* dS, bool spherical,
                   HMat& X) {
  HMat Tr;
  int m0 = n;
  if (spherical) spherical_transform(shells, n, Tr, m0);
  Dev dSp, dTr, dTmp;
  dSp.alloc((size_t)m0 * m0);
  if (spherical) {
    dTr.alloc((size_t)n * m0);
    dTmp.alloc((size_t)n * m0);
    upload(dTr.p, Tr.data(), (size_t)n * m0, st);
    bl.gemm(false, false, n, m0, n, 1.0, dS, n, dTr.p, m0, 0.0, dTmp.p, m0);      // S Tr
    bl.gemm(true, false, m0, m0, n, 1.0, dTr.p, m0, dTmp.p, m0, 0.0, dSp.p, m0);  // Tr^T S Tr
  } else {
    HIP_CHECK(hipMemcpyAsync(dSp.p, dS, (size_t)n * n * 8, hipMemcpyDeviceToDevice, st));
  }
  Dev dw;
  dw.alloc(m0);
  Eigh eig;
  eig.run(bl, dSp.p, dw.p, m0);
  std::vector<double> w(m0);
  download(w.data(), dw.p, m0, st);
  if (w[0] < -1e-8) throw HfgError{HFG_INTEGRAL, "overlap has negative eigenvalue"};
  int first = 0;
  while (first < m0 && w[first] < 1e-8) ++first;
  const int m = m0 - first;
  if (m <= 0) throw HfgError{HFG_INTEGRAL, "overlap has zero retained rank"};
  // U (row-major m0 x m0, columns = vectors) is the transpose of dSp's row-major view.
  HMat U((size_t)m0 * m0);
  download(U.data(), dSp.p, (size_t)m0 * m0, st);
  HMat Xp((size_t)m0 * m);
  for (int c = 0; c < m; ++c)
    for (int r = 0; r < m0; ++r) Xp[(size_t)r * m + c] = U[(size_t)(first + c) * m0 + r] / std::sqrt(w[first + c]);
  X = spherical ? hmm(Tr, Xp, n, m0, m) : Xp;
  return m;
}

int scf_rhf(const hfg_basis_view* basis, const double* zxyz, int natom, int charge, const hfg_scf_options* opt,
            hfg_scf_result* res, double* p_out, double* f_out, double* eps_out, const double* p_guess) {
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  std::memset(res, 0, sizeof *res);
  const double tol = opt->energy_tol;
  if (opt->max_iter <= 0 || basis->n_functions <= 0 || !(tol > 0) || !(opt->residual_tol > 0) ||
      !(opt->density_tol > 0))
    throw HfgError{HFG_INVALID_INPUT, "SCF requires a nonempty basis, positive tolerance and iteration limit"};
  const RhfInputs in = rhf_inputs(*basis, zxyz, natom, charge);
  const std::vector<std::array<double, 4>>& atoms = in.atoms;
  const int nocc = in.nocc;
  const int n = basis->n_functions;

  auto t_phase = clk::now();
  auto lap = [&]() {
    const auto now = clk::now();
    const double d = std::chrono::duration<double>(now - t_phase).count();
    t_phase = now;
    return d;
  };
  std::unique_ptr<Engine> eng_p(new Engine(*basis));
  Engine& eng = *eng_p;
  res->seconds_setup = lap();
  hipStream_t st = eng.stream;
  Blas bl(st);
  const size_t nn = (size_t)n * n;
  const double enuc = in.enuc;

  // S and H = T + V on the device (onee_dev.hip)
  Dev dS, dH, dX, dTmp, dTmp2, dF, dP, dPprev, dErr, dW, dC, dFu;
  dS.alloc(nn);
  dH.alloc(nn);
  one_electron_device(eng.shells, n, atoms, dS.p, nullptr, nullptr, dH.p, st);
  res->seconds_one_electron = lap();
  dTmp.alloc(nn);
  dTmp2.alloc(nn);
  dF.alloc(nn);
  dP.alloc(nn);
  dPprev.alloc(nn);
  dErr.alloc(nn);
  dW.alloc(n);
  dC.alloc(nn);
  dFu.alloc(nn);
  Eigh eig;
  int m = 0;
  {
    HMat X;
    m = orthogonalizer(bl, st, eng.shells, n, dS.p, opt->spherical != 0, X);
    if (nocc > m) throw HfgError{HFG_INVALID_INPUT, "more occupied orbitals than retained overlap rank"};
    dX.alloc((size_t)n * m);
    upload(dX.p, X.data(), (size_t)n * m, st);
  }
  const int tb = 256;
  auto grid = [&](size_t cnt) { return dim3((unsigned)((cnt + tb - 1) / tb)); };
  double t_fock = 0;
  int builds = 0;
  const bool trace = std::getenv("HFG_TRACE") != nullptr;
  auto fock = [&](const double* P, double* F) {
    const auto t0 = clk::now();
    eng.jk_device(P, nullptr, nullptr, dH.p, F);
    HIP_CHECK(hipStreamSynchronize(st));
    const double dt = std::chrono::duration<double>(clk::now() - t0).count();
    t_fock += dt;
    if (trace) fprintf(stderr, "hfg scf full build: %.2f ms\n", dt * 1e3);
    ++builds;
  };
  // Incremental builds inside the iterations (opt->incremental_fock): G, the
  // two-electron part of the last build, and Pb, its density, are kept; a
  // build of P adds J - K/2 of dP = P - Pb, screened at 1e-12 / max|dP|.
  // Every 12th build, and the first, is full.
  const bool incremental = opt->incremental_fock != 0;
  Dev dG2, dPb;
  int since_full = -1;  // -1: no G yet
  auto fock_iter = [&](const double* P, double* F) {
    if (!incremental) return fock(P, F);
    if (since_full < 0 || since_full >= 12) {
      fock(P, F);
      if (!dG2.p) {
        dG2.alloc(nn);
        dPb.alloc(nn);
      }
      hipLaunchKernelGGL(k_sub, grid(nn), dim3(tb), 0, st, F, dH.p, dG2.p, nn);
      HIP_CHECK(hipMemcpyAsync(dPb.p, P, nn * 8, hipMemcpyDeviceToDevice, st));
      since_full = 0;
      return;
    }
    const auto t0 = clk::now();
    hipLaunchKernelGGL(k_sub, grid(nn), dim3(tb), 0, st, P, dPb.p, dPb.p, nn);  // dPb <- P - Pb
    const double dmax = bl.amax(nn, dPb.p);
    if (dmax > 0.0) eng.jk_device(dPb.p, nullptr, nullptr, dG2.p, dG2.p, dmax);
    hipLaunchKernelGGL(k_add, grid(nn), dim3(tb), 0, st, dH.p, dG2.p, F, nn);
    HIP_CHECK(hipMemcpyAsync(dPb.p, P, nn * 8, hipMemcpyDeviceToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
    const double dt = std::chrono::duration<double>(clk::now() - t0).count();
    t_fock += dt;
    if (trace) fprintf(stderr, "hfg scf incremental build: max|dP| %.3e, %.2f ms\n", dmax, dt * 1e3);
    ++builds;
    ++since_full;
  };
  // Inside the iterations the density needs only the projector onto the nocc
  // lowest eigenvectors of F' = X^T F X, not the eigenpairs: it comes from
  // SP2 purification (one m x m GEMM per step, about 20-30 steps), which on
  // the GPU costs a fraction of the tridiagonalisation-bound dsyevd.  The
  // steps stop once tr(X - X^2) < 1e-9, then three more (quadratic
  // convergence takes the error to rounding); without convergence in 100
  // steps (a vanishing HOMO-LUMO gap) the eigensolver is used.  The final
  // density, the orbital energies and the atomic SAD solves always use the
  // eigensolver.  HFG_SCF_EIGH=1 uses it everywhere.
  const bool use_sp2 = std::getenv("HFG_SCF_EIGH") == nullptr && nocc > 0 && nocc < m;
  Sp2 sp2;
  auto sp2_density = [&](double* P) {  // F' in dTmp2 (m x m); false: not converged
    if (!sp2.run(bl, st, dTmp2.p, m, nocc)) return false;
    // P = 2 X D X^T
    bl.gemm(false, false, n, m, m, 1.0, dX.p, m, sp2.x.p, m, 0.0, dTmp.p, m);
    bl.gemm(false, true, n, n, m, 2.0, dTmp.p, m, dX.p, m, 0.0, P, n);
    return true;
  };
  // diagonalise X^T F X, C = X C', P = 2 C_occ C_occ^T; eigenvalues into eps
  std::vector<double> eps(m);
  auto diag_density = [&](const double* F, double* P, bool want_eps) {
    bl.gemm(false, false, n, m, n, 1.0, F, n, dX.p, m, 0.0, dTmp.p, m);     // F X
    bl.gemm(true, false, m, m, n, 1.0, dX.p, m, dTmp.p, m, 0.0, dTmp2.p, m);  // X^T F X
    if (use_sp2 && !want_eps) {
      HIP_CHECK(hipMemcpyAsync(dC.p, dTmp2.p, (size_t)m * m * 8, hipMemcpyDeviceToDevice, st));
      if (sp2_density(P)) return;
      HIP_CHECK(hipMemcpyAsync(dTmp2.p, dC.p, (size_t)m * m * 8, hipMemcpyDeviceToDevice, st));
    }
    eig.run(bl, dTmp2.p, dW.p, m);
    // C = X V where the row-major view of dTmp2 is V^T
    bl.gemm(false, true, n, m, m, 1.0, dX.p, m, dTmp2.p, m, 0.0, dC.p, m);
    bl.gemm(false, true, n, n, nocc, 2.0, dC.p, m, dC.p, m, 0.0, P, n);
    if (want_eps) download(eps.data(), dW.p, m, st);
  };
  auto energy = [&](const double* P, const double* F) { return 0.5 * (bl.dot(nn, P, dH.p) + bl.dot(nn, P, F)); };

  // Initial density: SAD, else core Hamiltonian (src/scf.rs:452-471).  The
  // reference also diagonalises the SAD Fock matrix for the orbital-rotation
  // accelerators only; the Fock-extrapolation family used here never reads it.
  // A given density replaces the guess (run_scf_with_initial_density,
  // src/scf.rs:630-639).
  bool have_sad = true;
  if (p_guess) {
    upload(dP.p, p_guess, nn, st);
  } else {
    try {
      sad_guess_device(eng.shells, n, atoms, charge, dS.p, dP.p, bl, st);
    } catch (HfgError&) {
      have_sad = false;
    }
  }
  if (!have_sad) diag_density(dH.p, dP.p, false);
  HIP_CHECK(hipStreamSynchronize(st));
  res->seconds_guess = lap();

  // DIIS histories (device), src/diis.rs:72-137 and 289-359
  const int MAXS = 8;
  const int mm2 = m * m;
  Dev cF, cE, aF, aP, dG, dCoef;
  cF.alloc(MAXS * nn);
  cE.alloc((size_t)MAXS * mm2);
  aF.alloc(MAXS * nn);
  aP.alloc(MAXS * nn);
  dG.alloc(MAXS * MAXS);
  dCoef.alloc(MAXS);
  hipLaunchKernelGGL(k_zero, grid(MAXS * nn), dim3(tb), 0, st, cF.p, MAXS * nn);
  hipLaunchKernelGGL(k_zero, grid(MAXS * nn), dim3(tb), 0, st, aF.p, MAXS * nn);
  hipLaunchKernelGGL(k_zero, grid(MAXS * nn), dim3(tb), 0, st, aP.p, MAXS * nn);
  std::vector<int> cslot, aslot;  // logical order -> slot
  std::vector<double> aE;
  std::map<std::pair<int, int>, double> edot;  // slot-pair error inner products
  auto take_slot = [&](std::vector<int>& order) {
    int s;
    if ((int)order.size() >= MAXS) {
      s = order.front();
      order.erase(order.begin());
    } else {
      s = (int)order.size();
      for (int c = 0; c < MAXS; ++c)
        if (std::find(order.begin(), order.end(), c) == order.end()) {
          s = c;
          break;
        }
    }
    order.push_back(s);
    return s;
  };
  auto combine = [&](const Dev& hist, const std::vector<int>& order, const std::vector<double>& c, double* out) {
    std::vector<double> cs(MAXS, 0.0);
    for (size_t i = 0; i < order.size(); ++i) cs[order[i]] = c[i];
    upload(dCoef.p, cs.data(), MAXS, st);
    const double one = 1.0, zero = 0.0;
    RB_CHECK(rocblas_dgemv(bl.h, rocblas_operation_none, (rocblas_int)nn, MAXS, &one, hist.p, (rocblas_int)nn,
                           dCoef.p, 1, &zero, out, 1));
  };

  double prev_e = 0;
  bool conv = false;
  int iters = 0;
  HIP_CHECK(hipMemcpyAsync(dPprev.p, dP.p, nn * 8, hipMemcpyDeviceToDevice, st));
  for (int it = 1; it <= opt->max_iter; ++it) {
    iters = it;
    fock_iter(dP.p, dF.p);
    const double tot = energy(dP.p, dF.p) + enuc;
    const double de = std::fabs(tot - prev_e);
    prev_e = tot;
    // e = FPS - SPF = FPS - (FPS)^T
    bl.gemm(false, false, n, n, n, 1.0, dF.p, n, dP.p, n, 0.0, dTmp.p, n);
    bl.gemm(false, false, n, n, n, 1.0, dTmp.p, n, dS.p, n, 0.0, dTmp2.p, n);
    hipLaunchKernelGGL(k_antisym, grid(nn), dim3(tb), 0, st, dTmp2.p, dErr.p, n);
    const double* err = dErr.p;
    size_t elen = nn;
    if (m < n) {
      bl.gemm(false, false, n, m, n, 1.0, dErr.p, n, dX.p, m, 0.0, dTmp.p, m);
      bl.gemm(true, false, m, m, n, 1.0, dX.p, m, dTmp.p, m, 0.0, dTmp2.p, m);
      err = dTmp2.p;
      elen = mm2;
    }
    const double en = bl.amax(elen, err);
    hipLaunchKernelGGL(k_sub, grid(nn), dim3(tb), 0, st, dP.p, dPprev.p, dTmp.p, nn);
    const double rms = bl.nrm2(nn, dTmp.p) / std::sqrt((double)nn);
    if (opt->verbose) fprintf(stderr, "  hfg iter %3d  E=%.12f  dE=%.3e  |e|=%.3e\n", it, tot, de, en);
    if (it > 1 && de < tol && en < opt->residual_tol && rms < opt->density_tol) {
      conv = true;
      break;
    }
    HIP_CHECK(hipMemcpyAsync(dPprev.p, dP.p, nn * 8, hipMemcpyDeviceToDevice, st));
    bool use_cdiis;
    if (opt->conv_algo == 1)
      use_cdiis = true;
    else if (opt->conv_algo == 2)
      use_cdiis = false;
    else
      use_cdiis = !(en > (opt->conv_algo == 3 ? 0.08 : 0.05) || it <= 2);
    bool ok = false;
    if (use_cdiis) {
      const int s = take_slot(cslot);
      HIP_CHECK(hipMemcpyAsync(cF.p + s * nn, dF.p, nn * 8, hipMemcpyDeviceToDevice, st));
      HIP_CHECK(hipMemcpyAsync(cE.p + (size_t)s * mm2, err, elen * 8, hipMemcpyDeviceToDevice, st));
      for (auto it2 = edot.begin(); it2 != edot.end();)
        if (it2->first.first == s || it2->first.second == s)
          it2 = edot.erase(it2);
        else
          ++it2;
      for (int o : cslot) edot[{std::min(s, o), std::max(s, o)}] = bl.dot(elen, cE.p + (size_t)s * mm2, cE.p + (size_t)o * mm2);
      const int k = (int)cslot.size();
      if (k >= 2) {
        HMat B((size_t)(k + 1) * (k + 1), 0.0);
        for (int i = 0; i < k; ++i) {
          for (int j = 0; j < k; ++j)
            B[(size_t)i * (k + 1) + j] = edot[{std::min(cslot[i], cslot[j]), std::max(cslot[i], cslot[j])}];
          B[(size_t)i * (k + 1) + k] = B[(size_t)k * (k + 1) + i] = -1.0;
        }
        std::vector<double> rhs(k + 1, 0.0), c;
        rhs[k] = -1.0;
        if (solve_small(B, rhs, k + 1, c)) {
          c.resize(k);
          combine(cF, cslot, c, dFu.p);
          ok = true;
        }
      }
    } else {
      const int s = take_slot(aslot);
      if ((int)aE.size() >= MAXS) aE.erase(aE.begin());
      aE.push_back(tot);
      HIP_CHECK(hipMemcpyAsync(aF.p + s * nn, dF.p, nn * 8, hipMemcpyDeviceToDevice, st));
      HIP_CHECK(hipMemcpyAsync(aP.p + s * nn, dP.p, nn * 8, hipMemcpyDeviceToDevice, st));
      const int k = (int)aslot.size();
      if (k >= 2) {
        // G[s][t] = P_s . F_t over all slots, stored column-major (dG[t * MAXS + s]):
        // one transposed GEMV per F_t.  (As one MAXS x MAXS GEMM over the n^2
        // inner dimension rocBLAS used a single workgroup, 25 ms at n = 400.)
        const double one = 1.0, zero = 0.0;
        for (int t = 0; t < MAXS; ++t)
          RB_CHECK(rocblas_dgemv(bl.h, rocblas_operation_transpose, (rocblas_int)nn, MAXS, &one, aP.p, (rocblas_int)nn,
                                 aF.p + (size_t)t * nn, 1, &zero, dG.p + t * MAXS, 1));
        std::vector<double> G(MAXS * MAXS);
        download(G.data(), dG.p, MAXS * MAXS, st);
        auto g = [&](int i, int j) { return G[(size_t)aslot[j] * MAXS + aslot[i]]; };  // P_i . F_j
        const int last = k - 1;
        HMat Q((size_t)k * k);
        std::vector<double> b(k);
        for (int i = 0; i < k; ++i) {
          b[i] = g(i, last) - g(last, last);
          for (int j = 0; j < k; ++j) Q[(size_t)i * k + j] = g(i, j) - g(i, last) - g(last, j) + g(last, last);
        }
        auto c = simplex_minimise(Q, b);
        combine(aF, aslot, c, dFu.p);
        ok = true;
      }
    }
    diag_density(ok ? dFu.p : dF.p, dP.p, false);
  }
  if (!conv) fock(dP.p, dF.p);
  diag_density(dF.p, dP.p, true);
  fock(dP.p, dF.p);
  const double fe = energy(dP.p, dF.p);
  res->total_energy = fe + enuc;
  res->electronic_energy = fe;
  res->nuclear_repulsion = enuc;
  res->iterations = iters;
  res->converged = conv;
  res->n_occ = nocc;
  res->fock_builds = builds;
  res->seconds_fock = t_fock;
  if (p_out) download(p_out, dP.p, nn, st);
  if (f_out) download(f_out, dF.p, nn, st);
  if (eps_out) std::memcpy(eps_out, eps.data(), m * sizeof(double));
  HIP_CHECK(hipStreamSynchronize(st));
  res->seconds_total = std::chrono::duration<double>(clk::now() - t_start).count();
  res->seconds_linalg = res->seconds_total - res->seconds_setup - res->seconds_one_electron - res->seconds_guess -
                        res->seconds_fock;
  return conv ? HFG_OK : HFG_NOT_CONVERGED;
}

}  // namespace hfg
