// capi.cpp -- the extern "C" boundary declared in include/hfg.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <string>

#include "../../include/hfg.h"
#include "engine.h"
#include "orient.h"
#include "rhf.h"

using namespace hfg;

namespace {
thread_local std::string g_err;
int set_err(int code, const std::string& m) {
  g_err = m;
  return code;
}
}  // namespace

int hfg::set_error(int code, const std::string& m) { return set_err(code, m); }

struct hfg_ctx {
  std::unique_ptr<Engine> e;
};

#define GUARD(...)                                         \
  try {                                                    \
    __VA_ARGS__                                            \
  } catch (HfgError & ex) {                                \
    return set_err(ex.code, ex.msg);                       \
  } catch (std::bad_alloc&) {                              \
    return set_err(HFG_DEVICE, "host allocation failed");  \
  } catch (std::exception & ex) {                          \
    return set_err(HFG_INTEGRAL, ex.what());               \
  }

static int nc_of(const Engine& e, int pair) {
  const PairInfo& p = e.pairs[pair];
  return ((p.la + 1) * (p.la + 2) / 2) * ((p.lb + 1) * (p.lb + 2) / 2);
}

namespace hfg {
void eri_mo(Engine& e, const double* C, int nmo, double* out);
void jk_device_noalloc(Engine& e, const double* dP, double* dJ, double* dK);
void rhf_gradient(Engine& e, const std::vector<std::array<double, 4>>& atoms, const double* P, const double* W,
                  double* grad);
int scf_rhf(const hfg_basis_view* basis, const double* zxyz, int natom, int charge, const hfg_scf_options* opt,
            hfg_scf_result* res, double* p_out, double* f_out, double* eps_out, const double* p_guess);
}

// The packed screened tensor of compute_shell_eri_tensor
// (src/shell_integrals.rs:156-240): surviving quartets ij >= kl in batches,
// blocks from the GPU block kernels, scattered by eri_index (src/scf.rs:221-235).
void hfg::eri_packed_host(Engine& e, double* out) {
  {
    const size_t n = e.n, np = n * (n + 1) / 2;
    const size_t len = np * (np + 1) / 2;
    std::memset(out, 0, len * sizeof(double));
    const size_t npair = e.pairs.size();
    auto cidx = [](size_t i, size_t j) { return i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i; };
    // batches of surviving quartets ij >= kl (src/shell_integrals.rs:225-235)
    size_t ij = 0, kl = 0;
    const size_t budget = 1 << 24;  // doubles per batch
    while (ij < npair) {
      std::vector<BlockTask> tasks;
      std::vector<int> cls;
      std::vector<std::pair<int, int>> which;
      long long off = 0;
      while (ij < npair && (size_t)off < budget) {
        if (e.Q[ij] * e.Q[kl] > e.screen) {
          const PairInfo& pb = e.pairs[ij];
          const PairInfo& pk = e.pairs[kl];
          int pcb = pair_class(pb.la, pb.lb), pck = pair_class(pk.la, pk.lb);
          int b = (int)ij, k = (int)kl;
          if (pck > pcb) {
            std::swap(b, k);
            std::swap(pcb, pck);
          }
          tasks.push_back({b, k, off});
          cls.push_back(quartet_class(pcb, pck));
          which.push_back({b, k});
          off += (long long)nc_of(e, b) * nc_of(e, k);
        }
        if (++kl > ij) {
          ++ij;
          kl = 0;
        }
      }
      // group by class for the launches
      std::vector<size_t> ord(tasks.size());
      for (size_t i = 0; i < ord.size(); ++i) ord[i] = i;
      std::stable_sort(ord.begin(), ord.end(), [&](size_t x, size_t y) { return cls[x] < cls[y]; });
      std::vector<BlockTask> t2;
      std::vector<int> c2;
      for (size_t i : ord) {
        t2.push_back(tasks[i]);
        c2.push_back(cls[i]);
      }
      std::vector<double> blk(off);
      e.blocks(t2, c2, blk);
      for (const BlockTask& t : t2) {
        const PairInfo& pb = e.pairs[t.bra];
        const PairInfo& pk = e.pairs[t.ket];
        const int na = (int)e.shells[pb.sh_a].scale.size(), nb = (int)e.shells[pb.sh_b].scale.size();
        const int nc_ = (int)e.shells[pk.sh_a].scale.size(), nd = (int)e.shells[pk.sh_b].scale.size();
        const double* v = blk.data() + t.off;
        for (int i = 0; i < na; ++i)
          for (int j = 0; j < nb; ++j) {
            const size_t mu = pb.ao_a + i, nu = pb.ao_b + j;
            const double fab = e.ao_scale[mu] * e.ao_scale[nu];
            for (int k = 0; k < nc_; ++k)
              for (int l = 0; l < nd; ++l) {
                const size_t la = pk.ao_a + k, si = pk.ao_b + l;
                out[cidx(cidx(mu, nu), cidx(la, si))] =
                    fab * e.ao_scale[la] * e.ao_scale[si] * v[((i * nb + j) * nc_ + k) * nd + l];
              }
          }
      }
    }
  }
}


extern "C" {

const char* hfg_last_error(void) { return g_err.c_str(); }

int hfg_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int hfg_ctx_create(const hfg_basis_view* basis, hfg_ctx** out) {
  if (!basis || !out) return set_err(HFG_INVALID_INPUT, "null argument");
  GUARD({
    auto* c = new hfg_ctx;
    try {
      c->e.reset(new Engine(*basis));
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
    return HFG_OK;
  })
}

void hfg_ctx_destroy(hfg_ctx* ctx) { delete ctx; }

int hfg_ctx_info(const hfg_ctx* ctx, int64_t* nf, int64_t* ns, int64_t* np, int64_t* nq) {
  if (!ctx) return set_err(HFG_INVALID_INPUT, "null context");
  if (nf) *nf = ctx->e->n;
  if (ns) *ns = (int64_t)ctx->e->shells.size();
  if (np) *np = (int64_t)ctx->e->pairs.size();
  if (nq) *nq = ctx->e->n_quartets;
  return HFG_OK;
}

int hfg_schwarz(hfg_ctx* ctx, double* q) {
  if (!ctx || !q) return set_err(HFG_INVALID_INPUT, "null argument");
  std::memcpy(q, ctx->e->Q.data(), ctx->e->Q.size() * sizeof(double));
  return HFG_OK;
}

int hfg_shell_quartet(hfg_ctx* ctx, int sa, int sb, int sc, int sd, double* out) {
  if (!ctx || !out) return set_err(HFG_INVALID_INPUT, "null argument");
  GUARD({
    Engine& e = *ctx->e;
    const int ns = (int)e.shells.size();
    for (int s : {sa, sb, sc, sd})
      if (s < 0 || s >= ns) return set_err(HFG_INVALID_INPUT, "one shell per quartet slot is required");
    Orientation o = orient(e.shells, sa, sb, sc, sd);
    auto pid = [](int a, int b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; };
    // kernel-orientation pairs must match the stored orientation of each pair
    int bra = pid(o.k[0], o.k[1]), ket = pid(o.k[2], o.k[3]);
    const PairInfo& pb = e.pairs[bra];
    const PairInfo& pk = e.pairs[ket];
    // stored pair orientation puts the higher-L shell first; equal L keeps
    // hf-rs order (a >= b).  Re-derive the slot mapping from it.
    Orientation o2 = o;
    o2.k[0] = pb.sh_a;
    o2.k[1] = pb.sh_b;
    o2.k[2] = pk.sh_a;
    o2.k[3] = pk.sh_b;
    if (o2.k[0] != o.k[0]) std::swap(o2.pos[0], o2.pos[1]);
    if (o2.k[2] != o.k[2]) std::swap(o2.pos[2], o2.pos[3]);
    size_t nt = 1;
    for (int j = 0; j < 4; ++j) nt *= e.shells[o2.k[j]].scale.size();
    std::vector<double> blk(nt);
    e.blocks({BlockTask{bra, ket, 0}}, {o.cls}, blk);
    unorient(e.shells, o2, blk.data(), out);
    return HFG_OK;
  })
}

int hfg_eri_packed(hfg_ctx* ctx, double* out, size_t len) {
  if (!ctx || !out) return set_err(HFG_INVALID_INPUT, "null argument");
  GUARD({
    const size_t n = ctx->e->n, np = n * (n + 1) / 2;
    if (len != np * (np + 1) / 2) return set_err(HFG_INVALID_INPUT, "packed length must be np(np+1)/2");
    eri_packed_host(*ctx->e, out);
    return HFG_OK;
  })
}

int hfg_eri_mo(hfg_ctx* ctx, const double* c, int n_mo, double* out, size_t len) {
  if (!ctx || !c || !out) return set_err(HFG_INVALID_INPUT, "null argument");
  const size_t mp = (size_t)n_mo * (n_mo + 1) / 2;
  if (n_mo <= 0 || len != mp * mp) return set_err(HFG_INVALID_INPUT, "output length must be (n_mo(n_mo+1)/2)^2");
  GUARD({
    eri_mo(*ctx->e, c, n_mo, out);
    return HFG_OK;
  })
}

int hfg_contract_jk_device(hfg_ctx* ctx, const double* d_p, double* d_j, double* d_k) {
  if (!ctx || !d_p) return set_err(HFG_INVALID_INPUT, "null argument");
  GUARD({
    ctx->e->jk_device(d_p, d_j, d_k);
    ctx->e->sync();
    return HFG_OK;
  })
}

int hfg_contract_jk_device_noalloc(hfg_ctx* ctx, const double* d_p, double* d_j, double* d_k) {
  if (!ctx || !d_p) return set_err(HFG_INVALID_INPUT, "null argument");
  GUARD({
    jk_device_noalloc(*ctx->e, d_p, d_j, d_k);
    ctx->e->sync();
    return HFG_OK;
  })
}

int hfg_rhf_gradient(hfg_ctx* ctx, const double* zxyz, int natom, const double* p, const double* w, double* grad) {
  if (!ctx || !zxyz || !p || !w || !grad) return set_err(HFG_INVALID_INPUT, "null argument");
  if (natom <= 0) return set_err(HFG_INVALID_INPUT, "gradient needs at least one atom");
  GUARD({
    std::vector<std::array<double, 4>> atoms(natom);
    for (int a = 0; a < natom; ++a)
      for (int k = 0; k < 4; ++k) {
        atoms[a][k] = zxyz[4 * a + k];
        if (!std::isfinite(atoms[a][k])) return set_err(HFG_INVALID_INPUT, "non-finite geometry");
      }
    rhf_gradient(*ctx->e, atoms, p, w, grad);
    return HFG_OK;
  })
}

// Nuclei (Z, x, y, z) of a call, checked finite.
static std::vector<std::array<double, 4>> nuclei(const double* zxyz, int natom) {
  std::vector<std::array<double, 4>> atoms(natom);
  for (int a = 0; a < natom; ++a)
    for (int k = 0; k < 4; ++k) {
      atoms[a][k] = zxyz[4 * a + k];
      if (!std::isfinite(atoms[a][k])) throw HfgError{HFG_INVALID_INPUT, "non-finite geometry"};
    }
  return atoms;
}

int hfg_one_electron(hfg_ctx* ctx, const double* zxyz, int natom, double* s, double* t, double* v) {
  if (!ctx || !zxyz || !s || !t || !v) return set_err(HFG_INVALID_INPUT, "null argument");
  if (natom <= 0) return set_err(HFG_INVALID_INPUT, "one-electron matrices need at least one nucleus");
  GUARD({
    Engine& e = *ctx->e;
    const auto atoms = nuclei(zxyz, natom);
    const size_t nn = (size_t)e.n * e.n;
    DevBuf d;
    d.alloc(3 * nn * sizeof(double));
    double* dS = d.as<double>();
    one_electron_device(e.shells, e.n, atoms, dS, dS + nn, dS + 2 * nn, nullptr, e.stream);
    HIP_CHECK(hipMemcpyAsync(s, dS, nn * 8, hipMemcpyDeviceToHost, e.stream));
    HIP_CHECK(hipMemcpyAsync(t, dS + nn, nn * 8, hipMemcpyDeviceToHost, e.stream));
    HIP_CHECK(hipMemcpyAsync(v, dS + 2 * nn, nn * 8, hipMemcpyDeviceToHost, e.stream));
    e.sync();
    return HFG_OK;
  })
}

int hfg_sad_density(hfg_ctx* ctx, const double* zxyz, int natom, int charge, double* p) {
  if (!ctx || !zxyz || !p) return set_err(HFG_INVALID_INPUT, "null argument");
  if (natom <= 0) return set_err(HFG_INVALID_INPUT, "SAD needs at least one atom");
  GUARD({
    Engine& e = *ctx->e;
    const auto atoms = nuclei(zxyz, natom);
    const size_t nn = (size_t)e.n * e.n;
    DevBuf d;
    d.alloc(2 * nn * sizeof(double));
    double* dS = d.as<double>();
    one_electron_device(e.shells, e.n, atoms, dS, nullptr, nullptr, nullptr, e.stream);
    Blas bl(e.stream);
    sad_guess_device(e.shells, e.n, atoms, charge, dS, dS + nn, bl, e.stream);
    HIP_CHECK(hipMemcpyAsync(p, dS + nn, nn * 8, hipMemcpyDeviceToHost, e.stream));
    e.sync();
    return HFG_OK;
  })
}

int hfg_contract_jk(hfg_ctx* ctx, const double* p, double* j, double* k) {
  if (!ctx || !p || !j || !k) return set_err(HFG_INVALID_INPUT, "null argument");
  GUARD({
    Engine& e = *ctx->e;
    const size_t nn = (size_t)e.n * e.n;
    for (size_t x = 0; x < nn; ++x)
      if (!std::isfinite(p[x])) return set_err(HFG_INVALID_INPUT, "non-finite density");
    HIP_CHECK(hipMemcpyAsync(e.d_D.p, p, nn * sizeof(double), hipMemcpyHostToDevice, e.stream));
    e.jk_device(e.d_D.as<double>(), e.d_J.as<double>(), e.d_K.as<double>());
    HIP_CHECK(hipMemcpyAsync(j, e.d_J.p, nn * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    HIP_CHECK(hipMemcpyAsync(k, e.d_K.p, nn * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    e.sync();
    return HFG_OK;
  })
}

int hfg_contract_jk_multi(hfg_ctx* ctx, const double* p_total, int nspin, const double* const* p_spin, double* j,
                          double* const* k_spin) {
  if (!ctx || !p_total || !j || nspin < 0 || (nspin > 0 && (!p_spin || !k_spin)))
    return set_err(HFG_INVALID_INPUT, "null argument or negative spin count");
  for (int s = 0; s < nspin; ++s)
    if (!p_spin[s] || !k_spin[s]) return set_err(HFG_INVALID_INPUT, "null spin density or exchange output");
  GUARD({
    Engine& e = *ctx->e;
    const size_t nn = (size_t)e.n * e.n;
    auto finite = [&](const double* m) {
      for (size_t x = 0; x < nn; ++x)
        if (!std::isfinite(m[x])) return false;
      return true;
    };
    if (!finite(p_total)) return set_err(HFG_INVALID_INPUT, "non-finite density");
    bool total_k = false;  // a spin density given as the total itself (RHF) shares the total's build
    for (int s = 0; s < nspin; ++s) {
      if (p_spin[s] == p_total) total_k = true;
      else if (!finite(p_spin[s])) return set_err(HFG_INVALID_INPUT, "non-finite spin density");
    }
    HIP_CHECK(hipMemcpyAsync(e.d_D.p, p_total, nn * sizeof(double), hipMemcpyHostToDevice, e.stream));
    e.jk_device(e.d_D.as<double>(), e.d_J.as<double>(), total_k ? e.d_K.as<double>() : nullptr);
    HIP_CHECK(hipMemcpyAsync(j, e.d_J.p, nn * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    for (int s = 0; s < nspin; ++s)
      if (p_spin[s] == p_total) HIP_CHECK(hipMemcpyAsync(k_spin[s], e.d_K.p, nn * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    e.sync();
    for (int s = 0; s < nspin; ++s) {
      if (p_spin[s] == p_total) continue;
      HIP_CHECK(hipMemcpyAsync(e.d_D.p, p_spin[s], nn * sizeof(double), hipMemcpyHostToDevice, e.stream));
      e.jk_device(e.d_D.as<double>(), nullptr, e.d_K.as<double>());
      HIP_CHECK(hipMemcpyAsync(k_spin[s], e.d_K.p, nn * sizeof(double), hipMemcpyDeviceToHost, e.stream));
      e.sync();
    }
    return HFG_OK;
  })
}

int hfg_build_fock(hfg_ctx* ctx, const double* h, const double* p, double* f) {
  if (!ctx || !h || !p || !f) return set_err(HFG_INVALID_INPUT, "null argument");
  GUARD({
    Engine& e = *ctx->e;
    const size_t nn = (size_t)e.n * e.n;
    HIP_CHECK(hipMemcpyAsync(e.d_D.p, p, nn * sizeof(double), hipMemcpyHostToDevice, e.stream));
    HIP_CHECK(hipMemcpyAsync(e.d_H.p, h, nn * sizeof(double), hipMemcpyHostToDevice, e.stream));
    e.jk_device(e.d_D.as<double>(), nullptr, nullptr, e.d_H.as<double>(), e.d_F.as<double>());
    HIP_CHECK(hipMemcpyAsync(f, e.d_F.p, nn * sizeof(double), hipMemcpyDeviceToHost, e.stream));
    e.sync();
    return HFG_OK;
  })
}

int hfg_set_density(hfg_ctx* ctx, const double* p) {
  if (!ctx || !p) return set_err(HFG_INVALID_INPUT, "null argument");
  GUARD({
    Engine& e = *ctx->e;
    HIP_CHECK(hipMemcpyAsync(e.d_D.p, p, (size_t)e.n * e.n * sizeof(double), hipMemcpyHostToDevice, e.stream));
    e.sync();
    return HFG_OK;
  })
}

int hfg_jk_resident(hfg_ctx* ctx) {
  if (!ctx) return set_err(HFG_INVALID_INPUT, "null context");
  GUARD({
    Engine& e = *ctx->e;
    e.jk_device(e.d_D.as<double>(), e.d_J.as<double>(), e.d_K.as<double>());
    e.sync();
    return HFG_OK;
  })
}

int hfg_get_jk(hfg_ctx* ctx, double* j, double* k) {
  if (!ctx) return set_err(HFG_INVALID_INPUT, "null context");
  GUARD({
    Engine& e = *ctx->e;
    const size_t nn = (size_t)e.n * e.n * sizeof(double);
    if (j) HIP_CHECK(hipMemcpyAsync(j, e.d_J.p, nn, hipMemcpyDeviceToHost, e.stream));
    if (k) HIP_CHECK(hipMemcpyAsync(k, e.d_K.p, nn, hipMemcpyDeviceToHost, e.stream));
    e.sync();
    return HFG_OK;
  })
}

int hfg_profile(hfg_ctx* ctx, int enable) {
  if (!ctx) return set_err(HFG_INVALID_INPUT, "null context");
  GUARD({
    ctx->e->set_profile(enable != 0);
    return HFG_OK;
  })
}

int hfg_graph_launches(const hfg_ctx* ctx, int64_t* launches) {
  if (!ctx || !launches) return set_err(HFG_INVALID_INPUT, "null argument");
  *launches = ctx->e->graph_launches;
  return HFG_OK;
}

int hfg_class_stats(hfg_ctx* ctx, int cls, int32_t* l4, double* flops, int64_t* quartets, double* ms,
                    int64_t* launches) {
  if (!ctx || cls < 0 || cls >= HFG_NCLASS) return set_err(HFG_INVALID_INPUT, "class index out of range");
  Engine& e = *ctx->e;
  int l[4];
  class_angular(cls, l);
  if (l4)
    for (int k = 0; k < 4; ++k) l4[k] = l[k];
  if (flops) *flops = e.class_flops[cls];
  if (quartets) *quartets = e.class_quartets[cls];
  if (ms) *ms = e.class_ms[cls];
  if (launches) *launches = e.class_launches[cls];
  return HFG_OK;
}

int hfg_fock_flops(hfg_ctx* ctx, double* flops, double* prim_quartets) {
  if (!ctx) return set_err(HFG_INVALID_INPUT, "null context");
  double f = 0;
  for (double x : ctx->e->class_flops) f += x;
  if (flops) *flops = f;
  if (prim_quartets) *prim_quartets = ctx->e->prim_quartets;
  return HFG_OK;
}

double hfg_md_flops(int la, int lb, int lc, int ld) { return md_flops(la, lb, lc, ld); }

int hfg_persistent_blocks(void) { return persistent_blocks(); }

int hfg_calibration_copy(size_t bytes, double* ms) {
  if (!ms || bytes < 64) return set_err(HFG_INVALID_INPUT, "invalid calibration size");
  GUARD({
    *ms = calibration_copy(bytes);
    return HFG_OK;
  })
}

int hfg_scf_rhf(const hfg_basis_view* basis, const double* zxyz, int natom, int charge, const hfg_scf_options* opt,
                hfg_scf_result* res, double* p_out, double* f_out, double* eps_out) {
  if (!basis || !zxyz || !opt || !res) return set_err(HFG_INVALID_INPUT, "null argument");
  GUARD({ return scf_rhf(basis, zxyz, natom, charge, opt, res, p_out, f_out, eps_out, nullptr); })
}

int hfg_scf_rhf_guess(const hfg_basis_view* basis, const double* zxyz, int natom, int charge,
                      const hfg_scf_options* opt, const double* p_guess, hfg_scf_result* res, double* p_out,
                      double* f_out, double* eps_out) {
  if (!basis || !zxyz || !opt || !res || !p_guess) return set_err(HFG_INVALID_INPUT, "null argument");
  const size_t nn = (size_t)basis->n_functions * basis->n_functions;
  for (size_t x = 0; x < nn; ++x)
    if (!std::isfinite(p_guess[x])) return set_err(HFG_INVALID_INPUT, "non-finite initial density");
  GUARD({ return scf_rhf(basis, zxyz, natom, charge, opt, res, p_out, f_out, eps_out, p_guess); })
}

}  // extern "C"
