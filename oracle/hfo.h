/* hfo.h -- C interface of the CPU oracle (test infrastructure only).
 *
 * The oracle is an FP64 CPU restatement of the hf-rs Fock-build path
 * (src/integrals.rs, src/shell_integrals.rs, src/scf.rs, src/df.rs,
 * src/diis.rs, src/sad.rs, src/matrix.rs, src/ao.rs, src/basis.rs).  It is
 * the checker that the HIP engine is compared with, and the CPU baseline that
 * bench.py times.  Nothing in the product library links or calls it.
 *
 * All matrices are row-major double arrays.  Functions return 0 on success
 * and a nonzero status otherwise; hfo_last_error() holds the message.
 */
#ifndef HFO_H
#define HFO_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct hfo_basis hfo_basis;

const char* hfo_last_error(void);

/* Boys function F_0..F_m(T) (src/integrals.rs:70-101). */
int hfo_boys(int m, double t, double* out);

/* Basis from Gaussian-94 text and atoms (src/basis.rs:296-496).
 * symbols: whitespace-separated element symbols; xyz: bohr, 3*natom. */
int hfo_basis_build(const char* g94, const char* symbols, const double* xyz, int natom,
                    hfo_basis** out);
void hfo_basis_free(hfo_basis* b);
int hfo_basis_nfunc(const hfo_basis* b);
int hfo_basis_nprim_total(const hfo_basis* b);
/* Export as the flat Cartesian-function arrays of the C ABI (one row per
 * ContractedGaussian, src/basis.rs:89-98). */
int hfo_basis_export(const hfo_basis* b, int32_t* atom_index, double* center, int32_t* lxyz,
                     int32_t* prim_offset, double* exponent, double* coefficient);
/* Build a basis directly from flat arrays (same layout as export). */
int hfo_basis_from_arrays(int nfunc, const int32_t* atom_index, const double* center,
                          const int32_t* lxyz, const int32_t* prim_offset,
                          const double* exponent, const double* coefficient, hfo_basis** out);
int hfo_nshells(const hfo_basis* b, int32_t* shell_start /* nshell+1 or NULL */);

/* One contracted Cartesian shell block [a][b][c][d] (src/shell_integrals.rs:138-151). */
int hfo_shell_quartet(const hfo_basis* b, int sa, int sb, int sc, int sd, double* out);
/* Scalar contracted ERI through the pair engine (src/integrals.rs:1062-1072). */
double hfo_contracted_eri(const hfo_basis* b, int i, int j, int k, int l);
/* Packed 8-fold screened tensor (src/shell_integrals.rs:156-240). */
int hfo_eri_packed(const hfo_basis* b, double* out, size_t len);
/* Schwarz bounds per shell pair a>=b, in (a,b) order (src/shell_integrals.rs:165-173). */
int hfo_schwarz(const hfo_basis* b, double* q);
/* contract_jk on a packed tensor (src/scf.rs:337-364); K for the given density. */
int hfo_contract_jk(const double* p, const double* eri, int n, double* j, double* k);
/* Direct J/K: the same screened shell quartets as the packed tensor, digested
 * without storing it.  sample_stride>1 keeps only bra pairs ij with
 * ij % sample_stride == sample_offset (CPU-baseline sampling). */
int hfo_direct_jk(const hfo_basis* b, const double* p, double* j, double* k, int sample_stride,
                  int sample_offset, double* seconds);

/* One-electron matrices S, T, V (src/scf.rs:247-271). */
int hfo_one_electron(const hfo_basis* b, const double* zxyz, int natom, double* s, double* t,
                     double* v);
/* Symmetric eigensolver with ascending values and the sign convention of
 * src/matrix.rs:150-190. */
int hfo_eigh(const double* a, int n, double* w, double* v);
/* Real spherical transform, Cartesian rows (src/ao.rs:166-217); returns columns. */
int hfo_spherical_transform(const hfo_basis* b, double* t, int* ncols);

typedef struct {
  int max_iter;
  double energy_tol, residual_tol, density_tol;
  int conv_algo;   /* 0 Adiissoscf(default) 1 Cdiis 2 Adiis 3 AdiisCdiis */
  int spherical;   /* BasisKind */
  int direct;      /* 1: direct J/K each iteration, 0: in-core packed tensor */
  int verbose;
} hfo_scf_options;

typedef struct {
  double total_energy, electronic_energy, nuclear_repulsion;
  int iterations, converged, n_occ, fock_builds;
  double seconds_eri, seconds_fock, seconds_total;
} hfo_scf_result;

/* RHF (src/scf.rs:410-627).  zxyz: per atom (Z, x, y, z).  Optional outputs
 * (may be NULL): density P, Fock F, orbital energies. */
int hfo_scf_rhf(const hfo_basis* b, const double* zxyz, int natom, int charge,
                const hfo_scf_options* opt, hfo_scf_result* res, double* p_out, double* f_out,
                double* eps_out);
/* SAD guess density (src/sad.rs:217-259). */
int hfo_sad_density(const hfo_basis* b, const double* zxyz, int natom, int charge,
                    const double* s, double* p_out);
int hfo_atomic_hf(const hfo_basis* b, double z, double* energy, double* density);

/* Density fitting (src/df.rs). */
int hfo_df_metric(const hfo_basis* aux, double* metric);
int hfo_df_three_center(const hfo_basis* b, const hfo_basis* aux, double* three /* naux x np */);
/* B tensor rank, then B (rank x np). */
int hfo_df_build(const hfo_basis* b, const hfo_basis* aux, double cutoff, int* rank,
                 double** b_out);
void hfo_free(void* p);
int hfo_df_jk(const double* bt, int rank, int n, const double* cocc, int nocc, double* j,
              double* k);
typedef struct {
  double total_energy;
  int iterations, converged, n_aux;
  double residual_norm;
} hfo_dfhf_result;
int hfo_df_rhf(const hfo_basis* b, const hfo_basis* aux, const double* zxyz, int natom,
               int charge, int max_iter, double e_tol, double r_tol, double d_tol, double cutoff,
               hfo_dfhf_result* res, double* p_out, double* f_out, double* j_out, double* k_out);

/* Pivoted Cholesky of the AO ERI supermatrix to residual diagonal threshold
 * (src/cholesky.rs:44-192): rank, largest remaining residual diagonal, the
 * pivots (packed pair indices, first max_pivots of them) and, if b_out is not
 * null, the factors (rank x np, packed; free with hfo_free). */
int hfo_cholesky_eri(const hfo_basis* b, double threshold, double budget, int* rank,
                     double* max_residual, int* pivots, int max_pivots, double** b_out);
/* CD-HF (src/hf_variants/cholesky_df.rs:43-95): Cholesky factors to delta in
 * the factorized RHF driver; res->n_aux is the number of Cholesky vectors. */
int hfo_cd_rhf(const hfo_basis* b, const double* zxyz, int natom, int charge, double delta,
               int max_iter, double tol, hfo_dfhf_result* res, double* p_out);

/* Analytic RHF nuclear gradient (natom x 3, Eh/bohr) for density p and
 * energy-weighted density w (n x n each): compute_gradient_from_density
 * (src/gradient.rs:287-512), without its screening of small products. */
int hfo_rhf_gradient(const hfo_basis* b, const double* zxyz, int natom, const double* p,
                     const double* w, double* grad);

/* Appendix-A style McMurchie-Davidson flop model of one primitive quartet of
 * angular momenta (la lb|lc ld), mirroring the loops of
 * src/shell_integrals.rs:101-133 and src/integrals.rs:70-101,259-346,905-971. */
double hfo_md_flops(int la, int lb, int lc, int ld);
/* Algorithmic flops of one direct Fock build: sum over surviving unique shell
 * quartets of Ka*Kb*Kc*Kd*flops(la,lb,lc,ld); also returns quartet counts. */
int hfo_fock_flops(const hfo_basis* b, double* flops, double* quartets, double* prim_quartets);
/* Same, restricted to bra pairs ij % stride == offset (the hfo_direct_jk sample). */
int hfo_fock_flops_sample(const hfo_basis* b, int stride, int offset, double* flops, double* quartets,
                          double* prim_quartets);

#ifdef __cplusplus
}
#endif
#endif
