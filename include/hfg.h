/* hfg.h -- C ABI of the MI355X Fock-build engine for hf-rs.
 *
 * Plain C types only: caller-owned, row-major FP64 buffers, int status codes
 * (0 = success) and a thread-local message in hfg_last_error().  One context
 * holds one basis on one GPU, with its own HIP stream; calls on one context
 * are not reentrant, calls on different contexts are.
 *
 * Each entry point names the hf-rs function (Rust, src/...) whose body it
 * replaces when hf-rs is built against this library.
 */
#ifndef HFG_H
#define HFG_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, mirroring hf-rs's HfError variants (src/error.rs:5-15). */
enum {
  HFG_OK = 0,
  HFG_INVALID_INPUT = 1, /* HfError::InvalidInput */
  HFG_INTEGRAL = 2,      /* HfError::Integral (numerical failure) */
  HFG_DEVICE = 3,        /* HIP runtime / device failure */
  HFG_NOT_CONVERGED = 4, /* SCF did not converge (src/scf.rs:600-604) */
  HFG_UNSUPPORTED = 5    /* angular momentum beyond the compiled classes */
};

/* A basis as hf-rs holds it: one row per Cartesian ContractedGaussian
 * (src/basis.rs:89-98), shells contiguous in hf-rs component order
 * (src/basis.rs:170-188, 430-436), coefficients already normalised. */
typedef struct {
  int32_t n_functions;
  const int32_t* atom_index;  /* [n] */
  const double* center;       /* [3n] bohr */
  const int32_t* lxyz;        /* [3n] */
  const int32_t* prim_offset; /* [n+1] */
  const double* exponent;     /* [prim_offset[n]] */
  const double* coefficient;  /* [prim_offset[n]] */
} hfg_basis_view;

typedef struct hfg_ctx hfg_ctx;

/* Message of the last failing call on this thread. */
const char* hfg_last_error(void);
/* Number of visible GPUs (0 when none); never allocates device memory. */
int hfg_device_count(void);

/* Build the per-molecule state: shells, primitive-pair data in HBM, Schwarz
 * bounds and the screened shell-quartet work lists.  Replaces the shell-pair
 * and bound set-up of compute_shell_eri_tensor (src/shell_integrals.rs:156-173);
 * the shell grouping and its errors follow shell_ranges (src/ao.rs:133-164). */
int hfg_ctx_create(const hfg_basis_view* basis, hfg_ctx** out);
void hfg_ctx_destroy(hfg_ctx* ctx);

/* Sizes: functions, shells, shell pairs, surviving unique shell quartets. */
int hfg_ctx_info(const hfg_ctx* ctx, int64_t* n_functions, int64_t* n_shells, int64_t* n_pairs,
                 int64_t* n_quartets);
/* Schwarz bounds Q_ab = sqrt(max (ab|ab)) in hf-rs pair order a>=b
 * (src/shell_integrals.rs:165-173); length n_shells*(n_shells+1)/2. */
int hfg_schwarz(hfg_ctx* ctx, double* q);

/* One contracted Cartesian block [a][b][c][d] of shells (sa,sb,sc,sd).
 * Replaces compute_shell_quartet (src/shell_integrals.rs:138-151). */
int hfg_shell_quartet(hfg_ctx* ctx, int sa, int sb, int sc, int sd, double* out);

/* Packed 8-fold screened tensor of length np(np+1)/2, np = n(n+1)/2, indexed
 * by eri_index (src/scf.rs:221-235).  Replaces compute_eri_tensor
 * (src/scf.rs:275-278) for callers that keep HFResult.eri. */
int hfg_eri_packed(hfg_ctx* ctx, double* out, size_t len);

/* Two-electron integrals in the MO basis of the n_mo columns of C (n x n_mo,
 * row-major): (pq|rs) at out[rs * np + pq], np = n_mo(n_mo+1)/2 and pq, rs
 * compound indices p >= q (len = np * np).  The AO tensor is the screened
 * packed one (hfg_eri_packed); both half transformations run on the device.
 * n <= 200.  Replaces transform_eri_ao_to_mo (src/mo_integrals.rs:156-265),
 * the reader of HFResult.eri (src/mo_integrals.rs:505). */
int hfg_eri_mo(hfg_ctx* ctx, const double* c, int n_mo, double* out, size_t len);

/* Direct J and K for a symmetric total density P (n x n), over the unique
 * shell quartets with Q_ab*Q_cd > 1e-12.  Replaces contract_jk
 * (src/scf.rs:337-364) with the packed tensor never formed. */
int hfg_contract_jk(hfg_ctx* ctx, const double* p, double* j, double* k);
/* J of the total density and K of each spin density, as
 * contract_jk(total, spins, eri, n) -> (J, Vec<K>) (src/scf.rs:337-364):
 * RHF passes spins = {P} (the same pointer as p_total: one build), UHF
 * {P_alpha, P_beta} (one build per distinct density).  All n x n. */
int hfg_contract_jk_multi(hfg_ctx* ctx, const double* p_total, int nspin, const double* const* p_spin, double* j,
                          double* const* k_spin);
/* F = H + J - K/2.  Replaces build_fock_matrix (src/scf.rs:366-369). */
int hfg_build_fock(hfg_ctx* ctx, const double* h, const double* p, double* f);
/* Same as hfg_contract_jk on device pointers (HBM-resident P, J, K), on the
 * context's stream; returns after the results are complete. */
int hfg_contract_jk_device(hfg_ctx* ctx, const double* d_p, double* d_j, double* d_k);

/* The same as hfg_contract_jk_device with no device allocation in the call:
 * the build's scratch memory (the molecule's table, the job counters and the
 * per-XCC J/K accumulators) lives in one file-scope device array of the
 * library, about 16.8 MB per GPU, kept until the process ends.  For callers
 * that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc); no counterpart in the reference.  Limit:
 * n <= 256 basis functions (HFG_INVALID_INPUT beyond).  Every call on one GPU
 * shares the array: calls that overlap on one GPU race. */
int hfg_contract_jk_device_noalloc(hfg_ctx* ctx, const double* d_p, double* d_j, double* d_k);

/* S, T and V (n x n each, row-major, Cartesian, the basis's component
 * normalisation) of the context's basis for natom nuclei (Z, x, y, z; bohr),
 * computed on the device.  Replaces compute_one_electron_integrals
 * (src/scf.rs:246-271; H = T + V).  Shells up to d, as the context. */
int hfg_one_electron(hfg_ctx* ctx, const double* zxyz, int natom, double* s, double* t, double* v);
/* The SAD initial density (n x n) of the context's molecule: the cached
 * fractional-occupation atomic densities, superposed and scaled to the
 * electron count with the device overlap.  Replaces compute_sad_density
 * (src/sad.rs:217-259).  HFG_INVALID_INPUT when the basis names an atom
 * beyond natom or the electron count is not positive. */
int hfg_sad_density(hfg_ctx* ctx, const double* zxyz, int natom, int charge, double* p);

/* Analytic RHF nuclear gradient, compute_gradient_from_density
 * (src/gradient.rs:287-512): grad (natom x 3, Eh/bohr) of the context's
 * molecule for density p and energy-weighted density w (n x n each; for a
 * converged RHF w = 2 C_occ eps C_occ^T = 1/2 p F p); zxyz holds natom rows of
 * (Z, x, y, z) in bohr, atom indices as in the basis.  The two-electron
 * derivative integrals and their contraction run on the device; the
 * one-electron derivative terms and the nuclear repulsion on the host. */
int hfg_rhf_gradient(hfg_ctx* ctx, const double* zxyz, int natom, const double* p, const double* w,
                     double* grad);

/* Resident-buffer form used by throughput measurements: the density is
 * uploaded once, J and K stay in HBM until fetched.  (Same arithmetic as
 * hfg_contract_jk; no reference counterpart: hf-rs has no device memory.) */
int hfg_set_density(hfg_ctx* ctx, const double* p);
int hfg_jk_resident(hfg_ctx* ctx);
int hfg_get_jk(hfg_ctx* ctx, double* j, double* k);

/* Instrumentation (no reference counterpart; hf-rs has no profiler hooks,
 * src/scf.rs:486-604 only prints an iteration table).  hfg_profile(1)
 * brackets every shell-quartet class launch with HIP events on the context's
 * stream and accumulates their durations; hfg_class_stats reports, per class
 * index 0..20, the angular momenta (la lb lc ld, kernel orientation), the
 * algorithmic FLOP count of one launch (McMurchie-Davidson model: sum over the
 * class's surviving quartets of Ka*Kb*Kc*Kd*flops(la,lb,lc,ld)), quartets per
 * launch, total measured milliseconds and number of launches. */
int hfg_profile(hfg_ctx* ctx, int enable);
/* Fock builds of the context that replayed a captured HIP graph of their
 * launches (no reference counterpart; HFG_GRAPH=0 at context creation turns
 * the capture off). */
int hfg_graph_launches(const hfg_ctx* ctx, int64_t* launches);
int hfg_class_stats(hfg_ctx* ctx, int cls, int32_t* l4, double* flops_per_launch, int64_t* quartets,
                    double* ms_total, int64_t* launches);
/* Algorithmic FLOPs of one Fock build (sum over classes) and per primitive
 * quartet of class (la lb|lc ld): 20 + 8 + 16(L+1) + 2(L+1) + 3(C(L+4,4)-(L+1))
 * + sum over Cartesian component quartets of 2 sum_ax (lab+1)(lcd+1)
 * + 3 prod_ax (lab+lcd+1) + 3, the loop counts of src/shell_integrals.rs:101-133
 * and src/integrals.rs:70-101, 259-346, 905-971. */
int hfg_fock_flops(hfg_ctx* ctx, double* flops, double* prim_quartets);
double hfg_md_flops(int la, int lb, int lc, int ld);
/* Calibration for hardware-counter traffic: copies `bytes` bytes between two
 * fresh device buffers with a plain streaming kernel (k_copy_calibration) and
 * returns its time in milliseconds.  Instrumentation, no reference counterpart. */
int hfg_calibration_copy(size_t bytes, double* ms);
/* Workgroups of one persistent class launch (4 per compute unit by default,
 * HFG_PERSISTENT_BLOCKS overrides); instrumentation, no reference counterpart. */
int hfg_persistent_blocks(void);

/* RHF options and result (ScfOptions, src/api.rs:323-410; HFResult,
 * src/scf.rs:14-33). */
typedef struct {
  int32_t max_iter;
  double energy_tol, residual_tol, density_tol;
  int32_t conv_algo; /* 0 ADIIS->CDIIS (default), 1 CDIIS, 2 ADIIS, 3 ADIIS->CDIIS at 0.08 */
  int32_t spherical; /* BasisKind::Spherical for the orthogonaliser */
  int32_t verbose;
  /* 1: inside the iterations, build J and K of the density change dP only
   * (F = H + G_prev + G(dP)) with the Schwarz threshold divided by max|dP|
   * when that is below 1; the first build and the final rebuild are full.
   * 0: every build full, as the reference. */
  int32_t incremental_fock;
} hfg_scf_options;

typedef struct {
  double total_energy, electronic_energy, nuclear_repulsion;
  int32_t iterations, converged, n_occ, fock_builds;
  double seconds_total, seconds_fock;
  /* phase breakdown: engine set-up (pairs, Schwarz bounds, work lists),
   * one-electron matrices, initial guess, eigensolver + DIIS */
  double seconds_setup, seconds_one_electron, seconds_guess, seconds_linalg;
} hfg_scf_result;

/* Restricted Hartree-Fock with device-resident Fock builds.  zxyz: per atom
 * (nuclear charge, x, y, z) in bohr.  Optional outputs P, F (n x n) and
 * orbital energies may be NULL.  Replaces run_scf_in_space_with_convergence
 * (src/scf.rs:410-627). */
int hfg_scf_rhf(const hfg_basis_view* basis, const double* zxyz, int natom, int charge,
                const hfg_scf_options* opt, hfg_scf_result* res, double* p_out, double* f_out,
                double* eps_out);

/* The same from a given initial density P (n x n) instead of SAD.  Replaces
 * run_scf_with_initial_density (src/scf.rs:630-639). */
int hfg_scf_rhf_guess(const hfg_basis_view* basis, const double* zxyz, int natom, int charge,
                      const hfg_scf_options* opt, const double* p_guess, hfg_scf_result* res, double* p_out,
                      double* f_out, double* eps_out);

/* Density fitting with a Coulomb metric.  hfg_df_build computes the metric
 * (P|Q), its canonical orthogonaliser X (eigenvalues below metric_cutoff
 * dropped), the 3-index integrals (P|mu nu) with the unit function of
 * src/df.rs:30-40 and B = X^T (P|mu nu), kept in HBM.  Replaces
 * DfTensor::build (src/df.rs:89-127; Cartesian auxiliary space). */
typedef struct hfg_df hfg_df;
int hfg_df_build(const hfg_basis_view* basis, const hfg_basis_view* aux, double metric_cutoff,
                 hfg_df** out);
void hfg_df_destroy(hfg_df* df);
int hfg_df_info(const hfg_df* df, int64_t* n, int64_t* n_aux, int64_t* rank);
/* J and K of the RHF density P = 2 C_occ C_occ^T from occupied coefficients
 * (n x nocc).  Replaces DfTensor::coulomb_exchange_rhf (src/df.rs:133-174). */
int hfg_df_jk(hfg_df* df, const double* cocc, int nocc, double* j, double* k);
/* The same with the precision of the exchange contraction chosen: 0 FP64 (as
 * hfg_df_jk), 1 FP32 operands with FP32 accumulation, 2 each operand split
 * into two BF16 parts with three BF16 products and FP32 accumulation, both
 * on the matrix cores; J stays FP64.  device_ms (may be NULL): device time
 * of the J/K contraction.  No reference counterpart (the reference's DF-K is
 * FP64, src/df.rs:143-173): its error against precision 0 is the caller's to
 * report. */
int hfg_df_jk_ex(hfg_df* df, const double* cocc, int nocc, double* j, double* k, int precision,
                 double* device_ms);

/* DFHFOptions (src/hf_variants/ri_df_hf.rs:21-45) and DFHFResult (:47-64),
 * with two fields that have no reference counterpart: k_precision selects
 * the exchange contraction of the early iterations (0 FP64 throughout, as
 * the reference; 1 FP32; 2 split BF16, as in hfg_df_jk_ex), used until the
 * projected residual falls below switch_residual; from then on every build
 * is FP64, the DIIS history restarts, and convergence is only accepted on
 * FP64 builds (at least two of them). */
typedef struct {
  int32_t max_iter;
  double energy_tol, residual_tol, density_tol, metric_cutoff;
  int32_t k_precision;
  double switch_residual;
} hfg_dfhf_options;
typedef struct {
  double total_energy;
  int32_t iterations, converged, n_aux, fock_builds;
  double seconds_total, seconds_fock;
} hfg_dfhf_result;
/* DF-RHF.  Replaces run_df_hf_with_aux (src/hf_variants/ri_df_hf.rs:122-155)
 * and run_factorized_rhf (:160-263).  Optional outputs may be NULL. */
int hfg_df_rhf(const hfg_basis_view* basis, const hfg_basis_view* aux, const double* zxyz, int natom,
               int charge, const hfg_dfhf_options* opt, hfg_dfhf_result* res, double* p_out,
               double* f_out, double* j_out, double* k_out);

/* Batches: J and K of many molecules in fused launches (one launch per
 * shell-quartet class for the whole batch).  Each molecule keeps its own
 * engine (as hfg_ctx_create); the results per molecule equal hfg_contract_jk's.
 * No reference counterpart: the reference builds each molecule's Fock matrix
 * on its own (src/scf.rs:366-369). */
typedef struct hfg_batch hfg_batch;
int hfg_batch_create(const hfg_basis_view* bases, int count, hfg_batch** out);
void hfg_batch_destroy(hfg_batch* batch);
/* count: molecules; n_max: largest basis; quartets: surviving shell quartets
 * over the batch; flops: algorithmic FLOPs of one build of the batch. */
int hfg_batch_info(const hfg_batch* batch, int64_t* count, int64_t* n_max, int64_t* quartets, double* flops);
/* Density of molecule mol (n x n row-major), kept in HBM. */
int hfg_batch_set_density(hfg_batch* batch, int mol, const double* p);
/* J and K of every molecule's resident density, into HBM. */
int hfg_batch_jk(hfg_batch* batch);
/* The same J and K, molecule after molecule with each one's own launches
 * (the unbatched path, for comparison). */
int hfg_batch_jk_each(hfg_batch* batch);
/* Copies molecule mol's J and K (either may be NULL) to the host. */
int hfg_batch_get_jk(hfg_batch* batch, int mol, double* j, double* k);
/* Instrumentation of the fused builds, as hfg_profile / hfg_class_stats for
 * one molecule: per class, the algorithmic FLOPs and quartets of one build of
 * the whole batch, and the measured milliseconds over the profiled builds.
 * enable = 2 profiles exclusively: the profiled builds run every class on one
 * stream, so that each class's time is its own and not shared with the
 * classes that otherwise overlap it on the side stream. */
int hfg_batch_profile(hfg_batch* batch, int enable);
int hfg_batch_class_stats(hfg_batch* batch, int cls, int32_t* l4, double* flops_per_launch, int64_t* quartets,
                          double* ms_total, int64_t* launches);

/* Pivoted Cholesky factors of the AO two-electron integrals
 * (pivoted_cholesky_eri, src/cholesky.rs:44-192): the pivot is the largest
 * residual diagonal (ties within 1e-12 relative resolve to the smallest packed
 * pair index); it stops when the largest residual is at most threshold.  The
 * factors stay in HBM behind a DF handle (rank = number of vectors), so
 * hfg_df_jk / hfg_df_jk_ex / hfg_df_info apply to them as to fitted B.
 * budget_bytes bounds the packed factors as CD_HF_FACTOR_BUDGET does
 * (src/hf_variants/cholesky_df.rs:22; HFG_INVALID_INPUT beyond it).  pivots
 * (optional) receives the first max_pivots pivots, packed AO-pair indices. */
typedef struct {
  int32_t rank;
  double max_residual; /* largest remaining residual diagonal (CholeskyInfo) */
} hfg_cd_info;
int hfg_cd_build(const hfg_basis_view* basis, double threshold, double budget_bytes, hfg_df** out,
                 hfg_cd_info* info, int32_t* pivots, int32_t max_pivots);
/* CD-HF (run_cd_hf_with_options, src/hf_variants/cholesky_df.rs:43-95): the
 * factors to delta in the factorized RHF driver of hfg_df_rhf (whose
 * metric_cutoff field is unused here); res->n_aux = the number of vectors. */
int hfg_cd_rhf(const hfg_basis_view* basis, const double* zxyz, int natom, int charge, double delta,
               double budget_bytes, const hfg_dfhf_options* opt, hfg_dfhf_result* res, hfg_cd_info* info,
               double* p_out, double* f_out);

/* Served and batched RHF (no reference counterpart: the reference solves one
 * molecule per call of run_scf_in_space_with_convergence, src/scf.rs:410-627,
 * or run_scf_with_initial_density, :630-639).  A server holds `capacity`
 * slots; each admitted request occupies one and every server iteration
 * advances all of them by one RHF iteration with the Fock builds of all slots
 * fused into one launch per shell-quartet class and the DIIS/purification
 * linear algebra batched over the slots.  A finished slot is refilled from the
 * queue of requests whose arrival time has passed, while the resident HBM
 * footprint stays within the budget (continuous batching).  A request is a
 * trajectory of geometry steps of one molecule: each step is an RHF solve
 * with the given options, started from the previous step's density (SAD for
 * the first); the per-molecule arithmetic is that of hfg_scf_rhf. */
typedef struct hfg_server hfg_server;
typedef struct {
  int32_t capacity;     /* slots: molecules resident at once */
  double hbm_budget_gb; /* bound on the resident footprint; <= 0: 80% of free HBM at creation */
} hfg_server_options;
typedef struct {
  double energy; /* total RHF energy of the step (Eh) */
  int32_t iterations, converged, fock_builds, n;
} hfg_step_result;
typedef struct {
  double arrival_s, admitted_s, finished_s; /* seconds from the start of hfg_server_run */
  int32_t steps, converged, status, n;      /* status: HFG_* of the request */
} hfg_request_result;
typedef struct {
  int64_t fused_builds, iterations;
  /* seconds: the run; seconds_setup: of it, waiting for slot set-ups (engine,
   * one-electron matrices, orthogonaliser, guess: they run on worker threads
   * while the slots already set up iterate) with no slot able to iterate;
   * seconds_sync: waiting for the per-iteration results */
  double seconds, seconds_setup, seconds_sync, hbm_budget_bytes;
  /* the class-launch configuration the autotuner chose (-1: still tuning)
   * out of jk_candidates */
  int32_t jk_config, jk_candidates;
} hfg_server_stats_t;
int hfg_server_create(const hfg_server_options* opt, hfg_server** out);
void hfg_server_destroy(hfg_server* server);
/* Queues a request: nsteps geometries of one molecule (bases[k] and zxyz[k],
 * natom atoms each, copied), arriving arrival_s seconds after the start of
 * the run.  flags: HFG_SERVE_KEEP_DENSITY keeps the last step's density for
 * hfg_server_result; HFG_SERVE_GRADIENT (the analytic nuclear gradient of
 * every step, compute_rhf_gradient, src/gradient.rs:272-285, for
 * hfg_server_gradient) is switched off in this build and returns
 * HFG_UNSUPPORTED. */
#define HFG_SERVE_KEEP_DENSITY 1
#define HFG_SERVE_GRADIENT 2
int hfg_server_submit(hfg_server* server, double arrival_s, int nsteps, const hfg_basis_view* bases,
                      const double* const* zxyz, int natom, int charge, const hfg_scf_options* opt, int flags,
                      int* request_id);
/* Serves every queued request (arrival times against the wall clock from the
 * start of this call) and returns when all have finished. */
int hfg_server_run(hfg_server* server);
/* A request's timing and per-step results (steps: nsteps entries, may be
 * NULL), and the last step's density (n x n, may be NULL; needs keep_density). */
int hfg_server_result(hfg_server* server, int request_id, hfg_request_result* result, hfg_step_result* steps,
                      double* p_out);
/* The nuclear gradients of a finished request submitted with
 * HFG_SERVE_GRADIENT: nsteps x natom x 3 (Eh/bohr), step-major. */
int hfg_server_gradient(hfg_server* server, int request_id, double* grad);
int hfg_server_stats(const hfg_server* server, hfg_server_stats_t* stats);

/* Host-side basis construction from Gaussian-94 text, mirroring build_basis
 * (src/basis.rs:296-496).  The arrays are owned by the returned handle. */
typedef struct hfg_basis hfg_basis;
int hfg_basis_from_g94(const char* g94, const char* symbols, const double* xyz, int natom,
                       hfg_basis** out);
void hfg_basis_free(hfg_basis* b);
int hfg_basis_get_view(const hfg_basis* b, hfg_basis_view* view);

#ifdef __cplusplus
}
#endif
#endif
