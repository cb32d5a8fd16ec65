/*
 * progress_gpu.h -- C-ABI of the MI355X density-matrix solver library
 * (libprogress_gpu.so).
 *
 * Entry points keep the names, argument order and meaning of PROGRESS's
 * C interface (src/prg_c_interface.F90 bind(C) wrappers, prototypes in
 * src/prg_progress_mod.h).  Each declaration cites the reference wrapper it
 * replaces.  Differences that are deliberate:
 *   - the reference returns nothing and calls `stop` on failure; here the
 *     routines still return void (same ABI) but record a status that
 *     prg_last_error_code()/prg_last_error() report, and never exit;
 *   - `sp2conv` is a NUL-terminated string compared like Fortran ("Rel",
 *     "Abs", trailing blanks ignored); the reference's bind(C) wrapper passes
 *     one character by value, so through its C path "Rel" never matches;
 *   - arrays (pp, vv) are plain pointers with the lengths stated per call;
 *   - out-parameters the reference loses through `value` (mu, occiter,
 *     totns) are pointers.
 *
 * The matrix handle: BML (lanl/bml) is not linked, so the
 * library carries its own bml_matrix_t holding a dense double-precision
 * column-major matrix with a lazily synchronised HBM mirror.  The handful of
 * bml_* functions below have BML's names and C signatures (bml.h) for the
 * dense / double_real / sequential case; other types and precisions are
 * accepted only where stated and are stored dense.
 */
#ifndef PROGRESS_GPU_H
#define PROGRESS_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ BML subset */
typedef enum { type_uninitialized, dense, ellpack, ellsort, ellblock, csr } bml_matrix_type_t;
typedef enum {
  precision_uninitialized,
  single_real,
  double_real,
  single_complex,
  double_complex
} bml_matrix_precision_t;
typedef enum { sequential, distributed, graph_distributed } bml_distribution_mode_t;
typedef enum { dense_row_major, dense_column_major } bml_dense_order_t;

typedef struct bml_matrix_t bml_matrix_t;

/* N x N zero / identity matrix; any matrix_type is stored dense; only
 * double_real precision is supported (others: NULL + error). */
bml_matrix_t *bml_zero_matrix(bml_matrix_type_t matrix_type, bml_matrix_precision_t precision,
                              int N, int M, bml_distribution_mode_t distrib_mode);
bml_matrix_t *bml_identity_matrix(bml_matrix_type_t matrix_type,
                                  bml_matrix_precision_t precision, int N, int M,
                                  bml_distribution_mode_t distrib_mode);
/* Sparse types (ellpack, ellsort, ellblock, csr) keep only the entries with
 * |a| > threshold, as the BML formats do; every type is stored dense. */
bml_matrix_t *bml_import_from_dense(bml_matrix_type_t matrix_type,
                                    bml_matrix_precision_t precision, bml_dense_order_t order,
                                    int N, int M, void *A, double threshold,
                                    bml_distribution_mode_t distrib_mode);
/* Returns a malloc'ed N*N double array the caller frees with free(). */
void *bml_export_to_dense(bml_matrix_t *A, bml_dense_order_t order);
bml_matrix_t *bml_copy_new(bml_matrix_t *A);
void bml_copy(bml_matrix_t *A, bml_matrix_t *B); /* B <- A */
bml_matrix_t *bml_transpose_new(bml_matrix_t *A);
void bml_add(bml_matrix_t *A, bml_matrix_t *B, double alpha, double beta, double threshold);
void bml_scale(void *scale_factor, bml_matrix_t *A, bml_matrix_t *B); /* B <- s A */
double bml_trace(bml_matrix_t *A);
int bml_get_N(bml_matrix_t *A);
int bml_get_M(bml_matrix_t *A);
double bml_get_element(bml_matrix_t *A, int i, int j); /* 0-based */
void bml_read_bml_matrix(bml_matrix_t *A, char *filename); /* Matrix Market */
void bml_write_bml_matrix(bml_matrix_t *A, char *filename);
void bml_deallocate(bml_matrix_t **A);
/* BML accessors the graph-partitioned SP2 callers use (examples/gptest/
 * gptest.F90:68,82,100,131,144): thresholding (|a| <= threshold -> 0),
 * Tr(A B), ||A - B||_F, and the Gershgorin interval as a malloc'ed {min, max}
 * pair the caller frees. */
void bml_threshold(bml_matrix_t *A, double threshold);
double bml_trace_mult(bml_matrix_t *A, bml_matrix_t *B);
double bml_fnorm2(bml_matrix_t *A, bml_matrix_t *B);
void *bml_gershgorin(bml_matrix_t *A);

/* ------------------------------------------------------------------ library */
/* src/prg_c_interface.F90:51-61 */
void prg_version(void);
void prg_progress_init(void);
void prg_progress_shutdown(void);

/* Status of the last call on this thread (0 = success) and its message. */
int prg_last_error_code(void);
const char *prg_last_error(void);

/* ------------------------------------------------------------------ normalize */
/* src/prg_c_interface.F90:849-854 (prg_normalize_mod.F90:32-46) */
void prg_normalize(bml_matrix_t *h_bml);
/* src/prg_c_interface.F90:856-865 (prg_normalize_mod.F90:59-72) */
void prg_normalize_fermi(bml_matrix_t *h_bml, double h1, double hN, double mu);
/* src/prg_c_interface.F90:867-876 (prg_normalize_mod.F90:83-94) */
void prg_normalize_implicit_fermi(bml_matrix_t *h_bml, double cnst, double mu);
/* Gershgorin interval [gbnd[0], gbnd[1]] (bml_gershgorin as used at
 * src/prg_normalize_mod.F90:41). */
void prg_gershgorin(bml_matrix_t *h_bml, double *gbnd);

/* ------------------------------------------------------------------ SP2 */
/* src/prg_c_interface.F90:1073-1089 (prg_sp2_mod.F90:49-135) */
void prg_sp2_basic(bml_matrix_t *h_bml, bml_matrix_t *rho_bml, double threshold, double bndfil,
                   int minsp2iter, int maxsp2iter, const char *sp2conv, double idemtol,
                   int verbose);
/* src/prg_c_interface.F90:1091-1111 (prg_sp2_mod.F90:138-246; prototype
 * src/prg_progress_mod.h:503-513): the basic branch rule for exactly
 * maxsp2iter + 1 iterations (no convergence exit, as in the reference), D = 2X.
 * rhofull is read only for the verbose > 1 diagnostic (the final
 * ||rhofull/2 - X||_F / ||rhofull||_F); it may be NULL. */
void prg_sp2_basic_tcore(bml_matrix_t *h_bml, bml_matrix_t *rho_bml, bml_matrix_t *rhofull_bml,
                         double threshold, double bndfil, int minsp2iter, int maxsp2iter,
                         const char *sp2conv, double idemtol, int verbose);
/* src/sp2_tensorcore/src/prg_sp2_tensorcore.cu:131-141: the reference's
 * mixed-precision SP2 entry (host float H, host double D = X, not 2X).  Runs
 * this library's FP16 hi+lo split SP2 with the reference's sigma rule,
 * noise-floor stop and FP64 refinement, Nocc = int(bndfil N).  Differences:
 * Gershgorin bounds instead of the reference's hard-coded (-27.55, 25.0); up
 * to two FP64 SP2 polishing steps in place of its single 2T^2 - T^4.  eps,
 * minsp2iter, sp2conv and idemtol are accepted and, as there, not used. */
void prg_sp2_tensorcore(int N, float *H, double *D, float eps, float bndfil, int minsp2iter,
                        int maxsp2iter, char sp2conv, float idemtol, int verbose);
/* src/prg_c_interface.F90:1113-1129 (prg_sp2_mod.F90:263-380) */
void prg_sp2_alg2(bml_matrix_t *h_bml, bml_matrix_t *rho_bml, double threshold, double bndfil,
                  int minsp2iter, int maxsp2iter, const char *sp2conv, double idemtol,
                  int verbose);
/* src/prg_c_interface.F90:1131-1151 (prg_sp2_mod.F90:395-533).
 * pp, vv: caller arrays of length maxsp2iter; vv gets one entry per
 * computed product (vv[*icount] holds the discarded last product when the
 * loop ends in the limdiff branch, as in the reference). */
void prg_sp2_alg2_genseq(bml_matrix_t *h_bml, bml_matrix_t *rho_bml, double threshold,
                         double bndfil, int minsp2iter, int maxsp2iter, const char *sp2conv,
                         double idemtol, int *pp, int *icount, double *vv, int verbose);
/* src/prg_c_interface.F90:1153-1167 (prg_sp2_mod.F90:543-627); pp, vv length >= *icount */
void prg_sp2_alg2_seq(bml_matrix_t *h_bml, bml_matrix_t *rho_bml, double threshold, int *pp,
                      int *icount, double *vv, int verbose);
/* src/prg_c_interface.F90:1169-1184 (prg_sp2_mod.F90:638-722) */
void prg_prg_sp2_alg2_seq_inplace(bml_matrix_t *rho_bml, double threshold, int ppsize, int *pp,
                                  int *icount, double *vv, double mineval, double maxeval);
/* src/prg_c_interface.F90:1186-1204 (prg_sp2_mod.F90:734-864) */
void prg_sp2_alg1(bml_matrix_t *h_bml, bml_matrix_t *rho_bml, double threshold, double bndfil,
                  int minsp2iter, int maxsp2iter, const char *sp2conv, double idemtol,
                  int verbose);
/* src/prg_c_interface.F90:1206-1224 (prg_sp2_mod.F90:879-1000) */
void prg_sp2_alg1_genseq(bml_matrix_t *h_bml, bml_matrix_t *rho_bml, double threshold,
                         double bndfil, int minsp2iter, int maxsp2iter, const char *sp2conv,
                         double idemtol, int *pp, int *icount, double *vv);
/* src/prg_c_interface.F90:1226-1239 (prg_sp2_mod.F90:1010-1089) */
void prg_sp2_alg1_seq(bml_matrix_t *h_bml, bml_matrix_t *rho_bml, double threshold, int *pp,
                      int *icount, double *vv);
/* src/prg_c_interface.F90:1241-1255 (prg_sp2_mod.F90:1100-1177) */
void prg_prg_sp2_alg1_seq_inplace(bml_matrix_t *rho_bml, double threshold, int ppsize, int *pp,
                                  int *icount, double *vv, double mineval, double maxeval);
/* src/prg_c_interface.F90:1257-1273 (prg_sp2_mod.F90:1190-1246); vv is
 * accumulated (vv[i] += core-row sum of (X - X^2)^2), length >= *icount.
 * The reference header's prototype (src/prg_progress_mod.h:601-608) omits the
 * two matrices; this follows the Fortran wrapper. */
void prg_sp2_submatrix(bml_matrix_t *ham_bml, bml_matrix_t *rho_bml, double threshold, int *pp,
                       int *icount, double *vv, double mineval, double maxeval, int core_size);
/* src/prg_c_interface.F90:1275-1288 (prg_sp2_mod.F90:1259-1310) */
void prg_sp2_submatrix_inplace(bml_matrix_t *rho_bml, double threshold, int *pp, int *icount,
                               double *vv, double mineval, double maxeval, int core_size);

/* ------------------------------------------------------------------ truncated SP2 (finite T) */
/* src/prg_c_interface.F90:904-927 (prg_sp2_fermi_mod.F90:43-185).  sgnlist:
 * caller array of length >= nsteps (+1 = 2X - X^2, -1 = X^2), filled on the
 * first pass; mu, beta, h1, hN are outputs.  The trailing verbose is the
 * bind(C) wrapper's (the header at src/prg_progress_mod.h:435-448 omits it). */
void prg_sp2_fermi_init(bml_matrix_t *h_bml, int nsteps, double nocc, double tscale,
                        double threshold, double occErrLimit, double traceLimit,
                        bml_matrix_t *x_bml, double *mu, double *beta, double *h1, double *hN,
                        int *sgnlist, int verbose);
/* src/prg_c_interface.F90:929-955 (prg_sp2_fermi_mod.F90:203-387): *beta in is
 * the target inverse temperature; *nsteps is cut where the SP2 polynomial's
 * steepest slope first exceeds it. */
void prg_sp2_fermi_init_norecs(bml_matrix_t *h_bml, int *nsteps, double nocc, double tscale,
                               double threshold, double occErrLimit, double traceLimit,
                               bml_matrix_t *x_bml, double *mu, double *beta, double *h1,
                               double *hN, int *sgnlist, int verbose);
/* src/prg_c_interface.F90:957-979 (prg_sp2_fermi_mod.F90:405-470); x_bml gets
 * the occupation-corrected X (not doubled), *mu the converged value. */
void prg_sp2_fermi(bml_matrix_t *h_bml, int osteps, int nsteps, double nocc, double *mu,
                   double *beta, double *h1, double *hN, int *sgnlist, double threshold,
                   double eps, double traceLimit, bml_matrix_t *x_bml);
/* src/prg_c_interface.F90:981-995 (prg_sp2_fermi_mod.F90:480-531); GG, ee:
 * caller arrays of 10001 doubles (the mesh 0, 1e-4, ..., 1). */
void prg_sp2_entropy_function(double mu, double h1, double hN, int nsteps, const int *sgnlist,
                              double *GG, double *ee);
/* sp2_inverse (prg_sp2_fermi_mod.F90:720-745) and sp2_entropy_ts (:541-587);
 * no C symbols in the reference. */
double prg_sp2_inverse(double f, double mu, double h1, double hN, int nsteps, const int *sgnlist);
double prg_sp2_entropy_ts(bml_matrix_t *d0_bml, const double *GG, const double *ee);

/* ------------------------------------------------------------------ congruence transforms */
/* prg_orthogonalize (src/prg_nonortho_mod.F90:35-68): orthoA = Z^T A Z, and
 * prg_deorthogonalize (:81-120): A = Z orthoA Z^T.  No C symbols in the
 * reference; Fortran argument order (bml_type is accepted and unused: every
 * handle here is dense). */
void prg_orthogonalize(bml_matrix_t *A_bml, bml_matrix_t *zmat_bml, bml_matrix_t *orthoA_bml,
                       double threshold, const char *bml_type, int verbose);
void prg_deorthogonalize(bml_matrix_t *orthoA_bml, bml_matrix_t *zmat_bml, bml_matrix_t *a_bml,
                         double threshold, const char *bml_type, int verbose);

/* ------------------------------------------------------------------ homo-lumo / sequence */
/* New C symbols for src/prg_homolumo_mod.F90:22-89 and :97-147 (no bind(C)
 * wrapper in the reference).  pp/vv hold the reference's pp(1..), vv(1..). */
void prg_homolumogap(const double *vv, int imax, const int *pp, double mineval, double maxeval,
                     double *ehomo, double *elumo, double *egap, int verbose);
/* pp: caller array of length >= 100; *imax receives the sequence length */
void prg_sp2sequence(int *pp, int *imax, double mineval, double maxeval, double ehomo,
                     double elumo, double errlimit, int verbose);

/* ------------------------------------------------------------------ eigensolver densities */
/* src/prg_c_interface.F90:68-86 (prg_densitymatrix_mod.F90:36-101); rocSOLVER dsyevd */
void prg_build_density_T0(int norbs, bml_matrix_t *ham_bml, bml_matrix_t *rho_bml,
                          double threshold, double bndfil, double *eigenvalues_out);
/* src/prg_c_interface.F90:227-241 (prg_densitymatrix_mod.F90:509-582) */
void prg_build_density_T_fermi(bml_matrix_t *ham_bml, bml_matrix_t *rho_bml, double threshold,
                               double kbt, double ef, int verbose, double *drho);
/* src/prg_c_interface.F90:281-293 (prg_densitymatrix_mod.F90:817-854) */
void prg_get_eigenvalues(int norbs, bml_matrix_t *ham_bml, double *eigenvalues_out, int verbose);
/* src/prg_c_interface.F90:295-304 (prg_densitymatrix_mod.F90:862-880) */
void prg_check_idempotency(bml_matrix_t *mat_bml, double threshold, double *idempotency);
/* prg_get_density_matrix (prg_implicit_fermi_mod.F90:1056-1100): P = f(H), no factor 2 */
void prg_get_density_matrix(bml_matrix_t *ham_bml, bml_matrix_t *p_bml, double beta, double mu,
                            double threshold);
/* src/prg_c_interface.F90:535-549 (prg_implicit_fermi_mod.F90:1112-1186); mu by value as in
 * the reference; prg_test_density_matrix_mu() returns the converged value. */
void prg_test_density_matrix(bml_matrix_t *ham_bml, bml_matrix_t *p_bml, double beta, double mu,
                             double nocc, int osteps, double occErrLimit, double threshold);
double prg_test_density_matrix_mu(void);

/* ------------------------------------------------------------------ implicit Fermi */
/* src/prg_c_interface.F90:470-490 (prg_implicit_fermi_mod.F90:209-336); even
 * expansion order k (k > 2: prg_setup_linsys, :843-876; odd k is an error).
 * mu by value as in the reference; prg_implicit_fermi_last_mu() returns the update. */
void prg_implicit_fermi(bml_matrix_t *h_bml, bml_matrix_t *p_bml, int nsteps, int k, double nocc,
                        double mu, double beta, int method, int osteps, double occErrLimit,
                        double threshold, double tol);
double prg_implicit_fermi_last_mu(void);
/* src/prg_c_interface.F90:440-468 (prg_implicit_fermi_mod.F90:46-193); Inv_bml[nsteps] are
 * caller-owned and persist across calls (warm starts); occiter / totns by value as in the
 * reference header, prg_implicit_fermi_counters() returns the updated counts. */
void prg_implicit_fermi_save_inverse(bml_matrix_t **Inv_bml, bml_matrix_t *h_bml,
                                     bml_matrix_t *p_bml, int nsteps, double nocc, double *mu,
                                     double beta, double occErrLimit, double threshold, double tol,
                                     int SCF_IT, int occiter, int totns);
void prg_implicit_fermi_counters(int *occiter, int *totns);
/* src/prg_c_interface.F90:492-506 (prg_implicit_fermi_mod.F90:347-425) */
void prg_implicit_fermi_zero(bml_matrix_t *h_bml, bml_matrix_t *p_bml, int nsteps, double mu,
                             int method, double threshold, double tol);
/* src/prg_c_interface.F90:508-533 (prg_implicit_fermi_mod.F90:439-529) */
void prg_implicit_fermi_first_order_response(bml_matrix_t *H0_bml, bml_matrix_t *H1_bml,
                                             bml_matrix_t *P0_bml, bml_matrix_t *P1_bml,
                                             bml_matrix_t **Inv_bml, int nsteps, double mu0,
                                             double beta, double nocc, double threshold);
/* Declared at src/prg_progress_mod.h:240-258 without an implementation there
 * (prg_implicit_fermi_mod.F90:547-712); the declared arity: H1..H3 and P1..P3 as
 * separate handles (those above `order` may be NULL), mu0 and mu by value (the
 * Fortran routine zeroes mu(:) on entry, so mu is unused).  The updated
 * chemical potentials are returned by prg_implicit_fermi_response_mu. */
void prg_implicit_fermi_response(bml_matrix_t *H0_bml, bml_matrix_t *H1_bml, bml_matrix_t *H2_bml,
                                 bml_matrix_t *H3_bml, bml_matrix_t *P0_bml, bml_matrix_t *P1_bml,
                                 bml_matrix_t *P2_bml, bml_matrix_t *P3_bml, int nsteps,
                                 double mu0, double mu, double beta, double nocc, double occ_tol,
                                 double lin_tol, int order, double threshold);
/* mu0 and mu[0..2] after the last prg_implicit_fermi_response call. */
void prg_implicit_fermi_response_mu(double *mu0, double *mu);
/* The same routine with handle arrays and pointer out-parameters (what the
 * Fortran module calls): H_bml[order] = H1..H3, P_bml[order] = P1..P3, *mu0
 * updated, mu[order] receives the response chemical potentials.  No
 * counterpart in the reference's C header. */
void prg_gpu_implicit_fermi_response(bml_matrix_t *H0_bml, bml_matrix_t **H_bml,
                                     bml_matrix_t *P0_bml, bml_matrix_t **P_bml, int nsteps,
                                     double *mu0, double *mu, double beta, double nocc,
                                     double occ_tol, double lin_tol, int order, double threshold);

/* ------------------------------------------------------------------ response */
/* New C symbol for prg_compute_response_SP2 (prg_response_mod.F90:653-782), Fortran order:
 * rho = 2X, rsp = 2 Delta (first-order response to prt). */
void prg_compute_response_SP2(bml_matrix_t *ham_bml, bml_matrix_t *prt_bml, bml_matrix_t *rsp_bml,
                              bml_matrix_t *rho_bml, double lambda, double bndfil, int minsp2iter,
                              int maxsp2iter, const char *sp2conv, double idemtol,
                              double threshold, int verbose);
/* New C symbol for prg_canon_response (prg_response_mod.F90:947-1022): H1 and P1 in the
 * eigenbasis of H0, evals[HDIM] its eigenvalues (host). */
void prg_canon_response(bml_matrix_t *P1_bml, bml_matrix_t *H1_bml, double nocc, double beta,
                        const double *evals, double mu0, int m, int HDIM);
/* New C symbol for prg_canon_response_orig (prg_response_mod.F90:1038-1105): the
 * same result as prg_canon_response (thresh unused). */
void prg_canon_response_orig(bml_matrix_t *P1_bml, bml_matrix_t *H1_bml, double nocc, double beta,
                             const double *evals, double mu0, int m, double thresh, int HDIM);
/* New C symbol for prg_canon_response_p1_dpdmu (prg_response_mod.F90:1112-1183):
 * the recursion without the trace correction; dPdMu[HDIM] (host) receives
 * beta p0 (1 - p0).  Norbs and Q_bml are unused (may be 0 / NULL). */
void prg_canon_response_p1_dpdmu(bml_matrix_t *P1_bml, double *dPdMu, bml_matrix_t *H1_bml,
                                 int Norbs, double beta, bml_matrix_t *Q_bml, const double *evals,
                                 double mu0, int m, int HDIM);
/* New C symbol for prg_canon_response_vector (prg_response_mod.F90:849-931): the
 * exact Fermi response in the eigenbasis (element-wise divided differences),
 * canonical diagonal correction (m, thresh unused). */
void prg_canon_response_vector(bml_matrix_t *P1_bml, bml_matrix_t *H1_bml, double nocc,
                               double beta, const double *evals, double mu0, int m, double thresh,
                               int HDIM);
/* New C symbol for prg_compute_response_RS (prg_response_mod.F90:250-361): T = 0
 * Rayleigh-Schroedinger first-order response rsp = C f C1^T + C1 f C^T. */
void prg_compute_response_RS(bml_matrix_t *ham_bml, bml_matrix_t *prt_bml, bml_matrix_t *rsp_bml,
                             double lambda, double bndfil, double threshold, int verbose);
/* New C symbol for prg_compute_response_FD (prg_response_mod.F90:380-424):
 * (rho(H + delta H1) - rho(H - delta H1)) / (2 delta), T = 0 densities. */
void prg_compute_response_FD(bml_matrix_t *ham_bml, bml_matrix_t *prt_bml, bml_matrix_t *rsp_bml,
                             double delta, double bndfil, double threshold, int verbose);

/* ------------------------------------------------------------------ graph partitioning */
/* Flat C counterpart of graph_partitioning_t (src/prg_graph_mod.F90:58-140) for
 * one rank: 0-based node ids; lsize / llsize / core_halo_index are filled by
 * prg_subgraphSP2Loop (or prg_getPartitionHalosFromGraph); pp / maxIter /
 * mineval / maxeval are the caller's inputs to the loop (prg_sp2sequence,
 * prg_gershgorin); vv receives sqrt of the per-iteration core-row sums. */
typedef struct prg_graph_partitioning {
  int totalParts;
  int totalNodes;
  int *nnodesInPart;     /* [totalParts] core sizes */
  int **nodeInPart;      /* [totalParts][nnodesInPart[i]] core node ids */
  int *lsize;            /* [totalParts] core + halo sizes */
  int *llsize;           /* [totalParts] core sizes */
  int **core_halo_index; /* [totalParts][lsize[i]] core first, then halo */
  int pp[100];
  int maxIter;
  double ehomo, elumo, mineval, maxeval;
  double vv[100];
} prg_graph_partitioning_t;

/* src/prg_graph_mod.F90:359-399 */
void prg_equalPartition(prg_graph_partitioning_t *gp, int nodesPerPart, int nnodes);
/* src/prg_graph_mod.F90:880-929 */
void prg_filePartition(prg_graph_partitioning_t *gp, const char *partFile);
/* Partition from flat arrays (a partitioner's output); no reference counterpart. */
void prg_setPartition(prg_graph_partitioning_t *gp, int nparts, int nnodes, const int *sizes,
                      const int *nodes);
/* prg_destroyGraphPartitioning (src/prg_graph_mod.F90:268-290) */
void prg_destroyGraphPartitioning(prg_graph_partitioning_t *gp);
/* src/prg_subgraphloop_mod.F90:336-368 */
void prg_getPartitionHalosFromGraph(prg_graph_partitioning_t *gp, bml_matrix_t *g_bml,
                                    int djflag);
/* src/prg_graph_mod.F90:933-971 */
void prg_fnormGraph(prg_graph_partitioning_t *gp);
/* New C symbol for prg_subgraphSP2Loop (src/prg_subgraphloop_mod.F90:36-126):
 * all parts purified in lock-step by batched GPU kernels. */
void prg_subgraphSP2Loop(bml_matrix_t *h_bml, bml_matrix_t *g_bml, bml_matrix_t *rho_bml,
                         prg_graph_partitioning_t *gp, double threshold);
/* The build's graph partitioner in place of METIS (prg_metisPartition,
 * src/prg_partition_mod.F90:216-311, has no C symbol): recursive bisection of
 * breadth-first orderings of the graph of g_bml's nonzero off-diagonal
 * entries into nparts parts (cores in ascending node order).  No reference
 * counterpart as a C symbol. */
void prg_bfsPartition(prg_graph_partitioning_t *gp, bml_matrix_t *g_bml, int nparts);
/* src/prg_progress_mod.h:803-811 (src/prg_graphsolver_mod.F90:39-191): graph-
 * partitioned finite-temperature density (kT = 0.1, per-part batched
 * eigensolves) with the reference's Fermi-level update.  Ef by value as in the
 * reference header; prg_build_densityGP_T0_ef() returns the updated value. */
void prg_build_densityGP_T0(bml_matrix_t *ham_bml, bml_matrix_t *g_bml, bml_matrix_t *rho_bml,
                            double threshold, double bndfil, double Ef, int nparts, int verbose);
double prg_build_densityGP_T0_ef(void);
/* src/prg_progress_mod.h:814-820 (declared there without an implementation;
 * Fortran src/prg_graphsolver_mod.F90:202-331): the congruence transform
 * Z = S^-1/2 assembled from parts.  The graph g_bml is partitioned into
 * nparts parts (the library's BFS partitioner in place of METIS); for each
 * part the Loewdin factor U s^-1/2 U^T of the overlap's double-jump
 * core+halo block (prg_buildZdiag, src/prg_genz_mod.F90:176-276) is formed
 * from batched eigensolves and its core rows are written into zmat_bml
 * (bml_submatrix2matrix); other entries of zmat_bml keep their values.  A
 * block that is not positive definite records a numeric error (the
 * reference stops).  threshold is ignored (dense storage). */
void prg_build_zmatGP(bml_matrix_t *over_bml, bml_matrix_t *g_bml, bml_matrix_t *zmat_bml,
                      double threshold, int nparts, int verbose);
/* Device time (ms) of the last prg_subgraphSP2Loop phases. */
void prg_gpu_subgraph_timing(double *ms_gather, double *ms_loop, double *ms_scatter);
/* Kernel configuration of the last prg_subgraphSP2Loop: the tile panel width
 * (64 or 128), the number of K slabs its batched SP2 kernel prefetches
 * (1 or 2), the K slab depth (16 or 32), and how its iterations were
 * launched: 0 eager (PRG_BATCH_GRAPH=0), 1 captured into a hipGraph, 2 replay
 * of the graph captured by an earlier call with the same buffers, tile list
 * and branch sequence.  The configuration is timed per batch of part sizes
 * on first use among the compiled ones (64/1/16, 64/2/16, 64/1/32,
 * 128/1/16); PRG_BATCH_TILE, PRG_BATCH_DEPTH and PRG_BATCH_KSLAB restrict
 * the choice.  No counterpart in the reference. */
void prg_gpu_subgraph_config(int *panel, int *depth, int *kslab, int *graph);

/* ------------------------------------------------------------------ precision modes */
/* Selects how prg_sp2_* evaluate the X^2 products (default PRG_PREC_FP64).
 *   PRG_PREC_FP64      FP64 MFMA;
 *   PRG_PREC_F16SPLIT  FP16 hi+lo split products with FP32 accumulation;
 *   PRG_PREC_BF16SPLIT BF16 hi+lo split products with FP32 accumulation
 *                      (about 16 significant bits; same stopping and FP64
 *                      polishing as the FP16 split);
 *   PRG_PREC_INT8CRT   exact int8 MFMA products of X's residues modulo
 *                      pairwise-coprime moduli <= 256, Chinese-remainder
 *                      reconstruction in FP64 (11 moduli: X carried to
 *                      2^-42 absolute), FP64 bookkeeping.
 * No counterpart in the reference (whose CPU path is FP64 only). */
#define PRG_PREC_FP64 0
#define PRG_PREC_F16SPLIT 1
#define PRG_PREC_INT8CRT 2
#define PRG_PREC_BF16SPLIT 3
void prg_set_precision(int mode);
int prg_get_precision(void);
/* Number of moduli (6..14, default 11) of PRG_PREC_INT8CRT.  In the SP2
 * square X is scaled by 2^a with a = min(52, floor((log2 M + 1 - 0.001) / 2))
 * (the residues stand for X^2 - X, whose entries are at most 1/4): 8 -> 32,
 * 9 -> 36, 10 -> 40, 11 -> 44, 12 -> 47, 13 -> 51, 14 -> 52 bits; general
 * products use a = min(52, floor((log2 M - 1.001) / 2)), one bit less.
 * Returns a status. */
int prg_set_int8_moduli(int count);
int prg_get_int8_moduli(void);
/* Configuration the int8 CRT product kernel last ran with, coded
 * waves * 100 + ring stages (803, 804, 805: 8 waves, one item per
 * workgroup; 904: the persistent 8-wave kernel, one workgroup per CU; 944:
 * 904 with its LDS fragment reads interleaved among the MFMAs):
 * chosen per (padded size, moduli) by timing each on the first product, or
 * forced with the environment variable PRG_CRT_CONFIG.  0 before any int8
 * product.  Measurement support; no reference counterpart. */
int prg_gpu_int8_waves(void);
/* K slices of the FP64 SP2 step at size N: 1 for the fused one-kernel step
 * (sp2_step_kernel), s > 1 for the split product and its epilogue kernel
 * (fewer than 1024 upper-triangle 128 x 128 tiles).  Measurement support;
 * no reference counterpart. */
int prg_gpu_sp2_split_slices(int N);
/* Forces the int8 CRT product kernel configuration (one of the codes above;
 * 0 returns to timing them per size).  Returns a status (1 for an unknown
 * code).  Test and measurement support; no reference counterpart. */
int prg_gpu_set_int8_config(int config);

/* ------------------------------------------------------------------ device-pointer API */
/* prg_sp2_alg2 on device-resident data (no host copies): dH and dRho are
 * ld x ld column-major FP64 device buffers with ld = ((N+127)/128)*128 and
 * zeros outside the N x N block (dRho gets 2X in the block, zeros outside).
 * Returns a status; *iters receives the iteration count, pp (may be NULL,
 * length maxsp2iter) the branch sequence.  Runs on the library stream and
 * synchronises before returning.  No reference counterpart: the entry point
 * the benchmark and the Python binding use to keep inputs in HBM. */
int prg_gpu_sp2_alg2_device(int N, const double *dH, double *dRho, double bndfil, int minsp2iter,
                            int maxsp2iter, const char *sp2conv, double idemtol, int precision,
                            int *iters, int *pp);
/* prg_gpu_sp2_alg2_device in FP64 without any allocation: no counterpart in
 * the reference.  For callers that cannot allocate (code under HIP graph
 * capture, tight serving or MD loops, harnesses that forbid hipMalloc).  The
 * ping-pong X, partial sums, state and K-split accumulators live in a static
 * device array of 52.3 MiB held by every device the library runs on until
 * the process ends; N is limited to 1024 (larger N returns status 1, bad
 * argument).  prg_progress_init must have been called on the device (it
 * resolves the array's address; otherwise status 6, no device).  The array
 * keeps the last call's contents; calls that overlap on one device race.
 * Same arguments and results as prg_gpu_sp2_alg2_device with precision
 * PRG_PREC_FP64. */
int prg_gpu_sp2_alg2_device_noalloc(int N, const double *dH, double *dRho, double bndfil,
                                    int minsp2iter, int maxsp2iter, const char *sp2conv,
                                    double idemtol, int *iters, int *pp);
/* New C symbols for the linear solvers of prg_implicit_fermi_mod.F90 (Fortran
 * argument order; the auxiliary matrices are unused and may be NULL):
 * prg_conjgrad (:1006-1048) and prg_pcg (:934-996, Jacobi preconditioner)
 * solve A X = B with X in p_bml (initial guess on entry), stopping when
 * ||R||_F^2 <= cg_tol; prg_pcg leaves the residual in p2_bml as the reference
 * does.  prg_newtonschulz (:888-922) refines ai_bml towards A^-1 until
 * ||I - A Ai||_F <= tol and stores the iteration count in *num_iter. */
void prg_conjgrad(bml_matrix_t *A_bml, bml_matrix_t *p_bml, bml_matrix_t *p2_bml,
                  bml_matrix_t *tmp_bml, bml_matrix_t *d_bml, bml_matrix_t *w_bml, double cg_tol,
                  double threshold);
void prg_pcg(bml_matrix_t *A_bml, bml_matrix_t *p_bml, bml_matrix_t *p2_bml, bml_matrix_t *d_bml,
             bml_matrix_t *wtmp_bml, double cg_tol, double threshold);
void prg_newtonschulz(bml_matrix_t *a_bml, bml_matrix_t *ai_bml, bml_matrix_t *r_bml,
                      bml_matrix_t *tmp_bml, bml_matrix_t *d_bml, bml_matrix_t *I_bml, double tol,
                      double threshold, int *num_iter);
/* prg_finite_diff (prg_implicit_fermi_mod.F90:726-831) at the arity and types
 * of its declaration, src/prg_progress_mod.h:259-268: one perturbation handle
 * (H_list = H1) and one coefficient (mu_list = mu1) by value, so order must be
 * 1; the difference is printed as the reference does (top-left 10 x 10 times
 * 1000).  Orders 2-3: prg_gpu_finite_diff. */
void prg_finite_diff(bml_matrix_t *H0_bml, bml_matrix_t *H_list, double mu0, double mu_list,
                     double beta, int order, double lambda, double h, double threshold);
/* The extended form (what the Fortran module calls): finite-difference
 * responses of prg_get_density_matrix along H0 + sum l^i H_i, mu0 + sum l^i mu_i
 * (i = 1..order, order <= 3) at l = lambda, step h; out_list[k] (order k + 1;
 * out_list may be NULL: then the corners are printed) receives each.  No
 * counterpart in the reference's C header. */
void prg_gpu_finite_diff(bml_matrix_t *H0_bml, bml_matrix_t **H_list, double mu0,
                         const double *mu_list, double beta, int order, double lambda, double h,
                         double threshold, bml_matrix_t **out_list);
/* prg_implicit_fermi at a fixed chemical potential (one pass of the level
 * loop of src/prg_implicit_fermi_mod.F90:259-298, method 1: level-1 CG inverse
 * refined by Newton-Schulz to tol at every level) on device pointers:
 * dP = P (not scaled by 2).  Returns a status. */
int prg_gpu_implicit_fermi_device(int N, const double *dH, double *dP, int nsteps, double mu,
                                  double beta, double tol);
/* Gradient dG = dTr(P A)/dH of that fixed-mu map (A symmetric), by a reverse
 * sweep over the levels (every level's X and inverse held in HBM, (2 nsteps +
 * 1) N^2 doubles); canonical != 0: the derivative at fixed occupation with the
 * reference's dP/dmu ~ beta P (I - P) (src/prg_implicit_fermi_mod.F90:512-521),
 * so that Tr(H1 dG) equals Tr(A P1) for the first-order response P1 of
 * prg_implicit_fermi_first_order_response.  dP (may be NULL) receives P.
 * Exactly symmetric.  No reference counterpart.  Returns a status. */
int prg_gpu_implicit_fermi_adjoint_device(int N, const double *dH, const double *dA, double *dG,
                                          int nsteps, double mu, double beta, double tol,
                                          int canonical, double *dP);
/* Padded leading dimension used by the device-pointer API. */
int prg_gpu_padded_ld(int N);

/* Gershgorin interval of a device-resident H (padded layout) into bounds[2] (host). */
int prg_gpu_gershgorin_device(int N, const double *dH, double *bounds);
/* prg_prg_sp2_alg2_seq_inplace semantics on device pointers: replay the branch
 * sequence pp[icount] with the given normalisation bounds; dRho = 2 X. */
int prg_gpu_sp2_seq_device(int N, const double *dH, double *dRho, const int *pp, int icount,
                           double mineval, double maxeval);
/* Reverse-mode adjoint of the SP2 recursion (no reference counterpart; the
 * transpose of the tangent recursion of prg_compute_response_SP2,
 * prg_response_mod.F90:719-755): given dL/dD in dGout, writes dL/dH to dGrad for
 * the forward solve that took the branch sequence pp[iters] with normalisation
 * bounds (mineval, maxeval), both frozen.  Checkpoints every `ckpt` steps. */
int prg_gpu_sp2_adjoint_device(int N, const double *dH, const double *dGout, double *dGrad,
                               const int *pp, int iters, double mineval, double maxeval,
                               int ckpt);

/* Opt-in timing of the SP2 product steps with HIP events on the library
 * stream (enable resets the counters); prg_gpu_step_timing returns the summed
 * milliseconds and the number of timed steps (all kernels of a step), and
 * prg_gpu_int8_product_timing the same for the int8 CRT product kernel alone
 * (the step's other kernel is the FP64 reconstruction / update).  Measurement
 * support for the benchmark; no reference counterpart. */
void prg_gpu_timing(int enable);
int prg_gpu_step_timing(double *ms_sum, int *count);
int prg_gpu_int8_product_timing(double *ms_sum, int *count);
/* Products issued by the iterative solvers (implicit Fermi and its
 * responses, SP2-Fermi, the adjoint) since the last reset: count, algorithmic
 * flops (2 N^3 each), FP64 MFMA flops executed (symmetric products compute
 * the upper-triangle tiles only) and int8 MFMA operations executed by int8
 * CRT products.  reset != 0 clears the counters after reading.  Any pointer
 * may be NULL.  Measurement support; no reference counterpart. */
void prg_gpu_gemm_counters(int reset, double *products, double *alg_flops, double *exec_flops,
                           double *int8_ops);
/* Arithmetic of the symmetric products of commuting symmetric matrices in
 * the iterative solvers other than the SP2 loops -- implicit Fermi (X^2,
 * the Newton-Schulz and CG products, Inv X^2), its saved-inverse variant,
 * SP2-Fermi, the adjoint's forward recomputation: PRG_PREC_FP64 (default),
 * PRG_PREC_INT8CRT (exact int8 products modulo prg_get_int8_moduli() moduli,
 * each operand scaled by the power of two above its largest row 2-norm,
 * FP64 reconstruction) or PRG_PREC_BF16SPLIT (BF16 hi+lo split products
 * hi*hi + 2^-8 (hi*lo + lo*hi) with FP32 accumulation, FP64 output: about 16
 * significant bits, so Newton-Schulz and CG tolerances must sit above that
 * noise floor).  Other products stay FP64.  Sets the error state for any
 * other mode.  No counterpart in the reference. */
void prg_set_product_precision(int mode);
int prg_get_product_precision(void);
/* C = A B on device pointers (padded layout, prg_gpu_padded_ld(N)) for
 * symmetric A and B that commute (so that A B is symmetric), in the current
 * product precision; C is exactly symmetric and must not alias A or B.  The
 * product the implicit Fermi recursion issues as bml_multiply(A, B, C, 1, 0,
 * threshold) (src/prg_implicit_fermi_mod.F90:296).  Returns a status. */
int prg_gpu_symmetric_product_device(int N, const double *dA, const double *dB, double *dC);

#ifdef __cplusplus
}
#endif

#endif /* PROGRESS_GPU_H */
