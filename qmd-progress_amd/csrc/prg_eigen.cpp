// Eigendecomposition-based density matrices on the GPU (rocSOLVER dsyevd +
// MFMA GEMMs): the reference's diagonalisation routines
// (src/prg_densitymatrix_mod.F90, src/prg_implicit_fermi_mod.F90:1056-1186),
// which callers use as the exact counterpart of the purification solvers.
#include "prg_eigen.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "prg_bml.h"
#include "prg_ops.h"

namespace prg {

namespace {
rocblas_handle g_blas = nullptr;
std::mutex g_blas_mu;

int blas(rocblas_handle* h) {
  std::lock_guard<std::mutex> lk(g_blas_mu);
  if (!g_blas) {
    if (rocblas_create_handle(&g_blas) != rocblas_status_success) {
      set_error(ERR_HIP, "rocblas_create_handle failed");
      return ERR_HIP;
    }
  }
  if (rocblas_set_stream(g_blas, stream()) != rocblas_status_success) {
    set_error(ERR_HIP, "rocblas_set_stream failed");
    return ERR_HIP;
  }
  *h = g_blas;
  return OK;
}
}  // namespace

int eig(const DMat& H, DMat Q, double* d_w, std::vector<double>* h_w) {
  int st = ensure_device();
  if (st) return st;
  // Q = H on the n x n block, zeros elsewhere (the padding stays zero: the
  // solver touches only the n x n block)
  if ((st = ops::lin(H, 1.0, nullptr, 0.0, 0.0, Q))) return st;
  rocblas_handle h;
  if ((st = blas(&h))) return st;
  double* d_e = (double*)workspace("eig_e", sizeof(double) * (size_t)H.ld, &st);
  int* d_info = (int*)workspace("eig_info", sizeof(int) * 4, &st);
  if (st) return st;
  if (rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, H.n, Q.p, Q.ld, d_w, d_e,
                       d_info) != rocblas_status_success) {
    set_error(ERR_HIP, "rocsolver_dsyevd failed");
    return ERR_HIP;
  }
  int info = 0;
  PRG_HIP_CHECK(hipMemcpyAsync(&info, d_info, sizeof(int), hipMemcpyDeviceToHost, stream()));
  if (h_w) {
    h_w->resize(H.n);
    PRG_HIP_CHECK(hipMemcpyAsync(h_w->data(), d_w, sizeof(double) * H.n, hipMemcpyDeviceToHost,
                                 stream()));
  }
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  if (info != 0) {
    set_error(ERR_NUMERIC, "rocsolver_dsyevd did not converge");
    return ERR_NUMERIC;
  }
  return OK;
}

int eig_strided_batched(double* a, int m, int count, double* w) {
  int st = ensure_device();
  if (st) return st;
  if (count <= 0) return OK;
  rocblas_handle h;
  if ((st = blas(&h))) return st;
  double* d_e = (double*)workspace("eigb_e", sizeof(double) * (size_t)m * count, &st);
  int* d_info = (int*)workspace("eigb_info", sizeof(int) * (size_t)count, &st);
  if (st) return st;
  if (rocsolver_dsyevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, m, a, m,
                                       (rocblas_stride)m * m, w, m, d_e, m, d_info,
                                       count) != rocblas_status_success) {
    set_error(ERR_HIP, "rocsolver_dsyevd_strided_batched failed");
    return ERR_HIP;
  }
  std::vector<int> info(count);
  PRG_HIP_CHECK(hipMemcpyAsync(info.data(), d_info, sizeof(int) * count, hipMemcpyDeviceToHost,
                               stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  for (int k = 0; k < count; ++k)
    if (info[k] != 0) {
      set_error(ERR_NUMERIC, "batched eigensolver did not converge");
      return ERR_NUMERIC;
    }
  return OK;
}

int gemm_nt_strided_batched(const double* a, const double* b, double* c, int m, int count) {
  int st = ensure_device();
  if (st) return st;
  if (count <= 0) return OK;
  rocblas_handle h;
  if ((st = blas(&h))) return st;
  const double one = 1.0, zero = 0.0;
  const rocblas_stride s = (rocblas_stride)m * m;
  if (rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_transpose, m, m,
                                    m, &one, a, m, s, b, m, s, &zero, c, m, s,
                                    count) != rocblas_status_success) {
    set_error(ERR_HIP, "rocblas_dgemm_strided_batched failed");
    return ERR_HIP;
  }
  return OK;
}

int density_from_eig(const DMat& Q, const std::vector<double>& occ, DMat out) {
  int st = OK;
  const int n = Q.n;
  double* d_f = (double*)workspace("eig_occ", sizeof(double) * (size_t)Q.ld, &st);
  if (st) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(d_f, occ.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                               stream()));
  DMat QF, QT;
  if ((st = ops::temp("eig_qf", n, &QF)) || (st = ops::temp("eig_qt", n, &QT))) return st;
  if ((st = ops::scale_cols(Q, d_f, QF))) return st;
  if ((st = ops::transpose(Q, QT))) return st;
  // symmetric product of non-symmetric factors: always the FP64 kernel
  if ((st = ops::mm_symmetric_result(QF, QT, out))) return st;
  // the occupation vector must stay alive until the kernels consumed it
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  return OK;
}

int density_fermi(const DMat& H, double beta, double mu, DMat out, std::vector<double>* w_out) {
  int st = OK;
  DMat Q;
  double* d_w = (double*)workspace("eig_w", sizeof(double) * (size_t)H.ld, &st);
  if (st || (st = ops::temp("eig_q", H.n, &Q))) return st;
  std::vector<double> w;
  if ((st = eig(H, Q, d_w, &w))) return st;
  std::vector<double> f(H.n);
  for (int i = 0; i < H.n; ++i) f[i] = 1.0 / (1.0 + std::exp(beta * (w[i] - mu)));
  if (w_out) *w_out = w;
  return density_from_eig(Q, f, out);
}

std::vector<double> t0_occupations(int n, double bndfil) {
  // src/prg_densitymatrix_mod.F90:73-84 (indices 1-based there)
  const double nocc = n * bndfil;
  std::vector<double> occ(n, 0.0);
  for (int i = 1; i <= n; ++i)
    if ((double)i - nocc < 0.0001 || std::fabs((double)i - nocc) < 0.0001) occ[i - 1] = 2.0;
  if (std::fabs(nocc - (int)nocc) > 0.01 && (int)nocc < n) occ[(int)nocc] = 1.0;
  return occ;
}

}  // namespace prg

using namespace prg;

extern "C" {

// src/prg_c_interface.F90:68-86 (src/prg_densitymatrix_mod.F90:36-101)
void prg_build_density_T0(int norbs, bml_matrix_t* ham_bml, bml_matrix_t* rho_bml,
                          double threshold, double bndfil, double* eigenvalues_out) {
  (void)threshold;
  clear_error();
  if (check_pair(ham_bml, rho_bml, "prg_build_density_T0")) return;
  DMat H, R, Q;
  if (bml_dev_read(ham_bml, &H)) return;
  int st = OK;
  double* d_w = (double*)workspace("eig_w", sizeof(double) * (size_t)H.ld, &st);
  if (st || ops::temp("eig_q", H.n, &Q)) return;
  std::vector<double> w;
  if (eig(H, Q, d_w, &w)) return;
  if (eigenvalues_out)
    for (int i = 0; i < H.n && i < norbs; ++i) eigenvalues_out[i] = w[i];
  if (bml_dev_write(rho_bml, &R)) return;
  density_from_eig(Q, t0_occupations(H.n, bndfil), R);
}

// src/prg_c_interface.F90:227-241 (src/prg_densitymatrix_mod.F90:509-582)
void prg_build_density_T_fermi(bml_matrix_t* ham_bml, bml_matrix_t* rho_bml, double threshold,
                               double kbt, double ef, int verbose, double* drho) {
  (void)threshold;
  clear_error();
  if (check_pair(ham_bml, rho_bml, "prg_build_density_T_fermi")) return;
  if (verbose >= 1) std::printf("In get_density_t_Fermi ...\nEf = %g\n", ef);
  DMat H, R, Q;
  if (bml_dev_read(ham_bml, &H)) return;
  int st = OK;
  double* d_w = (double*)workspace("eig_w", sizeof(double) * (size_t)H.ld, &st);
  if (st || ops::temp("eig_q", H.n, &Q)) return;
  std::vector<double> w;
  if (eig(H, Q, d_w, &w)) return;
  std::vector<double> f(H.n), df(H.n);
  double tr_df = 0.0;
  for (int i = 0; i < H.n; ++i) {
    const double x = (w[i] - ef) / kbt;
    f[i] = (x > 100.0) ? 0.0 : 2.0 / (1.0 + std::exp(x));
    double d = 0.0;
    if (x > 100.0) {
      if (std::fabs(w[i] - ef) < 0.001) d = -1000.0;
    } else {
      const double g = 1.0 / (1.0 + std::exp(x));
      d = g * g * std::exp(x) / kbt;
    }
    tr_df += 2.0 * d;  // Tr(Q diag(2 f') Q^T) = sum of 2 f'
  }
  if (drho) *drho = tr_df;
  if (bml_dev_write(rho_bml, &R)) return;
  density_from_eig(Q, f, R);
}

// src/prg_c_interface.F90:281-293 (src/prg_densitymatrix_mod.F90:817-854)
void prg_get_eigenvalues(int norbs, bml_matrix_t* ham_bml, double* eigenvalues_out, int verbose) {
  clear_error();
  if (verbose >= 1) std::printf("In prg_get_eigenvalues ...\n");
  DMat H, Q;
  if (bml_dev_read(ham_bml, &H)) return;
  int st = OK;
  double* d_w = (double*)workspace("eig_w", sizeof(double) * (size_t)H.ld, &st);
  if (st || ops::temp("eig_q", H.n, &Q)) return;
  std::vector<double> w;
  if (eig(H, Q, d_w, &w)) return;
  for (int i = 0; i < H.n && i < norbs; ++i) eigenvalues_out[i] = w[i];
}

// src/prg_c_interface.F90:295-304 (src/prg_densitymatrix_mod.F90:862-880)
void prg_check_idempotency(bml_matrix_t* mat_bml, double threshold, double* idempotency) {
  (void)threshold;
  clear_error();
  DMat A, T;
  if (bml_dev_read(mat_bml, &A) || ops::temp("idem_t", A.n, &T)) return;
  ops::Sums s;
  // A^2 - A  (A symmetric is not assumed: full product)
  if (ops::mm(A, A, T, false, 1.0, -1.0, &A, 0.0, &s)) return;
  if (idempotency) *idempotency = std::sqrt(s.fro2);
}

// prg_get_density_matrix (src/prg_implicit_fermi_mod.F90:1056-1100): no factor 2
void prg_get_density_matrix(bml_matrix_t* ham_bml, bml_matrix_t* p_bml, double beta, double mu,
                            double threshold) {
  (void)threshold;
  clear_error();
  if (check_pair(ham_bml, p_bml, "prg_get_density_matrix")) return;
  DMat H, P;
  if (bml_dev_read(ham_bml, &H) || bml_dev_write(p_bml, &P)) return;
  density_fermi(H, beta, mu, P, nullptr);
}

// src/prg_c_interface.F90:535-549 (src/prg_implicit_fermi_mod.F90:1112-1186).
// The reference passes mu by value through C; prg_test_density_matrix_mu
// below returns the updated chemical potential.
static double g_test_mu = 0.0;
void prg_test_density_matrix(bml_matrix_t* ham_bml, bml_matrix_t* p_bml, double beta, double mu,
                             double nocc, int osteps, double occErrLimit, double threshold) {
  (void)threshold;
  clear_error();
  if (check_pair(ham_bml, p_bml, "prg_test_density_matrix")) return;
  DMat H, P, Q;
  if (bml_dev_read(ham_bml, &H) || bml_dev_write(p_bml, &P)) return;
  int st = OK;
  double* d_w = (double*)workspace("eig_w", sizeof(double) * (size_t)H.ld, &st);
  if (st || ops::temp("eig_q", H.n, &Q)) return;
  std::vector<double> w;
  if (eig(H, Q, d_w, &w)) return;
  double occ_err = 1000.0;
  int iter = 0;
  std::vector<double> f(H.n);
  while ((osteps == 0 && occ_err > occErrLimit) || (osteps > 0 && iter < osteps)) {
    ++iter;
    for (int i = 0; i < H.n; ++i) f[i] = 1.0 / (1.0 + std::exp(beta * (w[i] - mu)));
    if (density_from_eig(Q, f, P)) return;
    ops::Sums s;
    if (ops::sums_of(P, &s)) return;
    const double trp = s.tr;
    const double dpdmu = beta * (trp - s.fro2);
    occ_err = std::fabs(trp - nocc);
    if (occ_err > occErrLimit) mu = mu + (nocc - trp) / dpdmu;
    if (iter > 1000) {
      set_error(ERR_NOT_CONVERGED, "prg_test_density_matrix: chemical potential");
      return;
    }
  }
  g_test_mu = mu;
}

double prg_test_density_matrix_mu(void) { return g_test_mu; }

}  // extern "C"
