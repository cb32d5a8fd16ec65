// Density-matrix perturbation response (src/prg_response_mod.F90).
//
// prg_compute_response_SP2: the SP2 recursion carried together with its
// tangent, Delta <- {X, Delta} = X Delta + Delta X (or 2 Delta - {X, Delta}),
// with the same branch as X.  The anticommutator is one symmetric MFMA GEMM
// over a 2N-deep K (X Delta and Delta X accumulate in the same registers) with
// the branch update fused in its epilogue; X advances through the regular
// fused SP2 step.
//
// prg_canon_response: the implicit-Fermi first-order response in the
// eigenbasis of H0 is elementwise (every (k, j) entry follows its own scalar
// recursion driven by the eigenvalues e_k, e_j), so it runs as one
// elementwise kernel plus the canonical trace correction.
#include <cmath>
#include <vector>

#include "prg_bml.h"
#include "prg_eigen.h"
#include "prg_ops.h"
#include "prg_sp2.h"

namespace prg {
namespace {

constexpr int NT = 256;

// P1[k][j] follows m steps of (src/prg_response_mod.F90:974-1003):
//   dx = p0_k P1 + P1 p0_j ; iD0 = 1/(2(p0^2 - p0) + 1) ; p0 <- iD0 p0^2 ;
//   P1 <- iD0_k (dx + 2 (P1 - dx) p0_j)
// with p0 = 0.5 + cnst (e - mu0) at the start (as the reference initialises).
__global__ void canon_response_kernel(const double* __restrict__ h1, const double* __restrict__ e,
                                      double* __restrict__ p1, int n, int ld, int m, double cnst,
                                      double mu0, double* __restrict__ diag_part) {
  double dsum = 0.0;
  for (int col = blockIdx.x; col < n; col += gridDim.x) {
    for (int row = threadIdx.x; row < n; row += NT) {
      double pk = 0.5 + cnst * (e[row] - mu0);
      double pj = 0.5 + cnst * (e[col] - mu0);
      double v = -cnst * h1[row + (size_t)col * ld];
      for (int i = 0; i < m; ++i) {
        const double dx = pk * v + v * pj;
        const double idk = 1.0 / (2.0 * (pk * pk - pk) + 1.0);
        const double idj = 1.0 / (2.0 * (pj * pj - pj) + 1.0);
        pk = idk * pk * pk;
        pj = idj * pj * pj;
        v = idk * (dx + 2.0 * (v - dx) * pj);
      }
      p1[row + (size_t)col * ld] = v;
      if (row == col) dsum += v;
    }
  }
  __shared__ double red[NT];
  red[threadIdx.x] = dsum;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) diag_part[blockIdx.x] = red[0];
}

__global__ void canon_diag_kernel(double* __restrict__ p1, const double* __restrict__ dpdmu,
                                  int n, int ld, double mu1) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) p1[i + (size_t)i * ld] += mu1 * dpdmu[i];
}

// P1[j][i] = K_ji H1[j][i] with the exact Fermi kernel of
// prg_canon_response_vector (src/prg_response_mod.F90:849-931):
// K_ji = (f_i - f_j) / (e_i - e_j) for i != j, K_ii = df_i; per-block sums of
// the diagonal.  Degenerate pairs divide by zero as in the reference.
__global__ void canon_vector_kernel(const double* __restrict__ h1, const double* __restrict__ e,
                                    const double* __restrict__ f, const double* __restrict__ df,
                                    double* __restrict__ p1, int n, int ld,
                                    double* __restrict__ diag_part) {
  double dsum = 0.0;
  for (int col = blockIdx.x; col < n; col += gridDim.x) {
    for (int row = threadIdx.x; row < n; row += NT) {
      const size_t idx = row + (size_t)col * ld;
      double k;
      if (row == col) {
        k = df[row];
      } else {
        k = (f[col] - f[row]) / (e[col] - e[row]);
      }
      const double v = k * h1[idx];
      p1[idx] = v;
      if (row == col) dsum += v;
    }
  }
  __shared__ double red[NT];
  red[threadIdx.x] = dsum;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) diag_part[blockIdx.x] = red[0];
}

// V[i][j] <- V[i][j] / (e_j - e_i) off the diagonal, 0 on it
// (src/prg_response_mod.F90:303-315).
__global__ void rs_scale_kernel(double* __restrict__ v, const double* __restrict__ e, int n,
                                int ld) {
  for (int col = blockIdx.x; col < n; col += gridDim.x)
    for (int row = threadIdx.x; row < n; row += NT) {
      const size_t idx = row + (size_t)col * ld;
      v[idx] = (row == col) ? 0.0 : v[idx] / (e[col] - e[row]);
    }
}

// The implicit-Fermi response recursion of prg_canon_response on the device
// (P1 before the trace correction), its diagonal sum and dPdmu (host, n).
int canon_recursion(const DMat& H1, DMat P1, const double* evals, int n, int m, double beta,
                    double mu0, double* trace_p1, std::vector<double>& dpdmu) {
  const int ld = H1.ld;
  int st = OK;
  double* d_e = (double*)workspace("canon_e", sizeof(double) * 2 * (size_t)ld, &st);
  double* d_part = (double*)workspace("canon_part", sizeof(double) * 1024, &st);
  if (st) return st;
  const int nb = n < 1024 ? n : 1024;
  PRG_HIP_CHECK(hipMemcpyAsync(d_e, evals, sizeof(double) * n, hipMemcpyHostToDevice, stream()));
  const double cnst = beta / std::pow(2.0, m + 2);
  hipLaunchKernelGGL(canon_response_kernel, dim3(nb), dim3(NT), 0, stream(), H1.p, d_e, P1.p, n,
                     ld, m, cnst, mu0, d_part);
  PRG_HIP_CHECK(hipGetLastError());
  if ((st = ops::fetch(d_part, nb, 1, trace_p1))) return st;
  // dPdmu = beta p0 (1 - p0) with p0 after m steps
  dpdmu.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double p0 = 0.5 + cnst * (evals[i] - mu0);
    for (int k = 0; k < m; ++k) {
      const double id = 1.0 / (2.0 * (p0 * p0 - p0) + 1.0);
      p0 = id * p0 * p0;
    }
    dpdmu[i] = beta * p0 * (1.0 - p0);
  }
  return OK;
}

// p1[i][i] += shift * w[i] (w host, n)
int shift_diagonal(DMat P1, const std::vector<double>& w, double shift) {
  int st = OK;
  double* d_w = (double*)workspace("canon_w", sizeof(double) * (size_t)P1.ld, &st);
  if (st) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(d_w, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice,
                               stream()));
  const int n = (int)w.size();
  hipLaunchKernelGGL(canon_diag_kernel, dim3((n + NT - 1) / NT), dim3(NT), 0, stream(), P1.p, d_w,
                     n, P1.ld, shift);
  PRG_HIP_CHECK(hipGetLastError());
  // the host vector must stay alive until the copy is done
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  return OK;
}

// rho(H + delta H1) at T = 0 with the prg_build_density_T0 occupations
int density_t0_shifted(const DMat& H, const DMat& H1, double delta, double bndfil, DMat out) {
  DMat T, Q;
  int st;
  if ((st = ops::temp("rfd_h", H.n, &T)) || (st = ops::temp("rfd_q", H.n, &Q))) return st;
  double* d_w = (double*)workspace("rfd_w", sizeof(double) * (size_t)H.ld, &st);
  if (st) return st;
  if ((st = ops::lin(H, 1.0, &H1, delta, 0.0, T))) return st;
  if ((st = eig(T, Q, d_w, nullptr))) return st;
  return density_from_eig(Q, t0_occupations(H.n, bndfil), out);
}

}  // namespace
}  // namespace prg

using namespace prg;

extern "C" {

// New C symbol for prg_compute_response_SP2 (src/prg_response_mod.F90:653-782),
// Fortran argument order.  lambda and threshold are accepted and unused, as in
// the reference routine.
void prg_compute_response_SP2(bml_matrix_t* ham_bml, bml_matrix_t* prt_bml, bml_matrix_t* rsp_bml,
                              bml_matrix_t* rho_bml, double lambda, double bndfil, int minsp2iter,
                              int maxsp2iter, const char* sp2conv, double idemtol,
                              double threshold, int verbose) {
  (void)lambda;
  (void)threshold;
  (void)verbose;
  clear_error();
  if (check_pair(ham_bml, prt_bml, "prg_compute_response_SP2") ||
      check_pair(ham_bml, rsp_bml, "prg_compute_response_SP2") ||
      check_pair(ham_bml, rho_bml, "prg_compute_response_SP2"))
    return;
  DMat H, H1, R, R1;
  if (bml_dev_read(ham_bml, &H) || bml_dev_read(prt_bml, &H1) || bml_dev_write(rho_bml, &R) ||
      bml_dev_write(rsp_bml, &R1))
    return;
  Sp2Params p;
  p.variant = V_RESPONSE;
  p.occ = bndfil * (double)H.n;
  p.idemtol = idemtol;
  p.minit = minsp2iter;
  p.maxit = maxsp2iter > kMaxIter - 2 ? kMaxIter - 2 : maxsp2iter;
  p.rel = fortran_rel(sp2conv);
  p.precision = PRG_PREC_FP64;
  Sp2Result res;
  run_sp2(H, p, R, &res, &H1, &R1);
}

// New C symbol for prg_canon_response (src/prg_response_mod.F90:947-1022).
// H1_bml holds the perturbation in the eigenbasis of H0, evals (host, HDIM)
// the eigenvalues of H0; P1_bml receives the first-order response in that
// basis with the canonical trace correction.
void prg_canon_response(bml_matrix_t* P1_bml, bml_matrix_t* H1_bml, double nocc, double beta,
                        const double* evals, double mu0, int m, int HDIM) {
  (void)nocc;
  clear_error();
  if (check_pair(P1_bml, H1_bml, "prg_canon_response")) return;
  if (!evals || HDIM != H1_bml->N || m < 0) {
    set_error(ERR_ARG, "prg_canon_response: bad arguments");
    return;
  }
  DMat H1, P1;
  if (bml_dev_read(H1_bml, &H1) || bml_dev_write(P1_bml, &P1)) return;
  const int n = HDIM, ld = H1.ld;
  int st = OK;
  double* d_e = (double*)workspace("canon_e", sizeof(double) * 2 * (size_t)ld, &st);
  const int nb = n < 1024 ? n : 1024;
  double* d_part = (double*)workspace("canon_part", sizeof(double) * 1024, &st);
  if (st) return;
  if (hipMemcpyAsync(d_e, evals, sizeof(double) * n, hipMemcpyHostToDevice, stream()) !=
      hipSuccess) {
    set_error(ERR_HIP, "prg_canon_response: copy failed");
    return;
  }
  const double cnst = beta / std::pow(2.0, m + 2);
  hipLaunchKernelGGL(canon_response_kernel, dim3(nb), dim3(NT), 0, stream(), H1.p, d_e, P1.p, n,
                     ld, m, cnst, mu0, d_part);
  double tr = 0.0;
  if (ops::fetch(d_part, nb, 1, &tr)) return;
  // dPdmu = beta p0 (1 - p0) with p0 after m steps
  std::vector<double> dp(n);
  double sdp = 0.0;
  for (int i = 0; i < n; ++i) {
    double p0 = 0.5 + cnst * (evals[i] - mu0);
    for (int k = 0; k < m; ++k) {
      const double id = 1.0 / (2.0 * (p0 * p0 - p0) + 1.0);
      p0 = id * p0 * p0;
    }
    dp[i] = beta * p0 * (1.0 - p0);
    sdp += dp[i];
  }
  const double mu1 = -tr / sdp;
  if (hipMemcpyAsync(d_e + ld, dp.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream()) !=
      hipSuccess) {
    set_error(ERR_HIP, "prg_canon_response: copy failed");
    return;
  }
  hipLaunchKernelGGL(canon_diag_kernel, dim3((n + NT - 1) / NT), dim3(NT), 0, stream(), P1.p,
                     d_e + ld, n, ld, mu1);
  if (hipStreamSynchronize(stream()) != hipSuccess) set_error(ERR_HIP, "prg_canon_response");
}

// New C symbol for prg_canon_response_orig (src/prg_response_mod.F90:1038-1105):
// the same recursion and trace correction as prg_canon_response (thresh unused).
void prg_canon_response_orig(bml_matrix_t* P1_bml, bml_matrix_t* H1_bml, double nocc, double beta,
                             const double* evals, double mu0, int m, double thresh, int HDIM) {
  (void)thresh;
  prg_canon_response(P1_bml, H1_bml, nocc, beta, evals, mu0, m, HDIM);
}

// New C symbol for prg_canon_response_p1_dpdmu (src/prg_response_mod.F90:1112-1183):
// the recursion without the trace correction; dPdMu (host, HDIM) receives
// beta p0 (1 - p0).  Norbs and Q_bml are accepted and unused, as in the
// reference routine.
void prg_canon_response_p1_dpdmu(bml_matrix_t* P1_bml, double* dPdMu, bml_matrix_t* H1_bml,
                                 int Norbs, double beta, bml_matrix_t* Q_bml, const double* evals,
                                 double mu0, int m, int HDIM) {
  (void)Norbs;
  (void)Q_bml;
  clear_error();
  if (check_pair(P1_bml, H1_bml, "prg_canon_response_p1_dpdmu")) return;
  if (!evals || !dPdMu || HDIM != H1_bml->N || m < 0) {
    set_error(ERR_ARG, "prg_canon_response_p1_dpdmu: bad arguments");
    return;
  }
  DMat H1, P1;
  if (bml_dev_read(H1_bml, &H1) || bml_dev_write(P1_bml, &P1)) return;
  double tr = 0.0;
  std::vector<double> dp;
  if (canon_recursion(H1, P1, evals, HDIM, m, beta, mu0, &tr, dp)) return;
  for (int i = 0; i < HDIM; ++i) dPdMu[i] = dp[i];
  if (hipStreamSynchronize(stream()) != hipSuccess)
    set_error(ERR_HIP, "prg_canon_response_p1_dpdmu");
}

// New C symbol for prg_canon_response_vector (src/prg_response_mod.F90:849-931):
// the exact Fermi response in the eigenbasis, P1 = K o H1 (element-wise) with
// the canonical diagonal correction.  m and thresh are accepted and unused,
// as in the reference routine.
void prg_canon_response_vector(bml_matrix_t* P1_bml, bml_matrix_t* H1_bml, double nocc,
                               double beta, const double* evals, double mu0, int m, double thresh,
                               int HDIM) {
  (void)nocc;
  (void)m;
  (void)thresh;
  clear_error();
  if (check_pair(P1_bml, H1_bml, "prg_canon_response_vector")) return;
  if (!evals || HDIM != H1_bml->N) {
    set_error(ERR_ARG, "prg_canon_response_vector: bad arguments");
    return;
  }
  DMat H1, P1;
  if (bml_dev_read(H1_bml, &H1) || bml_dev_write(P1_bml, &P1)) return;
  const int n = HDIM, ld = H1.ld;
  std::vector<double> host(3 * (size_t)n);
  double trk = 0.0;
  for (int i = 0; i < n; ++i) {
    host[i] = evals[i];
    host[n + i] = 1.0 / (1.0 + std::exp(beta * (evals[i] - mu0)));
    const double t = std::exp((evals[i] - mu0) * beta);
    host[2 * n + i] = -t / ((1.0 + t) * (1.0 + t)) * beta;
    trk += host[2 * n + i];
  }
  int st = OK;
  double* d_v = (double*)workspace("canonv_vec", sizeof(double) * 3 * (size_t)ld, &st);
  double* d_part = (double*)workspace("canon_part", sizeof(double) * 1024, &st);
  if (st) return;
  if (hipMemcpyAsync(d_v, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice,
                     stream()) != hipSuccess) {
    set_error(ERR_HIP, "prg_canon_response_vector: copy failed");
    return;
  }
  const int nb = n < 1024 ? n : 1024;
  hipLaunchKernelGGL(canon_vector_kernel, dim3(nb), dim3(NT), 0, stream(), H1.p, d_v, d_v + n,
                     d_v + 2 * n, P1.p, n, ld, d_part);
  double trp = 0.0;
  if (ops::fetch(d_part, nb, 1, &trp)) return;
  const double dmu = std::fabs(trk) > 1e-12 ? trp / trk : 0.0;
  std::vector<double> df(host.begin() + 2 * n, host.end());
  shift_diagonal(P1, df, -dmu);
}

// New C symbol for prg_compute_response_RS (src/prg_response_mod.F90:250-361):
// T = 0 first-order response by Rayleigh-Schroedinger perturbation theory in
// the eigenbasis of H (rocSOLVER eigensolve, FP64 MFMA products).  lambda,
// threshold and verbose are accepted and unused (the reference prints two
// traces; this library prints nothing).
void prg_compute_response_RS(bml_matrix_t* ham_bml, bml_matrix_t* prt_bml, bml_matrix_t* rsp_bml,
                             double lambda, double bndfil, double threshold, int verbose) {
  (void)lambda;
  (void)threshold;
  (void)verbose;
  clear_error();
  if (check_pair(ham_bml, prt_bml, "prg_compute_response_RS") ||
      check_pair(ham_bml, rsp_bml, "prg_compute_response_RS"))
    return;
  DMat H, H1, R;
  if (bml_dev_read(ham_bml, &H) || bml_dev_read(prt_bml, &H1) || bml_dev_write(rsp_bml, &R))
    return;
  const int n = H.n;
  DMat Q, QT, X1, V, C1, A, T1, T2;
  if (ops::temp("rrs_q", n, &Q) || ops::temp("rrs_qt", n, &QT) || ops::temp("rrs_x1", n, &X1) ||
      ops::temp("rrs_v", n, &V) || ops::temp("rrs_c1", n, &C1) || ops::temp("rrs_a", n, &A) ||
      ops::temp("rrs_t1", n, &T1) || ops::temp("rrs_t2", n, &T2))
    return;
  int st = OK;
  double* d_w = (double*)workspace("rrs_w", sizeof(double) * (size_t)H.ld, &st);
  double* d_f = (double*)workspace("rrs_f", sizeof(double) * (size_t)H.ld, &st);
  if (st) return;
  if (eig(H, Q, d_w, nullptr) || ops::transpose(Q, QT)) return;
  if (ops::mm(QT, H1, X1, false) || ops::mm(X1, Q, V, false)) return;  // V = Q^T H1 Q
  hipLaunchKernelGGL(rs_scale_kernel, dim3(n < 1024 ? n : 1024), dim3(NT), 0, stream(), V.p, d_w,
                     n, V.ld);
  if (ops::mm(Q, V, C1, false)) return;  // C1 = Q V~
  // occupations: 2 for state i (1-based) <= nocc, 1 for a fractional state
  const double nocc = n * bndfil;
  std::vector<double> occ(n, 0.0);
  for (int i = 1; i <= n; ++i) occ[i - 1] = (i <= nocc) ? 2.0 : 0.0;
  if (std::fabs(nocc - (int)nocc) > 0.01) occ[(int)nocc] = 1.0;
  if (hipMemcpyAsync(d_f, occ.data(), sizeof(double) * n, hipMemcpyHostToDevice, stream()) !=
      hipSuccess) {
    set_error(ERR_HIP, "prg_compute_response_RS: copy failed");
    return;
  }
  if (ops::scale_cols(Q, d_f, A) || ops::transpose(C1, T1) || ops::transpose(A, T2)) return;
  // rho1 = (Q f) C1^T + C1 (Q f)^T
  if (ops::mm2(A, T1, C1, T2, R, false)) return;
  if (hipStreamSynchronize(stream()) != hipSuccess) set_error(ERR_HIP, "prg_compute_response_RS");
}

// New C symbol for prg_compute_response_FD (src/prg_response_mod.F90:380-424):
// (rho(H + delta H1) - rho(H - delta H1)) / (2 delta) with the T = 0 densities
// of prg_build_density_T0.  threshold and verbose are accepted and unused.
void prg_compute_response_FD(bml_matrix_t* ham_bml, bml_matrix_t* prt_bml, bml_matrix_t* rsp_bml,
                             double delta, double bndfil, double threshold, int verbose) {
  (void)threshold;
  (void)verbose;
  clear_error();
  if (check_pair(ham_bml, prt_bml, "prg_compute_response_FD") ||
      check_pair(ham_bml, rsp_bml, "prg_compute_response_FD"))
    return;
  if (!(delta != 0.0)) {
    set_error(ERR_ARG, "prg_compute_response_FD: delta must be non-zero");
    return;
  }
  DMat H, H1, R, RP, RM;
  if (bml_dev_read(ham_bml, &H) || bml_dev_read(prt_bml, &H1) || bml_dev_write(rsp_bml, &R))
    return;
  if (ops::temp("rfd_plus", H.n, &RP) || ops::temp("rfd_minus", H.n, &RM)) return;
  if (density_t0_shifted(H, H1, delta, bndfil, RP) ||
      density_t0_shifted(H, H1, -delta, bndfil, RM))
    return;
  if (ops::lin(RP, 0.5 / delta, &RM, -0.5 / delta, 0.0, R)) return;
  if (hipStreamSynchronize(stream()) != hipSuccess) set_error(ERR_HIP, "prg_compute_response_FD");
}

}  // extern "C"
