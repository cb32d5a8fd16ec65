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

}  // extern "C"
