// Recursive implicit Fermi-Dirac expansion (src/prg_implicit_fermi_mod.F90)
// on one GPU.
//
// Every matrix product is an FP64 MFMA GEMM (symmetric-tile variant whenever
// both factors are polynomials of the same normalised Hamiltonian, i.e. they
// commute and the product is symmetric); the scalar recurrences -- CG step
// lengths, Newton-Schulz residual norms, the chemical-potential Newton /
// bisection updates -- run on the host from reductions fused into the
// producing kernels, as in the reference.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "prg_bml.h"
#include "prg_eigen.h"
#include "prg_ops.h"

namespace prg {

using ops::Sums;

// prg_conjgrad (src/prg_implicit_fermi_mod.F90:1006-1048): solve A X = B,
// X holding the initial guess, until ||R||_F^2 <= tol.  `commuting`: X0, B
// are polynomials of A (products then symmetric).
int conjgrad(const DMat& A, DMat X, const DMat& B, double tol, bool commuting, int* iters) {
  DMat R, D, W;
  int st;
  ops::HoldOperand hold(A);  // A is not changed here
  if ((st = ops::temp("cg_r", A.n, &R)) || (st = ops::temp("cg_d", A.n, &D)) ||
      (st = ops::temp("cg_w", A.n, &W)))
    return st;
  Sums s;
  // R = B - A X
  if ((st = ops::mm(A, X, R, commuting, -1.0, 1.0, &B, 0.0, &s))) return st;
  double rn = s.fro2, rn_old = rn;
  int k = 0;
  while (rn > tol) {
    ++k;
    if (k == 1) {
      if ((st = ops::copy(R, D))) return st;
    } else {
      if ((st = ops::lin(D, rn / rn_old, &R, 1.0, 0.0, D))) return st;
    }
    if ((st = ops::mm(A, D, W, commuting, 1.0, 0.0, nullptr, 0.0, &s, &D))) return st;
    const double alpha = rn / s.dot;  // <D, A D> = Tr(D^T W)
    if ((st = ops::lin(X, 1.0, &D, alpha, 0.0, X))) return st;
    if ((st = ops::lin(R, 1.0, &W, -alpha, 0.0, R, &s))) return st;
    rn_old = rn;
    rn = s.fro2;
    if (!std::isfinite(rn)) {
      set_error(ERR_NUMERIC, "conjugate gradient: non-finite residual");
      return ERR_NUMERIC;
    }
    if (k > 50) {
      set_error(ERR_NOT_CONVERGED, "Conjugate gradient is not converging");
      return ERR_NOT_CONVERGED;
    }
  }
  if (iters) *iters = k;
  return OK;
}

// prg_newtonschulz (src/prg_implicit_fermi_mod.F90:888-922): Ai <- Ai + Ai R,
// R = I - A Ai, until ||R||_F <= tol; CG fallback when the error grows.
// `commuting`: Ai is a polynomial of A (true of every inverse the implicit
// Fermi recursion refines: A and the initial Ai are functions of H), so both
// products are symmetric and only their upper-triangle tiles are computed.
int newtonschulz(const DMat& A, DMat Ai, double tol, int* iters, bool commuting) {
  DMat R, T, I;
  int st;
  ops::HoldOperand hold(A);  // A is not changed here
  if ((st = ops::temp("ns_r", A.n, &R)) || (st = ops::temp("ns_t", A.n, &T)) ||
      (st = ops::temp("ns_i", A.n, &I)))
    return st;
  double err = 100000.0;
  int i = 0;
  Sums s;
  while (err > tol) {
    if ((st = ops::mm(A, Ai, R, commuting, -1.0, 0.0, nullptr, 1.0, &s))) return st;
    const double prev = err;
    err = std::sqrt(s.fro2);
    if (!std::isfinite(err)) {
      set_error(ERR_NUMERIC, "Newton-Schulz: non-finite residual");
      return ERR_NUMERIC;
    }
    if (prev + 0.01 < err) {
      if ((st = ops::identity(I))) return st;
      if ((st = conjgrad(A, Ai, I, 0.00001, commuting, nullptr))) return st;
    } else {
      if ((st = ops::mm(Ai, R, T, commuting, 1.0, 1.0, &Ai))) return st;
      if ((st = ops::copy(T, Ai))) return st;
    }
    ++i;
    if (i > 200) {
      set_error(ERR_NOT_CONVERGED, "Newton-Schulz is not converging");
      return ERR_NOT_CONVERGED;
    }
  }
  if (iters) *iters = i;
  return OK;
}

// prg_pcg (src/prg_implicit_fermi_mod.F90:934-996): conjugate gradient for
// A X = B preconditioned by M = diag(A)^-1, X holding the initial guess; B is
// overwritten with the residual, as the reference does with p2_bml.  Runs
// while ||R||_F^2 > tol; more than 100 iterations is "not converging".
// Inner products are bml_trace_mult, Tr(X Y).
int pcg(const DMat& A, DMat X, DMat B, double tol, int* iters) {
  DMat Z, D, W;
  int st;
  if ((st = ops::temp("pcg_z", A.n, &Z)) || (st = ops::temp("pcg_d", A.n, &D)) ||
      (st = ops::temp("pcg_w", A.n, &W)))
    return st;
  double* minv = (double*)workspace("pcg_minv", sizeof(double) * (size_t)A.ld, &st);
  if (st) return st;
  if ((st = ops::inv_diagonal(A, minv))) return st;
  Sums s;
  if ((st = ops::mm(A, X, B, false, -1.0, 1.0, &B, 0.0, &s))) return st;  // R = B - A X
  if ((st = ops::scale_rows(B, minv, Z))) return st;                         // Z = M R
  double r_new = 0.0, r_old = 0.0;
  if ((st = ops::trace_mult(Z, B, &r_new))) return st;
  if ((st = ops::copy(Z, D))) return st;
  double rn = s.fro2;
  int k = 0;
  while (rn > tol) {
    ++k;
    if (k != 1 && (st = ops::lin(D, r_new / r_old, &Z, 1.0, 0.0, D))) return st;
    if ((st = ops::mm(A, D, W, false))) return st;
    double num = 0.0, den = 0.0;
    if ((st = ops::trace_mult(B, Z, &num)) || (st = ops::trace_mult(D, W, &den))) return st;
    const double alpha = num / den;
    if ((st = ops::lin(X, 1.0, &D, alpha, 0.0, X))) return st;
    if ((st = ops::lin(B, 1.0, &W, -alpha, 0.0, B, &s))) return st;
    if ((st = ops::scale_rows(B, minv, Z))) return st;
    r_old = r_new;
    if ((st = ops::trace_mult(B, Z, &r_new))) return st;
    rn = s.fro2;
    if (!std::isfinite(rn)) {
      set_error(ERR_NUMERIC, "preconditioned conjugate gradient: non-finite residual");
      return ERR_NUMERIC;
    }
    if (k > 100) {
      set_error(ERR_NOT_CONVERGED, "PCG is not converging");
      return ERR_NOT_CONVERGED;
    }
  }
  if (iters) *iters = k;
  return OK;
}

namespace {

int g_last_occiter = 0, g_last_totns = 0;
double g_last_mu = 0.0;
double g_resp_mu0 = 0.0, g_resp_mu[3] = {0.0, 0.0, 0.0};

// P0 = (0.5 + cnst mu) I - cnst H   (prg_normalize_implicit_fermi)
int normalize_if(const DMat& H, double cnst, double mu, DMat P) {
  return ops::lin(H, -cnst, nullptr, 0.0, 0.5 + cnst * mu, P);
}

}  // namespace

// prg_setup_linsys (src/prg_implicit_fermi_mod.F90:843-876): the linear system
// of one level for even expansion order k, A X_next = B with B = X^k and
// A = X^k + (I - X)^k (k = 2: A = 2(X^2 - X) + I).  Every factor is a
// polynomial of X, so all products are symmetric.
int setup_linsys(const DMat& X, DMat A, DMat B, int k) {
  int st;
  if (k == 2) {
    if ((st = ops::mm(X, X, B, true))) return st;
    return ops::lin(B, 2.0, &X, -2.0, 1.0, A);
  }
  DMat X2, Y, T;
  if ((st = ops::temp("lin_x2", X.n, &X2)) || (st = ops::temp("lin_y", X.n, &Y)) ||
      (st = ops::temp("lin_t", X.n, &T)))
    return st;
  if ((st = ops::mm(X, X, X2, true))) return st;
  if ((st = ops::lin(X2, 1.0, &X, -2.0, 1.0, Y))) return st;  // (I - X)^2
  if ((st = ops::copy(X2, B)) || (st = ops::copy(Y, A))) return st;
  for (int i = 0; i < (k - 2) / 2; ++i) {
    if ((st = ops::mm(B, X2, T, true)) || (st = ops::copy(T, B))) return st;
    if ((st = ops::mm(A, Y, T, true)) || (st = ops::copy(T, A))) return st;
  }
  return ops::lin(A, 1.0, &B, 1.0, 0.0, A);
}

// prg_implicit_fermi (src/prg_implicit_fermi_mod.F90:209-336), even k.
int implicit_fermi(const DMat& H, DMat P, int nsteps, int k, double nocc, double* mu,
                   double beta, int method, int osteps, double occ_err_limit, double tol) {
  if (k < 2 || k % 2 != 0 || k > 64) {
    set_error(ERR_ARG, "prg_implicit_fermi: the expansion order k must be even (2..64)");
    return ERR_ARG;
  }
  if (nsteps < 1 || nsteps > 60) {
    set_error(ERR_ARG, "prg_implicit_fermi: nsteps out of range");
    return ERR_ARG;
  }
  DMat P2, Y, AI, I;
  int st;
  if ((st = ops::temp("if_p2", H.n, &P2)) || (st = ops::temp("if_y", H.n, &Y)) ||
      (st = ops::temp("if_ai", H.n, &AI)) || (st = ops::temp("if_i", H.n, &I)))
    return st;
  if (method == 1) {
    if ((st = ops::identity(AI)) || (st = ops::identity(I))) return st;
  }
  const double cnst = beta / (4.0 * std::pow((double)k, nsteps));
  double occ_err = 10000.0;
  int iter = 0;
  Sums s;
  while ((osteps == 0 && occ_err > occ_err_limit) || (osteps > 0 && iter < osteps)) {
    ++iter;
    if ((st = normalize_if(H, cnst, *mu, P))) return st;
    for (int i = 0; i < nsteps; ++i) {
      // k = 2: P2 = P^2, Y = 2(P2 - P) + I; otherwise P2 = P^k, Y = P^k + (I - P)^k
      if ((st = setup_linsys(P, Y, P2, k))) return st;
      if (method == 0) {
        if ((st = conjgrad(Y, P, P2, tol, true, nullptr))) return st;
      } else {
        // level-1 CG inverse (tol 0.9) refined by Newton-Schulz at every level;
        // the reference comments the refinement out and then diverges
        if (i == 0 && (st = conjgrad(Y, AI, I, 0.9, true, nullptr))) return st;
        if ((st = newtonschulz(Y, AI, tol, nullptr, true))) return st;
        if ((st = ops::mm(AI, P2, P, true))) return st;
      }
    }
    if ((st = ops::sums_of(P, &s))) return st;
    const double trp0 = s.tr;
    const double dpdmu = beta * (trp0 - s.fro2);
    occ_err = std::fabs(trp0 - nocc);
    if (occ_err > occ_err_limit) *mu = *mu + (nocc - trp0) / dpdmu;
    if (iter > 1000) {
      set_error(ERR_NOT_CONVERGED, "prg_implicit_fermi: chemical potential not converging");
      return ERR_NOT_CONVERGED;
    }
  }
  return OK;
}

// prg_implicit_fermi_save_inverse (src/prg_implicit_fermi_mod.F90:46-193).
int implicit_fermi_save_inverse(std::vector<DMat>& inv, const DMat& H, DMat P, int nsteps,
                                double nocc, double* mu, double beta, double occ_err_limit,
                                double tol, int scf_it, int* occiter, int* totns) {
  DMat P2, Y, I;
  int st;
  if ((st = ops::temp("ifs_p2", H.n, &P2)) || (st = ops::temp("ifs_y", H.n, &Y)) ||
      (st = ops::temp("ifs_i", H.n, &I)))
    return st;
  if ((st = ops::identity(I))) return st;
  double occ_err = 10.0, alpha = 1.0;
  int prev = 1, iter = 0, muadj = 0;
  const int maxiter = 30;
  const double cnst = beta / std::pow(2.0, nsteps + 2);
  if (scf_it == 1) {
    // the reference also runs a CG solve here whose result it discards
    // (it resets every inverse to the identity right after); only the reset
    // is observable
    alpha = 4.0;
    for (int i = 0; i < nsteps; ++i)
      if ((st = ops::identity(inv[i]))) return st;
  }
  Sums s;
  int ns_total = 0;
  while ((occ_err > occ_err_limit || muadj == 1) && iter < maxiter) {
    ++iter;
    muadj = 0;
    if ((st = normalize_if(H, cnst, *mu, P))) return st;
    for (int i = 0; i < nsteps; ++i) {
      if ((st = ops::mm(P, P, P2, true))) return st;
      if ((st = ops::lin(P2, 2.0, &P, -2.0, 1.0, Y))) return st;
      // the saved inverses are functions of this H (and commute with Y)
      // unless they come from an earlier SCF iteration's Hamiltonian
      const bool commuting = (scf_it == 1 || iter > 1);
      if (iter == 1 && (st = conjgrad(Y, inv[i], I, 0.001, commuting, nullptr))) return st;
      int nsit = 0;
      if ((st = newtonschulz(Y, inv[i], tol, &nsit, commuting))) return st;
      ns_total += nsit;
      if ((st = ops::mm(inv[i], P2, P, commuting))) return st;
    }
    if ((st = ops::sums_of(P, &s))) return st;
    const double trp0 = s.tr;
    const double dpdmu = beta * (trp0 - s.fro2);
    occ_err = std::fabs(trp0 - nocc);
    if (occ_err > 1.0) {
      if (nocc - trp0 < 0.0 && prev == -1) {
        prev = -1;
      } else if (nocc - trp0 > 0.0 && prev == 1) {
        prev = 1;
      } else if (nocc - trp0 > 0.0 && prev == -1) {
        prev = 1;
        alpha = alpha / 2;
      } else {
        prev = -1;
        alpha = alpha / 2;
      }
      *mu = *mu + prev * alpha;
      muadj = 1;
    } else if (occ_err > occ_err_limit) {
      *mu = *mu + (nocc - trp0) / dpdmu;
      muadj = 1;
    }
  }
  if (occiter) *occiter += iter;
  if (totns) *totns += ns_total;
  if (iter >= maxiter)
    std::printf("Could not converge chemical potential in prg_implicit_fermi_save_inverse\n");
  return ops::lin(P, 2.0, nullptr, 0.0, 0.0, P);
}

// prg_implicit_fermi_zero (src/prg_implicit_fermi_mod.F90:347-425).  Method 1
// follows the intended algorithm (P2 = P^2 before the solve); the reference's
// method-1 branch multiplies by a P2 that it never computed.
int implicit_fermi_zero(const DMat& H, DMat P, int nsteps, double mu, int method, double tol) {
  DMat P2, Y, AI, I;
  int st;
  if ((st = ops::temp("if_p2", H.n, &P2)) || (st = ops::temp("if_y", H.n, &Y)) ||
      (st = ops::temp("if_ai", H.n, &AI)) || (st = ops::temp("if_i", H.n, &I)))
    return st;
  double* d_b = (double*)workspace("ifz_bounds", 2 * sizeof(double), &st);
  if (st || (st = gershgorin(H, d_b))) return st;
  double gb[2];
  PRG_HIP_CHECK(hipMemcpyAsync(gb, d_b, sizeof gb, hipMemcpyDeviceToHost, stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  const double cnst = 0.5 * std::fmin(1.0 / (mu - gb[0]), 1.0 / (gb[1] - mu));
  if ((st = normalize_if(H, cnst, mu, P))) return st;
  if (method == 1 && ((st = ops::identity(AI)) || (st = ops::identity(I)))) return st;
  for (int i = 0; i < nsteps; ++i) {
    if ((st = ops::mm(P, P, P2, true))) return st;
    if ((st = ops::lin(P2, 2.0, &P, -2.0, 1.0, Y))) return st;
    if (method == 0) {
      if ((st = conjgrad(Y, P, P2, tol, true, nullptr))) return st;
    } else {
      if (i == 0 && (st = conjgrad(Y, AI, I, 0.9, true, nullptr))) return st;
      if ((st = ops::mm(AI, P2, P, true))) return st;
    }
  }
  return OK;
}

// prg_implicit_fermi_first_order_response (src/prg_implicit_fermi_mod.F90:439-529).
int implicit_fermi_first_order_response(const DMat& H0, const DMat& H1, DMat P0, DMat P1,
                                        const std::vector<DMat>& inv, int nsteps, double mu0,
                                        double beta) {
  DMat C0, C, B, T;
  int st;
  if ((st = ops::temp("r1_c0", H0.n, &C0)) || (st = ops::temp("r1_c", H0.n, &C)) ||
      (st = ops::temp("r1_b", H0.n, &B)) || (st = ops::temp("r1_t", H0.n, &T)))
    return st;
  const double cnst = beta / std::pow(2.0, 2 + nsteps);
  if ((st = normalize_if(H0, cnst, mu0, P0))) return st;
  if ((st = ops::lin(H1, -cnst, nullptr, 0.0, 0.0, P1))) return st;
  for (int i = 0; i < nsteps; ++i) {
    if ((st = ops::mm(P0, P0, C0, true))) return st;              // C0 = P0^2
    if ((st = ops::mm2(P0, P1, P1, P0, C, true))) return st;      // C = P0 P1 + P1 P0
    if ((st = ops::lin(P1, 2.0, &C, -2.0, 0.0, B))) return st;    // B = 2 (P1 - C)
    if ((st = ops::mm(inv[i], C0, P0, false))) return st;         // P0 = Inv C0
    if ((st = ops::mm(B, P0, T, false, 1.0, 1.0, &C))) return st; // T = C + B P0
    if ((st = ops::mm(inv[i], T, P1, false))) return st;          // P1 = Inv T
  }
  // dPdmu = beta P0 (I - P0); trace correction mu1 = -Tr P1 / Tr dPdmu
  Sums s0, s1;
  if ((st = ops::mm(P0, P0, C0, false, -beta, beta, &P0, 0.0, &s0))) return st;
  if ((st = ops::sums_of(P1, &s1))) return st;
  const double dpdmu_tr = s0.tr;
  if (std::fabs(dpdmu_tr) > 1e-8) {
    const double mu1 = -s1.tr / dpdmu_tr;
    if ((st = ops::lin(P1, 1.0, &C0, mu1, 0.0, P1))) return st;
  }
  return OK;
}

// prg_implicit_fermi_response (src/prg_implicit_fermi_mod.F90:547-712), orders
// 1..3, with its own chemical-potential loop (at most 50 passes).  The level-1
// inverse (CG, tol 0.01) is refined by Newton-Schulz at every level; the
// reference comments that refinement out and its recursion diverges.
int implicit_fermi_response(const DMat& H0, const std::vector<DMat>& Hs, DMat P0,
                            std::vector<DMat>& Ps, int nsteps, double* mu0, double* mu,
                            double beta, double nocc, double occ_tol, double lin_tol, int order) {
  if (order < 1 || order > 3) {
    set_error(ERR_ARG, "prg_implicit_fermi_response: order must be 1..3");
    return ERR_ARG;
  }
  const int n = H0.n;
  DMat C0, T, Ti, I, Cm[3], Bm[3], Tmp;
  int st;
  if ((st = ops::temp("rs_c0", n, &C0)) || (st = ops::temp("rs_t", n, &T)) ||
      (st = ops::temp("rs_ti", n, &Ti)) || (st = ops::temp("rs_i", n, &I)) ||
      (st = ops::temp("rs_tmp", n, &Tmp)))
    return st;
  const char* cn[3] = {"rs_c1", "rs_c2", "rs_c3"};
  const char* bn[3] = {"rs_b1", "rs_b2", "rs_b3"};
  for (int j = 0; j < order; ++j)
    if ((st = ops::temp(cn[j], n, &Cm[j])) || (st = ops::temp(bn[j], n, &Bm[j]))) return st;
  if ((st = ops::identity(I)) || (st = ops::lin(I, 0.0, nullptr, 0.0, 0.0, Ti))) return st;
  const double cnst = beta / std::pow(2.0, 2 + nsteps);
  for (int j = 0; j < order; ++j) mu[j] = 0.0;
  double occ_err = 10000.0;
  int k = 0;
  while (occ_err > occ_tol) {
    ++k;
    if ((st = normalize_if(H0, cnst, *mu0, P0))) return st;
    for (int j = 0; j < order; ++j)  // P(j) = -cnst (H(j) - mu(j) I)
      if ((st = ops::lin(Hs[j], -cnst, nullptr, 0.0, cnst * mu[j], Ps[j]))) return st;
    for (int i = 0; i < nsteps; ++i) {
      if ((st = ops::mm(P0, P0, C0, true))) return st;
      if ((st = ops::mm2(P0, Ps[0], Ps[0], P0, Cm[0], true))) return st;
      if ((st = ops::lin(Ps[0], 2.0, &Cm[0], -2.0, 0.0, Bm[0]))) return st;
      if (order > 1) {
        if ((st = ops::mm2(P0, Ps[1], Ps[1], P0, Tmp, true))) return st;
        if ((st = ops::mm(Ps[0], Ps[0], Cm[1], true, 1.0, 1.0, &Tmp))) return st;
        if ((st = ops::lin(Ps[1], 2.0, &Cm[1], -2.0, 0.0, Bm[1]))) return st;
      }
      if (order > 2) {
        if ((st = ops::mm2(Ps[0], Ps[1], Ps[1], Ps[0], Tmp, true))) return st;
        if ((st = ops::mm2(P0, Ps[2], Ps[2], P0, Cm[2], true, 1.0, 1.0, &Tmp))) return st;
        if ((st = ops::lin(Ps[2], 2.0, &Cm[2], -2.0, 0.0, Bm[2]))) return st;
      }
      if ((st = ops::lin(C0, 2.0, &P0, -2.0, 1.0, T))) return st;  // T = 2 P0^2 - 2 P0 + I
      if (i == 0 && (st = conjgrad(T, Ti, I, 0.01, false, nullptr))) return st;
      if ((st = newtonschulz(T, Ti, lin_tol, nullptr, false))) return st;
      if ((st = ops::mm(Ti, C0, P0, false))) return st;
      if ((st = ops::mm(Bm[0], P0, Tmp, false, 1.0, 1.0, &Cm[0]))) return st;
      if ((st = ops::mm(Ti, Tmp, Ps[0], false))) return st;
      if (order > 1) {
        if ((st = ops::mm(Bm[1], P0, Cm[1], false, 1.0, 1.0, &Cm[1]))) return st;
        if ((st = ops::mm(Bm[0], Ps[0], Tmp, false, 1.0, 1.0, &Cm[1]))) return st;
        if ((st = ops::mm(Ti, Tmp, Ps[1], false))) return st;
      }
      if (order > 2) {
        if ((st = ops::mm(Bm[2], P0, Cm[2], false, 1.0, 1.0, &Cm[2]))) return st;
        if ((st = ops::mm(Bm[1], Ps[0], Cm[2], false, 1.0, 1.0, &Cm[2]))) return st;
        if ((st = ops::mm(Bm[0], Ps[1], Tmp, false, 1.0, 1.0, &Cm[2]))) return st;
        if ((st = ops::mm(Ti, Tmp, Ps[2], false))) return st;
      }
    }
    // Pmu = beta P0 (I - P0)
    Sums sm, s0;
    if ((st = ops::mm(P0, P0, Tmp, false, -beta, beta, &P0, 0.0, &sm))) return st;
    if ((st = ops::sums_of(P0, &s0))) return st;
    occ_err = std::fabs(s0.tr - nocc);
    *mu0 = *mu0 + (nocc - s0.tr) / sm.tr;
    for (int j = 0; j < order; ++j) {
      Sums sj;
      if ((st = ops::sums_of(Ps[j], &sj))) return st;
      mu[j] = mu[j] - sj.tr / sm.tr;
      occ_err += std::fabs(sj.tr);
    }
    if (k > 50) {
      std::printf("Chemical potential is not converging\n");
      break;
    }
  }
  return OK;
}

// ---------------------------------------------------------------- adjoint
// Fixed-mu forward of prg_implicit_fermi, method 1, k = 2
// (src/prg_implicit_fermi_mod.F90:259-298 for one mu, the same operations as
// implicit_fermi), keeping every level: X[0] = P0, X[i + 1] = the P after
// level i + 1, Inv[i] = that level's converged inverse of Y = 2(X^2 - X) + I.
int implicit_fermi_levels(const DMat& H, int nsteps, double mu, double beta, double tol,
                          std::vector<DMat>& X, std::vector<DMat>& Inv) {
  if (nsteps < 1 || nsteps > 60) {
    set_error(ERR_ARG, "implicit Fermi adjoint: nsteps out of range");
    return ERR_ARG;
  }
  int st;
  X.resize(nsteps + 1);
  Inv.resize(nsteps);
  for (int i = 0; i <= nsteps; ++i)
    if ((st = ops::temp(("ifa_x" + std::to_string(i)).c_str(), H.n, &X[i]))) return st;
  for (int i = 0; i < nsteps; ++i)
    if ((st = ops::temp(("ifa_inv" + std::to_string(i)).c_str(), H.n, &Inv[i]))) return st;
  DMat P2, Y, AI, I;
  if ((st = ops::temp("if_p2", H.n, &P2)) || (st = ops::temp("if_y", H.n, &Y)) ||
      (st = ops::temp("if_ai", H.n, &AI)) || (st = ops::temp("if_i", H.n, &I)))
    return st;
  if ((st = ops::identity(AI)) || (st = ops::identity(I))) return st;
  const double cnst = beta / (4.0 * std::pow(2.0, nsteps));
  if ((st = normalize_if(H, cnst, mu, X[0]))) return st;
  for (int i = 0; i < nsteps; ++i) {
    if ((st = ops::mm(X[i], X[i], P2, true))) return st;
    if ((st = ops::lin(P2, 2.0, &X[i], -2.0, 1.0, Y))) return st;
    if (i == 0 && (st = conjgrad(Y, AI, I, 0.9, true, nullptr))) return st;
    if ((st = newtonschulz(Y, AI, tol, nullptr, true))) return st;
    if ((st = ops::copy(AI, Inv[i]))) return st;
    if ((st = ops::mm(AI, P2, X[i + 1], true))) return st;
  }
  return OK;
}

// G = dTr(P A)/dH for the fixed-mu recursion above (A symmetric): the reverse
// sweep.  Level i maps X to F = Y^-1 X^2; with W = Y^-1 Gb and U = W - 2 W F
// the adjoint of X is X Us + Us X + W F + (W F)^T, Us = (U + U^T) / 2 (the
// symmetric part: only symmetric perturbations of H are meaningful), and
// X0 = (1/2 + c mu) I - c H gives G = -c Gb0.  canonical: the derivative at
// fixed occupation with the reference's dP/dmu ~ beta P (I - P)
// (src/prg_implicit_fermi_mod.F90:512-521), i.e. A replaced by A - kappa I,
// kappa = Tr(A beta P (I - P)) / Tr(beta P (I - P)).  Every level's X and
// inverse stay in HBM: (2 nsteps + 1) N^2 doubles.  P (optional) receives
// the forward result.
int implicit_fermi_adjoint(const DMat& H, const DMat& A, DMat G, int nsteps, double mu,
                           double beta, double tol, bool canonical, DMat* P) {
  std::vector<DMat> X, Inv;
  int st;
  if ((st = implicit_fermi_levels(H, nsteps, mu, beta, tol, X, Inv))) return st;
  DMat Gb, W, WF, U, Ut, V, T1, T2;
  if ((st = ops::temp("ifa_gb", H.n, &Gb)) || (st = ops::temp("ifa_w", H.n, &W)) ||
      (st = ops::temp("ifa_wf", H.n, &WF)) || (st = ops::temp("ifa_u", H.n, &U)) ||
      (st = ops::temp("ifa_ut", H.n, &Ut)) || (st = ops::temp("ifa_v", H.n, &V)) ||
      (st = ops::temp("ifa_t1", H.n, &T1)) || (st = ops::temp("ifa_t2", H.n, &T2)))
    return st;
  if ((st = ops::copy(A, Gb))) return st;
  const DMat& Pn = X[nsteps];
  if (canonical) {
    ops::Sums sa, sx2, sx;
    if ((st = ops::mm(A, Pn, T1, false, 1.0, 0.0, nullptr, 0.0, &sa))) return st;  // Tr(A P)
    if ((st = ops::mm(Pn, Pn, T2, false, 1.0, 0.0, nullptr, 0.0, &sx2, &A))) return st;
    if ((st = ops::sums_of(Pn, &sx))) return st;
    const double trcm = beta * (sx.tr - sx2.tr);  // Tr(beta P (I - P))
    if (std::fabs(trcm) > 1e-8) {
      const double kappa = beta * (sa.tr - sx2.dot) / trcm;
      if ((st = ops::lin(A, 1.0, nullptr, 0.0, -kappa, Gb))) return st;
    }
  }
  for (int i = nsteps - 1; i >= 0; --i) {
    if ((st = ops::mm(Inv[i], Gb, W, false))) return st;          // W = Y^-1 Gb
    if ((st = ops::mm(W, X[i + 1], WF, false))) return st;        // W F
    if ((st = ops::lin(W, 1.0, &WF, -2.0, 0.0, U))) return st;    // U = W - 2 W F
    if ((st = ops::transpose(U, Ut))) return st;
    if ((st = ops::lin(U, 0.5, &Ut, 0.5, 0.0, U))) return st;     // Us
    if ((st = ops::mm(X[i], U, V, false))) return st;             // X Us
    if ((st = ops::transpose(V, Ut))) return st;
    if ((st = ops::lin(V, 1.0, &Ut, 1.0, 0.0, T1))) return st;    // X Us + Us X
    if ((st = ops::transpose(WF, Ut))) return st;
    if ((st = ops::lin(WF, 1.0, &Ut, 1.0, 0.0, T2))) return st;   // W F + (W F)^T
    if ((st = ops::lin(T1, 1.0, &T2, 1.0, 0.0, Gb))) return st;
  }
  const double cnst = beta / (4.0 * std::pow(2.0, nsteps));
  if ((st = ops::lin(Gb, -cnst, nullptr, 0.0, 0.0, G))) return st;
  if (P && (st = ops::copy(Pn, *P))) return st;
  return OK;
}

}  // namespace prg

using namespace prg;

namespace {
bool dev_inv(bml_matrix_t** inv, int nsteps, std::vector<DMat>& out, bool write) {
  out.resize(nsteps);
  for (int i = 0; i < nsteps; ++i) {
    if (!inv || !inv[i]) {
      set_error(ERR_ARG, "null inverse handle");
      return false;
    }
    if (bml_dev_read(inv[i], &out[i])) return false;
    if (write) inv[i]->host_valid = false;
  }
  return true;
}
}  // namespace

extern "C" {

// New C symbols for the linear solvers of src/prg_implicit_fermi_mod.F90, with
// the Fortran argument order; the auxiliary matrices are accepted and unused
// (the library keeps its own).  prg_conjgrad (:1006-1048) and prg_pcg
// (:934-996) solve A X = B (X in p_bml, initial guess on entry); prg_pcg
// overwrites p2_bml with the residual as the reference does.  prg_newtonschulz
// (:888-922) refines ai_bml towards A^-1 until ||I - A Ai||_F <= tol (CG
// fallback when the error grows) and reports the iteration count.
void prg_conjgrad(bml_matrix_t* A_bml, bml_matrix_t* p_bml, bml_matrix_t* p2_bml,
                  bml_matrix_t* tmp_bml, bml_matrix_t* d_bml, bml_matrix_t* w_bml, double cg_tol,
                  double threshold) {
  (void)tmp_bml;
  (void)d_bml;
  (void)w_bml;
  (void)threshold;
  clear_error();
  if (check_pair(A_bml, p_bml, "prg_conjgrad") || check_pair(A_bml, p2_bml, "prg_conjgrad")) return;
  DMat A, X, B;
  if (bml_dev_read(A_bml, &A) || bml_dev_read(p2_bml, &B) || bml_dev_read(p_bml, &X)) return;
  p_bml->host_valid = false;
  conjgrad(A, X, B, cg_tol, false, nullptr);
}

void prg_pcg(bml_matrix_t* A_bml, bml_matrix_t* p_bml, bml_matrix_t* p2_bml, bml_matrix_t* d_bml,
             bml_matrix_t* wtmp_bml, double cg_tol, double threshold) {
  (void)d_bml;
  (void)wtmp_bml;
  (void)threshold;
  clear_error();
  if (check_pair(A_bml, p_bml, "prg_pcg") || check_pair(A_bml, p2_bml, "prg_pcg")) return;
  DMat A, X, B;
  if (bml_dev_read(A_bml, &A) || bml_dev_read(p_bml, &X) || bml_dev_read(p2_bml, &B)) return;
  p_bml->host_valid = false;
  p2_bml->host_valid = false;
  pcg(A, X, B, cg_tol, nullptr);
}

void prg_newtonschulz(bml_matrix_t* a_bml, bml_matrix_t* ai_bml, bml_matrix_t* r_bml,
                      bml_matrix_t* tmp_bml, bml_matrix_t* d_bml, bml_matrix_t* I_bml, double tol,
                      double threshold, int* num_iter) {
  (void)r_bml;
  (void)tmp_bml;
  (void)d_bml;
  (void)I_bml;
  (void)threshold;
  clear_error();
  if (check_pair(a_bml, ai_bml, "prg_newtonschulz")) return;
  DMat A, Ai;
  if (bml_dev_read(a_bml, &A) || bml_dev_read(ai_bml, &Ai)) return;
  ai_bml->host_valid = false;
  int it = 0;
  if (newtonschulz(A, Ai, tol, &it, false)) return;
  if (num_iter) *num_iter = it;
}

// prg_finite_diff (src/prg_implicit_fermi_mod.F90:726-831), extended form:
// finite-difference response of prg_get_density_matrix along the series
// H(l) = H0 + sum_i l^i H_i, mu(l) = mu0 + sum_i l^i mu_i at l = lambda with
// step h: order 1 the forward difference (D(l + h) - D(l)) / h, order 2
// (D(l + h) - 2 D(l) + D(l - h)) / (2 h^2), order 3 (D(l + 2h) - D(l - 2h)) /
// (12 h^3) - (D(l + h) - D(l - h)) / (6 h^3).  The reference only prints them
// (top-left 10 x 10 times 1000); here out_list[k] (order k + 1) receives each
// when out_list is given, and with out_list NULL the same corners are printed.
void prg_gpu_finite_diff(bml_matrix_t* H0_bml, bml_matrix_t** H_list, double mu0,
                         const double* mu_list, double beta, int order, double lambda, double h,
                         double threshold, bml_matrix_t** out_list) {
  (void)threshold;
  clear_error();
  if (!H0_bml || !H_list || !mu_list || order < 1 || order > 3 || !(h != 0.0)) {
    set_error(ERR_ARG, "prg_finite_diff: bad arguments");
    return;
  }
  DMat H0;
  if (bml_dev_read(H0_bml, &H0)) return;
  std::vector<DMat> Hs(order);
  for (int i = 0; i < order; ++i) {
    if (check_pair(H0_bml, H_list[i], "prg_finite_diff") || bml_dev_read(H_list[i], &Hs[i])) return;
  }
  const int n = H0.n;
  DMat T, D0, D1p, D1m, D2p, D2m, F;
  if (ops::temp("fdf_t", n, &T) || ops::temp("fdf_d0", n, &D0) || ops::temp("fdf_d1p", n, &D1p) ||
      ops::temp("fdf_d1m", n, &D1m) || ops::temp("fdf_d2p", n, &D2p) ||
      ops::temp("fdf_d2m", n, &D2m) || ops::temp("fdf_f", n, &F))
    return;
  auto density_at = [&](double l, DMat out) -> int {
    int st;
    double mu = mu0, lp = 1.0;
    if ((st = ops::copy(H0, T))) return st;
    for (int i = 0; i < order; ++i) {
      lp *= l;
      if ((st = ops::lin(T, 1.0, &Hs[i], lp, 0.0, T))) return st;
      mu += lp * mu_list[i];
    }
    return density_fermi(T, beta, mu, out, nullptr);
  };
  auto emit = [&](int k, const DMat& src) -> int {
    if (out_list) {
      if (!out_list[k] || out_list[k]->N != n) {
        set_error(ERR_ARG, "prg_finite_diff: bad output handle");
        return ERR_ARG;
      }
      DMat o;
      if (bml_dev_write(out_list[k], &o)) return last_error_code();
      return ops::copy(src, o);
    }
    const int c = n < 10 ? n : 10;
    std::vector<double> blk((size_t)c * c);
    PRG_HIP_CHECK(hipMemcpy2DAsync(blk.data(), sizeof(double) * c, src.p, sizeof(double) * src.ld,
                                   sizeof(double) * c, c, hipMemcpyDeviceToHost, stream()));
    PRG_HIP_CHECK(hipStreamSynchronize(stream()));
    std::printf("Finite diff - Order %d * 1000\n", k + 1);
    for (int r = 0; r < c; ++r) {
      for (int q = 0; q < c; ++q) std::printf(" %12.6f", 1000.0 * blk[r + (size_t)q * c]);
      std::printf("\n");
    }
    return OK;
  };
  const double l = lambda;
  if (density_at(l, D0) || density_at(l + h, D1p)) return;
  if (ops::lin(D1p, 1.0 / h, &D0, -1.0 / h, 0.0, F) || emit(0, F)) return;
  if (order > 1) {
    if (density_at(l - h, D1m)) return;
    if (ops::lin(D0, -1.0 / (h * h), &D1m, 0.5 / (h * h), 0.0, F) ||
        ops::lin(F, 1.0, &D1p, 0.5 / (h * h), 0.0, F) || emit(1, F))
      return;
  }
  if (order > 2) {
    if (density_at(l - 2.0 * h, D2m) || density_at(l + 2.0 * h, D2p)) return;
    const double h3 = h * h * h;
    if (ops::lin(D2p, 1.0 / (12.0 * h3), &D2m, -1.0 / (12.0 * h3), 0.0, F) ||
        ops::lin(F, 1.0, &D1p, -1.0 / (6.0 * h3), 0.0, F) ||
        ops::lin(F, 1.0, &D1m, 1.0 / (6.0 * h3), 0.0, F) || emit(2, F))
      return;
  }
  if (hipStreamSynchronize(stream()) != hipSuccess) set_error(ERR_HIP, "prg_finite_diff");
}

// prg_finite_diff at the arity and types of its declaration in the
// reference's C header (src/prg_progress_mod.h:259-268): one perturbation
// handle H_list (= H1) and one chemical-potential coefficient mu_list (= mu1)
// by value, so only order 1 can be carried; the difference is printed as the
// reference does (top-left 10 x 10 times 1000).  Orders 2 and 3 need the
// lists of prg_gpu_finite_diff (an argument error here).
void prg_finite_diff(bml_matrix_t* H0_bml, bml_matrix_t* H_list, double mu0, double mu_list,
                     double beta, int order, double lambda, double h, double threshold) {
  clear_error();
  if (order != 1 || !H_list) {
    set_error(ERR_ARG, "prg_finite_diff: the reference C prototype carries order 1 only "
                       "(use prg_gpu_finite_diff for orders 2-3)");
    return;
  }
  bml_matrix_t* hl[1] = {H_list};
  const double ml[1] = {mu_list};
  prg_gpu_finite_diff(H0_bml, hl, mu0, ml, beta, 1, lambda, h, threshold, nullptr);
}

// prg_implicit_fermi at a fixed mu (one pass of src/prg_implicit_fermi_mod.F90:
// 259-298, method 1) on device pointers (padded layout): dP = P (not scaled
// by 2).
int prg_gpu_implicit_fermi_device(int N, const double* dH, double* dP, int nsteps, double mu,
                                  double beta, double tol) {
  clear_error();
  if (N <= 0 || !dH || !dP || dP == dH) {
    set_error(ERR_ARG, "prg_gpu_implicit_fermi_device: bad arguments");
    return ERR_ARG;
  }
  int st = ensure_device();
  if (st) return st;
  const int ld = padded(N);
  DMat H{N, ld, const_cast<double*>(dH)}, P{N, ld, dP};
  std::vector<DMat> X, Inv;
  if ((st = implicit_fermi_levels(H, nsteps, mu, beta, tol, X, Inv))) return st;
  if ((st = ops::copy(X[nsteps], P))) return st;
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  return OK;
}

int prg_gpu_implicit_fermi_adjoint_device(int N, const double* dH, const double* dA, double* dG,
                                          int nsteps, double mu, double beta, double tol,
                                          int canonical, double* dP) {
  clear_error();
  if (N <= 0 || !dH || !dA || !dG || dG == dH || dG == dA || (dP && (dP == dH || dP == dA))) {
    set_error(ERR_ARG, "prg_gpu_implicit_fermi_adjoint_device: bad arguments");
    return ERR_ARG;
  }
  int st = ensure_device();
  if (st) return st;
  const int ld = padded(N);
  DMat H{N, ld, const_cast<double*>(dH)}, A{N, ld, const_cast<double*>(dA)}, G{N, ld, dG};
  DMat P{N, ld, dP};
  if ((st = implicit_fermi_adjoint(H, A, G, nsteps, mu, beta, tol, canonical != 0,
                                   dP ? &P : nullptr)))
    return st;
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  return OK;
}

// src/prg_c_interface.F90:470-490 (mu by value, as the reference's C path)
void prg_implicit_fermi(bml_matrix_t* h_bml, bml_matrix_t* p_bml, int nsteps, int k, double nocc,
                        double mu, double beta, int method, int osteps, double occErrLimit,
                        double threshold, double tol) {
  (void)threshold;
  clear_error();
  if (check_pair(h_bml, p_bml, "prg_implicit_fermi")) return;
  DMat H, P;
  if (bml_dev_read(h_bml, &H) || bml_dev_write(p_bml, &P)) return;
  double m = mu;
  if (implicit_fermi(H, P, nsteps, k, nocc, &m, beta, method, osteps, occErrLimit, tol)) return;
  g_last_mu = m;
}

// Chemical potential after the last prg_implicit_fermi call (the reference's
// C wrapper passes mu by value and loses the update).
double prg_implicit_fermi_last_mu(void) { return g_last_mu; }

// src/prg_c_interface.F90:440-468; occiter / totns by value as in the
// reference header; prg_implicit_fermi_counters returns the updated counts.
void prg_implicit_fermi_save_inverse(bml_matrix_t** Inv_bml, bml_matrix_t* h_bml,
                                     bml_matrix_t* p_bml, int nsteps, double nocc, double* mu,
                                     double beta, double occErrLimit, double threshold, double tol,
                                     int SCF_IT, int occiter, int totns) {
  (void)threshold;
  clear_error();
  if (!mu) {
    set_error(ERR_ARG, "prg_implicit_fermi_save_inverse: null mu");
    return;
  }
  if (check_pair(h_bml, p_bml, "prg_implicit_fermi_save_inverse")) return;
  DMat H, P;
  std::vector<DMat> inv;
  if (bml_dev_read(h_bml, &H) || bml_dev_write(p_bml, &P)) return;
  if (!dev_inv(Inv_bml, nsteps, inv, true)) return;
  int oc = occiter, tn = totns;
  if (implicit_fermi_save_inverse(inv, H, P, nsteps, nocc, mu, beta, occErrLimit, tol, SCF_IT,
                                  &oc, &tn))
    return;
  g_last_occiter = oc;
  g_last_totns = tn;
}

void prg_implicit_fermi_counters(int* occiter, int* totns) {
  if (occiter) *occiter = g_last_occiter;
  if (totns) *totns = g_last_totns;
}

// src/prg_c_interface.F90:492-506
void prg_implicit_fermi_zero(bml_matrix_t* h_bml, bml_matrix_t* p_bml, int nsteps, double mu,
                             int method, double threshold, double tol) {
  (void)threshold;
  clear_error();
  if (check_pair(h_bml, p_bml, "prg_implicit_fermi_zero")) return;
  DMat H, P;
  if (bml_dev_read(h_bml, &H) || bml_dev_write(p_bml, &P)) return;
  implicit_fermi_zero(H, P, nsteps, mu, method, tol);
}

// src/prg_c_interface.F90:508-533
void prg_implicit_fermi_first_order_response(bml_matrix_t* H0_bml, bml_matrix_t* H1_bml,
                                             bml_matrix_t* P0_bml, bml_matrix_t* P1_bml,
                                             bml_matrix_t** Inv_bml, int nsteps, double mu0,
                                             double beta, double nocc, double threshold) {
  (void)threshold;
  (void)nocc;
  clear_error();
  if (check_pair(H0_bml, H1_bml, "first_order_response") ||
      check_pair(H0_bml, P0_bml, "first_order_response") ||
      check_pair(H0_bml, P1_bml, "first_order_response"))
    return;
  DMat H0, H1, P0, P1;
  std::vector<DMat> inv;
  if (bml_dev_read(H0_bml, &H0) || bml_dev_read(H1_bml, &H1) || bml_dev_write(P0_bml, &P0) ||
      bml_dev_write(P1_bml, &P1))
    return;
  if (!dev_inv(Inv_bml, nsteps, inv, false)) return;
  implicit_fermi_first_order_response(H0, H1, P0, P1, inv, nsteps, mu0, beta);
}

// prg_implicit_fermi_response (src/prg_implicit_fermi_mod.F90:547-712), the
// extended form: H_bml[0..order-1] are H1..H3, P_bml the outputs P1..P3; *mu0
// is updated in place and mu[0..order-1] receives the response chemical
// potentials.
void prg_gpu_implicit_fermi_response(bml_matrix_t* H0_bml, bml_matrix_t** H_bml,
                                     bml_matrix_t* P0_bml, bml_matrix_t** P_bml, int nsteps,
                                     double* mu0, double* mu, double beta, double nocc,
                                     double occ_tol, double lin_tol, int order,
                                     double threshold) {
  (void)threshold;
  clear_error();
  if (!mu0 || !mu || !H_bml || !P_bml || order < 1 || order > 3) {
    set_error(ERR_ARG, "prg_implicit_fermi_response: bad arguments");
    return;
  }
  DMat H0, P0;
  if (bml_dev_read(H0_bml, &H0) || bml_dev_write(P0_bml, &P0)) return;
  std::vector<DMat> Hs(order), Ps(order);
  for (int j = 0; j < order; ++j) {
    if (check_pair(H0_bml, H_bml[j], "response") || check_pair(H0_bml, P_bml[j], "response"))
      return;
    if (bml_dev_read(H_bml[j], &Hs[j]) || bml_dev_write(P_bml[j], &Ps[j])) return;
  }
  implicit_fermi_response(H0, Hs, P0, Ps, nsteps, mu0, mu, beta, nocc, occ_tol, lin_tol, order);
}

// Declared at src/prg_progress_mod.h:240-258 (no implementation in the
// reference's C interface): the header's arity, H1..H3 and P1..P3 as separate
// handles (H2, H3, P2, P3 may be NULL above `order`), mu0 and mu by value as
// declared.  The Fortran routine sets mu(:) to zero on entry, so the input mu
// is unused; the updated chemical potentials are returned by
// prg_implicit_fermi_response_mu().
void prg_implicit_fermi_response(bml_matrix_t* H0_bml, bml_matrix_t* H1_bml, bml_matrix_t* H2_bml,
                                 bml_matrix_t* H3_bml, bml_matrix_t* P0_bml, bml_matrix_t* P1_bml,
                                 bml_matrix_t* P2_bml, bml_matrix_t* P3_bml, int nsteps, double mu0,
                                 double mu, double beta, double nocc, double occ_tol,
                                 double lin_tol, int order, double threshold) {
  (void)mu;
  clear_error();
  bml_matrix_t* hs[3] = {H1_bml, H2_bml, H3_bml};
  bml_matrix_t* ps[3] = {P1_bml, P2_bml, P3_bml};
  if (order < 1 || order > 3) {
    set_error(ERR_ARG, "prg_implicit_fermi_response: order must be 1..3");
    return;
  }
  for (int j = 0; j < order; ++j)
    if (!hs[j] || !ps[j]) {
      set_error(ERR_ARG, "prg_implicit_fermi_response: missing H or P handle below order");
      return;
    }
  double m0 = mu0, mr[3] = {0.0, 0.0, 0.0};
  prg_gpu_implicit_fermi_response(H0_bml, hs, P0_bml, ps, nsteps, &m0, mr, beta, nocc, occ_tol,
                                  lin_tol, order, threshold);
  if (last_error_code() != OK) return;
  g_resp_mu0 = m0;
  for (int j = 0; j < 3; ++j) g_resp_mu[j] = mr[j];
}

// Chemical potentials after the last prg_implicit_fermi_response call (the
// declared C prototype passes mu0 and mu by value): *mu0 and mu[0..2].
void prg_implicit_fermi_response_mu(double* mu0, double* mu) {
  if (mu0) *mu0 = g_resp_mu0;
  if (mu)
    for (int j = 0; j < 3; ++j) mu[j] = g_resp_mu[j];
}

}  // extern "C"
