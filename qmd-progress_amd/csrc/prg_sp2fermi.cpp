// Truncated (finite-temperature) SP2: prg_sp2_fermi_init,
// prg_sp2_fermi_init_norecs and prg_sp2_fermi (src/prg_sp2_fermi_mod.F90),
// plus the host-side entropy helpers of the same module, and the congruence
// transforms callers apply around the solvers (src/prg_nonortho_mod.F90).
//
// The X^2 products, the tangent X X1 + X1 X and the traces run on the GPU
// through the FP64 MFMA product kernels (symmetric-tile products with fused
// trace / Frobenius reductions); the chemical-potential Newton steps and the
// branch decisions are host scalars, as in the reference.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "prg_bml.h"
#include "prg_eigen.h"
#include "prg_ops.h"

namespace prg {
namespace {

constexpr int kMaxOuter = 200;  // cap of the reference's open-ended mu loops

int bounds(const DMat& H, double* g) {
  int st = OK;
  double* d_b = (double*)workspace("sp2f_bounds", sizeof(double) * 4, &st);
  if (st || (st = gershgorin(H, d_b))) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(g, d_b, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  if (!std::isfinite(g[0]) || !std::isfinite(g[1]) || !(g[1] > g[0])) {
    set_error(ERR_NUMERIC, "SP2 Fermi: empty or non-finite spectral interval");
    return ERR_NUMERIC;
  }
  return OK;
}

// X = ((hN - mu) I - H) / (hN - h1)   (prg_normalize_fermi,
// src/prg_normalize_mod.F90:59-72)
int normalize_fermi(const DMat& H, double h1, double hN, double mu, DMat X) {
  const double inv = 1.0 / (hN - h1);
  return ops::lin(H, -inv, nullptr, 0.0, (hN - mu) * inv, X);
}

// max_j |f(j+1) - f(j)| / de (absmaxderivative, src/prg_sp2_fermi_mod.F90:754-768)
double absmaxderivative(const std::vector<double>& f, double de) {
  double m = -10000.0;
  for (size_t j = 0; j + 1 < f.size(); ++j) m = std::max(m, std::fabs(f[j + 1] - f[j]) / de);
  return m;
}

}  // namespace

// src/prg_sp2_fermi_mod.F90:43-185 (norecs = false) and :203-387 (norecs = true:
// nsteps is cut where the probe's steepest slope first exceeds the target beta).
int sp2_fermi_init(const DMat& H, int* nsteps, bool norecs, double nocc, double tscale,
                   double occ_lim, double trace_lim, DMat X, double* mu, double* beta, double* h1,
                   double* hN, int* sgn, int verbose) {
  int st = ensure_device();
  if (st) return st;
  const int n = H.n;
  if (*nsteps <= 0 || *nsteps > 1000) {
    set_error(ERR_ARG, "prg_sp2_fermi_init: nsteps must be in [1, 1000]");
    return ERR_ARG;
  }
  double g[2];
  if ((st = bounds(H, g))) return st;
  const double beta0 = *beta;
  *mu = 0.5 * (g[1] + g[0]);
  *h1 = tscale * g[0];
  *hN = tscale * g[1];
  DMat X1, X2, T;
  if ((st = ops::temp("sp2f_x1", n, &X1)) || (st = ops::temp("sp2f_x2", n, &X2)) ||
      (st = ops::temp("sp2f_t", n, &T)))
    return st;
  bool first = true;
  double occ_err = 1.0;
  int outer = 0;
  std::vector<double> probe(1000), probe2(1000);
  while (occ_err > occ_lim) {
    if (++outer > kMaxOuter) {
      set_error(ERR_NOT_CONVERGED, "prg_sp2_fermi_init: occupation loop did not converge");
      return ERR_NOT_CONVERGED;
    }
    if ((st = normalize_fermi(H, *h1, *hN, *mu, X))) return st;
    for (int i = 0; i < 1000; ++i) probe[i] = 1.0 - (i + 1) * 0.001;
    // X1 = -I / (hN - h1)
    if ((st = ops::identity(X1, -1.0 / (*hN - *h1)))) return st;
    for (int i = 0; i < *nsteps; ++i) {
      ops::Sums sx, s2;
      if ((st = ops::sums_of(X, &sx))) return st;
      if ((st = ops::mm(X, X, X2, true, 1.0, 0.0, nullptr, 0.0, &s2))) return st;
      const double tr0 = sx.tr, tr2 = s2.tr;
      if (first) sgn[i] = (std::fabs(tr2 - nocc) < std::fabs(2.0 * tr0 - tr2 - nocc)) ? -1 : 1;
      // T = X X1 + X1 X
      if ((st = ops::mm2(X, X1, X1, X, T, true))) return st;
      if (sgn[i] == 1) {
        if ((st = ops::lin(X1, 2.0, &T, -1.0, 0.0, X1))) return st;
        if ((st = ops::lin(X, 2.0, &X2, -1.0, 0.0, X))) return st;
      } else {
        if ((st = ops::copy(T, X1))) return st;
        if ((st = ops::copy(X2, X))) return st;
      }
      if (norecs) {
        for (int k = 0; k < 1000; ++k) {
          const double p2 = probe[k] * probe[k];
          probe[k] = (sgn[i] == 1) ? 2.0 * probe[k] - p2 : p2;
        }
        const double maxder = absmaxderivative(probe, 0.001);
        *beta = 4.0 * maxder / (g[1] - g[0]);
        if (*beta > beta0) {
          *nsteps = i + 1;
          break;
        }
      }
    }
    if (norecs && verbose >= 1) std::printf("prg_sp2_fermi_init_norecs: beta = %g\n", *beta);
    first = false;
    ops::Sums sx, s1;
    if ((st = ops::sums_of(X, &sx)) || (st = ops::sums_of(X1, &s1))) return st;
    occ_err = std::fabs(nocc - sx.tr);
    const double lambda = (std::fabs(s1.tr) > trace_lim) ? (nocc - sx.tr) / s1.tr : 0.0;
    *mu += lambda;
  }
  // beta = -Tr(X1) / Tr(X (I - X))
  ops::Sums sx, s1;
  if ((st = ops::sums_of(X, &sx)) || (st = ops::sums_of(X1, &s1))) return st;
  const double tr_x_ix = sx.tr - sx.fro2;  // X symmetric: Tr(X (I - X)) = Tr X - ||X||_F^2
  *beta = (std::fabs(tr_x_ix) > trace_lim) ? -s1.tr / tr_x_ix : -1000.0;
  return OK;
}

// src/prg_sp2_fermi_mod.F90:405-470
int sp2_fermi(const DMat& H, int osteps, int nsteps, double nocc, double* mu, double beta,
              double h1, double hN, const int* sgn, double eps, double trace_lim, DMat X) {
  int st = ensure_device();
  if (st) return st;
  const int n = H.n;
  if (nsteps < 0 || nsteps > 1000 || osteps < 0) {
    set_error(ERR_ARG, "prg_sp2_fermi: bad step counts");
    return ERR_ARG;
  }
  if (!(hN - h1 != 0.0)) {
    set_error(ERR_NUMERIC, "prg_sp2_fermi: h1 == hN");
    return ERR_NUMERIC;
  }
  DMat X2, DX;
  if ((st = ops::temp("sp2f_x2", n, &X2)) || (st = ops::temp("sp2f_dx", n, &DX))) return st;
  double occ_err = 1.0 + eps, lambda = 0.0;
  int iter = 0;
  while ((osteps == 0 && occ_err > eps) || (osteps > 0 && iter < osteps)) {
    if (++iter > kMaxOuter && osteps == 0) {
      set_error(ERR_NOT_CONVERGED, "prg_sp2_fermi: occupation loop did not converge");
      return ERR_NOT_CONVERGED;
    }
    if ((st = normalize_fermi(H, h1, hN, *mu, X))) return st;
    for (int i = 0; i < nsteps; ++i) {
      if ((st = ops::mm(X, X, X2, true))) return st;
      if (sgn[i] == 1) {
        if ((st = ops::lin(X, 2.0, &X2, -1.0, 0.0, X))) return st;
      } else {
        if ((st = ops::copy(X2, X))) return st;
      }
    }
    ops::Sums sx;
    if ((st = ops::sums_of(X, &sx))) return st;
    occ_err = std::fabs(nocc - sx.tr);
    // DX = -beta X (I - X)
    if ((st = ops::lin(X, -1.0, nullptr, 0.0, 1.0, X2))) return st;  // X2 = I - X
    ops::Sums sd;
    if ((st = ops::mm(X, X2, DX, true, -beta, 0.0, nullptr, 0.0, &sd))) return st;
    lambda = (std::fabs(sd.tr) > trace_lim) ? (nocc - sx.tr) / sd.tr : 0.0;
    *mu += lambda;
  }
  // occupation correction X <- X + lambda DX
  return ops::lin(X, 1.0, &DX, lambda, 0.0, X);
}

// sp2_inverse (src/prg_sp2_fermi_mod.F90:720-745): the energy whose SP2-Fermi
// occupation is f.
double sp2_inverse(double f, double mu, double h1, double hN, int nsteps, const int* sgn) {
  double e = f;
  for (int i = nsteps - 1; i >= 0; --i) {
    const double s = (double)sgn[i];
    const double c = 0.5 * (1.0 + s);
    e = c - s * std::sqrt(std::fabs(c - s * e));
  }
  return hN - mu - e * (hN - h1);
}

}  // namespace prg

using namespace prg;

extern "C" {

// src/prg_c_interface.F90:904-927 (the bind(C) wrapper carries a trailing
// verbose argument that the header at src/prg_progress_mod.h:435-448 omits;
// it is unused).  sgnlist: caller array of length >= nsteps.
void prg_sp2_fermi_init(bml_matrix_t* h_bml, int nsteps, double nocc, double tscale,
                        double threshold, double occErrLimit, double traceLimit,
                        bml_matrix_t* x_bml, double* mu, double* beta, double* h1, double* hN,
                        int* sgnlist, int verbose) {
  (void)threshold;
  clear_error();
  if (!mu || !beta || !h1 || !hN || !sgnlist) {
    set_error(ERR_ARG, "prg_sp2_fermi_init: null argument");
    return;
  }
  if (check_pair(h_bml, x_bml, "prg_sp2_fermi_init")) return;
  DMat H, X;
  if (bml_dev_read(h_bml, &H) || bml_dev_write(x_bml, &X)) return;
  int ns = nsteps;
  sp2_fermi_init(H, &ns, false, nocc, tscale, occErrLimit, traceLimit, X, mu, beta, h1, hN,
                 sgnlist, verbose);
}

// src/prg_c_interface.F90:929-955 (src/prg_sp2_fermi_mod.F90:203-387)
void prg_sp2_fermi_init_norecs(bml_matrix_t* h_bml, int* nsteps, double nocc, double tscale,
                               double threshold, double occErrLimit, double traceLimit,
                               bml_matrix_t* x_bml, double* mu, double* beta, double* h1,
                               double* hN, int* sgnlist, int verbose) {
  (void)threshold;
  clear_error();
  if (!nsteps || !mu || !beta || !h1 || !hN || !sgnlist) {
    set_error(ERR_ARG, "prg_sp2_fermi_init_norecs: null argument");
    return;
  }
  if (check_pair(h_bml, x_bml, "prg_sp2_fermi_init_norecs")) return;
  DMat H, X;
  if (bml_dev_read(h_bml, &H) || bml_dev_write(x_bml, &X)) return;
  sp2_fermi_init(H, nsteps, true, nocc, tscale, occErrLimit, traceLimit, X, mu, beta, h1, hN,
                 sgnlist, verbose);
}

// src/prg_c_interface.F90:957-979 (src/prg_sp2_fermi_mod.F90:405-470).  X is
// the occupation-corrected Fermi-like matrix (not doubled), as in the reference.
void prg_sp2_fermi(bml_matrix_t* h_bml, int osteps, int nsteps, double nocc, double* mu,
                   double* beta, double* h1, double* hN, int* sgnlist, double threshold,
                   double eps, double traceLimit, bml_matrix_t* x_bml) {
  (void)threshold;
  clear_error();
  if (!mu || !beta || !h1 || !hN || !sgnlist) {
    set_error(ERR_ARG, "prg_sp2_fermi: null argument");
    return;
  }
  if (check_pair(h_bml, x_bml, "prg_sp2_fermi")) return;
  DMat H, X;
  if (bml_dev_read(h_bml, &H) || bml_dev_write(x_bml, &X)) return;
  sp2_fermi(H, osteps, nsteps, nocc, mu, *beta, *h1, *hN, sgnlist, eps, traceLimit, X);
}

// src/prg_c_interface.F90:981-995 (src/prg_sp2_fermi_mod.F90:480-531): the
// entropy function GG on the mesh ee = 0, 1e-4, ..., 1 (10001 points; the
// reference's C wrapper declares 1001-element arrays for the 10001 values it
// copies, so callers here pass 10001).  Two-point Gauss quadrature per cell.
void prg_sp2_entropy_function(double mu, double h1, double hN, int nsteps, const int* sgnlist,
                              double* GG, double* ee) {
  clear_error();
  if (!sgnlist || !GG || !ee || nsteps < 0) {
    set_error(ERR_ARG, "prg_sp2_entropy_function: bad arguments");
    return;
  }
  const int N = 10001;
  const double dh = 0.0001, x1 = 0.5773502691896257, x2 = -x1;
  for (int i = 0; i < N; ++i) ee[i] = i * dh;
  GG[0] = 0.0;
  double acc = 0.0;
  for (int i = 0; i < N - 1; ++i) {
    const double a = ee[i], b = ee[i + 1];
    const double c = 0.5 * (b - a);
    acc += c * sp2_inverse(0.5 * ((b - a) * x1 + (b + a)), mu, h1, hN, nsteps, sgnlist);
    acc += c * sp2_inverse(0.5 * ((b - a) * x2 + (b + a)), mu, h1, hN, nsteps, sgnlist);
    GG[i + 1] = acc;
  }
  const double last = GG[N - 1];
  for (int i = 0; i < N; ++i) GG[i] -= last * ee[i];
}

// sp2_inverse (src/prg_sp2_fermi_mod.F90:720-745) as a C symbol.
double prg_sp2_inverse(double f, double mu, double h1, double hN, int nsteps, const int* sgnlist) {
  return sp2_inverse(f, mu, h1, hN, nsteps, sgnlist);
}

// sp2_entropy_ts (src/prg_sp2_fermi_mod.F90:541-587): TS from the
// eigenvalues of D0 (rocSOLVER) interpolated in GG on the mesh ee.
double prg_sp2_entropy_ts(bml_matrix_t* d0_bml, const double* GG, const double* ee) {
  clear_error();
  if (!d0_bml || !GG || !ee) {
    set_error(ERR_ARG, "prg_sp2_entropy_ts: null argument");
    return 0.0;
  }
  DMat D;
  if (bml_dev_read(d0_bml, &D)) return 0.0;
  DMat Q;
  if (ops::temp("sp2f_q", D.n, &Q)) return 0.0;
  int st = OK;
  double* d_w = (double*)workspace("sp2f_w", sizeof(double) * (size_t)D.ld, &st);
  if (st) return 0.0;
  std::vector<double> w;
  if (eig(D, Q, d_w, &w)) return 0.0;
  double ts = 0.0;
  for (double x : w) {
    const double hs = std::fabs(x);
    const int j = (int)std::floor(hs / 0.0001) + 1;  // 1-based mesh index
    if (j > 0 && j <= 10000)
      ts += ((hs - ee[j - 1]) * GG[j] + (ee[j] - hs) * GG[j - 1]) / (ee[j] - ee[j - 1]);
  }
  return ts;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Congruence transforms around the solvers (src/prg_nonortho_mod.F90): the
// FP64 MFMA products of the GEMM family (transpose kernel + two products).
namespace prg {
namespace {
int congruence(const DMat& A, const DMat& Z, DMat out, bool ortho) {
  int st = ensure_device();
  if (st) return st;
  const int n = A.n;
  DMat Zt, T;
  if ((st = ops::temp("nonortho_zt", n, &Zt)) || (st = ops::temp("nonortho_t", n, &T))) return st;
  if ((st = ops::transpose(Z, Zt))) return st;
  if (ortho) {
    // out = Z^T A Z  (src/prg_nonortho_mod.F90:58-64)
    if ((st = ops::mm(Zt, A, T, false))) return st;
    return ops::mm(T, Z, out, false);
  }
  // out = Z A Z^T  (src/prg_nonortho_mod.F90:101-115)
  if ((st = ops::mm(A, Zt, T, false))) return st;
  return ops::mm(Z, T, out, false);
}
}  // namespace
}  // namespace prg

extern "C" {

// prg_orthogonalize (src/prg_nonortho_mod.F90:35-68): orthoA = Z^T A Z.  No C
// symbol in the reference; arguments in the Fortran order (bml_type names the
// storage of a newly allocated output and is not needed for the dense handle).
void prg_orthogonalize(bml_matrix_t* A_bml, bml_matrix_t* zmat_bml, bml_matrix_t* orthoA_bml,
                       double threshold, const char* bml_type, int verbose) {
  (void)threshold;
  (void)bml_type;
  clear_error();
  if (check_pair(A_bml, zmat_bml, "prg_orthogonalize") ||
      check_pair(A_bml, orthoA_bml, "prg_orthogonalize"))
    return;
  if (verbose == 1) std::printf("In prg_orthogonalize ...\n");
  DMat A, Z, O;
  if (bml_dev_read(A_bml, &A) || bml_dev_read(zmat_bml, &Z) || bml_dev_write(orthoA_bml, &O))
    return;
  congruence(A, Z, O, true);
}

// prg_deorthogonalize (src/prg_nonortho_mod.F90:81-120): A = Z orthoA Z^T.
void prg_deorthogonalize(bml_matrix_t* orthoA_bml, bml_matrix_t* zmat_bml, bml_matrix_t* a_bml,
                         double threshold, const char* bml_type, int verbose) {
  (void)threshold;
  (void)bml_type;
  clear_error();
  if (check_pair(orthoA_bml, zmat_bml, "prg_deorthogonalize") ||
      check_pair(orthoA_bml, a_bml, "prg_deorthogonalize"))
    return;
  if (verbose == 1) std::printf("In prg_deorthogonalize ...\n");
  DMat O, Z, A;
  if (bml_dev_read(orthoA_bml, &O) || bml_dev_read(zmat_bml, &Z) || bml_dev_write(a_bml, &A))
    return;
  congruence(O, Z, A, false);
}

}  // extern "C"
