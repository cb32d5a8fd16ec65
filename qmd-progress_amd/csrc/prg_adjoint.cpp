// Reverse-mode adjoint of the SP2 recursion: dL/dH for a loss of the density.
//
// Forward (with the branch sequence pp and the Gershgorin bounds frozen):
//   X_0 = (hi I - H)/(hi - lo),  X_{k+1} = X_k^2 (pp_k = 1) or 2 X_k - X_k^2,
//   D = 2 X_K.
// Backward, with G_K = 2 dL/dD:
//   S = X_k G_{k+1} + G_{k+1} X_k
//   G_k = S (pp_k = 1) or 2 G_{k+1} - S (pp_k = 0)
//   dL/dH = -G_0 / (hi - lo).
// X_k for the backward sweep are recomputed segment by segment from
// checkpoints stored every `ckpt` steps, so memory is (K/ckpt + ckpt) matrices
// instead of K.  Each backward step is one dual-product GEMM over a 2N-deep K
// (X G and G X accumulate together); it is the transpose of the tangent
// recursion of prg_compute_response_SP2 (src/prg_response_mod.F90:719-755), and
// since the Frechet derivative of a matrix function of a symmetric matrix is
// self-adjoint, dL/dH for L = Tr(D A) equals that response with H1 = A.
#include <cmath>
#include <string>
#include <vector>

#include "prg_ops.h"
#include "prg_sp2.h"

namespace prg {

int sp2_adjoint(const DMat& H, const DMat& Gout, DMat grad, const int* pp, int iters, double lo,
                double hi, int ckpt) {
  int st = ensure_device();
  if (st) return st;
  if (iters < 0 || iters > kMaxIter || ckpt < 1 || !(hi > lo)) {
    set_error(ERR_ARG, "sp2_adjoint: bad arguments");
    return ERR_ARG;
  }
  const int n = H.n, ld = H.ld;
  const size_t mbytes = sizeof(double) * (size_t)ld * ld;
  const int nck = iters / ckpt + 1;
  double* ck = (double*)workspace("adj_ckpt", mbytes * nck, &st);
  double* seg = (double*)workspace("adj_seg", mbytes * ckpt, &st);
  if (st) return st;
  auto ckm = [&](int c) { return DMat{n, ld, ck + (size_t)c * ld * ld}; };
  auto segm = [&](int j) { return DMat{n, ld, seg + (size_t)j * ld * ld}; };
  DMat X, Y, G, S;
  if ((st = ops::temp("adj_x", n, &X)) || (st = ops::temp("adj_y", n, &Y)) ||
      (st = ops::temp("adj_g", n, &G)) || (st = ops::temp("adj_s", n, &S)))
    return st;
  // forward with checkpoints
  const double inv = 1.0 / (hi - lo);
  if ((st = ops::lin(H, -inv, nullptr, 0.0, hi * inv, X))) return st;
  for (int k = 0; k < iters; ++k) {
    if (k % ckpt == 0 && (st = ops::copy(X, ckm(k / ckpt)))) return st;
    if ((st = ops::mm(X, X, Y, true))) return st;
    if (pp[k]) {
      if ((st = ops::copy(Y, X))) return st;
    } else {
      if ((st = ops::lin(X, 2.0, &Y, -1.0, 0.0, X))) return st;
    }
  }
  // backward
  if ((st = ops::lin(Gout, 2.0, nullptr, 0.0, 0.0, G))) return st;
  // A symmetric G stays symmetric (S = X G + (X G)^T), and then only the
  // upper-triangle tiles of S are computed; any other G takes full products.
  ops::Sums asym;
  if ((st = ops::transpose(G, S)) || (st = ops::lin(G, 1.0, &S, -1.0, 0.0, S, &asym))) return st;
  const bool sym = (asym.fro2 == 0.0);
  int seg_start = -1;
  for (int k = iters - 1; k >= 0; --k) {
    const int c = k / ckpt;
    if (c * ckpt != seg_start) {
      // recompute X_{c*ckpt} .. X_{min(iters, (c+1)*ckpt) - 1} into the segment buffer
      seg_start = c * ckpt;
      if ((st = ops::copy(ckm(c), segm(0)))) return st;
      for (int j = 1; j < ckpt && seg_start + j < iters; ++j) {
        const DMat prev = segm(j - 1);
        if ((st = ops::mm(prev, prev, Y, true))) return st;
        if (pp[seg_start + j - 1]) {
          if ((st = ops::copy(Y, segm(j)))) return st;
        } else {
          if ((st = ops::lin(prev, 2.0, &Y, -1.0, 0.0, segm(j)))) return st;
        }
      }
    }
    const DMat Xk = segm(k - seg_start);
    // S = X_k G + G X_k
    if ((st = ops::mm2(Xk, G, G, Xk, S, sym))) return st;
    if (pp[k]) {
      if ((st = ops::copy(S, G))) return st;
    } else {
      if ((st = ops::lin(G, 2.0, &S, -1.0, 0.0, G))) return st;
    }
  }
  return ops::lin(G, -inv, nullptr, 0.0, 0.0, grad);
}

}  // namespace prg

using namespace prg;

extern "C" {

int prg_gpu_gershgorin_device(int N, const double* dH, double* bounds) {
  clear_error();
  if (N <= 0 || !dH || !bounds) {
    set_error(ERR_ARG, "prg_gpu_gershgorin_device: bad arguments");
    return ERR_ARG;
  }
  int st = ensure_device();
  if (st) return st;
  DMat H{N, padded(N), const_cast<double*>(dH)};
  double* d = (double*)workspace("gersh_out", 2 * sizeof(double), &st);
  if (st || (st = gershgorin(H, d))) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(bounds, d, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  return OK;
}

int prg_gpu_sp2_seq_device(int N, const double* dH, double* dRho, const int* pp, int icount,
                           double mineval, double maxeval) {
  clear_error();
  if (N <= 0 || !dH || !dRho || !pp || icount < 0 || icount > kMaxIter) {
    set_error(ERR_ARG, "prg_gpu_sp2_seq_device: bad arguments");
    return ERR_ARG;
  }
  const int ld = padded(N);
  DMat H{N, ld, const_cast<double*>(dH)}, R{N, ld, dRho};
  Sp2Params p;
  p.variant = V_SEQ_ALG2;
  p.icount = icount;
  p.pp_in = pp;
  p.use_gershgorin = false;
  p.mineval = mineval;
  p.maxeval = maxeval;
  p.precision = PRG_PREC_FP64;
  Sp2Result r;
  return run_sp2(H, p, R, &r);
}

int prg_gpu_sp2_adjoint_device(int N, const double* dH, const double* dGout, double* dGrad,
                               const int* pp, int iters, double mineval, double maxeval,
                               int ckpt) {
  clear_error();
  if (N <= 0 || !dH || !dGout || !dGrad || !pp) {
    set_error(ERR_ARG, "prg_gpu_sp2_adjoint_device: bad arguments");
    return ERR_ARG;
  }
  const int ld = padded(N);
  DMat H{N, ld, const_cast<double*>(dH)}, G{N, ld, const_cast<double*>(dGout)},
      R{N, ld, dGrad};
  const int st = sp2_adjoint(H, G, R, pp, iters, mineval, maxeval, ckpt);
  if (st == OK) {
    if (hipStreamSynchronize(stream()) != hipSuccess) {
      set_error(ERR_HIP, "prg_gpu_sp2_adjoint_device: stream failure");
      return ERR_HIP;
    }
  }
  return st;
}

}  // extern "C"
