// Host-side wrappers of the device matrix operations (see prg_ops.h).
#include "prg_ops.h"

#include <cstring>
#include <string>

#include "prg_sp2.h"

namespace prg {

int transpose_kernel_launch(const DMat& A, DMat out);
int scale_cols_kernel_launch(const DMat& A, const double* d_f, DMat out);
int scale_rows_kernel_launch(const DMat& A, const double* d_f, DMat out);
int inv_diag_kernel_launch(const DMat& A, double* d_out);
int trace_mult_kernel_launch(const DMat& X, const DMat& Y, double* d_part, int* nblocks);

namespace ops {

// Products issued through mm / mm2 since the last reset: count, algorithmic
// flops (2 n^3 each), FP64 MFMA flops executed (symmetric products compute
// the upper-triangle 128 x 128 tiles only) and int8 MFMA operations executed
// (int8 CRT products: upper-triangle 256 x 256 tiles of the padded size, once
// per modulus).  BF16 split products count in the first two only.
namespace {
double g_products = 0.0, g_alg_flops = 0.0, g_exec_flops = 0.0, g_int8_ops = 0.0;
int g_product_precision = PRG_PREC_FP64;

void count_products(int nprod, const DMat& C, bool sym) {
  const double n = C.n, ld = C.ld, nt = C.ld / kTile;
  g_products += nprod;
  g_alg_flops += nprod * 2.0 * n * n * n;
  g_exec_flops += nprod * (sym ? nt * (nt + 1) / 2 * 2.0 * kTile * kTile * ld : 2.0 * ld * ld * ld);
}

void count_crt_product(const DMat& C, int np) {
  const double n = C.n, pld = ((C.ld + 255) / 256) * 256, nt = pld / 256;
  g_products += 1;
  g_alg_flops += 2.0 * n * n * n;
  g_int8_ops += np * nt * (nt + 1) / 2 * 2.0 * 256.0 * 256.0 * pld;
}
}  // namespace

void gemm_counters(double* products, double* alg_flops, double* exec_flops, bool reset,
                   double* int8_ops) {
  if (products) *products = g_products;
  if (alg_flops) *alg_flops = g_alg_flops;
  if (exec_flops) *exec_flops = g_exec_flops;
  if (int8_ops) *int8_ops = g_int8_ops;
  if (reset) g_products = g_alg_flops = g_exec_flops = g_int8_ops = 0.0;
}

int product_precision() { return g_product_precision; }

int set_product_precision(int mode) {
  if (mode != PRG_PREC_FP64 && mode != PRG_PREC_INT8CRT && mode != PRG_PREC_BF16SPLIT) {
    set_error(ERR_ARG,
              "prg_set_product_precision: PRG_PREC_FP64, PRG_PREC_INT8CRT or PRG_PREC_BF16SPLIT");
    return ERR_ARG;
  }
  g_product_precision = mode;
  return OK;
}

HoldOperand::HoldOperand(const DMat& A) : p(A.p) {
  crt_hold_operand(p);
  split_hold_operand(p);
}
HoldOperand::~HoldOperand() {
  crt_release_operand(p);
  split_release_operand(p);
}

int temp(const char* name, int n, DMat* out) {
  int st = OK;
  const int ld = padded(n);
  double* p = (double*)workspace(name, sizeof(double) * (size_t)ld * ld, &st);
  if (st) return st;
  out->n = n;
  out->ld = ld;
  out->p = p;
  return OK;
}

int fetch(const double* d_part, int nb, int nq, double* host) {
  int st = OK;
  double* d_out = (double*)workspace("ops_sums", 8 * sizeof(double), &st);
  double* h = (double*)host_staging(sizeof(Sp2State), &st);
  if (st) return st;
  if ((st = reduce_partials(d_part, nb, nq, d_out))) return st;
  PRG_HIP_CHECK(hipMemcpyAsync(h, d_out, nq * sizeof(double), hipMemcpyDeviceToHost, stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  std::memcpy(host, h, nq * sizeof(double));
  return OK;
}

// Partials of one launch: 128 x 128 tiles of the FP64 kernels, 64 x 64
// blocks of the int8 CRT epilogue, or the elementwise kernels' blocks.
static double* part_buf(int ld, int* st) {
  const int nb = ld / 64;
  return (double*)workspace("ops_part", sizeof(double) * 4 * (size_t)(nb * nb + 2048), st);
}

int mm(const DMat& A, const DMat& B, DMat C, bool sym, double alpha, double beta,
       const DMat* Cin, double gamma, Sums* sums, const DMat* dotE) {
  int st = OK;
  double* part = sums ? part_buf(C.ld, &st) : nullptr;
  if (st) return st;
  int nt = 0;
  if (sym && g_product_precision == PRG_PREC_INT8CRT) {
    const int np = crt_moduli();
    if ((st = crt_gemm(A, B, C, alpha, beta, Cin, gamma, part, &nt, dotE, np))) return st;
    count_crt_product(C, np);
  } else if (sym && g_product_precision == PRG_PREC_BF16SPLIT) {
    if ((st = split_gemm(A, B, C, alpha, beta, Cin, gamma, part, &nt, dotE))) return st;
    g_products += 1;
    g_alg_flops += 2.0 * C.n * (double)C.n * C.n;
  } else {
    if ((st = gemm(1, &A, &B, alpha, beta, Cin, gamma, C, sym ? 1 : 0, part, &nt, dotE)))
      return st;
    count_products(1, C, sym);
  }
  if (sums) {
    double h[4];
    if ((st = fetch(part, nt, 4, h))) return st;
    sums->tr = h[0];
    sums->fro2 = h[1];
    sums->dot = h[2];
  }
  return OK;
}

int mm_symmetric_result(const DMat& A, const DMat& B, DMat C) {
  int nt = 0;
  const int st = gemm(1, &A, &B, 1.0, 0.0, nullptr, 0.0, C, 1, nullptr, &nt, nullptr);
  if (!st) count_products(1, C, true);
  return st;
}

int mm2(const DMat& A0, const DMat& B0, const DMat& A1, const DMat& B1, DMat C, bool sym,
        double alpha, double beta, const DMat* Cin, double gamma, Sums* sums) {
  int st = OK;
  double* part = sums ? part_buf(C.ld, &st) : nullptr;
  if (st) return st;
  const DMat A[2] = {A0, A1}, B[2] = {B0, B1};
  int nt = 0;
  if ((st = gemm(2, A, B, alpha, beta, Cin, gamma, C, sym ? 1 : 0, part, &nt))) return st;
  count_products(2, C, sym);
  if (sums) {
    double h[4];
    if ((st = fetch(part, nt, 4, h))) return st;
    sums->tr = h[0];
    sums->fro2 = h[1];
    sums->dot = h[2];
  }
  return OK;
}

int lin(const DMat& A, double a, const DMat* B, double b, double c, DMat out, Sums* sums) {
  int st = OK;
  double* part = sums ? part_buf(out.ld, &st) : nullptr;
  if (st) return st;
  int nb = 0;
  if (B) st = axpby(A, a, *B, b, c, out, part, &nb);
  else st = affine(A, a, c, out, part, &nb);
  if (st) return st;
  if (sums) {
    double h[2];
    if ((st = fetch(part, nb, 2, h))) return st;
    sums->tr = h[0];
    sums->fro2 = h[1];
  }
  return OK;
}

int copy(const DMat& src, DMat dst) { return lin(src, 1.0, nullptr, 0.0, 0.0, dst); }
int identity(DMat dst, double diag) { return set_identity(dst, diag); }
int sums_of(const DMat& A, Sums* s) {
  DMat T;
  int st = temp("ops_sums_tmp", A.n, &T);
  if (st) return st;
  return lin(A, 1.0, nullptr, 0.0, 0.0, T, s);
}
int transpose(const DMat& A, DMat out) { return transpose_kernel_launch(A, out); }
int scale_cols(const DMat& A, const double* d_f, DMat out) {
  return scale_cols_kernel_launch(A, d_f, out);
}
int scale_rows(const DMat& A, const double* d_f, DMat out) {
  return scale_rows_kernel_launch(A, d_f, out);
}
int inv_diagonal(const DMat& A, double* d_out) { return inv_diag_kernel_launch(A, d_out); }
int trace_mult(const DMat& X, const DMat& Y, double* out) {
  int st = OK;
  double* part = (double*)workspace("ops_trmult", sizeof(double) * 1024, &st);
  if (st) return st;
  int nb = 0;
  if ((st = trace_mult_kernel_launch(X, Y, part, &nb))) return st;
  return fetch(part, nb, 1, out);
}

}  // namespace ops
}  // namespace prg

extern "C" {

// Counters of the FP64 products issued by the iterative solvers (implicit
// Fermi, responses, SP2-Fermi, adjoint): measurement support.
void prg_gpu_gemm_counters(int reset, double* products, double* alg_flops, double* exec_flops,
                           double* int8_ops) {
  prg::ops::gemm_counters(products, alg_flops, exec_flops, reset != 0, int8_ops);
}

int prg_gpu_symmetric_product_device(int N, const double* dA, const double* dB, double* dC) {
  prg::clear_error();
  if (N <= 0 || !dA || !dB || !dC || dC == dA || dC == dB) {
    prg::set_error(prg::ERR_ARG, "prg_gpu_symmetric_product_device: bad arguments");
    return prg::ERR_ARG;
  }
  int st = prg::ensure_device();
  if (st) return st;
  const int ld = prg::padded(N);
  prg::DMat A{N, ld, const_cast<double*>(dA)}, B{N, ld, const_cast<double*>(dB)}, Cm{N, ld, dC};
  if ((st = prg::ops::mm(A, B, Cm, true))) return st;
  PRG_HIP_CHECK(hipStreamSynchronize(prg::stream()));
  return prg::OK;
}

void prg_set_product_precision(int mode) {
  prg::clear_error();
  prg::ops::set_product_precision(mode);
}

int prg_get_product_precision(void) { return prg::ops::product_precision(); }

}  // extern "C"
