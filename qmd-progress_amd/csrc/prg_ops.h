// Host-side matrix operations on device matrices: thin wrappers over the
// GEMM / elementwise kernels that also return the fused reductions (trace,
// squared Frobenius norm, <C, E>) to the host.  Used by the iterative
// solvers whose scalar recurrences (CG step lengths, Newton-Schulz residuals,
// chemical-potential updates) the reference runs on the host.
#pragma once

#include "prg_internal.h"

namespace prg {
namespace ops {

struct Sums {
  double tr = 0.0;
  double fro2 = 0.0;
  double dot = 0.0;
};

// Products issued through mm / mm2 since the last reset (count, 2 n^3 per
// product, FP64 MFMA flops executed, int8 MFMA operations executed).
void gemm_counters(double* products, double* alg_flops, double* exec_flops, bool reset,
                   double* int8_ops = nullptr);
// Arithmetic of mm's symmetric products: PRG_PREC_FP64 (default) or
// PRG_PREC_INT8CRT (crt_gemm with the int8 CRT moduli count).
int product_precision();
int set_product_precision(int mode);

// Scope during which A is not modified: int8 CRT products reuse its residue
// planes (no effect on FP64 products).
struct HoldOperand {
  explicit HoldOperand(const DMat& A);
  ~HoldOperand();
  const double* p;
};

// n x n temporary held in a named workspace slot (contents undefined).
int temp(const char* name, int n, DMat* out);
// C = alpha A B + beta Cin + gamma I   (sym: A and B symmetric and commuting,
// so the product is symmetric; such products follow product_precision())
int mm(const DMat& A, const DMat& B, DMat C, bool sym, double alpha = 1.0, double beta = 0.0,
       const DMat* Cin = nullptr, double gamma = 0.0, Sums* sums = nullptr,
       const DMat* dotE = nullptr);
// C = A B for factors that are not symmetric but whose product is (FP64,
// upper-triangle tiles mirrored)
int mm_symmetric_result(const DMat& A, const DMat& B, DMat C);
// C = alpha (A0 B0 + A1 B1) + beta Cin + gamma I
int mm2(const DMat& A0, const DMat& B0, const DMat& A1, const DMat& B1, DMat C, bool sym,
        double alpha = 1.0, double beta = 0.0, const DMat* Cin = nullptr, double gamma = 0.0,
        Sums* sums = nullptr);
// out = a A + b B + c I   (B may be null; out may alias A or B)
int lin(const DMat& A, double a, const DMat* B, double b, double c, DMat out,
        Sums* sums = nullptr);
int copy(const DMat& src, DMat dst);
int identity(DMat dst, double diag = 1.0);
int sums_of(const DMat& A, Sums* s);
// out = A^T (out must not alias A)
int transpose(const DMat& A, DMat out);
// out[:, j] = A[:, j] * f[j]  (f: device vector of length n)
int scale_cols(const DMat& A, const double* d_f, DMat out);
// out[i, :] = f[i] * A[i, :]  (f: device vector of length n)
int scale_rows(const DMat& A, const double* d_f, DMat out);
// d_out[i] = 1 / A[i][i]
int inv_diagonal(const DMat& A, double* d_out);
// *out = Tr(X Y) = sum_ij X[i][j] Y[j][i] (bml_trace_mult), to the host
int trace_mult(const DMat& X, const DMat& Y, double* out);
// Reduce [nb][nq] device partials and copy the nq sums to the host.
int fetch(const double* d_part, int nb, int nq, double* host);

}  // namespace ops
}  // namespace prg
