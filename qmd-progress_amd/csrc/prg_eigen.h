// Eigendecomposition helpers (rocSOLVER) shared by the density routines.
#pragma once

#include <vector>

#include "prg_internal.h"

namespace prg {

// Eigenvalues (ascending, device d_w and optionally host h_w) and
// eigenvectors (columns of Q) of the symmetric n x n block of H.
int eig(const DMat& H, DMat Q, double* d_w, std::vector<double>* h_w);
// out = Q diag(occ) Q^T
int density_from_eig(const DMat& Q, const std::vector<double>& occ, DMat out);
// out = Q diag(1/(1+exp(beta (e - mu)))) Q^T
int density_fermi(const DMat& H, double beta, double mu, DMat out, std::vector<double>* w_out);
std::vector<double> t0_occupations(int n, double bndfil);

// Batched symmetric eigensolves of `count` m x m column-major matrices stored
// m*m apart from a (overwritten by the eigenvectors); w receives m ascending
// eigenvalues per matrix.
int eig_strided_batched(double* a, int m, int count, double* w);
// c_k = a_k b_k^T for `count` m x m matrices stored m*m apart.
int gemm_nt_strided_batched(const double* a, const double* b, double* c, int m, int count);

}  // namespace prg
