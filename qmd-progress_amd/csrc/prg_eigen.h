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

}  // namespace prg
