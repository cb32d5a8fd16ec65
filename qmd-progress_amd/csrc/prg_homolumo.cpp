// HOMO-LUMO gap estimate and SP2 sequence generation
// (src/prg_homolumo_mod.F90).  Both are O(iterations) scalar recursions on
// the host, exactly as in the reference; they feed the graph-partitioned loop
// with the shared branch sequence.
#include <cmath>
#include <cstdio>

#include "prg_internal.h"

extern "C" {

// src/prg_homolumo_mod.F90:22-89.  vv/pp are the reference's vv(1..imax),
// pp(1..imax) stored from index 0.
void prg_homolumogap(const double* vv, int imax, const int* pp, double mineval, double maxeval,
                     double* ehomo, double* elumo, double* egap, int verbose) {
  prg::clear_error();
  if (!vv || !pp || !ehomo || !elumo || !egap || imax < 1) {
    prg::set_error(prg::ERR_ARG, "prg_homolumogap: bad arguments");
    return;
  }
  double x_a = 0.0, x_b = 1.0;
  int i = imax;
  while (i >= 1 && vv[i - 1] <= 0.0) --i;
  const double g0 = 6.0 - 4.0 * std::sqrt(2.0);
  const double hgamma = g0 * (1.0 - g0);
  if (verbose >= 1) std::printf("In prg_homolumogap ...\n");
  while (i >= 1 && vv[i - 1] < hgamma) {
    const double precomp = std::sqrt(1.0 - 4.0 * vv[i - 1]);
    double y_a = 0.5 * (1.0 + precomp);
    double y_b = 0.5 * (1.0 - precomp);
    for (int j = i - 1; j >= 1; --j) {
      if (pp[j - 1] > 0) {
        y_a = std::sqrt(y_a);
        y_b = std::sqrt(y_b);
      } else {
        y_a = 1.0 - std::sqrt(1.0 - y_a);
        y_b = 1.0 - std::sqrt(1.0 - y_b);
      }
    }
    x_a = std::fmax(x_a, y_a);
    x_b = std::fmin(x_b, y_b);
    --i;
    if (i < 1 && verbose >= 1) std::printf("prg_homolumogap error: i < 1, i = %d\n", i);
  }
  *ehomo = maxeval - x_a * (maxeval - mineval);
  *elumo = maxeval - x_b * (maxeval - mineval);
  *egap = *elumo - *ehomo;
}

// src/prg_homolumo_mod.F90:97-147, including its parenthesisation of the
// second operand of the branch comparison.  pp must hold 100 entries.
void prg_sp2sequence(int* pp, int* imax, double mineval, double maxeval, double ehomo,
                     double elumo, double errlimit, int verbose) {
  prg::clear_error();
  if (!pp || !imax || !(maxeval - mineval != 0.0)) {
    prg::set_error(prg::ERR_ARG, "prg_sp2sequence: bad arguments");
    return;
  }
  double eh = (maxeval - ehomo) / (maxeval - mineval);
  double el = (maxeval - elumo) / (maxeval - mineval);
  double error = 1.0;
  int it = 0;
  while (error > errlimit) {
    ++it;
    const double lhs = std::fabs(1.0 - eh * eh) + std::fabs(el * el);
    const double rhs = std::fabs(1.0 - (2.0 * eh - eh * eh) + std::fabs(2.0 * el - el * el));
    pp[it - 1] = (lhs < rhs) ? 1 : 0;
    const double sgm = 1.0 - 2.0 * pp[it - 1];
    eh = eh + sgm * (eh - eh * eh);
    el = el + sgm * (el - el * el);
    error = std::fabs(1.0 - eh) + std::fabs(el);
    if (it >= 100) {
      error = 0.0;
      if (verbose >= 1) std::printf("prg_sp2sequence error: SP2 not converging\n");
    }
  }
  *imax = it;
}

}  // extern "C"
