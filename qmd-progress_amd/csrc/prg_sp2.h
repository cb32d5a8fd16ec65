// SP2 driver interface shared by the C-ABI wrappers, the response routines
// and the subgraph loop.
#pragma once

#include "prg_bml.h"

namespace prg {

struct Sp2Params {
  int variant = V_ALG2;
  double occ = 0.0;
  double idemtol = 1e-10;
  int minit = 0;
  int maxit = 100;
  int rel = 0;
  int absconv = 0;
  int icount = 0;
  int core = 0;
  const int* pp_in = nullptr;
  const double* vv_in = nullptr;
  bool use_gershgorin = true;  // false: normalise with (mineval, maxeval)
  bool want_bounds = false;
  double mineval = 0.0;
  double maxeval = 0.0;
  int precision = 0;
  int refine_max = 2;  // FP64 polishing steps after a low-precision loop
  double threshold = 0.0;  // sparse-format thresholding of X (0: dense)
};

struct Sp2Result {
  int iter = 0;
  int refine_steps = 0;
  int status = 0;
  double trx = 0.0;
  double mineval = 0.0;
  double maxeval = 0.0;
  int pp[kMaxIter] = {};
  double vv[kMaxIter] = {};
};

// Buffers a caller hands to run_sp2 instead of the library workspace (the
// no-allocation entry point): device memory for the ping-pong X, the partial
// sums, the bounds, the state, the K-split accumulators and the Gershgorin
// scratch, and three host state images.  With them the loop reads the state
// header back synchronously (no events) and keeps no per-shape history.
struct Sp2Buffers {
  double* x0;
  double* x1;
  double* part;
  double* bounds;
  Sp2State* state;
  double* split_acc;
  double* gersh;
  Sp2State* host;
};

// Runs one SP2 variant: out = 2 X_final.  For V_RESPONSE, H1 is the
// perturbation and out1 receives rho1 = 2 Delta_final.  bufs (FP64 only, no
// response): caller-provided buffers, see Sp2Buffers.
int run_sp2(const DMat& H, const Sp2Params& p, DMat out, Sp2Result* res,
            const DMat* H1 = nullptr, DMat* out1 = nullptr, const Sp2Buffers* bufs = nullptr);

// Mixed-precision fused SP2 step (prec = PRG_PREC_F16SPLIT or PRG_PREC_INT8CRT).
int sp2_step_mixed(const DMat& x0, const DMat& x1, Sp2State* d_state, double* d_part,
                   int* ntiles, int prec);
// Builds the low-precision representation of X0 (after normalisation).
int sp2_mixed_prepare(const DMat& x0, int prec);

// int8 Chinese-remainder products (prg_gemm_crt.hip)
int crt_prepare(const DMat& x0, int np);
int crt_step(const DMat& x0, const DMat& x1, Sp2State* d_state, double* d_part, int* ntiles,
             int np);
int crt_moduli();
// C = alpha A B + beta Cin + gamma I (gamma on the n x n diagonal) for
// symmetric, commuting A and B, from exact int8 products modulo np moduli
// (np in [6, 14]); part (if not null) receives per-64x64-block [Tr C,
// ||C||_F^2, <C, dotE>, 0], *nparts blocks.  C is exactly symmetric.
int crt_gemm(const DMat& A, const DMat& B, DMat C, double alpha, double beta, const DMat* Cin,
             double gamma, double* part, int* nparts, const DMat* dotE, int np);
// C = alpha A B + beta Cin + gamma I for symmetric, commuting A and B from
// BF16 hi+lo split products with FP32 accumulation (hi*hi + 2^-8 (hi*lo +
// lo*hi)); part (if not null) receives per-128x128-tile [Tr C, ||C||_F^2,
// <C, dotE>, 0], *nparts tiles.  C is exactly symmetric.
int split_gemm(const DMat& A, const DMat& B, DMat C, double alpha, double beta, const DMat* Cin,
               double gamma, double* part, int* nparts, const DMat* dotE);
// Hold operand p across split_gemm calls (its planes are formed once).
void split_hold_operand(const double* p);
void split_release_operand(const double* p);
// Hold operand p across crt_gemm calls (its planes are formed once); the
// caller must not change the matrix until the matching release.  Holds of the
// same pointer nest; a hold of another pointer while one is active is ignored.
void crt_hold_operand(const double* p);
void crt_release_operand(const double* p);
int crt_last_waves();  // configuration code of the last int8 CRT product kernel (0: none yet)
int set_crt_config(int config);  // force a product kernel configuration (0: autotune)
int crt_update_blocks(int ld);  // partial-sum slots one int8 CRT step writes
// out = scale X[cur] after an int8 CRT solve (from the residue planes when no
// FP64 master is kept, from the master's lower triangle otherwise)
int crt_finish(const DMat& x0, const DMat& x1, const Sp2State* d_state, double scale, DMat out,
               int np);
// HIP-event timing of the int8 CRT product kernel alone: enable/reset, start of
// a solve, sum the first `executed` steps of the solve, read the totals.
void crt_timing(bool enable);
void crt_timing_begin();
int crt_timing_collect(int executed);
void crt_timing_get(double* ms_sum, int* count);
int set_crt_moduli(int np);

int precision_mode();
void set_precision_mode(int m);
bool fortran_rel(const char* s);
bool fortran_abs(const char* s);

}  // namespace prg
