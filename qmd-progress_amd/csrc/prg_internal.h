// Internal declarations of the MI355X PROGRESS density-matrix library.
//
// Device storage: every dense N x N matrix lives in HBM as FP64, column major,
// with a leading dimension padded to a multiple of the 128-wide GEMM tile and
// zeros outside the N x N block.  Zero padding keeps every product, trace and
// Frobenius sum of the N x N block exact, so no kernel carries bounds checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace prg {

// ---------------------------------------------------------------- errors
enum Status : int {
  OK = 0,
  ERR_ARG = 1,          // bad argument (sizes, null handles, bad options)
  ERR_HIP = 2,          // HIP runtime failure
  ERR_NOT_CONVERGED = 3,// iteration cap reached (the reference calls stop)
  ERR_NUMERIC = 4,      // NaN/Inf or a zero-width spectral interval
  ERR_UNSUPPORTED = 5,  // option outside what this library implements
  ERR_NO_DEVICE = 6,    // no GPU visible
};

void set_error(int code, const std::string& msg);
void clear_error();
int last_error_code();
const char* last_error_message();

#define PRG_HIP_CHECK(expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::prg::set_error(::prg::ERR_HIP, std::string(#expr) + ": " +            \
                                           hipGetErrorString(e_));            \
      return ::prg::ERR_HIP;                                                  \
    }                                                                         \
  } while (0)

// ---------------------------------------------------------------- layout
constexpr int kTile = 128;  // GEMM output tile edge; padding granule
inline int padded(int n) { return ((n + kTile - 1) / kTile) * kTile; }

// A device matrix view: n x n block of an ld x ld zero-padded FP64 buffer.
struct DMat {
  int n = 0;
  int ld = 0;
  double* p = nullptr;
};

// ---------------------------------------------------------------- runtime
hipStream_t stream();
int ensure_device();  // initialises the device once; returns Status
bool device_ready();  // ensure_device has run and succeeded
// Resolves the no-allocation scratch array on the current device (once per
// device); called by prg_progress_init.
int noalloc_setup();

// Scratch buffer cache keyed by name; grows on demand and is reused across
// calls on the library stream (calls are serialised on that stream).
void* workspace(const char* name, size_t bytes, int* status);
// Incremented whenever a workspace buffer is freed (growth or
// release_workspaces).  Host-side caches of what was uploaded into a
// workspace record it and upload again once it has changed.
unsigned workspace_generation();
// Pinned host staging buffer for small device->host result copies.
void* host_staging(size_t bytes, int* status);
void release_workspaces();

// ---------------------------------------------------------------- SP2 state
constexpr int kMaxIter = 512;

// Variants of the SP2 control flow (src/prg_sp2_mod.F90).
enum Sp2Variant : int {
  V_ALG2 = 0,         // prg_sp2_alg2
  V_ALG2_GENSEQ = 1,  // prg_sp2_alg2_genseq
  V_ALG1 = 2,         // prg_sp2_alg1
  V_ALG1_GENSEQ = 3,  // prg_sp2_alg1_genseq
  V_BASIC = 4,        // prg_sp2_basic
  V_SEQ_ALG2 = 5,     // prg_sp2_alg2_seq / prg_prg_sp2_alg2_seq_inplace
  V_SEQ_ALG1 = 6,     // prg_sp2_alg1_seq / prg_prg_sp2_alg1_seq_inplace
  V_SUBMATRIX = 7,    // prg_sp2_submatrix / _inplace (vv accumulates core rows)
  V_RESPONSE = 8,     // prg_compute_response_SP2 (alg2 control flow)
  V_BASIC_FIXED = 9,  // prg_sp2_basic_tcore (basic branch rule, maxit + 1 steps, no early stop)
};

// Device-resident SP2 bookkeeping.  All FP64 scalar work of the loop (traces,
// branch choice, stopping rules) happens on the device in sp2_plan/absorb.
struct Sp2State {
  double occ, idemtol;
  double trx;        // trace of X carried by the variant's recursion
  double trx_fresh;  // trace of the current X, summed from its diagonal
  double fro2;       // ||X||_F^2 of the current X (= Tr X^2 for symmetric X)
  double trx_next;   // trace after the planned update
  double trx2;       // Tr(X^2) (alg2) or Tr(X - X^2) (alg1) of the planned step
  double limdiff;
  double e0, e1, e2; // idempotency error history
  double x2_trace;   // diagonal sum of the last computed product (diagnostic)
  int variant, minit, maxit, rel, absconv;
  int iter;          // completed iterations (reference's iter / icount)
  int planned;       // 1: a product is to be computed for step iter+1
  int apply;         // 1: the planned product updates X; 0: measure only
  int sigma;         // 1: X <- X^2 ; 0: X <- 2X - X^2
  int last;          // stop after the planned step
  int done;
  int status;        // Status
  int cur;           // which ping-pong buffer holds X
  int icount;        // sequence length for replay variants
  int core;          // core rows for V_SUBMATRIX
  // low-precision products: stop on the noise floor of the idempotency error
  // (e = Tr X - Tr X^2; src/sp2_tensorcore/src/prg_sp2_tensorcore.cu:308-317),
  // then polish with up to refine_max FP64 steps on the alg2 rule
  int noise_stop;
  int refine_max;
  int refine_phase;  // 1 while the FP64 polishing steps run
  int refine_steps;  // polishing steps taken
  double ie0, ie1, ie2;  // idempotency errors of the last three X
  // sparse BML formats (ellpack, ellsort, ellblock, csr): entries with
  // |a| <= threshold are dropped from every product and sum, as those
  // formats do; 0 for dense storage (BML dense ignores thresholds)
  double threshold;
  int pp[kMaxIter];
  double vv[kMaxIter];
};

// ---------------------------------------------------------------- kernels
// Gershgorin interval of the n x n block (row sums, as bml_gershgorin).
// scratch (optional, device): gershgorin_scratch_doubles(n) doubles instead
// of the library workspace.
int gershgorin(const DMat& h, double* d_bounds /*2 doubles, device*/, double* scratch = nullptr);
size_t gershgorin_scratch_doubles(int n);
// X = alpha*H + beta*I with alpha/beta computed on the device from the
// Gershgorin bounds (mode 0: SP2 map (hi I - H)/(hi-lo)), or given scalars.
// Also writes Tr(X) and ||X||_F^2 partials into d_part (per block) and the
// number of blocks to *nblocks.
int normalize_gershgorin(const DMat& h, const double* d_bounds, DMat x, double* d_part,
                         int* nblocks);
int normalize_bounds(const DMat& h, double lo, double hi, DMat x, double* d_part, int* nblocks);
int affine(const DMat& a, double alpha, double beta_identity, DMat out, double* d_part, int* nblocks);
// out = a*A + b*B (+ c I); optional partial sums of Tr(out) and ||out||^2.
int axpby(const DMat& A, double a, const DMat& B, double b, double c, DMat out, double* d_part,
          int* nblocks);
int scale_copy(const DMat& src, double s, DMat dst);
int set_identity(DMat dst, double diag);
int copy_block(const DMat& src, DMat dst);
// Deterministic sums of per-block partials: partials[nb][q] -> out[q].
int reduce_partials(const double* d_part, int nblocks, int nq, double* d_out);

// GEMM family (FP64 MFMA 16x16x4).
// C = alpha * sum_p A_p B_p + beta * Cin + gamma * I ; npair in {1, 2}.
// sym != 0: the product is symmetric; only upper-triangle tiles are computed
// and mirrored (Cin must then be symmetric too).  d_part (optional) receives
// per-tile [Tr(C), ||C||^2, <C, dotE>, 0].  C must not alias A_p or B_p.
int gemm(int npair, const DMat* A, const DMat* B, double alpha, double beta, const DMat* Cin,
         double gamma, DMat C, int sym, double* d_part, int* ntiles,
         const DMat* dotE = nullptr);

// Fused SP2 step: Y = X*X, then per the device state's planned branch
// Xn = Y or 2X - Y; epilogue partials [||Xn||^2, Tr Xn, idem, Tr Y].  Below
// 1024 upper-triangle tiles the product is split along K
// (sp2_split_count(ld) slices) into split_acc (sp2_split_bytes(ld) bytes; the
// library workspace when null) and summed in slice order by a second kernel.
// thresholded: the sparse formats' thresholding of the state's threshold
// (product and update in two kernels; split_acc then holds one slice per
// tile at least).
int sp2_step(const DMat& x0, const DMat& x1, Sp2State* d_state, double* d_part, int* ntiles,
             double* split_acc = nullptr, bool thresholded = false);
int sp2_split_count(int ld);
// Workgroups per tile of the split product's epilogue (4, or 8 at small ld);
// each writes one record of partial sums.
int sp2_epilogue_parts(int ld);
size_t sp2_split_bytes(int ld);
// Fused response step: Y = X D + D X with the same branch: Dn = Y or 2D - Y.
int sp2_response_step(const DMat& x0, const DMat& x1, const DMat& d0, const DMat& d1,
                      Sp2State* d_state, double* d_part, int* ntiles);
// Device-side SP2 bookkeeping kernels.
int sp2_init_state(Sp2State* d_state, const double* d_part, int nblocks);
int sp2_absorb_plan(Sp2State* d_state, const double* d_part, int ntiles);
// out = scale X[cur]; lower_only: only X's lower triangle (and its diagonal
// 64 x 64 blocks) is current, the upper triangle is rebuilt by symmetry.
int sp2_finish(const DMat& x0, const DMat& x1, const Sp2State* d_state, double scale, DMat out,
               bool lower_only = false);
// Switches a finished low-precision loop into its FP64 polishing phase.
int sp2_begin_refine(Sp2State* d_state);


}  // namespace prg
