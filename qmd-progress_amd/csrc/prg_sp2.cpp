// SP2 purification drivers (src/prg_sp2_mod.F90) on one GPU.
//
// The whole iteration runs on the device: one fused product-and-update launch
// (two for the K-split product of small matrices) and one single-workgroup
// bookkeeping launch per step, enqueued in chunks.  The host reads the state
// header of a chunk (an asynchronous copy into pinned memory) only after the
// next chunk is enqueued, so no scalar crosses PCIe per iteration and the
// device does not wait for the host.  Once the loop has stopped, the
// launches still queued exit at their first instruction.
#include "prg_sp2.h"
#include "prg_ops.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

namespace prg {

namespace {
// Static device scratch of the no-allocation SP2 entry point
// (prg_gpu_sp2_alg2_device_noalloc, FP64, N <= 1024): the ping-pong X, the
// K-split accumulators, the partial sums, the bounds, the state and the
// Gershgorin scratch, each at a 256-byte aligned offset.  Every device the
// library runs on holds its own copy for the life of the process; calls that
// overlap on one device share it and race.
constexpr int kNoallocMaxLd = 1024;
constexpr size_t kMatBytes = sizeof(double) * (size_t)kNoallocMaxLd * kNoallocMaxLd;
// K-split accumulators: room for 288 partial 128 x 128 tiles (36 MiB), at
// least what the split rule asks for with ld <= 1024; checked per call
constexpr size_t kSplitBytes = sizeof(double) * 36 * 8 * 128 * 128;
constexpr size_t kPartBytes = 128 * 1024;
constexpr size_t kBoundsBytes = 256;
constexpr size_t kStateBytes = 16 * 1024;
constexpr size_t kGershBytes = 160 * 1024;
constexpr size_t kOffX0 = 0;
constexpr size_t kOffX1 = kOffX0 + kMatBytes;
constexpr size_t kOffSplit = kOffX1 + kMatBytes;
constexpr size_t kOffPart = kOffSplit + kSplitBytes;
constexpr size_t kOffBounds = kOffPart + kPartBytes;
constexpr size_t kOffState = kOffBounds + kBoundsBytes;
constexpr size_t kOffGersh = kOffState + kStateBytes;
constexpr size_t kScratchBytes = kOffGersh + kGershBytes;  // 52.3 MiB
static_assert(sizeof(Sp2State) <= kStateBytes, "state slot too small");
static_assert(kOffSplit % 256 == 0 && kOffPart % 256 == 0 && kOffState % 256 == 0 &&
                  kOffGersh % 256 == 0,
              "scratch offsets must stay 256-byte aligned");

__device__ __align__(256) char scratch[kScratchBytes];

// Device address of `scratch` per device (host code uses these, never the
// array's name), and the host state images of the no-allocation loop.
constexpr int kMaxDevices = 64;
char* g_scratch_ptr[kMaxDevices] = {};
Sp2State g_noalloc_host[3];

int g_precision = PRG_PREC_FP64;

// Optional per-launch timing of the product kernel (HIP events on the library
// stream), used by the benchmark's roofline figure.
bool g_timing = false;
double g_step_ms_sum = 0.0;
int g_step_count = 0;
std::vector<hipEvent_t> g_events;

// Events of the pipelined header read-back (two slots).
hipEvent_t g_head_ev[2] = {nullptr, nullptr};

// Iteration count of the last solve per (n, variant, precision): the chunk
// lengths of the next solve of that shape.
struct ShapeKey {
  int n, variant, precision;
  bool operator<(const ShapeKey& o) const {
    if (n != o.n) return n < o.n;
    if (variant != o.variant) return variant < o.variant;
    return precision < o.precision;
  }
};
std::map<ShapeKey, int> g_iterations;

int expected_iterations(const ShapeKey& k) {
  auto it = g_iterations.find(k);
  return it == g_iterations.end() ? 0 : it->second;
}

void record_iterations(const ShapeKey& k, int iters) { g_iterations[k] = iters; }

bool fortran_eq(const char* s, const char* lit) {
  if (!s) return false;
  size_t n = std::strlen(s);
  while (n > 0 && s[n - 1] == ' ') --n;
  return n == std::strlen(lit) && std::strncmp(s, lit, n) == 0;
}
}  // namespace

int noalloc_setup() {
  int dev = 0;
  PRG_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error(ERR_ARG, "no-allocation scratch: device index out of range");
    return ERR_ARG;
  }
  if (g_scratch_ptr[dev]) return OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    set_error(ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return ERR_HIP;
  }
  g_scratch_ptr[dev] = (char*)ptr;
  return OK;
}

int precision_mode() { return g_precision; }
void set_precision_mode(int m) { g_precision = m; }

int run_sp2(const DMat& H, const Sp2Params& p_in, DMat out, Sp2Result* res, const DMat* H1,
            DMat* out1, const Sp2Buffers* bufs) {
  // the sparse formats' thresholding runs with FP64 products
  Sp2Params p = p_in;
  if (p.threshold > 0.0) p.precision = PRG_PREC_FP64;
  int st = bufs ? OK : ensure_device();
  if (st) return st;
  const int n = H.n, ld = H.ld;
  if (n <= 0 || ld != padded(n) || out.ld != ld || out.n != n) {
    set_error(ERR_ARG, "run_sp2: bad matrix layout");
    return ERR_ARG;
  }
  const int nt = ld / kTile;
  const int ntiles = nt * (nt + 1) / 2;
  const size_t mat_bytes = sizeof(double) * (size_t)ld * ld;
  if (bufs && (H1 || p.variant == V_RESPONSE || p.precision != PRG_PREC_FP64)) {
    set_error(ERR_ARG, "run_sp2: caller buffers serve FP64 solves without response only");
    return ERR_ARG;
  }
  double *x0, *x1, *part, *bounds;
  Sp2State *dst, *hst;
  if (bufs) {
    x0 = bufs->x0;
    x1 = bufs->x1;
    part = bufs->part;
    bounds = bufs->bounds;
    dst = bufs->state;
    hst = bufs->host;
  } else {
    x0 = (double*)workspace("sp2_x0", mat_bytes, &st);
    x1 = (double*)workspace("sp2_x1", mat_bytes, &st);
    // partial sums: one slot per product tile (the int8 CRT update writes one
    // per 64 x 64 block) plus the normalisation's
    // (the K-split epilogue writes four or eight per tile)
    const int nslots = std::max(8 * ntiles, crt_update_blocks(ld));
    part = (double*)workspace("sp2_part", sizeof(double) * 4 * (size_t)(nslots + 2048), &st);
    bounds = (double*)workspace("sp2_bounds", sizeof(double) * 4, &st);
    dst = (Sp2State*)workspace("sp2_state", sizeof(Sp2State), &st);
    // pinned host memory: the state image, then two header slots that the
    // pipelined loop below reads back alternately
    hst = (Sp2State*)host_staging(3 * sizeof(Sp2State), &st);
  }
  if (st) return st;
  DMat X0{n, ld, x0}, X1{n, ld, x1};
  const bool response = (p.variant == V_RESPONSE);
  DMat D0{n, ld, nullptr}, D1{n, ld, nullptr};
  if (response) {
    D0.p = (double*)workspace("sp2_d0", mat_bytes, &st);
    D1.p = (double*)workspace("sp2_d1", mat_bytes, &st);
    if (st) return st;
  }

  // X0 = (hi I - H)/(hi - lo)
  int nb = 0;
  if (p.use_gershgorin) {
    if ((st = gershgorin(H, bounds, bufs ? bufs->gersh : nullptr))) return st;
    if ((st = normalize_gershgorin(H, bounds, X0, part, &nb))) return st;
  } else {
    if ((st = normalize_bounds(H, p.mineval, p.maxeval, X0, part, &nb))) return st;
  }
  double gb[2] = {p.mineval, p.maxeval};
  if (p.use_gershgorin) {
    // 16-byte read-back: validates the interval (zero width / NaN input)
    PRG_HIP_CHECK(hipMemcpyAsync(hst, bounds, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()));
    PRG_HIP_CHECK(hipStreamSynchronize(stream()));
    gb[0] = ((double*)hst)[0];
    gb[1] = ((double*)hst)[1];
  }
  if (!std::isfinite(gb[0]) || !std::isfinite(gb[1]) || !(gb[1] > gb[0])) {
    set_error(ERR_NUMERIC, "SP2: spectral interval is empty or not finite (zero or NaN input?)");
    return ERR_NUMERIC;
  }
  if (res) {
    res->mineval = gb[0];
    res->maxeval = gb[1];
  }
  if (response) {
    // Delta0 = -H1/(hi - lo)  (src/prg_response_mod.F90:694-700)
    if (!(gb[1] - gb[0] != 0.0)) {
      set_error(ERR_NUMERIC, "response: zero-width spectral interval");
      return ERR_NUMERIC;
    }
    if ((st = affine(*H1, -1.0 / (gb[1] - gb[0]), 0.0, D0, nullptr, nullptr))) return st;
  }

  if (p.precision != PRG_PREC_FP64 && !response) {
    if ((st = sp2_mixed_prepare(X0, p.precision))) return st;
  }
  std::memset(hst, 0, sizeof(Sp2State));
  hst->occ = p.occ;
  hst->idemtol = p.idemtol;
  hst->variant = p.variant;
  hst->minit = p.minit;
  hst->maxit = p.maxit;
  hst->rel = p.rel;
  hst->absconv = p.absconv;
  hst->icount = p.icount;
  hst->core = p.core;
  hst->threshold = p.threshold > 0.0 ? p.threshold : 0.0;
  const bool lowp = (p.precision != PRG_PREC_FP64) && !response;
  // FP32-accumulated products stop on their noise floor and are polished in
  // FP64; the int8 CRT products carry X to 2^-42 and keep the FP64 rules.
  hst->noise_stop = lowp &&
                    (p.precision == PRG_PREC_F16SPLIT || p.precision == PRG_PREC_BF16SPLIT) &&
                    (p.variant == V_ALG2);
  hst->refine_max = hst->noise_stop ? p.refine_max : 0;
  if (p.pp_in)
    for (int i = 0; i < p.icount && i < kMaxIter; ++i) hst->pp[i] = p.pp_in[i];
  if (p.vv_in)
    for (int i = 0; i < p.icount && i < kMaxIter; ++i) hst->vv[i] = p.vv_in[i];
  PRG_HIP_CHECK(hipMemcpyAsync(dst, hst, sizeof(Sp2State), hipMemcpyHostToDevice, stream()));
  if ((st = sp2_init_state(dst, part, nb))) return st;

  const int maxsteps = (p.variant == V_SEQ_ALG2 || p.variant == V_SEQ_ALG1 ||
                        p.variant == V_SUBMATRIX)
                           ? p.icount
                           : p.maxit + 2;
  const bool replay = (maxsteps == p.icount) && p.variant != V_ALG2 && p.variant != V_ALG1;
  const size_t head = offsetof(Sp2State, pp);
  int prec = p.precision;
  const bool timing = g_timing && !bufs;
  if (timing) crt_timing_begin();
  if (timing && (int)g_events.size() < 2 * (maxsteps + 1)) {
    const size_t old = g_events.size();
    g_events.resize(2 * (maxsteps + 1));
    for (size_t e = old; e < g_events.size(); ++e) PRG_HIP_CHECK(hipEventCreate(&g_events[e]));
  }
  if (!bufs && !g_head_ev[0]) {
    PRG_HIP_CHECK(hipEventCreateWithFlags(&g_head_ev[0], hipEventDisableTiming));
    PRG_HIP_CHECK(hipEventCreateWithFlags(&g_head_ev[1], hipEventDisableTiming));
  }
  // Chunked, pipelined loop: chunk k + 1 is enqueued before the host waits
  // for the header of chunk k (copied asynchronously into a pinned slot), so
  // the device never idles on the host.  Once the state says done, the
  // launches still queued exit at their first instruction.  Chunk lengths
  // follow the iteration count of the last solve of the same shape (one long
  // chunk, then short ones), otherwise 16 steps (8 from n = 4096).
  const ShapeKey key{n, p.variant, p.precision};
  const int expect = (replay || bufs) ? 0 : expected_iterations(key);
  Sp2State* slot[2] = {hst + 1, hst + 2};
  int phase0_steps = 0;
  for (int phase = 0; phase < 2; ++phase) {
    int steps = 0;
    int phase_steps = maxsteps;
    if (phase == 1) {
      // FP64 polishing steps after a low-precision loop
      if (!hst->noise_stop || hst->refine_max <= 0) break;
      prec = PRG_PREC_FP64;
      if ((st = sp2_begin_refine(dst))) return st;
      phase_steps = hst->refine_max + 1;
    }
    int chunk_no = 0;
    bool waiting = false;  // a header copy is in flight for the previous chunk
    bool done = false;
    while (steps < phase_steps && !done) {
      int len;
      if (phase == 0 && replay)
        len = maxsteps;
      else if (phase == 0 && expect > 0)
        len = (chunk_no == 0) ? std::max(1, expect - 2) : 4;
      else
        len = n >= 4096 ? 8 : 16;
      for (int c = 0; c < len && steps < phase_steps; ++c, ++steps) {
        int nt_used = 0;
        const bool timed = timing && phase == 0;
        if (timed) PRG_HIP_CHECK(hipEventRecord(g_events[2 * steps], stream()));
        if (response) {
          if ((st = sp2_response_step(X0, X1, D0, D1, dst, part, &nt_used))) return st;
        }
        if (prec == PRG_PREC_FP64 || response) {
          st = sp2_step(X0, X1, dst, part, &nt_used, bufs ? bufs->split_acc : nullptr,
                        p.threshold > 0.0);
        } else if (prec == PRG_PREC_F16SPLIT || prec == PRG_PREC_BF16SPLIT ||
                   prec == PRG_PREC_INT8CRT) {
          st = sp2_step_mixed(X0, X1, dst, part, &nt_used, prec);
        } else {
          set_error(ERR_ARG, "unknown precision mode");
          return ERR_ARG;
        }
        if (st) return st;
        if (timed) PRG_HIP_CHECK(hipEventRecord(g_events[2 * steps + 1], stream()));
        if ((st = sp2_absorb_plan(dst, part, nt_used))) return st;
      }
      const int sl = chunk_no & 1;
      PRG_HIP_CHECK(hipMemcpyAsync(slot[sl], dst, head, hipMemcpyDeviceToHost, stream()));
      if (bufs) {
        // caller buffers: a plain synchronous read-back, no events
        PRG_HIP_CHECK(hipStreamSynchronize(stream()));
        done = slot[sl]->done != 0;
        ++chunk_no;
        continue;
      }
      PRG_HIP_CHECK(hipEventRecord(g_head_ev[sl], stream()));
      if (waiting) {
        PRG_HIP_CHECK(hipEventSynchronize(g_head_ev[sl ^ 1]));
        done = slot[sl ^ 1]->done != 0;
      }
      waiting = true;
      ++chunk_no;
    }
    if (!done && waiting) {
      // the last chunk's header
      PRG_HIP_CHECK(hipEventSynchronize(g_head_ev[(chunk_no - 1) & 1]));
      done = slot[(chunk_no - 1) & 1]->done != 0;
    }
    if (phase == 0) phase0_steps = steps;
  }
  // the int8 CRT update keeps X in its residue planes (or the lower
  // triangle of an FP64 master at 12 or more moduli) between steps
  if (p.precision == PRG_PREC_INT8CRT && !response) {
    if ((st = crt_finish(X0, X1, dst, 2.0, out, crt_moduli()))) return st;
  } else if ((st = sp2_finish(X0, X1, dst, 2.0, out))) {
    return st;
  }
  if (response && out1) {
    // rho1 = 2 Delta, Delta in the ping-pong slot named by the state
    if ((st = sp2_finish(D0, D1, dst, 2.0, *out1))) return st;
  }
  PRG_HIP_CHECK(hipMemcpyAsync(hst, dst, sizeof(Sp2State), hipMemcpyDeviceToHost, stream()));
  PRG_HIP_CHECK(hipStreamSynchronize(stream()));
  if (!replay && !bufs) record_iterations(key, hst->iter);
  if (!hst->done) {
    set_error(ERR_NOT_CONVERGED, "SP2 loop did not finish within the step budget");
    return ERR_NOT_CONVERGED;
  }
  if (timing) {
    // products of the main (phase 0) loop
    const int main_iters = hst->iter - hst->refine_steps;
    const int executed = main_iters < phase0_steps ? main_iters : phase0_steps;
    for (int e = 0; e < executed; ++e) {
      float ms = 0.f;
      PRG_HIP_CHECK(hipEventElapsedTime(&ms, g_events[2 * e], g_events[2 * e + 1]));
      g_step_ms_sum += ms;
      ++g_step_count;
    }
    if ((st = crt_timing_collect(executed))) return st;
  }
  if (res) {
    res->refine_steps = hst->refine_steps;
    res->iter = hst->iter;
    res->status = hst->status;
    res->trx = hst->trx;
    std::memcpy(res->pp, hst->pp, sizeof(res->pp));
    std::memcpy(res->vv, hst->vv, sizeof(res->vv));
  }
  if (hst->status != OK) {
    set_error(hst->status, "SP2 purification is not converging");
    return hst->status;
  }
  return OK;
}

}  // namespace prg

using namespace prg;

namespace {

Sp2Params base_params(int variant, int n, double bndfil, int minit, int maxit,
                      const char* sp2conv, double idemtol) {
  Sp2Params p;
  p.variant = variant;
  p.occ = bndfil * (double)n;
  p.idemtol = idemtol;
  p.minit = minit;
  p.maxit = maxit > kMaxIter - 2 ? kMaxIter - 2 : maxit;
  p.rel = fortran_rel(sp2conv);
  p.absconv = fortran_abs(sp2conv);
  p.precision = precision_mode();
  return p;
}

void report(const char* who, int verbose, const Sp2Result& r) {
  if (verbose >= 1) std::printf("%s: iterations = %d, trace(X) = %.15g\n", who, r.iter, r.trx);
}

// Shared body of the H -> rho variants.
// The threshold applies when rho is stored in a sparse BML format (ellpack,
// ellsort, ellblock, csr): those formats drop |a| <= threshold from every
// product and sum.  Dense storage ignores it, as BML dense does.
double format_threshold(const bml_matrix_t* rho, double threshold) {
  return (rho && rho->type != dense && rho->type != type_uninitialized && threshold > 0.0)
             ? threshold
             : 0.0;
}

void sp2_h_to_rho(const char* who, int variant, bml_matrix_t* h, bml_matrix_t* rho, double bndfil,
                  int minit, int maxit, const char* sp2conv, double idemtol, int verbose,
                  int* pp, int* icount, double* vv, int ppsize, double threshold) {
  clear_error();
  if (check_pair(h, rho, who)) return;
  DMat H, R;
  if (bml_dev_read(h, &H)) return;
  if (h == rho) {
    // in-place call: stage the input in a workspace first
    int st = OK;
    double* tmp = (double*)workspace("sp2_hcopy", sizeof(double) * (size_t)H.ld * H.ld, &st);
    if (st) return;
    DMat T{H.n, H.ld, tmp};
    if (copy_block(H, T)) return;
    H = T;
  }
  if (bml_dev_write(rho, &R)) return;
  Sp2Params p = base_params(variant, h->N, bndfil, minit, maxit, sp2conv, idemtol);
  p.threshold = format_threshold(rho, threshold);
  Sp2Result r;
  const int st = run_sp2(H, p, R, &r);
  if (st != OK && st != ERR_NOT_CONVERGED) return;
  report(who, verbose, r);
  if (icount) *icount = r.iter;
  // vv[iter] is written only when the loop ended on a discarded product
  const int nrec_full = (r.vv[r.iter] != 0.0) ? r.iter + 1 : r.iter;
  const int nrec = nrec_full < ppsize ? nrec_full : ppsize;
  if (pp)
    for (int i = 0; i < r.iter && i < ppsize; ++i) pp[i] = r.pp[i];
  if (vv)
    for (int i = 0; i < nrec; ++i) vv[i] = r.vv[i];
}

void sp2_replay(const char* who, int variant, bml_matrix_t* h, bml_matrix_t* rho, int* pp,
                int* icount, double* vv, bool gersh, double mineval, double maxeval, int core,
                double threshold) {
  clear_error();
  if (!icount || !pp || !vv) {
    set_error(ERR_ARG, std::string(who) + ": null pp / icount / vv");
    return;
  }
  if (*icount < 0 || *icount > kMaxIter) {
    set_error(ERR_ARG, std::string(who) + ": icount out of range");
    return;
  }
  if (check_pair(h, rho, who)) return;
  DMat H, R;
  if (bml_dev_read(h, &H)) return;
  if (h == rho) {
    int st = OK;
    double* tmp = (double*)workspace("sp2_hcopy", sizeof(double) * (size_t)H.ld * H.ld, &st);
    if (st) return;
    DMat T{H.n, H.ld, tmp};
    if (copy_block(H, T)) return;
    H = T;
  }
  if (bml_dev_write(rho, &R)) return;
  Sp2Params p;
  p.variant = variant;
  p.icount = *icount;
  p.pp_in = pp;
  p.vv_in = (variant == V_SUBMATRIX) ? vv : nullptr;
  p.use_gershgorin = gersh;
  p.mineval = mineval;
  p.maxeval = maxeval;
  p.core = core;
  p.precision = precision_mode();
  p.threshold = format_threshold(rho, threshold);
  Sp2Result r;
  if (run_sp2(H, p, R, &r) != OK) return;
  for (int i = 0; i < *icount; ++i) vv[i] = r.vv[i];
}

}  // namespace

extern "C" {

void prg_set_precision(int mode) {
  clear_error();
  if (mode != PRG_PREC_FP64 && mode != PRG_PREC_F16SPLIT && mode != PRG_PREC_INT8CRT &&
      mode != PRG_PREC_BF16SPLIT) {
    set_error(ERR_ARG, "prg_set_precision: unknown mode");
    return;
  }
  set_precision_mode(mode);
}
int prg_get_precision(void) { return precision_mode(); }
int prg_set_int8_moduli(int count) {
  clear_error();
  return set_crt_moduli(count);
}
int prg_get_int8_moduli(void) { return crt_moduli(); }
int prg_gpu_int8_waves(void) { return crt_last_waves(); }
int prg_gpu_sp2_split_slices(int N) { return N > 0 ? sp2_split_count(padded(N)) : 0; }
int prg_gpu_set_int8_config(int config) {
  clear_error();
  return set_crt_config(config);
}

void prg_sp2_basic(bml_matrix_t* h_bml, bml_matrix_t* rho_bml, double threshold, double bndfil,
                   int minsp2iter, int maxsp2iter, const char* sp2conv, double idemtol,
                   int verbose) {
  sp2_h_to_rho("prg_sp2_basic", V_BASIC, h_bml, rho_bml, bndfil, minsp2iter, maxsp2iter, sp2conv,
               idemtol, verbose, nullptr, nullptr, nullptr, 0, threshold);
}

// src/prg_c_interface.F90:1091-1111 (Fortran src/prg_sp2_mod.F90:138-246): the
// basic branch rule for exactly maxsp2iter + 1 iterations (the reference
// comments its convergence exit out), D = 2X.  rhofull is the reference
// density its per-iteration printout compares against; here, with
// verbose > 1, the final ||rhofull/2 - X||_F / ||rhofull||_F is printed.
void prg_sp2_basic_tcore(bml_matrix_t* h_bml, bml_matrix_t* rho_bml, bml_matrix_t* rhofull_bml,
                         double threshold, double bndfil, int minsp2iter, int maxsp2iter,
                         const char* sp2conv, double idemtol, int verbose) {
  sp2_h_to_rho("prg_sp2_basic_tcore", V_BASIC_FIXED, h_bml, rho_bml, bndfil, minsp2iter,
               maxsp2iter, sp2conv, idemtol, verbose, nullptr, nullptr, nullptr, 0, threshold);
  if (prg_last_error_code() != OK || verbose <= 1 || !rhofull_bml) return;
  if (check_pair(rho_bml, rhofull_bml, "prg_sp2_basic_tcore")) return;
  DMat R, F, T;
  if (bml_dev_read(rho_bml, &R) || bml_dev_read(rhofull_bml, &F) ||
      ops::temp("tcore_diff", R.n, &T))
    return;
  ops::Sums d, f;
  if (ops::lin(F, 0.5, &R, -0.5, 0.0, T, &d) || ops::sums_of(F, &f)) return;
  std::printf("prg_sp2_basic_tcore: ||rhofull/2 - X||_F / ||rhofull||_F = %.6e\n",
              f.fro2 > 0.0 ? std::sqrt(d.fro2 / f.fro2) : 0.0);
}

// src/sp2_tensorcore/src/prg_sp2_tensorcore.cu:131-141 (Fortran binding
// src/sp2_tensorcore/prg_sp2_tensorcore_mod.F90): host single-precision H in,
// host double D = X out (not 2X).  Runs this library's FP16 hi+lo split SP2
// (the reference routine's split products, sigma rule, noise-floor stop and
// FP64 refinement) with Nocc = int(bndfil N) as the reference counts it.
// Differences: Gershgorin bounds instead of the reference's hard-coded
// (-27.55, 25.0), and up to two FP64 SP2 polishing steps in place of its
// single 2T^2 - T^4 step.  eps, minsp2iter, sp2conv and idemtol are accepted
// as in the reference, which does not use them on this path either.
void prg_sp2_tensorcore(int N, float* H, double* D, float eps, float bndfil, int minsp2iter,
                        int maxsp2iter, char sp2conv, float idemtol, int verbose) {
  (void)eps;
  (void)sp2conv;
  clear_error();
  if (N <= 0 || !H || !D || maxsp2iter < 1) {
    set_error(ERR_ARG, "prg_sp2_tensorcore: bad arguments");
    return;
  }
  const int nocc = (int)(bndfil * (float)N);
  bml_matrix_t* h = bml_zero_matrix(dense, double_real, N, N, sequential);
  bml_matrix_t* r = h ? bml_zero_matrix(dense, double_real, N, N, sequential) : nullptr;
  if (!h || !r) {
    if (h) bml_deallocate(&h);
    return;
  }
  for (size_t i = 0; i < (size_t)N * N; ++i) h->host[i] = (double)H[i];
  int st = OK;
  DMat Hd, Rd;
  if (!(st = bml_dev_read(h, &Hd)) && !(st = bml_dev_write(r, &Rd))) {
    Sp2Params p = base_params(V_ALG2, N, (double)nocc / (double)N, minsp2iter, maxsp2iter, "Abs",
                              (double)idemtol);
    p.precision = PRG_PREC_F16SPLIT;
    Sp2Result res;
    st = run_sp2(Hd, p, Rd, &res);
    if (st == OK || st == ERR_NOT_CONVERGED) {
      report("prg_sp2_tensorcore", verbose, res);
      if (!(st = bml_host_sync(r)))
        for (size_t i = 0; i < (size_t)N * N; ++i) D[i] = 0.5 * r->host[i];
    }
  }
  bml_deallocate(&h);
  bml_deallocate(&r);
}

void prg_sp2_alg2(bml_matrix_t* h_bml, bml_matrix_t* rho_bml, double threshold, double bndfil,
                  int minsp2iter, int maxsp2iter, const char* sp2conv, double idemtol,
                  int verbose) {
  sp2_h_to_rho("prg_sp2_alg2", V_ALG2, h_bml, rho_bml, bndfil, minsp2iter, maxsp2iter, sp2conv,
               idemtol, verbose, nullptr, nullptr, nullptr, 0, threshold);
}

void prg_sp2_alg2_genseq(bml_matrix_t* h_bml, bml_matrix_t* rho_bml, double threshold,
                         double bndfil, int minsp2iter, int maxsp2iter, const char* sp2conv,
                         double idemtol, int* pp, int* icount, double* vv, int verbose) {
  sp2_h_to_rho("prg_sp2_alg2_genseq", V_ALG2_GENSEQ, h_bml, rho_bml, bndfil, minsp2iter,
               maxsp2iter, sp2conv, idemtol, verbose, pp, icount, vv, maxsp2iter, threshold);
}

void prg_sp2_alg2_seq(bml_matrix_t* h_bml, bml_matrix_t* rho_bml, double threshold, int* pp,
                      int* icount, double* vv, int verbose) {
  (void)verbose;
  sp2_replay("prg_sp2_alg2_seq", V_SEQ_ALG2, h_bml, rho_bml, pp, icount, vv, true, 0, 0, 0,
             threshold);
}

void prg_prg_sp2_alg2_seq_inplace(bml_matrix_t* rho_bml, double threshold, int ppsize, int* pp,
                                  int* icount, double* vv, double mineval, double maxeval) {
  (void)ppsize;
  sp2_replay("prg_prg_sp2_alg2_seq_inplace", V_SEQ_ALG2, rho_bml, rho_bml, pp, icount, vv, false,
             mineval, maxeval, 0, threshold);
}

void prg_sp2_alg1(bml_matrix_t* h_bml, bml_matrix_t* rho_bml, double threshold, double bndfil,
                  int minsp2iter, int maxsp2iter, const char* sp2conv, double idemtol,
                  int verbose) {
  sp2_h_to_rho("prg_sp2_alg1", V_ALG1, h_bml, rho_bml, bndfil, minsp2iter, maxsp2iter, sp2conv,
               idemtol, verbose, nullptr, nullptr, nullptr, 0, threshold);
}

void prg_sp2_alg1_genseq(bml_matrix_t* h_bml, bml_matrix_t* rho_bml, double threshold,
                         double bndfil, int minsp2iter, int maxsp2iter, const char* sp2conv,
                         double idemtol, int* pp, int* icount, double* vv) {
  sp2_h_to_rho("prg_sp2_alg1_genseq", V_ALG1_GENSEQ, h_bml, rho_bml, bndfil, minsp2iter,
               maxsp2iter, sp2conv, idemtol, 0, pp, icount, vv, maxsp2iter, threshold);
}

void prg_sp2_alg1_seq(bml_matrix_t* h_bml, bml_matrix_t* rho_bml, double threshold, int* pp,
                      int* icount, double* vv) {
  sp2_replay("prg_sp2_alg1_seq", V_SEQ_ALG1, h_bml, rho_bml, pp, icount, vv, true, 0, 0, 0,
             threshold);
}

void prg_prg_sp2_alg1_seq_inplace(bml_matrix_t* rho_bml, double threshold, int ppsize, int* pp,
                                  int* icount, double* vv, double mineval, double maxeval) {
  (void)ppsize;
  sp2_replay("prg_prg_sp2_alg1_seq_inplace", V_SEQ_ALG1, rho_bml, rho_bml, pp, icount, vv, false,
             mineval, maxeval, 0, threshold);
}

void prg_sp2_submatrix(bml_matrix_t* ham_bml, bml_matrix_t* rho_bml, double threshold, int* pp,
                       int* icount, double* vv, double mineval, double maxeval, int core_size) {
  sp2_replay("prg_sp2_submatrix", V_SUBMATRIX, ham_bml, rho_bml, pp, icount, vv, false, mineval,
             maxeval, core_size, threshold);
}

void prg_sp2_submatrix_inplace(bml_matrix_t* rho_bml, double threshold, int* pp, int* icount,
                               double* vv, double mineval, double maxeval, int core_size) {
  sp2_replay("prg_sp2_submatrix_inplace", V_SUBMATRIX, rho_bml, rho_bml, pp, icount, vv, false,
             mineval, maxeval, core_size, threshold);
}

void prg_gershgorin(bml_matrix_t* h_bml, double* gbnd) {
  clear_error();
  if (!gbnd) {
    set_error(ERR_ARG, "prg_gershgorin: null output");
    return;
  }
  DMat H;
  if (bml_dev_read(h_bml, &H)) return;
  int st = OK;
  double* b = (double*)workspace("gersh_out", 2 * sizeof(double), &st);
  if (st || gershgorin(H, b)) return;
  if (hipMemcpyAsync(gbnd, b, 2 * sizeof(double), hipMemcpyDeviceToHost, stream()) != hipSuccess ||
      hipStreamSynchronize(stream()) != hipSuccess)
    set_error(ERR_HIP, "prg_gershgorin: copy failed");
}

void prg_normalize(bml_matrix_t* h_bml) {
  clear_error();
  DMat H;
  if (bml_dev_read(h_bml, &H)) return;
  int st = OK;
  double* b = (double*)workspace("gersh_out", 2 * sizeof(double), &st);
  if (st || gershgorin(H, b)) return;
  h_bml->host_valid = false;
  normalize_gershgorin(H, b, H, nullptr, nullptr);
}

void prg_normalize_fermi(bml_matrix_t* h_bml, double h1, double hN, double mu) {
  clear_error();
  DMat H;
  if (bml_dev_read(h_bml, &H)) return;
  h_bml->host_valid = false;
  const double inv = 1.0 / (hN - h1);
  affine(H, -inv, (hN - mu) * inv, H, nullptr, nullptr);
}

void prg_normalize_implicit_fermi(bml_matrix_t* h_bml, double cnst, double mu) {
  clear_error();
  DMat H;
  if (bml_dev_read(h_bml, &H)) return;
  h_bml->host_valid = false;
  affine(H, -cnst, 0.5 + cnst * mu, H, nullptr, nullptr);
}

int prg_gpu_padded_ld(int N) { return padded(N); }

void prg_gpu_timing(int enable) {
  g_timing = enable != 0;
  g_step_ms_sum = 0.0;
  g_step_count = 0;
  crt_timing(enable != 0);
}

int prg_gpu_step_timing(double* ms_sum, int* count) {
  if (ms_sum) *ms_sum = g_step_ms_sum;
  if (count) *count = g_step_count;
  return OK;
}

int prg_gpu_int8_product_timing(double* ms_sum, int* count) {
  crt_timing_get(ms_sum, count);
  return OK;
}

int prg_gpu_sp2_alg2_device_noalloc(int N, const double* dH, double* dRho, double bndfil,
                                    int minsp2iter, int maxsp2iter, const char* sp2conv,
                                    double idemtol, int* iters, int* pp) {
  clear_error();
  if (N <= 0 || !dH || !dRho) {
    set_error(ERR_ARG, "prg_gpu_sp2_alg2_device_noalloc: bad arguments");
    return ERR_ARG;
  }
  const int ld = padded(N);
  if (ld > kNoallocMaxLd) {
    set_error(ERR_ARG, "prg_gpu_sp2_alg2_device_noalloc: N exceeds the static scratch "
                       "(N <= 1024)");
    return ERR_ARG;
  }
  if (!device_ready()) {
    set_error(ERR_NO_DEVICE, "prg_gpu_sp2_alg2_device_noalloc: call prg_progress_init first");
    return ERR_NO_DEVICE;
  }
  int st = noalloc_setup();
  if (st) return st;
  int dev = 0;
  PRG_HIP_CHECK(hipGetDevice(&dev));
  char* base = g_scratch_ptr[dev];
  if (sp2_split_bytes(ld) > kSplitBytes || gershgorin_scratch_doubles(N) * 8 > kGershBytes) {
    set_error(ERR_ARG, "prg_gpu_sp2_alg2_device_noalloc: scratch layout too small");
    return ERR_ARG;
  }
  Sp2Buffers b;
  b.x0 = (double*)(base + kOffX0);
  b.x1 = (double*)(base + kOffX1);
  b.split_acc = (double*)(base + kOffSplit);
  b.part = (double*)(base + kOffPart);
  b.bounds = (double*)(base + kOffBounds);
  b.state = (Sp2State*)(base + kOffState);
  b.gersh = (double*)(base + kOffGersh);
  b.host = g_noalloc_host;
  DMat H{N, ld, const_cast<double*>(dH)}, R{N, ld, dRho};
  Sp2Params p = base_params(V_ALG2, N, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol);
  p.precision = PRG_PREC_FP64;
  Sp2Result r;
  st = run_sp2(H, p, R, &r, nullptr, nullptr, &b);
  if (iters) *iters = r.iter;
  if (pp)
    for (int i = 0; i < r.iter && i < maxsp2iter; ++i) pp[i] = r.pp[i];
  return st;
}

int prg_gpu_sp2_alg2_device(int N, const double* dH, double* dRho, double bndfil, int minsp2iter,
                            int maxsp2iter, const char* sp2conv, double idemtol, int precision,
                            int* iters, int* pp) {
  clear_error();
  if (N <= 0 || !dH || !dRho) {
    set_error(ERR_ARG, "prg_gpu_sp2_alg2_device: bad arguments");
    return ERR_ARG;
  }
  if (precision != PRG_PREC_FP64 && precision != PRG_PREC_F16SPLIT &&
      precision != PRG_PREC_INT8CRT && precision != PRG_PREC_BF16SPLIT) {
    set_error(ERR_ARG, "prg_gpu_sp2_alg2_device: unknown precision");
    return ERR_ARG;
  }
  const int ld = padded(N);
  DMat H{N, ld, const_cast<double*>(dH)}, R{N, ld, dRho};
  Sp2Params p = base_params(V_ALG2, N, bndfil, minsp2iter, maxsp2iter, sp2conv, idemtol);
  p.precision = precision;
  Sp2Result r;
  const int st = run_sp2(H, p, R, &r);
  if (iters) *iters = r.iter;
  if (pp)
    for (int i = 0; i < r.iter && i < maxsp2iter; ++i) pp[i] = r.pp[i];
  return st;
}

}  // extern "C"

namespace prg {
bool fortran_rel(const char* s) { return fortran_eq(s, "Rel"); }
bool fortran_abs(const char* s) { return fortran_eq(s, "Abs"); }
}  // namespace prg
