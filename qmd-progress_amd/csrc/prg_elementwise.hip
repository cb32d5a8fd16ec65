// HBM-bound kernels of the density-matrix path (gfx950) and the device-side
// FP64 bookkeeping of the SP2 loop.
//
// All dense kernels sweep the ld x ld zero-padded column-major buffers with
// one thread per row element of a column (64 consecutive doubles per wave per
// load: fully coalesced) and grid-stride over columns.  Per-block partial sums
// are written to a [nblocks][nq] array and reduced in a fixed order, so every
// trace / Frobenius value is bitwise reproducible run to run.
#include "prg_internal.h"

#include <cmath>

namespace prg {
namespace {

constexpr int NT = 256;

template <int NQ>
__device__ __forceinline__ void block_partials(double (&v)[NQ], double* dst) {
  __shared__ double red[NT / 64][NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    double x = v[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    v[q] = x;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) red[w][q] = v[q];
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    double s = 0.0;
    for (int ww = 0; ww < NT / 64; ++ww) s += red[ww][threadIdx.x];
    dst[threadIdx.x] = s;
  }
}

// Row sums of |h_ij| (j != i) over a chunk of columns, one thread per row.
__global__ void gersh_rows_kernel(const double* __restrict__ h, int n, int ld, int cchunk,
                                  double* __restrict__ rs) {
  const int i = blockIdx.x * NT + threadIdx.x;
  const int c0 = blockIdx.y * cchunk;
  const int c1 = min(n, c0 + cchunk);
  if (i >= n) return;
  double s = 0.0;
  for (int j = c0; j < c1; ++j) {
    const double v = h[i + (size_t)j * ld];
    if (j != i) s += fabs(v);
  }
  rs[(size_t)blockIdx.y * n + i] = s;
}

// Per row: lo_i = h_ii - r_i, hi_i = h_ii + r_i; block min / max.
__global__ void gersh_minmax_kernel(const double* __restrict__ h, int n, int ld, int nchunks,
                                    const double* __restrict__ rs, double* __restrict__ bmm) {
  __shared__ double slo[NT], shi[NT];
  const int i = blockIdx.x * NT + threadIdx.x;
  double lo = INFINITY, hi = -INFINITY;
  if (i < n) {
    double r = 0.0;
    for (int c = 0; c < nchunks; ++c) r += rs[(size_t)c * n + i];
    const double d = h[i + (size_t)i * ld];
    lo = d - r;
    hi = d + r;
  }
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + s]);
      shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    bmm[2 * blockIdx.x] = slo[0];
    bmm[2 * blockIdx.x + 1] = shi[0];
  }
}

__global__ void gersh_final_kernel(const double* __restrict__ bmm, int nb, double* __restrict__ out) {
  __shared__ double slo[NT], shi[NT];
  double lo = INFINITY, hi = -INFINITY;
  for (int b = threadIdx.x; b < nb; b += NT) {
    lo = fmin(lo, bmm[2 * b]);
    hi = fmax(hi, bmm[2 * b + 1]);
  }
  slo[threadIdx.x] = lo;
  shi[threadIdx.x] = hi;
  __syncthreads();
  for (int s = NT / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      slo[threadIdx.x] = fmin(slo[threadIdx.x], slo[threadIdx.x + s]);
      shi[threadIdx.x] = fmax(shi[threadIdx.x], shi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = slo[0];
    out[1] = shi[0];
  }
}

// out = a*A + b*B + c*I on the n x n block, zero elsewhere in the ld x ld
// buffer; coefficients either given or derived on the device:
//   coef_mode 0: given (a, b, c)
//   coef_mode 1: SP2 normalisation from bounds [lo, hi]:
//                a = -1/(hi-lo), c = hi/(hi-lo), b = 0
// Partials per block: [Tr(out), ||out||^2].
struct AxpbyArgs {
  const double* A;
  const double* B;
  double* out;
  int n, ld;
  double a, b, c;
  int coef_mode;
  const double* bounds;
  double* part;
};

__global__ void axpby_kernel(AxpbyArgs g) {
  double a = g.a, b = g.b, c = g.c;
  if (g.coef_mode == 1) {
    const double lo = g.bounds[0], hi = g.bounds[1];
    const double inv = 1.0 / (hi - lo);
    a = -inv;
    c = hi * inv;
    b = 0.0;
  }
  double sums[2] = {0.0, 0.0};
  const int ld = g.ld;
  for (int col = blockIdx.x; col < ld; col += gridDim.x) {
    for (int row = threadIdx.x; row < ld; row += NT) {
      const size_t idx = row + (size_t)col * ld;
      double v = 0.0;
      if (row < g.n && col < g.n) {
        v = a * g.A[idx];
        if (g.B) v += b * g.B[idx];
        if (row == col) v += c;
      }
      g.out[idx] = v;
      sums[1] += v * v;
      if (row == col) sums[0] += v;
    }
  }
  if (g.part) block_partials<2>(sums, g.part + 2 * (size_t)blockIdx.x);
}

__global__ void scale_copy_kernel(const double* __restrict__ src, int ld_src,
                                  double* __restrict__ dst, int ld_dst, int n, double s) {
  for (int col = blockIdx.x; col < n; col += gridDim.x)
    for (int row = threadIdx.x; row < n; row += NT)
      dst[row + (size_t)col * ld_dst] = s * src[row + (size_t)col * ld_src];
}

__global__ void identity_kernel(double* __restrict__ dst, int ld, int n, double diag) {
  for (int col = blockIdx.x; col < ld; col += gridDim.x)
    for (int row = threadIdx.x; row < ld; row += NT)
      dst[row + (size_t)col * ld] = (row == col && row < n) ? diag : 0.0;
}

// Fixed-order reduction of [nb][nq] partials into out[nq].
__global__ void reduce_partials_kernel(const double* __restrict__ part, int nb, int nq,
                                       double* __restrict__ out) {
  __shared__ double red[NT];
  for (int q = 0; q < nq; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += NT) s += part[(size_t)b * nq + q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
    __syncthreads();
  }
}

// ------------------------------------------------------------ SP2 control
__device__ void block_sum(const double* part, int nb, int nq, double* out /*shared*/) {
  __shared__ double red[NT];
  for (int q = 0; q < nq; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += NT) s += part[(size_t)b * nq + q];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = red[0];
    __syncthreads();
  }
}

// Decide the next SP2 step from the current X's traces (runs in thread 0).
// The branch rules restate src/prg_sp2_mod.F90: alg2 :319-363, genseq
// :469-513, alg1 :801-849, alg1_genseq :941-982, basic :89-119, replay
// variants :581-600 / :1060-1072 / :1288-1300, and the response loop of
// src/prg_response_mod.F90:730-770.
__device__ void sp2_plan(Sp2State* s) {
  s->planned = 0;
  s->apply = 0;
  s->last = 0;
  if (s->done) return;
  const double occ = s->occ, tol = s->idemtol;
  const int v = s->variant;
  const int k = s->iter + 1;
  if (!isfinite(s->fro2) || !isfinite(s->trx) || !isfinite(s->trx_fresh)) {
    s->status = ERR_NUMERIC;
    s->done = 1;
    return;
  }
  switch (v) {
    case V_ALG2:
    case V_ALG2_GENSEQ:
    case V_RESPONSE: {
      if (s->iter >= s->maxit) { s->done = 1; return; }
      const double trx = (v == V_ALG2_GENSEQ) ? s->trx_fresh : s->trx;
      const double trx2 = s->fro2;
      if (s->refine_phase && s->refine_steps >= s->refine_max) { s->done = 1; return; }
      if (s->noise_stop && !s->refine_phase) {
        // low-precision products: stop where the idempotency error stops
        // falling quadratically (numerical noise floor)
        const double e = trx - trx2;
        const int it0 = s->iter;  // 0-based index of the step being planned
        const bool alt = it0 > 2 && s->pp[it0 - 1] != s->pp[it0 - 2];
        s->ie2 = s->ie1;
        s->ie1 = s->ie0;
        s->ie0 = e;
        if (e <= 0.0 || (alt && e >= 4.5 * s->ie2 * s->ie2)) {
          s->done = 1;
          return;
        }
      }
      const double tr2xx2 = 2.0 * trx - trx2;
      const double limdiff = fabs(trx2 - occ) - fabs(tr2xx2 - occ);
      s->trx2 = trx2;
      s->limdiff = limdiff;
      int stop = 0;
      if (limdiff >= tol) {
        s->sigma = 0;
        s->trx_next = tr2xx2;
      } else if (limdiff < -tol) {
        s->sigma = 1;
        s->trx_next = trx2;
      } else {
        stop = 1;
        s->trx_next = trx;
      }
      const double e_new = fabs(s->trx_next - trx);
      const int keff = (stop && v != V_RESPONSE) ? k - 1 : k;
      int last = stop;
      if (s->rel && keff >= s->minit && (e_new >= s->e1 || e_new < tol)) last = 1;
      if (!stop && k == s->maxit) {
        s->status = ERR_NOT_CONVERGED;
        last = 1;
      }
      if (s->refine_phase && s->refine_steps + 1 >= s->refine_max) last = 1;
      s->last = last;
      s->apply = !stop;
      s->planned = !stop || (v == V_ALG2_GENSEQ);
      if (!s->planned) {
        s->trx = s->trx_next;
        s->done = 1;
      }
      break;
    }
    case V_ALG1:
    case V_ALG1_GENSEQ: {
      if (s->iter >= s->maxit) { s->done = 1; return; }
      const double trx = s->trx;
      const double trx2 = s->trx_fresh - s->fro2;  // Tr(X - X^2)
      const double limdiff = fabs(trx - trx2 - occ) - fabs(trx + trx2 - occ);
      s->trx2 = trx2;
      s->limdiff = limdiff;
      int stop = 0;
      if (limdiff >= tol) {
        s->sigma = 0;
        s->trx_next = trx + trx2;
      } else if (limdiff < -tol) {
        s->sigma = 1;
        s->trx_next = trx - trx2;
      } else if (v == V_ALG1 && limdiff == 0.0 && k == 1) {
        s->sigma = 1;
        s->trx_next = trx - trx2;
      } else {
        stop = 1;
        s->trx_next = trx;
      }
      const double e_new = fabs(trx2);
      const int keff = stop ? k - 1 : k;
      int last = stop;
      if (s->rel && keff >= s->minit && (s->e1 <= e_new || e_new < tol)) last = 1;
      if (v == V_ALG1_GENSEQ && s->absconv && fabs(limdiff) < tol) last = 1;
      if (!stop && k == s->maxit) {
        s->status = ERR_NOT_CONVERGED;
        last = 1;
      }
      s->last = last;
      s->apply = !stop;
      s->planned = !stop || (v == V_ALG1_GENSEQ);
      if (!s->planned) s->done = 1;
      break;
    }
    case V_BASIC:
    case V_BASIC_FIXED: {
      // V_BASIC_FIXED (src/prg_sp2_mod.F90:138-246): the same branch rule run
      // for iterations 0..maxit with the convergence exit commented out
      // (:214) and only a message at maxit (:216-219)
      const int it = s->iter;  // reference loop index, starting at 0
      if (it > s->maxit) { s->done = 1; return; }
      const double trx = s->trx_fresh;
      s->trx2 = s->fro2;
      if (trx - occ <= 0.0) {
        s->sigma = 0;
        s->trx_next = 2.0 * trx - s->fro2;
      } else {
        s->sigma = 1;
        s->trx_next = s->fro2;
      }
      int last = (v == V_BASIC_FIXED) ? (it == s->maxit)
                                      : (fabs(occ - s->trx_next) < tol && it > s->minit);
      if (!last && it == s->maxit) {
        s->status = ERR_NOT_CONVERGED;
        last = 1;
      }
      s->last = last;
      s->apply = 1;
      s->planned = 1;
      break;
    }
    case V_SEQ_ALG2:
    case V_SEQ_ALG1:
    case V_SUBMATRIX: {
      if (s->iter >= s->icount) { s->done = 1; return; }
      s->sigma = s->pp[s->iter] != 0;
      s->trx_next = s->sigma ? s->fro2 : 2.0 * s->trx_fresh - s->fro2;
      s->apply = 1;
      s->planned = 1;
      s->last = (k == s->icount);
      break;
    }
    default:
      s->status = ERR_ARG;
      s->done = 1;
  }
}

// Fixed-order block sum of nb records of 4 partials (32 bytes each): one
// pass over the records, wave shuffles, then the NT / 64 wave totals.
__device__ void block_sum4(const double* part, int nb, double* out /*shared*/) {
  __shared__ double red[NT / 64][4];
  double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int b = threadIdx.x; b < nb; b += NT) {
    const double2 lo = *(const double2*)(part + 4 * (size_t)b);
    const double2 hi = *(const double2*)(part + 4 * (size_t)b + 2);
    a[0] += lo.x;
    a[1] += lo.y;
    a[2] += hi.x;
    a[3] += hi.y;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a[q] += __shfl_xor(a[q], off);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[w][q] = a[q];
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double t = 0.0;
#pragma unroll
    for (int ww = 0; ww < NT / 64; ++ww) t += red[ww][threadIdx.x];
    out[threadIdx.x] = t;
  }
  __syncthreads();
}

// Absorb the partials of the product just computed, then plan the next step.
__global__ void sp2_absorb_plan_kernel(Sp2State* s, const double* __restrict__ part, int ntiles) {
  __shared__ double sums[4];
  __shared__ int planned;
  if (threadIdx.x == 0) planned = s->planned;
  __syncthreads();
  if (!planned) return;
  block_sum4(part, ntiles, sums);
  if (threadIdx.x != 0) return;
  const double fro2 = sums[0], trf = sums[1], idem = sums[2];
  s->x2_trace = sums[3];
  const int v = s->variant;
  const int k = s->iter + 1;
  const int idx = k - 1;
  if (v == V_ALG2_GENSEQ || v == V_ALG1_GENSEQ || v == V_SEQ_ALG2 || v == V_SEQ_ALG1) {
    if (idx < kMaxIter) s->vv[idx] = sqrt(idem);
  } else if (v == V_SUBMATRIX) {
    if (idx < kMaxIter) s->vv[idx] += idem;
  }
  double trx_old = (v == V_ALG2_GENSEQ) ? s->trx_fresh : s->trx;
  if (v == V_BASIC || v == V_BASIC_FIXED) trx_old = s->trx_fresh;
  if (s->apply) {
    if (v != V_SEQ_ALG2 && v != V_SEQ_ALG1 && v != V_SUBMATRIX && idx < kMaxIter)
      s->pp[idx] = s->sigma;
    s->iter = k;
    s->cur ^= 1;
    s->fro2 = fro2;
    s->trx_fresh = trf;
    s->trx = s->trx_next;
  } else {
    s->trx = s->trx_next;
  }
  // idempotency-error history (src/prg_sp2_mod.F90:352-354, :830-832)
  double e_new;
  if (v == V_ALG1 || v == V_ALG1_GENSEQ) e_new = fabs(s->trx2);
  else e_new = fabs(s->trx - trx_old);
  s->e2 = s->e1;
  s->e1 = s->e0;
  s->e0 = e_new;
  if (s->refine_phase && s->apply) s->refine_steps += 1;
  if (s->last) s->done = 1;
  sp2_plan(s);
}

__global__ void sp2_begin_refine_kernel(Sp2State* s) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (s->refine_max <= 0 || s->status != OK) return;
  s->refine_phase = 1;
  s->refine_steps = 0;
  s->done = 0;
  s->e0 = s->e1 = s->e2 = 0.0;
  sp2_plan(s);
}

__global__ void sp2_init_kernel(Sp2State* s, const double* __restrict__ part, int nb) {
  __shared__ double sums[2];
  block_sum(part, nb, 2, sums);
  if (threadIdx.x != 0) return;
  s->trx_fresh = sums[0];
  s->fro2 = sums[1];
  s->trx = sums[0];
  s->iter = 0;
  s->cur = 0;
  s->done = 0;
  s->status = OK;
  s->e0 = s->e1 = s->e2 = 0.0;
  sp2_plan(s);
}

__global__ void sp2_finish_kernel(const double* __restrict__ x0, const double* __restrict__ x1,
                                  const Sp2State* __restrict__ s, int ld, double scale,
                                  double* __restrict__ out, int ld_out, int n) {
  const double* x = s->cur ? x1 : x0;
  for (int col = blockIdx.x; col < n; col += gridDim.x)
    for (int row = threadIdx.x; row < n; row += NT)
      out[row + (size_t)col * ld_out] = scale * x[row + (size_t)col * ld];
}

// Same as sp2_finish_kernel for an X whose lower triangle (and diagonal
// 64 x 64 blocks) alone is current: 64 x 64 output tiles on or below the
// diagonal are copied, the others are the transposes of their mirror tiles
// (through LDS, padded rows).
__global__ void sp2_finish_lower_kernel(const double* __restrict__ x0,
                                        const double* __restrict__ x1,
                                        const Sp2State* __restrict__ s, int ld, double scale,
                                        double* __restrict__ out, int ld_out, int n) {
  __shared__ double t[64][65];
  const double* x = s->cur ? x1 : x0;
  const int br = blockIdx.x * 64, bc = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4 threads
  if (br >= bc) {
    for (int j = ty; j < 64; j += 4) {
      const int row = br + tx, col = bc + j;
      if (row < n && col < n) out[row + (size_t)col * ld_out] = scale * x[row + (size_t)col * ld];
    }
    return;
  }
  // out(br + i, bc + j) = x(bc + j, br + i)
  for (int j = ty; j < 64; j += 4) {
    const int row = bc + tx, col = br + j;
    t[j][tx] = (row < n && col < n) ? x[row + (size_t)col * ld] : 0.0;
  }
  __syncthreads();
  for (int j = ty; j < 64; j += 4) {
    const int row = br + tx, col = bc + j;
    if (row < n && col < n) out[row + (size_t)col * ld_out] = scale * t[tx][j];
  }
}

int grid_cols(int ld) { return ld < 2048 ? ld : 2048; }

// 64 x 64 tiles through LDS (padded rows): coalesced reads and writes.
__global__ void transpose_kernel(const double* __restrict__ a, double* __restrict__ out, int ld) {
  __shared__ double t[64][65];
  const int bx = blockIdx.x * 64, by = blockIdx.y * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4 threads
  for (int j = ty; j < 64; j += 4) t[j][tx] = a[(bx + tx) + (size_t)(by + j) * ld];
  __syncthreads();
  for (int j = ty; j < 64; j += 4) out[(by + tx) + (size_t)(bx + j) * ld] = t[tx][j];
}

__global__ void scale_cols_kernel(const double* __restrict__ a, const double* __restrict__ f,
                                  double* __restrict__ out, int n, int ld) {
  for (int col = blockIdx.x; col < ld; col += gridDim.x) {
    const double s = col < n ? f[col] : 0.0;
    for (int row = threadIdx.x; row < ld; row += NT)
      out[row + (size_t)col * ld] = s * a[row + (size_t)col * ld];
  }
}

// out[r][c] = f[r] a[r][c] (rows beyond n scaled by 0)
__global__ void scale_rows_kernel(const double* __restrict__ a, const double* __restrict__ f,
                                  double* __restrict__ out, int n, int ld) {
  for (int col = blockIdx.x; col < ld; col += gridDim.x)
    for (int row = threadIdx.x; row < ld; row += NT)
      out[row + (size_t)col * ld] = (row < n ? f[row] : 0.0) * a[row + (size_t)col * ld];
}

// out[i] = 1 / a[i][i], i < n
__global__ void inv_diag_kernel(const double* __restrict__ a, double* __restrict__ out, int n,
                                int ld) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) out[i] = 1.0 / a[i + (size_t)i * ld];
}

// per-block partial sums of Tr(X Y) = sum_ij X[i][j] Y[j][i]: one column j of
// X (coalesced) against row j of Y
__global__ void trace_mult_kernel(const double* __restrict__ x, const double* __restrict__ y,
                                  int n, int ld, double* __restrict__ part) {
  double s = 0.0;
  for (int j = blockIdx.x; j < n; j += gridDim.x)
    for (int i = threadIdx.x; i < n; i += NT) s += x[i + (size_t)j * ld] * y[j + (size_t)i * ld];
  __shared__ double red[NT];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace

int scale_rows_kernel_launch(const DMat& A, const double* d_f, DMat out) {
  hipLaunchKernelGGL(scale_rows_kernel, dim3(grid_cols(A.ld)), dim3(NT), 0, stream(), A.p, d_f,
                     out.p, A.n, A.ld);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int inv_diag_kernel_launch(const DMat& A, double* d_out) {
  hipLaunchKernelGGL(inv_diag_kernel, dim3((A.n + NT - 1) / NT), dim3(NT), 0, stream(), A.p,
                     d_out, A.n, A.ld);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int trace_mult_kernel_launch(const DMat& X, const DMat& Y, double* d_part, int* nblocks) {
  const int nb = X.n < 1024 ? X.n : 1024;
  hipLaunchKernelGGL(trace_mult_kernel, dim3(nb), dim3(NT), 0, stream(), X.p, Y.p, X.n, X.ld,
                     d_part);
  PRG_HIP_CHECK(hipGetLastError());
  *nblocks = nb;
  return OK;
}

int transpose_kernel_launch(const DMat& A, DMat out) {
  if (A.ld != out.ld || A.p == out.p) {
    set_error(ERR_ARG, "transpose: layout mismatch or aliasing");
    return ERR_ARG;
  }
  const int nb = A.ld / 64;
  hipLaunchKernelGGL(transpose_kernel, dim3(nb, nb), dim3(NT), 0, stream(), A.p, out.p, A.ld);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int scale_cols_kernel_launch(const DMat& A, const double* d_f, DMat out) {
  hipLaunchKernelGGL(scale_cols_kernel, dim3(grid_cols(A.ld)), dim3(NT), 0, stream(), A.p, d_f,
                     out.p, A.n, A.ld);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

namespace {
}  // namespace

namespace {
constexpr int kGershChunk = 512;
}

size_t gershgorin_scratch_doubles(int n) {
  const size_t nchunks = (n + kGershChunk - 1) / kGershChunk;
  const size_t rblocks = (n + NT - 1) / NT;
  return nchunks * n + 2 * rblocks;
}

int gershgorin(const DMat& h, double* d_bounds, double* scratch) {
  const int n = h.n;
  const int cchunk = kGershChunk;
  const int nchunks = (n + cchunk - 1) / cchunk;
  const int rblocks = (n + NT - 1) / NT;
  int st = OK;
  double* rs = scratch ? scratch
                       : (double*)workspace("gersh_rs", sizeof(double) * (size_t)nchunks * n, &st);
  double* bmm = scratch ? scratch + (size_t)nchunks * n
                        : (double*)workspace("gersh_bmm", sizeof(double) * 2 * rblocks, &st);
  if (st) return st;
  hipLaunchKernelGGL(gersh_rows_kernel, dim3(rblocks, nchunks), dim3(NT), 0, stream(), h.p, n,
                     h.ld, cchunk, rs);
  hipLaunchKernelGGL(gersh_minmax_kernel, dim3(rblocks), dim3(NT), 0, stream(), h.p, n, h.ld,
                     nchunks, rs, bmm);
  hipLaunchKernelGGL(gersh_final_kernel, dim3(1), dim3(NT), 0, stream(), bmm, rblocks, d_bounds);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

static int launch_axpby(const AxpbyArgs& g, int* nblocks) {
  const int nb = grid_cols(g.ld);
  if (nblocks) *nblocks = nb;
  hipLaunchKernelGGL(axpby_kernel, dim3(nb), dim3(NT), 0, stream(), g);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int normalize_gershgorin(const DMat& h, const double* d_bounds, DMat x, double* d_part,
                         int* nblocks) {
  AxpbyArgs g{h.p, nullptr, x.p, h.n, x.ld, 0.0, 0.0, 0.0, 1, d_bounds, d_part};
  if (h.ld != x.ld) {
    set_error(ERR_ARG, "normalize: layout mismatch");
    return ERR_ARG;
  }
  return launch_axpby(g, nblocks);
}

int normalize_bounds(const DMat& h, double lo, double hi, DMat x, double* d_part, int* nblocks) {
  if (!(hi - lo != 0.0) || !std::isfinite(hi - lo)) {
    set_error(ERR_NUMERIC, "normalize: zero-width or non-finite spectral interval");
    return ERR_NUMERIC;
  }
  const double inv = 1.0 / (hi - lo);
  AxpbyArgs g{h.p, nullptr, x.p, h.n, x.ld, -inv, 0.0, hi * inv, 0, nullptr, d_part};
  return launch_axpby(g, nblocks);
}

int affine(const DMat& a, double alpha, double beta_identity, DMat out, double* d_part,
           int* nblocks) {
  AxpbyArgs g{a.p, nullptr, out.p, a.n, out.ld, alpha, 0.0, beta_identity, 0, nullptr, d_part};
  return launch_axpby(g, nblocks);
}

int axpby(const DMat& A, double a, const DMat& B, double b, double c, DMat out, double* d_part,
          int* nblocks) {
  AxpbyArgs g{A.p, B.p, out.p, A.n, out.ld, a, b, c, 0, nullptr, d_part};
  return launch_axpby(g, nblocks);
}

int scale_copy(const DMat& src, double s, DMat dst) {
  hipLaunchKernelGGL(scale_copy_kernel, dim3(grid_cols(src.n)), dim3(NT), 0, stream(), src.p,
                     src.ld, dst.p, dst.ld, src.n, s);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int copy_block(const DMat& src, DMat dst) { return scale_copy(src, 1.0, dst); }

int set_identity(DMat dst, double diag) {
  hipLaunchKernelGGL(identity_kernel, dim3(grid_cols(dst.ld)), dim3(NT), 0, stream(), dst.p,
                     dst.ld, dst.n, diag);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int reduce_partials(const double* d_part, int nblocks, int nq, double* d_out) {
  hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(NT), 0, stream(), d_part, nblocks, nq,
                     d_out);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int sp2_init_state(Sp2State* d_state, const double* d_part, int nblocks) {
  hipLaunchKernelGGL(sp2_init_kernel, dim3(1), dim3(NT), 0, stream(), d_state, d_part, nblocks);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int sp2_absorb_plan(Sp2State* d_state, const double* d_part, int ntiles) {
  hipLaunchKernelGGL(sp2_absorb_plan_kernel, dim3(1), dim3(NT), 0, stream(), d_state, d_part,
                     ntiles);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int sp2_begin_refine(Sp2State* d_state) {
  hipLaunchKernelGGL(sp2_begin_refine_kernel, dim3(1), dim3(64), 0, stream(), d_state);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int sp2_finish(const DMat& x0, const DMat& x1, const Sp2State* d_state, double scale, DMat out,
               bool lower_only) {
  if (lower_only) {
    const int nb = (out.n + 63) / 64;
    hipLaunchKernelGGL(sp2_finish_lower_kernel, dim3(nb, nb), dim3(NT), 0, stream(), x0.p, x1.p,
                       d_state, x0.ld, scale, out.p, out.ld, out.n);
  } else {
    hipLaunchKernelGGL(sp2_finish_kernel, dim3(grid_cols(out.n)), dim3(NT), 0, stream(), x0.p,
                       x1.p, d_state, x0.ld, scale, out.p, out.ld, out.n);
  }
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

}  // namespace prg
