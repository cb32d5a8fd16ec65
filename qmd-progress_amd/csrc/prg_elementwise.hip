// HBM-bound kernels of the density-matrix path (gfx950) and the device-side
// FP64 bookkeeping of the SP2 loop.
//
// All dense kernels sweep the ld x ld zero-padded column-major buffers with
// one thread per row element of a column (64 consecutive doubles per wave per
// load: fully coalesced) and grid-stride over columns.  Per-block partial sums
// are written to a [nblocks][nq] array and reduced in a fixed order, so every
// trace / Frobenius value is bitwise reproducible run to run.
#include "prg_internal.h"
#include "prg_plan.h"

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

// Row sums of |h_ij| (j != i) over a chunk of columns, one thread per row
// (consecutive threads read consecutive entries of a column).  The chunks are
// short so that the grid covers the CUs at small N; the loop is unrolled so
// that eight loads of a thread are in flight together.
__global__ void gersh_rows_kernel(const double* __restrict__ h, int n, int ld, int cchunk,
                                  double* __restrict__ rs) {
  const int i = blockIdx.x * NT + threadIdx.x;
  const int c0 = blockIdx.y * cchunk;
  const int c1 = min(n, c0 + cchunk);
  if (i >= n) return;
  double s = 0.0;
#pragma unroll 8
  for (int j = c0; j < c1; ++j) {
    const double v = h[i + (size_t)j * ld];
    s += (j != i) ? fabs(v) : 0.0;
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
#pragma unroll 8
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

// Absorb the partials of the product just computed, then plan the next step.
__global__ void sp2_absorb_plan_kernel(Sp2State* s, const double* __restrict__ part, int ntiles) {
  __shared__ int planned;
  if (threadIdx.x == 0) planned = s->planned;
  __syncthreads();
  if (!planned) return;
  sp2_absorb_plan_block<NT>(s, part, ntiles);
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
constexpr int kGershChunk = 64;
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
