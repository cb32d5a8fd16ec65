// Batched SP2 over the core+halo subgraphs of a graph partitioning
// (src/prg_subgraphloop_mod.F90:36-126 with prg_sp2_submatrix_inplace,
// src/prg_sp2_mod.F90:1259-1310).
//
// All parts share the branch sequence pp and the normalisation bounds, so
// they advance in lock-step: one launch per iteration processes every part.
// Parts are packed one after another in HBM, each zero-padded to a multiple
// of 64.  The work list holds the upper-triangle 64 x 64 tiles of every part,
// largest parts first; consecutive workgroups (per XCD) walk consecutive
// tiles, so a part's panels stay in one XCD's L2.  Each tile runs the FP64
// MFMA product and the fused update X <- X + (1 - 2 pp)(X - X^2) with the
// core-row sum of (X - X^2)^2 for vv.
#include "prg_graph.h"
#include "prg_tile.h"

namespace prg {
namespace {

using tile::block_store_partials;
using tile::xcd_remap;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int TB = 64, BK = 16, NT = 256;
constexpr int LROW = TB + 2;

struct BatchArgs {
  const double* x;      // packed current X
  double* xn;           // packed next X
  const PartDesc* parts;
  const int4* tiles;    // (part, I, J, 0)
  int sigma;            // 1: X^2 ; 0: 2X - X^2
  double* part_sums;    // per tile: core-row sum of (X - X^2)^2
};

__global__ __launch_bounds__(NT, 4) void batched_sp2_kernel(BatchArgs g) {
  __shared__ double As[BK][LROW];
  __shared__ double Bs[BK][LROW];
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int4 tl = g.tiles[t];
  const PartDesc pd = g.parts[tl.x];
  const int I = tl.y, J = tl.z;
  const int m = pd.m;
  const double* X = g.x + pd.offset;
  double* Xn = g.xn + pd.offset;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = tid >> 6, wm = w & 1, wn = w >> 1;
  // loaders: A slab = rows I*64..+63, k0..k0+15 ; B slab = k0..k0+15, cols J*64..+63
  const int ar = (tid & 31) * 2, ak = tid >> 5;  // rows ar, ar+1 ; k = ak, ak + 8
  const int bj = tid >> 2, bk = (tid & 3) * 4;   // column bj ; k = bk .. bk+3
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  d2 ra[2], rb[2];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      ra[q] = *(const d2*)(X + (size_t)(I * TB + ar) + (size_t)(k0 + ak + 8 * q) * m);
      rb[q] = *(const d2*)(X + (size_t)(J * TB + bj) * m + k0 + bk + 2 * q);
    }
  };
  load(0);
  for (int k0 = 0; k0 < m; k0 += BK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      As[ak + 8 * q][ar] = ra[q].x;
      As[ak + 8 * q][ar + 1] = ra[q].y;
      Bs[bk + 2 * q][bj] = rb[q].x;
      Bs[bk + 2 * q + 1][bj] = rb[q].y;
    }
    __syncthreads();
    if (k0 + BK < m) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kr][wm * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kr][wn * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
  }
  // epilogue: lane holds (X^2)[row][col], row = lane%16, col = lane/16 + 4r
  const bool mirror = (I != J);
  const double ca = g.sigma ? 1.0 : -1.0, cb = g.sigma ? 0.0 : 2.0;
  const int core = pd.core;
  double s[1] = {0.0};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int row = I * TB + wm * 32 + i * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = J * TB + wn * 32 + j * 16 + (lane >> 4) + 4 * r;
        const size_t idx = row + (size_t)col * m;
        const double y = acc[i][j][r];
        const double x = X[idx];
        const double z = fma(cb, x, ca * y);
        const double d = x - y;
        const double wi = ((row < core) ? 1.0 : 0.0) + ((mirror && col < core) ? 1.0 : 0.0);
        s[0] += wi * d * d;
        Xn[idx] = z;
        Xn[mirror ? col + (size_t)row * m : idx] = z;
      }
    }
  }
  block_store_partials<1>(s, g.part_sums + t);
}

// X_p = (maxeval I - H[idx, idx]) / (maxeval - mineval), zero padded.
__global__ void gather_normalize_kernel(const double* __restrict__ h, int ldh,
                                        const PartDesc* __restrict__ parts,
                                        const int* __restrict__ idx_all, double* __restrict__ x,
                                        double mineval, double maxeval) {
  const PartDesc pd = parts[blockIdx.y];
  const int* idx = idx_all + pd.idx_offset;
  double* X = x + pd.offset;
  const double inv = 1.0 / (maxeval - mineval);
  for (int c = blockIdx.x; c < pd.m; c += gridDim.x) {
    const int gc = c < pd.n ? idx[c] : 0;
    for (int r = threadIdx.x; r < pd.m; r += NT) {
      double v = 0.0;
      if (r < pd.n && c < pd.n) {
        v = -h[idx[r] + (size_t)gc * ldh] * inv;
        if (r == c) v += maxeval * inv;
      }
      X[r + (size_t)c * pd.m] = v;
    }
  }
}

// rho[idx[r], idx[c]] = scale X_p[r][c] for the core rows r < core.
__global__ void scatter_core_kernel(const double* __restrict__ x,
                                    const PartDesc* __restrict__ parts,
                                    const int* __restrict__ idx_all, double* __restrict__ rho,
                                    int ldr, double scale) {
  const PartDesc pd = parts[blockIdx.y];
  const int* idx = idx_all + pd.idx_offset;
  const double* X = x + pd.offset;
  for (int c = blockIdx.x; c < pd.n; c += gridDim.x) {
    const int gc = idx[c];
    for (int r = threadIdx.x; r < pd.core; r += NT)
      rho[idx[r] + (size_t)gc * ldr] = scale * X[r + (size_t)c * pd.m];
  }
}

// X_p = H[idx, idx] with the padding rows/columns decoupled: pad on their
// diagonal, zeros elsewhere.
__global__ void gather_pad_kernel(const double* __restrict__ h, int ldh,
                                  const PartDesc* __restrict__ parts,
                                  const int* __restrict__ idx_all, double* __restrict__ x,
                                  double pad) {
  const PartDesc pd = parts[blockIdx.y];
  const int* idx = idx_all + pd.idx_offset;
  double* X = x + pd.offset;
  for (int c = blockIdx.x; c < pd.m; c += gridDim.x) {
    const int gc = c < pd.n ? idx[c] : 0;
    for (int r = threadIdx.x; r < pd.m; r += NT) {
      double v = 0.0;
      if (r < pd.n && c < pd.n)
        v = h[idx[r] + (size_t)gc * ldh];
      else if (r == c)
        v = pad;
      X[r + (size_t)c * pd.m] = v;
    }
  }
}

// Occupations of one part from its eigenvalues w (m per part, ascending):
// U[:, j] = V[:, j] * 2 f(w_j) and, per part, sum_j 2 df/def(w_j)
// (src/prg_densitymatrix_mod.F90:540-578 with fermi / dfermi at :886-916).
__device__ __forceinline__ double fermi_occ(double e, double ef, double kbt) {
  const double t = (e - ef) / kbt;
  return t > 100.0 ? 0.0 : 1.0 / (1.0 + exp(t));
}
__device__ __forceinline__ double fermi_docc(double e, double ef, double kbt) {
  const double t = (e - ef) / kbt;
  if (t > 100.0) return fabs(e - ef) < 0.001 ? -1000.0 : 0.0;
  const double f = 1.0 / (1.0 + exp(t));
  return f * f * exp(t) / kbt;
}

__global__ void fermi_scale_kernel(const double* __restrict__ v, const double* __restrict__ w,
                                   const PartDesc* __restrict__ parts,
                                   const size_t* __restrict__ w_offset, double* __restrict__ u,
                                   double ef, double kbt, double* __restrict__ dnocc) {
  const PartDesc pd = parts[blockIdx.y];
  const double* W = w + w_offset[blockIdx.y];
  const double* V = v + pd.offset;
  double* U = u + pd.offset;
  for (int c = blockIdx.x; c < pd.m; c += gridDim.x) {
    const double occ = 2.0 * fermi_occ(W[c], ef, kbt);
    for (int r = threadIdx.x; r < pd.m; r += NT) U[r + (size_t)c * pd.m] = occ * V[r + (size_t)c * pd.m];
  }
  if (blockIdx.x == 0) {
    __shared__ double red[NT];
    double s = 0.0;
    for (int c = threadIdx.x; c < pd.m; c += NT) s += 2.0 * fermi_docc(W[c], ef, kbt);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) dnocc[blockIdx.y] = red[0];
  }
}

// vv[it] += sum of the tile partials (fixed order).
__global__ void vv_accumulate_kernel(const double* __restrict__ part, int ntiles,
                                     double* __restrict__ vv, int it) {
  __shared__ double red[NT];
  double s = 0.0;
  for (int b = threadIdx.x; b < ntiles; b += NT) s += part[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) vv[it] += red[0];
}

__global__ void zero_kernel(double* __restrict__ p, size_t count) {
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < count; i += (size_t)gridDim.x * NT)
    p[i] = 0.0;
}

}  // namespace

int batched_gather(const DMat& H, const BatchPlan& plan, double* x, double mineval,
                   double maxeval) {
  hipLaunchKernelGGL(gather_normalize_kernel, dim3(64, plan.nparts), dim3(NT), 0, stream(), H.p,
                     H.ld, plan.d_parts, plan.d_idx, x, mineval, maxeval);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int batched_step(const BatchPlan& plan, const double* x, double* xn, int sigma, double* d_part,
                 double* d_vv, int it) {
  BatchArgs g;
  g.x = x;
  g.xn = xn;
  g.parts = plan.d_parts;
  g.tiles = plan.d_tiles;
  g.sigma = sigma;
  g.part_sums = d_part;
  hipLaunchKernelGGL(batched_sp2_kernel, dim3(plan.ntiles), dim3(NT), 0, stream(), g);
  hipLaunchKernelGGL(vv_accumulate_kernel, dim3(1), dim3(NT), 0, stream(), d_part, plan.ntiles,
                     d_vv, it);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int batched_scatter(const BatchPlan& plan, const double* x, DMat rho, double scale) {
  hipLaunchKernelGGL(zero_kernel, dim3(2048), dim3(NT), 0, stream(), rho.p,
                     (size_t)rho.ld * rho.ld);
  hipLaunchKernelGGL(scatter_core_kernel, dim3(64, plan.nparts), dim3(NT), 0, stream(), x,
                     plan.d_parts, plan.d_idx, rho.p, rho.ld, scale);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int batched_gather_pad(const DMat& H, const BatchPlan& plan, double* x, double pad) {
  hipLaunchKernelGGL(gather_pad_kernel, dim3(64, plan.nparts), dim3(NT), 0, stream(), H.p, H.ld,
                     plan.d_parts, plan.d_idx, x, pad);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int batched_fermi_scale(const BatchPlan& plan, const double* v, const double* w,
                        const size_t* d_w_offset, double* u, double ef, double kbt,
                        double* d_dnocc) {
  hipLaunchKernelGGL(fermi_scale_kernel, dim3(64, plan.nparts), dim3(NT), 0, stream(), v, w,
                     plan.d_parts, d_w_offset, u, ef, kbt, d_dnocc);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

}  // namespace prg
