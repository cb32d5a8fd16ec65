// Device helpers shared by the GEMM kernels: XCD-aware tile scheduling and
// deterministic workgroup reductions.
#pragma once

#include <hip/hip_runtime.h>

namespace prg {
namespace tile {

constexpr int kThreads = 256;

// Workgroups are dispatched round-robin over the 8 XCDs; give XCD x the
// contiguous logical range [x*nb/8, (x+1)*nb/8).
__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int per = nb / 8, rem = nb % 8;
  const int xcd = b % 8, idx = b / 8;
  return (xcd < rem) ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

// Upper-triangle tile t -> (I, J), I <= J, column-by-column.
__device__ __forceinline__ void sym_tile(int t, int& I, int& J) {
  int j = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((j + 1) * (j + 2) / 2 <= t) ++j;
  while (j * (j + 1) / 2 > t) --j;
  J = j;
  I = t - j * (j + 1) / 2;
}

// Full tile t -> (I, J) with groups of 8 tile-rows sharing B panels.
__device__ __forceinline__ void full_tile(int t, int nt, int& I, int& J) {
  const int G = 8;
  const int group = t / (G * nt);
  const int first = group * G;
  const int gsize = min(nt - first, G);
  const int in = t % (G * nt);
  I = first + in % gsize;
  J = in / gsize;
}

// Deterministic workgroup sum of NQ per-thread values (BT threads); threads
// 0 .. NQ - 1 store them.
template <int NQ, int BT = kThreads>
__device__ __forceinline__ void block_store_partials(double (&v)[NQ], double* dst) {
  __shared__ double red[BT / 64][NQ];
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
#pragma unroll
    for (int ww = 0; ww < BT / 64; ++ww) s += red[ww][threadIdx.x];
    dst[threadIdx.x] = s;
  }
}


// One entry of the SP2 update z from x (the current X) and y (X^2) as the
// sparse BML formats compute it, each product and sum thresholded (|a| <=
// thr dropped): alg1 forms d = thr(X - X^2) then thr(X -+ d)
// (src/prg_sp2_mod.F90:783-784, :811-823), the other variants thr(X^2) then
// thr(X^2) or thr(2X - thr(X^2)) (:450-453, :328-337).  y is replaced by the
// thresholded X^2 (X - d for alg1) that the idempotency sums read.  apply =
// 0 leaves X unchanged.
__device__ __forceinline__ double thr_drop(double v, double thr) {
  return fabs(v) > thr ? v : 0.0;
}
struct SparseUpdate {
  double z, y;
};
__device__ __forceinline__ SparseUpdate sp2_sparse_update(double x, double y, int apply,
                                                          int sigma, bool alg1, double thr) {
  // alg1: d = thr(X - X^2); X^2 seen by the sums as X - d
  const double d = thr_drop(x - y, thr);
  const double yt = alg1 ? x - d : thr_drop(y, thr);
  const double up = alg1 ? (sigma ? x - d : x + d) : (sigma ? yt : 2.0 * x - yt);
  return SparseUpdate{apply ? thr_drop(up, thr) : x, yt};
}

}  // namespace tile
}  // namespace prg
