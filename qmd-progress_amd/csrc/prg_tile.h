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

// Deterministic workgroup sum of NQ per-thread values; thread 0 stores them.
template <int NQ>
__device__ __forceinline__ void block_store_partials(double (&v)[NQ], double* dst) {
  __shared__ double red[kThreads / 64][NQ];
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
    for (int ww = 0; ww < kThreads / 64; ++ww) s += red[ww][threadIdx.x];
    dst[threadIdx.x] = s;
  }
}


}  // namespace tile
}  // namespace prg
