// Mixed-precision SP2 products for gfx950.
//
// PRG_PREC_F16SPLIT: X (symmetric, |x| <= 1) is carried as an FP64 master and
// two FP16 planes, hi = fp16(x) and lo = fp16((x - hi) * 2^11), so that
// x ~= hi + 2^-11 lo (about 22 significant bits).  One fused kernel computes
//   X^2 ~= hi*hi + 2^-11 (hi*lo + lo*hi)
// with v_mfma_f32_16x16x32_f16 (FP32 accumulation, two accumulator sets),
// applies the SP2 branch update in FP64 against the master copy, writes the
// next master and its hi/lo planes, and emits the same trace / Frobenius
// partial sums as the FP64 kernel, so the FP64 bookkeeping and stopping rules
// are shared.  The dropped lo*lo term is 2^-22 relative.
//
// PRG_PREC_BF16SPLIT is the same kernel on BF16 planes, hi = bf16(x) and
// lo = bf16((x - hi) * 2^8) (about 16 significant bits), with
// v_mfma_f32_16x16x32_bf16.
//
// Because the planes are symmetric, the A operand (rows of the I panel) and
// the B operand (columns of the J panel) are both read as contiguous column
// chunks of a plane.
#include "prg_sp2.h"
#include "prg_tile.h"

namespace prg {
namespace {

using tile::block_store_partials;
using tile::sym_tile;
using tile::xcd_remap;

typedef float f4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));

// The two 16-bit split formats: element type, lo scale (2^(mantissa bits)),
// and the 16x16x32 MFMA with FP32 accumulation.
template <typename T>
struct Split;
template <>
struct Split<_Float16> {
  typedef _Float16 v8 __attribute__((ext_vector_type(8)));
  static constexpr double kScale = 2048.0;
  static __device__ __forceinline__ f4 mfma(v8 a, v8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Split<__bf16> {
  typedef __bf16 v8 __attribute__((ext_vector_type(8)));
  static constexpr double kScale = 256.0;
  static __device__ __forceinline__ f4 mfma(v8 a, v8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

constexpr int BM = 128, BK = 64, NT = 256;
constexpr int HROW = BK + 8;  // padded LDS row (halves): 144 bytes

struct F16Args {
  const double* buf[2];         // FP64 master ping-pong
  const void* hi[2];            // hi planes ping-pong (T)
  const void* lo[2];            // lo planes ping-pong (T)
  int ld;
  Sp2State* st;
  double* part;
};

template <typename T>
__global__ __launch_bounds__(NT, 1) void sp2_f16split_kernel(F16Args g) {
  typedef typename Split<T>::v8 h8;
  constexpr double kScale = Split<T>::kScale;
  __shared__ T As[BM][HROW];
  __shared__ T Bs[BM][HROW];
  __shared__ int s_flags[4];
  if (threadIdx.x == 0) {
    s_flags[0] = g.st->planned;
    s_flags[1] = g.st->apply;
    s_flags[2] = g.st->sigma;
    s_flags[3] = g.st->cur;
  }
  __syncthreads();
  if (!s_flags[0]) return;
  const int apply = s_flags[1], sigma = s_flags[2], cur = s_flags[3];
  const int ld = g.ld;
  const double* X = g.buf[cur];
  double* Xn = const_cast<double*>(g.buf[cur ^ 1]);
  const T* hi = (const T*)g.hi[cur];
  const T* lo = (const T*)g.lo[cur];
  T* hin = (T*)const_cast<void*>(g.hi[cur ^ 1]);
  T* lon = (T*)const_cast<void*>(g.lo[cur ^ 1]);

  const int t = xcd_remap(blockIdx.x, gridDim.x);
  int I, J;
  sym_tile(t, I, J);
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = tid >> 6, wm = w & 1, wn = w >> 1;

  f4 acc_hh[4][4], acc_x[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      acc_hh[m][n] = f4{0.f, 0.f, 0.f, 0.f};
      acc_x[m][n] = f4{0.f, 0.f, 0.f, 0.f};
    }

  // three segments over K: (hi,hi) -> acc_hh ; (hi,lo), (lo,hi) -> acc_x
  const int ksteps = ld / BK;
  const int total = 3 * ksteps;
  const int c = tid >> 1, kh = (tid & 1) * 32;
  i4 ra[4], rb[4];
  auto load = [&](int s) {
    const int seg = s / ksteps;
    const int k0 = (s - seg * ksteps) * BK;
    const T* Ap = (seg == 2) ? lo : hi;
    const T* Bp = (seg == 1) ? lo : hi;
    const T* pa = Ap + (size_t)(I * BM + c) * ld + k0 + kh;
    const T* pb = Bp + (size_t)(J * BM + c) * ld + k0 + kh;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ra[q] = *(const i4*)(pa + 8 * q);
      rb[q] = *(const i4*)(pb + 8 * q);
    }
  };
  load(0);
  for (int s = 0; s < total; ++s) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      *(i4*)(&As[c][kh + 8 * q]) = ra[q];
      *(i4*)(&Bs[c][kh + 8 * q]) = rb[q];
    }
    __syncthreads();
    if (s + 1 < total) load(s + 1);
    const bool cross = s >= ksteps;
#pragma unroll
    for (int kk = 0; kk < BK / 32; ++kk) {
      const int kc = kk * 32 + 8 * (lane >> 4);
      h8 a[4], b[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) a[m] = *(const h8*)(&As[wm * 64 + m * 16 + (lane & 15)][kc]);
#pragma unroll
      for (int n = 0; n < 4; ++n) b[n] = *(const h8*)(&Bs[wn * 64 + n * 16 + (lane & 15)][kc]);
      if (cross) {
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int n = 0; n < 4; ++n)
            acc_x[m][n] = Split<T>::mfma(b[n], a[m], acc_x[m][n]);
      } else {
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int n = 0; n < 4; ++n)
            acc_hh[m][n] = Split<T>::mfma(b[n], a[m], acc_hh[m][n]);
      }
    }
  }

  // Epilogue.  Operands were swapped, so this lane holds (X^2)[row][col]
  // with row = lane%16 and col = 4*(lane/16) + r within each 16 x 16 block.
  // Off-diagonal tiles are mirrored; inside a diagonal tile the upper half is
  // mirrored onto the lower one (the hi.lo and lo.hi terms of (r, c) and
  // (c, r) accumulate in different orders), so X_next is exactly symmetric.
  const double ca = apply ? (sigma ? 1.0 : -1.0) : 0.0;
  const double cb = apply ? (sigma ? 0.0 : 2.0) : 1.0;
  // prg_sp2_submatrix: the idempotency sum runs over the first `core` rows
  // (and, by symmetry, the mirrored columns) only
  const int core_lim = (g.st->variant == V_SUBMATRIX) ? g.st->core : 0x7fffffff;
  double sums[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < 4; ++m) {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int row = I * BM + wm * 64 + m * 16 + (lane & 15);
      double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = J * BM + wn * 64 + n * 16 + 4 * (lane >> 4) + r;
        if (row > col) continue;  // lower half of a diagonal tile
        const bool mirror = (row != col);
        const double we = mirror ? 2.0 : 1.0;
        const size_t idx = row + (size_t)col * ld;
        const size_t tdx = mirror ? col + (size_t)row * ld : idx;
        const double y = (double)acc_hh[m][n][r] + (1.0 / kScale) * (double)acc_x[m][n][r];
        const double x = X[idx];
        const double z = fma(cb, x, ca * y);
        const double dxy = x - y;
        const bool diag = (row == col);
        const double wc = (row < core_lim ? 1.0 : 0.0) + ((mirror && col < core_lim) ? 1.0 : 0.0);
        q0 += we * z * z;
        q2 += wc * dxy * dxy;
        q1 += diag ? z : 0.0;
        q3 += diag ? y : 0.0;
        const T h = (T)z;
        const T l = (T)((z - (double)h) * kScale);
        Xn[idx] = z;
        Xn[tdx] = z;
        hin[idx] = h;
        hin[tdx] = h;
        lon[idx] = l;
        lon[tdx] = l;
      }
      sums[0] += q0;
      sums[1] += q1;
      sums[2] += q2;
      sums[3] += q3;
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  block_store_partials<4>(sums, g.part + 4 * (size_t)t);
}

struct SplitGemmArgs {
  const void* ahi;
  const void* alo;
  const void* bhi;
  const void* blo;
  const double* Cin;  // beta term, or null
  const double* E;    // <C, E> partial, or null
  double* C;
  double* part;       // per tile [Tr C, ||C||_F^2, <C, E>, 0], or null
  double alpha, beta, gamma;
  int ld, n;
};

// C = alpha A B + beta Cin + gamma I (gamma on the n x n diagonal) for
// symmetric, commuting A and B from their hi/lo planes: one workgroup per
// upper-triangle 128 x 128 tile, the tile mirrored; inside a diagonal tile
// the upper half is mirrored onto the lower one, so C is exactly symmetric.
template <typename T>
__global__ __launch_bounds__(NT, 1) void split_gemm_kernel(SplitGemmArgs g) {
  constexpr double kScale = Split<T>::kScale;
  __shared__ T As[BM][HROW];
  __shared__ T Bs[BM][HROW];
  const int ld = g.ld;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  int I, J;
  sym_tile(t, I, J);
  const T* ahi = (const T*)g.ahi;
  const T* alo = (const T*)g.alo;
  const T* bhi = (const T*)g.bhi;
  const T* blo = (const T*)g.blo;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = tid >> 6, wm = w & 1, wn = w >> 1;
  // the main loop of sp2_f16split_kernel on the operand planes, three
  // segments over K: (hi, hi) -> acc_hh; (hi, lo), (lo, hi) -> acc_x.  (An
  // inline copy: behind a shared function the compiler moved the
  // accumulators between AGPRs and VGPRs inside the loop, a third slower.)
  typedef typename Split<T>::v8 h8;
  f4 acc_hh[4][4], acc_x[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      acc_hh[m][n] = f4{0.f, 0.f, 0.f, 0.f};
      acc_x[m][n] = f4{0.f, 0.f, 0.f, 0.f};
    }
  const int ksteps = ld / BK;
  const int total = 3 * ksteps;
  const int c = tid >> 1, kh = (tid & 1) * 32;
  i4 ra[4], rb[4];
  auto load = [&](int s) {
    const int seg = s / ksteps;
    const int k0 = (s - seg * ksteps) * BK;
    const T* Ap = (seg == 2) ? alo : ahi;
    const T* Bp = (seg == 1) ? blo : bhi;
    const T* pa = Ap + (size_t)(I * BM + c) * ld + k0 + kh;
    const T* pb = Bp + (size_t)(J * BM + c) * ld + k0 + kh;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ra[q] = *(const i4*)(pa + 8 * q);
      rb[q] = *(const i4*)(pb + 8 * q);
    }
  };
  load(0);
  for (int s = 0; s < total; ++s) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      *(i4*)(&As[c][kh + 8 * q]) = ra[q];
      *(i4*)(&Bs[c][kh + 8 * q]) = rb[q];
    }
    __syncthreads();
    if (s + 1 < total) load(s + 1);
    const bool cross = s >= ksteps;
#pragma unroll
    for (int kk = 0; kk < BK / 32; ++kk) {
      const int kc = kk * 32 + 8 * (lane >> 4);
      h8 a[4], b[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) a[m] = *(const h8*)(&As[wm * 64 + m * 16 + (lane & 15)][kc]);
#pragma unroll
      for (int n = 0; n < 4; ++n) b[n] = *(const h8*)(&Bs[wn * 64 + n * 16 + (lane & 15)][kc]);
      if (cross) {
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int n = 0; n < 4; ++n)
            acc_x[m][n] = Split<T>::mfma(b[n], a[m], acc_x[m][n]);
      } else {
#pragma unroll
        for (int m = 0; m < 4; ++m)
#pragma unroll
          for (int n = 0; n < 4; ++n)
            acc_hh[m][n] = Split<T>::mfma(b[n], a[m], acc_hh[m][n]);
      }
    }
  }
  double sums[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < 4; ++m) {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int row = I * BM + wm * 64 + m * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = J * BM + wn * 64 + n * 16 + 4 * (lane >> 4) + r;
        if (row > col) continue;  // lower half of a diagonal tile
        const bool mirror = (row != col), diag = !mirror;
        const size_t idx = row + (size_t)col * ld;
        double v = g.alpha * ((double)acc_hh[m][n][r] + (1.0 / kScale) * (double)acc_x[m][n][r]);
        if (g.Cin) v += g.beta * g.Cin[idx];
        v += (diag && row < g.n) ? g.gamma : 0.0;
        const double we = mirror ? 2.0 : 1.0;
        sums[0] += diag ? v : 0.0;
        sums[1] += we * v * v;
        if (g.E) sums[2] += we * v * g.E[idx];
        g.C[idx] = v;
        if (mirror) g.C[col + (size_t)row * ld] = v;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  if (g.part) block_store_partials<4>(sums, g.part + 4 * (size_t)t);
}

// hi/lo planes of an FP64 matrix (whole ld x ld buffer, padding stays zero).
template <typename T>
__global__ void split_f16_kernel(const double* __restrict__ x, T* __restrict__ hi,
                                 T* __restrict__ lo, size_t count) {
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < count; i += (size_t)gridDim.x * NT) {
    const double z = x[i];
    const T h = (T)z;
    hi[i] = h;
    lo[i] = (T)((z - (double)h) * Split<T>::kScale);
  }
}

}  // namespace

int split_f16(const DMat& x, void* hi, void* lo, bool bf16) {
  const size_t count = (size_t)x.ld * x.ld;
  if (bf16)
    hipLaunchKernelGGL(split_f16_kernel<__bf16>, dim3(2048), dim3(NT), 0, stream(), x.p,
                       (__bf16*)hi, (__bf16*)lo, count);
  else
    hipLaunchKernelGGL(split_f16_kernel<_Float16>, dim3(2048), dim3(NT), 0, stream(), x.p,
                       (_Float16*)hi, (_Float16*)lo, count);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int sp2_step_mixed(const DMat& x0, const DMat& x1, Sp2State* d_state, double* d_part,
                   int* ntiles, int prec) {
  if (prec == PRG_PREC_INT8CRT) return crt_step(x0, x1, d_state, d_part, ntiles, crt_moduli());
  if (prec != PRG_PREC_F16SPLIT && prec != PRG_PREC_BF16SPLIT) {
    set_error(ERR_ARG, "unknown precision mode");
    return ERR_ARG;
  }
  const bool bf16 = (prec == PRG_PREC_BF16SPLIT);
  const int ld = x0.ld;
  const size_t plane = 2 * (size_t)ld * ld;  // bytes of one 16-bit plane
  int st = OK;
  char* planes = (char*)workspace(bf16 ? "bf16_planes" : "f16_planes", 4 * plane, &st);
  if (st) return st;
  F16Args g;
  g.buf[0] = x0.p;
  g.buf[1] = x1.p;
  g.hi[0] = planes;
  g.lo[0] = planes + plane;
  g.hi[1] = planes + 2 * plane;
  g.lo[1] = planes + 3 * plane;
  g.ld = ld;
  g.st = d_state;
  g.part = d_part;
  const int nt = ld / BM;
  const int nb = nt * (nt + 1) / 2;
  if (ntiles) *ntiles = nb;
  if (bf16)
    hipLaunchKernelGGL(sp2_f16split_kernel<__bf16>, dim3(nb), dim3(NT), 0, stream(), g);
  else
    hipLaunchKernelGGL(sp2_f16split_kernel<_Float16>, dim3(nb), dim3(NT), 0, stream(), g);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int sp2_mixed_prepare(const DMat& x0, int prec) {
  if (prec == PRG_PREC_INT8CRT) return crt_prepare(x0, crt_moduli());
  if (prec != PRG_PREC_F16SPLIT && prec != PRG_PREC_BF16SPLIT) return OK;
  const bool bf16 = (prec == PRG_PREC_BF16SPLIT);
  const int ld = x0.ld;
  const size_t plane = 2 * (size_t)ld * ld;
  int st = OK;
  char* planes = (char*)workspace(bf16 ? "bf16_planes" : "f16_planes", 4 * plane, &st);
  if (st) return st;
  return split_f16(x0, planes, planes + plane, bf16);
}

// One operand that a caller promises not to change while held (the matrix of
// a Newton-Schulz or CG solve): its planes are formed once and reused by
// every split product that names it until the hold ends.
namespace {
struct SplitHeld {
  const double* key = nullptr;
  int depth = 0;
  bool valid = false;
  int ld = 0;
};
SplitHeld g_split_held;
}  // namespace

void split_hold_operand(const double* p) {
  if (g_split_held.depth == 0) {
    g_split_held.key = p;
    g_split_held.valid = false;
  }
  if (g_split_held.key == p) ++g_split_held.depth;
}

void split_release_operand(const double* p) {
  if (g_split_held.depth > 0 && g_split_held.key == p && --g_split_held.depth == 0) {
    g_split_held.key = nullptr;
    g_split_held.valid = false;
  }
}

int split_gemm(const DMat& A, const DMat& B, DMat C, double alpha, double beta, const DMat* Cin,
               double gamma, double* part, int* nparts, const DMat* dotE) {
  const int ld = C.ld;
  if (A.ld != ld || B.ld != ld || (Cin && Cin->ld != ld) || (dotE && dotE->ld != ld) ||
      ld % BM != 0) {
    set_error(ERR_ARG,
              "BF16 split products: operand leading dimensions differ or are not multiples of 128");
    return ERR_ARG;
  }
  const size_t plane = 2 * (size_t)ld * ld;  // bytes of one BF16 plane
  const bool same = (A.p == B.p);
  const bool held_a = g_split_held.depth > 0 && A.p == g_split_held.key;
  const bool held_b = g_split_held.depth > 0 && B.p == g_split_held.key && !same;
  int st = OK;
  char* pa = (char*)workspace(held_a ? "splitg_planes_held" : "splitg_planes_a", 2 * plane, &st);
  char* pb = same ? pa
                  : (char*)workspace(held_b ? "splitg_planes_held" : "splitg_planes_b",
                                     2 * plane, &st);
  if (st) return st;
  if ((held_a || held_b) && g_split_held.ld != ld) g_split_held.valid = false;
  if (!(held_a && g_split_held.valid) && (st = split_f16(A, pa, pa + plane, true))) return st;
  if (!same && !(held_b && g_split_held.valid) && (st = split_f16(B, pb, pb + plane, true)))
    return st;
  if (held_a || held_b) {
    g_split_held.valid = true;
    g_split_held.ld = ld;
  }
  SplitGemmArgs g;
  g.ahi = pa;
  g.alo = pa + plane;
  g.bhi = pb;
  g.blo = pb + plane;
  g.Cin = (Cin && beta != 0.0) ? Cin->p : nullptr;
  g.E = dotE ? dotE->p : nullptr;
  g.C = C.p;
  g.part = part;
  g.alpha = alpha;
  g.beta = beta;
  g.gamma = gamma;
  g.ld = ld;
  g.n = C.n;
  const int nt = ld / BM;
  const int nb = nt * (nt + 1) / 2;
  if (nparts) *nparts = nb;
  hipLaunchKernelGGL(split_gemm_kernel<__bf16>, dim3(nb), dim3(NT), 0, stream(), g);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

}  // namespace prg
