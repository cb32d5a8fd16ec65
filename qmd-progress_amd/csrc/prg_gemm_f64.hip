// FP64 MFMA GEMM family for gfx950 (v_mfma_f64_16x16x4_f64).
//
// Tiling: 128 x 128 output tile per 256-thread workgroup (4 waves, 2 x 2),
// each wave owns a 64 x 64 sub-tile = 4 x 4 MFMA blocks of 16 x 16.  The K
// loop stages 16-deep slabs of A and B through LDS with a register prefetch
// of the next slab.  Workgroup ids are remapped so that each of the 8 XCDs
// walks a contiguous range of output tiles (its own L2 keeps the shared
// panels).  For symmetric products (X*X, X*D + D*X, products of commuting
// symmetric matrices) only the upper-triangle tiles are computed; the
// epilogue writes each tile and its transpose, so the result is exactly
// symmetric and the executed flops are half of the algorithmic 2 N^3.
//
// Epilogues are fused: the SP2 branch update (X <- X^2 or 2X - X^2, decided on
// the device before the product from ||X||_F^2 = Tr X^2), the trace and
// Frobenius partial sums for the next branch decision, and the idempotency
// sums ||X - X^2||_F^2 (genseq) or its core-row restriction (submatrix).
#include "prg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include "prg_tile.h"

namespace prg {
namespace {

using tile::xcd_remap;
using tile::sym_tile;
using tile::full_tile;
using tile::block_store_partials;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

constexpr int BM = 128, BN = 128, BK = 16, NT = 256;
constexpr int LROW = BM + 4;  // padded LDS row length in doubles

struct Operands {
  const double* A[2];
  const double* B[2];
  int npair;
};

// Accumulate acc += sum_p A_p[I-panel, :] * B_p[:, J-panel] for one tile.
// A is read as column-major A[i + k*ld], B as column-major B[k + j*ld].
// K runs over [kbeg, kbeg + kdim) of each product.
__device__ __forceinline__ void mainloop(const Operands& op, int ld, int kbeg, int kdim, int I,
                                         int J, double (*As)[LROW], double (*Bs)[LROW],
                                         d4 acc[4][4]) {
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int w = tid >> 6, wm = w & 1, wn = w >> 1;
  const int ar = (tid & 63) * 2, ak = tid >> 6;  // A: rows ar,ar+1; k = ak + 4q
  const int bj = tid >> 1, bk = (tid & 1) * 8;   // B: column bj; k = bk .. bk+7
  const int ksteps = kdim / BK;
  const int total = ksteps * op.npair;
  d2 ra[4], rb[4];

  auto load = [&](int s) {
    const int p = s / ksteps;
    const int k0 = kbeg + (s - p * ksteps) * BK;
    const double* A = p ? op.A[1] : op.A[0];
    const double* B = p ? op.B[1] : op.B[0];
    const double* Ap = A + (size_t)I * BM + ar;
    const double* Bp = B + (size_t)(J * BN + bj) * ld + k0 + bk;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      ra[q] = *(const d2*)(Ap + (size_t)(k0 + ak + 4 * q) * ld);
      rb[q] = *(const d2*)(Bp + 2 * q);
    }
  };

  load(0);
  for (int s = 0; s < total; ++s) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      As[ak + 4 * q][ar] = ra[q].x;
      As[ak + 4 * q][ar + 1] = ra[q].y;
      Bs[bk + 2 * q][bj] = rb[q].x;
      Bs[bk + 2 * q + 1][bj] = rb[q].y;
    }
    __syncthreads();
    if (s + 1 < total) load(s + 1);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kr = kk * 4 + (lane >> 4);
      double a[4], b[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) a[m] = As[kr][wm * 64 + m * 16 + (lane & 15)];
#pragma unroll
      for (int n = 0; n < 4; ++n) b[n] = Bs[kr][wn * 64 + n * 16 + (lane & 15)];
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int n = 0; n < 4; ++n)
          acc[m][n] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[n], a[m], acc[m][n], 0, 0, 0);
    }
  }
}

// Element coordinates of accumulator entry (m, n, r) held by this lane.
// v_mfma_f64_16x16x4 returns D[lane/16 + 4r][lane%16] in register r; the
// main loop feeds the B fragment as the instruction's first operand, so the
// block it returns is (A B)^T and lane holds (A B)[lane%16][lane/16 + 4r]:
// 16 consecutive lanes own 16 consecutive rows of one column (a 128-byte
// segment of the column-major output).
#define PRG_ROW(I, m) ((I) * BM + wm * 64 + (m) * 16 + (lane & 15))
#define PRG_COL(J, n, r) ((J) * BN + wn * 64 + (n) * 16 + (lane >> 4) + 4 * (r))

// ------------------------------------------------------------- general GEMM
struct GemmArgs {
  Operands op;
  int ld, kdim, nt, n, sym;
  double alpha, beta, gamma;
  const double* Cin;
  double* C;
  const double* dotE;  // optional: partial <C, E> (elementwise product sum)
  double* part;        // per tile [Tr C, ||C||^2, <C, E>, 0] or null
};

__global__ __launch_bounds__(NT, 2) void gemm_kernel(GemmArgs g) {
  __shared__ double As[BK][LROW];
  __shared__ double Bs[BK][LROW];
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  int I, J;
  if (g.sym) sym_tile(t, I, J); else full_tile(t, g.nt, I, J);
  d4 acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = d4{0.0, 0.0, 0.0, 0.0};
  mainloop(g.op, g.ld, 0, g.kdim, I, J, As, Bs, acc);

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6, wm = w & 1, wn = w >> 1;
  const int ld = g.ld;
  // Symmetric products: off-diagonal tiles are mirrored; inside a diagonal
  // tile the upper half is mirrored onto the lower one, so that the output is
  // exactly symmetric also for two-product sums (A0 B0 + A1 B1 accumulates its
  // (r, c) and (c, r) entries in different orders).
  double sums[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < 4; ++m) {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int row = PRG_ROW(I, m);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = PRG_COL(J, n, r);
        if (g.sym && row > col) continue;  // lower half of a diagonal tile
        const bool mirror = g.sym && (row != col);
        const double wgt = mirror ? 2.0 : 1.0;
        const size_t idx = row + (size_t)col * ld;
        double v = acc[m][n][r] * g.alpha;
        if (g.beta != 0.0) v += g.beta * g.Cin[idx];
        const bool diag = (row == col);
        v += (diag && row < g.n) ? g.gamma : 0.0;
        sums[0] += diag ? v : 0.0;
        sums[1] += wgt * v * v;
        if (g.dotE) sums[2] += wgt * v * g.dotE[idx];
        g.C[idx] = v;
        if (mirror) g.C[col + (size_t)row * ld] = v;
      }
    }
  }
  if (g.part) block_store_partials<4>(sums, g.part + 4 * (size_t)t);
}

// ------------------------------------------------------------- SP2 step
struct Sp2Args {
  const double* buf[2];  // ping-pong X buffers
  const double* dbuf[2]; // ping-pong D buffers (response step)
  int ld, kdim, nt;
  int split;             // K slices of the split product (1: fused kernel)
  Sp2State* st;
  double* part;  // per tile [||Xn||^2, Tr Xn, idem, Tr Y]
  double* acc;   // split product: per (tile, slice) partial accumulators
};

// The SP2 update fused into the epilogue of tile (I, J) of Y = X^2: z =
// cb x + ca y with (ca, cb) = (1, 0) for X^2, (-1, 2) for 2X - X^2, (0, 1)
// when the product is only measured (X unchanged); the tile's partial sums
// [||Xn||^2, Tr Xn, ||X - X^2||^2 (core rows for submatrix), Tr Y] go to
// part[4 t].
// MN of the tile's 16-row MFMA groups, starting at group m0, and NN of its
// column groups, starting at n0 (MN = NN = 4: the whole tile); the partial
// sums go to part[4 slot].
template <bool THR, int MN, int NN = 4>
__device__ __forceinline__ void sp2_epilogue_impl(const Sp2Args& g, const Sp2State* st, int slot,
                                                  int I, int J, const d4 (&acc)[MN][NN], int m0,
                                                  int apply, int sigma, int cur, int n0 = 0) {
  const int core = st->core;
  const int variant = st->variant;
  const double* X = g.buf[cur];
  double* Xn = const_cast<double*>(g.buf[cur ^ 1]);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6, wm = w & 1, wn = w >> 1;
  const int ld = g.ld;
  const bool mirror = (I != J);
  const bool core_only = (variant == V_SUBMATRIX);
  const double ca = apply ? (sigma ? 1.0 : -1.0) : 0.0;
  const double cb = apply ? (sigma ? 0.0 : 2.0) : 1.0;
  const double wz = mirror ? 2.0 : 1.0;
  const int core_lim = core_only ? core : 0x7fffffff;
  const double thr = st->threshold;
  const bool alg1 = (variant == V_ALG1 || variant == V_ALG1_GENSEQ || variant == V_SEQ_ALG1);
  double sums[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int m = 0; m < MN; ++m) {
#pragma unroll
    for (int n = 0; n < NN; ++n) {
      const int row = PRG_ROW(I, m0 + m);
      const double wrow = (row < core_lim) ? 1.0 : 0.0;
      double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = PRG_COL(J, n0 + n, r);
        const size_t idx = row + (size_t)col * ld;
        double y = acc[m][n][r];
        const double x = X[idx];
        double z;
        if (THR) {
          const tile::SparseUpdate u = tile::sp2_sparse_update(x, y, apply, sigma, alg1, thr);
          z = u.z;
          y = u.y;
        } else {
          z = fma(cb, x, ca * y);
        }
        const double dxy = x - y;
        const double wi = wrow + ((mirror && col < core_lim) ? 1.0 : 0.0);
        const bool diag = (row == col);
        q0 += z * z;
        q2 += wi * dxy * dxy;
        q1 += diag ? z : 0.0;
        q3 += diag ? y : 0.0;
        Xn[idx] = z;
        // transposed copy; on diagonal tiles the same lane rewrites its own
        // element with the same value (no branch, no cross-lane race)
        Xn[mirror ? col + (size_t)row * ld : idx] = z;
      }
      sums[0] += wz * q0;
      sums[1] += q1;
      sums[2] += q2;
      sums[3] += q3;
      // one 16 x 16 block at a time: keeps the X loads of the epilogue from
      // being hoisted together (register pressure)
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  block_store_partials<4>(sums, g.part + 4 * (size_t)slot);
}

__global__ __launch_bounds__(NT, 2) void sp2_step_kernel(Sp2Args g) {
  __shared__ double As[BK][LROW];
  __shared__ double Bs[BK][LROW];
  __shared__ int s_flags[4];
  if (threadIdx.x == 0) {
    s_flags[0] = g.st->planned;
    s_flags[1] = g.st->apply;
    s_flags[2] = g.st->sigma;
    s_flags[3] = g.st->cur;
  }
  __syncthreads();
  if (!s_flags[0]) return;
  const int cur = s_flags[3];
  const double* X = g.buf[cur];
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  int I, J;
  sym_tile(t, I, J);
  d4 acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = d4{0.0, 0.0, 0.0, 0.0};
  Operands op;
  op.A[0] = X; op.B[0] = X; op.A[1] = X; op.B[1] = X; op.npair = 1;
  mainloop(op, g.ld, 0, g.kdim, I, J, As, Bs, acc);
  sp2_epilogue_impl<false, 4>(g, g.st, t, I, J, acc, 0, s_flags[1], s_flags[2], cur);
}

// Split product for small matrices (too few upper-triangle tiles to fill the
// CUs): workgroup b computes K slice b % split of tile b / split and stores
// its accumulators, lane-interleaved (entry q of thread tid at
// [q][tid], d4 each), in acc[(tile * split + slice) * 128 * 128].
__global__ __launch_bounds__(NT, 2) void sp2_split_product_kernel(Sp2Args g) {
  __shared__ double As[BK][LROW];
  __shared__ double Bs[BK][LROW];
  __shared__ int s_flags[2];
  if (threadIdx.x == 0) {
    s_flags[0] = g.st->planned;
    s_flags[1] = g.st->cur;
  }
  __syncthreads();
  if (!s_flags[0]) return;
  const double* X = g.buf[s_flags[1]];
  const int t = blockIdx.x / g.split, slice = blockIdx.x % g.split;
  // slice's K steps: [slice ks / split, (slice + 1) ks / split)
  const int ks = g.kdim / BK;
  const int k0 = (slice * ks) / g.split, k1 = ((slice + 1) * ks) / g.split;
  int I, J;
  sym_tile(t, I, J);
  d4 acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = d4{0.0, 0.0, 0.0, 0.0};
  Operands op;
  op.A[0] = X; op.B[0] = X; op.A[1] = X; op.B[1] = X; op.npair = 1;
  mainloop(op, g.ld, k0 * BK, (k1 - k0) * BK, I, J, As, Bs, acc);
  d4* dst = (d4*)(g.acc + (size_t)blockIdx.x * BM * BN) + threadIdx.x;
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) dst[(4 * m + n) * NT] = acc[m][n];
}

// Sums the split slices of one 16-row group of every wave's quadrant of a
// tile (all four of its 16-column groups for Q = 4, two for Q = 8) in slice
// order (deterministic) and applies the SP2 epilogue to it: workgroup b
// handles tile b / Q, part b % Q, so Q workgroups share a tile's partial
// reads (Q = 8 when four per tile would leave CUs idle).  THR: with the
// sparse formats' thresholding (the fused kernel above serves dense storage
// only, so that its epilogue stays straight-line code).  Partial sums per
// workgroup, part[4 b].
template <bool THR, int Q>
__global__ __launch_bounds__(NT, 2) void sp2_split_epilogue_kernel(Sp2Args g) {
  __shared__ int s_flags[4];
  if (threadIdx.x == 0) {
    s_flags[0] = g.st->planned;
    s_flags[1] = g.st->apply;
    s_flags[2] = g.st->sigma;
    s_flags[3] = g.st->cur;
  }
  __syncthreads();
  if (!s_flags[0]) return;
  static_assert(Q == 4 || Q == 8, "four or eight workgroups per tile");
  constexpr int NN = 16 / Q;  // 16-column groups per workgroup
  const int t = blockIdx.x / Q, q = blockIdx.x % Q;
  const int mg = q / (Q / 4), n0 = (q % (Q / 4)) * NN;
  int I, J;
  sym_tile(t, I, J);
  d4 acc[1][NN];
  const d4* src = (const d4*)(g.acc + (size_t)t * g.split * BM * BN) + threadIdx.x;
#pragma unroll
  for (int n = 0; n < NN; ++n) acc[0][n] = src[(4 * mg + n0 + n) * NT];
  for (int sl = 1; sl < g.split; ++sl) {
    const d4* sp = src + (size_t)sl * (BM * BN / 4);
#pragma unroll
    for (int n = 0; n < NN; ++n) acc[0][n] += sp[(4 * mg + n0 + n) * NT];
  }
  sp2_epilogue_impl<THR, 1, NN>(g, g.st, blockIdx.x, I, J, acc, mg, s_flags[1], s_flags[2],
                                s_flags[3], n0);
}

// Response step: Y = X D + D X (two products accumulated over a 2N-deep K),
// Dn = Y (sigma=1) or 2D - Y (sigma=0).  X itself is advanced by sp2_step.
__global__ __launch_bounds__(NT, 2) void sp2_response_kernel(Sp2Args g) {
  __shared__ double As[BK][LROW];
  __shared__ double Bs[BK][LROW];
  __shared__ int s_flags[4];
  if (threadIdx.x == 0) {
    s_flags[0] = g.st->planned;
    s_flags[1] = g.st->apply;
    s_flags[2] = g.st->sigma;
    s_flags[3] = g.st->cur;
  }
  __syncthreads();
  if (!s_flags[0] || !s_flags[1]) return;
  const int sigma = s_flags[2], cur = s_flags[3];
  const double ca = sigma ? 1.0 : -1.0, cb = sigma ? 0.0 : 2.0;
  const double* X = g.buf[cur];
  const double* D = g.dbuf[cur];
  double* Dn = const_cast<double*>(g.dbuf[cur ^ 1]);
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  int I, J;
  sym_tile(t, I, J);
  d4 acc[4][4];
#pragma unroll
  for (int m = 0; m < 4; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[m][n] = d4{0.0, 0.0, 0.0, 0.0};
  Operands op;
  op.A[0] = X; op.B[0] = D; op.A[1] = D; op.B[1] = X; op.npair = 2;
  mainloop(op, g.ld, 0, g.kdim, I, J, As, Bs, acc);
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6, wm = w & 1, wn = w >> 1;
  const int ld = g.ld;
#pragma unroll
  for (int m = 0; m < 4; ++m) {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int row = PRG_ROW(I, m);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = PRG_COL(J, n, r);
        // X D + D X accumulates its (r, c) and (c, r) entries in different
        // orders: inside a diagonal tile the upper half is mirrored
        if (row > col) continue;
        const size_t idx = row + (size_t)col * ld;
        const double z = fma(cb, D[idx], ca * acc[m][n][r]);
        Dn[idx] = z;
        Dn[col + (size_t)row * ld] = z;
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }
}

}  // namespace

int gemm(int npair, const DMat* A, const DMat* B, double alpha, double beta, const DMat* Cin,
         double gamma, DMat C, int sym, double* d_part, int* ntiles, const DMat* dotE) {
  if (npair < 1 || npair > 2) {
    set_error(ERR_ARG, "gemm: npair must be 1 or 2");
    return ERR_ARG;
  }
  const int ld = C.ld;
  for (int p = 0; p < npair; ++p) {
    if (A[p].ld != ld || B[p].ld != ld || !A[p].p || !B[p].p) {
      set_error(ERR_ARG, "gemm: operand layout mismatch");
      return ERR_ARG;
    }
  }
  if (beta != 0.0 && (!Cin || Cin->ld != ld || !Cin->p)) {
    set_error(ERR_ARG, "gemm: beta != 0 needs an input C of the same layout");
    return ERR_ARG;
  }
  if (ld % kTile != 0 || ld <= 0) {
    set_error(ERR_ARG, "gemm: leading dimension must be a positive multiple of 128");
    return ERR_ARG;
  }
  GemmArgs g;
  for (int p = 0; p < 2; ++p) {
    g.op.A[p] = A[p < npair ? p : 0].p;
    g.op.B[p] = B[p < npair ? p : 0].p;
  }
  g.op.npair = npair;
  g.ld = ld;
  g.kdim = ld;
  g.nt = ld / BM;
  g.n = C.n;
  g.sym = sym;
  g.alpha = alpha;
  g.beta = beta;
  g.gamma = gamma;
  g.Cin = Cin ? Cin->p : nullptr;
  g.C = C.p;
  g.dotE = dotE ? dotE->p : nullptr;
  g.part = d_part;
  const int nb = sym ? g.nt * (g.nt + 1) / 2 : g.nt * g.nt;
  if (ntiles) *ntiles = nb;
  hipLaunchKernelGGL(gemm_kernel, dim3(nb), dim3(NT), 0, stream(), g);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

// K slices of the SP2 square for a padded size ld.  Below 1024 upper-triangle
// tiles the tiles leave CUs idle (36 at ld = 1024, 528 at 4096, for 256 CUs),
// so the K range is cut into s slices (uneven, at least 4 K steps of 16
// each, s <= 16) chosen by a makespan model: ceil(nb s / 256) slices per CU
// of 1/s of a tile's MFMA time each (128 x 128 x ld at ~0.24 TFLOP/s per
// CU), plus the partial tiles' round trip through HBM (2 x 128 KB per slice
// at ~5 TB/s) and the extra launch (~5 us).  PRG_SP2_SPLIT=s forces a count.
int sp2_split_count(int ld) {
  const int nt = ld / BM, nb = nt * (nt + 1) / 2;
  const int ks = ld / BK;
  const int smax = std::min(16, ks / 4);
  static int forced = -1;
  if (forced < 0) {
    const char* e = std::getenv("PRG_SP2_SPLIT");
    forced = e ? std::max(0, std::atoi(e)) : 0;
  }
  if (forced > 0) return std::max(1, std::min(forced, smax));
  if (nb >= 1024) return 1;
  const double tile_s = 2.0 * BM * BN * (double)ld / 0.24e12;
  int best = 1;
  double best_t = 1e30;
  for (int sl = 1; sl <= smax; ++sl) {
    const double rounds = std::ceil((double)nb * sl / 256.0);
    double t = rounds * tile_s / sl;
    if (sl > 1) t += 2.0 * nb * sl * (BM * BN * 8.0) / 5.0e12 + 5e-6;
    if (t < best_t * 0.97) {  // a split must win by 3 %
      best_t = t;
      best = sl;
    }
  }
  return best;
}

int sp2_epilogue_parts(int ld) {
  const int nt = ld / BM, nb = nt * (nt + 1) / 2;
  return (4 * nb < 256) ? 8 : 4;
}

size_t sp2_split_bytes(int ld) {
  const int split = sp2_split_count(ld);
  if (split == 1) return 0;
  const int nt = ld / BM, nb = nt * (nt + 1) / 2;
  return sizeof(double) * (size_t)nb * split * BM * BN;
}

int sp2_step(const DMat& x0, const DMat& x1, Sp2State* d_state, double* d_part, int* ntiles,
             double* split_acc, bool thresholded) {
  Sp2Args g;
  g.buf[0] = x0.p;
  g.buf[1] = x1.p;
  g.dbuf[0] = g.dbuf[1] = nullptr;
  g.ld = x0.ld;
  g.kdim = x0.ld;
  g.nt = x0.ld / BM;
  g.st = d_state;
  g.part = d_part;
  g.split = sp2_split_count(g.ld);
  g.acc = nullptr;
  const int nb = g.nt * (g.nt + 1) / 2;
  if (ntiles) *ntiles = nb;
  if (g.split > 1 || thresholded) {
    int st = OK;
    const size_t bytes = sizeof(double) * (size_t)nb * g.split * BM * BN;
    g.acc = split_acc ? split_acc : (double*)workspace("sp2_split_acc", bytes, &st);
    if (st) return st;
    hipLaunchKernelGGL(sp2_split_product_kernel, dim3(nb * g.split), dim3(NT), 0, stream(), g);
    // eight epilogue workgroups per tile while four per tile would leave
    // half the 256 CUs idle
    const int q = sp2_epilogue_parts(g.ld);
    void (*epi)(Sp2Args) =
        thresholded ? (q == 8 ? sp2_split_epilogue_kernel<true, 8> : sp2_split_epilogue_kernel<true, 4>)
                    : (q == 8 ? sp2_split_epilogue_kernel<false, 8>
                              : sp2_split_epilogue_kernel<false, 4>);
    hipLaunchKernelGGL(epi, dim3(q * nb), dim3(NT), 0, stream(), g);
    if (ntiles) *ntiles = q * nb;  // partial-sum slots: q per tile
  } else {
    hipLaunchKernelGGL(sp2_step_kernel, dim3(nb), dim3(NT), 0, stream(), g);
  }
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int sp2_response_step(const DMat& x0, const DMat& x1, const DMat& d0, const DMat& d1,
                      Sp2State* d_state, double* d_part, int* ntiles) {
  Sp2Args g;
  g.buf[0] = x0.p;
  g.buf[1] = x1.p;
  g.dbuf[0] = d0.p;
  g.dbuf[1] = d1.p;
  g.ld = x0.ld;
  g.kdim = x0.ld;
  g.nt = x0.ld / BM;
  g.st = d_state;
  g.part = d_part;
  g.split = 1;
  g.acc = nullptr;
  const int nb = g.nt * (g.nt + 1) / 2;
  if (ntiles) *ntiles = nb;
  hipLaunchKernelGGL(sp2_response_kernel, dim3(nb), dim3(NT), 0, stream(), g);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

}  // namespace prg
