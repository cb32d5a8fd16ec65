// Batched SP2 over the core+halo subgraphs of a graph partitioning
// (src/prg_subgraphloop_mod.F90:36-126 with prg_sp2_submatrix_inplace,
// src/prg_sp2_mod.F90:1259-1310).
//
// All parts share the branch sequence pp and the normalisation bounds, so
// they advance in lock-step: one launch per iteration processes every part,
// and the iteration launches are replayed from a captured hipGraph.  Parts
// are packed one after another in HBM, each zero-padded to a multiple of 64.
// The work list holds the upper-triangle tiles of every part, largest parts
// first; consecutive workgroups (per XCD) walk consecutive tiles, so a part's
// panels stay in one XCD's L2.  A part is cut into panels of 64 or of 128
// (the last one 64 wide when the size is an odd multiple of 64), giving
// 64 x 64, 128 x 64 and 128 x 128 tiles; the panel width is picked per batch
// by timing both (prg_graph.cpp).  Each tile runs the FP64 MFMA product and
// the fused update X <- X + (1 - 2 pp)(X - X^2) with the core-row sum of
// (X - X^2)^2 for vv.
#include "prg_graph.h"
#include "prg_lds_dma.h"
#include "prg_tile.h"

namespace prg {
namespace {

using tile::block_store_partials;
using tile::xcd_remap;

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int NT = 256;

struct BatchArgs {
  const double* x;      // packed current X
  double* xn;           // packed next X
  const PartDesc* parts;
  const int4* tiles;    // (part, first row, first column, shape: 0 64x64, 1 128x64, 3 128x128)
  int sigma;            // 1: X^2 ; 0: 2X - X^2
  double* part_sums;    // per tile: core-row sum of (X - X^2)^2
  const int* queue_off; // queue q holds tiles [queue_off[q], queue_off[q + 1]), q < 8
  int* tickets;         // 8 ticket counters, zero at launch
  // split-precision kernels: 16-bit hi / lo planes of the current and the
  // next X, packed like x (same element offsets)
  const void* hi;
  const void* lo;
  void* hin;
  void* lon;
};

// One TM x TN tile of X^2 (rows r0.., columns c0..) of one part and its fused
// SP2 update.  Four waves (2 x 2), each (TM/2) x (TN/2) = (TM/32) x (TN/32)
// MFMA blocks of 16 x 16; 16-deep K slabs of both panels staged through LDS
// with a register prefetch PF slabs ahead.  In a diagonal tile the wave of
// the strictly lower quadrant idles and the upper off-diagonal wave writes
// its block and the transpose.
template <int TM, int TN, int PF, int BK>
__device__ __forceinline__ void sp2_tile(const BatchArgs& g, int t, int4 tl, double* smem) {
  constexpr int LA = TM + (TM == 128 ? 4 : 2), LB = TN + (TN == 128 ? 4 : 2);
  constexpr int MI = TM / 32, NJ = TN / 32;
  constexpr int AR = TM / 2, AKS = NT / AR, QA = BK / AKS;    // A slab loaders
  constexpr int TPC = NT / TN, KPT = BK / TPC, QB = KPT / 2;  // B slab loaders
  static_assert(QA >= 1 && QB >= 1 && AKS * QA == BK && TPC * KPT == BK, "tile loaders");
  double (*As)[LA] = reinterpret_cast<double (*)[LA]>(smem);
  double (*Bs)[LB] = reinterpret_cast<double (*)[LB]>(smem + BK * LA);
  const PartDesc pd = g.parts[tl.x];
  const int r0 = tl.y, c0 = tl.z;
  const int m = pd.m;
  const double* X = g.x + pd.offset;
  double* Xn = g.xn + pd.offset;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = tid >> 6, wm = w & 1, wn = w >> 1;
  const bool diag = (r0 == c0);
  const bool idle = diag && wm == 1 && wn == 0;
  const bool mirror = !diag || (wm == 0 && wn == 1);
  const int ar = (tid % AR) * 2, ak = tid / AR;     // rows ar, ar+1 ; k = ak + AKS q
  const int bj = tid / TPC, bk = (tid % TPC) * KPT;  // column bj ; k = bk .. bk+KPT-1
  d4 acc[MI][NJ];
#pragma unroll
  for (int a = 0; a < MI; ++a)
#pragma unroll
    for (int b = 0; b < NJ; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
  // PF slabs in flight in registers: slab k is loaded PF slabs before it is
  // staged in LDS (m is a multiple of 64, so of PF * BK)
  d2 ra[PF][QA], rb[PF][QB];
  auto load = [&](int slot, int k0) {
#pragma unroll
    for (int q = 0; q < QA; ++q)
      ra[slot][q] = *(const d2*)(X + (size_t)(r0 + ar) + (size_t)(k0 + ak + AKS * q) * m);
#pragma unroll
    for (int q = 0; q < QB; ++q)
      rb[slot][q] = *(const d2*)(X + (size_t)(c0 + bj) * m + k0 + bk + 2 * q);
  };
#pragma unroll
  for (int p = 0; p < PF; ++p) load(p, p * BK);
  for (int k0 = 0; k0 < m; k0 += PF * BK) {
#pragma unroll
    for (int p = 0; p < PF; ++p) {
      const int kc = k0 + p * BK;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < QA; ++q) {
        As[ak + AKS * q][ar] = ra[p][q].x;
        As[ak + AKS * q][ar + 1] = ra[p][q].y;
      }
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        Bs[bk + 2 * q][bj] = rb[p][q].x;
        Bs[bk + 2 * q + 1][bj] = rb[p][q].y;
      }
      __syncthreads();
      if (kc + PF * BK < m) load(p, kc + PF * BK);
      if (!idle) {
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
          const int kr = kk * 4 + (lane >> 4);
          double a[MI], b[NJ];
#pragma unroll
          for (int i = 0; i < MI; ++i) a[i] = As[kr][wm * (TM / 2) + i * 16 + (lane & 15)];
#pragma unroll
          for (int j = 0; j < NJ; ++j) b[j] = Bs[kr][wn * (TN / 2) + j * 16 + (lane & 15)];
#pragma unroll
          for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
        }
      }
    }
  }
  // epilogue: lane holds (X^2)[row][col], row = lane%16, col = lane/16 + 4r
  const double ca = g.sigma ? 1.0 : -1.0, cb = g.sigma ? 0.0 : 2.0;
  const int core = pd.core;
  double s[1] = {0.0};
  if (!idle) {
#pragma unroll
    for (int i = 0; i < MI; ++i) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int row = r0 + wm * (TM / 2) + i * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = c0 + wn * (TN / 2) + j * 16 + (lane >> 4) + 4 * r;
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
  }
  block_store_partials<1>(s, g.part_sums + t);
}

// Persistent work-queue launch: one workgroup per resident slot.  The tile
// list is split into 8 queues, one per XCD (whole parts, balanced by work);
// workgroup b (dispatched to XCD b mod 8) takes tickets from its own queue,
// then from the others in turn, and exits when every queue is drained.
// Tickets only grow, so every workgroup reaches its exit; the tiles are
// independent and each stores its own partial, so results do not depend on
// which workgroup ran which tile.
// WIDE: the tile list holds 128 x 128, 128 x 64 and 64 x 64 tiles (two
// workgroups per CU); otherwise 64 x 64 tiles only.  PF: slabs of prefetch,
// BK: slab depth.
template <bool WIDE, int PF, int BK>
__global__ __launch_bounds__(NT, WIDE ? 2 : 4) void batched_sp2_kernel(BatchArgs g) {
  constexpr int LDS = WIDE ? 2 * BK * (128 + 4) : 2 * BK * (64 + 2);
  __shared__ double smem[LDS];
  __shared__ int s_ticket;
  const int home = blockIdx.x & 7;
  for (int q = 0; q < 8; ++q) {
    const int qi = (home + q) & 7;
    const int first = g.queue_off[qi], count = g.queue_off[qi + 1] - first;
    for (;;) {
      __syncthreads();  // every thread has read the previous ticket
      if (threadIdx.x == 0) s_ticket = atomicAdd(&g.tickets[qi], 1);
      __syncthreads();
      const int k = s_ticket;
      if (k >= count) break;
      const int t = first + k;
      const int4 tl = g.tiles[t];
      if constexpr (WIDE) {
        if (tl.w == 3)
          sp2_tile<128, 128, PF, BK>(g, t, tl, smem);
        else if (tl.w == 1)
          sp2_tile<128, 64, PF, BK>(g, t, tl, smem);
        else
          sp2_tile<64, 64, PF, BK>(g, t, tl, smem);
      } else {
        sp2_tile<64, 64, PF, BK>(g, t, tl, smem);
      }
    }
  }
}

// ------------------------------------------------------------------------
// Split-precision variant (PRG_PREC_F16SPLIT / PRG_PREC_BF16SPLIT): X is
// carried as the FP64 master plus two 16-bit planes, hi = t(x) and
// lo = t((x - hi) s) with s = 2^11 (FP16) or 2^8 (BF16), so that
//   X^2 ~= hi hi + (hi lo + lo hi) / s
// runs as three 16x16x32 MFMA products with FP32 accumulation (two
// accumulator sets); the update, the next planes and the core-row sums are
// formed in FP64 from the master, as the FP64 tile does.  64 x 64 tiles of
// 64-wide panels; 64-deep K slabs of the four operand planes (A hi, A lo,
// B hi, B lo; wave w stages plane w) through LDS with a register prefetch of
// the next slab.  X symmetric: row r of a panel is column r in HBM,
// contiguous in k.
template <typename T>
struct SplitT;
template <>
struct SplitT<_Float16> {
  typedef _Float16 v8 __attribute__((ext_vector_type(8)));
  static constexpr double kScale = 2048.0;
  static __device__ __forceinline__ f4 mfma(v8 a, v8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};
template <>
struct SplitT<__bf16> {
  typedef __bf16 v8 __attribute__((ext_vector_type(8)));
  static constexpr double kScale = 256.0;
  static __device__ __forceinline__ f4 mfma(v8 a, v8 b, f4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};

constexpr int SBK = 64;          // K slab (halves)
constexpr int SROW = SBK + 8;    // padded LDS row: 144 bytes
constexpr int kSplitLds = 4 * 64 * SROW * 2;  // bytes: four 64-row planes

template <typename T>
__device__ __forceinline__ void split_tile(const BatchArgs& g, int t, int4 tl, T* smem) {
  typedef typename SplitT<T>::v8 h8;
  constexpr double kScale = SplitT<T>::kScale;
  T (*Sp)[SROW] = reinterpret_cast<T (*)[SROW]>(smem);  // plane w: rows 64 w .. 64 w + 63
  const PartDesc pd = g.parts[tl.x];
  const int r0 = tl.y, c0 = tl.z;
  const int m = pd.m;
  const double* X = g.x + pd.offset;
  double* Xn = g.xn + pd.offset;
  const T* hi = (const T*)g.hi + pd.offset;
  const T* lo = (const T*)g.lo + pd.offset;
  T* hin = (T*)g.hin + pd.offset;
  T* lon = (T*)g.lon + pd.offset;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = tid >> 6, wm = w & 1, wn = w >> 1;
  const bool diag = (r0 == c0);
  const bool idle = diag && wm == 1 && wn == 0;
  // loader: wave w stages plane w (A hi, A lo, B hi, B lo); lane reads the
  // 16-byte chunk (lane & 7) of rows (lane >> 3) + 8 q, q < 8
  const T* src = (w & 1) ? lo : hi;
  const int lrow0 = (w >> 1) ? c0 : r0;
  const int lc = 8 * (lane & 7), lr = lane >> 3;
  i4 pre[8];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q)
      pre[q] = *(const i4*)(src + (size_t)(lrow0 + lr + 8 * q) * m + k0 + lc);
  };
  f4 acc_hh[2][2], acc_x[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      acc_hh[i][j] = f4{0.f, 0.f, 0.f, 0.f};
      acc_x[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    }
  load(0);
  for (int k0 = 0; k0 < m; k0 += SBK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 8; ++q) *(i4*)(&Sp[64 * w + lr + 8 * q][lc]) = pre[q];
    __syncthreads();
    if (k0 + SBK < m) load(k0 + SBK);
    if (!idle) {
#pragma unroll 1
      for (int kk = 0; kk < SBK / 32; ++kk) {
        const int kc = kk * 32 + 8 * (lane >> 4);
        h8 bh[2], bl[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int rb = wn * 32 + j * 16 + (lane & 15);
          bh[j] = *(const h8*)(&Sp[128 + rb][kc]);
          bl[j] = *(const h8*)(&Sp[192 + rb][kc]);
        }
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          const int ra = wm * 32 + i * 16 + (lane & 15);
          const h8 ah = *(const h8*)(&Sp[ra][kc]);
          const h8 al = *(const h8*)(&Sp[64 + ra][kc]);
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc_hh[i][j] = SplitT<T>::mfma(bh[j], ah, acc_hh[i][j]);
            acc_x[i][j] = SplitT<T>::mfma(bl[j], ah, acc_x[i][j]);
            acc_x[i][j] = SplitT<T>::mfma(bh[j], al, acc_x[i][j]);
          }
        }
      }
    }
  }
  // epilogue: operands were swapped, so this lane holds (X^2)[row][col] with
  // row = lane % 16, col = 4 (lane / 16) + r inside each 16 x 16 block.  In
  // the diagonal blocks of a diagonal tile only row <= col is formed and
  // mirrored (the hi lo and lo hi sums of (r, c) and (c, r) accumulate in
  // different orders), so X_next is exactly symmetric.
  const double ca = g.sigma ? 1.0 : -1.0, cb = g.sigma ? 0.0 : 2.0;
  const int core = pd.core;
  double s[1] = {0.0};
  if (!idle) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int row = r0 + wm * 32 + i * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = c0 + wn * 32 + j * 16 + 4 * (lane >> 4) + r;
          if (row > col) continue;
          const bool mir = (row != col);
          const size_t idx = row + (size_t)col * m;
          const size_t tdx = mir ? col + (size_t)row * m : idx;
          const double y = (double)acc_hh[i][j][r] + (1.0 / kScale) * (double)acc_x[i][j][r];
          const double x = X[idx];
          const double z = fma(cb, x, ca * y);
          const double d = x - y;
          const double wi = ((row < core) ? 1.0 : 0.0) + ((mir && col < core) ? 1.0 : 0.0);
          s[0] += wi * d * d;
          const T h = (T)z;
          const T l = (T)((z - (double)h) * kScale);
          Xn[idx] = z;
          Xn[tdx] = z;
          hin[idx] = h;
          hin[tdx] = h;
          lon[idx] = l;
          lon[tdx] = l;
        }
        // one 16 x 16 block at a time (register pressure)
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  block_store_partials<1>(s, g.part_sums + t);
}

// Split-precision tile of 128-wide panels: TM x TN (128 x 128, 128 x 64 or
// 64 x 64) with four waves of (TM/2) x (TN/2), so that each pair of fragment
// reads feeds up to NJ x 3 MFMAs (twice the reuse of the 64 x 64 tile's
// 32 x 32 waves).  32-deep K slabs of the four operand planes through LDS
// (A hi, A lo: TM rows; B hi, B lo: TN rows), loaded by all threads, eight
// halves per 16-byte chunk, with a register prefetch of the next slab.  An
// LDS row is 64 bytes, four chunks; chunk c of row r sits in slot
// c ^ ((r >> 1) & 3), which puts the 16-byte reads of 8 consecutive rows
// (and the stores of 2 rows) on distinct banks.  Products, update and sums
// as split_tile.
constexpr int WBK = 32;          // K slab of the wide split tiles (halves)
constexpr int kSplitWideLds = 512 * WBK * 2;  // bytes: 2 x 128 + 2 x 128 rows

__device__ __forceinline__ int wide_slot(int row, int chunk) {
  return chunk ^ ((row >> 1) & 3);
}

template <typename T, int TM, int TN>
__device__ __forceinline__ void split_tile_wide(const BatchArgs& g, int t, int4 tl, T* smem) {
  typedef typename SplitT<T>::v8 h8;
  constexpr double kScale = SplitT<T>::kScale;
  constexpr int MI = TM / 32, NJ = TN / 32;      // 16 x 16 blocks per wave
  constexpr int ROWS = 2 * TM + 2 * TN;          // LDS rows: A hi, A lo, B hi, B lo
  constexpr int NQ = ROWS * (WBK / 8) / NT;      // 16-byte chunks per thread
  static_assert(NQ * NT == ROWS * (WBK / 8) && (NT / (WBK / 8)) == 64, "wide split loader");
  T (*Sp)[WBK] = reinterpret_cast<T (*)[WBK]>(smem);
  const PartDesc pd = g.parts[tl.x];
  const int r0 = tl.y, c0 = tl.z;
  const int m = pd.m;
  const double* X = g.x + pd.offset;
  double* Xn = g.xn + pd.offset;
  const T* hi = (const T*)g.hi + pd.offset;
  const T* lo = (const T*)g.lo + pd.offset;
  T* hin = (T*)g.hin + pd.offset;
  T* lon = (T*)g.lon + pd.offset;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = tid >> 6, wm = w & 1, wn = w >> 1;
  const bool diag = (r0 == c0);
  const bool idle = diag && wm == 1 && wn == 0;
  // loader: chunk q of thread tid is chunk tid % 4 (halves 8 (tid % 4) ..)
  // of LDS row 64 q + tid / 4; each group of 64 rows lies in one plane, and
  // 64 q leaves the slot swizzle unchanged
  const int lr = tid >> 2, lc = 8 * (tid & 3);
  const int ls = 8 * wide_slot(lr, tid & 3);
  i4 pre[NQ];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int row = 64 * q + lr;  // compile-time plane per q
      const T* src;
      int prow;
      if (64 * q < TM) { src = hi; prow = r0 + row; }
      else if (64 * q < 2 * TM) { src = lo; prow = r0 + row - TM; }
      else if (64 * q < 2 * TM + TN) { src = hi; prow = c0 + row - 2 * TM; }
      else { src = lo; prow = c0 + row - 2 * TM - TN; }
      pre[q] = *(const i4*)(src + (size_t)prow * m + k0 + lc);
    }
  };
  f4 acc_hh[MI][NJ], acc_x[MI][NJ];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      acc_hh[i][j] = f4{0.f, 0.f, 0.f, 0.f};
      acc_x[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    }
  load(0);
  for (int k0 = 0; k0 < m; k0 += WBK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) *(i4*)(&Sp[64 * q + lr][ls]) = pre[q];
    __syncthreads();
    if (k0 + WBK < m) load(k0 + WBK);
    if (!idle) {
      // this lane's K chunk lane / 16 of rows whose offsets inside their
      // plane are multiples of 16 plus lane % 16: one slot for all of them
      const int ks = 8 * wide_slot(lane & 15, lane >> 4);
      h8 bh[NJ], bl[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int rb = wn * (TN / 2) + j * 16 + (lane & 15);
        bh[j] = *(const h8*)(&Sp[2 * TM + rb][ks]);
        bl[j] = *(const h8*)(&Sp[2 * TM + TN + rb][ks]);
      }
#pragma unroll
      for (int i = 0; i < MI; ++i) {
        const int ra = wm * (TM / 2) + i * 16 + (lane & 15);
        const h8 ah = *(const h8*)(&Sp[ra][ks]);
        const h8 al = *(const h8*)(&Sp[TM + ra][ks]);
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          acc_hh[i][j] = SplitT<T>::mfma(bh[j], ah, acc_hh[i][j]);
          acc_x[i][j] = SplitT<T>::mfma(bl[j], ah, acc_x[i][j]);
          acc_x[i][j] = SplitT<T>::mfma(bh[j], al, acc_x[i][j]);
        }
      }
    }
  }
  // epilogue as split_tile: this lane holds (X^2)[row][col] with row = lane
  // % 16, col = 4 (lane / 16) + r inside each 16 x 16 block; in the diagonal
  // quadrants of a diagonal tile only row <= col is formed and mirrored
  const double ca = g.sigma ? 1.0 : -1.0, cb = g.sigma ? 0.0 : 2.0;
  const int core = pd.core;
  double s[1] = {0.0};
  if (!idle) {
#pragma unroll
    for (int i = 0; i < MI; ++i) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int row = r0 + wm * (TM / 2) + i * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = c0 + wn * (TN / 2) + j * 16 + 4 * (lane >> 4) + r;
          if (row > col) continue;
          const bool mir = (row != col);
          const size_t idx = row + (size_t)col * m;
          const size_t tdx = mir ? col + (size_t)row * m : idx;
          const double y = (double)acc_hh[i][j][r] + (1.0 / kScale) * (double)acc_x[i][j][r];
          const double x = X[idx];
          const double z = fma(cb, x, ca * y);
          const double d = x - y;
          const double wi = ((row < core) ? 1.0 : 0.0) + ((mir && col < core) ? 1.0 : 0.0);
          s[0] += wi * d * d;
          const T h = (T)z;
          const T l = (T)((z - (double)h) * kScale);
          Xn[idx] = z;
          Xn[tdx] = z;
          hin[idx] = h;
          hin[tdx] = h;
          lon[idx] = l;
          lon[tdx] = l;
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  block_store_partials<1>(s, g.part_sums + t);
}

// Persistent work-queue launch of the wide split tiles (shapes as in
// batched_sp2_kernel<true, ...>: 3 = 128 x 128, 1 = 128 x 64, 0 = 64 x 64).
template <typename T>
__global__ __launch_bounds__(NT, 2) void batched_split_wide_kernel(BatchArgs g) {
  __shared__ __attribute__((aligned(16))) T smem[kSplitWideLds / sizeof(T)];
  __shared__ int s_ticket;
  const int home = blockIdx.x & 7;
  for (int q = 0; q < 8; ++q) {
    const int qi = (home + q) & 7;
    const int first = g.queue_off[qi], count = g.queue_off[qi + 1] - first;
    for (;;) {
      __syncthreads();  // every thread has read the previous ticket
      if (threadIdx.x == 0) s_ticket = atomicAdd(&g.tickets[qi], 1);
      __syncthreads();
      const int k = s_ticket;
      if (k >= count) break;
      const int t = first + k;
      const int4 tl = g.tiles[t];
      if (tl.w == 3)
        split_tile_wide<T, 128, 128>(g, t, tl, smem);
      else if (tl.w == 1)
        split_tile_wide<T, 128, 64>(g, t, tl, smem);
      else
        split_tile_wide<T, 64, 64>(g, t, tl, smem);
    }
  }
}

// Ring form of the wide split tile: 512 threads, 8 waves in a 2 x 4 grid of
// (TM/2) x (TN/4), and the 32-deep operand slabs copied global -> LDS
// directly (dma16) into a 4-stage ring, 3 slabs ahead of the MFMAs, with no
// register staging.  A slab is ROWS = 2 TM + 2 TN rows of 64 bytes (A hi,
// A lo, B hi, B lo), ROWS / 16 chunks of 1 KB; chunk c = w + 8 q goes to
// bytes 1024 c of its ring slot, lane l writing row 16 c + l / 4, slot l % 4,
// and loading global K-chunk (l % 4) ^ ((l / 8) & 3): the slot swizzle of
// split_tile_wide, so the 16-byte fragment reads of 8 consecutive rows fall
// on distinct banks.  Each step waits for its slab in every wave (one
// barrier), refills the slot just read with the slab 4 ahead, reads the next
// slab's fragments and runs this slab's MFMAs.  Products, update and sums as
// split_tile.
constexpr int RNT = 512;                // threads of the ring kernel
constexpr int RSTAGE = 4;               // ring slots
constexpr int kRingSlot = 512 * 64;     // bytes: up to 512 rows of 64 bytes
// LDS of the ring kernel: the ring, reused after a tile's main loop to stage
// X_next of an off-diagonal tile for its mirrored store, 128 columns of 129
// doubles (one more than the ring holds)
constexpr int kRingLds = 128 * 129 * 8;
static_assert(kRingLds >= RSTAGE * kRingSlot, "ring kernel LDS");

template <typename T, int TM, int TN>
__device__ __forceinline__ void split_tile_ring(const BatchArgs& g, int t, int4 tl,
                                                const int8_t (*ring)[kRingSlot], double* stage) {
  typedef typename SplitT<T>::v8 h8;
  constexpr double kScale = SplitT<T>::kScale;
  constexpr int ROWS = 2 * TM + 2 * TN;
  constexpr int NQ = ROWS / 16 / 8;       // DMA chunks per wave per slab
  constexpr int WMR = TM / 2, WNC = TN / 4;
  constexpr int MI = WMR / 16, NJ = WNC / 16;
  static_assert(NQ * 16 * 8 == ROWS && MI >= 1 && NJ >= 1, "ring tile shape");
  const PartDesc pd = g.parts[tl.x];
  const int r0 = tl.y, c0 = tl.z;
  const int m = pd.m;
  const double* X = g.x + pd.offset;
  double* Xn = g.xn + pd.offset;
  const T* hi = (const T*)g.hi + pd.offset;
  const T* lo = (const T*)g.lo + pd.offset;
  T* hin = (T*)g.hin + pd.offset;
  T* lon = (T*)g.lon + pd.offset;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = tid >> 6, wm = w & 1, wn = w >> 1;
  const bool diag = (r0 == c0);
  // a wave whose rows all lie below its columns in a diagonal tile
  const bool idle = diag && (wm * WMR >= (wn + 1) * WNC);
  // the previous tile's epilogue stores are done before this tile's copies
  // are counted (the waits below count this wave's vector-memory operations)
  lds_dma::wait_vm_barrier<0>();
  // DMA: the lane's byte offset inside a chunk's 16 rows of one plane
  const int lrow = lane >> 2;
  const uint32_t voff =
      (uint32_t)(lrow * m * (int)sizeof(T) + 16 * ((lane & 3) ^ ((lrow >> 1) & 3)));
  const uint32_t ring0 = lds_dma::lds_addr(&ring[0][0]);
  auto issue = [&](int s) {
    const uint32_t slot = ring0 + kRingSlot * (s % RSTAGE);
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int row = 16 * (w + 8 * q);  // first LDS row of the chunk (wave-uniform)
      const T* src;
      int prow;
      if (row < TM) { src = hi; prow = r0 + row; }
      else if (row < 2 * TM) { src = lo; prow = r0 + row - TM; }
      else if (row < 2 * TM + TN) { src = hi; prow = c0 + row - 2 * TM; }
      else { src = lo; prow = c0 + row - 2 * TM - TN; }
      const uint64_t base =
          lds_dma::uniform64((uint64_t)(uintptr_t)(src + (size_t)prow * m + (size_t)s * 32));
      lds_dma::dma16(base, voff, __builtin_amdgcn_readfirstlane(slot + 64 * row));
    }
  };
  // fragment reads: row (lane % 16) of a 16-row block, K-chunk lane / 16,
  // whose slot is swizzled by the row's bits 1-2 (block starts are multiples
  // of 16 rows)
  const int ks = 16 * ((lane >> 4) ^ (((lane & 15) >> 1) & 3));
  const int ra0 = wm * WMR + (lane & 15), rb0 = 2 * TM + wn * WNC + (lane & 15);
  struct Frags {
    h8 ah[MI], al[MI], bh[NJ], bl[NJ];
  };
  auto read_frags = [&](int s, Frags& f) {
    const int8_t* st = ring[s % RSTAGE];
#pragma unroll
    for (int i = 0; i < MI; ++i) {
      f.ah[i] = *(const h8*)(st + 64 * (ra0 + 16 * i) + ks);
      f.al[i] = *(const h8*)(st + 64 * (ra0 + TM + 16 * i) + ks);
    }
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      f.bh[j] = *(const h8*)(st + 64 * (rb0 + 16 * j) + ks);
      f.bl[j] = *(const h8*)(st + 64 * (rb0 + TN + 16 * j) + ks);
    }
  };
  f4 acc_hh[MI][NJ], acc_x[MI][NJ];
#pragma unroll
  for (int i = 0; i < MI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      acc_hh[i][j] = f4{0.f, 0.f, 0.f, 0.f};
      acc_x[i][j] = f4{0.f, 0.f, 0.f, 0.f};
    }
  const int ksteps = m / 32;  // m is a multiple of 64
#pragma unroll
  for (int s = 0; s < RSTAGE; ++s)
    if (s < ksteps) issue(s);
  if (RSTAGE - 1 >= ksteps)
    lds_dma::wait_vm_barrier<0>();
  else
    lds_dma::wait_vm_barrier<NQ * (RSTAGE - 1)>();
  Frags fa, fb;
  if (!idle) read_frags(0, fa);
  auto step = [&](int s, Frags& use, Frags& next) {
    // slab s + 1 has landed in every wave and every wave's reads of slab s
    // are done: its slot takes slab s + RSTAGE
    if (s + RSTAGE - 1 < ksteps)
      lds_dma::wait_vm_barrier<NQ * (RSTAGE - 2)>();
    else
      lds_dma::wait_vm_barrier<0>();
    if (s + RSTAGE < ksteps) issue(s + RSTAGE);
    if (!idle) {
      // past the last slab this reads a stale slot that is never used
      read_frags(s + 1, next);
#pragma unroll
      for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) {
          acc_hh[i][j] = SplitT<T>::mfma(use.bh[j], use.ah[i], acc_hh[i][j]);
          acc_x[i][j] = SplitT<T>::mfma(use.bl[j], use.ah[i], acc_x[i][j]);
          acc_x[i][j] = SplitT<T>::mfma(use.bh[j], use.al[i], acc_x[i][j]);
        }
    }
  };
  // ksteps is even
  for (int s = 0; s < ksteps; s += 2) {
    step(s, fa, fb);
    step(s + 1, fb, fa);
  }
  // every wave is done with the ring before the next tile's prologue (and
  // before an off-diagonal tile stages X_next in it)
  lds_dma::wait_vm_barrier<0>();
  const double ca = g.sigma ? 1.0 : -1.0, cb = g.sigma ? 0.0 : 2.0;
  const int core = pd.core;
  constexpr int SP = TM + 1;  // staging pitch (doubles): column c of the tile at c * SP
  double sum[1] = {0.0};
  if (!idle) {
#pragma unroll
    for (int i = 0; i < MI; ++i) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int row = r0 + wm * WMR + i * 16 + (lane & 15);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int col = c0 + wn * WNC + j * 16 + 4 * (lane >> 4) + r;
          if (row > col) continue;
          const bool mir = (row != col);
          const size_t idx = row + (size_t)col * m;
          const double y = (double)acc_hh[i][j][r] + (1.0 / kScale) * (double)acc_x[i][j][r];
          const double x = X[idx];
          const double z = fma(cb, x, ca * y);
          const double d = x - y;
          const double wi = ((row < core) ? 1.0 : 0.0) + ((mir && col < core) ? 1.0 : 0.0);
          sum[0] += wi * d * d;
          const T h = (T)z;
          const T l = (T)((z - (double)h) * kScale);
          Xn[idx] = z;
          hin[idx] = h;
          lon[idx] = l;
          if (diag) {
            // inside a diagonal tile the upper half of the planes is mirrored
            // directly (the FP64 master is kept in the upper triangle only:
            // every reader takes that half)
            const size_t tdx = mir ? col + (size_t)row * m : idx;
            hin[tdx] = h;
            lon[tdx] = l;
          } else {
            stage[(col - c0) * SP + (row - r0)] = z;
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
    }
  }
  if (!diag) {
    // the mirror block of the hi / lo planes (rows c0 .., columns r0 ..; the
    // operand copies read whole rows of a plane, the FP64 master is read in
    // its upper triangle only): thread tid takes tile column c = tid % TN,
    // i.e. memory row c0 + c, for every (RNT / TN)-th tile row, so that
    // consecutive lanes store consecutive addresses
    __syncthreads();
    const int c = tid % TN;
#pragma unroll 4
    for (int rr = tid / TN; rr < TM; rr += RNT / TN) {
      const double z = stage[c * SP + rr];
      const T h = (T)z;
      const T l = (T)((z - (double)h) * kScale);
      const size_t tdx = (c0 + c) + (size_t)(r0 + rr) * m;
      hin[tdx] = h;
      lon[tdx] = l;
    }
  }
  block_store_partials<1, RNT>(sum, g.part_sums + t);
}

template <typename T>
__global__ __launch_bounds__(RNT, 1) void batched_split_ring_kernel(BatchArgs g) {
  __shared__ __attribute__((aligned(1024))) int8_t lds[kRingLds];
  const int8_t(*ring)[kRingSlot] = (const int8_t(*)[kRingSlot])lds;
  double* stage = (double*)lds;
  __shared__ int s_ticket;
  const int home = blockIdx.x & 7;
  for (int q = 0; q < 8; ++q) {
    const int qi = (home + q) & 7;
    const int first = g.queue_off[qi], count = g.queue_off[qi + 1] - first;
    for (;;) {
      __syncthreads();  // every thread has read the previous ticket
      if (threadIdx.x == 0) s_ticket = atomicAdd(&g.tickets[qi], 1);
      __syncthreads();
      const int k = s_ticket;
      if (k >= count) break;
      const int t = first + k;
      const int4 tl = g.tiles[t];
      if (tl.w == 3)
        split_tile_ring<T, 128, 128>(g, t, tl, ring, stage);
      else if (tl.w == 1)
        split_tile_ring<T, 128, 64>(g, t, tl, ring, stage);
      else
        split_tile_ring<T, 64, 64>(g, t, tl, ring, stage);
    }
  }
}

// Persistent work-queue launch of split tiles (the queue walk of
// batched_sp2_kernel over the 64-wide tile list).
template <typename T>
__global__ __launch_bounds__(NT, 4) void batched_split_kernel(BatchArgs g) {
  __shared__ __attribute__((aligned(16))) T smem[kSplitLds / sizeof(T)];
  __shared__ int s_ticket;
  const int home = blockIdx.x & 7;
  for (int q = 0; q < 8; ++q) {
    const int qi = (home + q) & 7;
    const int first = g.queue_off[qi], count = g.queue_off[qi + 1] - first;
    for (;;) {
      __syncthreads();  // every thread has read the previous ticket
      if (threadIdx.x == 0) s_ticket = atomicAdd(&g.tickets[qi], 1);
      __syncthreads();
      const int k = s_ticket;
      if (k >= count) break;
      const int t = first + k;
      split_tile<T>(g, t, g.tiles[t], smem);
    }
  }
}

// hi / lo planes of the packed FP64 X (padding stays zero).
template <typename T>
__global__ void split_planes_kernel(const double* __restrict__ x, T* __restrict__ hi,
                                    T* __restrict__ lo, size_t count) {
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < count; i += (size_t)gridDim.x * NT) {
    const double z = x[i];
    const T h = (T)z;
    hi[i] = h;
    lo[i] = (T)((z - (double)h) * SplitT<T>::kScale);
  }
}

typedef void (*BatchKernel)(BatchArgs);

// Threads per workgroup of the configured kernel.
int batch_threads(const BatchPlan& p) {
  return (p.split && p.wide && p.depth == RSTAGE && p.kslab == WBK) ? RNT : NT;
}

// The compiled configurations: (panel, slabs of prefetch, slab depth); the
// split-precision kernels have three, (64, 1, 64), (128, 1, 32) and the ring
// form (128, 4, 32).
BatchKernel batch_kernel(const BatchPlan& p) {
  if (p.split) {
    if (!p.wide && p.depth == 1 && p.kslab == SBK)
      return p.split == 2 ? batched_split_kernel<__bf16> : batched_split_kernel<_Float16>;
    if (p.wide && p.depth == 1 && p.kslab == WBK)
      return p.split == 2 ? batched_split_wide_kernel<__bf16>
                          : batched_split_wide_kernel<_Float16>;
    if (p.wide && p.depth == RSTAGE && p.kslab == WBK)
      return p.split == 2 ? batched_split_ring_kernel<__bf16>
                          : batched_split_ring_kernel<_Float16>;
    return nullptr;
  }
  if (p.wide) return (p.depth == 1 && p.kslab == 16) ? batched_sp2_kernel<true, 1, 16> : nullptr;
  if (p.kslab == 32) return p.depth == 1 ? batched_sp2_kernel<false, 1, 32> : nullptr;
  if (p.kslab == 16)
    return p.depth == 2 ? batched_sp2_kernel<false, 2, 16>
                        : (p.depth == 1 ? batched_sp2_kernel<false, 1, 16> : nullptr);
  return nullptr;
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
    // X is read in its upper triangle: the split kernels keep only that half
    // of the FP64 master current (the FP64 kernels write both)
    for (int r = threadIdx.x; r < pd.core; r += NT)
      rho[idx[r] + (size_t)gc * ldr] =
          scale * X[std::min(r, c) + (size_t)std::max(r, c) * pd.m];
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

// Loewdin factor of each part: U_p = V_p diag(w_p^-1/2) (the congruence
// transform Z_p = V_p w_p^-1/2 V_p^T of prg_buildZdiag, formed by the U V^T
// product after this scaling); wmin[p] = the part's smallest eigenvalue
// (padding eigenvalues are the gather's positive pad).
__global__ void invsqrt_scale_kernel(const double* __restrict__ v, const double* __restrict__ w,
                                     const PartDesc* __restrict__ parts,
                                     const size_t* __restrict__ w_offset, double* __restrict__ u,
                                     double* __restrict__ wmin) {
  const PartDesc pd = parts[blockIdx.y];
  const double* W = w + w_offset[blockIdx.y];
  const double* V = v + pd.offset;
  double* U = u + pd.offset;
  for (int c = blockIdx.x; c < pd.m; c += gridDim.x) {
    const double f = W[c] > 0.0 ? 1.0 / sqrt(W[c]) : 0.0;
    for (int r = threadIdx.x; r < pd.m; r += NT) U[r + (size_t)c * pd.m] = f * V[r + (size_t)c * pd.m];
  }
  if (blockIdx.x == 0) {
    __shared__ double red[NT];
    double s = INFINITY;
    for (int c = threadIdx.x; c < pd.m; c += NT) s = fmin(s, W[c]);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] = fmin(red[threadIdx.x], red[threadIdx.x + st]);
      __syncthreads();
    }
    if (threadIdx.x == 0) wmin[blockIdx.y] = red[0];
  }
}

// vv[it] += sum of iteration it's tile partials (fixed order); block it.
__global__ void vv_accumulate_kernel(const double* __restrict__ part, int ntiles,
                                     double* __restrict__ vv) {
  __shared__ double red[NT];
  const double* p = part + (size_t)blockIdx.x * ntiles;
  double s = 0.0;
  for (int b = threadIdx.x; b < ntiles; b += NT) s += p[b];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) vv[blockIdx.x] += red[0];
}

__global__ void zero_int_kernel(int* __restrict__ p, int count) {
  for (int i = threadIdx.x; i < count; i += NT) p[i] = 0;
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
                 int* d_tickets, int parity) {
  BatchArgs g;
  g.x = x;
  g.xn = xn;
  g.hi = plan.planes[2 * parity];
  g.lo = plan.planes[2 * parity + 1];
  g.hin = plan.planes[2 * (parity ^ 1)];
  g.lon = plan.planes[2 * (parity ^ 1) + 1];
  if (plan.split && (!g.hi || !g.lo || !g.hin || !g.lon)) {
    set_error(ERR_ARG, "batched SP2: split planes not set");
    return ERR_ARG;
  }
  g.parts = plan.d_parts;
  g.tiles = plan.d_tiles;
  g.sigma = sigma;
  g.part_sums = d_part;
  g.queue_off = plan.d_queue_off;
  g.tickets = d_tickets;
  BatchKernel kernel = batch_kernel(plan);
  if (!kernel) {
    set_error(ERR_ARG, "batched SP2: no kernel for this configuration");
    return ERR_ARG;
  }
  if (plan.grid <= 0 || plan.grid % 8) {
    set_error(ERR_ARG, "batched SP2: launch slots not set");
    return ERR_ARG;
  }
  const int grid = plan.grid;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(batch_threads(plan)), 0, stream(), g);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

bool batched_config_valid(const BatchPlan& plan) { return batch_kernel(plan) != nullptr; }

int batched_slots(BatchPlan* plan) {
  BatchKernel kernel = batch_kernel(*plan);
  if (!kernel) {
    set_error(ERR_ARG, "batched SP2: no kernel for this configuration");
    return ERR_ARG;
  }
  int nb = 0, cus = 0, dev = 0;
  PRG_HIP_CHECK(
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, batch_threads(*plan), 0));
  PRG_HIP_CHECK(hipGetDevice(&dev));
  PRG_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // one workgroup per resident slot, a multiple of 8 (workgroup b runs on XCD b mod 8)
  plan->grid = ((cus * std::max(nb, 1) + 7) / 8) * 8;
  return OK;
}

int batched_split_prepare(const BatchPlan& plan, const double* x) {
  if (plan.split == 2)
    hipLaunchKernelGGL(split_planes_kernel<__bf16>, dim3(2048), dim3(NT), 0, stream(), x,
                       (__bf16*)plan.planes[0], (__bf16*)plan.planes[1], plan.total);
  else
    hipLaunchKernelGGL(split_planes_kernel<_Float16>, dim3(2048), dim3(NT), 0, stream(), x,
                       (_Float16*)plan.planes[0], (_Float16*)plan.planes[1], plan.total);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int batched_zero_tickets(int* d_tickets, int count) {
  hipLaunchKernelGGL(zero_int_kernel, dim3(1), dim3(NT), 0, stream(), d_tickets, count);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int batched_vv(const BatchPlan& plan, const double* d_part, int iters, double* d_vv) {
  if (iters <= 0) return OK;
  hipLaunchKernelGGL(vv_accumulate_kernel, dim3(iters), dim3(NT), 0, stream(), d_part,
                     plan.ntiles, d_vv);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int batched_scatter(const BatchPlan& plan, const double* x, DMat rho, double scale, bool zero) {
  if (zero)
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

int batched_invsqrt_scale(const BatchPlan& plan, const double* v, const double* w,
                          const size_t* d_w_offset, double* u, double* d_wmin) {
  hipLaunchKernelGGL(invsqrt_scale_kernel, dim3(64, plan.nparts), dim3(NT), 0, stream(), v, w,
                     plan.d_parts, d_w_offset, u, d_wmin);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

}  // namespace prg
