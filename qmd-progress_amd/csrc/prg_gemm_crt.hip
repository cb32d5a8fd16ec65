// FP64-accurate SP2 products from exact int8 MFMA arithmetic (gfx950).
//
// PRG_PREC_INT8CRT.  X is symmetric with its spectrum in [0, 1], so every row
// has ||X_i||_2 <= 1.  With A = rint(2^a X) (an integer matrix) every entry of
// the integer product C = A A^T obeys |C_ij| <= ||A_i|| ||A_j|| < M/2 when
// 2^(2a + 1.001) <= M, M being the product of the moduli below.  C is then fixed by
// its residues modulo each p_k (Chinese remainder theorem), and each residue
// is one exact int8 product:
//   C mod p_k = (A mod p_k)(A mod p_k)^T mod p_k,
// with the residues stored as symmetric int8 values in [-128, 127] and
// v_mfma_i32_16x16x64_i8 accumulating in int32 (|sum| <= ld * 128^2 < 2^31 for
// ld < 2^17).  The reconstruction needs no wide integers:
//   C / M = sum_k ((C mod p_k) * q_k mod p_k) / p_k   (mod 1),
//   q_k = (M / p_k)^-1 mod p_k,
// summed in FP64 per output element as the moduli are processed, so that
// X^2 ~= frac * M / 2^(2a).  The only rounding is that of A (2^-(a+1) per
// entry) and of the FP64 fraction sum (each term a symmetric remainder in
// [-1/2, 1/2], the sum reduced mod 1 once at the end).
//
// Two kernels per SP2 step:
//  * crt_product_kernel: one workgroup per (256 x 256 upper-triangle tile,
//    modulus) item computes that modulus's int8 product over the full K range
//    and stores ((C mod p) q mod p) as symmetric int8 values (64 KB per item);
//  * crt_update_kernel: one workgroup per upper-triangle 64 x 64 block sums
//    the fractions over the moduli in FP64, applies the SP2 branch update
//    against the FP64 master X, writes the next master (lower triangle; the
//    final copy of the solve restores the upper one) and its residue planes
//    (both triangle positions) and emits the same trace / Frobenius partial
//    sums as the FP64 kernel, so the FP64 bookkeeping is shared.
// Items are independent, so the product grid is tiles x moduli (5808
// workgroups at N = 8192) and the 256 x 256 tiles halve the operand bytes per
// MFMA of a 128 x 128 tile.  Residue planes are stored in 64-column blocks
// (element (r, c) at (c / 64) * pld * 64 + r * 64 + c % 64, pld = ld rounded
// up to 256), so a 256-row x 64-byte operand stage is one contiguous 16 KB
// panel and a 64 x 64 block of a plane is one contiguous 4 KB.
#include "prg_sp2.h"
#include "prg_lds_dma.h"
#include "prg_tile.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <type_traits>
#include <vector>


namespace prg {
namespace {

using tile::xcd_remap;

typedef int i4 __attribute__((ext_vector_type(4)));

constexpr int BK = 64;                 // K bytes per operand stage
constexpr int TM = 256;                // product tile edge
constexpr int SB = 64;                 // update block edge (and plane column block)
constexpr int NT = 256;
constexpr int kPanel = TM * BK;        // one operand panel of a stage (16 KB)
constexpr int kStage = 2 * kPanel;     // A and B panels (32 KB)
constexpr int kItemBytes = TM * TM;    // product residues of one (tile, modulus) item
constexpr int kPitch = SB + 4;         // LDS row pitch of the update kernel's residues
constexpr int kMaxModuli = 14;

// Pairwise coprime moduli <= 256, largest first.
constexpr int kModuli[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239,
                                     233, 229, 227, 223, 217, 211, 199};

constexpr int kRecon = 6;  // moduli that rebuild A itself (M6 = 2^47.8)

constexpr bool moduli_odd_after_first() {
  for (int k = 1; k < kMaxModuli; ++k)
    if (kModuli[k] % 2 == 0) return false;
  return kModuli[0] == 256;
}
static_assert(moduli_odd_after_first(), "the residue shortcuts assume 256 first, then odd moduli");

struct CrtConst {
  double p[kMaxModuli];    // moduli
  double inv[kMaxModuli];  // 1 / p
  double q[kMaxModuli];    // (M / p)^-1 mod p
  double q6[kRecon];       // (M6 / p)^-1 mod p over the first kRecon moduli
  double m6;               // M6, the product of the first kRecon moduli
  double scale;            // 2^a
  double iscale;           // 2^-a
  double cscale;           // M / 2^(2a)
  double icscale;          // 2^(2a) / M
  int recon;               // X is carried by its residues alone (a <= 45)
  int shift;               // SP2 square: the residues stand for C - 2^a A
};

struct ProductArgs {
  const int8_t* planes[2];  // residue planes ping-pong, plane k at k * pld * pld
  const int8_t* planes_b;   // B operand's planes (general products), or null: B = A
  int8_t* res;              // product residues, item (t, k) at (t * np + k) * kItemBytes
  const int4* items;        // (modulus, I, J, upper-triangle index of (I, J))
  const Sp2State* st;       // SP2 state (planned, current buffer), or null: planes[0]
  const double* fold;       // (p, 1/p, q, q/p) per modulus
  int pld, np;
  int nitems;               // items in the list (the persistent kernel's range)
};

struct UpdateArgs {
  const double* buf[2];     // FP64 master ping-pong (ld x ld, column-major)
  int8_t* planes[2];
  const int8_t* res;
  const Sp2State* st;
  double* part;
  int ld, pld;
  CrtConst c;
};

// The LDS DMA helpers (dma16, wait_vm_barrier, lds_addr, uniform64).  M0 is
// clobbered by dma16; no compiler-generated code in this file keeps a value
// in M0 (checked in the generated assembly).
using namespace lds_dma;

// Symmetric residue of the integer-valued double v (|v| <= 2^53) modulo p,
// in [-128, 127]: v - p rint(v / p), formed exactly by the fma.  v * inv
// carries a relative error of 2^-52, so for the largest |v| the quotient may
// round to a neighbouring multiple; the two range checks move such a result
// back by one p (and map the p = 256 remainder 128 to -128).
__device__ __forceinline__ int sym_residue(double v, double p, double inv) {
  double m = fma(-p, rint(v * inv), v);
  m = (m > 127.0) ? m - p : m;
  m = (m < -128.0) ? m + p : m;
  return (int)m;
}

// The same for an odd modulus p and |v| / p < 2^38 (the product residues
// C q, |v| < 2^39, and X_next scaled by 2^a <= 2^45): v * inv is then within
// 2^-14 of v / p, while the exact quotient of an integer v stays at least
// 1 / (2p) >= 2^-9 from a half-integer, so rint gives the nearest multiple
// and v - p rint(.) lies in [-(p - 1) / 2, (p - 1) / 2]: no correction.
__device__ __forceinline__ int sym_residue_odd(double v, double p, double inv) {
  return (int)fma(-p, rint(v * inv), v);
}

// 1.5 * 2^52: for an integer m with |m| < 2^51, the low 32 bits of the double
// kMagic + m (exact) are m as a two's-complement int32, so a residue formed
// as fma(-p, t, v + kMagic) is read without a conversion.  The int8 stored
// for m is the low byte of that dword (+128 and -128, the two readings of a
// p = 256 tie, share the byte 0x80).
constexpr double kMagic = 6755399441055744.0;

// The low bytes of four int32 values as one dword (e0 lowest).  v_perm picks
// bytes of its operand pair: selector byte s < 4 takes byte s of the second
// operand, 4 + s byte s of the first, and 0x0c gives zero.
__device__ __forceinline__ uint32_t pack_low_bytes(int e0, int e1, int e2, int e3) {
  const uint32_t lo = __builtin_amdgcn_perm((uint32_t)e1, (uint32_t)e0, 0x0c0c0400u);
  const uint32_t hi = __builtin_amdgcn_perm((uint32_t)e3, (uint32_t)e2, 0x0c0c0400u);
  return __builtin_amdgcn_perm(hi, lo, 0x05040100u);
}

// Transpose of a 4 x 4 byte block given as its rows a, b, c, d (byte j of a
// row is column j): col[j] = (a_j, b_j, c_j, d_j).
__device__ __forceinline__ void transpose_4x4(uint32_t a, uint32_t b, uint32_t c, uint32_t d,
                                              uint32_t (&col)[4]) {
  const uint32_t ab01 = __builtin_amdgcn_perm(b, a, 0x05010400u);  // a0 b0 a1 b1
  const uint32_t ab23 = __builtin_amdgcn_perm(b, a, 0x07030602u);  // a2 b2 a3 b3
  const uint32_t cd01 = __builtin_amdgcn_perm(d, c, 0x05010400u);
  const uint32_t cd23 = __builtin_amdgcn_perm(d, c, 0x07030602u);
  col[0] = __builtin_amdgcn_perm(cd01, ab01, 0x05040100u);
  col[1] = __builtin_amdgcn_perm(cd01, ab01, 0x07060302u);
  col[2] = __builtin_amdgcn_perm(cd23, ab23, 0x05040100u);
  col[3] = __builtin_amdgcn_perm(cd23, ab23, 0x07060302u);
}

// One (tile, modulus) item: C = A_I B_J^T mod p over the full K range (B = A
// for the SP2 square; for a general product of symmetric A and B, B_J^T is
// B's column panel J).  W = 8
// waves in a 2 x 4 grid of 128 x 64 (two waves per SIMD) or W = 4 waves in a
// 2 x 2 grid of 128 x 128 (one wave per SIMD, a third fewer LDS fragment
// reads per MFMA); 16 x 16 MFMA blocks, int32 accumulators.  Operand stages
// are copied global -> LDS directly (global_load_lds, 16 bytes per lane),
// NSTAGE - 1 stages ahead of the MFMAs.  A stage holds the A panel (256 rows
// x 64 bytes) then the B panel, as thirty-two 1 KB chunks of 16 rows; within
// a row the 16-byte chunk g sits in slot g ^ ((row >> 1) & 3), which keeps
// the ds_read_b128 of 8 consecutive rows on distinct banks.  The ring is the
// kernel's only LDS object.
// Per-workgroup constants of the product kernels (W waves, NSTAGE-slot ring).
template <int W>
struct ProductLane {
  static constexpr int WN = W / 2;          // waves along the tile columns
  static constexpr int MT = 8;              // 16-row MFMA blocks per wave
  static constexpr int NTL = TM / WN / 16;  // 16-column MFMA blocks per wave
  static constexpr int CPW = 32 / W;        // 1 KB DMA chunks per wave per stage
  int lane, w, wm, wn, drow, kc, fo, offa;
  uint32_t dst0, voff;
  __device__ ProductLane(const int8_t* ring0) {
    const int tid = threadIdx.x;
    lane = tid & 63;
    w = tid >> 6;
    wm = w & 1;
    wn = w >> 1;
    // DMA chunks of wave w: c = w + W q (q < CPW) is chunk c % 16 of panel
    // c / 16 (A = rows of tile I, then B = rows of tile J): panel rows
    // 16 (c % 16) + lane / 4, slot lane % 4.  c / 16 = W q / 16 and the row
    // offset from 16 w + lane / 4 are compile-time; multiples of 16 rows
    // leave the swizzle unchanged.
    drow = 16 * w + (lane >> 2);
    kc = 16 * ((lane & 3) ^ ((drow >> 1) & 3));
    voff = (uint32_t)(drow * BK + kc);
    dst0 = lds_addr(ring0) + 1024 * w;
    // LDS read offsets of this lane's fragments: row r = lane % 16 of each
    // 16-row block, 16-byte chunk lane / 16; the block offsets are multiples
    // of 8 rows, so the swizzle depends on the lane only.
    fo = (lane & 15) * BK + 16 * ((lane >> 4) ^ (((lane & 15) >> 1) & 3));
    offa = wm * 128 * BK + fo;
  }
};

// Operand panels of one (tile, modulus) item: the wave-uniform addresses of
// their first rows (each lane adds its ProductLane::voff).
struct ItemSrc {
  uint64_t a;
  uint64_t b;
  bool diag;  // I == J and B = A: one panel, loaded once
};

template <int W>
__device__ __forceinline__ ItemSrc item_src(const ProductArgs& g, int cur, const int4& it,
                                            const ProductLane<W>&) {
  const size_t plane = (size_t)g.pld * g.pld;
  const int8_t* R = g.planes[cur] + (size_t)it.x * plane;
  const int8_t* RB = g.planes_b ? g.planes_b + (size_t)it.x * plane : R;
  ItemSrc s;
  s.a = uniform64((uint64_t)(uintptr_t)(R + (size_t)it.y * TM * BK));
  s.b = uniform64((uint64_t)(uintptr_t)(RB + (size_t)it.z * TM * BK));
  s.diag = (it.y == it.z) && !g.planes_b;
  return s;
}

// Stage s of an item into ring slot s % NSTAGE: NQ 1 KB chunks per wave
// (half of them, the A panel only, for a diagonal item).
template <int W, int NSTAGE, bool DIAG>
__device__ __forceinline__ void issue_stage(const ItemSrc& src, size_t kstride,
                                            const ProductLane<W>& L, int s) {
  constexpr int NQ = DIAG ? ProductLane<W>::CPW / 2 : ProductLane<W>::CPW;
  const size_t off = (size_t)s * kstride;
  const uint32_t base = __builtin_amdgcn_readfirstlane(L.dst0 + kStage * (s % NSTAGE));
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int panel = (W * q) >> 4;
    const int rows = 16 * W * q - TM * panel;
    dma16((panel ? src.b : src.a) + off + (size_t)rows * BK, L.voff, base + 1024 * W * q);
  }
}

// The first NSTAGE stages of an item (its prologue).
template <int W, int NSTAGE>
__device__ __forceinline__ void issue_prologue(const ItemSrc& src, size_t kstride,
                                               const ProductLane<W>& L, int ksteps) {
#pragma unroll
  for (int s = 0; s < NSTAGE; ++s) {
    if (s < ksteps) {
      if (src.diag)
        issue_stage<W, NSTAGE, true>(src, kstride, L, s);
      else
        issue_stage<W, NSTAGE, false>(src, kstride, L, s);
    }
  }
}

// One item's K loop after its prologue was issued: every stage lands through
// the ring, NSTAGE - 1 stages ahead of the MFMAs.  all_landed: the prologue
// is complete already (wait for everything once instead of for stage 0).
// Iteration s: wait until stage s + 1 has landed in every wave (and this
// wave's reads of stage s are done), refill the slot of stage s with stage
// s + NSTAGE, then run the MFMAs of stage s while reading stage s + 1.  The
// first step starts from zero accumulators (peeled, so that the loop carries
// only MFMA results and they stay in AGPRs).
template <int W, int NSTAGE, bool DIAG, bool IL = false>
__device__ __forceinline__ void item_mainloop(
    const ItemSrc& src, size_t kstride, const ProductLane<W>& L, int ksteps,
    const int8_t (*ring)[kStage], bool all_landed,
    i4 (&acc)[ProductLane<W>::MT][ProductLane<W>::NTL]) {
  constexpr int MT = ProductLane<W>::MT, NTL = ProductLane<W>::NTL, WN = ProductLane<W>::WN;
  constexpr int NQ = DIAG ? ProductLane<W>::CPW / 2 : ProductLane<W>::CPW;
  const int offb = (DIAG ? 0 : kPanel) + L.wn * (TM / WN) * BK + L.fo;
  i4 fa[MT + NTL], fb[MT + NTL];  // A fragments, then B fragments
  auto read_frags = [&](int s, i4 (&f)[MT + NTL]) {
    const int8_t* st = ring[s % NSTAGE];
#pragma unroll
    for (int m = 0; m < MT; ++m) f[m] = *(const i4*)(st + L.offa + 1024 * m);
#pragma unroll
    for (int n = 0; n < NTL; ++n) f[MT + n] = *(const i4*)(st + offb + 1024 * n);
  };
  // IL: the interleaved issue order.  PRE MFMAs of the step go before its
  // barrier: they read only fragments already in registers, so a wave that
  // reaches the barrier early leaves work queued in its SIMD's MFMA pipe
  // while it waits.  After the barrier, each of the MT + NTL fragment reads
  // of stage s + 1 is followed by two MFMAs of stage s, so that the reads
  // overlap the MFMAs and have completed by the next barrier.
  constexpr int PRE = IL ? 8 : 0;
  static_assert(PRE + 2 * (MT + NTL) <= MT * NTL, "schedule groups");
  auto step = [&](int s, i4 (&use)[MT + NTL], i4 (&next)[MT + NTL], bool first) {
    const i4 zero = {0, 0, 0, 0};
    if constexpr (PRE > 0) {
#pragma unroll
      for (int i = 0; i < PRE; ++i) {
        const int m = i / NTL, n = i % NTL;
        acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(use[MT + n], use[m],
                                                          first ? zero : acc[m][n], 0, 0, 0);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    if (s + NSTAGE - 1 < ksteps)
      wait_vm_barrier<NQ * (NSTAGE - 2)>();
    else
      wait_vm_barrier<0>();
    if (s + NSTAGE < ksteps) issue_stage<W, NSTAGE, DIAG>(src, kstride, L, s + NSTAGE);
    // after the last stage this reads a stale slot that is never used
    read_frags(s + 1, next);
#pragma unroll
    for (int i = PRE; i < MT * NTL; ++i) {
      const int m = i / NTL, n = i % NTL;
      acc[m][n] = __builtin_amdgcn_mfma_i32_16x16x64_i8(use[MT + n], use[m],
                                                        first ? zero : acc[m][n], 0, 0, 0);
    }
    if constexpr (IL) {
#pragma unroll
      for (int i = 0; i < MT + NTL; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
        __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);  // MFMA
      }
      __builtin_amdgcn_sched_group_barrier(0x008, MT * NTL - PRE - 2 * (MT + NTL), 0);
    }
  };
  if (all_landed || NSTAGE - 1 >= ksteps)
    wait_vm_barrier<0>();
  else
    wait_vm_barrier<NQ * (NSTAGE - 1)>();
  read_frags(0, fa);
  // ksteps = pld / 64 is a multiple of 4
  step(0, fa, fb, true);
  step(1, fb, fa, false);
  for (int s = 2; s < ksteps; s += 2) {
    step(s, fa, fb, false);
    step(s + 1, fb, fa, false);
  }
}

// (C q mod p) as symmetric int8.  Operands were swapped, so this lane holds
// C[row][col] with row = lane % 16 and col = 4 (lane / 16) + r inside each
// 16 x 16 block; the four bytes of r form one dword.  Records follow the
// 8-wave layout whatever W is: the 64-column group g = (wn * NTL + n) / 4
// and row half wm name record wave wm + 2 g; (record wave, m, lane) holds
// the four blocks n % 4 of that group as one 16-byte store (1 KB contiguous
// per record wave and m).
template <int W>
__device__ __forceinline__ void item_epilogue(
    const ProductArgs& g, const int4& it, const ProductLane<W>& L,
    const i4 (&acc)[ProductLane<W>::MT][ProductLane<W>::NTL]) {
  constexpr int MT = ProductLane<W>::MT, NTL = ProductLane<W>::NTL;
  const int k = it.x, t = it.w;
  // p, q and q / p of modulus k: C q = v is exact (|v| < 2^39), and v / p
  // is read as C (q / p), within 2^-19 of v / p, which for odd p lies at
  // least 2^-9 from a half-integer (for p = 256 both are exact), so the
  // rounded quotient is rint(v / p) and fma(-p, t, v + kMagic) the
  // symmetric residue (a p = 256 tie, +128, stores the byte of -128)
  const double p = g.fold[4 * k], qm = g.fold[4 * k + 2], qmi = g.fold[4 * k + 3];
  i4* dst = (i4*)(g.res + ((size_t)t * g.np + k) * kItemBytes) + L.lane;
#pragma unroll
  for (int m = 0; m < MT; ++m) {
#pragma unroll
    for (int h = 0; h < NTL / 4; ++h) {
      i4 out;
#pragma unroll
      for (int n4 = 0; n4 < 4; ++n4) {
        const int n = 4 * h + n4;
        int e[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double a = (double)acc[m][n][r];
          e[r] = __double2loint(fma(-p, rint(a * qmi), fma(a, qm, kMagic)));
        }
        out[n4] = (int)pack_low_bytes(e[0], e[1], e[2], e[3]);
      }
      const int rw = L.wm + 2 * (L.wn * (NTL / 4) + h);
      dst[(rw * 8 + m) * 64] = out;
    }
  }
}

// One (tile, modulus) item per workgroup: C = A_I B_J^T mod p over the full K
// range (B = A for the SP2 square; for a general product of symmetric A and
// B, B_J^T is B's column panel J).  W = 8 waves in a 2 x 4 grid of 128 x 64
// (two waves per SIMD) or W = 4 waves in a 2 x 2 grid of 128 x 128 (one wave
// per SIMD, a third fewer LDS fragment reads per MFMA); 16 x 16 MFMA blocks,
// int32 accumulators.  Operand stages are copied global -> LDS directly
// (global_load_lds, 16 bytes per lane), NSTAGE - 1 stages ahead of the
// MFMAs.  A stage holds the A panel (256 rows x 64 bytes) then the B panel,
// as thirty-two 1 KB chunks of 16 rows; within a row the 16-byte chunk g
// sits in slot g ^ ((row >> 1) & 3), which keeps the ds_read_b128 of 8
// consecutive rows on distinct banks.  The ring is the kernel's only LDS
// object.
template <int W, int NSTAGE>
__global__ __launch_bounds__(64 * W, 1) void crt_product_kernel(ProductArgs g) {
  __shared__ __attribute__((aligned(1024))) int8_t ring[NSTAGE][kStage];
  int cur = 0;
  if (g.st) {
    const volatile Sp2State* stv = g.st;
    if (!stv->planned) return;
    cur = stv->cur;
  }
  const ProductLane<W> L(&ring[0][0]);
  const int4 it = g.items[xcd_remap(blockIdx.x, gridDim.x)];
  const ItemSrc src = item_src<W>(g, cur, it, L);
  const size_t kstride = (size_t)g.pld * BK;  // one 64-column block of a plane
  const int ksteps = g.pld / BK;
  i4 acc[ProductLane<W>::MT][ProductLane<W>::NTL];
  issue_prologue<W, NSTAGE>(src, kstride, L, ksteps);
  if (src.diag)
    item_mainloop<W, NSTAGE, true>(src, kstride, L, ksteps, ring, false, acc);
  else
    item_mainloop<W, NSTAGE, false>(src, kstride, L, ksteps, ring, false, acc);
  item_epilogue<W>(g, it, L, acc);
}

// Persistent form: one workgroup per CU walks its XCD's share of the item
// list (the same contiguous eighths as xcd_remap gives the one-item form, in
// strides of the XCD's workgroup count) and issues the next item's prologue
// before the current item's epilogue, so that the residue arithmetic and
// stores overlap the next operand loads.  Each item starts with a full wait
// (the prologue and the previous item's residue stores), so loads and
// stores never share a partial vmcnt wait.  Every workgroup exits after its
// last item.
template <int W, int NSTAGE, bool IL = false>
__global__ __launch_bounds__(64 * W, 1) void crt_product_persistent_kernel(ProductArgs g) {
  __shared__ __attribute__((aligned(1024))) int8_t ring[NSTAGE][kStage];
  int cur = 0;
  if (g.st) {
    const volatile Sp2State* stv = g.st;
    if (!stv->planned) return;
    cur = stv->cur;
  }
  const ProductLane<W> L(&ring[0][0]);
  // XCD x = blockIdx.x % 8 owns items [lo, hi) (xcd_remap's split of the
  // list); its workgroups take lo + j, lo + j + nx, ...
  const int nitems = g.nitems, x = blockIdx.x & 7, j = blockIdx.x >> 3;
  const int nx = gridDim.x >> 3;
  const int per = nitems / 8, rem = nitems % 8;
  const int lo = (x < rem) ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
  const int hi = lo + per + (x < rem ? 1 : 0);
  int idx = lo + j;
  if (idx >= hi) return;
  const size_t kstride = (size_t)g.pld * BK;
  const int ksteps = g.pld / BK;
  int4 it = g.items[idx];
  ItemSrc src = item_src<W>(g, cur, it, L);
  issue_prologue<W, NSTAGE>(src, kstride, L, ksteps);
  i4 acc[ProductLane<W>::MT][ProductLane<W>::NTL];
  while (true) {
    if (src.diag)
      item_mainloop<W, NSTAGE, true, IL>(src, kstride, L, ksteps, ring, true, acc);
    else
      item_mainloop<W, NSTAGE, false, IL>(src, kstride, L, ksteps, ring, true, acc);
    // every wave has read its last stage (no loads are outstanding): the
    // ring is free for the next item's prologue
    wait_vm_barrier<0>();
    const int nidx = idx + nx;
    const bool more = nidx < hi;
    int4 nit = it;
    if (more) {
      nit = g.items[nidx];
      src = item_src<W>(g, cur, nit, L);
      issue_prologue<W, NSTAGE>(src, kstride, L, ksteps);
    }
    item_epilogue<W>(g, it, L, acc);
    if (!more) break;
    idx = nidx;
    it = nit;
  }
}

// X entry from its residues r_k (symmetric int8 of A mod p_k, k < kRecon):
// A/M6 = sum_k ((r_k q6_k) mod p_k) / p_k (mod 1), X = A 2^-a.
__device__ __forceinline__ double crt_rebuild(const int (&r)[kRecon], const CrtConst& c) {
  double S = 0.0;
#pragma unroll
  for (int k = 0; k < kRecon; ++k) {
    const double v = (double)r[k] * c.q6[k];  // exact: |v| < 2^15
    const double md = fma(-c.p[k], rint(v * c.inv[k]), v);
    S = fma(md, c.inv[k], S);
  }
  return rint((S - rint(S)) * c.m6) * c.iscale;
}

// One upper-triangle 64 x 64 block (I, J) of X^2: sum the moduli's fractions,
// apply the SP2 update, write X_next at (J, I) (the lower triangle), and the residue
// planes of X_next (through LDS, one modulus at a time, as contiguous 4 KB
// blocks in both positions).  The block lies in product tile (I / 4, J / 4),
// in the share of product wave (I % 4) / 2 + 2 (J % 4), MFMA rows
// 4 (I % 2) .. + 3.  Thread (b, lane) holds elements (row, col0(n) + r) with
// row = 64 I + 16 b + lane % 16, col0(n) = 64 J + 16 n + 4 (lane / 16).
template <int NP, bool RECON>
__global__ __launch_bounds__(NT, 3) void crt_update_kernel(UpdateArgs g) {
  __shared__ __attribute__((aligned(16))) int8_t rb[2][SB][kPitch];
  __shared__ double red[NT / 64][4];
  const volatile Sp2State* stv = g.st;
  if (!stv->planned) return;
  const int apply = stv->apply, sigma = stv->sigma, cur = stv->cur;
  // prg_sp2_submatrix: the idempotency sum runs over the first `core` rows
  // (and, by symmetry, the mirrored columns) only
  const int core_lim = (stv->variant == V_SUBMATRIX) ? stv->core : 0x7fffffff;
  const int ld = g.ld, pld = g.pld;
  int I, J;
  tile::sym_tile(blockIdx.x, I, J);
  const int bi = I >> 2, bj = J >> 2;
  const int t = bj * (bj + 1) / 2 + bi;
  const int wave = ((I & 3) >> 1) + 2 * (J & 3);
  const int tid = threadIdx.x, lane = tid & 63, b = tid >> 6;
  const int m = 4 * (I & 1) + b;
  const i4* src = (const i4*)(g.res + (size_t)t * NP * kItemBytes) + (wave * 8 + m) * 64 + lane;
  const int rl = 16 * b + (lane & 15);
  const int row = SB * I + rl;
  const int cl0 = 4 * (lane >> 4);
  const double* X = g.buf[cur];
  double* Xn = const_cast<double*>(g.buf[cur ^ 1]);

  // X at (row, col0 + r): with RECON rebuilt from the first kRecon residue
  // planes of X (4 bytes per plane and n at J * pld * 64 + row * 64 +
  // 16 n + cl0), otherwise read from the FP64 master as (col0 + r, row),
  // 32 contiguous bytes per n
  double xs[16];
  if (RECON) {
    const int8_t* Pc = g.planes[cur] + (size_t)J * pld * SB + (size_t)row * SB + cl0;
    const size_t plane = (size_t)pld * pld;
    int rw[kRecon][4];
#pragma unroll
    for (int k = 0; k < kRecon; ++k)
#pragma unroll
      for (int n = 0; n < 4; ++n) rw[k][n] = *(const int*)(Pc + (size_t)k * plane + 16 * n);
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int rk[kRecon];
#pragma unroll
        for (int k = 0; k < kRecon; ++k) rk[k] = (int)(int8_t)(rw[k][n] >> (8 * r));
        xs[4 * n + r] = crt_rebuild(rk, g.c);
      }
  } else {
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const double2* xp = (const double2*)(X + SB * J + 16 * n + cl0 + (size_t)row * ld);
      const double2 u = xp[0], v = xp[1];
      xs[4 * n] = u.x;
      xs[4 * n + 1] = u.y;
      xs[4 * n + 2] = v.x;
      xs[4 * n + 3] = v.y;
    }
  }
  // S = sum_k ((C q_k) mod p_k) / p_k, bytes (n, r) of the 16-byte record
  i4 d[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) d[k] = src[k * (kItemBytes / 16)];
  double S[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) S[i] = 0.0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const double inv = g.c.inv[k];
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        S[4 * n + r] = fma((double)(int8_t)(d[k][n] >> (8 * r)), inv, S[4 * n + r]);
  }

  const bool mirror = (I != J);
  const double ca = apply ? (sigma ? 1.0 : -1.0) : 0.0;
  const double cb = apply ? (sigma ? 0.0 : 2.0) : 1.0;
  double zz[16];
  double q0 = 0.0, q1 = 0.0, q2 = 0.0, q3 = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int col0 = SB * J + 16 * n + cl0;
    double z4[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double x = xs[4 * n + r];
      // X~ = A 2^-a, the operand the product squared (the master rounded to
      // the residues' grid when an FP64 master is kept)
      const double xa = RECON ? x : rint(x * g.c.scale) * g.c.iscale;
      // (C - 2^a A) / M = S - X~ 2^(2a) / M (mod 1) lies in (-1/2, 1/2), and
      // X~^2 = (C - 2^a A) / 2^(2a) + X~
      const double sr = fma(-xa, g.c.icscale, S[4 * n + r]);
      const double y = fma(sr - rint(sr), g.c.cscale, xa);
      const double z = fma(cb, x, ca * y);
      const double dxy = x - y;
      const bool diag = (row == col0 + r);
      // weight 1 + mirror for the whole matrix; core rows only for submatrix
      const double wi = (row < core_lim ? 1.0 : 0.0) + ((mirror && col0 + r < core_lim) ? 1.0 : 0.0);
      q0 += z * z;
      q2 += wi * dxy * dxy;
      q1 += diag ? z : 0.0;
      q3 += diag ? y : 0.0;
      z4[r] = z;
      zz[4 * n + r] = rint(z * g.c.scale);
    }
    // without RECON, X_next at (col0 + r, row) only: the lower triangle,
    // plus the whole of each diagonal block.  These are the positions the
    // next step reads; the solve's final copy rebuilds the upper triangle by
    // symmetry.  With RECON the residue planes below are X_next.
    if (!RECON) {
      double2* tp = (double2*)(Xn + col0 + (size_t)row * ld);
      tp[0] = double2{z4[0], z4[1]};
      tp[1] = double2{z4[2], z4[3]};
    }
  }
  // Residue planes of X_next, one modulus at a time through a double-buffered
  // LDS block rb[row][col] (4 columns per dword).  Block (I, J) of a plane:
  // rows SB I + [0, 64), columns SB J + [0, 64) at J * pld * 64 +
  // (SB I + row) * 64 + col; block (J, I) holds the transpose, written by
  // each thread as one transposed 4 x 4 byte block (rows qr .. qr + 3,
  // columns qc .. qc + 3 of rb), 64 contiguous bytes per 16 lanes.
  int8_t* P = g.planes[cur ^ 1];
  const size_t plane = (size_t)pld * pld;
  const int dr = tid >> 2, dc = 16 * (tid & 3);     // direct: row dr, bytes dc..dc+15
  const int qr = 4 * (tid & 15), qc = 4 * (tid >> 4);
  int8_t* pd = P + (size_t)J * pld * SB + (size_t)(SB * I + dr) * SB + dc;
  int8_t* pt = P + (size_t)I * pld * SB + (size_t)(SB * J + qc) * SB + qr;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const double p = g.c.p[k], inv = g.c.inv[k];
    int8_t (*buf)[kPitch] = rb[k & 1];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      int e[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (RECON && k == 0) {
          // |zz| <= 2^45 (a <= 45); p = kModuli[0] = 256, the one even
          // modulus: the quotient is exact and its tie (+128 or -128) is
          // the byte 0x80 either way, so no fold is needed
          e[r] = __double2loint(fma(-p, rint(zz[4 * n + r] * inv), zz[4 * n + r] + kMagic));
        } else if (RECON) {
          e[r] = sym_residue_odd(zz[4 * n + r], p, inv);
        } else {
          e[r] = sym_residue(zz[4 * n + r], p, inv);
        }
      }
      *(uint32_t*)&buf[rl][16 * n + cl0] = pack_low_bytes(e[0], e[1], e[2], e[3]);
    }
    // the writes of modulus k - 1 from this buffer's twin are done by every
    // thread once all pass this barrier; modulus k + 1 fills the twin next
    __syncthreads();
    const int* s4 = (const int*)&buf[dr][dc];
    *(i4*)(pd + (size_t)k * plane) = i4{s4[0], s4[1], s4[2], s4[3]};
    if (mirror) {
      uint32_t col[4];
      transpose_4x4(*(const uint32_t*)&buf[qr][qc], *(const uint32_t*)&buf[qr + 1][qc],
                    *(const uint32_t*)&buf[qr + 2][qc], *(const uint32_t*)&buf[qr + 3][qc], col);
#pragma unroll
      for (int j = 0; j < 4; ++j) *(uint32_t*)(pt + (size_t)k * plane + (size_t)j * SB) = col[j];
    }
  }
  // block sums (fixed order); off-diagonal blocks stand for their mirror too
  const double wz = mirror ? 2.0 : 1.0;
  double sums[4] = {wz * q0, q1, q2, q3};
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    double x = sums[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    sums[q] = x;
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) red[b][q] = sums[q];
  }
  __syncthreads();
  if (tid < 4) {
    double a = 0.0;
#pragma unroll
    for (int ww = 0; ww < NT / 64; ++ww) a += red[ww][tid];
    g.part[4 * (size_t)blockIdx.x + tid] = a;
  }
}

// Residue planes of an FP64 symmetric matrix in the 64-column block layout
// (pld x pld; entries outside the ld x ld buffer are zero).  The padding of
// the second plane buffer is cleared too: the update kernel writes only
// inside ld x ld.
template <int NP>
__global__ void crt_split_kernel(const double* __restrict__ x, int8_t* __restrict__ p0,
                                 int8_t* __restrict__ p1, int ld, int pld, CrtConst c) {
  const size_t count = (size_t)pld * pld;
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < count; i += (size_t)gridDim.x * NT) {
    const size_t q = i / SB;
    const int r = (int)(q % pld);
    const int col = (int)(q / pld) * SB + (int)(i % SB);
    const bool in = (r < ld) && (col < ld);
    // X is symmetric: (r, col) is read as (col, r), contiguous along col
    const double v = in ? rint(x[col + (size_t)r * ld] * c.scale) : 0.0;
#pragma unroll
    for (int k = 0; k < NP; ++k) p0[(size_t)k * count + i] = (int8_t)sym_residue(v, c.p[k], c.inv[k]);
    if (!in) {
#pragma unroll
      for (int k = 0; k < NP; ++k) p1[(size_t)k * count + i] = 0;
    }
  }
}

// out = scale X for the n x n block, X rebuilt from the first kRecon residue
// planes of the current buffer (used when the update kernel keeps no FP64
// master).  A thread rebuilds rows r0 .. r0 + 3 of one column, reading
// element (r, col) as (col, r): 4 contiguous bytes per plane.
__global__ void crt_finish_kernel(const int8_t* __restrict__ p0, const int8_t* __restrict__ p1,
                                  const Sp2State* __restrict__ s, int pld, CrtConst c,
                                  double scale, double* __restrict__ out, int ld_out, int n) {
  const int8_t* P = s->cur ? p1 : p0;
  const size_t plane = (size_t)pld * pld;
  const int rq = (n + 3) / 4;
  const size_t total = (size_t)rq * n;
  for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < total; i += (size_t)gridDim.x * NT) {
    const int r0 = 4 * (int)(i % rq), col = (int)(i / rq);
    const size_t off = (size_t)(r0 / SB) * pld * SB + (size_t)col * SB + (r0 % SB);
    int w[kRecon];
#pragma unroll
    for (int k = 0; k < kRecon; ++k) w[k] = *(const int*)(P + (size_t)k * plane + off);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int row = r0 + j;
      if (row >= n) break;
      int rk[kRecon];
#pragma unroll
      for (int k = 0; k < kRecon; ++k) rk[k] = (int)(int8_t)(w[k] >> (8 * j));
      out[row + (size_t)col * ld_out] = scale * crt_rebuild(rk, c);
    }
  }
}

// ---------------------------------------------------------------- general products
// C = alpha A B + beta Cin + gamma I for symmetric, commuting A and B (so that
// A B is symmetric).  A is scaled by 2^a / sA and B by 2^a / sB, sA and sB
// being powers of two above the largest row 2-norm of each, found on the
// device; then every row of the integer operands has 2-norm at most 2^a (plus
// sqrt(ld) / 2 from rounding), and the bound of the SP2 square applies to
// the integer product unchanged.  C = (C~ / M mod 1) M sA sB / 2^(2a).

// Per-block maxima of the column sums of squares of the n x n block of x
// (for a symmetric matrix, its squared row 2-norms); one wave per column.
__global__ void crt_colnorm_kernel(const double* __restrict__ x, int ld, int n,
                                   double* __restrict__ part) {
  __shared__ double red[NT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double best = 0.0;
  for (int c = blockIdx.x * (NT / 64) + w; c < n; c += gridDim.x * (NT / 64)) {
    const double* col = x + (size_t)c * ld;
    double s = 0.0;
    for (int r = 2 * lane; r + 1 < n; r += 128) {
      const double2 u = *(const double2*)(col + r);
      s = fma(u.x, u.x, s);
      s = fma(u.y, u.y, s);
    }
    if ((n & 1) && lane == 0) s = fma(col[n - 1], col[n - 1], s);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    best = fmax(best, s);
  }
  if (lane == 0) red[w] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) m = fmax(m, red[k]);
    part[blockIdx.x] = m;
  }
}

// out = (s, 1 / s), s the power of two above sqrt(max part) (1 for a zero
// matrix; non-finite input gives a non-finite s, which the outputs carry).
__global__ void crt_scale_kernel(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double red[NT];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += NT) m = fmax(m, part[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int st = NT / 2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + st]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double norm = sqrt(red[0]);
    double s = 1.0;
    if (!(norm < INFINITY)) {
      s = norm;  // inf or NaN
    } else if (norm > 0.0) {
      int e = 0;
      frexp(norm, &e);  // norm = f 2^e, f in [0.5, 1)
      s = ldexp(1.0, e);
    }
    out[0] = s;
    out[1] = 1.0 / s;
  }
}

// Residue planes of rint(2^a x / s) (s = sc[0], 1 / s = sc[1]) in the
// 64-column block layout, zero outside the ld x ld buffer.  A thread forms 16
// consecutive bytes of every plane: one row, 16 columns of a 64-column block
// (read as 128 contiguous bytes of the transposed position; ld is a multiple
// of 64, so the 16 columns are all inside or all outside).
template <int NP>
__global__ void crt_split_scaled_kernel(const double* __restrict__ x, int8_t* __restrict__ p0,
                                        int ld, int pld, CrtConst c,
                                        const double* __restrict__ sc) {
  const size_t count = (size_t)pld * pld;
  const size_t nchunk = count / 16;
  const double f = c.scale * sc[1];
  for (size_t ch = (size_t)blockIdx.x * NT + threadIdx.x; ch < nchunk;
       ch += (size_t)gridDim.x * NT) {
    const size_t i = 16 * ch;
    const size_t q = i / SB;
    const int r = (int)(q % pld);
    const int col0 = (int)(q / pld) * SB + (int)(i % SB);
    double v[16];
    if (r < ld && col0 < ld) {
      const double2* xp = (const double2*)(x + col0 + (size_t)r * ld);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const double2 u = xp[j];
        v[2 * j] = rint(u.x * f);
        v[2 * j + 1] = rint(u.y * f);
      }
    } else {
#pragma unroll
      for (int j = 0; j < 16; ++j) v[j] = 0.0;
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      i4 out;
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        int b4 = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) b4 |= (sym_residue(v[4 * w + j], c.p[k], c.inv[k]) & 0xff) << (8 * j);
        out[w] = b4;
      }
      *(i4*)(p0 + (size_t)k * count + i) = out;
    }
  }
}

struct GemmEpiArgs {
  const int8_t* res;
  const double* sa;   // (sA, 1 / sA)
  const double* sb;   // (sB, 1 / sB)
  const double* Cin;  // beta term, or null
  const double* E;    // <C, E> partial, or null
  double* C;
  double* part;       // per block [Tr C, ||C||_F^2, <C, E>, 0]
  double alpha, beta, gamma;
  int ld, n;
  CrtConst c;
};

// One upper-triangle 64 x 64 block (I, J) of C (same residue record mapping
// as crt_update_kernel): the CRT sum over the moduli, C at (row, col) and its
// mirror; inside a diagonal block the upper half is mirrored onto the lower
// one, so C is exactly symmetric.
template <int NP>
__global__ __launch_bounds__(NT) void crt_gemm_epilogue_kernel(GemmEpiArgs g) {
  __shared__ double red[NT / 64][4];
  const int ld = g.ld;
  int I, J;
  tile::sym_tile(blockIdx.x, I, J);
  const int bi = I >> 2, bj = J >> 2;
  const int t = bj * (bj + 1) / 2 + bi;
  const int wave = ((I & 3) >> 1) + 2 * (J & 3);
  const int tid = threadIdx.x, lane = tid & 63, b = tid >> 6;
  const int m = 4 * (I & 1) + b;
  const i4* src = (const i4*)(g.res + (size_t)t * NP * kItemBytes) + (wave * 8 + m) * 64 + lane;
  const int row = SB * I + 16 * b + (lane & 15);
  const int cl0 = 4 * (lane >> 4);
  i4 d[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) d[k] = src[k * (kItemBytes / 16)];
  double S[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) S[i] = 0.0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    const double inv = g.c.inv[k];
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        S[4 * n + r] = fma((double)(int8_t)(d[k][n] >> (8 * r)), inv, S[4 * n + r]);
  }
  const double ys = g.c.cscale * g.sa[0] * g.sb[0];
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;
#pragma unroll
  for (int n = 0; n < 4; ++n) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = SB * J + 16 * n + cl0 + r;
      if (I == J && row > col) continue;
      const bool mirror = (row != col);
      const double w = mirror ? 2.0 : 1.0;
      const double sr = S[4 * n + r];
      const size_t idx = row + (size_t)col * ld;
      double v = g.alpha * ((sr - rint(sr)) * ys);
      if (g.Cin) v += g.beta * g.Cin[idx];
      const bool diag = (row == col);
      v += (diag && row < g.n) ? g.gamma : 0.0;
      q0 += diag ? v : 0.0;
      q1 += w * v * v;
      if (g.E) q2 += w * v * g.E[idx];
      g.C[idx] = v;
      if (mirror) g.C[col + (size_t)row * ld] = v;
    }
  }
  double sums[3] = {q0, q1, q2};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    double x = sums[q];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    sums[q] = x;
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 3; ++q) red[b][q] = sums[q];
  }
  __syncthreads();
  if (g.part && tid < 4) {
    double a = 0.0;
    if (tid < 3) {
#pragma unroll
      for (int ww = 0; ww < NT / 64; ++ww) a += red[ww][tid];
    }
    g.part[4 * (size_t)blockIdx.x + tid] = a;
  }
}

int g_moduli = 11;

// Modular inverse of a mod p (p > 1, gcd(a, p) = 1).
long long inv_mod(long long a, long long p) {
  long long t = 0, nt = 1, r = p, nr = ((a % p) + p) % p;
  while (nr) {
    const long long qq = r / nr;
    long long tmp = t - qq * nt;
    t = nt;
    nt = tmp;
    tmp = r - qq * nr;
    r = nr;
    nr = tmp;
  }
  return t < 0 ? t + p : t;
}

// square: the constants of the SP2 square X^2 (spectrum of X in [0, 1]);
// otherwise those of a general product of row-norm-scaled operands.
CrtConst make_const(int np, bool square) {
  CrtConst c{};
  double log2m = 0.0;
  for (int k = 0; k < np; ++k) log2m += std::log2((double)kModuli[k]);
  for (int k = 0; k < np; ++k) {
    const long long p = kModuli[k];
    long long mk = 1;  // (M / p) mod p
    for (int j = 0; j < np; ++j)
      if (j != k) mk = (mk * (kModuli[j] % p)) % p;
    c.p[k] = (double)p;
    c.inv[k] = 1.0 / (double)p;
    c.q[k] = (double)inv_mod(mk, p);
  }
  // General product: |C_ij| <= ||A_i|| ||A_j|| <= 2^(2a) (1 + 1e-10) for
  // operand rows of 2-norm at most 2^a (the rounding of A adds at most
  // sqrt(ld) / 2 to a row norm); 2^(2a + 1.001) <= M keeps |C| < M / 2.
  // SP2 square: C - 2^a A = 2^(2a) (X~^2 - X~), and X~ = A 2^-a has its
  // spectrum in [0, 1] up to the rounding of A (the SP2 maps keep [0, 1]),
  // where x^2 - x lies in [-1/4, 0]; an entry of a symmetric matrix is at
  // most its spectral norm, so |C_ij - 2^a A_ij| <= 2^(2a) / 4 and
  // 2^(2a - 1 + 0.001) <= M keeps it below M / 2: one more bit of scale.
  // a <= 52 (X carries 53 bits).
  const double slack = square ? -1.001 : 1.001;
  const int a = std::min(52, (int)std::floor((log2m - slack) / 2.0));
  c.scale = std::ldexp(1.0, a);
  c.iscale = std::ldexp(1.0, -a);
  c.cscale = std::exp2(log2m - 2.0 * a);
  c.icscale = std::exp2(2.0 * a - log2m);
  c.shift = square ? 1 : 0;
  // A = rint(2^a X) itself from its first kRecon residues: |A| <= 2^a < M6 / 2
  // and the FP64 fraction sum is off by at most kRecon 2^-53, i.e. < 0.1 after
  // scaling by M6 = 2^47.8, so rint recovers A exactly for a <= 45
  long long m6 = 1;
  for (int k = 0; k < kRecon; ++k) m6 *= kModuli[k];
  for (int k = 0; k < kRecon; ++k) {
    const long long p = kModuli[k];
    c.q6[k] = (double)inv_mod((m6 / p) % p, p);
  }
  c.m6 = (double)m6;
  c.recon = (a <= 45) ? 1 : 0;
  return c;
}

inline int plane_ld(int ld) { return ((ld + TM - 1) / TM) * TM; }

// Product kernel configuration, coded W * 100 + NSTAGE: 8 waves with a 3-,
// 4- or 5-stage ring (96, 128, 160 KB of LDS); 904 is the persistent 8-wave
// kernel with a 4-stage ring, and 944 the same with the interleaved issue
// order of item_mainloop (IL).  (Other read / MFMA patterns measured within
// 1.5 % of 944 or slower, or spilled registers; the interleaved form with a
// 5-stage ring, 160 KB, was 6 % slower than 944 at N = 8192.)  (The
// 4-wave layouts, the 3-stage persistent form and the 5-stage one without
// the interleaved order compile with register spills and are not offered.)
// PRG_CRT_CONFIG forces one; otherwise the first product of each
// (pld, moduli) pair times every configuration (twice each, the minimum
// counts) and keeps the fastest.  The product kernel only writes its residue
// buffer, so the trial launches leave no trace and the results do not depend
// on the choice.
constexpr int kConfigs[] = {803, 804, 805, 904, 944};
constexpr int kNumConfigs = sizeof(kConfigs) / sizeof(kConfigs[0]);
int g_forced = -1;                        // -1: not read yet; 0: autotune
std::vector<std::array<int, 3>> g_tuned;  // (pld, np, config)
int g_last_config = 0;

int forced_config() {
  if (g_forced < 0) {
    const char* e = std::getenv("PRG_CRT_CONFIG");
    const int v = e ? std::atoi(e) : 0;
    g_forced = 0;
    for (int c : kConfigs)
      if (c == v) g_forced = v;
  }
  return g_forced;
}

// Workgroups of the persistent kernel: one per CU (its ring takes most of
// the LDS), a multiple of 8 so that every XCD gets the same count.
int persistent_grid(int nitems) {
  static int cus = 0;
  if (!cus) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n < 8)
      n = 256;
    cus = n;
  }
  const int need = ((nitems + 7) / 8) * 8;
  return std::max(8, std::min((cus / 8) * 8, need));
}

void launch_product(const ProductArgs& g, int nitems, int config) {
  switch (config) {
    case 904:
      hipLaunchKernelGGL((crt_product_persistent_kernel<8, 4>), dim3(persistent_grid(nitems)),
                         dim3(512), 0, stream(), g);
      break;
    case 944:
      hipLaunchKernelGGL((crt_product_persistent_kernel<8, 4, true>),
                         dim3(persistent_grid(nitems)), dim3(512), 0, stream(), g);
      break;
    case 803:
      hipLaunchKernelGGL((crt_product_kernel<8, 3>), dim3(nitems), dim3(512), 0, stream(), g);
      break;
    case 805:
      hipLaunchKernelGGL((crt_product_kernel<8, 5>), dim3(nitems), dim3(512), 0, stream(), g);
      break;
    default:
      hipLaunchKernelGGL((crt_product_kernel<8, 4>), dim3(nitems), dim3(512), 0, stream(), g);
      break;
  }
}

int run_product(const ProductArgs& g, int nitems, int* used) {
  int config = forced_config();
  if (!config)
    for (const auto& t : g_tuned)
      if (t[0] == g.pld && t[1] == g.np) config = t[2];
  if (config) {
    launch_product(g, nitems, config);
    *used = config;
    return OK;
  }
  hipEvent_t ev[2];
  PRG_HIP_CHECK(hipEventCreate(&ev[0]));
  PRG_HIP_CHECK(hipEventCreate(&ev[1]));
  float best[kNumConfigs];
  for (int c = 0; c < kNumConfigs; ++c) best[c] = 1e30f;
  for (int rep = 0; rep < 2; ++rep)
    for (int c = 0; c < kNumConfigs; ++c) {
      PRG_HIP_CHECK(hipEventRecord(ev[0], stream()));
      launch_product(g, nitems, kConfigs[c]);
      PRG_HIP_CHECK(hipEventRecord(ev[1], stream()));
      PRG_HIP_CHECK(hipEventSynchronize(ev[1]));
      float ms = 0.f;
      PRG_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
      best[c] = std::min(best[c], ms);
    }
  PRG_HIP_CHECK(hipEventDestroy(ev[0]));
  PRG_HIP_CHECK(hipEventDestroy(ev[1]));
  int bi = 0;
  for (int c = 1; c < kNumConfigs; ++c)
    if (best[c] < best[bi]) bi = c;
  g_tuned.push_back({g.pld, g.np, kConfigs[bi]});
  *used = kConfigs[bi];  // the trials computed the residues already
  return OK;
}

#define PRG_CRT_DISPATCH(NPV, CALL) \
  switch (NPV) {                    \
    case 6: CALL(6); break;         \
    case 7: CALL(7); break;         \
    case 8: CALL(8); break;         \
    case 9: CALL(9); break;         \
    case 10: CALL(10); break;       \
    case 11: CALL(11); break;       \
    case 12: CALL(12); break;       \
    case 13: CALL(13); break;       \
    case 14: CALL(14); break;       \
    default: break;                 \
  }

template <int NP>
void launch_update(const UpdateArgs& g, int nblocks) {
  if (g.c.recon)
    hipLaunchKernelGGL((crt_update_kernel<NP, true>), dim3(nblocks), dim3(NT), 0, stream(), g);
  else
    hipLaunchKernelGGL((crt_update_kernel<NP, false>), dim3(nblocks), dim3(NT), 0, stream(), g);
}

template <int NP>
void launch_split_scaled(const double* x, int8_t* p, int ld, int pld, const CrtConst& c,
                         const double* sc) {
  hipLaunchKernelGGL(crt_split_scaled_kernel<NP>, dim3(2048), dim3(NT), 0, stream(), x, p, ld,
                     pld, c, sc);
}

template <int NP>
void launch_gemm_epilogue(const GemmEpiArgs& g, int nblocks) {
  hipLaunchKernelGGL(crt_gemm_epilogue_kernel<NP>, dim3(nblocks), dim3(NT), 0, stream(), g);
}

template <int NP>
void launch_split(const double* x, int8_t* p0, int8_t* p1, int ld, int pld, const CrtConst& c) {
  hipLaunchKernelGGL(crt_split_kernel<NP>, dim3(2048), dim3(NT), 0, stream(), x, p0, p1, ld, pld,
                     c);
}

// Product items: the upper-triangle tiles in groups of G x G tiles (group
// rows SI <= group columns SJ, tiles column by column inside a group), split
// into 8 contiguous parts, one per XCD; each part lists its tiles once per
// modulus.  The 32 workgroups resident on an XCD then work on one modulus and
// a compact block of tiles (8 A panels and 4 B panels shared through that
// XCD's L2), and all XCDs move through the moduli together, so one modulus's
// planes (pld^2 bytes) are what the shared last-level cache holds.
// PRG_CRT_GROUP=2..32 sets G (default 8).
int group_size() {
  static int g = -1;
  if (g < 0) {
    const char* e = std::getenv("PRG_CRT_GROUP");
    const int v = e ? std::atoi(e) : 0;
    g = (v >= 2 && v <= 32) ? v : 8;
  }
  return g;
}

const int4* item_list(int nt, int np, int* st) {
  // the upload is kept while the workspace buffer stays the same one
  static int cached_nt = 0, cached_np = 0;
  static unsigned cached_gen = 0;
  const int nb = nt * (nt + 1) / 2;
  int4* d = (int4*)workspace("crt_items", sizeof(int4) * (size_t)nb * np, st);
  const unsigned gen = workspace_generation();
  if (*st || (cached_nt == nt && cached_np == np && cached_gen == gen)) return d;
  const int G = group_size();
  std::vector<int2> tiles;
  tiles.reserve(nb);
  const int ng = (nt + G - 1) / G;
  for (int SJ = 0; SJ < ng; ++SJ)
    for (int SI = 0; SI <= SJ; ++SI)
      for (int J = SJ * G; J < std::min(nt, SJ * G + G); ++J)
        for (int I = SI * G; I < std::min(nt, SI * G + G); ++I)
          if (I <= J) tiles.push_back(int2{I, J});
  std::vector<int4> h;
  h.reserve((size_t)nb * np);
  for (int x = 0; x < 8; ++x) {
    const int lo = (int)((long long)nb * x / 8), hi = (int)((long long)nb * (x + 1) / 8);
    for (int k = 0; k < np; ++k)
      for (int i = lo; i < hi; ++i)
        h.push_back(int4{k, tiles[i].x, tiles[i].y, tiles[i].y * (tiles[i].y + 1) / 2 + tiles[i].x});
  }
  // on the library stream, after the workspace's zero fill, and complete
  // before the host vector goes away
  if ((int)tiles.size() != nb ||
      hipMemcpyAsync(d, h.data(), sizeof(int4) * h.size(), hipMemcpyHostToDevice, stream()) !=
          hipSuccess ||
      hipStreamSynchronize(stream()) != hipSuccess) {
    set_error(ERR_HIP, "int8 CRT products: item list upload failed");
    *st = ERR_HIP;
    return nullptr;
  }
  cached_nt = nt;
  cached_np = np;
  cached_gen = gen;
  return d;
}

// Device copy of the per-modulus constants the product kernel indexes at run
// time (refreshed when the number of moduli changes).
double* fold_constants(int np, int* st) {
  static int cached = 0;
  static unsigned cached_gen = 0;
  double* d = (double*)workspace("crt_fold", sizeof(double) * 4 * kMaxModuli, st);
  const unsigned gen = workspace_generation();
  if (*st || (cached == np && cached_gen == gen)) return d;
  const CrtConst c = make_const(np, false);  // p, 1/p and q only
  double h[4 * kMaxModuli] = {};
  for (int k = 0; k < np; ++k) {
    h[4 * k] = c.p[k];
    h[4 * k + 1] = c.inv[k];
    h[4 * k + 2] = c.q[k];
    h[4 * k + 3] = c.q[k] * c.inv[k];
  }
  if (hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, stream()) != hipSuccess ||
      hipStreamSynchronize(stream()) != hipSuccess) {
    set_error(ERR_HIP, "int8 CRT products: constant upload failed");
    *st = ERR_HIP;
    return nullptr;
  }
  cached = np;
  cached_gen = gen;
  return d;
}

// Optional HIP-event timing of the product kernel alone (one event pair per
// step of the current solve; the caller names how many steps computed).
bool g_timing = false;
std::vector<hipEvent_t> g_ev;
int g_ev_used = 0;
double g_ms_sum = 0.0;
int g_ms_count = 0;

}  // namespace

void crt_timing(bool enable) {
  g_timing = enable;
  g_ev_used = 0;
  g_ms_sum = 0.0;
  g_ms_count = 0;
}

void crt_timing_begin() { g_ev_used = 0; }

int crt_timing_collect(int executed) {
  if (!g_timing) return OK;
  const int n = std::min(executed, g_ev_used);
  for (int e = 0; e < n; ++e) {
    float ms = 0.f;
    PRG_HIP_CHECK(hipEventElapsedTime(&ms, g_ev[2 * e], g_ev[2 * e + 1]));
    g_ms_sum += ms;
    ++g_ms_count;
  }
  g_ev_used = 0;
  return OK;
}

void crt_timing_get(double* ms_sum, int* count) {
  if (ms_sum) *ms_sum = g_ms_sum;
  if (count) *count = g_ms_count;
}

int crt_moduli() { return g_moduli; }

int set_crt_moduli(int np) {
  if (np < 6 || np > kMaxModuli) {
    set_error(ERR_ARG, "int8 CRT products: the number of moduli must be in [6, 14]");
    return ERR_ARG;
  }
  g_moduli = np;
  return OK;
}

int crt_update_blocks(int ld) {
  const int nb = ld / SB;
  return nb * (nb + 1) / 2;
}

int crt_prepare(const DMat& x0, int np) {
  int st = OK;
  const int ld = x0.ld, pld = plane_ld(ld);
  const size_t plane = (size_t)pld * pld;
  int8_t* planes = (int8_t*)workspace("crt_planes", 2 * (size_t)np * plane, &st);
  if (st) return st;
  if (!fold_constants(np, &st)) return st;
  const CrtConst c = make_const(np, true);
#define PRG_CALL(K) launch_split<K>(x0.p, planes, planes + (size_t)np * plane, ld, pld, c)
  PRG_CRT_DISPATCH(np, PRG_CALL)
#undef PRG_CALL
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int crt_step(const DMat& x0, const DMat& x1, Sp2State* d_state, double* d_part, int* ntiles,
             int np) {
  const int ld = x0.ld, pld = plane_ld(ld);
  if (ld % SB != 0 || pld >= (1 << 17)) {
    set_error(ERR_ARG, "int8 CRT products: matrix too large for int32 accumulation");
    return ERR_ARG;
  }
  const size_t plane = (size_t)pld * pld;
  const int nt = pld / TM;
  const int nb = nt * (nt + 1) / 2;
  int st = OK;
  int8_t* planes = (int8_t*)workspace("crt_planes", 2 * (size_t)np * plane, &st);
  int8_t* res = (int8_t*)workspace("crt_residues", (size_t)nb * np * kItemBytes, &st);
  if (st) return st;
  ProductArgs pa;
  pa.planes[0] = planes;
  pa.planes[1] = planes + (size_t)np * plane;
  pa.planes_b = nullptr;
  pa.res = res;
  pa.st = d_state;
  pa.pld = pld;
  pa.np = np;
  pa.nitems = nb * np;
  pa.fold = fold_constants(np, &st);
  if (st) return st;
  pa.items = item_list(nt, np, &st);
  if (st) return st;
  UpdateArgs ua;
  ua.buf[0] = x0.p;
  ua.buf[1] = x1.p;
  ua.planes[0] = planes;
  ua.planes[1] = planes + (size_t)np * plane;
  ua.res = res;
  ua.st = d_state;
  ua.part = d_part;
  ua.ld = ld;
  ua.pld = pld;
  ua.c = make_const(np, true);
  const int nblocks = crt_update_blocks(ld);
  if (ntiles) *ntiles = nblocks;
  const bool timed = g_timing;
  if (timed) {
    if ((int)g_ev.size() < 2 * (g_ev_used + 1)) {
      const size_t old = g_ev.size();
      g_ev.resize(2 * (g_ev_used + 1) + 64);
      for (size_t e = old; e < g_ev.size(); ++e) PRG_HIP_CHECK(hipEventCreate(&g_ev[e]));
    }
    PRG_HIP_CHECK(hipEventRecord(g_ev[2 * g_ev_used], stream()));
  }
  if ((st = run_product(pa, nb * np, &g_last_config))) return st;
  PRG_HIP_CHECK(hipGetLastError());
  if (timed) {
    PRG_HIP_CHECK(hipEventRecord(g_ev[2 * g_ev_used + 1], stream()));
    ++g_ev_used;
  }
#define PRG_CALL(K) launch_update<K>(ua, nblocks)
  PRG_CRT_DISPATCH(np, PRG_CALL)
#undef PRG_CALL
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int crt_finish(const DMat& x0, const DMat& x1, const Sp2State* d_state, double scale, DMat out,
               int np) {
  const CrtConst c = make_const(np, true);
  if (!c.recon) return sp2_finish(x0, x1, d_state, scale, out, true);
  const int ld = x0.ld, pld = plane_ld(ld);
  const size_t plane = (size_t)pld * pld;
  int st = OK;
  int8_t* planes = (int8_t*)workspace("crt_planes", 2 * (size_t)np * plane, &st);
  if (st) return st;
  hipLaunchKernelGGL(crt_finish_kernel, dim3(2048), dim3(NT), 0, stream(), planes,
                     planes + (size_t)np * plane, d_state, pld, c, scale, out.p, out.ld, out.n);
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

int crt_last_waves() { return g_last_config; }

int set_crt_config(int config) {
  forced_config();  // the environment is read first, so that this call wins
  if (config == 0) {
    g_forced = 0;
    return OK;
  }
  for (int c : kConfigs)
    if (c == config) {
      g_forced = config;
      return OK;
    }
  set_error(ERR_ARG, "int8 CRT products: unknown kernel configuration");
  return ERR_ARG;
}

// One operand that a caller promises not to change while held (the matrix of
// a Newton-Schulz or CG solve): its residue planes and scale are formed once
// and reused by every product that names it until the hold ends.
namespace {
struct HeldOperand {
  const double* key = nullptr;
  int depth = 0;
  bool valid = false;
  int np = 0, ld = 0;
};
HeldOperand g_held;
}  // namespace

void crt_hold_operand(const double* p) {
  if (g_held.depth == 0) {
    g_held.key = p;
    g_held.valid = false;
  }
  if (g_held.key == p) ++g_held.depth;
}

void crt_release_operand(const double* p) {
  if (g_held.depth > 0 && g_held.key == p && --g_held.depth == 0) {
    g_held.key = nullptr;
    g_held.valid = false;
  }
}

int crt_gemm(const DMat& A, const DMat& B, DMat C, double alpha, double beta, const DMat* Cin,
             double gamma, double* part, int* nparts, const DMat* dotE, int np) {
  const int ld = C.ld, pld = plane_ld(ld);
  if (A.ld != ld || B.ld != ld || (Cin && Cin->ld != ld) || (dotE && dotE->ld != ld)) {
    set_error(ERR_ARG, "int8 CRT products: operand leading dimensions differ");
    return ERR_ARG;
  }
  if (ld % SB != 0 || pld >= (1 << 17)) {
    set_error(ERR_ARG, "int8 CRT products: matrix too large for int32 accumulation");
    return ERR_ARG;
  }
  if (np < 6 || np > kMaxModuli) {
    set_error(ERR_ARG, "int8 CRT products: the number of moduli must be in [6, 14]");
    return ERR_ARG;
  }
  const size_t plane = (size_t)pld * pld;
  const int nt = pld / TM;
  const int nb = nt * (nt + 1) / 2;
  const bool same = (A.p == B.p);
  const bool held_a = g_held.depth > 0 && A.p == g_held.key;
  const bool held_b = g_held.depth > 0 && B.p == g_held.key && !same;
  int st = OK;
  int8_t* pa_planes = (int8_t*)workspace(held_a ? "crtg_planes_held" : "crtg_planes_a",
                                         (size_t)np * plane, &st);
  int8_t* pb_planes = same ? nullptr
                           : (int8_t*)workspace(held_b ? "crtg_planes_held" : "crtg_planes_b",
                                                (size_t)np * plane, &st);
  int8_t* res = (int8_t*)workspace("crt_residues", (size_t)nb * np * kItemBytes, &st);
  double* sc = (double*)workspace("crtg_scales", sizeof(double) * 6, &st);
  double* npart = (double*)workspace("crtg_norm_part", sizeof(double) * 1024, &st);
  if (st) return st;
  if ((held_a || held_b) && (g_held.np != np || g_held.ld != ld)) g_held.valid = false;
  const double* fold = fold_constants(np, &st);
  if (st) return st;
  const int4* items = item_list(nt, np, &st);
  if (st) return st;
  const CrtConst c = make_const(np, false);
  const int nnb = std::max(1, std::min(1024, (C.n + NT / 64 - 1) / (NT / 64)));
  auto scale_split = [&](const DMat& X, int8_t* planes, double* s) -> int {
    hipLaunchKernelGGL(crt_colnorm_kernel, dim3(nnb), dim3(NT), 0, stream(), X.p, ld, X.n, npart);
    hipLaunchKernelGGL(crt_scale_kernel, dim3(1), dim3(NT), 0, stream(), npart, nnb, s);
#define PRG_CALL(K) launch_split_scaled<K>(X.p, planes, ld, pld, c, s)
    PRG_CRT_DISPATCH(np, PRG_CALL)
#undef PRG_CALL
    PRG_HIP_CHECK(hipGetLastError());
    return OK;
  };
  // scale slots: A at sc, B at sc + 2, the held operand at sc + 4
  double* sa = held_a ? sc + 4 : sc;
  double* sb = same ? sa : (held_b ? sc + 4 : sc + 2);
  if (!(held_a && g_held.valid) && (st = scale_split(A, pa_planes, sa))) return st;
  if (!same && !(held_b && g_held.valid) && (st = scale_split(B, pb_planes, sb))) return st;
  if (held_a || held_b) {
    g_held.valid = true;
    g_held.np = np;
    g_held.ld = ld;
  }
  ProductArgs pa;
  pa.planes[0] = pa_planes;
  pa.planes[1] = pa_planes;
  pa.planes_b = pb_planes;
  pa.res = res;
  pa.items = items;
  pa.st = nullptr;
  pa.fold = fold;
  pa.pld = pld;
  pa.np = np;
  pa.nitems = nb * np;
  if ((st = run_product(pa, nb * np, &g_last_config))) return st;
  PRG_HIP_CHECK(hipGetLastError());
  GemmEpiArgs ea;
  ea.res = res;
  ea.sa = sa;
  ea.sb = sb;
  ea.Cin = (Cin && beta != 0.0) ? Cin->p : nullptr;
  ea.E = dotE ? dotE->p : nullptr;
  ea.C = C.p;
  ea.part = part;
  ea.alpha = alpha;
  ea.beta = beta;
  ea.gamma = gamma;
  ea.ld = ld;
  ea.n = C.n;
  ea.c = c;
  const int nblocks = crt_update_blocks(ld);
  if (nparts) *nparts = nblocks;
#define PRG_CALL(K) launch_gemm_epilogue<K>(ea, nblocks)
  PRG_CRT_DISPATCH(np, PRG_CALL)
#undef PRG_CALL
  PRG_HIP_CHECK(hipGetLastError());
  return OK;
}

}  // namespace prg
