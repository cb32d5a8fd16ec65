// Device-side FP64 bookkeeping of the SP2 loop: the branch decision for the
// next step and the absorption of a step's partial sums, for the one-block
// kernels that run them.  (A last-workgroup absorb inside the split epilogue
// was slower: the device-scope fences it needs write back and invalidate the
// per-XCD L2s, 10.8 us per step at N = 1024 against 4.6 us for this launch.)
#pragma once

#include "prg_internal.h"

#include <cmath>

namespace prg {

// Decide the next SP2 step from the current X's traces (runs in thread 0).
// The branch rules restate src/prg_sp2_mod.F90: alg2 :319-363, genseq
// :469-513, alg1 :801-849, alg1_genseq :941-982, basic :89-119, replay
// variants :581-600 / :1060-1072 / :1288-1300, and the response loop of
// src/prg_response_mod.F90:730-770.
__device__ inline void sp2_plan(Sp2State* s) {
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
// pass over the records, wave shuffles, then the BT / 64 wave totals (BT threads).
template <int BT>
__device__ inline void block_sum4(const double* part, int nb, double* out /*shared*/) {
  __shared__ double red[BT / 64][4];
  double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 4
  for (int b = threadIdx.x; b < nb; b += BT) {
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
    for (int ww = 0; ww < BT / 64; ++ww) t += red[ww][threadIdx.x];
    out[threadIdx.x] = t;
  }
  __syncthreads();
}

// Absorb the partials of the product just computed (ntiles records of 4),
// then plan the next step.  Called by every thread of one BT-thread block of
// a launch whose step was planned.
template <int BT>
__device__ inline void sp2_absorb_plan_block(Sp2State* s, const double* part, int ntiles) {
  __shared__ double sums[4];
  block_sum4<BT>(part, ntiles, sums);
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

}  // namespace prg
