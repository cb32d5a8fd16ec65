// Direct global -> LDS copies (global_load_lds_dwordx4) and the explicit
// waits that go with them, shared by the kernels that stage operands through
// an LDS ring.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

// dma16 sets M0 itself in inline assembly; LLVM treats M0 as reserved and
// does not honour a clobber of it, so a kernel that uses dma16 must not have
// the compiler keep a value in M0 across it (the including files check their
// generated assembly), and the reserved-register clobber warning is silenced.
#pragma clang diagnostic ignored "-Winline-asm"

namespace prg {
namespace lds_dma {

// LDS byte address of a __shared__ object.
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// 16 bytes per lane from global memory at gbase + voff straight into LDS at
// lds_base + 16 * lane (global_load_lds_dwordx4 with a scalar base: gbase
// and lds_base must be wave-uniform, voff is the lane's 32-bit offset).
// Written as inline assembly so that the compiler does not add its
// conservative vmcnt(0) waits before every LDS read: the waits are explicit
// (wait_vm_barrier).
__device__ __forceinline__ void dma16(uint64_t gbase, uint32_t voff, uint32_t lds_base) {
  asm volatile("s_mov_b32 m0, %2\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %1"
               :
               : "v"(voff), "s"(gbase), "s"(lds_base)
               : "memory", "m0");
}

// A 64-bit value that is uniform across the wave, moved to scalar registers.
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// Waits until at most N vector-memory operations of this wave are pending and
// its LDS reads have completed, then joins the workgroup barrier.
template <int N>
__device__ __forceinline__ void wait_vm_barrier() {
  asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" : : "n"(N) : "memory");
}

}  // namespace lds_dma
}  // namespace prg
