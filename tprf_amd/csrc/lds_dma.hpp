// lds_dma.hpp -- gfx950 global -> LDS loads and the waits around them
// (device code only: included by the .hip files, not by the host library).
#pragma once

#include "common.hpp"

namespace tprf {

// Direct global -> LDS load (global_load_lds_dwordx4): lane l of a wave writes
// the 16 bytes at its own `src` to `dst` + 16 l; `dst` must be the same for the
// whole wave.  Completion is counted by vmcnt like any vector memory load.
__device__ __forceinline__ void dma16(const c2 *src, c2 *dst) {
  __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)dst, 16, 0, 0);
}
// s_waitcnt vmcnt(K): vmcnt in bits 3:0 and 15:14, expcnt (6:4) and lgkmcnt (11:8) at their maxima.
template <int K>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(K >= 0 && K < 64, "vmcnt holds 6 bits");
  __builtin_amdgcn_s_waitcnt((K & 15) | ((K >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Workgroup barrier that leaves vector memory loads in flight: waits for this
// wave's LDS reads and writes (lgkmcnt), then s_barrier.  __syncthreads() would
// also wait for every in-flight global -> LDS load (its release fence covers
// them), which would stall the prefetch two steps ahead.  The "memory" clobber
// keeps the compiler from moving LDS accesses across it.
__device__ __forceinline__ void barrier_keep_vm() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

}  // namespace tprf
