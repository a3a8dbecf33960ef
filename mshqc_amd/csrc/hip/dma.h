// Direct global -> LDS loads (gfx950 global_load_lds_dwordx4) and the stage
// barrier of the DMA pipelines built on them.
#pragma once
#include <hip/hip_runtime.h>

// Stage barrier of the DMA pipelines: wait until at most N of this wave's
// vector-memory operations are outstanding (they retire in issue order, so
// every older LDS DMA load has landed) and all of its LDS accesses have
// returned, then a workgroup barrier.  The empty asm statements keep the
// compiler from moving LDS or global accesses across it.  s_waitcnt
// immediate: vmcnt in bits [3:0] and [15:14], expcnt [6:4] (left at its
// maximum), lgkmcnt [11:8] = 0.
template <int N>
__device__ __forceinline__ void dma_barrier() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | 0x0070);
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Direct global -> LDS load of 1 KiB by one wave, 16 bytes per lane: the
// instruction writes lane l's bytes at dst + 16*l.  `dst` must be wave-uniform.
__device__ __forceinline__ void dma16(const void* src, char* dst) {
    __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
}
