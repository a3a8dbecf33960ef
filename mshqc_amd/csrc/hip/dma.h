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

// The same 1 KiB load as a buffer instruction (buffer_load_dwordx4 ... lds):
// lane l reads 16 bytes at base + voff (voff per lane, base wave-uniform and
// held in the scalar resource) and writes them at dst + 16*l.  Unlike the
// global form, which is a FLAT-encoded instruction that the compiler's wait
// analysis counts against both vmcnt and lgkmcnt, a buffer load is counted by
// vmcnt alone, so the LDS fragment reads after it keep partial lgkmcnt waits
// instead of lgkmcnt(0).  The resource covers 4 GiB from base (raw buffer,
// stride 0, 32-bit data format in word 3 as gfx9 expects); voff + soff stays
// below.
// soff is a wave-uniform byte offset added to every lane's (an SGPR operand).
__device__ __forceinline__ void dma16b(const void* base, unsigned voff, unsigned soff, char* dst) {
    __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, -1, 0x00020000);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(rs, (__attribute__((address_space(3))) void*)dst, 16, voff, soff, 0, 0);
}
