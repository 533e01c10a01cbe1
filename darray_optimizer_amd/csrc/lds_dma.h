// Direct global -> LDS copies (global_load_lds) for the kernels that stage
// operand chunks without passing them through registers.
#pragma once

#include <cstdint>

// 16-byte (or 4-byte) direct global -> LDS copy by one wave from the uniform
// address gbase (scalar registers) plus lane L's byte offset voff; lane L's
// value lands at lds + L * size, lds the same for all lanes of the wave.
// Written as inline assembly (M0 = LDS base, one wait state, then the load)
// so that the compiler does not make every later LDS read wait for the copy:
// the copies of a step go to a buffer nobody reads in that step, and
// wait_vm() + a workgroup barrier publish them.  The scalar
// base leaves no per-lane address arithmetic to the vector ALU.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ const void* uniform_ptr(const void* p) {
    const uintptr_t v = (uintptr_t)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (const void*)(((uintptr_t)hi << 32) | lo);
}
__device__ __forceinline__ void dma16(const void* gbase, unsigned voff, void* lds) {
    unsigned base = (unsigned)(uintptr_t)(__attribute__((address_space(3))) void*)lds;
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, %2"
                 :
                 : "v"(voff), "s"(__builtin_amdgcn_readfirstlane(base)), "s"(uniform_ptr(gbase))
                 : "memory", "m0");
}
__device__ __forceinline__ void dma4(const void* gbase, unsigned voff, void* lds) {
    unsigned base = (unsigned)(uintptr_t)(__attribute__((address_space(3))) void*)lds;
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dword %0, %2"
                 :
                 : "v"(voff), "s"(__builtin_amdgcn_readfirstlane(base)), "s"(uniform_ptr(gbase))
                 : "memory", "m0");
}
#pragma clang diagnostic pop
// Wait for this wave's outstanding global (and global -> LDS) loads only.
__device__ __forceinline__ void wait_vm() { __builtin_amdgcn_s_waitcnt(0x0F70); }
