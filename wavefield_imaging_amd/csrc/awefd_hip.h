/*
 * Device helpers shared by the gfx950 kernels of libawefd: XCD-aware block
 * order, buffer-descriptor loads and stores, and DPP lane shifts.
 */
#ifndef AWEFD_HIP_H
#define AWEFD_HIP_H

#include <hip/hip_runtime.h>

namespace awefd {

constexpr int WAVE = 64;

/* Blocks are dispatched round-robin over the 8 XCDs (block b -> XCD b%8).
 * Give each XCD a contiguous range of tiles so that tiles sharing halos
 * share an L2. */
__device__ __forceinline__ int xcd_remap(int b, int n)
{
    const int per = n >> 3, rem = n & 7, xcd = b & 7, slot = b >> 3;
    return xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
}

typedef __amdgpu_buffer_rsrc_t rsrc_t;
/* an offset beyond every descriptor's range: loads return 0, stores drop */
constexpr unsigned OOB = 0x80000000u;

__device__ __forceinline__ rsrc_t mk_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}
/* voffset: per-lane bytes; soffset: wave-uniform bytes (both range-checked).
 * POL is the cache policy (2: non-temporal). */
template <int POL = 0>
__device__ __forceinline__ float bld(rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, POL));
}
template <int POL = 0>
__device__ __forceinline__ void bst(rsrc_t r, unsigned voff, unsigned soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, POL);
}
/* DPP wave shifts: lane i receives lane i-1 (shr) or lane i+1 (shl); the
 * lane without a source (0 for shr, 63 for shl) keeps `old`. */
__device__ __forceinline__ float shr1(float old, float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float shl1(float old, float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                                 __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, false));
}

} // namespace awefd

#endif
