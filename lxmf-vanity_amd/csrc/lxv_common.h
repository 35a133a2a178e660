// Shared definitions of the device code. The field and hash code is written
// once as __host__ __device__ functions: on the GPU it uses the gfx950
// builtins, on the host the same arithmetic in portable C++ so the library can
// check the device algorithms on the CPU (lxv_debug_* entry points).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LXV_HD __host__ __device__ __forceinline__
#define LXV_D __device__ __forceinline__
#else
#define LXV_HD inline
#define LXV_D inline
#endif

namespace lxv {

// Rotate right: v_alignbit_b32 on gfx950.
LXV_HD uint32_t rotr32(uint32_t x, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(x, x, n);
#else
    return (x >> n) | (x << ((32 - n) & 31));
#endif
}

// (hi:lo) >> n, the low 32 bits: v_alignbit_b32.
LXV_HD uint32_t funnel32(uint32_t hi, uint32_t lo, uint32_t n) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, n);
#else
    return n ? (lo >> n) | (hi << (32 - n)) : lo;
#endif
}

// Byte swap: v_perm_b32 with selector 0x00010203.
LXV_HD uint32_t bswap32(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(0u, x, 0x00010203u);
#else
    return (x >> 24) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | (x << 24);
#endif
}

}  // namespace lxv
