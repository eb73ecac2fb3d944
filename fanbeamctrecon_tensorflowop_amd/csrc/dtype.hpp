// dtype.hpp -- storage types of images and sinograms in device memory.
//
// Arithmetic is always FP32 (sums in FP32/FP64); BF16 is a storage format only:
// widened exactly on load, rounded to nearest even on store.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fbct {

using bf16_t = uint16_t;  // raw bfloat16 bits

__device__ __forceinline__ float load_f(const float *p, size_t i)
{
    return p[i];
}

__device__ __forceinline__ float load_f(const bf16_t *p, size_t i)
{
    return __uint_as_float(static_cast<unsigned>(p[i]) << 16);
}
// Element at a 32-bit byte offset from a base held in scalar registers: the address is
// formed by the load itself (base + zero-extended offset), no 64-bit add per element.
__device__ __forceinline__ float load_f_at(const float *p, unsigned byte_off)
{
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(p) + byte_off);
}
__device__ __forceinline__ float load_f_at(const bf16_t *p, unsigned byte_off)
{
    const bf16_t b = *reinterpret_cast<const bf16_t *>(reinterpret_cast<const char *>(p) + byte_off);
    return __uint_as_float(static_cast<unsigned>(b) << 16);
}

// A loaded storage element widened to FP32 (exact).  Kept apart from the load where a
// prefetched value waits in registers: widening at the load would wait for it there.
__device__ __forceinline__ float widen(float v)
{
    return v;
}
__device__ __forceinline__ float widen(bf16_t v)
{
    return __uint_as_float(static_cast<unsigned>(v) << 16);
}

// *(base + byte_off): a 32-bit offset from the base (a scalar base takes it as the lane
// offset of the load)
template <typename T>
__device__ __forceinline__ T load_at(const T *base, unsigned byte_off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

// p, known to be the same in every lane and to point to global memory: a scalar base
// the loads through it can take as their address (32-bit lane offsets, global loads).
template <typename T>
__device__ __forceinline__ const T *uniform_global(const T *p)
{
    const unsigned long long a = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
    const auto *g = reinterpret_cast<const __attribute__((address_space(1))) T *>(((unsigned long long)hi << 32) | lo);
    return (const T *)g;
}

__device__ __forceinline__ bf16_t f32_to_bf16(float v)
{
    const unsigned u = __float_as_uint(v);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))  // NaN stays a quiet NaN
        return static_cast<bf16_t>((u >> 16) | 0x0040u);
    return static_cast<bf16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ void store_f(float *p, size_t i, float v)
{
    p[i] = v;
}

__device__ __forceinline__ void store_f(bf16_t *p, size_t i, float v)
{
    p[i] = f32_to_bf16(v);
}

}  // namespace fbct
