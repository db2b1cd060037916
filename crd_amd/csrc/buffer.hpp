// Raw buffer addressing for the gfx950 kernels: the base address sits in a 128-bit resource in scalar registers and
// every access takes a 32-bit byte offset per lane plus a scalar offset, so the accesses of a thread share one
// address VGPR and cost no vector address arithmetic. Accesses at or past num_records bytes from the base are not
// performed (loads return 0). Offsets must stay below 4 GiB from the base.
#pragma once

#include "dft.hpp"

namespace crdfft::dev
{

typedef float f2v __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t num_records = 0xFFFFFFFFu)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(num_records), 0x00020000);
}

// The builtins' own 8-byte vector type (a generic vector, converted as a whole: its elements are not accessed).
using b64_t = decltype(__builtin_amdgcn_raw_buffer_load_b64(make_rsrc(nullptr), 0, 0, 0));
static_assert(sizeof(b64_t) == 8, "64-bit buffer access");
constexpr int kAuxNt = 2; // cache policy of the buffer instructions: non-temporal

template <int AUX> __device__ __forceinline__ cf bld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    const b64_t u = __builtin_amdgcn_raw_buffer_load_b64(r, static_cast<int>(voff), static_cast<int>(soff), AUX);
    const f2v f = __builtin_bit_cast(f2v, u);
    return cf{f.x, f.y};
}

template <int AUX> __device__ __forceinline__ float bld32(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    const auto u = __builtin_amdgcn_raw_buffer_load_b32(r, static_cast<int>(voff), static_cast<int>(soff), AUX);
    static_assert(sizeof(u) == 4, "32-bit buffer access");
    return __builtin_bit_cast(float, u);
}

typedef float f4v __attribute__((ext_vector_type(4)));

template <int AUX> __device__ __forceinline__ f4v bld128(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    const auto u = __builtin_amdgcn_raw_buffer_load_b128(r, static_cast<int>(voff), static_cast<int>(soff), AUX);
    static_assert(sizeof(u) == 16, "128-bit buffer access");
    return __builtin_bit_cast(f4v, u);
}

template <int AUX> __device__ __forceinline__ void bst32(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, static_cast<int>(voff),
                                          static_cast<int>(soff), AUX);
}

template <int AUX> __device__ __forceinline__ void bst128(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, f4v v)
{
    using u4 = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u4, v), r, static_cast<int>(voff), static_cast<int>(soff),
                                           AUX);
}

template <int AUX> __device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff, cf v)
{
    f2v f;
    f.x = v.x;
    f.y = v.y;
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(b64_t, f), r, static_cast<int>(voff),
                                          static_cast<int>(soff), AUX);
}

} // namespace crdfft::dev
