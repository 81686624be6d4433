// buffer_chunk.hpp -- buffer access to a workgroup's z-chunk of one array
// (device helpers of the multi-step adjoint kernels).
#ifndef SCAFES_BUFFER_CHUNK_HPP_
#define SCAFES_BUFFER_CHUNK_HPP_

#include <hip/hip_runtime.h>

#include <cstdint>

namespace scafes {
namespace chunk {

using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 family: 32-bit data format, raw (stride 0) buffer

// Buffer access to a workgroup's z-chunk of one array: the resource is
// built once per workgroup (base = the lowest plane the workgroup touches,
// kRecords bytes); the plane enters each access as the instruction's
// scalar offset (one scalar multiply per plane and plane stride instead of a
// 64-bit base and a resource per array and plane).
constexpr uint32_t kRecords = 0x7FFFFFFFu;

__device__ inline Rsrc chunk_rsrc(const void* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)kRecords, kRsrcWord3);
}

// The plane offset is uniform; readfirstlane states it for accesses inside
// divergent branches, where the compiler would otherwise keep it in a
// vector register and wrap the access in a waterfall loop.
template <typename T>
__device__ inline T cload(Rsrc r, uint32_t voff, uint32_t soff)
{
    soff = __builtin_amdgcn_readfirstlane(soff);
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}

template <typename T>
__device__ inline void cstore(T v, Rsrc r, uint32_t voff, uint32_t soff)
{
    soff = __builtin_amdgcn_readfirstlane(soff);
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)voff, (int)soff, 0);
    } else {
        typedef int v2i __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v), r, (int)voff, (int)soff, 0);
    }
}

// voffset of an access a lane does not make: beyond kRecords, so the range
// check answers a load with 0 and drops a store, whatever the plane offset.
constexpr uint32_t kSkip = 0x80000000u;

constexpr int gcd_c(int a, int b) { return b == 0 ? a : gcd_c(b, a % b); }

}  // namespace chunk
}  // namespace scafes
#endif
