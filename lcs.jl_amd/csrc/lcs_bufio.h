// Range-checked plane access through buffer resources.
//
// A Plane describes one z-plane of a halo array: its base address and its
// size in bytes, or size 0 for a plane outside the array (or one a kernel
// chooses to skip).  Loads at a byte offset past the size return 0 and stores
// there are dropped, in hardware: on the GPU a Plane is a buffer resource and
// every access is a raw buffer load or store, so a kernel reads and writes
// halo rings, ragged tile edges and out-of-range planes without a branch.
// Straight-line memory code lets the compiler count outstanding loads
// exactly, so loads issued one or two planes ahead stay in flight instead of
// being waited for at the next branch.
//
// kOff0 is the offset of "no element": past every plane.
//
// The same functions run on the host (for the thread-by-thread emulation of
// the tile kernels) with the same semantics, range checks in software.
#pragma once

#include <cstdint>
#include <cstring>

#include <hip/hip_runtime.h>

namespace lcs {

constexpr uint32_t kOff0 = 0x80000000u;

struct Plane {
#if defined(__HIP_DEVICE_COMPILE__)
    __amdgpu_buffer_rsrc_t r;
#else
    const unsigned char *base;
    uint32_t bytes;
#endif
};

// word 3 of a raw (stride 0) gfx9 buffer resource: 32-bit data format
constexpr int kRsrcWord3 = 0x00020000;

__host__ __device__ inline Plane make_plane(const void *base, uint32_t bytes)
{
    Plane p;
#if defined(__HIP_DEVICE_COMPILE__)
    p.r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, kRsrcWord3);
#else
    p.base = static_cast<const unsigned char *>(base);
    p.bytes = bytes;
#endif
    return p;
}

template <typename T>
__host__ __device__ inline T pload(const Plane &p, uint32_t off)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit elements");
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(p.r, (int)off, 0, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(p.r, (int)off, 0, 0));
    }
#else
    T v = T(0);
    if ((uint64_t)off + sizeof(T) <= p.bytes) std::memcpy(&v, p.base + off, sizeof(T));
    return v;
#endif
}

template <typename T>
__host__ __device__ inline void pstore(const Plane &p, uint32_t off, T v)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit elements");
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), p.r, (int)off, 0, 0);
    } else {
        using V2 = uint32_t __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), p.r, (int)off, 0, 0);
    }
#else
    if ((uint64_t)off + sizeof(T) <= p.bytes) std::memcpy(const_cast<unsigned char *>(p.base) + off, &v, sizeof(T));
#endif
}

}  // namespace lcs
