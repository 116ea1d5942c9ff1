// Storage formats of the derivative "words" between the FD kernel and the
// RHS kernels.  The reference format is compressed_float3 (one dithered
// 32-bit word per cell, tensors.h:305-394); the precision study also stores
// them as three fp32 (12 bytes) or three fp16 (6 bytes, round to nearest
// even) values per cell, undithered.
#pragma once
#include "am_device.hpp"

namespace am {

enum { WORD_CF3 = 0, WORD_F32 = 1, WORD_F16 = 2 };

// bytes per cell of one word array
__host__ __device__ constexpr int word_bytes(int fmt) { return fmt == WORD_F32 ? 12 : fmt == WORD_F16 ? 6 : 4; }

// o: the cell's index in the word array (below 2^30: the cf3 store
// addresses it as a 32-bit byte offset from the array's base)
template <int FMT>
__device__ __forceinline__ void word_put(void *base, uint32_t o, float x, float y, float z, int32_t &rng)
{
    if (FMT == WORD_CF3) {
        *reinterpret_cast<uint32_t *>(static_cast<char *>(base) + 4u * o) = cf3_pack(x, y, z, rng);
    } else if (FMT == WORD_F32) {
        float *p = static_cast<float *>(base) + 3 * (size_t)o;
        p[0] = x; p[1] = y; p[2] = z;
    } else {
        // The fp32 values are materialised before the conversion: left to
        // itself the compiler fuses (half)(a * b) into one mixed-precision
        // multiply-add with a +0 addend (v_fma_mixlo_f16), which rounds the
        // exact product once instead of the fp32 product, and turns a -0
        // product into +0.  The empty asm only pins the value in a register.
        asm volatile("" : "+v"(x), "+v"(y), "+v"(z));
        _Float16 *p = static_cast<_Float16 *>(base) + 3 * (size_t)o;
        p[0] = (_Float16)x; p[1] = (_Float16)y; p[2] = (_Float16)z;
    }
}

template <int FMT>
__device__ __forceinline__ float3 word_get(const void *base, size_t o)
{
    if (FMT == WORD_CF3) return cf3_unpack(static_cast<const uint32_t *>(base)[o]);
    if (FMT == WORD_F32) {
        const float *p = static_cast<const float *>(base) + 3 * o;
        return make_float3(p[0], p[1], p[2]);
    }
    const _Float16 *p = static_cast<const _Float16 *>(base) + 3 * o;
    return make_float3((float)p[0], (float)p[1], (float)p[2]);
}

}  // namespace am
