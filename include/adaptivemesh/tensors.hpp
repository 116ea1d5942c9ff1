// Host-side numeric types of the AdaptiveMesh drop-in API: fast_prng and the
// undithered compressed_float3 pack used for boundary descriptors.  Same
// names, state and bit layout as the reference (tensors.h:276-394); the
// dithered device pack lives in the HIP kernels.
#pragma once
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>

namespace am {

// tensors.h:276-304
struct fast_prng {
    int32_t seed;
    explicit fast_prng(int32_t s = 12894712) : seed(s) {}
    int32_t generate_int()
    {
        int32_t v = (int32_t)((uint32_t)seed + 1124781675u);
        v ^= v >> 7;
        seed = (int32_t)((uint32_t)v * 1246477263u);
        return seed;
    }
    int32_t generate_int(int32_t min_incl, int32_t max_excl)
    {
        const int32_t s = generate_int();
        return (int32_t)(((uint64_t)(uint32_t)s * (uint64_t)(uint32_t)(max_excl - min_incl)) >> 32) + min_incl;
    }
    float generate_float() { return (float)generate_int() * 2.3283064E-10f; }
    float generate_float01() { return (float)(uint32_t)generate_int() * 2.3283064E-10f; }
};

// compressed_float3: 5-bit shared exponent (offset 15) in bits 27-31, three
// sign+8-bit-mantissa lanes below (tensors.h:254-274, 305-394).
struct compressed_float3 {
    uint32_t data = 0;

    static uint32_t bits(float f)
    {
        uint32_t u;
        memcpy(&u, &f, 4);
        return u;
    }
    static float from_bits(uint32_t u)
    {
        float f;
        memcpy(&f, &u, 4);
        return f;
    }
    compressed_float3() = default;
    // undithered constructor (tensors.h:374-393)
    compressed_float3(float x, float y, float z)
    {
        uint32_t e = std::max(std::max((bits(x) >> 23) & 255u, std::max((bits(y) >> 23) & 255u, (bits(z) >> 23) & 255u)),
                              112u) - 112u;
        e = std::min(e, 31u);
        const float sc = (32768u / (1.f / 128u)) / (float)(1u << e);
        x *= sc; y *= sc; z *= sc;
        x = std::fmax(-255.f, std::fmin(x, 255.f));
        y = std::fmax(-255.f, std::fmin(y, 255.f));
        z = std::fmax(-255.f, std::fmin(z, 255.f));
        data = e << 27;
        data |= (uint32_t)(std::fabs(x) + .5f) + ((uint32_t)(x < 0.f) << 8);
        data |= ((uint32_t)(std::fabs(y) + .5f) << 9) + ((uint32_t)(y < 0.f) << 17);
        data |= ((uint32_t)(std::fabs(z) + .5f) << 18) + ((uint32_t)(z < 0.f) << 26);
    }
    // operator float3 (tensors.h:335-348)
    void unpack(float *o) const
    {
        const float base = from_bits((127u - 15u + (data >> 27)) << 23);
        o[0] = base * ((float)(data & 255u) * (1.f - 2.f * (float)((data >> 8) & 1u)) * (1.f / 128u));
        o[1] = base * ((float)((data >> 9) & 255u) * (1.f - 2.f * (float)((data >> 17) & 1u)) * (1.f / 128u));
        o[2] = base * ((float)((data >> 18) & 255u) * (1.f - 2.f * (float)((data >> 26) & 1u)) * (1.f / 128u));
    }
};

}  // namespace am
