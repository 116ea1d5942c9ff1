// Device-side building blocks shared by every AdaptiveMesh kernel for gfx950:
// the fast_prng step and per-thread seeding, the compressed_float3 codec, and
// the octree / boundary decoders.  Every routine is written so that, compiled
// with -ffp-contract=off and IEEE division, it produces the same bits as the
// reference's routine on an IEEE host:
//   fast_prng                  tensors.h:276-304
//   compressed_float3          tensors.h:305-394
//   octree_abs_pos::depth      AMR.h:27-29
//   domain_boundary accessors  AMR.h:87-99
//   active_depth               AMR.h:640-644
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_common.hpp"

namespace am {

// ---------------------------------------------------------------- PRNG
__device__ __forceinline__ int32_t prng_next(int32_t &s)
{
    int32_t v = (int32_t)((uint32_t)s + 1124781675u);
    v ^= v >> 7;
    v = (int32_t)((uint32_t)v * 1246477263u);
    s = v;
    return v;
}
// seed = gs + x*2178614; seed ^= next()*y; seed ^= next()*z  (AMR.h:1160-1162)
__device__ __forceinline__ int32_t thread_seed(int32_t gs, uint32_t x, uint32_t y, uint32_t z)
{
    int32_t s = (int32_t)((uint32_t)gs + x * 2178614u);
    uint32_t t = (uint32_t)prng_next(s) * y;
    s = (int32_t)((uint32_t)s ^ t);
    t = (uint32_t)prng_next(s) * z;
    s = (int32_t)((uint32_t)s ^ t);
    return s;
}

// thread_seed split for a thread that keeps x and y and walks z: the part
// that depends on (gs, x, y) -- the state after the second draw -- and the
// final mix with z, so thread_seed(gs, x, y, z) == thread_seed_z(
// thread_seed_xy(gs, x, y), z) bit for bit
__device__ __forceinline__ int32_t thread_seed_xy(int32_t gs, uint32_t x, uint32_t y)
{
    int32_t s = (int32_t)((uint32_t)gs + x * 2178614u);
    const uint32_t t = (uint32_t)prng_next(s) * y;
    s = (int32_t)((uint32_t)s ^ t);
    prng_next(s);
    return s;
}
__device__ __forceinline__ int32_t thread_seed_z(int32_t sxy, uint32_t z)
{
    return (int32_t)((uint32_t)sxy ^ ((uint32_t)sxy * z));
}

// ------------------------------------------------- compressed_float3 codec
__device__ __forceinline__ float cf3_clamp(float f) { return fmaxf(-255.f, fminf(f, 255.f)); }

__device__ __forceinline__ uint32_t cf3_exponent(float x, float y, float z)
{
    const uint32_t ex = (__float_as_uint(x) >> 23) & 255u;
    const uint32_t ey = (__float_as_uint(y) >> 23) & 255u;
    const uint32_t ez = (__float_as_uint(z) >> 23) & 255u;
    const uint32_t m = max(max(ex, max(ey, ez)), 112u) - 112u;
    return min(m, 31u);
}
__device__ __forceinline__ uint32_t cf3_assemble(uint32_t e, float x, float y, float z)
{
    x = cf3_clamp(x); y = cf3_clamp(y); z = cf3_clamp(z);
    uint32_t w = e << 27;
    w |= (uint32_t)__fadd_rn(fabsf(x), .5f) | ((uint32_t)(x < 0.f) << 8);
    w |= ((uint32_t)__fadd_rn(fabsf(y), .5f) << 9) | ((uint32_t)(y < 0.f) << 17);
    w |= ((uint32_t)__fadd_rn(fabsf(z), .5f) << 18) | ((uint32_t)(z < 0.f) << 26);
    return w;
}
// The scale 2^(22-e) is exact, so the product below is the reference's
// result.x *= ((1<<15)/(1/128)) / float(1<<e) (tensors.h:361-363).
__device__ __forceinline__ float cf3_scale(uint32_t e) { return __uint_as_float((127u + 22u - e) << 23); }

// One dither term of the pack, generate_float() * 0.95 (tensors.h:297,
// 365-367), with one rounding instead of two: generate_float() is float(v)
// times 2.3283064E-10f, which is exactly 2^-32, so that product is exact and
// the reference's result is float(v) * 0.95f correctly rounded, then scaled
// by 2^-32 -- the same value as float(v) times the float 0.95f * 2^-32
// correctly rounded (no product here is subnormal: |float(v)| >= 1 or 0).
__device__ __forceinline__ float cf3_dither(int32_t &rng)
{
    constexpr float K = 0.95f * 2.3283064E-10f;  // exact: a power-of-two scaling of 0.95f
    return __fmul_rn((float)prng_next(rng), K);
}
// The dithered pack's exponent as a field, e << 23: the largest exponent
// field of the three components, less 112 << 23, clamped to [0, 31 << 23]
// (cf3_exponent with the shifts folded away; NaN and inf have the largest
// field, so they give 31 as in the reference).
__device__ __forceinline__ uint32_t cf3_exponent_field(float x, float y, float z)
{
    constexpr uint32_t EXP = 0x7f800000u, LOW = 112u << 23;
    const uint32_t m = max(max(__float_as_uint(x) & EXP, __float_as_uint(y) & EXP), __float_as_uint(z) & EXP);
    return min(max(m, LOW) - LOW, 31u << 23);
}
// One component of a dithered word, the 8-bit magnitude with the sign at bit
// 8 inverted (cf3_pack flips the three sign bits at the end).  c = -clamp(x)
// comes from one median: v_med3_f32(-x, -255, 255) equals
// -fmaxf(-255, fminf(x, 255)) for every finite and infinite x and every
// quiet NaN (x here is a sum, never a signalling NaN).  The reference tests
// x < 0 after the clamp (tensors.h:368-371); the sign bit of clamp(x) is that
// test because a dithered value is never -0 (the dither is +0 or nonzero,
// and an exact zero sum rounds to +0) and NaN clamps to +255.
__device__ __forceinline__ uint32_t cf3_dithered_component(float x)
{
    const float c = __builtin_amdgcn_fmed3f(-x, -255.f, 255.f);
    const uint32_t m = (uint32_t)__fadd_rn(fabsf(c), .5f);
    return m | ((__float_as_uint(c) >> 31) << 8);
}
// dithered pack (tensors.h:350-373): three draws, x then y then z
__device__ __forceinline__ uint32_t cf3_pack(float x, float y, float z, int32_t &rng)
{
    const uint32_t e23 = cf3_exponent_field(x, y, z);
    const float sc = __uint_as_float((127u + 22u) * (1u << 23) - e23);  // 2^(22-e), exact
    x = __fmul_rn(x, sc); y = __fmul_rn(y, sc); z = __fmul_rn(z, sc);
    x = __fadd_rn(x, cf3_dither(rng));
    y = __fadd_rn(y, cf3_dither(rng));
    z = __fadd_rn(z, cf3_dither(rng));
    const uint32_t w = (e23 << 4) | cf3_dithered_component(x) | (cf3_dithered_component(y) << 9) |
                       (cf3_dithered_component(z) << 18);
    return w ^ ((1u << 8) | (1u << 17) | (1u << 26));
}
// undithered pack (tensors.h:374-393)
__device__ __forceinline__ uint32_t cf3_pack_plain(float x, float y, float z)
{
    const uint32_t e = cf3_exponent(x, y, z);
    const float sc = cf3_scale(e);
    return cf3_assemble(e, __fmul_rn(x, sc), __fmul_rn(y, sc), __fmul_rn(z, sc));
}
// unpack (tensors.h:335-348).  The reference's chain
// 2^(e-15) * ((m * (1 - 2s)) * (1/128)) rounds nowhere: every factor is a
// power of two or an 8-bit integer, so its value is exactly +-m * 2^(e-22),
// a normal float for m > 0.  That value is formed here with one exact
// conversion, one exact product and the sign bit copied in (zero keeps the
// sign, -0 for s = 1 as in the reference): 5 ALU operations per component
// instead of 7, bit-identical for all 2^32 words.
__device__ __forceinline__ float cf3_component(uint32_t m8, uint32_t sign31, float scale)
{
    return __uint_as_float(__float_as_uint((float)m8 * scale) | (sign31 & 0x80000000u));
}
__device__ __forceinline__ float3 cf3_unpack(uint32_t w)
{
    const float scale = __uint_as_float(((w >> 27) + 105u) << 23);  // 2^(e-22)
    float3 r;
    r.x = cf3_component(w & 255u, w << 23, scale);
    r.y = cf3_component((w >> 9) & 255u, w << 14, scale);
    r.z = cf3_component((w >> 18) & 255u, w << 5, scale);
    return r;
}

// ------------------------------------------------------------ octree
__device__ __forceinline__ int pos_depth(uint32_t d) { return (~d) ? (int)(d >> 27) : -1; }
__device__ __forceinline__ bool active_depth(uint32_t sub, uint32_t depth, uint32_t maxd)
{
    return (sub & ((1u << (maxd - depth)) - 1u)) == 0u;
}
__device__ __forceinline__ int bnd_target(uint32_t nd)
{
    return (int)((nd & 16777215u) | ((nd & (1u << 23)) ? 4278190080u : 0u));
}
__device__ __forceinline__ uint32_t bnd_ddiff(uint32_t nd) { return (nd >> 24) & 15u; }
__device__ __forceinline__ int bnd_depth(uint32_t nd) { return ((nd >> 28) == 15u) ? -1 : (int)(nd >> 28); }

}  // namespace am
