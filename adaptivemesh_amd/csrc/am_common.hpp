// Internal definitions shared by the host launchers and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/adaptivemesh.h"

namespace am {

// Centred FD coefficients for half-width q = 1, 2, 3 (AMR.h:1144-1150); the
// order-2q stencil starts at index q*q + q - 1.
struct FdCoef {
    float d1[15];
    float d2[15];
};
__host__ __device__ constexpr FdCoef fd_coef()
{
    return FdCoef{{-.5f, 0.f, .5f, 1.f / 12, -2.f / 3, 0.f, 2.f / 3, -1.f / 12, -1.f / 60, 3.f / 20, -3.f / 4, 0.f,
                   3.f / 4, -3.f / 20, 1.f / 60},
                  {1.f, -2.f, 1.f, -1.f / 12, 4.f / 3, -5.f / 2, 4.f / 3, -1.f / 12, 1.f / 90, -3.f / 20, 3.f / 2,
                   -49.f / 18, 3.f / 2, -3.f / 20, 1.f / 90}};
}

// Per-depth scale factors computed on the host in IEEE float exactly as the
// reference computes them in-kernel: inv_dx[d] = float(1u<<d) / outer_dx
// (AMR.h:1214, 1263), dt[d] is outer_dt / float(1u<<d).
struct DepthScales {
    float inv_dx[9];
    float dt[9];
};

// ------------------------------------------------------- tile layouts
// Floats per slot of a field array (am_tile_layout): the padded tile T^3,
// or for the interior-first layout T^3 rounded up to 32 floats, so every
// slot, interior row and slab row starts 64-byte aligned (S % 16 == 0).
template <int S, int P, int L>
__host__ __device__ constexpr uint32_t tile_stride()
{
    constexpr uint32_t T = S + 2 * P;
    return L == 0 ? T * T * T : (T * T * T + 31u) / 32u * 32u;
}

// Offset within its slot of padded cell (x, y, z), 0 <= x, y, z < T, in the
// interior-first layout: interior, then face slabs (face f = +x +y +z -x -y
// -z, layer l counted outward, cell (a, b) = the other two axes in x, y, z
// order), then the twelve edges (the interior axis i and its coordinate,
// then the two ghost coordinates, each coded 0..2P-1: c < P as c, c >= P + S
// as c - S), then the eight corners.
template <int S, int P>
__host__ __device__ __forceinline__ uint32_t interior_first_cell(uint32_t x, uint32_t y, uint32_t z)
{
    constexpr uint32_t uS = S, uP = P, S2 = uS * uS, F0 = uS * S2, E0 = F0 + 6u * uP * S2, G = 2u * uP,
                       C0 = E0 + 3u * uS * G * G;
    const uint32_t c[3] = {x, y, z};
    uint32_t inner[3], code[3], n_out = 0, out_axis = 0, in_axis = 0;
    for (int k = 0; k < 3; k++) {
        inner[k] = c[k] - uP;  // wraps for c < P
        const bool in = inner[k] < uS;
        code[k] = c[k] < uP ? c[k] : c[k] - uS;
        if (!in) {
            n_out++;
            out_axis = (uint32_t)k;
        } else {
            in_axis = (uint32_t)k;
        }
    }
    if (n_out == 0) return (inner[2] * uS + inner[1]) * uS + inner[0];
    if (n_out == 1) {
        const uint32_t k = out_axis, high = c[k] >= uP + uS;
        const uint32_t f = k + (high ? 0u : 3u), l = high ? c[k] - uP - uS : uP - 1u - c[k];
        const uint32_t a = inner[k == 0 ? 1 : 0], b = inner[k == 2 ? 1 : 2];
        return F0 + ((f * uP + l) * uS + b) * uS + a;
    }
    if (n_out == 2) {
        const uint32_t g0 = code[in_axis == 0 ? 1 : 0], g1 = code[in_axis == 2 ? 1 : 2];
        return E0 + ((in_axis * uS + inner[in_axis]) * G + g0) * G + g1;
    }
    return C0 + (code[0] * G + code[1]) * G + code[2];
}

// Offset within its slot of padded cell (x, y, z) in layout L.
template <int S, int P, int L>
__host__ __device__ __forceinline__ uint32_t tile_cell(uint32_t x, uint32_t y, uint32_t z)
{
    constexpr uint32_t T = S + 2 * P;
    if constexpr (L == 0)
        return (z * T + y) * T + x;
    else
        return interior_first_cell<S, P>(x, y, z);
}

// Offset within its slot of interior cell (x, y, z), 0 <= x, y, z < S.
template <int S, int P, int L>
__host__ __device__ __forceinline__ uint32_t tile_interior(uint32_t x, uint32_t y, uint32_t z)
{
    constexpr uint32_t T = S + 2 * P;
    if constexpr (L == 0)
        return ((z + P) * T + (y + P)) * T + (x + P);
    else
        return (z * S + y) * S + x;
}

// Interior-first layout: offset of ghost layer l of face f, in-face cell (a, b).
template <int S, int P>
__host__ __device__ __forceinline__ uint32_t slab_cell(uint32_t f, uint32_t l, uint32_t a, uint32_t b)
{
    return (uint32_t)(S * S * S) + ((f * P + l) * S + b) * S + a;
}

int set_error(int code, const char *fmt, ...);
int check_launch(const char *what);
DepthScales depth_scales(const am_geometry &g);
bool aligned16(const void *p);
// the tricubic stage fill of every face of mesh->tiles, or of the listed
// faces (what the learned fill starts from; gate: see LearnedGate)
struct LearnedGate;
int ghost_stage_fill(const am_field_batch *b, const am_mesh *mesh, const int32_t *faces, int n_faces,
                     const am_geometry *g, float stage_frac, int equal_mode, void *stream, LearnedGate gate);

}  // namespace am

// Dispatch on the tile geometry (S, P, q = fd_order / 2).  FN is a function
// template taking <S, P, Q> and returning int.
#define AM_DISPATCH_GEOMETRY(g, FN, ...)                                                                 \
    do {                                                                                                 \
        const int _s = (g).size_domain, _p = (g).padding_domain, _q = (g).fd_order / 2;                  \
        if (_s == 16 && _p == 3 && _q == 3) return FN<16, 3, 3>(__VA_ARGS__);                            \
        if (_s == 16 && _p == 3 && _q == 2) return FN<16, 3, 2>(__VA_ARGS__);                            \
        if (_s == 16 && _p == 3 && _q == 1) return FN<16, 3, 1>(__VA_ARGS__);                            \
        if (_s == 16 && _p == 2 && _q == 2) return FN<16, 2, 2>(__VA_ARGS__);                            \
        if (_s == 16 && _p == 2 && _q == 1) return FN<16, 2, 1>(__VA_ARGS__);                            \
        if (_s == 28 && _p == 2 && _q == 2) return FN<28, 2, 2>(__VA_ARGS__);                            \
        if (_s == 28 && _p == 2 && _q == 1) return FN<28, 2, 1>(__VA_ARGS__);                            \
        return ::am::set_error(AM_ERR_UNSUPPORTED_GEOMETRY, "no kernel for S=%d P=%d fd_order=%d", _s, _p, \
                               (g).fd_order);                                                            \
    } while (0)

// Dispatch on the tile layout: FN<S, P, L> for L = 0 (padded) or 1
// (interior-first); used inside a function templated on S and P.
#define AM_DISPATCH_LAYOUT(layout, FN, ...)                                                              \
    do {                                                                                                 \
        if ((layout) == AM_LAYOUT_PADDED) return FN<S, P, 0>(__VA_ARGS__);                               \
        if ((layout) == AM_LAYOUT_INTERIOR_FIRST) return FN<S, P, 1>(__VA_ARGS__);                       \
        return ::am::set_error(AM_ERR_INVALID, "unknown tile layout %d", (int)(layout));                 \
    } while (0)

// Dispatch on (S, P) only, for kernels that do not use the FD order.
#define AM_DISPATCH_TILE(g, FN, ...)                                                                     \
    do {                                                                                                 \
        const int _s = (g).size_domain, _p = (g).padding_domain;                                         \
        if (_s == 16 && _p == 3) return FN<16, 3>(__VA_ARGS__);                                          \
        if (_s == 16 && _p == 2) return FN<16, 2>(__VA_ARGS__);                                          \
        if (_s == 28 && _p == 2) return FN<28, 2>(__VA_ARGS__);                                          \
        return ::am::set_error(AM_ERR_UNSUPPORTED_GEOMETRY, "no kernel for S=%d P=%d", _s, _p);          \
    } while (0)
