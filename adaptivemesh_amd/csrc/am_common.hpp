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

int set_error(int code, const char *fmt, ...);
int check_launch(const char *what);
DepthScales depth_scales(const am_geometry &g);
bool aligned16(const void *p);
// the tricubic stage fill of every face (what the learned fill starts from)
int ghost_stage_fill(const am_field_batch *b, const am_mesh *mesh, const am_geometry *g, float stage_frac,
                     int equal_mode, void *stream);

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

// Dispatch on (S, P) only, for kernels that do not use the FD order.
#define AM_DISPATCH_TILE(g, FN, ...)                                                                     \
    do {                                                                                                 \
        const int _s = (g).size_domain, _p = (g).padding_domain;                                         \
        if (_s == 16 && _p == 3) return FN<16, 3>(__VA_ARGS__);                                          \
        if (_s == 16 && _p == 2) return FN<16, 2>(__VA_ARGS__);                                          \
        if (_s == 28 && _p == 2) return FN<28, 2>(__VA_ARGS__);                                          \
        return ::am::set_error(AM_ERR_UNSUPPORTED_GEOMETRY, "no kernel for S=%d P=%d", _s, _p);          \
    } while (0)
