// Interpolation helpers shared by the ghost-fill kernels (am_mesh.hip and
// the learned fill): the reference's quadratic and trilinear forms
// (AMR.h:673-699, 767-785), the 4-point Lagrange tricubic form, and the
// per-cell tricubic + quadratic-time ghost value, all with contraction off
// so every kernel that uses them gives the same bits.
#pragma once
#include "am_device.hpp"

#pragma clang fp contract(off)

namespace am {

struct FieldBatchDev {
    float *old_[AM_MAX_VARS];
    float *new_[AM_MAX_VARS];
    const float *older[AM_MAX_VARS];
};

// The learned fill's per-tile gate as the tricubic coarse-face fill sees
// it.  mode 0: plain fill.  mode 1 (the fill before the network): every
// listed face except those the network writes (target one level coarser,
// not the root), clearing resid[slot] of each tile with a listed face, the
// largest residual the network's blocks then record.  mode 2 (after the
// network): only the network's faces -- all their cells for a tile whose
// resid exceeds threshold, only the cells the network leaves for a tile
// within it; stats[0] / stats[1] count those faces.
struct LearnedGate {
    float *resid = nullptr;
    float threshold = 0.f;
    unsigned long long *stats = nullptr;
    int mode = 0;
};

// padded-tile buffers prolonged by one k_copy_to_child launch
struct ChildBufs {
    float *buf[3 * AM_MAX_VARS];
};

// __interpolate_dat, quadratic 10-point form (AMR.h:673-699)
template <int T>
__device__ __forceinline__ float interp_lagrange(const float *__restrict__ d, float rx, float ry, float rz, size_t n)
{
    constexpr int T2 = T * T;
    float f = d[n], fx = d[n + 1], fy = d[n + T], fz = d[n + T2];
    float fxx = d[n - 1], fyy = d[n - T], fzz = d[n - T2];
    float fxy = d[n + T + 1], fxz = d[n + T2 + 1], fyz = d[n + T2 + T];
    fxy = fxy + ((f - fx) - fy);
    fxz = fxz + ((f - fx) - fz);
    fyz = fyz + ((f - fy) - fz);
    fx = (fx - fxx) * .5f;
    fy = (fy - fyy) * .5f;
    fz = (fz - fzz) * .5f;
    fxx = (fx + fxx) - f;
    fyy = (fy + fyy) - f;
    fzz = (fz + fzz) - f;
    return f + (fx + fxx * rx) * rx + (fy + fyy * ry) * ry + (fz + fzz * rz) * rz + fxy * rx * ry + fxz * rx * rz +
           fyz * ry * rz;
}
// __interpolate_dat, trilinear form (AMR.h:767-785)
template <int T>
__device__ __forceinline__ float interp_trilinear(const float *__restrict__ d, float rx, float ry, float rz, size_t n)
{
    constexpr int T2 = T * T;
    float r00 = d[n];
    r00 = r00 * (1.f - rx) + d[n + 1] * rx;
    float r01 = d[n + T];
    r01 = r01 * (1.f - rx) + d[n + T + 1] * rx;
    r00 = r00 * (1.f - ry) + r01 * ry;
    float r10 = d[n + T2];
    r10 = r10 * (1.f - rx) + d[n + T2 + 1] * rx;
    float r11 = d[n + T2 + T];
    r11 = r11 * (1.f - rx) + d[n + T2 + T + 1] * rx;
    r10 = r10 * (1.f - ry) + r11 * ry;
    return r00 * (1.f - rz) + r10 * rz;
}
// 4-point Lagrange weights on nodes -1, 0, 1, 2 (exact on cubics)
__device__ __forceinline__ void lag4(float t, float *w)
{
    w[0] = -t * (t - 1.f) * (t - 2.f) / 6.f;
    w[1] = (t + 1.f) * (t - 1.f) * (t - 2.f) / 2.f;
    w[2] = -(t + 1.f) * t * (t - 2.f) / 2.f;
    w[3] = (t + 1.f) * t * (t - 1.f) / 6.f;
}
template <int T>
__device__ __forceinline__ float interp_tricubic(const float *__restrict__ d, const float *wx, const float *wy,
                                                 const float *wz, size_t n)
{
    constexpr int T2 = T * T;
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float ay = 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float *r = d + n + (ptrdiff_t)(k - 1) * T2 + (ptrdiff_t)(j - 1) * T - 1;
            const float ax = r[0] * wx[0] + r[1] * wx[1] + r[2] * wx[2] + r[3] * wx[3];
            ay = ay + ax * wy[j];
        }
        acc = acc + ay * wz[k];
    }
    return acc;
}


// The tricubic ghost value of one field at the read position rp (target
// padded coordinates) of a fine ghost cell: 4-point Lagrange in space on
// the clamped nodes, then quadratic in time through older/old/new at ty
// (linear through old/new when older is absent or ddiff == 0).
template <int T, int P>
__device__ __forceinline__ size_t tricubic_nodes(int tgt, const float *rp_in, float *wx, float *wy, float *wz)
{
    constexpr uint32_t uP = P;
    float rp[3] = {rp_in[0], rp_in[1], rp_in[2]};
    uint32_t ri[3];
#pragma unroll
    for (int k = 0; k < 3; k++)
        ri[k] = (tgt == 0) ? min((uint32_t)fmaxf(rp[k], 1.f), (uint32_t)T - 3u)
                           : min((uint32_t)fmaxf(rp[k], (float)(uP + 1u)), (uint32_t)T - uP - 3u);
#pragma unroll
    for (int k = 0; k < 3; k++) rp[k] -= (float)ri[k];
    lag4(rp[0], wx); lag4(rp[1], wy); lag4(rp[2], wz);
    return ri[0] + ((size_t)ri[1] + (size_t)ri[2] * T) * T;
}
template <int T>
__device__ __forceinline__ float tricubic_time(const float *dold, const float *dnew, const float *dolder,
                                               uint32_t ddiff, float ty, const float *wx, const float *wy,
                                               const float *wz, size_t rn)
{
    const float o = interp_tricubic<T>(dold, wx, wy, wz, rn);
    const float nw = interp_tricubic<T>(dnew, wx, wy, wz, rn);
    if (dolder && ddiff > 0) {
        const float od = interp_tricubic<T>(dolder, wx, wy, wz, rn);
        const float s_ = ty;
        const float c0 = .5f * s_ * (s_ - 1.f), c1 = (1.f - s_) * (1.f + s_), c2 = .5f * s_ * (s_ + 1.f);
        return od * c0 + o * c1 + nw * c2;
    }
    return o * (1.f - ty) + nw * ty;
}

// ------------------------------------------------------------------
// The same forms on a tile of either layout: a node (x, y, z) in padded
// coordinates of the tile starting at `base`, and its neighbours by offset.
// Each form reads the same cells in the same order and runs the same
// arithmetic as its padded-only counterpart above, so values are identical.
template <int S, int P, int L>
struct Cells {
    size_t base;
    uint32_t x, y, z;
    __device__ __forceinline__ size_t at(int dx, int dy, int dz) const
    {
        if constexpr (L == 0) {
            return base + tile_cell<S, P, 0>((uint32_t)((int)x + dx), (uint32_t)((int)y + dy), (uint32_t)((int)z + dz));
        } else {
            // every read of the fills lies in the tile; the clamp keeps a read
            // position outside it (which the padded layout would take from
            // the neighbouring row) inside the slot
            constexpr int T1 = S + 2 * P - 1;
            return base + tile_cell<S, P, 1>((uint32_t)min(max((int)x + dx, 0), T1),
                                             (uint32_t)min(max((int)y + dy, 0), T1),
                                             (uint32_t)min(max((int)z + dz, 0), T1));
        }
    }
};

template <int S, int P, int L>
__device__ __forceinline__ float interp_lagrange_c(const float *__restrict__ d, const Cells<S, P, L> &c, float rx,
                                                   float ry, float rz)
{
    float f = d[c.at(0, 0, 0)], fx = d[c.at(1, 0, 0)], fy = d[c.at(0, 1, 0)], fz = d[c.at(0, 0, 1)];
    float fxx = d[c.at(-1, 0, 0)], fyy = d[c.at(0, -1, 0)], fzz = d[c.at(0, 0, -1)];
    float fxy = d[c.at(1, 1, 0)], fxz = d[c.at(1, 0, 1)], fyz = d[c.at(0, 1, 1)];
    fxy = fxy + ((f - fx) - fy);
    fxz = fxz + ((f - fx) - fz);
    fyz = fyz + ((f - fy) - fz);
    fx = (fx - fxx) * .5f;
    fy = (fy - fyy) * .5f;
    fz = (fz - fzz) * .5f;
    fxx = (fx + fxx) - f;
    fyy = (fy + fyy) - f;
    fzz = (fz + fzz) - f;
    return f + (fx + fxx * rx) * rx + (fy + fyy * ry) * ry + (fz + fzz * rz) * rz + fxy * rx * ry + fxz * rx * rz +
           fyz * ry * rz;
}
template <int S, int P, int L>
__device__ __forceinline__ float interp_trilinear_c(const float *__restrict__ d, const Cells<S, P, L> &c, float rx,
                                                    float ry, float rz)
{
    float r00 = d[c.at(0, 0, 0)];
    r00 = r00 * (1.f - rx) + d[c.at(1, 0, 0)] * rx;
    float r01 = d[c.at(0, 1, 0)];
    r01 = r01 * (1.f - rx) + d[c.at(1, 1, 0)] * rx;
    r00 = r00 * (1.f - ry) + r01 * ry;
    float r10 = d[c.at(0, 0, 1)];
    r10 = r10 * (1.f - rx) + d[c.at(1, 0, 1)] * rx;
    float r11 = d[c.at(0, 1, 1)];
    r11 = r11 * (1.f - rx) + d[c.at(1, 1, 1)] * rx;
    r10 = r10 * (1.f - ry) + r11 * ry;
    return r00 * (1.f - rz) + r10 * rz;
}
template <int S, int P, int L>
__device__ __forceinline__ float interp_tricubic_c(const float *__restrict__ d, const Cells<S, P, L> &c,
                                                   const float *wx, const float *wy, const float *wz)
{
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float ay = 0.f;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float ax = d[c.at(-1, j - 1, k - 1)] * wx[0] + d[c.at(0, j - 1, k - 1)] * wx[1] +
                             d[c.at(1, j - 1, k - 1)] * wx[2] + d[c.at(2, j - 1, k - 1)] * wx[3];
            ay = ay + ax * wy[j];
        }
        acc = acc + ay * wz[k];
    }
    return acc;
}
template <int S, int P, int L>
__device__ __forceinline__ float tricubic_time_c(const float *dold, const float *dnew, const float *dolder,
                                                 uint32_t ddiff, float ty, const float *wx, const float *wy,
                                                 const float *wz, const Cells<S, P, L> &c)
{
    const float o = interp_tricubic_c<S, P, L>(dold, c, wx, wy, wz);
    const float nw = interp_tricubic_c<S, P, L>(dnew, c, wx, wy, wz);
    if (dolder && ddiff > 0) {
        const float od = interp_tricubic_c<S, P, L>(dolder, c, wx, wy, wz);
        const float s_ = ty;
        const float c0 = .5f * s_ * (s_ - 1.f), c1 = (1.f - s_) * (1.f + s_), c2 = .5f * s_ * (s_ + 1.f);
        return od * c0 + o * c1 + nw * c2;
    }
    return o * (1.f - ty) + nw * ty;
}

}  // namespace am
