// Classical RK4 combine for one interior cell of NV fields and the
// equal-depth ghost push, shared by every stage kernel.
//
// Stage s of y' = k(y) with accumulator acc (interior-only layout
// slot * S^3 + (z * S + y) * S + x, index `w`; state, old_ and new_ are
// padded tiles, index `r`):
//   s = 0: acc  = k;   new = y_n + dt/2 k
//   s = 1: acc += 2k;  new = y_n + dt/2 k
//   s = 2: acc += 2k;  new = y_n + dt k
//   s = 3:             new = y_n + dt/6 (acc + k)
// The stage-0 input is y_n itself (read from `state`), later stages read y_n
// from `old_`.
//
// Push: an interior cell within P of face f is also a ghost cell of f's
// target when that target has the same depth; it lands in the opposite face
// slab of the target, shifted by S cells along f's axis.  Writing it here
// makes equal-depth ghosts exact copies without a separate fill pass.  Only
// cells within `layers` (<= P) of the face are pushed: the innermost ghost
// layers, all that stencils of half-width `layers` read.
#pragma once
#include "am_device.hpp"

namespace am {

template <int NV, int STAGE>
__device__ __forceinline__ void rk4_combine(const float *const *state, const float *const *old_, float *const *acc,
                                            float *const *new_, float (&k)[NV], size_t r, size_t w, float dt)
{
#pragma unroll
    for (int v = 0; v < NV; v++) {
        const float yn = STAGE == 0 ? state[v][r] : old_[v][r];
        float nv;
        if (STAGE == 0) {
            acc[v][w] = k[v];
            nv = yn + (.5f * dt) * k[v];
        } else if (STAGE == 1 || STAGE == 2) {
            acc[v][w] = acc[v][w] + 2.f * k[v];
            nv = yn + (STAGE == 1 ? .5f * dt : dt) * k[v];
        } else {
            nv = yn + (dt * (1.f / 6.f)) * (acc[v][w] + k[v]);
        }
        new_[v][r] = nv;
        k[v] = nv;  // the stage output, for the push
    }
}

// L: tile layout (0 padded, 1 interior-first).  In the interior-first layout
// the cell lands in the target's slab of the opposite face, layer = its
// distance from the shared face, so the cells of one row of this tile go to
// one contiguous slab row of the target.
template <int NV, int S, int P, int L = 0>
__device__ __forceinline__ void rk4_push(float *const *new_, const float (&y)[NV], const uint64_t *boundaries, int node,
                                         int x, int y_, int z, int layers, int axes)
{
    constexpr int T = S + 2 * P;
    constexpr size_t TS = tile_stride<S, P, L>();
    const int c3[3] = {x, y_, z};
#pragma unroll
    for (int f = 0; f < 6; f++) {
        const int ax = f % 3;
        const bool near = (f < 3 ? c3[ax] >= S - layers : c3[ax] < layers) && ((axes >> ax) & 1);
        if (!near) continue;
        const uint32_t nd = (uint32_t)boundaries[(size_t)node * 6 + f];
        const int tgt = bnd_target(nd);
        if (tgt < 0 || bnd_depth(nd) <= 0 || bnd_ddiff(nd) != 0) continue;
        size_t dst;
        if constexpr (L == 0) {
            int p3[3] = {x + P, y_ + P, z + P};
            p3[ax] += f < 3 ? -S : S;
            dst = (size_t)tgt * TS + ((size_t)p3[2] * T + p3[1]) * T + p3[0];
        } else {
            const uint32_t l = f < 3 ? (uint32_t)(S - 1 - c3[ax]) : (uint32_t)c3[ax];
            const uint32_t a = (uint32_t)c3[ax == 0 ? 1 : 0], bq = (uint32_t)c3[ax == 2 ? 1 : 2];
            dst = (size_t)tgt * TS + slab_cell<S, P>((uint32_t)((f + 3) % 6), l, a, bq);
        }
#pragma unroll
        for (int v = 0; v < NV; v++) new_[v][dst] = y[v];
    }
}

}  // namespace am
