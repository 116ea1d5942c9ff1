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

// The same combine with the stage's inputs already loaded (yn[v] = y_n and
// ac[v] = acc of the cell, stages 1-3), the cell addressed by its byte
// offsets ro in the field arrays and wo in the accumulator arrays.
template <int NV, int STAGE>
__device__ __forceinline__ void rk4_combine_at(float *const *acc, float *const *new_, float (&k)[NV],
                                               const float (&yn)[NV], const float (&ac)[NV], uint32_t ro,
                                               uint32_t wo, float dt)
{
#pragma unroll
    for (int v = 0; v < NV; v++) {
        float nv;
        float *ap = reinterpret_cast<float *>(reinterpret_cast<char *>(acc[v]) + wo);
        if (STAGE == 0) {
            *ap = k[v];
            nv = yn[v] + (.5f * dt) * k[v];
        } else if (STAGE == 1 || STAGE == 2) {
            *ap = ac[v] + 2.f * k[v];
            nv = yn[v] + (STAGE == 1 ? .5f * dt : dt) * k[v];
        } else {
            nv = yn[v] + (dt * (1.f / 6.f)) * (ac[v] + k[v]);
        }
        *reinterpret_cast<float *>(reinterpret_cast<char *>(new_[v]) + ro) = nv;
        k[v] = nv;  // the stage output, for the push
    }
}

// L: tile layout (0 padded, 1 interior-first).  In the interior-first layout
// the cell lands in the target's slab of the opposite face, layer = its
// distance from the shared face, so the cells of one row of this tile go to
// one contiguous slab row of the target.
// B32: address the fields with 32-bit byte offsets on their scalar bases, so
// no lane does 64-bit arithmetic per field (the caller's field arrays stay
// below 4 GiB)
template <int NV, int S, int P, int L = 0, bool B32 = false>
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
        if constexpr (B32) {
            const uint32_t db = 4u * (uint32_t)dst;
#pragma unroll
            for (int v = 0; v < NV; v++) *reinterpret_cast<float *>(reinterpret_cast<char *>(new_[v]) + db) = y[v];
        } else {
#pragma unroll
            for (int v = 0; v < NV; v++) new_[v][dst] = y[v];
        }
    }
}

// The x faces of the push for a workgroup that holds one whole z-plane of an
// interior-first tile (S * S == 256 threads, thread = x + S * y): the cells
// of an x face are a column of the plane, one lane in 16, while the target's
// slab row (layer l, row b = z) holds them as S consecutive floats.  They go
// through LDS (xbuf: 2 * P * NV * S floats) and leave as whole rows, one
// field per wave: 64 lanes = all the rows of one field at two layers.  Every
// thread of the workgroup must call this (it holds two barriers).  The
// fields are addressed with 32-bit byte offsets, as rk4_push<..., true>.
template <int NV, int S, int P>
__device__ __forceinline__ void rk4_push_x_rows(float *const *new_, const float (&y)[NV], const uint64_t *boundaries,
                                                int node, int x, int y_, int z, int layers, int axes, float *xbuf)
{
    static_assert(S * S == 256, "one z-plane per 256-thread workgroup");
    constexpr size_t TS = tile_stride<S, P, 1>();
    int tgt[2] = {-1, -1};  // side 0: the +x face, side 1: the -x face (uniform per workgroup)
    if (axes & 1) {
#pragma unroll
        for (int sd = 0; sd < 2; sd++) {
            const uint32_t nd = (uint32_t)boundaries[(size_t)node * 6 + (sd ? 3 : 0)];
            const int t = bnd_target(nd);
            if (t >= 0 && bnd_depth(nd) > 0 && bnd_ddiff(nd) == 0) tgt[sd] = t;
        }
    }
    if (tgt[0] < 0 && tgt[1] < 0) return;
    const int sd = x >= S - layers ? 0 : (x < layers ? 1 : -1);
    if (sd >= 0 && tgt[sd] >= 0) {
        const int l = sd == 0 ? S - 1 - x : x;  // layer of the target's slab: distance from the shared face
#pragma unroll
        for (int v = 0; v < NV; v++) xbuf[((sd * P + l) * NV + v) * S + y_] = y[v];
    }
    __syncthreads();
    const int lane = (int)threadIdx.x & 63, wave = (int)threadIdx.x >> 6;
    const int n_row = 2 * layers * S;  // (side, layer, row) of one field
    for (int v0 = 0; v0 < NV; v0 += 4) {
        const int v = __builtin_amdgcn_readfirstlane(v0 + wave);
        if (v >= NV) break;
        float *dst = new_[v];
        for (int i = lane; i < n_row; i += 64) {
            const int s2 = i / (layers * S), l = (i / S) % layers, yy = i % S;
            if (tgt[s2] < 0) continue;
            const uint32_t db =
                4u * (uint32_t)((size_t)tgt[s2] * TS + slab_cell<S, P>(s2 == 0 ? 3u : 0u, (uint32_t)l, (uint32_t)yy, (uint32_t)z));
            *reinterpret_cast<float *>(reinterpret_cast<char *>(dst) + db) = xbuf[((s2 * P + l) * NV + v) * S + yy];
        }
    }
    __syncthreads();  // xbuf is refilled by the workgroup's next item
}

}  // namespace am
