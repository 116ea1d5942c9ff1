// Refinement criterion of the SimpleWave plugin for gfx950: per listed tile
// and octant, the largest scaled gradient and diagonal-Hessian magnitude of
// its cells (__amr_criterion_1 and __amr_criterion_2, SimpleWave.h:91-144).
//
// One workgroup per (tile, octant) reduces the (S/2)^3 cells of the octant.
// A cell contributes max(|d1| dx_d, |d2| dx_d^2) with dx_d = outer_dx / 2^d,
// written as the reference writes it: length(v) * outer_dx / float(1 << d)
// and length(v) * outer_dx * outer_dx / float(1 << 2d), length = sqrtf of
// x*x + y*y + z*z, contraction off.  The maximum does not depend on the
// order of the reduction, so every value equals the reference's two-pass
// result bit for bit.  A tile that is unused or not active at `substep`
// yields NaN (the reference's marker for "no domain").
#include "am_device.hpp"

#pragma clang fp contract(off)

namespace am {

template <int S>
__global__ void __launch_bounds__(256) k_amr_criterion(const uint32_t *__restrict__ d1,
                                                       const uint32_t *__restrict__ d2,
                                                       const uint32_t *__restrict__ positions,
                                                       const int32_t *__restrict__ tiles, uint32_t substep,
                                                       uint32_t maxd, float outer_dx, float *__restrict__ out)
{
    constexpr int H = S / 2, H3 = H * H * H;
    const int ti = blockIdx.x >> 3, oct = blockIdx.x & 7;
    const int node = tiles[ti];
    const int depth = pos_depth(positions[node]);
    if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) {
        if (threadIdx.x == 0) out[(size_t)node * 8 + oct] = __builtin_nanf("");
        return;  // uniform per workgroup
    }
    const int ox = (oct & 1) * H, oy = ((oct >> 1) & 1) * H, oz = (oct >> 2) * H;
    const float s1 = (float)(1u << depth), s2 = (float)(1u << (2 * depth));
    float crit = 0.f;
    for (int i = threadIdx.x; i < H3; i += 256) {
        const int x = ox + i % H, y = oy + (i / H) % H, z = oz + i / (H * H);
        const size_t w = (size_t)node * (S * S * S) + ((size_t)z * S + y) * S + x;
        float3 v = cf3_unpack(d1[w]);
        crit = fmaxf(crit, sqrtf(v.x * v.x + v.y * v.y + v.z * v.z) * outer_dx / s1);
        v = cf3_unpack(d2[w]);
        crit = fmaxf(crit, sqrtf(v.x * v.x + v.y * v.y + v.z * v.z) * outer_dx * outer_dx / s2);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) crit = fmaxf(crit, __shfl_down(crit, off, 64));
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = crit;
    __syncthreads();
    if (threadIdx.x == 0) out[(size_t)node * 8 + oct] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

template <int S, int P>
static int launch_criterion(const uint32_t *d1, const uint32_t *d2, const am_mesh *mesh, const am_geometry *g,
                            float *out, void *stream)
{
    hipLaunchKernelGGL((k_amr_criterion<S>), dim3(mesh->n_tiles * 8), dim3(256), 0, (hipStream_t)stream, d1, d2,
                       mesh->positions, mesh->tiles, mesh->substep, mesh->max_depth, g->outer_dx, out);
    return check_launch("k_amr_criterion");
}

}  // namespace am

extern "C" int am_amr_criterion(const uint32_t *d1, const uint32_t *d2diag, const am_mesh *mesh,
                                const am_geometry *g, float *out, void *stream)
{
    if (!d1 || !d2diag || !mesh || !g || !out || !mesh->tiles || !mesh->positions)
        return am::set_error(AM_ERR_INVALID, "am_amr_criterion: null argument");
    if (mesh->n_tiles < 0) return am::set_error(AM_ERR_INVALID, "am_amr_criterion: negative tile count");
    if (mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_criterion, d1, d2diag, mesh, g, out, stream);
}
