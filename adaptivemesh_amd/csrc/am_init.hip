// The reference's wave-test initial data (__init_temp, kernel.cu:9-43) for
// gfx950: two opposite bumps f = mul (10 / (|x - a|^2 + 10) - 10 / (|x + a|^2
// + 10)), a = (-25, 1/2, 0), written into old and new at every cell of a
// depth-0 tile and at the interior cells of deeper tiles.  Cell positions
// follow the reference: the tile centre (AMR.h:33-43, accumulated in float)
// plus (i + 1/2 - T/2) outer_dx / 2^depth per axis.  One thread per padded
// cell; contraction off, so the values equal the IEEE restatement.
#include "am_device.hpp"

#pragma clang fp contract(off)

namespace am {

template <int S, int P>
__global__ void __launch_bounds__(256) k_init_temp(float *__restrict__ old_d, float *__restrict__ new_d,
                                                   const uint32_t *__restrict__ positions,
                                                   const int32_t *__restrict__ tiles, float outer_size,
                                                   float outer_dx, float mul)
{
    constexpr int T = S + 2 * P, T3 = T * T * T;
    const int node = tiles[blockIdx.y];
    const int cell = blockIdx.x * 256 + (int)threadIdx.x;
    if (cell >= T3) return;
    const uint32_t pos = positions[node];
    const int depth = pos_depth(pos);
    if (depth == -1) return;
    const int x = cell % T, y = (cell / T) % T, z = cell / (T * T);
    if (depth > 0 && (x < P || x >= S + P || y < P || y >= S + P || z < P || z >= S + P)) return;
    float c[3] = {0.f, 0.f, 0.f};
    uint32_t s = pos & 134217727u;
    for (int i = 0; i < depth; i++, s >>= 3) {
        const float w = (float)(1 << i);
        c[0] += ((s & 1u) ? .25f : -.25f) / w;
        c[1] += ((s & 2u) ? .25f : -.25f) / w;
        c[2] += ((s & 4u) ? .25f : -.25f) / w;
    }
    const float scale = (float)(1u << depth);
    const int idx[3] = {x, y, z};
    float p[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float a = (float)idx[k] + .5f;
        a = a - (float)T * .5f;
        a = a * outer_dx;
        a = a / scale;
        p[k] = c[k] * outer_size + a;
    }
    const float dm[3] = {p[0] + 25.f, p[1] - .5f, p[2]};  // x - a
    const float dp[3] = {p[0] - 25.f, p[1] + .5f, p[2]};  // x + a
    const float q1 = dm[0] * dm[0] + dm[1] * dm[1] + dm[2] * dm[2];
    const float q2 = dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2];
    const float v = mul * (10.f / (q1 + 10.f) - 10.f / (q2 + 10.f));
    const size_t o = (size_t)node * T3 + cell;
    old_d[o] = v;
    new_d[o] = v;
}

template <int S, int P>
static int launch_init_temp(float *old_d, float *new_d, const am_mesh *mesh, const am_geometry *g, float mul,
                            void *stream)
{
    constexpr int T3 = (S + 2 * P) * (S + 2 * P) * (S + 2 * P);
    hipLaunchKernelGGL((k_init_temp<S, P>), dim3((T3 + 255) / 256, mesh->n_tiles), dim3(256), 0,
                       (hipStream_t)stream, old_d, new_d, mesh->positions, mesh->tiles, g->outer_size, g->outer_dx,
                       mul);
    return check_launch("k_init_temp");
}

}  // namespace am

extern "C" int am_init_temp(float *old_field, float *new_field, const am_mesh *mesh, const am_geometry *g, float mul,
                            void *stream)
{
    if (!old_field || !new_field || !mesh || !g || !mesh->positions || !mesh->tiles)
        return am::set_error(AM_ERR_INVALID, "am_init_temp: null argument");
    if (mesh->n_tiles < 0 || mesh->n_tiles > 65535)
        return am::set_error(AM_ERR_INVALID, "am_init_temp: %d tiles (0..65535)", mesh->n_tiles);
    if (mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_init_temp, old_field, new_field, mesh, g, mul, stream);
}
