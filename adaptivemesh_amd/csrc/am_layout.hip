// Tile layout conversion for gfx950: one field's tiles between the
// reference's padded layout and the interior-first layout of the RK4
// drivers.  No counterpart in the reference.  One thread per padded cell,
// consecutive threads on consecutive padded cells, so the padded side is
// read or written in whole rows and the interior-first side in the runs its
// map keeps contiguous (interior rows, slab rows).
#include <algorithm>

#include "am_common.hpp"

namespace am {

template <int S, int P>
__global__ void __launch_bounds__(256) k_layout_convert(const float *__restrict__ src, float *__restrict__ dst,
                                                         int64_t n_cells, int to_interior_first)
{
    constexpr uint32_t T = S + 2 * P, T3 = T * T * T;
    constexpr uint32_t TS = tile_stride<S, P, 1>();
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n_cells; i += (int64_t)gridDim.x * 256) {
        const int64_t slot = i / T3;
        const uint32_t c = (uint32_t)(i - slot * T3);
        const uint32_t z = c / (T * T), y = (c / T) % T, x = c % T;
        const int64_t padded = i, inner = slot * TS + interior_first_cell<S, P>(x, y, z);
        if (to_interior_first)
            dst[inner] = src[padded];
        else
            dst[padded] = src[inner];
    }
}

template <int S, int P>
static int launch_convert(const float *src, float *dst, int32_t n_slots, int to_interior_first, void *stream)
{
    constexpr int64_t T = S + 2 * P;
    const int64_t n = (int64_t)n_slots * T * T * T;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL((k_layout_convert<S, P>), dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, src, dst,
                       n, to_interior_first);
    return check_launch("k_layout_convert");
}

template <int S, int P>
static int cell_offsets(int32_t layout, uint32_t *out)
{
    constexpr uint32_t T = S + 2 * P;
    for (uint32_t z = 0; z < T; z++)
        for (uint32_t y = 0; y < T; y++)
            for (uint32_t x = 0; x < T; x++)
                out[(z * T + y) * T + x] =
                    layout == AM_LAYOUT_PADDED ? tile_cell<S, P, 0>(x, y, z) : tile_cell<S, P, 1>(x, y, z);
    return AM_OK;
}

template <int S, int P>
static int64_t stride_of(int32_t layout)
{
    return layout == AM_LAYOUT_PADDED ? (int64_t)tile_stride<S, P, 0>() : (int64_t)tile_stride<S, P, 1>();
}

}  // namespace am

extern "C" int64_t am_tile_stride(const am_geometry *g, int32_t layout)
{
    if (!g || (layout != AM_LAYOUT_PADDED && layout != AM_LAYOUT_INTERIOR_FIRST)) return -1;
    const int s = g->size_domain, p = g->padding_domain;
    if (s == 16 && p == 3) return am::stride_of<16, 3>(layout);
    if (s == 16 && p == 2) return am::stride_of<16, 2>(layout);
    if (s == 28 && p == 2) return am::stride_of<28, 2>(layout);
    return -1;
}

extern "C" int am_tile_layout_convert(const float *src, float *dst, int32_t n_slots, const am_geometry *g,
                                      int32_t from, int32_t to, void *stream)
{
    if (!src || !dst || !g || n_slots < 0) return am::set_error(AM_ERR_INVALID, "am_tile_layout_convert: bad argument");
    const bool ok_from = from == AM_LAYOUT_PADDED || from == AM_LAYOUT_INTERIOR_FIRST;
    const bool ok_to = to == AM_LAYOUT_PADDED || to == AM_LAYOUT_INTERIOR_FIRST;
    if (!ok_from || !ok_to || from == to)
        return am::set_error(AM_ERR_INVALID, "am_tile_layout_convert: layouts %d -> %d", from, to);
    if (n_slots == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_convert, src, dst, n_slots, to == AM_LAYOUT_INTERIOR_FIRST ? 1 : 0, stream);
}

extern "C" int am_tile_cell_offsets(const am_geometry *g, int32_t layout, uint32_t *out)
{
    if (!g || !out || (layout != AM_LAYOUT_PADDED && layout != AM_LAYOUT_INTERIOR_FIRST))
        return am::set_error(AM_ERR_INVALID, "am_tile_cell_offsets: bad argument");
    AM_DISPATCH_TILE(*g, am::cell_offsets, layout, out);
}
