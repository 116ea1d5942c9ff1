// Slice rasteriser and RGBA8 encoder for gfx950: a z-slice of one field
// through the whole octree as an image (__rasterise_to_image, AMR.h:1313-1353,
// with the quadratic __interpolate_dat that AMR.h:4 selects and the tile-edge
// boxes of DISPLAY_BOUNDING_BOXES), and the float4 -> RGBA8 packing of
// __encode_img / ___rgba (image_process.h:10-54).
//
// One thread per pixel, 16x16 pixels per workgroup.  Each thread walks the
// octree from the root to the leaf holding its point (children[] is a few KB
// and stays in L2), then reads the 10-point stencil of that tile.  Output
// only: not on the timestep path.  Contraction is off so the pixels equal the
// IEEE restatement bit for bit.
#include "am_interp.hpp"

#pragma clang fp contract(off)

namespace am {

template <int S, int P>
__global__ void __launch_bounds__(256) k_rasterise(const float *__restrict__ data, const int32_t *__restrict__ children,
                                                   uint32_t width, uint32_t height, float z_slice, int32_t boxes,
                                                   float4 *__restrict__ raster)
{
    constexpr int T = S + 2 * P;
    const uint32_t px = blockIdx.x * 16u + (threadIdx.x & 15u), py = blockIdx.y * 16u + (threadIdx.x >> 4u);
    if (px >= width || py >= height) return;
    float tx = (float)px / (float)width - .5f, ty = (float)py / (float)height - .5f, tz = z_slice;
    uint32_t cur = 0u, depth = 0u;
    int child = children[(tx > 0.f) | ((ty > 0.f) << 1) | ((tz > 0.f) << 2)];
    // a tree built by the octree is at most depth_limit deep; the bound only
    // guards against a damaged child table
    while (child != -1 && depth < 16u) {
        cur = (uint32_t)child;
        float v = tx * 2.f;
        tx = (v - floorf(v)) - .5f;
        v = ty * 2.f;
        ty = (v - floorf(v)) - .5f;
        v = tz * 2.f;
        tz = (v - floorf(v)) - .5f;
        child = children[cur * 8u + ((tx > 0.f) | ((ty > 0.f) << 1) | ((tz > 0.f) << 2))];
        depth++;
    }
    float4 out;
    const float edge = .5f - 2.f * (float)(1u << depth) / (float)width;
    if (boxes && fmaxf(fabsf(tx), fabsf(ty)) >= edge) {
        out = make_float4(1.f, 0.f, 0.f, 1.f);
    } else {
        float cx = (tx + .5f) * (float)S + (float)P, cy = (ty + .5f) * (float)S + (float)P,
              cz = (tz + .5f) * (float)S + (float)P;
        const uint32_t ix = (uint32_t)(cx + .5f), iy = (uint32_t)(cy + .5f), iz = (uint32_t)(cz + .5f);
        cx -= (float)ix;
        cy -= (float)iy;
        cz -= (float)iz;
        const size_t n = (((size_t)cur * T + iz) * T + iy) * T + ix;
        const float v = interp_lagrange<T>(data, cx, cy, cz, n) + .5f;
        out = make_float4(v, v, v, 1.f);
    }
    raster[(size_t)py * width + px] = out;
}

__device__ __forceinline__ uint32_t channel(float v) { return (uint32_t)(fminf(fmaxf(v, 0.f), 1.f) * 255.f); }

__global__ void __launch_bounds__(256) k_encode_rgba8(const float4 *__restrict__ raster, uint32_t *__restrict__ pixels,
                                                      int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 v = raster[i];
    uint32_t p = 4294902015u;  // NaN: magenta
    if (!(isnan(v.x) || isnan(v.y) || isnan(v.z) || isnan(v.w)))
        p = (channel(v.w) << 24) | (channel(v.z) << 16) | (channel(v.y) << 8) | channel(v.x);
    pixels[i] = p;
}

// ___rgba(float) and ___rgba(float3) (image_process.h:22-38): a grey pixel
// per float, an RGB pixel per decoded compressed_float3 word
__global__ void __launch_bounds__(256) k_encode_rgba8_f32(const float *__restrict__ img, uint32_t *__restrict__ pixels,
                                                          int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float v = img[i];
    uint32_t p = 4294902015u;
    if (!isnan(v)) {
        const uint32_t c = channel(v);
        p = (255u << 24) | (c << 16) | (c << 8) | c;
    }
    pixels[i] = p;
}
__global__ void __launch_bounds__(256) k_encode_rgba8_cf3(const uint32_t *__restrict__ words,
                                                          uint32_t *__restrict__ pixels, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float3 v = cf3_unpack(words[i]);
    uint32_t p = 4294902015u;
    if (!(isnan(v.x) || isnan(v.y) || isnan(v.z)))
        p = (255u << 24) | (channel(v.z) << 16) | (channel(v.y) << 8) | channel(v.x);
    pixels[i] = p;
}

template <int S, int P>
static int launch_rasterise(const float *field, const am_mesh *mesh, uint32_t width, uint32_t height, float z_slice,
                            int32_t boxes, float *raster, void *stream)
{
    hipLaunchKernelGGL((k_rasterise<S, P>), dim3((width + 15u) / 16u, (height + 15u) / 16u), dim3(256), 0,
                       (hipStream_t)stream, field, mesh->children, width, height, z_slice, boxes,
                       reinterpret_cast<float4 *>(raster));
    return check_launch("k_rasterise");
}

}  // namespace am

extern "C" int am_rasterise_slice(const float *field, const am_mesh *mesh, const am_geometry *g, uint32_t width,
                                  uint32_t height, float z_slice, int32_t boxes, float *raster, void *stream)
{
    if (!field || !mesh || !g || !raster || !mesh->children)
        return am::set_error(AM_ERR_INVALID, "am_rasterise_slice: null argument");
    if (width == 0 || height == 0 || width > 32768u || height > 32768u)
        return am::set_error(AM_ERR_INVALID, "am_rasterise_slice: image %ux%u (1-32768 per side)", width, height);
    if (!(z_slice >= -.5f && z_slice < .5f))
        return am::set_error(AM_ERR_INVALID, "am_rasterise_slice: z_slice outside [-0.5, 0.5)");
    if (reinterpret_cast<uintptr_t>(raster) % 16u) return am::set_error(AM_ERR_MISALIGNED, "raster not 16-byte aligned");
    AM_DISPATCH_TILE(*g, am::launch_rasterise, field, mesh, width, height, z_slice, boxes, raster, stream);
}

extern "C" int am_encode_rgba8(const float *raster, uint32_t *pixels, int64_t n, void *stream)
{
    if (!raster || !pixels || n < 0) return am::set_error(AM_ERR_INVALID, "am_encode_rgba8: bad argument");
    if (reinterpret_cast<uintptr_t>(raster) % 16u) return am::set_error(AM_ERR_MISALIGNED, "raster not 16-byte aligned");
    if (n == 0) return AM_OK;
    hipLaunchKernelGGL(am::k_encode_rgba8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       reinterpret_cast<const float4 *>(raster), pixels, n);
    return am::check_launch("k_encode_rgba8");
}

extern "C" int am_encode_rgba8_f32(const float *image, uint32_t *pixels, int64_t n, void *stream)
{
    if (!image || !pixels || n < 0) return am::set_error(AM_ERR_INVALID, "am_encode_rgba8_f32: bad argument");
    if (n == 0) return AM_OK;
    hipLaunchKernelGGL(am::k_encode_rgba8_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       image, pixels, n);
    return am::check_launch("k_encode_rgba8_f32");
}

extern "C" int am_encode_rgba8_cf3(const uint32_t *words, uint32_t *pixels, int64_t n, void *stream)
{
    if (!words || !pixels || n < 0) return am::set_error(AM_ERR_INVALID, "am_encode_rgba8_cf3: bad argument");
    if (n == 0) return AM_OK;
    hipLaunchKernelGGL(am::k_encode_rgba8_cf3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       words, pixels, n);
    return am::check_launch("k_encode_rgba8_cf3");
}
