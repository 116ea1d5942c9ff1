// Image output of the plugins: the reference's rasterise_to_image
// (AMR.h:1355-1364) and save_image (image_process.h:55-62), over the C ABI
// (am_rasterise_slice, am_encode_rgba8, am_write_png).
//
//   smart_gpu_cpu_buffer<float4> temp(1024u * 1024u);   // any 16-byte pixel
//   smart_gpu_cpu_buffer<uint32_t> temp2(1024u * 1024u);
//   rasterise_to_image(amr.sim_data.new_field, temp, amr, 1024u, 1024u);
//   save_image(temp2, temp, 1024u, 1024u, "test_0.png");
//
// The plugin supplies the tile geometry (`sim_data.geom`).  Unlike the
// reference, z_slice must lie in [-1/2, 1/2) (a point outside the root cube
// would read outside its tile) and the tile-edge boxes can be switched off.
#pragma once
#include <stdint.h>

#include <stdexcept>

#include "../adaptivemesh_host.h"
#include "AMR.hpp"

namespace am {

// a host-side name for the 16-byte raster pixel (HIP's float4 works alike)
struct rgba_f32 {
    float r, g, b, a;
};

// the device side of a smart_gpu_cpu_buffer is made on first use
template <class B>
auto ensure_device(B &b, int) -> decltype(b.ensure_gpu(), void())
{
    b.ensure_gpu();
}
template <class B>
void ensure_device(B &, long)
{
}

template <class AMR_t, class FieldBuf, class RasterBuf>
void rasterise_to_image(const FieldBuf &data, RasterBuf &raster_image, AMR_t &amr, uint32_t width, uint32_t height,
                        float z_slice = 1e-9f, bool boxes = true, void *stream = nullptr)
{
    static_assert(sizeof(*raster_image.gpu_buffer_ptr) == 16, "raster pixels are 4 floats");
    if (raster_image.total_size() < (size_t)width * height * 16u) throw std::runtime_error("raster buffer too small");
    ensure_device(raster_image, 0);
    const am_mesh m = amr.mesh_all();
    check(am_rasterise_slice(static_cast<const float *>(data.gpu_buffer_ptr), &m, &amr.sim_data.geom, width, height,
                             z_slice, boxes ? 1 : 0, reinterpret_cast<float *>(raster_image.gpu_buffer_ptr), stream),
          "rasterise_to_image");
}

// ___rgba per pixel type (image_process.h:9-38): 4-float raster pixels,
// scalar fields (grey) and compressed_float3 words (RGB)
inline int encode_rgba8(const float *img, uint32_t *px, int64_t n, void *stream)
{
    return am_encode_rgba8_f32(img, px, n, stream);
}
inline int encode_rgba8(const compressed_float3 *img, uint32_t *px, int64_t n, void *stream)
{
    return am_encode_rgba8_cf3(reinterpret_cast<const uint32_t *>(img), px, n, stream);
}
template <class Pixel16>
int encode_rgba8(const Pixel16 *img, uint32_t *px, int64_t n, void *stream)
{
    static_assert(sizeof(Pixel16) == 16, "save_image encodes float, compressed_float3 or 4-float pixels");
    return am_encode_rgba8(reinterpret_cast<const float *>(img), px, n, stream);
}

// temp: RGBA8 staging with a device side (smart_gpu_cpu_buffer<uint32_t>);
// image: a raster, a float field or a word array (any buffer with
// gpu_buffer_ptr), its first width * height elements as rows of `width`
template <class PixelBuf, class RasterBuf>
void save_image(PixelBuf &temp, const RasterBuf &image, uint32_t width, uint32_t height, const char *filename,
                void *stream = nullptr)
{
    const size_t n = (size_t)width * height;
    if (temp.dedicated_len < n) throw std::runtime_error("save_image: pixel buffer too small");
    temp.ensure_gpu();
    check(encode_rgba8(image.gpu_buffer_ptr, temp.gpu_buffer_ptr, (int64_t)n, stream), "save_image");
    temp.copy_to_cpu(stream);
    check(am_write_png(filename, temp.cpu_buffer_ptr, width, height), "save_image");
}

// A slice of `field` straight to host RGBA8 pixels (the drivers' _raster
// calls): scratch raster and pixel buffers live for the call only.
template <class AMR_t>
void raster_to_host(AMR_t &amr, const float *field, uint32_t width, uint32_t height, float z_slice, bool boxes,
                    uint32_t *host, void *stream)
{
    smart_gpu_buffer<float> raster(4 * (size_t)width * height);
    smart_gpu_buffer<uint32_t> px((size_t)width * height);
    int rc = AM_OK;
    {
        const am_mesh m = amr.mesh_all();
        rc = am_rasterise_slice(field, &m, &amr.sim_data.geom, width, height, z_slice, boxes ? 1 : 0,
                                raster.gpu_buffer_ptr, stream);
        if (!rc) rc = am_encode_rgba8(raster.gpu_buffer_ptr, px.gpu_buffer_ptr, (int64_t)width * height, stream);
        if (!rc) rc = am_memcpy_d2h(host, px.gpu_buffer_ptr, (size_t)width * height * 4u, stream);
    }
    if (stream) am_stream_sync(stream);
    raster.destroy();
    px.destroy();
    check(rc, "raster");
}

}  // namespace am
