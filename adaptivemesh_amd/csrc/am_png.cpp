// PNG writer for RGBA8 images, the part of lodepng the reference uses
// (lodepng_encode32_file, called by save_image, image_process.h:55-62).
// The image data goes into stored (uncompressed) deflate blocks inside one
// zlib stream, so the writer needs no compressor: a 1024x1024 image is a
// 4 MB file that any PNG reader decodes.  Chunks and checksums follow the
// PNG specification (ISO/IEC 15948): CRC-32 over type + data, Adler-32 over
// the raw scanlines, each scanline prefixed with filter type 0.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "am_common.hpp"

namespace {

struct crc_table {
    uint32_t t[256];
    crc_table()
    {
        for (uint32_t n = 0; n < 256; n++) {
            uint32_t c = n;
            for (int k = 0; k < 8; k++) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[n] = c;
        }
    }
};

uint32_t crc32(uint32_t crc, const unsigned char *p, size_t n)
{
    static const crc_table table;  // built once, thread-safe
    crc = ~crc;
    for (size_t i = 0; i < n; i++) crc = table.t[(crc ^ p[i]) & 255u] ^ (crc >> 8);
    return ~crc;
}

void put_be32(std::vector<unsigned char> &o, uint32_t v)
{
    o.push_back((unsigned char)(v >> 24));
    o.push_back((unsigned char)(v >> 16));
    o.push_back((unsigned char)(v >> 8));
    o.push_back((unsigned char)v);
}

void chunk(std::vector<unsigned char> &o, const char *type, const std::vector<unsigned char> &data)
{
    put_be32(o, (uint32_t)data.size());
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put_be32(o, crc32(0u, o.data() + start, o.size() - start));
}

}  // namespace

extern "C" int am_write_png(const char *path, const uint32_t *pixels, uint32_t width, uint32_t height)
{
    if (!path || !pixels || width == 0 || height == 0 || width > 32768u || height > 32768u)
        return am::set_error(AM_ERR_INVALID, "am_write_png: bad argument");
    // raw scanlines: filter byte 0, then RGBA bytes (pixel = R | G<<8 | B<<16 | A<<24)
    const size_t row = (size_t)width * 4u + 1u, raw_n = row * height;
    std::vector<unsigned char> raw(raw_n);
    for (uint32_t y = 0; y < height; y++) {
        raw[y * row] = 0;
        for (uint32_t x = 0; x < width; x++) {
            const uint32_t p = pixels[(size_t)y * width + x];
            unsigned char *d = &raw[y * row + 1 + 4 * (size_t)x];
            d[0] = (unsigned char)p;
            d[1] = (unsigned char)(p >> 8);
            d[2] = (unsigned char)(p >> 16);
            d[3] = (unsigned char)(p >> 24);
        }
    }
    // zlib stream: header, stored blocks of at most 65535 bytes, Adler-32
    std::vector<unsigned char> z;
    z.reserve(raw_n + raw_n / 65535u * 5u + 16u);
    z.push_back(0x78);
    z.push_back(0x01);
    for (size_t at = 0; at < raw_n;) {
        const size_t len = std::min<size_t>(65535u, raw_n - at);
        z.push_back(at + len == raw_n ? 1 : 0);
        z.push_back((unsigned char)len);
        z.push_back((unsigned char)(len >> 8));
        z.push_back((unsigned char)~len);
        z.push_back((unsigned char)(~len >> 8));
        z.insert(z.end(), raw.begin() + at, raw.begin() + at + len);
        at += len;
    }
    uint32_t a = 1, b = 0;
    for (size_t i = 0; i < raw_n; i++) {
        a = (a + raw[i]) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);

    std::vector<unsigned char> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, width);
    put_be32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, filter method 0, no interlace
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});

    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return am::set_error(AM_ERR_IO, "am_write_png: cannot create %s", tmp.c_str());
    const bool ok = fwrite(out.data(), 1, out.size(), f) == out.size();
    if (fclose(f) != 0 || !ok || rename(tmp.c_str(), path) != 0) {
        remove(tmp.c_str());
        return am::set_error(AM_ERR_IO, "am_write_png: cannot write %s", path);
    }
    return AM_OK;
}
