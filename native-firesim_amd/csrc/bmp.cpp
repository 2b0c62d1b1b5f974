// bmp.cpp - 32-bit top-down BGRA BMP output.
//
// Same file format as the reference's writeBmp (src/main.cpp:2511-2536):
// BITMAPFILEHEADER (14 bytes) + BITMAPINFOHEADER (40 bytes), biHeight negative
// (top-down rows), 32 bits per pixel, BI_RGB, pixels as little-endian
// 0xAARRGGBB words. Written field by field so the layout does not depend on
// Windows headers or struct packing.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/firesim.h"

namespace {
void put16(std::vector<unsigned char>& b, uint16_t v) {
    b.push_back(static_cast<unsigned char>(v & 0xff));
    b.push_back(static_cast<unsigned char>(v >> 8));
}
void put32(std::vector<unsigned char>& b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back(static_cast<unsigned char>((v >> (8 * i)) & 0xff));
}
}  // namespace

extern "C" int fs_write_bmp(const char* path, const uint32_t* bgra, int32_t width, int32_t height) {
    if (!path || !bgra || width <= 0 || height <= 0) return FS_ERR_INVALID_ARGUMENT;
    const uint64_t pixelBytes = static_cast<uint64_t>(width) * static_cast<uint64_t>(height) * 4u;
    if (pixelBytes > 0xFFFFFFFFull - 54u) return FS_ERR_INVALID_ARGUMENT;
    std::vector<unsigned char> h;
    h.reserve(54);
    put16(h, 0x4d42);                                   // bfType "BM"
    put32(h, static_cast<uint32_t>(54 + pixelBytes));   // bfSize
    put16(h, 0);                                        // bfReserved1
    put16(h, 0);                                        // bfReserved2
    put32(h, 54);                                       // bfOffBits
    put32(h, 40);                                       // biSize
    put32(h, static_cast<uint32_t>(width));             // biWidth
    put32(h, static_cast<uint32_t>(-height));           // biHeight (top-down)
    put16(h, 1);                                        // biPlanes
    put16(h, 32);                                       // biBitCount
    put32(h, 0);                                        // biCompression = BI_RGB
    put32(h, static_cast<uint32_t>(pixelBytes));        // biSizeImage
    put32(h, 0);                                        // biXPelsPerMeter
    put32(h, 0);                                        // biYPelsPerMeter
    put32(h, 0);                                        // biClrUsed
    put32(h, 0);                                        // biClrImportant
    FILE* f = std::fopen(path, "wb");
    if (!f) return FS_ERR_IO;
    bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    std::vector<unsigned char> row(static_cast<size_t>(width) * 4);
    for (int32_t y = 0; ok && y < height; ++y) {
        for (int32_t x = 0; x < width; ++x) {
            const uint32_t v = bgra[static_cast<size_t>(y) * width + x];
            for (int i = 0; i < 4; ++i) row[static_cast<size_t>(x) * 4 + i] = static_cast<unsigned char>((v >> (8 * i)) & 0xff);
        }
        ok = std::fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? FS_OK : FS_ERR_IO;
}
