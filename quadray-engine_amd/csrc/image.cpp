/*
 * image.cpp -- BMP texture files and frame dumps, the file formats of
 * QuadRay's core/engine/rtimag.cpp (load_image 47-186, save_image 204-361),
 * restated: 24-bit uncompressed images, little-endian header fields read and
 * written byte by byte, rows padded to 4 bytes, pixels 0x00RRGGBB.
 */
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "image.h"

namespace qr {

namespace {

std::mutex g_mu;
std::string g_root = "../";     /* QuadRay's RT_PATH default */

uint32_t le16(const unsigned char *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8);
}

uint32_t le32(const unsigned char *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) |
           ((uint32_t)p[3] << 24);
}

void put16(unsigned char *p, uint32_t v)
{
    p[0] = (unsigned char)(v & 0xFF);
    p[1] = (unsigned char)((v >> 8) & 0xFF);
}

void put32(unsigned char *p, uint32_t v)
{
    for (int k = 0; k < 4; k++) p[k] = (unsigned char)((v >> (8 * k)) & 0xFF);
}

}  /* namespace */

void set_data_root(const std::string &root)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_root = root;
}

std::string data_root()
{
    std::lock_guard<std::mutex> lk(g_mu);
    return g_root;
}

/* QuadRay accepts signature "BM", one plane, 24 bits per pixel and no
 * compression; the pixel data starts at the offset stored at byte 10; rows
 * follow one another as stored (a positive height, bottom-up in the BMP
 * sense, keeps the file's first row as texel row 0), each padded to 4 bytes,
 * the last row's padding included; the texel's top byte is 0.  A width or
 * height of 0, which QuadRay takes for an empty texture, fails here, and a
 * negative width is read as its magnitude */
bool bmp_load(const std::string &path, std::vector<uint32_t> &texels, int &x_dim, int &y_dim)
{
    FILE *f = fopen(path.c_str(), "rb");
    if (f == nullptr) return false;
    unsigned char h[34];
    bool ok = fread(h, 1, sizeof(h), f) == sizeof(h);
    uint32_t off = 0, planes = 0, depth = 0, comp = 0, sig = 0;
    int32_t w = 0, hgt = 0;
    if (ok)
    {
        sig = le16(h);
        off = le32(h + 10);
        w = (int32_t)le32(h + 18);
        hgt = (int32_t)le32(h + 22);
        planes = le16(h + 26);
        depth = le16(h + 28);
        comp = le32(h + 30);
        ok = sig == 0x4D42 && planes == 1 && depth == 24 && comp == 0 && w != 0 && hgt != 0;
    }
    const int64_t xd = w < 0 ? -(int64_t)w : w, yd = hgt < 0 ? -(int64_t)hgt : hgt;
    ok = ok && xd <= 65536 && yd <= 65536 && fseek(f, (long)off, SEEK_SET) == 0;
    if (ok)
    {
        texels.assign((size_t)(xd * yd), 0u);
        const size_t row = (size_t)xd * 3, pad = ((row + 3) / 4) * 4 - row;
        std::vector<unsigned char> buf(row + pad);
        for (int64_t y = 0; y < yd && ok; y++)
        {
            ok = fread(buf.data(), 1, row + pad, f) == row + pad;
            for (int64_t x = 0; x < xd && ok; x++)
            {
                const unsigned char *p = buf.data() + 3 * x;
                texels[(size_t)(y * xd + x)] =
                    (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16);
            }
        }
    }
    fclose(f);
    if (!ok)
    {
        texels.clear();
        return false;
    }
    x_dim = (int)xd;
    y_dim = (int)yd;
    return true;
}

/* QuadRay's header: size fields count 3 bytes per pixel without the row
 * padding, 4000 pixels per metre, x_dim and y_dim stored as given (a negative
 * y_dim marks top-down rows, as frames are saved); row y of the source at
 * px + y * stride, padding bytes 0 */
bool bmp_save(const std::string &path, const uint32_t *px, int x_dim, int y_dim,
              ptrdiff_t stride)
{
    const uint32_t w = (uint32_t)(x_dim < 0 ? -x_dim : x_dim);
    const uint32_t hh = (uint32_t)(y_dim < 0 ? -y_dim : y_dim);
    FILE *f = fopen(path.c_str(), "wb");
    if (f == nullptr) return false;
    unsigned char h[54];
    memset(h, 0, sizeof(h));
    const uint32_t size = 54 + w * hh * 3;
    put16(h, 0x4D42);
    put32(h + 2, size);
    put32(h + 10, 54);
    put32(h + 14, 40);
    put32(h + 18, (uint32_t)x_dim);
    put32(h + 22, (uint32_t)y_dim);
    put16(h + 26, 1);
    put16(h + 28, 24);
    put32(h + 34, size - 54);
    put32(h + 38, 4000);
    put32(h + 42, 4000);
    bool ok = fwrite(h, 1, sizeof(h), f) == sizeof(h);
    const size_t row = (size_t)w * 3, pad = ((row + 3) / 4) * 4 - row;
    std::vector<unsigned char> buf(row + pad, 0);
    for (uint32_t y = 0; y < hh && ok; y++)
    {
        const uint32_t *src = px + (ptrdiff_t)y * stride;
        for (uint32_t x = 0; x < w; x++)
        {
            buf[3 * x + 0] = (unsigned char)(src[x] & 0xFF);
            buf[3 * x + 1] = (unsigned char)((src[x] >> 8) & 0xFF);
            buf[3 * x + 2] = (unsigned char)((src[x] >> 16) & 0xFF);
        }
        ok = fwrite(buf.data(), 1, row + pad, f) == row + pad;
    }
    ok = fclose(f) == 0 && ok;
    return ok;
}

/* rt_Scene::save_frame's name (core/engine/engine.cpp:3863-3888): scrXXX.bmp
 * below 1000, scrXXX-Y.bmp from there */
std::string frame_name(int index)
{
    char name[20];
    if (index < 1000) strncpy(name, "scrXXX.bmp", sizeof(name));
    else strncpy(name, "scrXXX-Y.bmp", sizeof(name));
    name[5] = (char)('0' + index % 10);
    index /= 10;
    name[4] = (char)('0' + index % 10);
    index /= 10;
    name[3] = (char)('0' + index % 10);
    if (index >= 10)
    {
        index /= 10;
        index -= 1;
        name[7] = (char)('0' + index % 10);
    }
    return name;
}

}  /* namespace qr */
