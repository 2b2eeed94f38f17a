// 8-bit RGBA PNG output of a host RGBA32F image (SavePngImage,
// src/FileIO.cpp:167-194, which writes through stb_image_write with a
// vertical flip).  Filter type 0 on every row and zlib's default
// compression; the pixels are the reference's, the compressed bytes need not
// be.
#include "internal.hpp"

#include <zlib.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

void put32(std::vector<uint8_t> &v, uint32_t x)
{
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}

void chunk(std::vector<uint8_t> &out, const char *type, const std::vector<uint8_t> &data)
{
    put32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put32(out, (uint32_t)crc32(0L, out.data() + start, (uInt)(out.size() - start)));
}

// std::filesystem::path::replace_extension(".png")
std::string png_path(std::string p)
{
    for (char &c : p)
        if (c == '\\') c = '/';
    const size_t slash = p.find_last_of('/');
    const size_t dot = p.find_last_of('.');
    const size_t name = slash == std::string::npos ? 0 : slash + 1;
    // a leading dot names a file ("/a/.b" keeps ".b" as its stem)
    if (dot != std::string::npos && dot > name && dot + 1 <= p.size()) p = p.substr(0, dot);
    return p + ".png";
}

bool make_dirs(const std::string &file)
{
    const size_t slash = file.find_last_of('/');
    if (slash == std::string::npos || slash == 0) return true;
    const std::string dir = file.substr(0, slash);
    for (size_t i = 1; i <= dir.size(); ++i) {
        if (i == dir.size() || dir[i] == '/') {
            const std::string part = dir.substr(0, i);
            if (mkdir(part.c_str(), 0777) != 0 && errno != EEXIST) return false;
        }
    }
    return true;
}

uint8_t to8(float v)
{
    const float c = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);  // glm::clamp; NaN passes through as in min/max
    return (uint8_t)(c * 255.0f + 0.5f);
}

}  // namespace

extern "C" int ymrt_save_png(const char *path, const float *rgba, uint32_t width, uint32_t height)
{
    YM_REQUIRE(path != nullptr && path[0] != '\0', "empty path");
    YM_REQUIRE(width > 0 && height > 0 && rgba != nullptr, "empty image");
    YM_REQUIRE((uint64_t)width * height <= (1ull << 28), "image too large");
    const std::string file = png_path(path);
    YM_REQUIRE(make_dirs(file), "cannot create the folders of " + file);
    const size_t stride = (size_t)width * 4 + 1;
    std::vector<uint8_t> raw(stride * height);
    for (uint32_t y = 0; y < height; ++y) {
        uint8_t *row = &raw[(size_t)y * stride];
        row[0] = 0;
        const float *src = rgba + (size_t)(height - 1 - y) * width * 4;  // flipped on write
        for (size_t i = 0; i < (size_t)width * 4; ++i) row[1 + i] = to8(src[i]);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    YM_REQUIRE(compress(z.data(), &zlen, raw.data(), (uLong)raw.size()) == Z_OK, "compression failed");
    z.resize(zlen);
    std::vector<uint8_t> png = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ihdr;
    put32(ihdr, width);
    put32(ihdr, height);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, filter set 0, no interlace
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE *f = std::fopen(file.c_str(), "wb");
    YM_REQUIRE(f != nullptr, "cannot open " + file);
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    YM_REQUIRE(std::fclose(f) == 0 && ok, "cannot write " + file);
    return YMRT_OK;
}
