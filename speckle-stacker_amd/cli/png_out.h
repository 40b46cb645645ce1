// png_out.h -- 16-bit PNG output of the CLI's nine images, with the value
// mapping of the reference's Array2Mat / mapToColor (include/videoio.h:100-211):
//  * grey:   sample = (uint16)(x * 65535), x = converter(value)
//  * colour: heat map of x clamped to [0, 1]: blue up (x < 1/4), green up and
//            blue down (< 1/2), red up and green down (< 3/4), red with blue up
//  * signed symmetry: pixel (j, i) shows the array at signed index
//    (j + min_x, i + min_y), i.e. zero frequency in the centre.
// The grey conversion of values above 1 (the power spectrum is scaled to a
// maximum of 1000 before it is written, smip-cli/main.cpp:365-366) keeps the
// x86-64 behaviour of the reference's cast: truncation to a 32-bit integer,
// then its low 16 bits.
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

namespace smip_cli {

inline uint16_t grey16(double x)
{
    const double v = x * 65535.0;
    if (!(v > -2147483648.0 && v < 2147483648.0))
        return 0; // the 32-bit conversion's "integer indefinite" value has low bits 0
    return static_cast<uint16_t>(static_cast<int32_t>(v));
}

struct Rgb16 {
    uint16_t r, g, b;
};

inline Rgb16 heat16(double value)
{
    const double m = 65535.0;
    value = std::clamp(value, 0., 1.);
    if (value < 0.25)
        return { 0, 0, static_cast<uint16_t>(m * (4 * value)) };
    if (value < 0.5)
        return { 0, static_cast<uint16_t>(m * (4 * value - 1)), static_cast<uint16_t>(m * (2 - 4 * value)) };
    if (value < 0.75)
        return { static_cast<uint16_t>(m * (4 * value - 2)), static_cast<uint16_t>(m * (3 - 4 * value)), 0 };
    return { 65535, 0, static_cast<uint16_t>(m * (4 * value - 3)) };
}

inline void put_be32(std::vector<uint8_t>& o, uint32_t v)
{
    o.push_back(static_cast<uint8_t>(v >> 24));
    o.push_back(static_cast<uint8_t>(v >> 16));
    o.push_back(static_cast<uint8_t>(v >> 8));
    o.push_back(static_cast<uint8_t>(v));
}

inline void chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data)
{
    put_be32(o, static_cast<uint32_t>(data.size()));
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put_be32(o, static_cast<uint32_t>(crc32(0L, o.data() + start, static_cast<uInt>(o.size() - start))));
}

// samples: rows * cols * channels 16-bit values, row-major; channels 1 (grey) or 3 (RGB)
inline void write_png16(const std::string& path, uint32_t cols, uint32_t rows, int channels,
    const std::vector<uint16_t>& samples)
{
    std::vector<uint8_t> raw;
    raw.reserve(static_cast<size_t>(rows) * (1 + 2 * cols * channels));
    for (uint32_t y = 0; y < rows; ++y) {
        raw.push_back(0); // filter: none
        for (uint32_t x = 0; x < cols * (uint32_t)channels; ++x) {
            const uint16_t v = samples[static_cast<size_t>(y) * cols * channels + x];
            raw.push_back(static_cast<uint8_t>(v >> 8));
            raw.push_back(static_cast<uint8_t>(v));
        }
    }
    uLongf zlen = compressBound(static_cast<uLong>(raw.size()));
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), static_cast<uLong>(raw.size()), 6) != Z_OK)
        throw std::runtime_error("PNG compression failed");
    z.resize(zlen);
    std::vector<uint8_t> png = { 0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n' };
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, cols);
    put_be32(ihdr, rows);
    ihdr.push_back(16);
    ihdr.push_back(channels == 3 ? 2 : 0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    ihdr.push_back(0);
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f)
        throw std::runtime_error("cannot write " + path);
    const size_t w = std::fwrite(png.data(), 1, png.size(), f);
    std::fclose(f);
    if (w != png.size())
        throw std::runtime_error("short write to " + path);
}

// value(col, row) -> x for every pixel; centred = signed symmetry
inline void save_image(const std::string& path, uint32_t cols, uint32_t rows, bool colour, bool centred,
    const std::function<double(int64_t)>& value_at)
{
    std::vector<uint16_t> s(static_cast<size_t>(cols) * rows * (colour ? 3 : 1));
    const int64_t mx = -static_cast<int64_t>(cols >> 1), my = -static_cast<int64_t>(rows >> 1);
    for (uint32_t i = 0; i < rows; ++i)
        for (uint32_t j = 0; j < cols; ++j) {
            int64_t c = j, r = i;
            if (centred) {
                c = j + mx;
                r = i + my;
                c = c < 0 ? c + cols : c;
                r = r < 0 ? r + rows : r;
            }
            const double x = value_at(r * cols + c);
            const size_t o = static_cast<size_t>(i) * cols + j;
            if (colour) {
                const Rgb16 k = heat16(x);
                s[3 * o] = k.r;
                s[3 * o + 1] = k.g;
                s[3 * o + 2] = k.b;
            } else {
                s[o] = grey16(x);
            }
        }
    write_png16(path, cols, rows, colour ? 3 : 1, s);
}

} // namespace smip_cli
