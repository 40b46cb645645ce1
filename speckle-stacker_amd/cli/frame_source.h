// frame_source.h -- frame stacks for the CLI without OpenCV: animated GIF
// (decoded and composited here), NumPy .npy stacks and headerless raw stacks.
//
// Channel selection follows the reference's Mat2Array (include/videoio.h:71-98):
// a colour mask r=4, g=2, b=1, i=7 picks channel round(log2(mask)) of the
// pixel in B,G,R order, clamped to the channels present, so "i" and "r" both
// pick red and a one-channel stack always yields its only channel.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "smip_mi355x.h"

namespace smip_cli {

inline std::vector<uint8_t> read_file(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("file open error: " + path);
    return std::vector<uint8_t>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

inline int channel_index(int mask, int channels)
{
    long c = mask > 0 ? std::lround(std::log2(static_cast<double>(mask))) : 0;
    return static_cast<int>(std::clamp(c, 0L, static_cast<long>(channels) - 1L));
}

class FrameSource {
public:
    virtual ~FrameSource() = default;
    virtual uint64_t nframes() const = 0;
    virtual uint64_t ncols() const = 0;
    virtual uint64_t nrows() const = 0;
    virtual int32_t dtype() const = 0; // SMIP_DTYPE_U8 / U16 / F32
    // copies frames [first, first+count) of the selected channel into out
    virtual void read(uint64_t first, uint64_t count, void* out) = 0;
    size_t frame_bytes() const
    {
        const size_t s = dtype() == SMIP_DTYPE_U8 ? 1 : dtype() == SMIP_DTYPE_U16 ? 2 : 4;
        return s * ncols() * nrows();
    }
};

// ---------------------------------------------------------------- GIF (89a/87a)
class GifSource : public FrameSource {
public:
    GifSource(const std::string& path, int colour_mask)
    {
        decode(read_file(path));
        m_channel = channel_index(colour_mask, 3);
    }
    uint64_t nframes() const override { return m_frames.size(); }
    uint64_t ncols() const override { return m_w; }
    uint64_t nrows() const override { return m_h; }
    int32_t dtype() const override { return SMIP_DTYPE_U8; }
    void read(uint64_t first, uint64_t count, void* out) override
    {
        auto* o = static_cast<uint8_t*>(out);
        const size_t n = (size_t)m_w * m_h;
        // stored pixels are R,G,B; channel indices count B,G,R
        const int rgb = 2 - m_channel;
        for (uint64_t f = 0; f < count; ++f) {
            const auto& fr = m_frames.at(first + f);
            for (size_t k = 0; k < n; ++k)
                o[f * n + k] = fr[3 * k + rgb];
        }
    }
    // composited R,G,B frame (tests)
    const std::vector<uint8_t>& rgb_frame(size_t i) const { return m_frames.at(i); }

private:
    struct Reader {
        const std::vector<uint8_t>& b;
        size_t p = 0;
        uint8_t u8()
        {
            if (p >= b.size())
                throw std::runtime_error("GIF: truncated file");
            return b[p++];
        }
        uint16_t u16()
        {
            const uint16_t lo = u8();
            return static_cast<uint16_t>(lo | (u8() << 8));
        }
        void bytes(uint8_t* dst, size_t n)
        {
            if (p + n > b.size())
                throw std::runtime_error("GIF: truncated file");
            std::memcpy(dst, b.data() + p, n);
            p += n;
        }
        std::vector<uint8_t> sub_blocks()
        {
            std::vector<uint8_t> d;
            for (uint8_t len = u8(); len; len = u8()) {
                const size_t o = d.size();
                d.resize(o + len);
                bytes(d.data() + o, len);
            }
            return d;
        }
    };

    static std::vector<uint8_t> lzw(const std::vector<uint8_t>& data, int min_size, size_t npix)
    {
        if (min_size < 2 || min_size > 11)
            throw std::runtime_error("GIF: bad LZW code size");
        const int clear = 1 << min_size, eoi = clear + 1;
        std::vector<uint16_t> prefix(4096);
        std::vector<uint8_t> suffix(4096), first(4096);
        std::vector<uint16_t> len(4096);
        for (int c = 0; c < clear; ++c) {
            suffix[c] = first[c] = static_cast<uint8_t>(c);
            len[c] = 1;
        }
        std::vector<uint8_t> out;
        out.reserve(npix);
        int size = min_size + 1, next = clear + 2, old = -1;
        uint32_t acc = 0;
        int bits = 0;
        size_t pos = 0;
        auto emit = [&](int code) {
            const size_t o = out.size();
            out.resize(o + len[code]);
            for (int c = code, k = len[code] - 1; k >= 0; --k, c = prefix[c])
                out[o + k] = suffix[c];
        };
        while (out.size() < npix) {
            while (bits < size && pos < data.size()) {
                acc |= static_cast<uint32_t>(data[pos++]) << bits;
                bits += 8;
            }
            if (bits < size)
                break;
            const int code = static_cast<int>(acc & ((1u << size) - 1));
            acc >>= size;
            bits -= size;
            if (code == clear) {
                size = min_size + 1;
                next = clear + 2;
                old = -1;
                continue;
            }
            if (code == eoi)
                break;
            if (old < 0) {
                if (code >= clear)
                    throw std::runtime_error("GIF: bad first LZW code");
                emit(code);
                old = code;
                continue;
            }
            if (code < next) {
                emit(code);
                if (next < 4096) {
                    prefix[next] = static_cast<uint16_t>(old);
                    suffix[next] = first[code];
                    first[next] = first[old];
                    len[next] = static_cast<uint16_t>(len[old] + 1);
                    ++next;
                }
            } else if (code == next && next < 4096) {
                prefix[next] = static_cast<uint16_t>(old);
                suffix[next] = first[old];
                first[next] = first[old];
                len[next] = static_cast<uint16_t>(len[old] + 1);
                ++next;
                emit(code);
            } else {
                throw std::runtime_error("GIF: bad LZW code");
            }
            if (next == (1 << size) && size < 12)
                ++size;
            old = code;
        }
        out.resize(npix, 0);
        return out;
    }

    void decode(const std::vector<uint8_t>& file)
    {
        Reader r { file };
        uint8_t sig[6];
        r.bytes(sig, 6);
        if (std::memcmp(sig, "GIF87a", 6) && std::memcmp(sig, "GIF89a", 6))
            throw std::runtime_error("not a GIF file");
        m_w = r.u16();
        m_h = r.u16();
        const uint8_t packed = r.u8();
        const uint8_t bg = r.u8();
        r.u8(); // pixel aspect
        std::vector<uint8_t> gct;
        if (packed & 0x80) {
            gct.resize(3u << ((packed & 7) + 1));
            r.bytes(gct.data(), gct.size());
        }
        const size_t n = (size_t)m_w * m_h;
        std::vector<uint8_t> canvas(3 * n, 0);
        if (!gct.empty() && 3u * bg + 2 < gct.size())
            for (size_t k = 0; k < n; ++k)
                std::memcpy(&canvas[3 * k], &gct[3u * bg], 3);
        const std::vector<uint8_t> background = canvas;
        int disposal = 0, transparent = -1;
        for (;;) {
            const uint8_t tag = r.u8();
            if (tag == 0x3B)
                break;
            if (tag == 0x21) {
                const uint8_t label = r.u8();
                if (label == 0xF9) {
                    const auto gce = r.sub_blocks();
                    if (gce.size() >= 4) {
                        disposal = (gce[0] >> 2) & 7;
                        transparent = (gce[0] & 1) ? gce[3] : -1;
                    }
                } else {
                    r.sub_blocks();
                }
                continue;
            }
            if (tag != 0x2C)
                throw std::runtime_error("GIF: unexpected block");
            const int left = r.u16(), top = r.u16(), iw = r.u16(), ih = r.u16();
            const uint8_t ip = r.u8();
            std::vector<uint8_t> lct;
            if (ip & 0x80) {
                lct.resize(3u << ((ip & 7) + 1));
                r.bytes(lct.data(), lct.size());
            }
            const auto& pal = lct.empty() ? gct : lct;
            const int min_size = r.u8();
            const auto idx = lzw(r.sub_blocks(), min_size, (size_t)iw * ih);
            std::vector<int> rows(ih);
            if (ip & 0x40) { // interlaced: passes 0/8, 4/8, 2/4, 1/2
                int k = 0;
                for (int pass = 0; pass < 4; ++pass) {
                    static const int start[4] = { 0, 4, 2, 1 }, step[4] = { 8, 8, 4, 2 };
                    for (int y = start[pass]; y < ih; y += step[pass])
                        rows[k++] = y;
                }
            } else {
                for (int y = 0; y < ih; ++y)
                    rows[y] = y;
            }
            const std::vector<uint8_t> before = canvas;
            for (int k = 0; k < ih; ++k) {
                const int y = top + rows[k];
                if (y < 0 || y >= (int)m_h)
                    continue;
                for (int x = 0; x < iw; ++x) {
                    const int cx = left + x;
                    const int c = idx[(size_t)k * iw + x];
                    if (cx >= (int)m_w || c == transparent || 3u * c + 2 >= pal.size())
                        continue;
                    std::memcpy(&canvas[3 * ((size_t)y * m_w + cx)], &pal[3u * c], 3);
                }
            }
            m_frames.push_back(canvas);
            if (disposal == 2) { // restore the frame's rectangle to the background
                for (int y = std::max(top, 0); y < std::min<int>(top + ih, m_h); ++y)
                    for (int x = std::max(left, 0); x < std::min<int>(left + iw, m_w); ++x)
                        std::memcpy(&canvas[3 * ((size_t)y * m_w + x)], &background[3 * ((size_t)y * m_w + x)], 3);
            } else if (disposal == 3) {
                canvas = before;
            }
            disposal = 0;
            transparent = -1;
        }
        if (m_frames.empty())
            throw std::runtime_error("GIF: no frames");
    }

    uint32_t m_w = 0, m_h = 0;
    int m_channel = 2;
    std::vector<std::vector<uint8_t>> m_frames;
};

// ------------------------------------------------------------ .npy and raw
class RawSource : public FrameSource {
public:
    RawSource(const std::string& path, uint64_t offset, uint64_t ncols, uint64_t nrows, int32_t dtype)
        : m_f(path, std::ios::binary)
        , m_off(offset)
        , m_X(ncols)
        , m_Y(nrows)
        , m_dtype(dtype)
    {
        if (!m_f)
            throw std::runtime_error("file open error: " + path);
        m_f.seekg(0, std::ios::end);
        const uint64_t size = static_cast<uint64_t>(m_f.tellg());
        if (size < offset || frame_bytes() == 0)
            throw std::runtime_error("raw stack: bad geometry");
        m_n = (size - offset) / frame_bytes();
    }
    uint64_t nframes() const override { return m_n; }
    uint64_t ncols() const override { return m_X; }
    uint64_t nrows() const override { return m_Y; }
    int32_t dtype() const override { return m_dtype; }
    void read(uint64_t first, uint64_t count, void* out) override
    {
        m_f.seekg(static_cast<std::streamoff>(m_off + first * frame_bytes()));
        m_f.read(static_cast<char*>(out), static_cast<std::streamsize>(count * frame_bytes()));
        if (!m_f)
            throw std::runtime_error("raw stack: short read");
    }

private:
    std::ifstream m_f;
    uint64_t m_off, m_X, m_Y, m_n = 0;
    int32_t m_dtype;
};

// NumPy format 1.0/2.0/3.0 header: little-endian u1 / u2 / f4, C order, shape
// (frames, rows, cols) or (rows, cols)
inline std::unique_ptr<FrameSource> open_npy(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw std::runtime_error("file open error: " + path);
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "\x93NUMPY", 6))
        throw std::runtime_error("not a .npy file: " + path);
    uint32_t hl = 0;
    if (magic[6] == 1) {
        uint8_t b[2];
        f.read(reinterpret_cast<char*>(b), 2);
        hl = b[0] | (b[1] << 8);
    } else {
        uint8_t b[4];
        f.read(reinterpret_cast<char*>(b), 4);
        hl = b[0] | (b[1] << 8) | (b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    std::string h(hl, '\0');
    f.read(h.data(), hl);
    const uint64_t offset = (magic[6] == 1 ? 10 : 12) + (uint64_t)hl;
    auto field = [&](const std::string& key) {
        const auto k = h.find("'" + key + "'");
        if (k == std::string::npos)
            throw std::runtime_error(".npy header lacks " + key);
        return h.substr(h.find(':', k) + 1);
    };
    const std::string descr = field("descr");
    int32_t dtype = 0;
    if (descr.find("u1") != std::string::npos)
        dtype = SMIP_DTYPE_U8;
    else if (descr.find("<u2") != std::string::npos)
        dtype = SMIP_DTYPE_U16;
    else if (descr.find("<f4") != std::string::npos)
        dtype = SMIP_DTYPE_F32;
    else
        throw std::runtime_error(".npy dtype must be uint8, little-endian uint16 or float32");
    if (field("fortran_order").find("True") < field("fortran_order").find(','))
        throw std::runtime_error(".npy stack must be in C order");
    std::string shape = field("shape");
    shape = shape.substr(shape.find('(') + 1, shape.find(')') - shape.find('(') - 1);
    std::vector<uint64_t> dims;
    for (size_t p = 0; p < shape.size();) {
        while (p < shape.size() && !std::isdigit(static_cast<unsigned char>(shape[p])))
            ++p;
        if (p >= shape.size())
            break;
        dims.push_back(std::stoull(shape.substr(p)));
        while (p < shape.size() && std::isdigit(static_cast<unsigned char>(shape[p])))
            ++p;
    }
    if (dims.size() != 2 && dims.size() != 3)
        throw std::runtime_error(".npy stack must have shape (frames, rows, cols) or (rows, cols)");
    const uint64_t Y = dims[dims.size() - 2], X = dims.back();
    return std::make_unique<RawSource>(path, offset, X, Y, dtype);
}

} // namespace smip_cli
