// frame_source.h -- frame stacks for the CLI without OpenCV: animated GIF
// (decoded and composited here), PNG images and numbered PNG sequences (decoded
// here with zlib), NumPy .npy stacks and headerless raw stacks.
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

#include <sys/stat.h>
#include <zlib.h>

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

// ------------------------------------------------------- PNG and PNG sequences
// One PNG file, or a numbered sequence read as cv::VideoCapture's image-sequence
// backend reads one (src/videoio.cpp:13-20 opens whatever VideoCapture opens): a
// printf pattern such as "frame_%04d.png" starts at the first index in [0, 1000]
// whose file exists; a file name whose stem ends in digits ("frame_0007.png")
// starts at that number with the same zero padding; either continues until an
// index has no file.  A name without digits is a single frame.  Every frame
// must have the first one's size, colour type and bit depth.
//
// Decoding: all colour types (grey, grey + alpha, RGB, RGB + alpha, palette),
// bit depths 1-16, Adam7 interlacing and the five row filters; chunk CRCs are
// checked.  Grey samples below 8 bits are scaled to 0..255 as image readers do;
// 16-bit images give uint16 frames (the reference's Mat2Array reads 8-bit
// images only, include/videoio.h:75), others uint8.  Channel selection: a grey
// image always yields its grey value; colour images follow channel_index over
// B,G,R (alpha ignored).
class PngImage {
public:
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> plte; // R,G,B triples

    bool grey() const { return ctype == 0 || ctype == 4; }
    int channels() const { return ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4; }

    // Decodes the file; `sel` (an index in B,G,R order) picks the channel of a
    // colour image; writes w*h samples (uint8, or uint16 when depth == 16).
    void decode(const std::string& path, int sel, void* out)
    {
        const std::vector<uint8_t> f = read_file(path);
        std::vector<uint8_t> z = parse(f, path);
        const int spp = channels(); // samples per pixel in the file
        const size_t bpp_bits = (size_t)spp * depth;
        const size_t bpp = std::max<size_t>(1, bpp_bits / 8); // filter byte distance
        // passes: (x0, y0, dx, dy); one full pass when not interlaced
        static const int adam7[7][4] = { { 0, 0, 8, 8 }, { 4, 0, 8, 8 }, { 0, 4, 4, 8 }, { 2, 0, 4, 4 },
            { 0, 2, 2, 4 }, { 1, 0, 2, 2 }, { 0, 1, 1, 2 } };
        const int full[1][4] = { { 0, 0, 1, 1 } };
        const int npass = interlace ? 7 : 1;
        // inflated size
        size_t need = 0;
        for (int q = 0; q < npass; ++q) {
            const int* P = interlace ? adam7[q] : full[0];
            const size_t pw = w > (uint32_t)P[0] ? (w - P[0] + P[2] - 1) / P[2] : 0;
            const size_t ph = h > (uint32_t)P[1] ? (h - P[1] + P[3] - 1) / P[3] : 0;
            if (pw && ph)
                need += ph * (1 + (pw * bpp_bits + 7) / 8);
        }
        std::vector<uint8_t> raw(need);
        inflate_all(z, raw, path);
        size_t pos = 0;
        std::vector<uint8_t> prev, cur;
        for (int q = 0; q < npass; ++q) {
            const int* P = interlace ? adam7[q] : full[0];
            const size_t pw = w > (uint32_t)P[0] ? (w - P[0] + P[2] - 1) / P[2] : 0;
            const size_t ph = h > (uint32_t)P[1] ? (h - P[1] + P[3] - 1) / P[3] : 0;
            if (!pw || !ph)
                continue;
            const size_t rb = (pw * bpp_bits + 7) / 8;
            prev.assign(rb, 0);
            cur.resize(rb);
            for (size_t y = 0; y < ph; ++y) {
                const uint8_t ft = raw[pos++];
                std::memcpy(cur.data(), raw.data() + pos, rb);
                pos += rb;
                unfilter(ft, cur, prev, bpp, path);
                for (size_t x = 0; x < pw; ++x)
                    store(cur, x, spp, sel, out, (size_t)(P[1] + y * P[3]) * w + P[0] + x * P[2]);
                std::swap(prev, cur);
            }
        }
    }

    // header of the file only (IHDR, PLTE)
    void header(const std::string& path) { parse(read_file(path), path); }

private:
    static uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

    std::vector<uint8_t> parse(const std::vector<uint8_t>& f, const std::string& path)
    {
        static const uint8_t sig[8] = { 0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n' };
        if (f.size() < 8 || std::memcmp(f.data(), sig, 8))
            throw std::runtime_error("PNG: bad signature: " + path);
        std::vector<uint8_t> z;
        bool ihdr = false, iend = false;
        for (size_t p = 8; p + 12 <= f.size() && !iend;) {
            const uint32_t len = be32(&f[p]);
            if (p + 12 + (size_t)len > f.size())
                throw std::runtime_error("PNG: truncated chunk: " + path);
            const uint8_t* type = &f[p + 4];
            const uint8_t* body = &f[p + 8];
            const uint32_t crc = (uint32_t)crc32(crc32(0L, Z_NULL, 0), type, 4 + len);
            if (crc != be32(&f[p + 8 + len]))
                throw std::runtime_error("PNG: chunk CRC mismatch: " + path);
            if (!std::memcmp(type, "IHDR", 4)) {
                if (len != 13)
                    throw std::runtime_error("PNG: bad IHDR: " + path);
                w = be32(body);
                h = be32(body + 4);
                depth = body[8];
                ctype = body[9];
                interlace = body[12];
                const bool ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16))
                    || (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8))
                    || ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
                if (!ok || body[10] != 0 || body[11] != 0 || interlace > 1 || w == 0 || h == 0)
                    throw std::runtime_error("PNG: unsupported IHDR: " + path);
                ihdr = true;
            } else if (!std::memcmp(type, "PLTE", 4)) {
                plte.assign(body, body + len);
            } else if (!std::memcmp(type, "IDAT", 4)) {
                z.insert(z.end(), body, body + len);
            } else if (!std::memcmp(type, "IEND", 4)) {
                iend = true;
            }
            p += 12 + (size_t)len;
        }
        if (!ihdr || z.empty())
            throw std::runtime_error("PNG: missing IHDR or IDAT: " + path);
        if (ctype == 3 && (plte.empty() || plte.size() % 3))
            throw std::runtime_error("PNG: palette image without a valid PLTE: " + path);
        // Extents are checked before any frame buffer is sized from them: at most
        // 2^30 pixels (OpenCV's default CV_IO_MAX_IMAGE_PIXELS, the limit of the
        // reference's cv::VideoCapture path), and no more pixel bytes than the
        // compressed data can hold (deflate expands by at most 1032:1).
        if ((uint64_t)w * h > (uint64_t(1) << 30))
            throw std::runtime_error("PNG: image larger than 2^30 pixels: " + path);
        const uint64_t channels = ctype == 2 ? 3 : ctype == 4 ? 2 : ctype == 6 ? 4 : 1;
        if ((uint64_t)w * h * channels * depth / 8 > (uint64_t)z.size() * 1032 + 64)
            throw std::runtime_error("PNG: image data too short for its extents: " + path);
        return z;
    }

    static void inflate_all(std::vector<uint8_t>& z, std::vector<uint8_t>& raw, const std::string& path)
    {
        z_stream st {};
        if (inflateInit(&st) != Z_OK)
            throw std::runtime_error("PNG: zlib init failed");
        st.next_in = z.data();
        st.avail_in = (uInt)z.size();
        st.next_out = raw.data();
        st.avail_out = (uInt)raw.size();
        const int rc = inflate(&st, Z_FINISH);
        const size_t got = raw.size() - st.avail_out;
        inflateEnd(&st);
        if ((rc != Z_STREAM_END && rc != Z_BUF_ERROR) || got != raw.size())
            throw std::runtime_error("PNG: image data shorter than its header says: " + path);
    }

    static void unfilter(uint8_t ft, std::vector<uint8_t>& cur, const std::vector<uint8_t>& prev, size_t bpp,
        const std::string& path)
    {
        const size_t n = cur.size();
        switch (ft) {
        case 0:
            break;
        case 1: // Sub
            for (size_t i = bpp; i < n; ++i)
                cur[i] = (uint8_t)(cur[i] + cur[i - bpp]);
            break;
        case 2: // Up
            for (size_t i = 0; i < n; ++i)
                cur[i] = (uint8_t)(cur[i] + prev[i]);
            break;
        case 3: // Average
            for (size_t i = 0; i < n; ++i)
                cur[i] = (uint8_t)(cur[i] + (((i >= bpp ? cur[i - bpp] : 0) + prev[i]) >> 1));
            break;
        case 4: // Paeth
            for (size_t i = 0; i < n; ++i) {
                const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
                const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                cur[i] = (uint8_t)(cur[i] + (pa <= pb && pa <= pc ? a : pb <= pc ? b : c));
            }
            break;
        default:
            throw std::runtime_error("PNG: unknown row filter: " + path);
        }
    }

    // sample k (0-based within the row's samples) at `depth` bits
    uint32_t sample(const std::vector<uint8_t>& row, size_t k) const
    {
        if (depth == 8)
            return row[k];
        if (depth == 16)
            return (uint32_t)row[2 * k] << 8 | row[2 * k + 1];
        const size_t bit = k * depth;
        return (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1);
    }

    void store(const std::vector<uint8_t>& row, size_t x, int spp, int sel, void* out, size_t at) const
    {
        uint32_t v;
        if (ctype == 3) { // palette index -> R,G,B; sel counts B,G,R
            const size_t idx = sample(row, x);
            if (3 * idx + 2 >= plte.size())
                throw std::runtime_error("PNG: palette index beyond PLTE");
            v = plte[3 * idx + (2 - sel)];
        } else if (grey()) {
            v = sample(row, x * spp);
            if (depth < 8)
                v = v * 255u / ((1u << depth) - 1u);
        } else {
            v = sample(row, x * spp + (2 - sel));
        }
        if (depth == 16)
            static_cast<uint16_t*>(out)[at] = (uint16_t)v;
        else
            static_cast<uint8_t*>(out)[at] = (uint8_t)v;
    }
};

class PngSource : public FrameSource {
public:
    PngSource(const std::string& path, int colour_mask)
    {
        m_files = sequence(path);
        if (m_files.empty())
            throw std::runtime_error("file open error: " + path);
        m_first.header(m_files[0]);
        m_sel = channel_index(colour_mask, 3);
    }
    uint64_t nframes() const override { return m_files.size(); }
    uint64_t ncols() const override { return m_first.w; }
    uint64_t nrows() const override { return m_first.h; }
    int32_t dtype() const override { return m_first.depth == 16 ? SMIP_DTYPE_U16 : SMIP_DTYPE_U8; }
    void read(uint64_t first, uint64_t count, void* out) override
    {
        for (uint64_t f = 0; f < count; ++f) {
            PngImage im;
            const std::string& path = m_files.at(first + f);
            im.header(path);
            if (im.w != m_first.w || im.h != m_first.h || im.depth != m_first.depth || im.ctype != m_first.ctype)
                throw std::runtime_error("PNG sequence: " + path + " differs in size or pixel format from "
                    + m_files[0]);
            im.decode(path, m_sel, static_cast<uint8_t*>(out) + f * frame_bytes());
        }
    }
    const std::vector<std::string>& files() const { return m_files; }

    // The files of a path or pattern, in order.
    static std::vector<std::string> sequence(const std::string& path)
    {
        auto exists = [](const std::string& p) {
            struct stat sb;
            return ::stat(p.c_str(), &sb) == 0 && S_ISREG(sb.st_mode);
        };
        std::vector<std::string> out;
        Pattern pat;
        long start = 0;
        if (path.find('%') != std::string::npos) {
            pat = parse_pattern(path);
            start = -1;
            for (long i = 0; i <= 1000 && start < 0; ++i)
                if (exists(pat.render(i)))
                    start = i;
            if (start < 0)
                return out;
        } else {
            // "<prefix><digits>.<ext>": a sequence from that number, same width
            const size_t slash = path.find_last_of('/');
            const size_t base = slash == std::string::npos ? 0 : slash + 1;
            const size_t dot = path.rfind('.');
            const size_t stem_end = (dot == std::string::npos || dot < base) ? path.size() : dot;
            size_t d0 = stem_end;
            while (d0 > base && std::isdigit(static_cast<unsigned char>(path[d0 - 1])))
                --d0;
            if (!exists(path))
                return out;
            if (d0 == stem_end || stem_end - d0 > 9) {
                out.push_back(path);
                return out;
            }
            pat.pre = path.substr(0, d0);
            pat.post = path.substr(stem_end);
            pat.width = (int)(stem_end - d0);
            pat.zero = true;
            start = std::stol(path.substr(d0, stem_end - d0));
        }
        for (long i = start; exists(pat.render(i)); ++i)
            out.push_back(pat.render(i));
        return out;
    }

private:
    // "<pre>%[0][width][l]d<post>", with "%%" for a literal '%' elsewhere
    struct Pattern {
        std::string pre, post;
        int width = 0;
        bool zero = false;
        std::string render(long i) const
        {
            std::string d = std::to_string(i);
            if ((int)d.size() < width)
                d.insert(0, (size_t)(width - (int)d.size()), zero ? '0' : ' ');
            return pre + d + post;
        }
    };
    static Pattern parse_pattern(const std::string& p)
    {
        Pattern r;
        bool seen = false;
        std::string* cur = &r.pre;
        for (size_t k = 0; k < p.size(); ++k) {
            if (p[k] != '%') {
                *cur += p[k];
                continue;
            }
            if (k + 1 < p.size() && p[k + 1] == '%') {
                *cur += '%';
                ++k;
                continue;
            }
            size_t e = k + 1;
            if (e < p.size() && p[e] == '0') {
                r.zero = true;
                ++e;
            }
            while (e < p.size() && std::isdigit(static_cast<unsigned char>(p[e])))
                r.width = r.width * 10 + (p[e++] - '0');
            while (e < p.size() && p[e] == 'l')
                ++e;
            if (seen || e >= p.size() || (p[e] != 'd' && p[e] != 'u') || r.width > 32)
                throw std::runtime_error("image sequence pattern needs one %d conversion: " + p);
            seen = true;
            cur = &r.post;
            k = e;
        }
        if (!seen)
            throw std::runtime_error("image sequence pattern needs one %d conversion: " + p);
        return r;
    }

    std::vector<std::string> m_files;
    PngImage m_first;
    int m_sel = 2;
};

} // namespace smip_cli
