// exr_io.cpp -- OpenEXR scanline files: FP16/FP32 RGBA, NONE / ZIPS / ZIP.
#include "exr_io.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iterator>
#include <cstring>
#include <fstream>
#include <map>
#include <stdexcept>

namespace aur {

// ---------------------------------------------------------------------------
// IEEE binary16 conversion, round to nearest even
// ---------------------------------------------------------------------------
uint16_t float_to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const int exp = (int)((x >> 23) & 0xff);
    uint32_t mant = x & 0x7fffffu;
    if (exp == 255) return (uint16_t)(sign | 0x7c00u | (mant ? (0x200u | (mant >> 13)) : 0u));
    const int e = exp - 127 + 15;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        mant |= 0x800000u;
        const int shift = 14 - e;
        uint32_t h = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (h & 1u))) ++h;
        return (uint16_t)(sign | h);
    }
    uint32_t h = sign | ((uint32_t)e << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1fffu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) ++h;  // a carry rounds into the exponent (up to inf)
    return (uint16_t)h;
}

float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const int exp = (h >> 10) & 0x1f;
    uint32_t mant = h & 0x3ffu;
    uint32_t x;
    if (exp == 0) {
        if (!mant) {
            x = sign;
        } else {  // subnormal: normalise
            int e = -1;
            do {
                ++e;
                mant <<= 1;
            } while (!(mant & 0x400u));
            x = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3ffu) << 13);
        }
    } else if (exp == 31) {
        x = sign | 0x7f800000u | (mant << 13);
    } else {
        x = sign | ((uint32_t)(exp - 15 + 127) << 23) | (mant << 13);
    }
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// ---------------------------------------------------------------------------
// byte helpers (the format is little-endian)
// ---------------------------------------------------------------------------
namespace {

struct Out {
    std::vector<unsigned char> b;
    void u8(unsigned v) { b.push_back((unsigned char)v); }
    void i32(int32_t v) { for (int k = 0; k < 4; ++k) b.push_back((unsigned char)((uint32_t)v >> (8 * k))); }
    void u64(uint64_t v) { for (int k = 0; k < 8; ++k) b.push_back((unsigned char)(v >> (8 * k))); }
    void f32(float f) { uint32_t x; std::memcpy(&x, &f, 4); i32((int32_t)x); }
    void str0(const std::string &s) { b.insert(b.end(), s.begin(), s.end()); b.push_back(0); }
    void bytes(const void *p, size_t n) { const unsigned char *c = (const unsigned char *)p; b.insert(b.end(), c, c + n); }
    void attr(const std::string &name, const std::string &type, const std::vector<unsigned char> &v) {
        str0(name);
        str0(type);
        i32((int32_t)v.size());
        bytes(v.data(), v.size());
    }
};

struct In {
    const unsigned char *p, *end;
    void need(size_t n) const { if ((size_t)(end - p) < n) throw std::runtime_error("EXR file truncated"); }
    uint32_t u32() { need(4); uint32_t v = p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24); p += 4; return v; }
    int32_t i32() { return (int32_t)u32(); }
    uint64_t u64() { uint64_t lo = u32(); uint64_t hi = u32(); return lo | (hi << 32); }
    float f32() { uint32_t x = u32(); float f; std::memcpy(&f, &x, 4); return f; }
    std::string str0() {
        const unsigned char *s = p;
        while (p < end && *p) ++p;
        if (p >= end) throw std::runtime_error("EXR header: unterminated name");
        std::string r((const char *)s, (size_t)(p - s));
        ++p;
        return r;
    }
};

int lines_per_chunk(int compression) { return compression == EXR_ZIP ? 16 : 1; }

// OpenEXR's ZIP filter: bytes split into even / odd halves, then a delta
// predictor, then zlib.
std::vector<unsigned char> zip_pack(const std::vector<unsigned char> &raw) {
    const size_t n = raw.size();
    std::vector<unsigned char> t(n);
    size_t a = 0, b = (n + 1) / 2;
    for (size_t i = 0; i < n; ++i) t[(i & 1) ? b++ : a++] = raw[i];
    for (size_t i = n; i-- > 1;) t[i] = (unsigned char)((int)t[i] - (int)t[i - 1] + 128 + 256);
    uLongf out_n = compressBound((uLong)n);
    std::vector<unsigned char> out(out_n);
    if (compress(out.data(), &out_n, t.data(), (uLong)n) != Z_OK) throw std::runtime_error("zlib compress failed");
    out.resize(out_n);
    return out;
}

std::vector<unsigned char> zip_unpack(const unsigned char *src, size_t n, size_t raw_size) {
    std::vector<unsigned char> t(raw_size), raw(raw_size);
    uLongf got = (uLongf)raw_size;
    if (uncompress(t.data(), &got, src, (uLong)n) != Z_OK || got != raw_size)
        throw std::runtime_error("EXR: corrupt ZIP chunk");
    for (size_t i = 1; i < raw_size; ++i) t[i] = (unsigned char)((int)t[i - 1] + (int)t[i] - 128);
    size_t a = 0, b = (raw_size + 1) / 2;
    for (size_t i = 0; i < raw_size; ++i) raw[i] = t[(i & 1) ? b++ : a++];
    return raw;
}

}  // namespace

// ---------------------------------------------------------------------------
// writer
// ---------------------------------------------------------------------------
void exr_write(const std::string &path, int width, int height, const float *rgba, int compression, bool half,
               const std::vector<std::pair<std::string, std::string>> &meta) {
    if (width <= 0 || height <= 0) throw std::invalid_argument("EXR: empty image");
    if (compression != EXR_NONE && compression != EXR_ZIPS && compression != EXR_ZIP)
        throw std::invalid_argument("EXR: compression must be NONE, ZIPS or ZIP");
    Out o;
    o.i32(20000630);
    o.i32(2);
    // channel list, names in alphabetical order as the format requires
    static const char *names[4] = {"A", "B", "G", "R"};
    static const int comp[4] = {3, 2, 1, 0};
    {
        Out c;
        for (int k = 0; k < 4; ++k) {
            c.str0(names[k]);
            c.i32(half ? 1 : 2);
            c.u8(0); c.u8(0); c.u8(0); c.u8(0);
            c.i32(1);
            c.i32(1);
        }
        c.u8(0);
        o.attr("channels", "chlist", c.b);
    }
    o.attr("compression", "compression", {(unsigned char)compression});
    Out box;
    box.i32(0); box.i32(0); box.i32(width - 1); box.i32(height - 1);
    o.attr("dataWindow", "box2i", box.b);
    o.attr("displayWindow", "box2i", box.b);
    o.attr("lineOrder", "lineOrder", {0});
    Out f;
    f.f32(1.f);
    o.attr("pixelAspectRatio", "float", f.b);
    Out v2;
    v2.f32(0.f); v2.f32(0.f);
    o.attr("screenWindowCenter", "v2f", v2.b);
    o.attr("screenWindowWidth", "float", f.b);
    for (const auto &kv : meta) o.attr(kv.first, "string", std::vector<unsigned char>(kv.second.begin(), kv.second.end()));
    o.u8(0);  // end of header

    const int lpc = lines_per_chunk(compression);
    const int n_chunks = (height + lpc - 1) / lpc;
    const size_t table = o.b.size();
    for (int k = 0; k < n_chunks; ++k) o.u64(0);
    const int bpc = half ? 2 : 4;
    for (int k = 0; k < n_chunks; ++k) {
        const int y0 = k * lpc, y1 = std::min(height, y0 + lpc);
        std::vector<unsigned char> raw;
        raw.reserve((size_t)(y1 - y0) * width * 4 * bpc);
        for (int y = y0; y < y1; ++y)
            for (int ch = 0; ch < 4; ++ch)
                for (int x = 0; x < width; ++x) {
                    const float v = rgba[4 * ((size_t)y * width + x) + comp[ch]];
                    if (half) {
                        const uint16_t h = float_to_half(v);
                        raw.push_back((unsigned char)(h & 0xff));
                        raw.push_back((unsigned char)(h >> 8));
                    } else {
                        uint32_t u;
                        std::memcpy(&u, &v, 4);
                        for (int b = 0; b < 4; ++b) raw.push_back((unsigned char)(u >> (8 * b)));
                    }
                }
        std::vector<unsigned char> data = raw;
        if (compression != EXR_NONE) {
            std::vector<unsigned char> z = zip_pack(raw);
            if (z.size() < raw.size()) data.swap(z);  // otherwise stored raw, as the format allows
        }
        const uint64_t at = o.b.size();
        for (int b = 0; b < 8; ++b) o.b[table + 8 * k + b] = (unsigned char)(at >> (8 * b));
        o.i32(y0);
        o.i32((int32_t)data.size());
        o.bytes(data.data(), data.size());
    }
    // written next to the target and renamed, so readers never see a partial file
    const std::string tmp = path + "_tmp";
    {
        std::ofstream fh(tmp, std::ios::binary);
        if (!fh) throw std::runtime_error("cannot write " + tmp);
        fh.write((const char *)o.b.data(), (std::streamsize)o.b.size());
        if (!fh) throw std::runtime_error("write failed: " + tmp);
    }
    if (std::rename(tmp.c_str(), path.c_str()) != 0) throw std::runtime_error("cannot rename " + tmp + " to " + path);
}

// ---------------------------------------------------------------------------
// reader
// ---------------------------------------------------------------------------
ExrImage exr_read(const std::string &path) {
    std::ifstream fh(path, std::ios::binary);
    if (!fh) throw std::runtime_error("cannot open " + path);
    std::vector<unsigned char> buf((std::istreambuf_iterator<char>(fh)), std::istreambuf_iterator<char>());
    In in{buf.data(), buf.data() + buf.size()};
    if (in.i32() != 20000630) throw std::runtime_error(path + ": not an OpenEXR file");
    const uint32_t ver = in.u32();
    if ((ver & 0xff) != 2) throw std::runtime_error(path + ": unsupported EXR version");
    if (ver & 0x200) throw std::runtime_error(path + ": tiled EXR files are not supported");
    if (ver & 0x1000) throw std::runtime_error(path + ": multi-part EXR files are not supported");
    struct Chan { std::string name; int type; };
    std::vector<Chan> chans;
    int compression = -1, line_order = 0;
    int dw[4] = {0, 0, -1, -1};
    ExrImage img;
    for (;;) {
        const std::string name = in.str0();
        if (name.empty()) break;
        const std::string type = in.str0();
        const int32_t size = in.i32();
        if (size < 0) throw std::runtime_error(path + ": bad attribute size");
        in.need((size_t)size);
        In a{in.p, in.p + size};
        if (name == "channels") {
            for (;;) {
                const std::string cn = a.str0();
                if (cn.empty()) break;
                Chan c;
                c.name = cn;
                c.type = a.i32();
                a.need(4);
                a.p += 4;
                const int xs = a.i32(), ys = a.i32();
                if (xs != 1 || ys != 1) throw std::runtime_error(path + ": subsampled channels are not supported");
                chans.push_back(c);
            }
        } else if (name == "compression") {
            a.need(1);
            compression = a.p[0];
        } else if (name == "dataWindow") {
            for (int k = 0; k < 4; ++k) dw[k] = a.i32();
        } else if (name == "lineOrder") {
            a.need(1);
            line_order = a.p[0];
        } else if (type == "string") {
            img.meta.push_back({name, std::string((const char *)a.p, (size_t)size)});
        }
        in.p += size;
    }
    if (compression != EXR_NONE && compression != EXR_ZIPS && compression != EXR_ZIP)
        throw std::runtime_error(path + ": unsupported EXR compression " + std::to_string(compression));
    if (line_order != 0 && line_order != 1) throw std::runtime_error(path + ": unsupported line order");
    const int w = dw[2] - dw[0] + 1, h = dw[3] - dw[1] + 1;
    if (w <= 0 || h <= 0 || (long long)w * h > (1ll << 30)) throw std::runtime_error(path + ": bad data window");
    img.width = w;
    img.height = h;
    img.rgba.assign((size_t)w * h * 4, 0.f);
    for (size_t i = 0; i < (size_t)w * h; ++i) img.rgba[4 * i + 3] = 1.f;
    size_t pixel_bytes = 0;
    for (const Chan &c : chans) {
        if (c.type != 1 && c.type != 2 && c.type != 0) throw std::runtime_error(path + ": unknown channel type");
        pixel_bytes += c.type == 1 ? 2 : 4;
    }
    // a luminance-only file reads as grey (RgbaInputFile: R = G = B = Y)
    bool has_rgb = false;
    for (const Chan &c : chans) has_rgb = has_rgb || c.name == "R" || c.name == "G" || c.name == "B";
    const int lpc = lines_per_chunk(compression);
    const int n_chunks = (h + lpc - 1) / lpc;
    std::vector<uint64_t> offs(n_chunks);
    for (int k = 0; k < n_chunks; ++k) offs[k] = in.u64();
    for (int k = 0; k < n_chunks; ++k) {
        if (offs[k] + 8 > buf.size()) throw std::runtime_error(path + ": bad chunk offset");
        In c{buf.data() + offs[k], buf.data() + buf.size()};
        const int y0 = c.i32() - dw[1];
        const int32_t n = c.i32();
        if (n < 0) throw std::runtime_error(path + ": bad chunk size");
        c.need((size_t)n);
        if (y0 < 0 || y0 >= h) throw std::runtime_error(path + ": chunk outside the data window");
        const int y1 = std::min(h, y0 + lpc);
        const size_t raw_size = (size_t)(y1 - y0) * w * pixel_bytes;
        std::vector<unsigned char> raw;
        if ((size_t)n == raw_size) raw.assign(c.p, c.p + n);
        else if (compression == EXR_NONE) throw std::runtime_error(path + ": bad uncompressed chunk size");
        else raw = zip_unpack(c.p, (size_t)n, raw_size);
        const unsigned char *p = raw.data();
        for (int y = y0; y < y1; ++y)
            for (const Chan &ch : chans) {
                const int dst = ch.name == "R" ? 0 : ch.name == "G" ? 1 : ch.name == "B" ? 2 : ch.name == "A" ? 3
                              : (ch.name == "Y" && !has_rgb) ? 4 : -1;
                for (int x = 0; x < w; ++x) {
                    float v;
                    if (ch.type == 1) {
                        v = half_to_float((uint16_t)(p[0] | (p[1] << 8)));
                        p += 2;
                    } else {
                        uint32_t u = p[0] | (p[1] << 8) | (p[2] << 16) | ((uint32_t)p[3] << 24);
                        p += 4;
                        if (ch.type == 2) std::memcpy(&v, &u, 4);
                        else v = (float)u;
                    }
                    float *px = &img.rgba[4 * ((size_t)y * w + x)];
                    if (dst == 4) px[0] = px[1] = px[2] = v;
                    else if (dst >= 0) px[dst] = v;
                }
            }
    }
    return img;
}

}  // namespace aur
