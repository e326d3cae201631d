// image_io.cpp — WriteImage of the C++ API: linear .pfm and .hdr, and
// tone-mapped 8-bit .png and .jpg (Bulbit's WriteImage,
// src/util/image.cpp:219-266).
//
// Orientation: the film's row y = 0 is the bottom of the picture (the camera
// adds vertical * y / height to lower_left, src/camera/perspective_camera.cpp:40-52),
// so top-down formats write row height - 1 first, as the reference does with
// stbi_flip_vertically_on_write.  The PNG and JPEG encoders are this
// project's own: a zlib stream of stored deflate blocks, and a baseline
// JPEG with 4:4:4 sampling and the all-ones quantization that quality 100
// gives.
#include "bulbit/bulbit.h"

#include <algorithm>
#include <cctype>
#include <cstdlib>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <new>
#include <stdexcept>
#include <vector>

#include <zlib.h>

namespace bulbit
{

// sRGB_from_RGB (include/bulbit/color.h:31-50)
Vec3 sRGB_from_RGB(const Vec3& rgb)
{
    Float c[3];
    for (int i = 0; i < 3; ++i) {
        const Float comp = rgb[i];
        c[i] = comp <= 0.0031308f ? 12.92f * comp : (1.0f + 0.055f) * std::pow(comp, 1.0f / 2.4f) - 0.055f;
    }
    return Vec3(c[0], c[1], c[2]);
}

namespace
{

bool ends_with(const std::string& s, const char* ext)
{
    const size_t n = std::strlen(ext);
    return s.size() >= n && s.compare(s.size() - n, n, ext) == 0;
}

// tone-mapped 8-bit RGB in top-down row order (image.cpp:229-252)
std::vector<unsigned char> to_8bit(const Image3& image, ToneMappingCallback* callback)
{
    const int32 w = image.width, h = image.height;
    std::vector<unsigned char> px(size_t(w) * size_t(h) * 3);
    for (int32 y = 0; y < h; ++y)
        for (int32 x = 0; x < w; ++x) {
            const Spectrum& s = image(x, h - 1 - y);
            const Vec3 m = callback(Vec3(s.r, s.g, s.b));
            unsigned char* o = &px[(size_t(y) * size_t(w) + size_t(x)) * 3];
            for (int k = 0; k < 3; ++k)
                o[k] = (unsigned char)(std::min(double(std::clamp(m[k], 0.0f, 1.0f)) * 256.0, 255.0));
        }
    return px;
}

// ------------------------------------------------------------------ PNG

uint32_t crc32_update(uint32_t crc, const unsigned char* p, size_t n)
{
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
            table[i] = c;
        }
        init = true;
    }
    for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xff] ^ (crc >> 8);
    return crc;
}

void put_be32(std::vector<unsigned char>& v, uint32_t x)
{
    v.push_back((unsigned char)(x >> 24));
    v.push_back((unsigned char)(x >> 16));
    v.push_back((unsigned char)(x >> 8));
    v.push_back((unsigned char)x);
}

void png_chunk(std::vector<unsigned char>& out, const char* type, const std::vector<unsigned char>& data)
{
    put_be32(out, uint32_t(data.size()));
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    const uint32_t crc = crc32_update(0xffffffffu, &out[start], out.size() - start) ^ 0xffffffffu;
    put_be32(out, crc);
}

// zlib stream (RFC 1950) of stored deflate blocks (RFC 1951, BTYPE 00)
std::vector<unsigned char> zlib_stored(const std::vector<unsigned char>& raw)
{
    std::vector<unsigned char> z = { 0x78, 0x01 };
    size_t pos = 0;
    do {
        const size_t n = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((unsigned char)(n & 0xff));
        z.push_back((unsigned char)(n >> 8));
        z.push_back((unsigned char)(~n & 0xff));
        z.push_back((unsigned char)((~n >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + long(pos), raw.begin() + long(pos + n));
        pos += n;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0; // Adler-32
    for (unsigned char c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    return z;
}

bool write_png(const std::string& filename, int32 w, int32 h, const std::vector<unsigned char>& px)
{
    std::vector<unsigned char> raw;
    raw.reserve(size_t(h) * (size_t(w) * 3 + 1));
    for (int32 y = 0; y < h; ++y) {
        raw.push_back(0); // filter type None
        const unsigned char* row = &px[size_t(y) * size_t(w) * 3];
        raw.insert(raw.end(), row, row + size_t(w) * 3);
    }
    std::vector<unsigned char> out = { 0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n' };
    std::vector<unsigned char> ihdr;
    put_be32(ihdr, uint32_t(w));
    put_be32(ihdr, uint32_t(h));
    ihdr.insert(ihdr.end(), { 8, 2, 0, 0, 0 }); // 8-bit RGB, deflate, no filter set choice, no interlace
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", zlib_stored(raw));
    png_chunk(out, "IEND", {});
    FILE* f = std::fopen(filename.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0) && ok;
}

// ----------------------------------------------------------------- JPEG

// The Huffman tables of the JPEG standard's Annex K.3 (ITU-T T.81): code
// lengths 1..16 and symbol values, luminance and chrominance, DC and AC
const unsigned char kDcLumBits[16] = { 0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0 };
const unsigned char kDcChrBits[16] = { 0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0 };
const unsigned char kDcVals[12] = { 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11 };
const unsigned char kAcLumBits[16] = { 0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d };
const unsigned char kAcLumVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61, 0x07, 0x22, 0x71, 0x14,
    0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72, 0x82, 0x09,
    0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a,
    0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65,
    0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88,
    0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9,
    0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca,
    0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea,
    0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa
};
const unsigned char kAcChrBits[16] = { 0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77 };
const unsigned char kAcChrVals[162] = {
    0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07, 0x61, 0x71, 0x13, 0x22, 0x32,
    0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09, 0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1, 0x0a, 0x16,
    0x24, 0x34, 0xe1, 0x25, 0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36, 0x37, 0x38, 0x39,
    0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64,
    0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x82, 0x83, 0x84, 0x85, 0x86,
    0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7,
    0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8,
    0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9,
    0xea, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa
};
// zigzag position -> row-major index within the 8x8 block
const unsigned char kZigzag[64] = { 0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                    12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                    58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63 };

struct Huff
{
    unsigned short code[256] = {};
    unsigned char size[256] = {};
    // canonical codes from the code-length counts (T.81 Annex C)
    Huff(const unsigned char* bits, const unsigned char* vals)
    {
        unsigned short c = 0;
        int k = 0;
        for (int len = 1; len <= 16; ++len) {
            for (int i = 0; i < bits[len - 1]; ++i, ++k) {
                code[vals[k]] = c++;
                size[vals[k]] = (unsigned char)len;
            }
            c = (unsigned short)(c << 1);
        }
    }
};

struct BitWriter
{
    std::vector<unsigned char>& out;
    uint32_t acc = 0;
    int n = 0;
    void put(uint32_t bits, int len)
    {
        for (int i = len - 1; i >= 0; --i) {
            acc = (acc << 1) | ((bits >> i) & 1u);
            if (++n == 8) {
                out.push_back((unsigned char)acc);
                if (acc == 0xff) out.push_back(0); // byte stuffing
                acc = 0;
                n = 0;
            }
        }
    }
    void flush()
    {
        while (n != 0) put(1, 1); // pad with ones
    }
};

// magnitude category and the low bits of a coefficient (T.81 F.1.2)
void encode_value(int v, int& cat, uint32_t& bits)
{
    const int a = v < 0 ? -v : v;
    cat = 0;
    while ((a >> cat) != 0) ++cat;
    bits = uint32_t(v < 0 ? v + (1 << cat) - 1 : v);
}

void encode_block(BitWriter& bw, const float* blk, int& prev_dc, const Huff& dc, const Huff& ac)
{
    // forward DCT-II of the level-shifted block, straight from its definition
    static float basis[8][8];
    static bool init = false;
    if (!init) {
        for (int u = 0; u < 8; ++u)
            for (int x = 0; x < 8; ++x)
                basis[u][x] = float((u == 0 ? std::sqrt(0.125) : 0.5) * std::cos((2 * x + 1) * u * M_PI / 16.0));
        init = true;
    }
    float tmp[64], coef[64];
    for (int y = 0; y < 8; ++y)
        for (int u = 0; u < 8; ++u) {
            float s = 0.0f;
            for (int x = 0; x < 8; ++x) s += basis[u][x] * blk[y * 8 + x];
            tmp[y * 8 + u] = s;
        }
    for (int v = 0; v < 8; ++v)
        for (int u = 0; u < 8; ++u) {
            float s = 0.0f;
            for (int y = 0; y < 8; ++y) s += basis[v][y] * tmp[y * 8 + u];
            coef[v * 8 + u] = s;
        }
    int q[64]; // quantization by the all-ones tables: rounding only
    for (int i = 0; i < 64; ++i) q[i] = int(std::lround(coef[kZigzag[i]]));
    int cat;
    uint32_t bits;
    encode_value(q[0] - prev_dc, cat, bits);
    prev_dc = q[0];
    bw.put(dc.code[cat], dc.size[cat]);
    if (cat) bw.put(bits, cat);
    int run = 0;
    for (int i = 1; i < 64; ++i) {
        if (q[i] == 0) {
            ++run;
            continue;
        }
        while (run > 15) {
            bw.put(ac.code[0xf0], ac.size[0xf0]); // ZRL
            run -= 16;
        }
        encode_value(q[i], cat, bits);
        const int sym = (run << 4) | cat;
        bw.put(ac.code[sym], ac.size[sym]);
        bw.put(bits, cat);
        run = 0;
    }
    if (run) bw.put(ac.code[0x00], ac.size[0x00]); // EOB
}

void put_marker(std::vector<unsigned char>& o, unsigned char m, const std::vector<unsigned char>& payload)
{
    o.push_back(0xff);
    o.push_back(m);
    const size_t len = payload.size() + 2;
    o.push_back((unsigned char)(len >> 8));
    o.push_back((unsigned char)len);
    o.insert(o.end(), payload.begin(), payload.end());
}

bool write_jpg(const std::string& filename, int32 w, int32 h, const std::vector<unsigned char>& px)
{
    std::vector<unsigned char> o = { 0xff, 0xd8 };                                     // SOI
    put_marker(o, 0xe0, { 'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0 });        // APP0 JFIF 1.1
    std::vector<unsigned char> dqt;
    for (unsigned char t = 0; t < 2; ++t) {
        dqt.push_back(t);
        dqt.insert(dqt.end(), 64, 1);
    }
    put_marker(o, 0xdb, dqt);
    put_marker(o, 0xc0, { 8, (unsigned char)(h >> 8), (unsigned char)h, (unsigned char)(w >> 8), (unsigned char)w, 3,
                          1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1 });                       // SOF0, 4:4:4
    std::vector<unsigned char> dht;
    auto add_table = [&](unsigned char cls_id, const unsigned char* bits, const unsigned char* vals, int nvals) {
        dht.push_back(cls_id);
        dht.insert(dht.end(), bits, bits + 16);
        dht.insert(dht.end(), vals, vals + nvals);
    };
    add_table(0x00, kDcLumBits, kDcVals, 12);
    add_table(0x10, kAcLumBits, kAcLumVals, 162);
    add_table(0x01, kDcChrBits, kDcVals, 12);
    add_table(0x11, kAcChrBits, kAcChrVals, 162);
    put_marker(o, 0xc4, dht);
    put_marker(o, 0xda, { 3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0 });                  // SOS
    const Huff dc_l(kDcLumBits, kDcVals), ac_l(kAcLumBits, kAcLumVals);
    const Huff dc_c(kDcChrBits, kDcVals), ac_c(kAcChrBits, kAcChrVals);
    BitWriter bw{ o };
    int prev[3] = { 0, 0, 0 };
    for (int32 by = 0; by < h; by += 8)
        for (int32 bx = 0; bx < w; bx += 8) {
            float blk[3][64];
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x) {
                    // edge blocks repeat the last row and column
                    const int32 sx = std::min(bx + x, w - 1), sy = std::min(by + y, h - 1);
                    const unsigned char* p = &px[(size_t(sy) * size_t(w) + size_t(sx)) * 3];
                    const float r = p[0], g = p[1], b = p[2];
                    // JFIF YCbCr, level-shifted by 128
                    blk[0][y * 8 + x] = 0.299f * r + 0.587f * g + 0.114f * b - 128.0f;
                    blk[1][y * 8 + x] = -0.168736f * r - 0.331264f * g + 0.5f * b;
                    blk[2][y * 8 + x] = 0.5f * r - 0.418688f * g - 0.081312f * b;
                }
            encode_block(bw, blk[0], prev[0], dc_l, ac_l);
            encode_block(bw, blk[1], prev[1], dc_c, ac_c);
            encode_block(bw, blk[2], prev[2], dc_c, ac_c);
        }
    bw.flush();
    o.push_back(0xff);
    o.push_back(0xd9); // EOI
    FILE* f = std::fopen(filename.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
    return (std::fclose(f) == 0) && ok;
}

} // namespace

bool WriteImage(const Image3& image, const std::string& filename, ToneMappingCallback* callback)
{
    const int32 w = image.width, h = image.height;
    if (w <= 0 || h <= 0 || image.data.size() != size_t(w) * size_t(h)) {
        std::fprintf(stderr, "bulbit_amd: WriteImage: %s: empty or malformed image\n", filename.c_str());
        return false;
    }
    if (ends_with(filename, ".pfm")) {
        FILE* f = std::fopen(filename.c_str(), "wb");
        if (!f) return false;
        std::fprintf(f, "PF\n%d %d\n-1\n", w, h);
        // PFM stores rows bottom to top, as the film does: row 0 first
        for (int32 y = 0; y < h; ++y)
            for (int32 x = 0; x < w; ++x) {
                const Spectrum& s = image(x, y);
                const float v[3] = { s.r, s.g, s.b };
                std::fwrite(v, sizeof(float), 3, f);
            }
        return std::fclose(f) == 0;
    }
    if (ends_with(filename, ".hdr")) {
        FILE* f = std::fopen(filename.c_str(), "wb");
        if (!f) return false;
        std::fprintf(f, "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n", h, w);
        // -Y: the first scanline is the top of the picture, the film's last row
        for (int32 y = h - 1; y >= 0; --y)
            for (int32 x = 0; x < w; ++x) {
                const Spectrum& s = image(x, y);
                const float m = std::max(s.r, std::max(s.g, s.b));
                unsigned char rgbe[4] = { 0, 0, 0, 0 };
                if (m > 1e-32f) {
                    int e = 0;
                    const float scale = std::frexp(m, &e) * 256.0f / m;
                    rgbe[0] = (unsigned char)(std::max(0.0f, s.r) * scale);
                    rgbe[1] = (unsigned char)(std::max(0.0f, s.g) * scale);
                    rgbe[2] = (unsigned char)(std::max(0.0f, s.b) * scale);
                    rgbe[3] = (unsigned char)(e + 128);
                }
                std::fwrite(rgbe, 1, 4, f);
            }
        return std::fclose(f) == 0;
    }
    if (ends_with(filename, ".png") || ends_with(filename, ".jpg")) {
        if (!callback) callback = default_tonemapping_callback;
        const std::vector<unsigned char> px = to_8bit(image, callback);
        return ends_with(filename, ".png") ? write_png(filename, w, h, px) : write_jpg(filename, w, h, px);
    }
    std::fprintf(stderr, "bulbit_amd: WriteImage: %s: supported extensions are .png .jpg .hdr .pfm\n",
                 filename.c_str());
    return false;
}

// ------------------------------------------------------------ readers
//
// ReadImage1 / ReadImage3 (src/util/image.cpp:13-130) decode through
// stb_image (pinned at ae721c50, not vendored in the reference) with
// stbi_set_flip_vertically_on_load(true) and stbi_loadf: 8-bit samples c
// become pow(c / 255, gamma) with gamma 2.2 for colour and 1 for non-colour
// data (stbi__ldr_to_hdr), an alpha channel becomes c / 255, Radiance .hdr
// RGBE becomes r * 2^(e - 136) (stbi__hdr_convert), and every value is
// clamped at 0.  Decoded here: binary PNM (P5 grey, P6 RGB, 8 or 16 bits,
// 16-bit samples reduced to their high byte as stb does), PNG (below), JPEG
// (jpeg_decode.cpp), Radiance .hdr (flat and RLE scanlines), and PFM
// (beyond stb: linear floats, already bottom row first).

namespace jpeg
{
bool Decode(const uint8_t* data, size_t size, std::vector<uint8_t>& rgba, int& width, int& height, int& file_comps);
}

namespace
{

struct Decoded
{
    int32 width = 0, height = 0, comps = 0;
    int32 file_comps = 0;    // the file's own channel count (stb's *comp); 0: comps
    std::vector<float> data; // comps per pixel, row 0 at the bottom
};

float ldr_to_float(unsigned char c, bool non_color)
{
    const float x = c / 255.0f;
    return non_color ? float(std::pow(double(x), 1.0)) : float(std::pow(double(x), double(2.2f)));
}

std::string pnm_token(std::FILE* f)
{
    std::string t;
    int c = std::fgetc(f);
    while (c != EOF) {
        if (c == '#') {
            while (c != EOF && c != '\n') c = std::fgetc(f);
        } else if (!std::isspace(c)) {
            break;
        }
        c = std::fgetc(f);
    }
    while (c != EOF && !std::isspace(c)) {
        t.push_back(char(c));
        c = std::fgetc(f);
    }
    return t;
}

bool read_pnm(std::FILE* f, bool non_color, Decoded& out)
{
    const std::string magic = pnm_token(f);
    if (magic != "P5" && magic != "P6") return false;
    const int32 w = std::atoi(pnm_token(f).c_str()), h = std::atoi(pnm_token(f).c_str());
    const int32 maxval = std::atoi(pnm_token(f).c_str());
    if (w <= 0 || h <= 0 || maxval <= 0 || maxval > 65535) return false;
    const int32 comps = magic == "P6" ? 3 : 1;
    const int32 bytes = maxval > 255 ? 2 : 1;
    std::vector<unsigned char> raw(size_t(w) * h * comps * bytes);
    if (std::fread(raw.data(), 1, raw.size(), f) != raw.size()) return false;
    out.width = w;
    out.height = h;
    out.comps = comps;
    out.data.resize(size_t(w) * h * comps);
    for (int32 y = 0; y < h; ++y)
        for (int32 i = 0; i < w * comps; ++i) {
            const size_t src = (size_t(y) * w * comps + i) * bytes; // first byte: the high byte of 16-bit samples
            out.data[size_t(h - 1 - y) * w * comps + i] = ldr_to_float(raw[src], non_color);
        }
    return true;
}

bool read_hdr(std::FILE* f, Decoded& out)
{
    char line[256];
    if (!std::fgets(line, sizeof(line), f) || std::strncmp(line, "#?", 2) != 0) return false;
    for (;;) { // header lines up to the blank line
        if (!std::fgets(line, sizeof(line), f)) return false;
        if (line[0] == '\n' || line[0] == '\r') break;
        if (std::strncmp(line, "FORMAT=", 7) == 0 && !std::strstr(line, "32-bit_rle_rgbe")) return false;
    }
    int32 w = 0, h = 0;
    if (!std::fgets(line, sizeof(line), f) || std::sscanf(line, "-Y %d +X %d", &h, &w) != 2 || w <= 0 || h <= 0)
        return false;
    std::vector<unsigned char> rgbe(size_t(w) * h * 4);
    std::vector<unsigned char> row(size_t(w) * 4);
    for (int32 y = 0; y < h; ++y) {
        unsigned char* dst = rgbe.data() + size_t(y) * w * 4;
        int c0 = std::fgetc(f), c1 = std::fgetc(f), c2 = std::fgetc(f), c3 = std::fgetc(f);
        if (c3 == EOF) return false;
        if (w < 8 || w >= 32768 || c0 != 2 || c1 != 2 || (c2 & 0x80)) { // flat scanline
            dst[0] = (unsigned char)c0;
            dst[1] = (unsigned char)c1;
            dst[2] = (unsigned char)c2;
            dst[3] = (unsigned char)c3;
            if (w > 1 && std::fread(dst + 4, 1, size_t(w - 1) * 4, f) != size_t(w - 1) * 4) return false;
            continue;
        }
        if (((c2 << 8) | c3) != w) return false;
        for (int k = 0; k < 4; ++k) { // RLE: each component in turn
            int32 x = 0;
            while (x < w) {
                int count = std::fgetc(f);
                if (count == EOF) return false;
                if (count > 128) {
                    count -= 128;
                    const int v = std::fgetc(f);
                    if (v == EOF || x + count > w) return false;
                    for (int j = 0; j < count; ++j) row[size_t(x++) * 4 + k] = (unsigned char)v;
                } else {
                    if (count == 0 || x + count > w) return false;
                    for (int j = 0; j < count; ++j) {
                        const int v = std::fgetc(f);
                        if (v == EOF) return false;
                        row[size_t(x++) * 4 + k] = (unsigned char)v;
                    }
                }
            }
        }
        std::memcpy(dst, row.data(), row.size());
    }
    out.width = w;
    out.height = h;
    out.comps = 3;
    out.data.resize(size_t(w) * h * 3);
    for (int32 y = 0; y < h; ++y)
        for (int32 x = 0; x < w; ++x) {
            const unsigned char* p = rgbe.data() + (size_t(y) * w + x) * 4;
            float* o = out.data.data() + (size_t(h - 1 - y) * w + x) * 3;
            if (p[3]) {
                const float f1 = float(std::ldexp(1.0f, int(p[3]) - (128 + 8)));
                o[0] = p[0] * f1;
                o[1] = p[1] * f1;
                o[2] = p[2] * f1;
            } else {
                o[0] = o[1] = o[2] = 0.0f;
            }
        }
    return true;
}

bool read_pfm(std::FILE* f, Decoded& out)
{
    char magic[3] = { 0, 0, 0 };
    if (std::fscanf(f, "%2s", magic) != 1) return false;
    const int32 comps = std::strcmp(magic, "PF") == 0 ? 3 : (std::strcmp(magic, "Pf") == 0 ? 1 : 0);
    int32 w = 0, h = 0;
    float scale = 0.0f;
    if (!comps || std::fscanf(f, "%d %d %f", &w, &h, &scale) != 3 || w <= 0 || h <= 0) return false;
    std::fgetc(f); // the single whitespace before the data
    out.width = w;
    out.height = h;
    out.comps = comps;
    out.data.resize(size_t(w) * h * comps);
    if (std::fread(out.data.data(), 4, out.data.size(), f) != out.data.size()) return false;
    if (scale > 0) // big-endian data
        for (float& v : out.data) {
            uint32_t u;
            std::memcpy(&u, &v, 4);
            u = (u >> 24) | ((u >> 8) & 0xff00u) | ((u << 8) & 0xff0000u) | (u << 24);
            std::memcpy(&v, &u, 4);
        }
    return true;
}

// PNG (stb_image's stbi__parse_png_file / stbi__create_png_image /
// stbi__convert_format): every image is expanded to RGBA as stbi_loadf's
// STBI_rgb_alpha request does (grey replicated, alpha 255 where the file has
// none, palette entries looked up, tRNS colour keys turned into alpha 0),
// sub-byte greys scaled to 8 bits (x 0xff, 0x55, 0x11), 16-bit samples
// reduced to their high byte, Adam7 interlacing undone.  The reported
// channel count is the file's: 1 grey, 2 grey + alpha, 3 RGB or palette,
// 4 RGBA or palette with tRNS.  The zlib stream is inflated by zlib.
bool read_png(std::FILE* f, bool non_color, Decoded& out)
{
    std::vector<unsigned char> file;
    unsigned char buf[65536];
    for (size_t n; (n = std::fread(buf, 1, sizeof buf, f)) > 0;) file.insert(file.end(), buf, buf + n);
    static const unsigned char sig[8] = { 137, 80, 78, 71, 13, 10, 26, 10 };
    if (file.size() < 8 || std::memcmp(file.data(), sig, 8) != 0) return false;
    auto be32 = [](const unsigned char* p) {
        return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | uint32_t(p[3]);
    };
    uint32_t w = 0, h = 0;
    int depth = 0, color = -1, interlace = 0;
    std::vector<unsigned char> idat, palette; // palette: RGBA entries
    unsigned char trns_key[6] = { 0, 0, 0, 0, 0, 0 };
    bool has_key = false, pal_alpha = false;
    size_t pos = 8;
    for (;;) {
        if (pos + 8 > file.size()) return false;
        const uint32_t len = be32(&file[pos]);
        const std::string type(reinterpret_cast<const char*>(&file[pos + 4]), 4);
        if (pos + 12 + size_t(len) > file.size()) return false;
        const unsigned char* d = &file[pos + 8];
        if (type == "IHDR") {
            if (len != 13) return false;
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            color = d[9];
            interlace = d[12];
            if (w == 0 || h == 0 || w > (1u << 24) || h > (1u << 24) || d[10] != 0 || d[11] != 0 || interlace > 1)
                return false;
        } else if (type == "PLTE") {
            if (len % 3 || len > 768) return false;
            palette.assign(size_t(len / 3) * 4, 255);
            for (uint32_t i = 0; i < len / 3; ++i)
                for (int k = 0; k < 3; ++k) palette[i * 4 + k] = d[i * 3 + k];
        } else if (type == "tRNS") {
            if (color == 3) {
                if (palette.empty() || len > palette.size() / 4) return false;
                for (uint32_t i = 0; i < len; ++i) palette[i * 4 + 3] = d[i];
                pal_alpha = true;
            } else if (color == 0 || color == 2) {
                if (len != (color == 0 ? 2u : 6u)) return false;
                std::memcpy(trns_key, d, len);
                has_key = true;
            } else {
                return false;
            }
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        } else if (!(type[0] & 0x20)) {
            return false; // an unknown critical chunk
        }
        pos += 12 + size_t(len);
    }
    const int samples = color == 0 ? 1 : color == 2 ? 3 : color == 3 ? 1 : color == 4 ? 2 : color == 6 ? 4 : 0;
    if (!samples || idat.empty()) return false;
    const bool depth_ok = color == 0 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)
                          : color == 3 ? (depth == 1 || depth == 2 || depth == 4 || depth == 8)
                                       : (depth == 8 || depth == 16);
    if (!depth_ok || (color == 3 && palette.empty())) return false;

    // the 7 Adam7 passes (or the whole image): origin and step in x and y
    static const int ax0[7] = { 0, 4, 0, 2, 0, 1, 0 }, ay0[7] = { 0, 0, 4, 0, 2, 0, 1 };
    static const int adx[7] = { 8, 8, 4, 4, 2, 2, 1 }, ady[7] = { 8, 8, 8, 4, 4, 2, 2 };
    const int passes = interlace ? 7 : 1;
    const size_t bits_pp = size_t(samples) * size_t(depth);
    size_t raw_size = 0;
    uint32_t pw[7], ph[7];
    for (int p = 0; p < passes; ++p) {
        pw[p] = interlace ? (w > uint32_t(ax0[p]) ? (w - ax0[p] + adx[p] - 1) / adx[p] : 0) : w;
        ph[p] = interlace ? (h > uint32_t(ay0[p]) ? (h - ay0[p] + ady[p] - 1) / ady[p] : 0) : h;
        if (pw[p] && ph[p]) raw_size += size_t(ph[p]) * (1 + (size_t(pw[p]) * bits_pp + 7) / 8);
    }
    // deflate expands at most 1032:1, so IHDR sizes that the IDAT data cannot
    // fill are rejected before anything is allocated for them; zlib counts
    // its output in a 32-bit uInt
    if (raw_size == 0 || raw_size / 1032 > idat.size() + 64 || raw_size > size_t(UINT32_MAX)) return false;
    std::vector<unsigned char> raw(raw_size);
    z_stream zs;
    std::memset(&zs, 0, sizeof zs);
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = idat.data();
    zs.avail_in = uInt(idat.size());
    zs.next_out = raw.data();
    zs.avail_out = uInt(raw.size());
    const int zr = inflate(&zs, Z_FINISH);
    const size_t produced = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    if ((zr != Z_STREAM_END && zr != Z_BUF_ERROR) || produced != raw.size()) return false;

    // RGBA 8-bit, top row first
    std::vector<unsigned char> rgba(size_t(w) * h * 4);
    const size_t bpp = std::max<size_t>(1, bits_pp / 8); // filter byte distance
    static const unsigned char scale[5] = { 0, 0xff, 0x55, 0, 0x11 };
    size_t off = 0;
    for (int p = 0; p < passes; ++p) {
        if (!pw[p] || !ph[p]) continue;
        const size_t stride = (size_t(pw[p]) * bits_pp + 7) / 8;
        std::vector<unsigned char> prev(stride, 0), cur(stride);
        for (uint32_t y = 0; y < ph[p]; ++y) {
            const int filter = raw[off];
            const unsigned char* src = &raw[off + 1];
            off += 1 + stride;
            for (size_t i = 0; i < stride; ++i) { // undo the scanline filter (PNG spec section 9)
                const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
                int v = src[i];
                switch (filter) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: {
                    const int pa = std::abs(b - c), pb = std::abs(a - c), pc = std::abs(a + b - 2 * c);
                    v += (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                    break;
                }
                default: return false;
                }
                cur[i] = (unsigned char)v;
            }
            for (uint32_t x = 0; x < pw[p]; ++x) {
                unsigned char px[4] = { 0, 0, 0, 255 };
                uint16_t s16[4] = { 0, 0, 0, 0 };
                for (int k = 0; k < samples; ++k) {
                    const size_t bit = (size_t(x) * samples + k) * size_t(depth);
                    if (depth == 16) {
                        s16[k] = uint16_t((cur[bit / 8] << 8) | cur[bit / 8 + 1]);
                        px[k] = (unsigned char)(s16[k] >> 8);
                    } else if (depth == 8) {
                        px[k] = s16[k] = cur[bit / 8];
                    } else {
                        const int v = (cur[bit / 8] >> (8 - depth - int(bit % 8))) & ((1 << depth) - 1);
                        s16[k] = uint16_t(v);
                        px[k] = (unsigned char)(color == 3 ? v : v * scale[depth]);
                    }
                }
                unsigned char o[4];
                if (color == 3) {
                    if (size_t(px[0]) * 4 >= palette.size()) return false;
                    std::memcpy(o, &palette[size_t(px[0]) * 4], 4);
                } else if (color == 0 || color == 4) {
                    o[0] = o[1] = o[2] = px[0];
                    o[3] = color == 4 ? px[1] : 255;
                } else {
                    o[0] = px[0];
                    o[1] = px[1];
                    o[2] = px[2];
                    o[3] = color == 6 ? px[3] : 255;
                }
                if (has_key) { // a colour key compares the file's samples at their own depth
                    bool match = true;
                    for (int k = 0; k < samples; ++k)
                        match &= s16[k] == (depth == 16 ? uint16_t((trns_key[2 * k] << 8) | trns_key[2 * k + 1])
                                                        : uint16_t(trns_key[2 * k + 1]));
                    if (match) o[3] = 0;
                }
                const size_t X = interlace ? size_t(ax0[p]) + size_t(x) * adx[p] : x;
                const size_t Y = interlace ? size_t(ay0[p]) + size_t(y) * ady[p] : y;
                std::memcpy(&rgba[(Y * w + X) * 4], o, 4);
            }
            std::swap(prev, cur);
        }
    }
    out.width = int32(w);
    out.height = int32(h);
    out.comps = 4;
    out.file_comps = color == 3 ? (pal_alpha ? 4 : 3) : samples;
    out.data.resize(size_t(w) * h * 4);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x)
            for (int k = 0; k < 4; ++k) {
                const unsigned char c = rgba[(size_t(y) * w + x) * 4 + k];
                out.data[(size_t(h - 1 - y) * w + x) * 4 + k] = k == 3 ? c / 255.0f : ldr_to_float(c, non_color);
            }
    return true;
}

// JPEG: baseline and progressive, grey, YCbCr with any whole sampling
// ratios, Adobe RGB/CMYK/YCCK; expanded to RGBA with alpha 255 as
// STBI_rgb_alpha asks, reporting 1 channel for grey and 3 for colour
bool read_jpeg(std::FILE* f, bool non_color, Decoded& out)
{
    std::vector<unsigned char> file;
    unsigned char buf[65536];
    for (size_t n; (n = std::fread(buf, 1, sizeof buf, f)) > 0;) file.insert(file.end(), buf, buf + n);
    std::vector<uint8_t> rgba;
    int w = 0, h = 0, comps = 0;
    if (!jpeg::Decode(file.data(), file.size(), rgba, w, h, comps)) return false;
    out.width = w;
    out.height = h;
    out.comps = 4;
    out.file_comps = comps;
    out.data.resize(size_t(w) * size_t(h) * 4);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int k = 0; k < 4; ++k) {
                const unsigned char c = rgba[(size_t(y) * size_t(w) + size_t(x)) * 4 + size_t(k)];
                out.data[(size_t(h - 1 - y) * size_t(w) + size_t(x)) * 4 + size_t(k)] =
                    k == 3 ? c / 255.0f : ldr_to_float(c, non_color);
            }
    return true;
}

// The format comes from the file's first bytes, as stb_image decides it
// (each of its decoders tests a signature; the file name plays no part), so
// a PNG saved as .jpg reads as a PNG.  PFM, which stb does not read, is
// recognised by its own signature.
enum class Format { unknown, png, jpeg, hdr, pnm, pfm };

Format sniff(std::FILE* f)
{
    unsigned char h[8] = {};
    const size_t n = std::fread(h, 1, sizeof h, f);
    std::rewind(f);
    const auto space = [](unsigned char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; };
    if (n >= 8 && std::memcmp(h, "\x89PNG\r\n\x1a\n", 8) == 0) return Format::png;
    if (n >= 2 && h[0] == 0xFF && h[1] == 0xD8) return Format::jpeg; // start-of-image marker
    if (n >= 2 && h[0] == '#' && h[1] == '?') return Format::hdr;    // "#?RADIANCE", "#?RGBE"
    if (n >= 3 && h[0] == 'P' && (h[1] == '5' || h[1] == '6') && space(h[2])) return Format::pnm;
    if (n >= 3 && h[0] == 'P' && (h[1] == 'F' || h[1] == 'f') && space(h[2])) return Format::pfm;
    return Format::unknown;
}

bool decode(const std::string& filename, bool non_color, Decoded& out)
{
    std::FILE* f = std::fopen(filename.c_str(), "rb");
    if (!f) return false;
    bool ok = false;
    // a header whose sizes cannot be allocated fails the read (stb_image
    // returns NULL for it) instead of throwing out of ReadImage1/ReadImage3
    try {
        switch (sniff(f)) {
        case Format::png: ok = read_png(f, non_color, out); break;
        case Format::jpeg: ok = read_jpeg(f, non_color, out); break;
        case Format::hdr: ok = read_hdr(f, out); break;
        case Format::pnm: ok = read_pnm(f, non_color, out); break;
        case Format::pfm: ok = read_pfm(f, out); break;
        case Format::unknown: break;
        }
    } catch (const std::bad_alloc&) {
        ok = false;
    } catch (const std::length_error&) {
        ok = false;
    }
    std::fclose(f);
    return ok;
}

} // namespace

Image1 ReadImage1(const std::string& filename, int32 channel, bool non_color, std::function<Float(Float)> transform)
{
    Decoded d;
    if (!decode(filename, non_color, d)) {
        std::fprintf(stderr, "Failed to read image1: %s\n", filename.c_str());
        return {};
    }
    // a channel beyond the file's components (an alpha channel the file
    // does not have, for one) gives an empty image, as in the reference
    if (channel < 0 || channel >= (d.file_comps ? d.file_comps : d.comps)) return {};
    Image1 image(d.width, d.height);
    for (size_t i = 0; i < image.data.size(); ++i) {
        const float v = d.comps == 1 ? d.data[i] : d.data[i * size_t(d.comps) + size_t(channel)];
        const Float c = Float(std::fmax(0.0f, v));
        image.data[i] = transform ? transform(c) : c;
    }
    return image;
}

Image3 ReadImage3(const std::string& filename, bool non_color, std::function<Spectrum(Spectrum)> transform)
{
    Decoded d;
    if (!decode(filename, non_color, d)) {
        std::fprintf(stderr, "Failed to read image3: %s\n", filename.c_str());
        return {};
    }
    Image3 image(d.width, d.height);
    for (size_t i = 0; i < image.data.size(); ++i) {
        Spectrum s;
        for (int k = 0; k < 3; ++k) {
            const float v = d.comps == 1 ? d.data[i] : d.data[i * size_t(d.comps) + size_t(k)];
            s[k] = Float(std::fmax(0.0f, v));
        }
        image.data[i] = transform ? transform(s) : s;
    }
    return image;
}

} // namespace bulbit
