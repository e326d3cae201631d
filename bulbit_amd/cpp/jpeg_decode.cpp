// JPEG decoding for ReadImage1 / ReadImage3.
//
// The reference reads JPEG textures through stb_image (pinned at ae721c50,
// not vendored): src/util/image.cpp:24,102,157 call stbi_loadf, and the
// reference's own scenes load .jpg textures (cli/scene/head.cpp:12).  This
// decoder is written here from two sources:
//  - ITU-T T.81 for the bitstream: markers, Huffman tables (Annex C),
//    sequential Huffman scans (Annex F, SOF0 and SOF1, 8-bit), progressive
//    scans with spectral selection and successive approximation (Annex G,
//    SOF2), interleaved and single-component scans, restart intervals;
//  - stb_image's published sample reconstruction, so that the 8-bit values
//    come out as stbi_loadf sees them:
//    * the IJG "islow" integer IDCT (Loeffler-Ligtenberg-Moschytz) with
//      12-bit constants, two fraction bits kept between the column and the
//      row pass, and round-half-up descaling by 2^10 and 2^17;
//    * triangle ("fancy") chroma upsampling, 3/4 of the nearer and 1/4 of
//      the farther sample, for 2:1 ratios in either or both directions;
//      replication for other ratios;
//    * JFIF YCbCr -> RGB in 20-bit fixed point; Adobe RGB, CMYK and YCCK.
// Arithmetic coding, lossless and hierarchical frames and 12-bit samples
// are rejected, as stb rejects them.

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

namespace bulbit
{
namespace jpeg
{
namespace
{

// zigzag position -> row-major index within the 8x8 block, built from the
// walk along anti-diagonals that T.81 Figure A.6 draws
struct Zigzag
{
    uint8_t natural[64];
    Zigzag()
    {
        int k = 0;
        for (int s = 0; s < 15; ++s) {
            // even diagonals run up and to the right, odd ones down and to the left
            const int lo = std::max(0, s - 7), hi = std::min(s, 7);
            for (int i = lo; i <= hi; ++i) {
                const int row = (s & 1) ? i : s - i;
                natural[k++] = uint8_t(row * 8 + (s - row));
            }
        }
    }
};
const Zigzag kZz;

struct HuffTable
{
    bool defined = false;
    uint8_t symbols[256] = {};
    int32_t maxcode[17] = {}; // largest code of each length, -1 when there is none
    int32_t offset[17] = {};  // index into symbols of a code of each length, minus that code
    uint16_t fast[512] = {};  // 9-bit prefix -> (length << 8) | symbol; 0: longer code

    // canonical codes from the 16 code-length counts (T.81 Annex C)
    bool build(const uint8_t counts[16], const uint8_t* syms, int n)
    {
        std::memcpy(symbols, syms, size_t(n));
        std::memset(fast, 0, sizeof fast);
        int32_t code = 0, k = 0;
        for (int len = 1; len <= 16; ++len) {
            offset[len] = k - code;
            for (int i = 0; i < counts[len - 1]; ++i, ++k, ++code) {
                if (code >= (1 << len)) return false; // more codes than the length allows
                if (len <= 9) {
                    const int first = code << (9 - len), span = 1 << (9 - len);
                    for (int j = 0; j < span; ++j) fast[first + j] = uint16_t((len << 8) | syms[k]);
                }
            }
            maxcode[len] = counts[len - 1] ? code - 1 : -1;
            code <<= 1;
        }
        defined = true;
        return true;
    }
};

// Entropy-coded data: bytes 0xFF 0x00 stand for 0xFF; any other 0xFF xx is
// a marker, after which the reader supplies zero bits
struct BitReader
{
    const uint8_t* p = nullptr;
    const uint8_t* end = nullptr;
    uint64_t acc = 0; // bits, most significant first
    int nbits = 0;
    int marker = -1; // marker met in the data and not yet consumed (p stays on its 0xFF)

    void fill()
    {
        while (nbits <= 56) {
            uint32_t b = 0;
            if (marker < 0 && p < end) {
                if (*p != 0xFF) {
                    b = *p++;
                } else {
                    const uint8_t* q = p + 1;
                    while (q < end && *q == 0xFF) ++q; // fill bytes before a marker
                    if (q < end && *q == 0x00) {
                        b = 0xFF;
                        p = q + 1;
                    } else {
                        marker = q < end ? *q : 0xD9; // data that runs out reads as the end of the image
                        p = q - 1;
                    }
                }
            }
            acc |= uint64_t(b) << (56 - nbits);
            nbits += 8;
        }
    }
    uint32_t peek(int n)
    {
        if (nbits < n) fill();
        return uint32_t(acc >> (64 - n));
    }
    void skip(int n)
    {
        acc <<= n;
        nbits -= n;
    }
    uint32_t bits(int n)
    {
        if (n == 0) return 0;
        const uint32_t v = peek(n);
        skip(n);
        return v;
    }
    // a magnitude category's extra bits as a signed value (T.81 F.2.2.1 EXTEND)
    int32_t extend(int s)
    {
        if (s == 0) return 0;
        const int32_t v = int32_t(bits(s));
        return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v;
    }
    int decode(const HuffTable& t)
    {
        const uint32_t head = peek(16);
        const uint16_t f = t.fast[head >> 7];
        if (f) {
            skip(f >> 8);
            return f & 0xFF;
        }
        for (int len = 10; len <= 16; ++len) {
            const int32_t c = int32_t(head >> (16 - len));
            if (c <= t.maxcode[len]) {
                skip(len);
                return t.symbols[t.offset[len] + c];
            }
        }
        return -1;
    }
    // the next marker from where the data stands; p is left on its 0xFF
    int find_marker()
    {
        acc = 0;
        nbits = 0;
        if (marker < 0) {
            while (p + 1 < end && !(p[0] == 0xFF && p[1] != 0x00 && p[1] != 0xFF)) ++p;
            if (p + 1 < end) {
                marker = p[1];
            } else { // no marker before the data ends
                p = end;
                marker = 0xD9;
            }
        }
        return marker;
    }
};

struct Component
{
    int id = 0, h = 1, v = 1, tq = 0;
    int w = 0, rows = 0;   // samples: ceil(X h / hmax) by ceil(Y v / vmax)
    int bw = 0, bh = 0;    // blocks in the MCU-padded plane
    int td = 0, ta = 0;    // tables of the current scan
    int32_t pred = 0;      // DC predictor
    std::vector<uint8_t> plane;  // bw*8 by bh*8 samples
    std::vector<int16_t> coefs;  // progressive frames: 64 per block, natural order
};

// 12-bit fixed-point constant as stb_image rounds it (in float)
constexpr int fix12(float x) { return int(x * 4096.0f + 0.5f); }

// one 8-point islow IDCT: s[0..7] at stride, results written as
// even[i] +- odd[i] to out[i] and out[7 - i] before descaling by `shift`
// after adding `bias`.  64-bit sums: the same results as 32-bit ones
// wherever those do not overflow, and no overflow on crafted coefficients.
template <typename In, typename Out, typename Store>
inline void idct8(const In* s, int stride, int64_t bias, int shift, Out* out, int ostride, Store store)
{
    using i64 = int64_t;
    // even part: inputs 0, 2, 4, 6
    const i64 p2e = s[2 * stride], p3e = s[6 * stride];
    const i64 p1 = (p2e + p3e) * fix12(0.5411961f);
    const i64 t2e = p1 + p3e * fix12(-1.847759065f);
    const i64 t3e = p1 + p2e * fix12(0.765366865f);
    const i64 t0e = (i64(s[0]) + s[4 * stride]) * 4096;
    const i64 t1e = (i64(s[0]) - s[4 * stride]) * 4096;
    const i64 e0 = t0e + t3e + bias, e3 = t0e - t3e + bias;
    const i64 e1 = t1e + t2e + bias, e2 = t1e - t2e + bias;
    // odd part: inputs 7, 5, 3, 1
    i64 a0 = s[7 * stride], a1 = s[5 * stride], a2 = s[3 * stride], a3 = s[stride];
    const i64 z3 = a0 + a2, z4 = a1 + a3, z1 = a0 + a3, z2 = a1 + a2;
    const i64 z5 = (z3 + z4) * fix12(1.175875602f);
    a0 *= fix12(0.298631336f);
    a1 *= fix12(2.053119869f);
    a2 *= fix12(3.072711026f);
    a3 *= fix12(1.501321110f);
    const i64 m1 = z5 + z1 * fix12(-0.899976223f);
    const i64 m2 = z5 + z2 * fix12(-2.562915447f);
    const i64 m3 = z3 * fix12(-1.961570560f);
    const i64 m4 = z4 * fix12(-0.390180644f);
    const i64 o3 = a3 + m1 + m4, o2 = a2 + m2 + m3, o1 = a1 + m2 + m4, o0 = a0 + m1 + m3;
    store(out[0], (e0 + o3) >> shift);
    store(out[7 * ostride], (e0 - o3) >> shift);
    store(out[1 * ostride], (e1 + o2) >> shift);
    store(out[6 * ostride], (e1 - o2) >> shift);
    store(out[2 * ostride], (e2 + o1) >> shift);
    store(out[5 * ostride], (e2 - o1) >> shift);
    store(out[3 * ostride], (e3 + o0) >> shift);
    store(out[4 * ostride], (e3 - o0) >> shift);
}

// dequantized coefficients (natural order) -> 8x8 samples at dst
void idct_block(const int16_t* c, uint8_t* dst, int stride)
{
    int64_t tmp[64];
    for (int x = 0; x < 8; ++x) {
        const int16_t* col = c + x;
        if (!col[8] && !col[16] && !col[24] && !col[32] && !col[40] && !col[48] && !col[56]) {
            // a column with only its DC term: the column pass gives 4 x DC everywhere
            for (int y = 0; y < 8; ++y) tmp[y * 8 + x] = col[0] * 4;
            continue;
        }
        idct8(col, 8, 512, 10, tmp + x, 8, [](int64_t& o, int64_t v) { o = v; });
    }
    for (int y = 0; y < 8; ++y)
        idct8(tmp + y * 8, 1, 65536 + (128 << 17), 17, dst + y * stride, 1,
              [](uint8_t& o, int64_t v) { o = uint8_t(v < 0 ? 0 : (v > 255 ? 255 : v)); });
}

struct Decoder
{
    Decoder(const uint8_t* bytes, size_t n) : data(bytes), size(n) {}

    const uint8_t* data;
    size_t size;
    size_t pos = 0;
    uint16_t quant[4][64] = {}; // natural order
    HuffTable dc[4], ac[4];
    std::vector<Component> comps;
    int X = 0, Y = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false, frame = false, jfif = false;
    int adobe_transform = -1;
    int restart = 0;
    int eobrun = 0;

    int byte()
    {
        return pos < size ? data[pos++] : -1;
    }
    int be16()
    {
        const int a = byte(), b = byte();
        return (a < 0 || b < 0) ? -1 : (a << 8) | b;
    }

    bool read_dqt(size_t seg_end)
    {
        while (pos < seg_end) {
            const int pq_tq = byte();
            const int pq = pq_tq >> 4, tq = pq_tq & 15;
            if (pq > 1 || tq > 3 || pos + size_t(64 << pq) > seg_end) return false;
            for (int k = 0; k < 64; ++k) quant[tq][kZz.natural[k]] = uint16_t(pq ? be16() : byte());
        }
        return true;
    }
    bool read_dht(size_t seg_end)
    {
        while (pos < seg_end) {
            const int tc_th = byte();
            const int tc = tc_th >> 4, th = tc_th & 15;
            if (tc > 1 || th > 3 || pos + 16 > seg_end) return false;
            uint8_t counts[16];
            int n = 0;
            for (int i = 0; i < 16; ++i) n += counts[i] = uint8_t(byte());
            if (n > 256 || pos + size_t(n) > seg_end) return false;
            if (!(tc ? ac[th] : dc[th]).build(counts, data + pos, n)) return false;
            pos += size_t(n);
        }
        return true;
    }
    bool read_sof(size_t seg_end)
    {
        if (frame) return false;
        if (byte() != 8) return false; // 8-bit samples only
        Y = be16();
        X = be16();
        const int nf = byte();
        // a height of 0 (defined later by a DNL marker) is not supported
        if (X <= 0 || Y <= 0 || X > (1 << 24) || Y > (1 << 24)) return false;
        if (nf != 1 && nf != 3 && nf != 4) return false;
        if (pos + size_t(3 * nf) > seg_end) return false;
        comps.resize(size_t(nf));
        for (Component& c : comps) {
            c.id = byte();
            const int hv = byte();
            c.h = hv >> 4;
            c.v = hv & 15;
            c.tq = byte();
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return false;
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        mcux = (X + 8 * hmax - 1) / (8 * hmax);
        mcuy = (Y + 8 * vmax - 1) / (8 * vmax);
        for (Component& c : comps) {
            if (hmax % c.h || vmax % c.v) return false; // sampling ratios must be whole
            c.w = int((int64_t(X) * c.h + hmax - 1) / hmax);
            c.rows = int((int64_t(Y) * c.v + vmax - 1) / vmax);
            c.bw = mcux * c.h;
            c.bh = mcuy * c.v;
            c.plane.assign(size_t(c.bw) * c.bh * 64, 0);
            if (progressive) c.coefs.assign(size_t(c.bw) * c.bh * 64, 0);
        }
        frame = true;
        return true;
    }

    // ------------------------------------------------------------ blocks

    bool block_sequential(BitReader& br, Component& c, int bx, int by)
    {
        int16_t blk[64] = {};
        const uint16_t* q = quant[c.tq];
        const int t = br.decode(dc[c.td]);
        if (t < 0 || t > 15) return false;
        // the predictor and the products wrap (as 16-bit coefficients do) on crafted data
        c.pred = int32_t(uint32_t(c.pred) + uint32_t(br.extend(t)));
        blk[0] = int16_t(uint32_t(c.pred) * q[0]);
        for (int k = 1; k < 64;) {
            const int rs = br.decode(ac[c.ta]);
            if (rs < 0) return false;
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (rs != 0xF0) break; // end of block
                k += 16;              // sixteen zeros
                continue;
            }
            k += r;
            if (k > 63) return false;
            const int z = kZz.natural[k++];
            blk[z] = int16_t(uint32_t(br.extend(s)) * q[z]);
        }
        idct_block(blk, &c.plane[(size_t(by) * 8 * size_t(c.bw) + size_t(bx)) * 8], c.bw * 8);
        return true;
    }

    bool block_dc(BitReader& br, Component& c, int16_t* blk, int ah, int al)
    {
        if (ah == 0) {
            const int t = br.decode(dc[c.td]);
            if (t < 0 || t > 15) return false;
            c.pred = int32_t(uint32_t(c.pred) + uint32_t(br.extend(t)));
            blk[0] = int16_t(uint32_t(c.pred) << al);
        } else if (br.bits(1)) {
            blk[0] = int16_t(blk[0] + (1 << al));
        }
        return true;
    }

    bool block_ac_first(BitReader& br, Component& c, int16_t* blk, int ss, int se, int al)
    {
        if (eobrun > 0) {
            --eobrun;
            return true;
        }
        for (int k = ss; k <= se;) {
            const int rs = br.decode(ac[c.ta]);
            if (rs < 0) return false;
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r < 15) { // a run of 2^r + extra bits end-of-bands, this block the first
                    eobrun = (1 << r) - 1 + int(br.bits(r));
                    break;
                }
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) return false;
            blk[kZz.natural[k++]] = int16_t(uint32_t(br.extend(s)) << al);
        }
        return true;
    }

    // T.81 G.1.2.3: one correction bit for every coefficient already
    // nonzero, new coefficients of magnitude 1 << al placed among the zeros
    bool block_ac_refine(BitReader& br, Component& c, int16_t* blk, int ss, int se, int al)
    {
        const int bit = 1 << al;
        auto correct = [&](int16_t& v) {
            if (br.bits(1) && (v & bit) == 0) v = int16_t(v >= 0 ? v + bit : v - bit);
        };
        int k = ss;
        if (eobrun == 0) {
            for (; k <= se; ++k) {
                const int rs = br.decode(ac[c.ta]);
                if (rs < 0) return false;
                int r = rs >> 4;
                int value = 0;
                if (rs & 15) {
                    value = br.bits(1) ? bit : -bit;
                } else if (r < 15) {
                    eobrun = (1 << r) + int(br.bits(r)); // includes this block
                    break;
                }
                // pass over nonzero coefficients (correcting them) and r zeros
                for (; k <= se; ++k) {
                    int16_t& v = blk[kZz.natural[k]];
                    if (v != 0) {
                        correct(v);
                    } else {
                        if (r == 0) break;
                        --r;
                    }
                }
                if (value) {
                    if (k > se) return false;
                    blk[kZz.natural[k]] = int16_t(value);
                }
            }
        }
        if (eobrun > 0) {
            for (; k <= se; ++k) {
                int16_t& v = blk[kZz.natural[k]];
                if (v != 0) correct(v);
            }
            --eobrun;
        }
        return true;
    }

    // ------------------------------------------------------------- scans

    bool read_sos(size_t seg_end)
    {
        if (!frame) return false;
        const int ns = byte();
        if (ns < 1 || ns > 4 || pos + size_t(2 * ns + 3) > seg_end) return false;
        std::vector<Component*> scan;
        for (int i = 0; i < ns; ++i) {
            const int id = byte(), tables = byte();
            Component* found = nullptr;
            for (Component& c : comps)
                if (c.id == id) found = &c;
            if (!found) return false;
            found->td = tables >> 4;
            found->ta = tables & 15;
            if (found->td > 3 || found->ta > 3) return false;
            scan.push_back(found);
        }
        const int ss = byte(), se = byte(), a = byte();
        const int ah = a >> 4, al = a & 15;
        pos = seg_end;
        if (progressive) {
            if (ss > se || se > 63 || ah > 13 || al > 13) return false;
            if (ss == 0 && se != 0) return false; // DC and AC never share a scan
            if (ss > 0 && ns != 1) return false;  // AC scans hold one component
        } else if (ss != 0 || se != 63 || a != 0) {
            return false;
        }
        // the tables the scan will use must be defined
        for (Component* c : scan) {
            const bool needs_dc = !progressive || (ss == 0 && ah == 0);
            const bool needs_ac = !progressive || ss > 0;
            if ((needs_dc && !dc[c->td].defined) || (needs_ac && !ac[c->ta].defined)) return false;
        }
        for (Component& c : comps) c.pred = 0;
        eobrun = 0;

        BitReader br;
        br.p = data + pos;
        br.end = data + size;
        // one MCU: every block of each scan component (a single component's
        // scan has one block per MCU and covers only that component's blocks)
        const bool single = ns == 1;
        const int units_x = single ? (scan[0]->w + 7) / 8 : mcux;
        const int units_y = single ? (scan[0]->rows + 7) / 8 : mcuy;
        int todo = restart ? restart : -1;
        for (int uy = 0; uy < units_y; ++uy)
            for (int ux = 0; ux < units_x; ++ux) {
                for (Component* c : scan) {
                    const int nh = single ? 1 : c->h, nv = single ? 1 : c->v;
                    for (int v = 0; v < nv; ++v)
                        for (int h = 0; h < nh; ++h) {
                            const int bx = ux * nh + h, by = uy * nv + v;
                            bool ok;
                            if (!progressive) {
                                ok = block_sequential(br, *c, bx, by);
                            } else {
                                int16_t* blk = &c->coefs[(size_t(by) * size_t(c->bw) + size_t(bx)) * 64];
                                if (ss == 0)
                                    ok = block_dc(br, *c, blk, ah, al);
                                else if (ah == 0)
                                    ok = block_ac_first(br, *c, blk, ss, se, al);
                                else
                                    ok = block_ac_refine(br, *c, blk, ss, se, al);
                            }
                            if (!ok) return false;
                        }
                }
                if (todo > 0 && --todo == 0) {
                    // end of a restart interval: the rest of the byte is padding;
                    // an RSTn marker resets the predictors, anything else ends the scan
                    const int m = br.find_marker();
                    if (m < 0xD0 || m > 0xD7) {
                        pos = size_t(br.p - data);
                        return true;
                    }
                    br.p += 2;
                    br.marker = -1;
                    for (Component& c : comps) c.pred = 0;
                    eobrun = 0;
                    todo = restart;
                }
            }
        br.find_marker();
        pos = size_t(br.p - data);
        return true;
    }

    void finish_progressive()
    {
        for (Component& c : comps)
            for (int by = 0; by < c.bh; ++by)
                for (int bx = 0; bx < c.bw; ++bx) {
                    int16_t* blk = &c.coefs[(size_t(by) * size_t(c.bw) + size_t(bx)) * 64];
                    for (int i = 0; i < 64; ++i) blk[i] = int16_t(blk[i] * quant[c.tq][i]);
                    idct_block(blk, &c.plane[(size_t(by) * 8 * size_t(c.bw) + size_t(bx)) * 8], c.bw * 8);
                }
    }

    bool parse()
    {
        if (size < 4 || data[0] != 0xFF || data[1] != 0xD8) return false;
        pos = 2;
        bool scanned = false;
        for (;;) {
            int b = byte();
            if (b < 0) return scanned; // data that ends after a scan reads as the end of the image
            if (b != 0xFF) return false;
            int m;
            do m = byte();
            while (m == 0xFF);
            if (m < 0) return scanned;
            if (m == 0xD9) return scanned; // EOI
            if (m >= 0xD0 && m <= 0xD7) continue; // a stray RSTn has no segment
            const int len = be16();
            if (len < 2 || pos + size_t(len) - 2 > size) return false;
            const size_t seg_end = pos + size_t(len) - 2;
            bool ok = true;
            switch (m) {
            case 0xDB: ok = read_dqt(seg_end); break;
            case 0xC4: ok = read_dht(seg_end); break;
            case 0xC0:
            case 0xC1: ok = read_sof(seg_end); break;
            case 0xC2:
                progressive = true;
                ok = read_sof(seg_end);
                break;
            case 0xDD:
                if (len != 4) return false;
                restart = be16();
                break;
            case 0xDA:
                if (!read_sos(seg_end)) return false;
                scanned = true;
                continue; // read_sos leaves pos on the next marker
            case 0xE0:
                if (len >= 7 && std::memcmp(data + pos, "JFIF\0", 5) == 0) jfif = true;
                break;
            case 0xEE:
                if (len >= 14 && std::memcmp(data + pos, "Adobe", 5) == 0) adobe_transform = data[pos + 11];
                break;
            default:
                // other frame types (lossless, hierarchical, arithmetic coding) are not decoded
                if ((m >= 0xC3 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC)) return false;
                break;
            }
            if (!ok) return false;
            pos = seg_end;
        }
    }
};

// ------------------------------------------------------------ output

// the samples of output row `y` of component c, upsampled to X samples
struct RowStepper
{
    const Component* c;
    int hs, vs, ystep, ypos;
    const uint8_t* line0;
    const uint8_t* line1;
    std::vector<uint8_t> out;

    RowStepper(const Component& comp, int hmax, int vmax)
        : c(&comp), hs(hmax / comp.h), vs(vmax / comp.v), ystep(vs >> 1), ypos(0)
    {
        line0 = line1 = comp.plane.data();
        out.resize(size_t(comp.w) * size_t(hs) + 2);
    }
    static uint8_t div4(int x) { return uint8_t(x >> 2); }
    static uint8_t div16(int x) { return uint8_t(x >> 4); }

    const uint8_t* next()
    {
        const bool bottom = ystep >= (vs >> 1);
        const uint8_t* near = bottom ? line1 : line0;
        const uint8_t* far = bottom ? line0 : line1;
        const int w = c->w;
        const uint8_t* result = out.data();
        uint8_t* o = out.data();
        if (hs == 1 && vs == 1) {
            result = near;
        } else if (hs == 1 && vs == 2) {
            for (int i = 0; i < w; ++i) o[i] = div4(3 * near[i] + far[i] + 2);
        } else if (hs == 2 && vs == 1) {
            if (w == 1) {
                o[0] = o[1] = near[0];
            } else {
                o[0] = near[0];
                o[1] = div4(near[0] * 3 + near[1] + 2);
                for (int i = 1; i < w - 1; ++i) {
                    const int n = 3 * near[i] + 2;
                    o[2 * i] = div4(n + near[i - 1]);
                    o[2 * i + 1] = div4(n + near[i + 1]);
                }
                o[2 * (w - 1)] = div4(near[w - 1] * 3 + near[w - 2] + 2);
                o[2 * (w - 1) + 1] = near[w - 1];
            }
        } else if (hs == 2 && vs == 2) {
            // vertical 3:1 blend first (kept at 4x scale), then horizontal
            int t1 = 3 * near[0] + far[0];
            if (w == 1) {
                o[0] = o[1] = div4(t1 + 2);
            } else {
                o[0] = div4(t1 + 2);
                for (int i = 1; i < w; ++i) {
                    const int t0 = t1;
                    t1 = 3 * near[i] + far[i];
                    o[2 * i - 1] = div16(3 * t0 + t1 + 8);
                    o[2 * i] = div16(3 * t1 + t0 + 8);
                }
                o[2 * w - 1] = div4(t1 + 2);
            }
        } else {
            for (int i = 0; i < w; ++i)
                for (int j = 0; j < hs; ++j) o[i * hs + j] = near[i];
        }
        if (++ystep >= vs) {
            ystep = 0;
            line0 = line1;
            if (++ypos < c->rows) line1 += size_t(c->bw) * 8;
        }
        return result;
    }
};

constexpr int fix20(float x) { return int(x * 4096.0f + 0.5f) << 8; }

inline uint8_t clamp255(int v) { return uint8_t(v < 0 ? 0 : (v > 255 ? 255 : v)); }

// JFIF YCbCr -> RGB with 20 fraction bits and the rounding half added to Y;
// green's Cb term is truncated to a multiple of 2^16 before the sum
inline void ycc_to_rgb(int y, int cb, int cr, uint8_t* o)
{
    const int yf = (y << 20) + (1 << 19);
    cb -= 128;
    cr -= 128;
    const int r = yf + cr * fix20(1.40200f);
    const int g = yf + cr * -fix20(0.71414f) + int(uint32_t(cb * -fix20(0.34414f)) & 0xFFFF0000u);
    const int b = yf + cb * fix20(1.77200f);
    o[0] = clamp255(r >> 20);
    o[1] = clamp255(g >> 20);
    o[2] = clamp255(b >> 20);
}

// x * y / 255, rounded, in 8-bit arithmetic
inline uint8_t mul255(int x, int y)
{
    const int t = x * y + 128;
    return uint8_t((t + (t >> 8)) >> 8);
}

} // namespace

// Decodes a JPEG file held in memory to 8-bit RGBA, top row first.
// `file_comps` is the channel count stb_image reports: 1 for grey, 3 for
// colour (3 or 4 components).  Returns false for data it cannot decode.
bool Decode(const uint8_t* data, size_t size, std::vector<uint8_t>& rgba, int& width, int& height, int& file_comps)
{
    Decoder d(data, size);
    if (!d.parse() || !d.frame) return false;
    if (d.progressive) d.finish_progressive();
    const int X = d.X, Y = d.Y, n = int(d.comps.size());
    std::vector<RowStepper> steps;
    for (const Component& c : d.comps) steps.emplace_back(c, d.hmax, d.vmax);
    const bool rgb_ids = n == 3 && d.comps[0].id == 'R' && d.comps[1].id == 'G' && d.comps[2].id == 'B';
    const bool is_rgb = n == 3 && (rgb_ids || (d.adobe_transform == 0 && !d.jfif));
    rgba.assign(size_t(X) * size_t(Y) * 4, 255);
    for (int y = 0; y < Y; ++y) {
        const uint8_t* row[4];
        for (int k = 0; k < n; ++k) row[k] = steps[size_t(k)].next();
        uint8_t* o = &rgba[size_t(y) * size_t(X) * 4];
        for (int x = 0; x < X; ++x, o += 4) {
            if (n == 1) {
                o[0] = o[1] = o[2] = row[0][x];
            } else if (is_rgb) {
                o[0] = row[0][x];
                o[1] = row[1][x];
                o[2] = row[2][x];
            } else if (n == 4 && d.adobe_transform == 0) { // CMYK, stored inverted by Adobe
                o[0] = mul255(row[0][x], row[3][x]);
                o[1] = mul255(row[1][x], row[3][x]);
                o[2] = mul255(row[2][x], row[3][x]);
            } else {
                ycc_to_rgb(row[0][x], row[1][x], row[2][x], o);
                if (n == 4 && d.adobe_transform == 2) { // YCCK
                    o[0] = mul255(255 - o[0], row[3][x]);
                    o[1] = mul255(255 - o[1], row[3][x]);
                    o[2] = mul255(255 - o[2], row[3][x]);
                }
            }
        }
    }
    width = X;
    height = Y;
    file_comps = n >= 3 ? 3 : 1;
    return true;
}

} // namespace jpeg
} // namespace bulbit
