"""JPEG textures through ReadImage1 / ReadImage3 (src/util/image.cpp:13-130
read them with stb_image's stbi_loadf; the reference's own head scene loads
res/head/lambertian.jpg, cli/scene/head.cpp:12).

stb_image is not vendored in the reference and is absent here, so the
decoder is pinned by what the JPEG standard (ITU-T T.81) and the
mathematics fix, with files written here from chosen quantized
coefficients:
  - blocks with only a DC term decode exactly to 128 + DC / 8;
  - each of the 64 DCT basis functions decodes within one level of the
    T.81 A.3.3 inverse DCT evaluated in float64 (a wrong zigzag entry, a
    transposed block or a wrong constant fails);
  - the same coefficients coded as one interleaved baseline scan, as one
    scan per component, with restart markers, with the Annex K tables or
    with tables whose codes run to 16 bits, and as a progressive script
    using all four kinds of progressive scan, decode to the same image,
    bit for bit;
  - chroma upsampling, read through an RGB-coded file so that no colour
    conversion intervenes, is the bilinear interpolation at centred sample
    positions, rounded half up, for 2x1, 1x2 and 2x2 ratios, and pixel
    replication for others;
  - YCbCr -> RGB is within one level of the JFIF equations; CMYK is the
    rounded product c * k / 255.
Files that libjpeg-turbo writes (through PIL: baseline, progressive, 4:4:4,
4:2:2, 4:2:0, restart markers, grey) decode within a level or two of
libjpeg-turbo's own decoding of them.
"""

import io
import struct
import subprocess

import numpy as np
import pytest

from bulbit_amd import build

from test_cpp_api import read_pfm
from test_model_loader import _decode8, _exe

# T.81 Figure A.6: the zigzag sequence as row-major indices within a block
ZIGZAG = [0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5,
          12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14, 21, 28,
          35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
          58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63]

# T.81 Annex K.3 tables: code-length counts and symbols
K_DC_L = ([0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0], list(range(12)))
K_DC_C = ([0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0], list(range(12)))
K_AC_L = ([0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7D], bytes.fromhex(
    "01020300041105122131410613516107227114328191a1082342b1c11552d1f02433627282090a161718191a25262728292a3435363738"
    "393a434445464748494a535455565758595a636465666768696a737475767778797a838485868788898a92939495969798999aa2a3a4"
    "a5a6a7a8a9aab2b3b4b5b6b7b8b9bac2c3c4c5c6c7c8c9cad2d3d4d5d6d7d8d9dae1e2e3e4e5e6e7e8e9eaf1f2f3f4f5f6f7f8f9fa"))
K_AC_C = ([0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77], bytes.fromhex(
    "000102031104052131061241510761711322328108144291a1b1c109233352f0156272d10a162434e125f11718191a262728292a3536"
    "3738393a434445464748494a535455565758595a636465666768696a737475767778797a82838485868788898a92939495969798999a"
    "a2a3a4a5a6a7a8a9aab2b3b4b5b6b7b8b9bac2c3c4c5c6c7c8c9cad2d3d4d5d6d7d8d9dae2e3e4e5e6e7e8e9eaf2f3f4f5f6f7f8f9fa"))


# ------------------------------------------------------------------ writer

def _cat(v):
    return abs(int(v)).bit_length()


def _low_bits(v, c):
    return int(v) if v >= 0 else int(v) + (1 << c) - 1


class _Bits:
    def __init__(self):
        self.out = bytearray()
        self.acc = 0
        self.n = 0

    def put(self, v, n):
        for i in range(n - 1, -1, -1):
            self.acc = (self.acc << 1) | ((v >> i) & 1)
            self.n += 1
            if self.n == 8:
                self.out.append(self.acc)
                if self.acc == 0xFF:
                    self.out.append(0)  # byte stuffing
                self.acc = 0
                self.n = 0

    def flush(self):
        while self.n:
            self.put(1, 1)


def _table(symbols, style):
    """a canonical table over the symbols used: 'flat' gives every symbol the
    same length; 'deep' gives one symbol 1 bit and the rest 16 bits"""
    syms = sorted(set(symbols))
    counts = [0] * 16
    if style == "deep" and len(syms) > 1:
        counts[0], counts[15] = 1, len(syms) - 1
    else:
        counts[max(1, len(syms).bit_length()) - 1] = len(syms)
    return counts, syms


def _codes(counts, syms):
    codes, code, k = {}, 0, 0
    for length in range(1, 17):
        for _ in range(counts[length - 1]):
            codes[syms[k]] = (code, length)
            code += 1
            k += 1
        code <<= 1
    return codes


class _Prog:
    """end-of-band run state of a progressive AC scan (T.81 G.1.2.2)"""

    def __init__(self, ops, t):
        self.ops, self.t, self.run, self.be = ops, t, 0, []

    def flush(self):
        if self.run:
            r = self.run.bit_length() - 1
            self.ops += [("sym", self.t, r << 4), ("bits", self.run - (1 << r), r)]
            self.ops += [("bits", b, 1) for b in self.be]
            self.run, self.be = 0, []

    def count(self, bits=()):
        self.run += 1
        self.be += list(bits)
        if self.run == 0x7FFF:
            self.flush()


def _seq_block(ops, z, pred, td, ta):
    d = z[0] - pred
    c = _cat(d)
    ops += [("sym", td, c), ("bits", _low_bits(d, c), c)]
    run = 0
    for v in z[1:]:
        if v == 0:
            run += 1
            continue
        while run > 15:
            ops.append(("sym", ta, 0xF0))
            run -= 16
        c = _cat(v)
        ops += [("sym", ta, (run << 4) | c), ("bits", _low_bits(v, c), c)]
        run = 0
    if run:
        ops.append(("sym", ta, 0x00))
    return z[0]


def _ac_first(st, band, al):
    v = [(abs(x) >> al) * (1 if x >= 0 else -1) for x in band]
    nz = [i for i, x in enumerate(v) if x]
    if not nz:
        st.count()
        return
    st.flush()
    run = 0
    for x in v[:nz[-1] + 1]:
        if x == 0:
            run += 1
            continue
        while run > 15:
            st.ops.append(("sym", st.t, 0xF0))
            run -= 16
        c = _cat(x)
        st.ops += [("sym", st.t, (run << 4) | c), ("bits", _low_bits(x, c), c)]
        run = 0
    if nz[-1] < len(v) - 1:
        st.count()


def _ac_refine(st, band, al):
    a = [abs(x) >> al for x in band]
    last_new = max([i for i, x in enumerate(a) if x == 1], default=-1)
    r, br = 0, []
    for i, x in enumerate(a):
        if x == 0:
            r += 1
            continue
        while r > 15 and i <= last_new:
            st.flush()
            st.ops.append(("sym", st.t, 0xF0))
            st.ops += [("bits", b, 1) for b in br]
            r -= 16
            br = []
        if x > 1:
            br.append(x & 1)  # correction bit of a coefficient already nonzero
            continue
        st.flush()
        st.ops += [("sym", st.t, (r << 4) | 1), ("bits", 1 if band[i] > 0 else 0, 1)]
        st.ops += [("bits", b, 1) for b in br]
        r, br = 0, []
    if r > 0 or br:
        st.count(br)


def _segment(marker, payload):
    return bytes([0xFF, marker]) + struct.pack(">H", len(payload) + 2) + payload


def write_jpeg(width, height, comps, coefs, quant, scans=None, progressive=False, restart=0, tables="flat",
               app=b"", q16=False):
    """comps: [(id, h, v, quant table)]; coefs[i]: (blocks down, blocks across,
    64) quantized coefficients in row-major block order, over the MCU-padded
    grid; quant: {table: 64 values in row-major order}; scans: [(component
    indices, Ss, Se, Ah, Al)], default one interleaved sequential scan"""
    hmax = max(c[1] for c in comps)
    vmax = max(c[2] for c in comps)
    mcux, mcuy = -(-width // (8 * hmax)), -(-height // (8 * vmax))
    cw = [-(-width * c[1] // hmax) for c in comps]
    ch = [-(-height * c[2] // vmax) for c in comps]
    if scans is None:
        scans = [(list(range(len(comps))), 0, 63, 0, 0)]
    out = b"\xff\xd8" + app
    for t, q in quant.items():
        zq = [int(q[ZIGZAG[k]]) for k in range(64)]
        out += _segment(0xDB, bytes([(16 if q16 else 0) | t]) + (struct.pack(">64H", *zq) if q16 else bytes(zq)))
    sof = 0xC2 if progressive else 0xC0
    out += _segment(sof, struct.pack(">BHHB", 8, height, width, len(comps)) +
                    b"".join(bytes([c[0], (c[1] << 4) | c[2], c[3]]) for c in comps))
    if restart:
        out += _segment(0xDD, struct.pack(">H", restart))
    for idx, ss, se, ah, al in scans:
        ops = []
        single = len(idx) == 1
        ux = -(-cw[idx[0]] // 8) if single else mcux
        uy = -(-ch[idx[0]] // 8) if single else mcuy
        pred = {i: 0 for i in idx}
        st = _Prog(ops, ("ac", 0))
        n = 0
        for by in range(uy):
            for bx in range(ux):
                for j, i in enumerate(idx):
                    nh, nv = (1, 1) if single else comps[i][1:3]
                    for v in range(nv):
                        for h in range(nh):
                            blk = coefs[i][by * nv + v, bx * nh + h]
                            z = [int(blk[ZIGZAG[k]]) for k in range(64)]
                            if not progressive:
                                pred[i] = _seq_block(ops, z, pred[i], ("dc", j), ("ac", j))
                            elif ss == 0 and ah == 0:
                                d = (z[0] >> al) - pred[i]
                                pred[i] = z[0] >> al
                                c = _cat(d)
                                ops += [("sym", ("dc", j), c), ("bits", _low_bits(d, c), c)]
                            elif ss == 0:
                                ops.append(("bits", (z[0] >> al) & 1, 1))
                            elif ah == 0:
                                _ac_first(st, z[ss:se + 1], al)
                            else:
                                _ac_refine(st, z[ss:se + 1], al)
                n += 1
                if restart and n % restart == 0 and n < ux * uy:
                    st.flush()
                    ops.append(("rst", (n // restart - 1) % 8))
                    pred = {i: 0 for i in idx}
        st.flush()
        # tables for this scan, one pair per scan component
        used = {}
        for op in ops:
            if op[0] == "sym":
                used.setdefault(op[1], []).append(op[2])
        codes = {}
        dht = b""
        for (kind, j), syms in sorted(used.items()):
            if tables == "annexk":
                counts, vals = {("dc", 0): K_DC_L, ("dc", 1): K_DC_C, ("ac", 0): K_AC_L,
                                ("ac", 1): K_AC_C}[(kind, min(j, 1))]
                vals = list(vals)
            else:
                counts, vals = _table(syms, "deep" if (tables == "deep" and j % 2 == 0) else "flat")
            codes[(kind, j)] = _codes(counts, vals)
            dht += bytes([(16 if kind == "ac" else 0) | j]) + bytes(counts) + bytes(vals)
        if dht:
            out += _segment(0xC4, dht)
        out += _segment(0xDA, bytes([len(idx)]) + b"".join(bytes([comps[i][0], (j << 4) | j]) for j, i in enumerate(idx))
                        + bytes([ss, se, (ah << 4) | al]))
        bits = _Bits()
        for op in ops:
            if op[0] == "sym":
                bits.put(*codes[op[1]][op[2]])
            elif op[0] == "bits":
                bits.put(op[1], op[2])
            else:
                bits.flush()
                bits.out += bytes([0xFF, 0xD0 + op[1]])
        bits.flush()
        out += bytes(bits.out)
    return out + b"\xff\xd9"


# ------------------------------------------------------------------ reading

def _read(tmp_path, data, name="a.jpg", channel=None, non_color=True):
    """the decoded image as 8-bit values (non-colour reads are c / 255), top row first"""
    src = tmp_path / name
    src.write_bytes(data)
    out = tmp_path / "out.pfm"
    args = [_exe(build.CPP_CONVERT_IMAGE), str(src), str(out)] + (["non_color"] if non_color else [])
    if channel is not None:
        args.append(f"channel={channel}")
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    if r.returncode != 0:
        return r.returncode
    img = read_pfm(out)[::-1]
    if not non_color:
        return img
    c = np.rint(img.astype(np.float64) * 255.0)
    assert np.array_equal((c.astype(np.float32) / np.float32(255)), img)  # exactly c / 255
    return c.astype(int)


def _idct(block):
    """T.81 A.3.3 inverse DCT of a row-major 8x8 block of dequantized coefficients, in float64"""
    u = np.arange(8)
    basis = np.cos((2 * u[None, :] + 1) * u[:, None] * np.pi / 16) * np.where(u == 0, np.sqrt(0.5), 1.0)[:, None] / 2
    return basis.T @ block.reshape(8, 8) @ basis


def _plane(coefs, quant, width, height):
    """float64 samples of a component from its blocks"""
    bh, bw = coefs.shape[:2]
    p = np.zeros((bh * 8, bw * 8))
    for y in range(bh):
        for x in range(bw):
            p[y * 8:y * 8 + 8, x * 8:x * 8 + 8] = _idct(coefs[y, x] * quant) + 128
    return p[:height, :width]


rng = np.random.default_rng(5)


def _random_coefs(bh, bw, scale=1.0):
    """coefficients of a plausible picture: DC spread, AC falling with frequency"""
    c = np.zeros((bh, bw, 64), np.int64)
    c[..., 0] = rng.integers(-60, 60, (bh, bw))
    f = (np.arange(8)[:, None] + np.arange(8)[None, :]).reshape(64)[1:]
    amp = np.maximum(1, (40 * scale / (1 + f))).astype(int)
    c[..., 1:] = rng.integers(-amp, amp + 1, (bh, bw, 63)) * (rng.random((bh, bw, 63)) < 0.5)
    return c


ONES = np.ones(64, np.int64)


# ------------------------------------------------------------------ tests

def test_zigzag_table_is_the_standard_walk():
    # the anti-diagonal walk the decoder builds, written out here from T.81 Figure A.6
    assert sorted(ZIGZAG) == list(range(64))
    for k in range(1, 64):
        a, b = divmod(ZIGZAG[k - 1], 8), divmod(ZIGZAG[k], 8)
        assert abs(a[0] - b[0]) <= 1 and abs(a[1] - b[1]) <= 1
        assert sum(a) <= sum(b)


def test_dc_only_blocks_are_exact(tmp_path):
    # 16-bit quantization values too: DC 264 = 8 x 33 per step
    q = np.ones(64, np.int64)
    q[0] = 264
    c = np.zeros((3, 5, 64), np.int64)
    c[..., 0] = rng.integers(-3, 4, (3, 5))
    img = _read(tmp_path, write_jpeg(37, 21, [(1, 1, 1, 0)], [c], {0: q}, q16=True))
    want = np.clip(128 + 33 * np.repeat(np.repeat(c[..., 0], 8, 0), 8, 1), 0, 255)[:21, :37]
    assert np.array_equal(img[..., 0], want)
    assert np.array_equal(img[..., 1], want) and np.array_equal(img[..., 2], want)  # grey replicated


def test_every_basis_function_is_within_one_level_of_the_float_idct(tmp_path):
    c = np.zeros((8, 8, 64), np.int64)
    for k in range(64):
        blk = c[k // 8, k % 8]
        blk[ZIGZAG[k]] = 160 if k else 400  # one coefficient per block, in zigzag order
        blk[0] += 0 if k == 0 else -80
    for tables in ("annexk", "deep"):
        img = _read(tmp_path, write_jpeg(64, 64, [(1, 1, 1, 0)], [c], {0: ONES}, tables=tables))[..., 0]
        want = np.clip(np.floor(_plane(c, ONES, 64, 64) + 0.5), 0, 255)
        d = np.abs(img - want)
        assert d.max() <= 1, (tables, np.argwhere(d > 1)[:4])
        assert np.mean(d == 0) > 0.93


def test_random_blocks_are_within_one_level_of_the_float_idct(tmp_path):
    q = rng.integers(1, 9, 64)
    c = _random_coefs(4, 6)
    img = _read(tmp_path, write_jpeg(45, 27, [(1, 1, 1, 0)], [c], {0: q}, tables="annexk"))[..., 0]
    want = np.clip(np.floor(_plane(c, q, 45, 27) + 0.5), 0, 255)
    d = np.abs(img - want)
    assert d.max() <= 1 and np.mean(d == 0) > 0.95


def _colour_file(**kw):
    comps = [(1, 2, 2, 0), (2, 1, 1, 1), (3, 1, 1, 1)]
    w, h = 53, 35
    mx, my = -(-w // 16), -(-h // 16)
    coefs = [_random_coefs(my * 2, mx * 2), _random_coefs(my, mx, 0.5), _random_coefs(my, mx, 0.5)]
    quant = {0: rng.integers(1, 6, 64), 1: rng.integers(1, 9, 64)}
    return w, h, comps, coefs, quant


PROGRESSIVE = [([0, 1, 2], 0, 0, 0, 1), ([0], 1, 5, 0, 2), ([2], 1, 63, 0, 1), ([1], 1, 63, 0, 1),
               ([0], 6, 63, 0, 2), ([0], 1, 63, 2, 1), ([0, 1, 2], 0, 0, 1, 0), ([2], 1, 63, 1, 0),
               ([1], 1, 63, 1, 0), ([0], 1, 63, 1, 0)]


def test_every_coding_of_the_same_coefficients_decodes_the_same(tmp_path):
    w, h, comps, coefs, quant = _colour_file()
    base = _read(tmp_path, write_jpeg(w, h, comps, coefs, quant))
    variants = {
        "annex K tables": dict(tables="annexk"),
        "16-bit codes": dict(tables="deep"),
        "restart markers": dict(restart=2),
        "a scan per component": dict(scans=[([0], 0, 63, 0, 0), ([2], 0, 63, 0, 0), ([1], 0, 63, 0, 0)], restart=5),
        "progressive": dict(progressive=True, scans=PROGRESSIVE),
        "progressive, restarts": dict(progressive=True, scans=PROGRESSIVE, restart=3, tables="deep"),
        "progressive, spectral selection only": dict(progressive=True, scans=[
            ([0, 1, 2], 0, 0, 0, 0), ([0], 1, 9, 0, 0), ([0], 10, 63, 0, 0), ([1], 1, 63, 0, 0), ([2], 1, 63, 0, 0)]),
    }
    for name, kw in variants.items():
        got = _read(tmp_path, write_jpeg(w, h, comps, coefs, quant, **kw))
        assert np.array_equal(got, base), name


def test_progressive_scans_rebuild_the_coefficients(tmp_path):
    # a grey progressive file with successive approximation down to bit 0
    # decodes like its baseline coding, and both match the float IDCT
    c = _random_coefs(3, 4, 2.0)
    c[0, 0, 1:20] = 0  # long zero runs: ZRL and end-of-band runs in every scan kind
    c[1, 2, 5] = 300
    scans = [([0], 0, 0, 0, 3), ([0], 1, 63, 0, 3), ([0], 0, 0, 3, 2), ([0], 1, 63, 3, 2), ([0], 0, 0, 2, 1),
             ([0], 1, 63, 2, 1), ([0], 0, 0, 1, 0), ([0], 1, 63, 1, 0)]
    prog = _read(tmp_path, write_jpeg(30, 20, [(7, 1, 1, 0)], [c], {0: ONES}, scans=scans, progressive=True))
    base = _read(tmp_path, write_jpeg(30, 20, [(7, 1, 1, 0)], [c], {0: ONES}))
    assert np.array_equal(prog, base)
    want = np.clip(np.floor(_plane(c, ONES, 30, 20) + 0.5), 0, 255)
    assert np.abs(prog[..., 0] - want).max() <= 1


def _upsample(p, hs, vs, width, height):
    """bilinear at centred sample positions (3/4 of the nearer sample, 1/4 of
    the farther, edges clamped) for 2x1, 1x2 and 2x2; replication for every
    other pair of ratios; rounded half up"""
    fancy = (hs, vs) in ((2, 1), (1, 2), (2, 2))

    def axis(n_out, n_in, s):
        x = np.arange(n_out)
        if s != 2 or not fancy:
            i = x // s
            return i, i, np.zeros(n_out)
        pos = (x + 0.5) / 2 - 0.5
        i0 = np.floor(pos).astype(int)
        f = pos - i0
        return np.clip(i0, 0, n_in - 1), np.clip(i0 + 1, 0, n_in - 1), f

    y0, y1, fy = axis(height, p.shape[0], vs)
    x0, x1, fx = axis(width, p.shape[1], hs)
    a = p[y0][:, x0] * (1 - fx) + p[y0][:, x1] * fx
    b = p[y1][:, x0] * (1 - fx) + p[y1][:, x1] * fx
    return np.floor(a * (1 - fy)[:, None] + b * fy[:, None] + 0.5).astype(int)


@pytest.mark.parametrize("hs,vs", [(2, 2), (2, 1), (1, 2), (3, 1), (4, 2)])
def test_chroma_upsampling_is_centred_bilinear(tmp_path, hs, vs):
    # component ids R, G, B: the samples are colours and no conversion applies
    w, h = 47, 29
    comps = [(ord("R"), hs, vs, 0), (ord("G"), 1, 1, 0), (ord("B"), 1, 1, 0)]
    mx, my = -(-w // (8 * hs)), -(-h // (8 * vs))
    coefs = [_random_coefs(my * vs, mx * hs), _random_coefs(my, mx), _random_coefs(my, mx)]
    img = _read(tmp_path, write_jpeg(w, h, comps, coefs, {0: ONES}))
    cw, ch = -(-w // hs), -(-h // vs)
    for k in (1, 2):
        # the subsampled plane as the decoder reconstructs it: the same blocks as a grey image
        plane = _read(tmp_path, write_jpeg(cw, ch, [(1, 1, 1, 0)], [coefs[k]], {0: ONES}), "p.jpg")[..., 0]
        assert np.array_equal(img[..., k], _upsample(plane.astype(np.float64), hs, vs, w, h)), k
    full = _read(tmp_path, write_jpeg(w, h, [(1, 1, 1, 0)], [coefs[0]], {0: ONES}), "f.jpg")[..., 0]
    assert np.array_equal(img[..., 0], full)


def test_ycbcr_is_converted_by_the_jfif_equations(tmp_path):
    w, h = 40, 24
    comps = [(1, 1, 1, 0), (2, 1, 1, 0), (3, 1, 1, 0)]
    coefs = [_random_coefs(3, 5, 2.0) for _ in range(3)]
    img = _read(tmp_path, write_jpeg(w, h, comps, coefs, {0: ONES}))
    planes = [_read(tmp_path, write_jpeg(w, h, [(1, 1, 1, 0)], [c], {0: ONES}), "p.jpg")[..., 0].astype(np.float64)
              for c in coefs]
    y, cb, cr = planes[0], planes[1] - 128, planes[2] - 128
    want = np.stack([y + 1.402 * cr, y - 0.344136 * cb - 0.714136 * cr, y + 1.772 * cb], -1)
    want = np.clip(np.floor(want + 0.5), 0, 255)
    d = np.abs(img - want)
    assert d.max() <= 1 and np.mean(d == 0) > 0.9
    grey = (np.abs(cb) < 0.5) & (np.abs(cr) < 0.5)
    assert np.array_equal(img[grey], np.repeat(y[grey][:, None], 3, 1))  # the grey axis is exact


@pytest.mark.parametrize("transform", [0, 2])
def test_adobe_cmyk_and_ycck(tmp_path, transform):
    w, h = 24, 16
    comps = [(1, 1, 1, 0), (2, 1, 1, 0), (3, 1, 1, 0), (4, 1, 1, 0)]
    coefs = [_random_coefs(2, 3, 2.0) for _ in range(4)]
    adobe = _segment(0xEE, b"Adobe" + bytes([0, 100, 0, 0, 0, 0, transform]))
    img = _read(tmp_path, write_jpeg(w, h, comps, coefs, {0: ONES}, app=adobe))
    planes = [_read(tmp_path, write_jpeg(w, h, [(1, 1, 1, 0)], [c], {0: ONES}), "p.jpg")[..., 0] for c in coefs]
    k = planes[3]
    if transform == 0:
        colours = planes[:3]
    else:
        ycc = _read(tmp_path, write_jpeg(w, h, comps[:3], coefs[:3], {0: ONES}), "y.jpg")
        colours = [255 - ycc[..., i] for i in range(3)]
    for i in range(3):
        assert np.array_equal(img[..., i], np.floor(colours[i] * k / 255 + 0.5).astype(int)), i


def test_channels_and_gamma(tmp_path):
    c = _random_coefs(2, 2)
    grey = write_jpeg(16, 16, [(1, 1, 1, 0)], [c], {0: ONES})
    assert _read(tmp_path, grey, channel=1) == 3  # a grey file has one channel, as stb reports
    assert isinstance(_read(tmp_path, grey, channel=0), np.ndarray)
    w, h, comps, coefs, quant = _colour_file()
    colour = write_jpeg(w, h, comps, coefs, quant)
    assert _read(tmp_path, colour, channel=3) == 3  # no alpha channel
    lin = _read(tmp_path, colour, non_color=False)
    c8 = _read(tmp_path, colour)
    assert np.array_equal(lin, _decode8(c8.astype(np.uint8), False))  # pow(c / 255, 2.2)
    for ch in range(3):
        one = _read(tmp_path, colour, channel=ch)
        assert np.array_equal(one[..., 0], c8[..., ch])


def test_broken_and_unsupported_files_are_rejected(tmp_path):
    c = _random_coefs(2, 2)
    good = write_jpeg(16, 16, [(1, 1, 1, 0)], [c], {0: ONES})
    assert isinstance(_read(tmp_path, good), np.ndarray)
    assert _read(tmp_path, good[:40]) == 1  # cut inside the tables: no scan
    assert _read(tmp_path, good.replace(b"\xff\xc0", b"\xff\xc9", 1)) == 1  # arithmetic coding
    assert _read(tmp_path, good.replace(b"\xff\xc0\x00\x0b\x08", b"\xff\xc0\x00\x0b\x0c", 1)) == 1  # 12-bit
    assert _read(tmp_path, b"\xff\xd8\xff\xd9") == 1
    # data that stops inside the last scan still decodes, its missing bits read as zeros
    assert isinstance(_read(tmp_path, good[:-6]), np.ndarray)


# ------------------------------------------------ against libjpeg-turbo

def _picture(h, w, grey=False):
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    r = 128 + 100 * np.sin(x / 7.0) * np.cos(y / 5.0)
    g = 128 + 90 * np.cos((x + y) / 9.0)
    b = 255 * ((x // 6 + y // 4) % 2) * 0.7 + 30
    img = np.stack([r, g, b], -1) + rng.normal(0, 6, (h, w, 3))
    img = np.clip(np.rint(img), 0, 255).astype(np.uint8)
    return img[..., 1] if grey else img


@pytest.mark.parametrize("mode,sub,progressive,restart", [
    ("L", 0, False, 0), ("L", 0, True, 3), ("RGB", 0, False, 0), ("RGB", 1, True, 0), ("RGB", 2, False, 4),
    ("RGB", 2, True, 2)])
def test_libjpeg_written_files_decode_like_libjpeg(tmp_path, mode, sub, progressive, restart):
    pil = pytest.importorskip("PIL.Image")
    src = _picture(45, 71, mode == "L")
    buf = io.BytesIO()
    kw = dict(quality=88, progressive=progressive, optimize=True)
    if mode == "RGB":
        kw["subsampling"] = sub
    if restart:
        kw["restart_marker_blocks"] = restart
    pil.fromarray(src, mode).save(buf, "JPEG", **kw)
    ref = np.asarray(pil.open(io.BytesIO(buf.getvalue())).convert("RGB")).astype(int)
    got = _read(tmp_path, buf.getvalue())
    d = np.abs(got - ref)
    # libjpeg-turbo's islow IDCT keeps 13-bit constants (12 here), its colour
    # tables round each term, and its 2x2 upsampling alternates its bias
    assert d.max() <= (1 if mode == "L" else 3), d.max()
    assert d.mean() < 0.25, d.mean()


def test_the_format_comes_from_the_content_not_the_name(tmp_path):
    # stb_image picks its decoder from the file's signature: a PNG saved as
    # .jpg, a JPEG saved as .png, an HDR or PPM under any name read as they are
    import test_png
    from test_model_loader import _write_hdr, _write_pnm

    c = _random_coefs(2, 3)
    jpg = write_jpeg(20, 12, [(1, 1, 1, 0)], [c], {0: ONES})
    assert np.array_equal(_read(tmp_path, jpg, "a.png"), _read(tmp_path, jpg, "a.jpg"))
    rng = np.random.default_rng(4)
    test_png.write_png(tmp_path / "p.png", rng.integers(0, 256, (5, 6, 3)).astype(np.uint8), 2, 8)
    png = (tmp_path / "p.png").read_bytes()
    assert np.array_equal(_read(tmp_path, png, "p.jpg"), _read(tmp_path, png, "p2.png"))
    _write_hdr(tmp_path / "h.hdr", rng.integers(1, 256, (3, 4, 4)))
    _write_pnm(tmp_path / "i.ppm", rng.integers(0, 256, (3, 4, 3)).astype(np.uint8))
    for name, other in (("h.hdr", "h.img"), ("i.ppm", "i.texture")):
        data = (tmp_path / name).read_bytes()
        a = _read(tmp_path, data, "x" + name[1:], non_color=False)
        b = _read(tmp_path, data, other, non_color=False)
        assert np.array_equal(a, b), name
    assert _read(tmp_path, b"not an image at all", "n.png") == 1
