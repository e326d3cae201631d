"""Texel indices of ImageTexture stay inside the image for every coordinate.

ImageTexture::Evaluate (include/bulbit/textures.h:60-88) converts
floor(u * width - 0.5) to int32 and FilterTexCoord (include/bulbit/texture.h:26-46)
maps it into [0, n).  The kernels share bb_texcoord.h, compiled here for the
host with g++, which reduces coordinates whose int32 conversion or +1 would
overflow.  Checked: the plain conversion in the int range (against the
reference's formula on Python integers), far coordinates against the exact
integer's texel, and NaN / infinities / the int32 limits inside the image.
"""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPEAT, CLAMP = 0, 1


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    from bulbit_amd import capi

    assert capi.BB_TEXCOORD_REPEAT == REPEAT and capi.BB_TEXCOORD_CLAMP == CLAMP
    exe = str(tmp_path_factory.mktemp("texcoord") / "probe_texcoord")
    subprocess.run(["g++", "-std=c++17", "-O2", "-o", exe, os.path.join(HERE, "probe_texcoord.cpp")], check=True)
    return exe


def run(exe, cases):
    text = "".join(f"{struct.unpack('<I', struct.pack('<f', w))[0]:08x} {n} {f}\n" for w, n, f in cases)
    out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    return [tuple(int(x) for x in line.split()) for line in out if line.strip()]


def ref_texel(i, n, filt):
    """FilterTexCoord on an unbounded integer (texture.h:26-46)"""
    if filt == CLAMP:
        return min(max(i, 0), n - 1)
    return i % n if i >= 0 else n - ((-i) % n) - 1


def test_in_range_coordinates_match_the_reference(probe):
    rng = np.random.default_rng(3)
    cases = []
    for n in (1, 2, 3, 7, 64, 1000):
        for filt in (REPEAT, CLAMP):
            ws = np.concatenate([rng.uniform(-5 * n, 5 * n, 200), np.arange(-2 * n, 2 * n + 1) - 0.5,
                                 np.arange(-2 * n, 2 * n + 1, dtype=np.float64)]).astype(np.float32)
            cases += [(float(w), n, filt) for w in ws]
    for (w, n, filt), (i0, i1, t0, t1) in zip(cases, run(probe, cases)):
        f = math.floor(w)
        assert i0 == f and i1 == f + 1
        assert t0 == ref_texel(f, n, filt) and t1 == ref_texel(f + 1, n, filt)


def test_far_coordinates_read_the_exact_integers_texel(probe):
    cases = []
    for n in (1, 3, 7, 64, 1000, 4096):
        for filt in (REPEAT, CLAMP):
            for e in (30, 31, 32, 40, 60, 100, 127):
                for sign in (1.0, -1.0):
                    for k in (0, 1, 3):
                        w = float(np.float32(sign * 2.0 ** e * (1.0 + k * 2.0 ** -22)))
                        cases.append((w, n, filt))
            # negative multiples of n far out: the reference maps them to n - 1
            cases.append((float(np.float32(-(2.0 ** 40) * n)), n, filt))
    for (w, n, filt), (i0, i1, t0, t1) in zip(cases, run(probe, cases)):
        f = int(w)  # exact: far floats are integers
        assert 0 <= t0 < n and 0 <= t1 < n
        assert t0 == ref_texel(f, n, filt), (w, n, filt)
        assert t1 == ref_texel(f + 1, n, filt), (w, n, filt)
        assert abs(i0) <= 2 ** 30 and abs(i1) <= 2 ** 30 + 1


def test_non_finite_and_int32_limit_coordinates_stay_in_the_image(probe):
    specials = [math.nan, -math.nan, math.inf, -math.inf, float(2 ** 31), float(-2 ** 31), float(2 ** 31 - 128),
                float(np.float32(-2.0 ** 31 - 256)), 3.4e38, -3.4e38]
    cases = [(w, n, filt) for w in specials for n in (1, 5, 256) for filt in (REPEAT, CLAMP)]
    for (w, n, filt), (i0, i1, t0, t1) in zip(cases, run(probe, cases)):
        assert 0 <= t0 < n and 0 <= t1 < n, (w, n, filt, t0, t1)
        if math.isinf(w) and filt == CLAMP:
            assert t0 == t1 == (0 if w < 0 else n - 1)
