"""WriteImage of the C++ API (src/util/image.cpp:219-266): .png and .jpg
tone-mapped with sRGB_from_RGB (include/bulbit/color.h:31-50) and quantized
as uint8(min(clamp(v, 0, 1) * 256, 255)), linear .hdr (RGBE) and .pfm.  The
film's row 0 is the bottom of the picture, so top-down formats start with
the last row.  Files come from bulbit_amd/cpp/examples/write_images.cpp and
are decoded independently: PNG and JPEG with PIL, RGBE and PFM here."""
import os
import subprocess

import numpy as np
import pytest

from bulbit_amd import build

W, H = 37, 21


def _linear():
    """the image write_images.cpp builds, [y][x] with y = 0 the bottom row"""
    x = np.arange(W, dtype=np.float32)[None, :]
    y = np.arange(H, dtype=np.float32)[:, None]
    r = x / np.float32(W - 1) * np.float32(1.2) + 0 * y
    g = y / np.float32(H - 1) + 0 * x
    b = np.float32(0.01) * ((np.arange(W)[None, :] * 7 + np.arange(H)[:, None] * 13) % 11).astype(np.float32)
    return np.stack([r, g, b], -1).astype(np.float32)


def _srgb8(lin):
    """IEC 61966-2-1 transfer curve in float64, then the reference's 8-bit rule"""
    v = lin.astype(np.float64)
    s = np.where(v <= 0.0031308, 12.92 * v, 1.055 * np.power(np.clip(v, 0, None), 1 / 2.4) - 0.055)
    return np.minimum(np.clip(s, 0, 1) * 256.0, 255.0).astype(np.uint8)


@pytest.fixture(scope="module")
def written(tmp_path_factory):
    if not os.path.exists(build.CPP_WRITE_IMAGES):
        build.build_cpp()
    prefix = str(tmp_path_factory.mktemp("img") / "wi")
    r = subprocess.run([build.CPP_WRITE_IMAGES, prefix], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return prefix


def test_png_is_the_tone_mapped_image_top_row_first(written):
    from PIL import Image

    png = np.asarray(Image.open(written + ".png").convert("RGB"))
    assert png.shape == (H, W, 3)
    want = _srgb8(_linear())[::-1]  # top row of the picture first
    diff = np.abs(png.astype(int) - want.astype(int))
    assert diff.max() <= 1 and np.mean(diff == 0) > 0.99  # float32 powf vs float64 at a rounding edge
    assert np.all(png[0, :, 1] == 255) and np.all(png[-1, :, 1] == 0)  # g = y / (H - 1): 1 at the top
    assert np.all(png[:, -1, 0] == 255)  # r = 1.2 clips


def test_jpg_decodes_close_to_the_png(written):
    from PIL import Image

    jpg = Image.open(written + ".jpg")
    assert jpg.format == "JPEG" and jpg.size == (W, H)
    a = np.asarray(jpg.convert("RGB")).astype(int)
    b = np.asarray(Image.open(written + ".png").convert("RGB")).astype(int)
    d = np.abs(a - b)
    assert d.max() <= 4 and d.mean() < 1.0


def test_hdr_rgbe_holds_linear_radiance_top_row_first(written):
    data = open(written + ".hdr", "rb").read()
    head, body = data.split(b"\n\n", 1)
    assert head.startswith(b"#?RADIANCE") and b"FORMAT=32-bit_rle_rgbe" in head
    res, pix = body.split(b"\n", 1)
    assert res == b"-Y %d +X %d" % (H, W)
    rgbe = np.frombuffer(pix, np.uint8).reshape(H, W, 4).astype(np.float64)
    scale = np.where(rgbe[..., 3] > 0, np.ldexp(1.0, (rgbe[..., 3] - 136).astype(int)), 0.0)
    img = (rgbe[..., :3] + 0.5) * scale[..., None] * (rgbe[..., 3:] > 0)
    lin = _linear()[::-1].astype(np.float64)
    m = lin.max(-1, keepdims=True)
    assert np.all(np.abs(img - lin) <= m / 128 + 1e-30)


def test_pfm_holds_the_film_rows_bottom_first(written):
    with open(written + ".pfm", "rb") as f:
        assert f.readline().strip() == b"PF"
        assert tuple(map(int, f.readline().split())) == (W, H)
        assert float(f.readline()) < 0  # little-endian
        img = np.frombuffer(f.read(), "<f4").reshape(H, W, 3)  # PFM rows run bottom to top
    assert np.array_equal(img, _linear())
