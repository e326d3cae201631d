"""The C++ library API (bulbit/bulbit.h) over the C ABI.

The README's library example (README.md:136-190) is compiled unchanged apart
from taking its resolution, spp and output file from argv.  On a GPU its film
must equal, bit for bit, the C ABI render of the same scene built by the
Python mirror (same flattening, same RNG streams) and agree with the oracle.
Without a device it must stop with the library's error, not render anything.
"""
import os
import subprocess

import numpy as np
import pytest

from bulbit_amd import build, scenes

from conftest import pixel_agreement

EXAMPLE = build.CPP_EXAMPLE


def _example():
    if not os.path.exists(EXAMPLE):
        build.build_cpp()
    return EXAMPLE


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), dtype="<f4" if scale < 0 else ">f4")
    # PFM rows run bottom to top, like the film's rows (row 0 is the bottom)
    return data.reshape(h, w, 3)


def test_cpp_api_library_and_example_are_built():
    assert os.path.exists(build.CPP_LIB)
    assert os.path.exists(_example())


def test_readme_example_stops_loudly_without_a_device(bblib, tmp_path):
    import ctypes

    n = ctypes.c_int32(0)
    bblib.bbgpu_device_count(ctypes.byref(n))
    if n.value > 0:
        pytest.skip("a HIP device is present")
    out = tmp_path / "x.pfm"
    r = subprocess.run([_example(), "32", "18", "1", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "no HIP device" in r.stderr
    assert not out.exists()


@pytest.mark.gpu
def test_readme_example_equals_c_abi_render_and_oracle(tmp_path):
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    out = tmp_path / "readme.pfm"
    print("[cpp] readme_example 640 360 16", flush=True)
    r = subprocess.run([_example(), "640", "360", "16", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    img = read_pfm(out)

    sc, cam = scenes.readme_scene(width=640, height=360)
    g = GpuScene(sc)
    s, c, _ = g.render(cam, scenes.render_params(16))
    ref = s / c[..., None]
    assert np.array_equal(img.view(np.int32), ref.astype(np.float32).view(np.int32))

    o = OracleScene(sc)
    os_, oc, _ = o.render(cam, scenes.render_params(16), threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    frac = pixel_agreement(img, os_ / oc[..., None])
    print(f"[cpp] README example vs oracle: {frac:.6f} within 1e-3", flush=True)
    assert frac >= 0.999

    # the same render written as an sRGB PNG: top row of the picture first
    from PIL import Image
    from test_image_io import _srgb8

    png_out = tmp_path / "readme.png"
    r = subprocess.run([_example(), "640", "360", "16", str(png_out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    png = np.asarray(Image.open(png_out).convert("RGB")).astype(int)
    want = _srgb8(ref.astype(np.float32))[::-1].astype(int)
    d = np.abs(png - want)
    assert d.max() <= 1 and np.mean(d == 0) > 0.99
