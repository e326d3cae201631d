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


def textured_readme_scene(width, height):
    """The scene of examples/textured.cpp built through the Python mirror:
    the README scene, the sphere's reflectance an image texture and a normal
    map, plus a checker-textured ground quad with texture coordinates."""
    f = np.float32
    sc, cam = scenes.readme_scene(width=width, height=height)
    albedo = np.zeros((8, 16, 3), np.float32)
    bumps = np.zeros((8, 8, 3), np.float32)
    for y in range(8):
        for x in range(16):
            albedo[y, x] = (f(0.2) + f(0.04) * f(x), f(0.2) + f(0.08) * f(y), f(0.5))
        for x in range(8):
            bumps[y, x] = (0.7 if x % 2 else 0.3, 0.5, 1.0)
    sc.materials[0].tex_color = sc.add_texture(scenes.image_texture(albedo))
    sc.materials[0].tex_normal = sc.add_texture(scenes.image_texture(bumps))
    a = sc.add_texture(scenes.constant_texture((0.8, 0.8, 0.8)))
    b = sc.add_texture(scenes.constant_texture((0.1, 0.1, 0.15)))
    ground = scenes.diffuse((0.0, 0.0, 0.0))
    ground.tex_color = sc.add_texture(scenes.checker_texture(a, b, (8.0, 8.0), 3))
    m = sc.add_material(ground)
    sc.add_mesh([[-4, 0, -4], [4, 0, -4], [4, 0, 4], [-4, 0, 4]], np.tile([0.0, 1.0, 0.0], (4, 1)),
                np.tile([1.0, 0.0, 0.0], (4, 1)), [[0, 2, 1], [0, 3, 2]], m, uvs=[[0, 0], [1, 0], [1, 1], [0, 1]])
    return sc, cam


@pytest.mark.gpu
def test_textured_example_equals_c_abi_render_and_oracle(tmp_path):
    """ImageTexture, CheckerTexture and a normal map through the C++ API:
    the example's image equals the C ABI render of the same scene bit for
    bit, and agrees with the oracle."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    exe = build.CPP_TEXTURED
    if not os.path.exists(exe):
        build.build_cpp()
    out = tmp_path / "textured.pfm"
    print("[cpp] textured_example 320 180 8", flush=True)
    r = subprocess.run([exe, "320", "180", "8", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    img = read_pfm(out)
    sc, cam = textured_readme_scene(320, 180)
    s, c, _ = GpuScene(sc).render(cam, scenes.render_params(8))
    ref = s / c[..., None]
    assert np.array_equal(img.view(np.int32), ref.astype(np.float32).view(np.int32))
    os_, oc, _ = OracleScene(sc).render(cam, scenes.render_params(8), threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    frac = pixel_agreement(img, os_ / oc[..., None])
    print(f"[cpp] textured example vs oracle: {frac:.6f} within 1e-3", flush=True)
    assert frac >= 0.999
