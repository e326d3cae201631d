"""Scene ingestion in the C++ API: ReadImage3 against stb_image's published
conversions (rows flipped, 8-bit samples decoded as pow(c / 255, 2.2) for
colour and c / 255 for non-colour data, Radiance RGBE as r * 2^(e - 136)),
and LoadModel on an OBJ + MTL scene against the same scene built by hand
with the reference loader's rules (one vertex per face corner, meshes per
material in HashMap bucket order, CoordinateSystem tangents,
MetallicRoughness materials, every triangle an area light with the
material's Ke)."""
import os
import subprocess

import numpy as np
import pytest

from bulbit_amd import build, scenes

from conftest import pixel_agreement
from test_cpp_api import read_pfm

F = np.float32


def _exe(path):
    if not os.path.exists(path):
        build.build_cpp()
    return path


def _convert(tmp_path, src, non_color=False):
    out = tmp_path / (os.path.basename(str(src)) + ".pfm")
    args = [_exe(build.CPP_CONVERT_IMAGE), str(src), str(out)] + (["non_color"] if non_color else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return read_pfm(out)  # row 0 = bottom, as ReadImage returns it


def _decode8(c, non_color):
    """stbi__ldr_to_hdr: (float) pow(c / 255.0f, gamma) with gamma 2.2f or 1"""
    x = c.astype(F) / F(255)
    if non_color:
        return x
    return np.power(x.astype(np.float64), np.float64(F(2.2))).astype(F)


def _write_pnm(path, img, maxval=255):
    h, w = img.shape[:2]
    magic = b"P6" if img.ndim == 3 else b"P5"
    with open(path, "wb") as f:
        f.write(magic + b"\n# test\n%d %d\n%d\n" % (w, h, maxval))
        f.write(img.astype(">u2" if maxval > 255 else np.uint8).tobytes())


def test_read_pnm_gamma_flip_and_channels(tmp_path):
    rng = np.random.default_rng(1)
    rgb = rng.integers(0, 256, (5, 7, 3)).astype(np.uint8)
    _write_pnm(tmp_path / "a.ppm", rgb)
    got = _convert(tmp_path, tmp_path / "a.ppm")
    assert np.array_equal(got.view(np.int32), _decode8(rgb, False)[::-1].view(np.int32))
    got = _convert(tmp_path, tmp_path / "a.ppm", non_color=True)
    assert np.array_equal(got.view(np.int32), _decode8(rgb, True)[::-1].view(np.int32))
    grey = rng.integers(0, 256, (4, 6)).astype(np.uint8)
    _write_pnm(tmp_path / "g.pgm", grey)
    got = _convert(tmp_path, tmp_path / "g.pgm")
    assert np.array_equal(got, np.repeat(_decode8(grey, False)[::-1, :, None], 3, axis=2))
    # 16-bit samples keep their high byte, as stb's 8-bit path does
    wide = rng.integers(0, 65536, (3, 4, 3)).astype(np.uint16)
    _write_pnm(tmp_path / "w.ppm", wide, 65535)
    got = _convert(tmp_path, tmp_path / "w.ppm")
    assert np.array_equal(got, _decode8((wide >> 8).astype(np.uint8), False)[::-1])


def _write_hdr(path, rgbe, rle_rows=()):
    h, w = rgbe.shape[:2]
    with open(path, "wb") as f:
        f.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w))
        for y in range(h):
            if y not in rle_rows:
                f.write(rgbe[y].astype(np.uint8).tobytes())
                continue
            f.write(bytes([2, 2, w >> 8, w & 255]))
            for k in range(4):
                comp = rgbe[y, :, k].astype(np.uint8)
                x = 0
                while x < w:
                    run = 1
                    while x + run < w and run < 127 and comp[x + run] == comp[x]:
                        run += 1
                    if run >= 3:
                        f.write(bytes([128 + run, comp[x]]))
                        x += run
                    else:
                        n = min(w - x, 128)
                        f.write(bytes([n]) + comp[x:x + n].tobytes())
                        x += n


def test_read_radiance_hdr_flat_and_rle(tmp_path):
    rng = np.random.default_rng(2)
    w, h = 16, 6
    rgbe = rng.integers(0, 256, (h, w, 4)).astype(np.int64)
    rgbe[..., 3] = rng.integers(120, 140, (h, w))
    rgbe[2, 3:9, :] = rgbe[2, 3, :]  # a run for the RLE encoder
    rgbe[4, 5, 3] = 0  # exponent 0: black
    _write_hdr(tmp_path / "a.hdr", rgbe, rle_rows=(1, 2, 4))
    got = _convert(tmp_path, tmp_path / "a.hdr")
    f1 = np.ldexp(np.float32(1), (rgbe[..., 3] - 136).astype(np.int32)).astype(F)
    want = (rgbe[..., :3].astype(F) * f1[..., None]).astype(F)
    want[rgbe[..., 3] == 0] = 0
    assert np.array_equal(got, want[::-1])


def test_read_pfm_passes_through(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.uniform(0, 4, (5, 9, 3)).astype(F)
    p = tmp_path / "a.pfm"
    with open(p, "wb") as f:
        f.write(b"PF\n9 5\n-1.0\n" + img.astype("<f4").tobytes())  # stored bottom row first
    assert np.array_equal(_convert(tmp_path, p), img)


# ------------------------------------------------------------------ OBJ

OBJ = """# floor, a two-material box and a lamp
mtllib scene.mtl
v -2 0 -2
v 2 0 -2
v 2 0 2
v -2 0 2
vt 0 0
vt 1 0
vt 1 1
vt 0 1
vn 0 1 0
usemtl floor
f 1/1/1 4/4/1 3/3/1 2/2/1
o box
v -0.5 0 -0.5
v 0.5 0 -0.5
v 0 1.2 0
v 0 0 0.6
vn 0 0.4 -1
vn 0.9 0.4 0.5
vn -0.9 0.4 0.5
usemtl blue
f 5//2 7//2 6//2
f 6//3 7//3 8//3
usemtl red
f 8//4 7//4 5//4
o lamp
v -0.6 3 -0.6
v 0.6 3 -0.6
v 0 3 0.6
vn 0 -1 0
usemtl lamp
f -3//5 -2//5 -1//5
"""

MTL = """newmtl floor
Kd 0.5 0.5 0.5
map_Kd floor.ppm
bump -bm 1.0 floor_n.ppm
newmtl red
Kd 0.8 0.2 0.2
newmtl lamp
Ke 12 12 10
newmtl blue
Kd 0.2 0.3 0.8
"""

NAMES = ["floor", "red", "lamp", "blue"]


def _coordinate_system(n):
    """CoordinateSystem (include/bulbit/frame.h:105-112), float32"""
    n = n.astype(F)
    sign = np.copysign(F(1), n[2])
    a = F(-1) / (sign + n[2])
    b = n[0] * n[1] * a
    return np.array([F(1) + sign * (n[0] * n[0]) * a, sign * b, -sign * n[0]], F)


def _expected_scene(tmp_path, floor, floor_n, ty, width, height):
    """The scene LoadOBJ builds from OBJ/MTL above, constructed by hand"""
    v = np.array([[-2, 0, -2], [2, 0, -2], [2, 0, 2], [-2, 0, 2], [-0.5, 0, -0.5], [0.5, 0, -0.5], [0, 1.2, 0],
                  [0, 0, 0.6], [-0.6, 3, -0.6], [0.6, 3, -0.6], [0, 3, 0.6]], F)
    vt = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], F)
    vn = np.array([[0, 1, 0], [0, 0.4, -1], [0.9, 0.4, 0.5], [-0.9, 0.4, 0.5], [0, -1, 0]], F)
    # shapes -> {material: [(v, vt, vn) per corner]}; the quad is fanned from its first corner;
    # the box's groups come out in bucket order (red = 1 before blue = 3)
    shapes = [
        [(0, [(0, 0, 0), (3, 3, 0), (2, 2, 0), (0, 0, 0), (2, 2, 0), (1, 1, 0)])],
        [(1, [(7, None, 3), (6, None, 3), (4, None, 3)]),
         (3, [(4, None, 1), (6, None, 1), (5, None, 1), (5, None, 2), (6, None, 2), (7, None, 2)])],
        [(2, [(8, None, 4), (9, None, 4), (10, None, 4)])],
    ]
    sc = scenes.SceneData()
    t_floor = sc.add_texture(scenes.image_texture(floor))
    t_norm = sc.add_texture(scenes.image_texture(floor_n))
    kd = {"floor": (0.5, 0.5, 0.5), "red": (0.8, 0.2, 0.2), "lamp": (0, 0, 0), "blue": (0.2, 0.3, 0.8)}
    ke = {"floor": (0, 0, 0), "red": (0, 0, 0), "lamp": (12, 12, 10), "blue": (0, 0, 0)}
    mats = []
    for name in NAMES:
        m = scenes.metallic_roughness(kd[name], 0.0, 1.0)
        if name == "floor":
            m.tex_color, m.tex_normal = t_floor, t_norm
        mats.append(sc.add_material(m))
    for groups in shapes:
        for mid, corners in groups:
            pos = np.array([v[c[0]] for c in corners], F)
            pos[:, 1] = pos[:, 1] + F(ty)  # Mul(Mat4(Transform(0, ty, 0)), (p, 1)): y + ty
            nrm = np.array([vn[c[2]] for c in corners], F)
            tan = np.array([_coordinate_system(n) for n in nrm], F)
            uv = np.array([vt[c[1]] if c[1] is not None else (0, 0) for c in corners], F)
            prims = sc.add_mesh(pos, nrm, tan, np.arange(len(corners)).reshape(-1, 3), mats[mid], uvs=uv)
            for p in prims:
                sc.add_area_light(p, ke[NAMES[mid]])
    sc.add_infinite_light((0.2, 0.2, 0.2))
    cam = scenes.make_camera((0.0, 2.0, 6.0), (0.0, 0.5, 0.0), 40.0, 0.0,
                             scenes.dist32((0.0, 2.0, 6.0), (0.0, 0.5, 0.0)), width, height)
    return sc, cam


@pytest.mark.gpu
@pytest.mark.parametrize("texture", ["ppm", "jpg"])
def test_obj_model_equals_hand_built_scene(tmp_path, texture):
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    rng = np.random.default_rng(7)
    floor8 = rng.integers(20, 240, (4, 4, 3)).astype(np.uint8)
    norm8 = np.full((2, 2, 3), 128, np.uint8)
    norm8[0, 0] = (180, 128, 230)
    norm8[1, 1] = (90, 150, 220)
    _write_pnm(tmp_path / "floor_n.ppm", norm8)
    if texture == "ppm":
        _write_pnm(tmp_path / "floor.ppm", floor8)
        floor = _decode8(floor8, False)[::-1].copy()
    else:
        # a 4:2:0 JPEG written from chosen coefficients; the hand-built scene
        # takes the texture as ReadImage3 decodes it
        from test_jpeg import _random_coefs, write_jpeg
        coefs = [_random_coefs(2, 2), _random_coefs(1, 1, 0.5), _random_coefs(1, 1, 0.5)]
        (tmp_path / "floor.jpg").write_bytes(
            write_jpeg(13, 11, [(1, 2, 2, 0), (2, 1, 1, 0), (3, 1, 1, 0)], coefs, {0: np.full(64, 2)}))
        floor = _convert(tmp_path, tmp_path / "floor.jpg")
        assert floor.shape == (11, 13, 3)
    (tmp_path / "scene.obj").write_text(OBJ)
    (tmp_path / "scene.mtl").write_text(MTL.replace("floor.ppm", f"floor.{texture}"))
    out = tmp_path / "obj.pfm"
    print("[obj] load_obj 160 90 8", flush=True)
    r = subprocess.run([_exe(build.CPP_LOAD_OBJ), str(tmp_path / "scene.obj"), "160", "90", "8", str(out), "0.25"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    img = read_pfm(out)
    sc, cam = _expected_scene(tmp_path, floor, _decode8(norm8, True)[::-1].copy(), 0.25, 160, 90)
    s, c, _ = GpuScene(sc).render(cam, scenes.render_params(8))
    ref = s / c[..., None]
    assert np.array_equal(img.view(np.int32), ref.astype(np.float32).view(np.int32))
    os_, oc, _ = OracleScene(sc).render(cam, scenes.render_params(8), threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    frac = pixel_agreement(img, os_ / oc[..., None])
    print(f"[obj] OBJ scene vs oracle: {frac:.6f} within 1e-3, mean {img.mean():.4f}", flush=True)
    assert frac >= 0.999 and img.mean() > 0.01
