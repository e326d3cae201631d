"""Textures on the path: the oracle's Texture<T>::Evaluate pinned by closed
forms (texel centres, bilinear reproduction of linear fields, checker
parity, the clamp and repeat filters), and GPU renders and traces of
textured scenes (image, checker and constant textures in every material
slot, normal maps, alpha tests) against the oracle."""
import os

import numpy as np
import pytest

from bulbit_amd import capi, scenes

from conftest import pixel_agreement

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def _one_triangle_scene(textures):
    sc = scenes.SceneData()
    m = sc.add_material(scenes.diffuse((0.5, 0.5, 0.5)))
    sc.add_mesh([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.tile([0, 0, 1.0], (3, 1)), np.tile([1.0, 0, 0], (3, 1)),
                [[0, 1, 2]], m, uvs=[[0, 0], [1, 0], [0, 1]])
    slots = [sc.add_texture(t) for t in textures]
    return sc, slots


# ------------------------------------------------------------------ oracle

def test_texture_constant_and_texel_centres():
    from oracle import OracleScene

    rng = np.random.default_rng(1)
    img = rng.random((8, 16, 3)).astype(np.float32)
    sc, slots = _one_triangle_scene([scenes.constant_texture((0.25, 0.5, 0.75)), scenes.image_texture(img),
                                     scenes.constant_texture(0.3)])
    o = OracleScene(sc)
    assert np.array_equal(o.texture(0, (0.37, 0.81)), np.float32([0.25, 0.5, 0.75]))
    assert o.texture(2, (0.1, 0.2))[0] == np.float32(0.3)
    # at a texel centre (power-of-two sizes: w = i exactly) the image returns the texel
    for (i, j) in [(0, 0), (15, 7), (3, 5), (9, 2)]:
        uv = ((i + 0.5) / 16, (j + 0.5) / 8)
        assert np.array_equal(o.texture(1, uv), img[j, i]), (i, j)


def test_bilinear_reproduces_linear_fields():
    """A linear texel field a*i + b*j + c is reproduced exactly (to rounding)
    between texel centres: the bilinear weights sum to one and are linear in
    the position."""
    from oracle import OracleScene

    W, H = 32, 16
    jj, ii = np.mgrid[0:H, 0:W]
    field = (0.02 * ii + 0.05 * jj + 0.1).astype(np.float32)
    sc, _ = _one_triangle_scene([scenes.image_texture(field, capi.BB_TEXCOORD_CLAMP)])
    o = OracleScene(sc)
    rng = np.random.default_rng(2)
    for _ in range(200):
        u, v = rng.uniform(0.5 / W, 1 - 0.5 / W), rng.uniform(0.5 / H, 1 - 0.5 / H)
        x, y = u * W - 0.5, v * H - 0.5
        assert abs(o.texture(0, (u, v))[0] - (0.02 * x + 0.05 * y + 0.1)) < 1e-5
    # outside the texel-centre hull the clamp filter holds the edge texels
    assert abs(o.texture(0, (0.0, 0.0))[0] - field[0, 0]) < 1e-6
    assert abs(o.texture(0, (1.0, 1.0))[0] - field[H - 1, W - 1]) < 1e-6


def test_repeat_filter_wraps_inside_the_image():
    """With repeat every looked-up texel index lands in the image, so values
    stay within the image's range for any uv, and a constant image stays
    constant; clamp and repeat agree on the interior."""
    from oracle import OracleScene

    rng = np.random.default_rng(3)
    img = rng.random((8, 8)).astype(np.float32)
    sc, _ = _one_triangle_scene([scenes.image_texture(img, capi.BB_TEXCOORD_REPEAT),
                                 scenes.image_texture(img, capi.BB_TEXCOORD_CLAMP),
                                 scenes.image_texture(np.full((8, 8), 0.7, np.float32))])
    o = OracleScene(sc)
    for _ in range(300):
        uv = rng.uniform(-3, 3, 2)
        val = o.texture(0, uv)[0]
        assert img.min() - 1e-6 <= val <= img.max() + 1e-6
        assert abs(o.texture(2, uv)[0] - 0.7) < 1e-6
    for _ in range(100):
        uv = rng.uniform(1 / 16, 15 / 16, 2)
        assert o.texture(0, uv)[0] == o.texture(1, uv)[0]


def test_checker_parity():
    from oracle import OracleScene

    sc, slots = _one_triangle_scene([scenes.constant_texture((1.0, 0.0, 0.0)), scenes.constant_texture((0.0, 0.0, 1.0))])
    sc.add_texture(scenes.checker_texture(slots[0], slots[1], (4.0, 2.0), 3))
    o = OracleScene(sc)
    rng = np.random.default_rng(4)
    for _ in range(200):
        u, v = rng.uniform(0, 1, 2)
        odd = (int(u * 4) + int(v * 2)) % 2 == 1
        want = [1.0, 0.0, 0.0] if odd else [0.0, 0.0, 1.0]
        assert np.array_equal(o.texture(2, (u, v)), np.float32(want))


def test_library_rejects_bad_texture_slots(bblib):
    import ctypes

    sc, slots = _one_triangle_scene([scenes.constant_texture(0.5)])
    sc.materials[0].tex_color = slots[0]  # a 1-channel texture in a colour slot
    h = ctypes.c_void_p()
    n = ctypes.c_int32()
    bblib.bbgpu_device_count(ctypes.byref(n))
    st = bblib.bbgpu_scene_create(ctypes.byref(sc.desc()), ctypes.byref(h))
    assert st != 0 and not h.value
    if n.value > 0:  # with a device the description is checked: a 1-channel texture in a colour slot
        assert st == 1 and "channel" in capi.last_error(bblib)


# --------------------------------------------------------------------- GPU

def textured_scene(width=240, height=136, level=5, alpha=True):
    """The displaced icosphere with spherical texture coordinates, its face
    groups cycling through textured materials: image albedo, checker
    roughness on a conductor, image roughness and metallic on
    metallic-roughness, image roughness on a compensated dielectric, a normal
    map, and (with alpha) cut-out and fractional alpha; a checker-albedo
    ground; area lights and a dim environment."""
    sc, cam = scenes.icosphere_scene(width=width, height=height, level=level, n_lights=16)
    p = sc.positions.astype(np.float64) - np.array([0.0, 1.05, 0.0])
    r = np.linalg.norm(p, axis=1)
    u = (np.arctan2(p[:, 2], p[:, 0]) / (2 * np.pi)) % 1.0
    v = np.arccos(np.clip(p[:, 1] / np.maximum(r, 1e-9), -1, 1)) / np.pi
    uvs = np.stack([u * 3.0, v * 2.0], 1)
    n_ground = 4
    uvs[-(n_ground + 3 * 16):] = 0.0
    gl = len(sc.positions) - 3 * 16 - n_ground  # the ground's 4 vertices follow the sphere
    uvs[gl:gl + 4] = [[0, 0], [8, 0], [8, 8], [0, 8]]
    sc.uvs = uvs.astype(np.float32)
    rng = np.random.default_rng(11)
    img = lambda h, w, c: rng.uniform(0.05, 0.95, (h, w, c) if c > 1 else (h, w)).astype(np.float32)
    t_alb = sc.add_texture(scenes.image_texture(img(32, 64, 3)))
    t_rough = sc.add_texture(scenes.image_texture(img(16, 16, 1)))
    t_metal = sc.add_texture(scenes.image_texture((rng.random((8, 8)) > 0.5).astype(np.float32)))
    c_lo, c_hi = sc.add_texture(scenes.constant_texture(0.05)), sc.add_texture(scenes.constant_texture(0.5))
    t_check_r = sc.add_texture(scenes.checker_texture(c_lo, c_hi, (8.0, 8.0), 1))
    nm = np.zeros((32, 32, 3), np.float32)
    yy, xx = np.mgrid[0:32, 0:32] / 32.0
    nm[..., 0] = 0.5 + 0.3 * np.sin(2 * np.pi * xx * 3)
    nm[..., 1] = 0.5 + 0.3 * np.cos(2 * np.pi * yy * 2)
    nm[..., 2] = 0.9
    t_normal = sc.add_texture(scenes.image_texture(nm))
    a_img = np.where(rng.random((16, 16)) < 0.4, 0.0, 1.0).astype(np.float32)
    a_img[:4] = 0.5  # fractional rows: the hash decides
    t_alpha = sc.add_texture(scenes.image_texture(a_img))
    w1, w2 = sc.add_texture(scenes.constant_texture((0.8, 0.8, 0.8))), sc.add_texture(scenes.constant_texture((0.2, 0.3, 0.6)))
    t_check_c = sc.add_texture(scenes.checker_texture(w1, w2, (4.0, 4.0), 3))

    mats = []
    m = scenes.diffuse((1, 1, 1)); m.tex_color = t_alb; mats.append(m)
    m = scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.3); m.tex_u_roughness = m.tex_v_roughness = t_check_r; mats.append(m)
    m = scenes.metallic_roughness((0.9, 0.6, 0.3), 0.0, 0.5); m.tex_color = t_alb; m.tex_u_roughness = t_rough
    m.tex_v_roughness = t_rough; m.tex_metallic = t_metal; mats.append(m)
    m = scenes.dielectric(1.5, roughness=0.3, energy_compensation=1); m.tex_u_roughness = m.tex_v_roughness = t_rough; mats.append(m)
    m = scenes.diffuse((0.7, 0.7, 0.7)); m.tex_normal = t_normal; mats.append(m)
    m = scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.2, energy_compensation=1); m.tex_normal = t_normal; mats.append(m)
    m = scenes.diffuse((0.8, 0.7, 0.6)); m.tex_u_roughness = t_check_r; mats.append(m)  # EON where the roughness > 0
    if alpha:
        m = scenes.diffuse((0.6, 0.9, 0.6)); m.tex_alpha = t_alpha; mats.append(m)
        m = scenes.metallic_roughness((0.5, 0.5, 0.9), 1.0, 0.3); m.tex_alpha = t_alpha; m.tex_normal = t_normal; mats.append(m)
    base = len(sc.materials)
    sc.materials += mats
    n_ico = 20 * 4 ** level
    sc.prims[:n_ico, 2] = base + (np.arange(n_ico) // 256) % len(mats)
    g = scenes.diffuse((1, 1, 1))
    g.tex_color = t_check_c
    sc.materials.append(g)
    sc.prims[n_ico:n_ico + 2, 2] = len(sc.materials) - 1
    sc.add_infinite_light((0.05, 0.06, 0.08))
    return sc, cam


@pytest.mark.gpu
def test_textured_scene_trace_parity():
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    sc, cam = textured_scene()
    g, o = GpuScene(sc), OracleScene(sc)
    bmin, bmax = o.world_bounds()
    rays = scenes.random_rays(bmin, bmax, 300_000, seed=41)
    for any_hit in (False, True):
        gh, gc = g.trace(rays, any_hit=any_hit, count=True)
        oh, oc = o.trace(rays, any_hit=any_hit, threads=THREADS)
        same = gh["prim"] == oh["prim"]
        print(f"[tex] alpha trace any_hit={any_hit}: {np.mean(same):.7f} equal ids, "
              f"{np.sum(oh['prim'] >= 0)} hits", flush=True)
        assert np.array_equal(gh["prim"], oh["prim"])
        if not any_hit:
            for f in ("t", "u", "v"):
                assert np.array_equal(gh[f].view(np.int32), oh[f].view(np.int32)), f
        assert gc.as_dict() == oc.as_dict()


@pytest.mark.gpu
@pytest.mark.parametrize("alpha", [False, True], ids=["textures", "textures+alpha"])
def test_textured_scene_render_parity(alpha):
    """Against the oracle with glibc's float transcendentals (the
    reference's) the usual bar holds without alpha.  With alpha, HashFloat
    of the exact ray bits decides every fractional-alpha crossing, so the
    last-bit differences between glibc's sinf/cosf and correctly rounded
    ones (on the GPU) in sampled directions flip some of those decisions:
    the bar there is 99.5 %.  Against the oracle with correctly rounded
    transcendentals every pixel is bit-equal, alpha or not, which shows that
    nothing else differs."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene, correctly_rounded_libm

    sc, cam = textured_scene(alpha=alpha)
    g, o = GpuScene(sc), OracleScene(sc)
    p = scenes.render_params(8)
    gs, gc, _ = g.render(cam, p)
    os_, oc, _ = o.render(cam, p, threads=THREADS)
    gi, oi = gs / gc[..., None], os_ / oc[..., None]
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    print(f"[tex] textured render (alpha={alpha}) 240x136x8: {frac:.6f} within 1e-3, "
          f"{np.mean(gi == oi):.4f} bit-equal", flush=True)
    assert frac >= (0.995 if alpha else 0.999)
    try:
        correctly_rounded_libm(True)
        cs, cc, _ = o.render(cam, p, threads=THREADS)
    finally:
        correctly_rounded_libm(False)
    assert np.array_equal(cc, gc)
    assert np.array_equal(cs.view(np.int32), gs.view(np.int32)), "not bit-equal with correctly rounded libm"
    # the textures change the image against the same scene without them
    sc2, _ = textured_scene(alpha=alpha)
    for m in sc2.materials:
        m.tex_color = m.tex_u_roughness = m.tex_v_roughness = m.tex_metallic = m.tex_normal = m.tex_alpha = 0
    s2, c2, _ = GpuScene(sc2).render(cam, p)
    assert np.mean(np.abs(s2 / c2[..., None] - gi)) > 1e-3


@pytest.mark.gpu
def test_textured_sphere_parity():
    """Sphere texture coordinates (ComputeTexCoord of the object-space
    normal): the README sphere with an image albedo and a normal map."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    sc, cam = scenes.readme_scene(320, 180)
    rng = np.random.default_rng(5)
    t = sc.add_texture(scenes.image_texture(rng.uniform(0.1, 0.9, (16, 32, 3)).astype(np.float32)))
    nm = np.full((8, 8, 3), 0.5, np.float32)
    nm[..., 2] = 1.0
    nm[::2, :, 0] = 0.7
    tn = sc.add_texture(scenes.image_texture(nm))
    sc.materials[0].tex_color = t
    sc.materials[0].tex_normal = tn
    g, o = GpuScene(sc), OracleScene(sc)
    p = scenes.render_params(8)
    gs, gc, _ = g.render(cam, p)
    os_, oc, _ = o.render(cam, p, threads=THREADS)
    gi, oi = gs / gc[..., None], os_ / oc[..., None]
    frac = pixel_agreement(gi, oi)
    print(f"[tex] textured sphere 320x180x8: {frac:.6f} within 1e-3", flush=True)
    assert np.array_equal(gc, oc) and frac >= 0.999
