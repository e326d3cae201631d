"""Parity of the HIP path with the oracle, through the C ABI, on an MI355X.

* closest-hit primitive ids, hit distances and barycentrics: bit-exact on
  seeded random rays, camera rays and secondary rays; traversal counters equal;
* renders with identical RNG streams: >= 99.9% of pixels within 1e-3 relative
  (the residue comes from glibc sinf/cosf vs correctly rounded sin/cos);
  sample counts identical;
* at the benchmark's full size: determinism, sample-range splitting and
  wavefront-size invariance (properties that need no oracle run).
"""
import os

import numpy as np
import pytest

from bulbit_amd import capi, scenes

from conftest import pixel_agreement

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def log(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def readme():
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    sc, cam = scenes.readme_scene()
    return sc, cam, GpuScene(sc), OracleScene(sc)


@pytest.fixture(scope="module")
def ico6():
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    sc, cam = scenes.icosphere_scene(width=320, height=180, level=6)
    return sc, cam, GpuScene(sc), OracleScene(sc)


@pytest.fixture(scope="module")
def c2():
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    sc, cam = scenes.icosphere_scene()
    return sc, cam, GpuScene(sc), OracleScene(sc)


def _check_trace(g, o, rays, any_hit=False):
    """Both kernel variants against the oracle: the counting one (hits and
    RayCast's counters) and the one the renderer launches (hits)."""
    gh, gc = g.trace(rays, any_hit=any_hit, count=True)
    wh, _ = g.trace(rays, any_hit=any_hit, count=False)
    oh, oc = o.trace(rays, any_hit=any_hit, threads=THREADS)
    for h in (gh, wh):
        assert np.array_equal(h["prim"], oh["prim"])
        if not any_hit:
            for f in ("t", "u", "v"):
                assert np.array_equal(h[f].view(np.int32), oh[f].view(np.int32)), f
    assert gc.as_dict() == oc.as_dict()
    return gh


def test_trace_readme_sphere(readme):
    sc, cam, g, o = readme
    bmin, bmax = o.world_bounds()
    rays = scenes.random_rays(bmin, bmax, 200_000, seed=1)
    h = _check_trace(g, o, rays)
    assert 0.2 < np.mean(h["prim"] >= 0) < 0.8
    _check_trace(g, o, rays, any_hit=True)


def test_trace_icosphere_random_and_bounded(ico6):
    sc, cam, g, o = ico6
    bmin, bmax = o.world_bounds()
    rays = scenes.random_rays(bmin, bmax, 500_000, seed=2)
    _check_trace(g, o, rays)
    rng = np.random.default_rng(3)
    bounded = scenes.random_rays(bmin, bmax, 500_000, seed=4, t_max=rng.uniform(0, 10, 500_000))
    _check_trace(g, o, bounded)
    _check_trace(g, o, bounded, any_hit=True)


def test_trace_full_scene_rays(c2):
    """1.31M-primitive scene: 2M random rays plus camera and secondary rays."""
    sc, cam, g, o = c2
    bmin, bmax = o.world_bounds()
    rays = scenes.random_rays(bmin, bmax, 2_000_000, seed=5)
    h = _check_trace(g, o, rays)
    log(f"random rays hit fraction {np.mean(h['prim'] >= 0):.3f}")
    # camera rays through a grid of pixels and secondary rays from their hits
    from oracle import load_oracle
    import ctypes
    from bulbit_amd import capi

    lib = load_oracle()
    ys, xs = np.mgrid[0:cam.height:8, 0:cam.width:8]
    cr = np.zeros(xs.size, capi.RAY_DTYPE)
    ray = capi.BbRay()
    u = (ctypes.c_float * 2)(0.5, 0.5)
    for i, (x, y) in enumerate(zip(xs.ravel(), ys.ravel())):
        lib.orc_camera_ray(ctypes.byref(cam), int(x), int(y), u, u, ctypes.byref(ray))
        cr[i] = (tuple(ray.o), ray.t_min, tuple(ray.d), ray.t_max)
    h = _check_trace(g, o, cr)
    hit = h["prim"] >= 0
    sec = np.zeros(int(hit.sum()), capi.RAY_DTYPE)
    sec["o"] = cr["o"][hit] + cr["d"][hit] * h["t"][hit, None]
    d = np.random.default_rng(6).normal(size=(len(sec), 3)).astype(np.float32)
    sec["d"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    sec["t_min"] = np.float32(1e-4)
    sec["t_max"] = np.float32(np.inf)
    _check_trace(g, o, sec)
    _check_trace(g, o, sec, any_hit=True)


def test_trace_axis_aligned_rays_on_box_planes(ico6):
    """Rays with zero direction components whose origins lie exactly on BVH
    box planes: the slab test then multiplies 0 by an infinite inverse and
    gets NaN on that axis, which both forms of the test must ignore the same
    way (bounding_box.h:443-484)."""
    sc, cam, g, o = ico6
    nodes, _ = o.bvh()
    rng = np.random.default_rng(11)
    n = 200_000
    pick = nodes[rng.integers(0, len(nodes), n)]
    corner = np.where(rng.random((n, 3)) < 0.5, pick["bmin"], pick["bmax"]).astype(np.float32)
    lo, hi = o.world_bounds()
    free = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    keep = rng.random((n, 3)) < 0.5  # per axis: on the plane, or anywhere
    origin = np.where(keep, corner, free)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    zero = rng.integers(0, 3, n)
    d[np.arange(n), zero] = 0.0  # one zero component per ray
    two = rng.random(n) < 0.3    # and a second one for 30% of the rays
    d[np.arange(n)[two], (zero[two] + 1) % 3] = 0.0
    d = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros(n, capi.RAY_DTYPE)
    rays["o"] = origin
    rays["d"] = d.astype(np.float32)
    rays["t_min"] = np.float32(1e-4)
    rays["t_max"] = np.float32(np.inf)
    h = _check_trace(g, o, rays)
    log(f"plane rays hit fraction {np.mean(h['prim'] >= 0):.3f}")
    _check_trace(g, o, rays, any_hit=True)


def test_trace_accepts_host_buffers(ico6):
    """bbgpu_trace with rays and hits in host memory (staged by the library)
    equals the device-buffer call."""
    import ctypes

    sc, cam, g, o = ico6
    bmin, bmax = o.world_bounds()
    rays = np.ascontiguousarray(scenes.random_rays(bmin, bmax, 100_000, seed=12))
    dev_hits, _ = g.trace(rays)
    host_hits = np.zeros(len(rays), capi.HIT_DTYPE)
    capi.check(g.lib.bbgpu_trace(g.h, rays.ctypes.data, len(rays), 0, host_hits.ctypes.data, None, None),
               "bbgpu_trace", g.lib)
    assert np.array_equal(host_hits.view(np.int32), dev_hits.view(np.int32))


def test_render_accepts_host_film(readme):
    """bbgpu_render into a film in host memory (copied in, accumulated, copied
    back) equals the device-film render, and accumulates onto what is there."""
    import ctypes

    sc, cam, g, o = readme
    p = scenes.render_params(4)
    ds, dc, _ = g.render(cam, p)
    npix = cam.width * cam.height
    hs = np.zeros(3 * npix, np.float32)
    hc = np.zeros(npix, np.int32)
    for begin in (0, 2):
        q = scenes.render_params(2, sample_begin=begin)
        capi.check(g.lib.bbgpu_render(g.h, ctypes.byref(cam), ctypes.byref(q), hs.ctypes.data, hc.ctypes.data, None,
                                      None, 0), "bbgpu_render", g.lib)
    assert np.array_equal(hc.reshape(dc.shape), dc)
    assert np.array_equal(hs.reshape(ds.shape).view(np.int32), ds.view(np.int32))


def _render_pair(g, o, cam, params):
    gs, gc, _ = g.render(cam, params)
    os_, oc, _ = o.render(cam, params, threads=THREADS)
    return gs / gc[..., None], os_ / oc[..., None], gc, oc


def test_render_readme_scene_parity(readme):
    sc, cam, g, o = readme
    gi, oi, gc, oc = _render_pair(g, o, cam, scenes.render_params(16))
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"readme 640x360x16: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999


def test_render_mirror_parity():
    """MirrorMaterial on the README scene's spheres:
    within 1e-3 of the oracle, bit-equal with its libm correctly rounded."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene, correctly_rounded_libm

    sc, cam = scenes.readme_scene(320, 180)
    mid = sc.add_material(scenes.mirror((0.9, 0.8, 0.7)))
    sc.prims[:, 2] = np.where(sc.prims[:, 0] == capi.BB_SHAPE_SPHERE, mid, sc.prims[:, 2])
    g, o = GpuScene(sc), OracleScene(sc)
    p = scenes.render_params(16)
    gi, oi, gc, oc = _render_pair(g, o, cam, p)
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"mirror 320x180x16: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999
    gs, _, _ = g.render(cam, p)
    try:
        correctly_rounded_libm(True)
        cs, _, _ = o.render(cam, p, threads=THREADS)
    finally:
        correctly_rounded_libm(False)
    assert np.array_equal(cs.view(np.int32), gs.view(np.int32))


def test_render_point_lights_parity():
    """PointLight (delta, Li = I / d^2) and DirectionalLight beside the README scene's lights:
    within 1e-3 of the oracle, bit-equal with its libm correctly rounded."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene, correctly_rounded_libm

    sc, cam = scenes.readme_scene(320, 180)
    sc.add_point_light((1.5, 2.0, 1.0), (3.0, 2.5, 2.0))
    sc.add_point_light((-2.0, 0.6, 0.5), (0.5, 1.0, 2.0))
    sc.add_directional_light((0.3, 1.0, 0.4), (0.6, 0.5, 0.4))
    g, o = GpuScene(sc), OracleScene(sc)
    p = scenes.render_params(16)
    gi, oi, gc, oc = _render_pair(g, o, cam, p)
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"point lights 320x180x16: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999
    gs, _, _ = g.render(cam, p)
    try:
        correctly_rounded_libm(True)
        cs, _, _ = o.render(cam, p, threads=THREADS)
    finally:
        correctly_rounded_libm(False)
    assert np.array_equal(cs.view(np.int32), gs.view(np.int32))


@pytest.mark.parametrize("kind", ["tent", "gaussian"])
def test_render_with_tent_and_gaussian_filters(readme, kind):
    """TentFilter and GaussianFilter (the reference's default camera filter):
    the camera's filter sample moves every primary ray."""
    sc, cam, g, o = readme
    import copy

    cam2 = copy.copy(cam)
    scenes.set_filter(cam2, kind)
    gi, oi, gc, oc = _render_pair(g, o, cam2, scenes.render_params(8))
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"readme {kind} filter: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999
    # the filter changes the image (against the box filter)
    bs, bc, _ = g.render(cam, scenes.render_params(8))
    assert not np.array_equal(bs / bc[..., None], gi)


def test_render_icosphere_parity(ico6):
    sc, cam, g, o = ico6
    gi, oi, gc, oc = _render_pair(g, o, cam, scenes.render_params(8))
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"ico6 320x180x8: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999


def test_render_all_materials_parity():
    """Rough dielectric, smooth and rough conductor, metallic-roughness and
    Lambertian groups plus spot and environment lights."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    sc, cam = scenes.icosphere_scene(width=256, height=144, level=5, n_lights=16)
    rng = np.random.default_rng(8)
    mats = [scenes.dielectric(1.5, roughness=0.3), scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.0),
            scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.4), scenes.metallic_roughness((0.8, 0.3, 0.2), 0.0, 0.4),
            scenes.metallic_roughness((0.9, 0.8, 0.5), 1.0, 0.2), scenes.diffuse((0.7, 0.7, 0.7))]
    base = len(sc.materials)
    sc.materials += mats
    n_ico = 20 * 4 ** 5
    sc.prims[:n_ico, 2] = base + rng.integers(0, len(mats), n_ico // 64).repeat(64)
    sc.add_spot_light((2.0, 4.0, 2.0), scenes.normalize32(np.array([2.0, 3.0, 2.0], np.float32)), (40.0, 40.0, 40.0), 30.0,
                      20.0)
    sc.add_infinite_light((0.05, 0.06, 0.08))
    g, o = GpuScene(sc), OracleScene(sc)
    gi, oi, gc, oc = _render_pair(g, o, cam, scenes.render_params(8))
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"all-materials 256x144x8: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999


def test_render_compensated_conductors_parity():
    """ConductorMultiScatteringBxDF (energy_compensation, the reference's
    default for ConductorMaterial): rough gold and a rough R-parameterized
    conductor, against the oracle; the compensation lobe makes rough metal
    brighter than the plain conductor."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    def build(ec):
        sc, cam = scenes.icosphere_scene(width=192, height=108, level=5, n_lights=16)
        r_mat = scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.8, energy_compensation=ec)
        r_mat.type = capi.BB_MAT_CONDUCTOR_R
        r_mat.k[:] = [0.9, 0.6, 0.3]
        mats = [scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.5, energy_compensation=ec), r_mat,
                scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.0, energy_compensation=ec)]
        base = len(sc.materials)
        sc.materials += mats
        n_ico = 20 * 4 ** 5
        sc.prims[:n_ico, 2] = base + np.random.default_rng(9).integers(0, len(mats), n_ico // 64).repeat(64)
        return sc, cam

    sc, cam = build(1)
    g, o = GpuScene(sc), OracleScene(sc)
    gi, oi, gc, oc = _render_pair(g, o, cam, scenes.render_params(8))
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"compensated conductors 192x108x8: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999
    sc0, _ = build(0)
    s0, c0, _ = GpuScene(sc0).render(cam, scenes.render_params(8))
    assert gi.mean() > (s0 / c0[..., None]).mean()


def test_render_compensated_dielectrics_parity():
    """DielectricMultiScatteringBxDF (energy_compensation, the reference's
    default for DielectricMaterial) on rough glass and rough water, with the
    3-D albedo tables the library computes on the host, against the oracle;
    compensation adds the energy that the single-scattering interface loses."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    def build(ec):
        sc, cam = scenes.icosphere_scene(width=192, height=108, level=5, n_lights=16)
        mats = [scenes.dielectric(1.5, roughness=0.3, energy_compensation=ec),
                scenes.dielectric(1.5, roughness=0.7, energy_compensation=ec),
                scenes.dielectric(1.33, roughness=0.9, color=(0.9, 0.8, 0.6), energy_compensation=ec),
                scenes.dielectric(1.5, roughness=0.0, energy_compensation=ec)]
        base = len(sc.materials)
        sc.materials += mats
        n_ico = 20 * 4 ** 5
        sc.prims[:n_ico, 2] = base + np.random.default_rng(10).integers(0, len(mats), n_ico // 64).repeat(64)
        return sc, cam

    sc, cam = build(1)
    g, o = GpuScene(sc), OracleScene(sc)
    gi, oi, gc, oc = _render_pair(g, o, cam, scenes.render_params(8))
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"compensated dielectrics 192x108x8: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999
    sc0, _ = build(0)
    s0, c0, _ = GpuScene(sc0).render(cam, scenes.render_params(8))
    assert np.all(np.isfinite(gi)) and gi.mean() > (s0 / c0[..., None]).mean()


def test_large_scene_generator_parity():
    """The large-scene generator at a reduced size (256 level-3 spheres, 128
    log-normal lights): hit ids bit-exact on random and bounded rays (closest
    and any-hit), and a 12-bounce render of the metallic-roughness spheres
    within 1e-3 of the oracle."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    sc, cam = scenes.large_scene(width=200, height=112, n_spheres=256, level=3, n_lights=128)
    g, o = GpuScene(sc), OracleScene(sc)
    bmin, bmax = o.world_bounds()
    _check_trace(g, o, scenes.random_rays(bmin, bmax, 300_000, seed=31))
    rng = np.random.default_rng(32)
    bounded = scenes.random_rays(bmin, bmax, 300_000, seed=33, t_max=rng.uniform(0, 20, 300_000))
    _check_trace(g, o, bounded)
    _check_trace(g, o, bounded, any_hit=True)
    gi, oi, gc, oc = _render_pair(g, o, cam, scenes.render_params(8, max_bounces=12))
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"large-scene generator 200x112x8, 12 bounces: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999
    assert gi.mean() > 0.01  # the lights reach the camera


def test_render_eon_parity():
    """DiffuseMaterial with roughness > 0 (the EON BxDF) on the icosphere and
    the ground, beside Lambertian groups: within 1e-3 of the oracle, and
    bit-equal with the oracle's transcendentals correctly rounded."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene, correctly_rounded_libm

    sc, cam = scenes.icosphere_scene(width=240, height=136, level=5, n_lights=16)
    mats = []
    for r, col in ((0.2, (0.8, 0.5, 0.3)), (0.6, (0.3, 0.7, 0.4)), (1.0, (0.9, 0.9, 0.9)), (0.0, (0.5, 0.5, 0.8))):
        m = scenes.diffuse(col)
        m.u_roughness = m.v_roughness = r
        mats.append(m)
    base = len(sc.materials)
    sc.materials += mats
    n_ico = 20 * 4 ** 5
    sc.prims[:n_ico, 2] = base + (np.arange(n_ico) // 512) % len(mats)
    sc.prims[n_ico:n_ico + 2, 2] = base + 1  # rough ground
    sc.add_infinite_light((0.05, 0.05, 0.05))
    g, o = GpuScene(sc), OracleScene(sc)
    p = scenes.render_params(8)
    gs, gc, _ = g.render(cam, p)
    os_, oc, _ = o.render(cam, p, threads=THREADS)
    gi, oi = gs / gc[..., None], os_ / oc[..., None]
    assert np.array_equal(gc, oc)
    frac = pixel_agreement(gi, oi)
    log(f"EON 240x136x8: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999
    try:
        correctly_rounded_libm(True)
        cs, cc, _ = o.render(cam, p, threads=THREADS)
    finally:
        correctly_rounded_libm(False)
    assert np.array_equal(cs.view(np.int32), gs.view(np.int32))


@pytest.mark.parametrize("which", ["readme", "ico6", "large"])
def test_render_bit_exact_with_correctly_rounded_libm(readme, ico6, which):
    """The only difference between the GPU path and the oracle is libm: with
    the oracle's sampling transcendentals evaluated in double and rounded
    (as the GPU does), film sums and counts are bit-equal."""
    from oracle import correctly_rounded_libm

    if which == "large":
        from bulbit_amd.gpu import GpuScene
        from oracle import OracleScene

        sc, cam = scenes.large_scene(width=160, height=90, n_spheres=128, level=3, n_lights=64)
        g, o = GpuScene(sc), OracleScene(sc)
        p = scenes.render_params(4, max_bounces=12)
    else:
        sc, cam, g, o = readme if which == "readme" else ico6
        p = scenes.render_params(4)
    gs, gc, _ = g.render(cam, p)
    try:
        correctly_rounded_libm(True)
        os_, oc, _ = o.render(cam, p, threads=THREADS)
    finally:
        correctly_rounded_libm(False)
    assert np.array_equal(gc, oc)
    log(f"{which}: {np.mean(gs == os_):.6f} of film values bit-equal with correctly rounded libm")
    assert np.array_equal(gs.view(np.int32), os_.view(np.int32))


def test_render_full_scene_crop_parity(c2):
    """The benchmark frame itself: GPU image rows against the oracle on a
    strided set of rows at 2 spp."""
    sc, cam, g, o = c2
    params = scenes.render_params(2)
    gs, gc, _ = g.render(cam, params)
    rows = list(range(3, cam.height, 27))
    os_ = np.zeros_like(gs)
    oc = np.zeros_like(gc)
    for y in rows:
        o.render(cam, params, rect=(0, y, cam.width, y + 1), threads=THREADS, film=(os_, oc))
    gi = gs[rows] / gc[rows][..., None]
    oi = os_[rows] / oc[rows][..., None]
    assert np.array_equal(gc[rows], oc[rows])
    frac = pixel_agreement(gi, oi)
    log(f"C2 rows x 2 spp: {frac:.6f} within 1e-3, {np.mean(gi == oi):.4f} bit-equal")
    assert frac >= 0.999


def test_full_scene_properties(c2):
    """At 1920x1080: a 4-spp render is deterministic, equals two 2-spp renders
    accumulated with sample_begin 0 and 2, and does not depend on the
    wavefront size; every sample is finite here, so counts equal spp."""
    sc, cam, g, o = c2
    a, ca, _ = g.render(cam, scenes.render_params(4))
    b, cb, _ = g.render(cam, scenes.render_params(4))
    assert np.array_equal(a, b) and np.array_equal(ca, cb)
    assert np.all(ca == 4)
    film = g.new_film(cam)
    g.render(cam, scenes.render_params(2), film=film)
    s2, c2_, _ = g.render(cam, scenes.render_params(2, sample_begin=2), film=film)
    assert np.array_equal(s2, a) and np.array_equal(c2_, ca)
    c, cc, _ = g.render(cam, scenes.render_params(4, wavefront_paths=1 << 20))
    assert np.array_equal(c, a) and np.array_equal(cc, ca)
    assert np.isfinite(a).all() and a.min() >= 0
