"""Pins of the oracle (the CPU restatement of Bulbit's path tracer) against
published vectors, closed forms, invariants and brute force.  CPU only.

The reference has no tests or golden data and cannot be compiled here (it
needs <format>, absent from this image's libstdc++, and NanoVDB headers), so
each function of the oracle is pinned by an independent fact instead."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from bulbit_amd import capi, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")
F = ctypes.c_float


def farr(*v):
    return (F * len(v))(*v)


# ------------------------------------------------------------ RNG / hashing

def test_pcg32_published_demo_vector(orc):
    # pcg32-demo of the PCG reference implementation: pcg32_srandom_r(42, 54)
    out = (ctypes.c_uint32 * 6)()
    orc.orc_pcg32_outputs(42, 54, 6, out)
    assert list(out) == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_murmurhash64a_smhasher_verification(orc):
    # SMHasher's verification value of MurmurHash64A ("Murmur2B"): hash keys
    # {0..i-1} with seed 256 - i, hash the concatenated results with seed 0.
    key = (ctypes.c_uint8 * 256)(*range(256))
    hashes = np.zeros(256, np.uint64)
    for i in range(256):
        hashes[i] = orc.orc_murmur64a(key, i, 256 - i)
    buf = hashes.tobytes()
    final = orc.orc_murmur64a(buf, len(buf), 0)
    assert final & 0xFFFFFFFF == 0x1F0D3804


def test_pixel_hash_packs_sixteen_bytes(orc):
    # Hash(pixel, seed, sample) = MurmurHash64A over the 16 packed bytes
    b = np.array([3, 7, 11, 5], np.int32).tobytes()
    assert orc.orc_pixel_sample_hash(3, 7, 11, 5) == orc.orc_murmur64a(b, 16, 0)


def test_sampler_floats_are_pcg_outputs(orc):
    # NextFloat = min(1 - eps, float(u32) * 2^-32), stream seeded (hash, initseq 1)
    h = orc.orc_pixel_sample_hash(12, 34, 0, 9)
    raw = (ctypes.c_uint32 * 8)()
    orc.orc_pcg32_outputs(h, 1, 8, raw)
    fl = (F * 8)()
    orc.orc_sampler_floats(12, 34, 0, 9, 8, fl)
    expect = np.minimum(np.float32(1) - np.finfo(np.float32).eps,
                        np.array(raw, np.uint32).astype(np.float32) * np.float32(2.0 ** -32))
    assert np.array_equal(np.array(fl, np.float32), expect.astype(np.float32))


# ------------------------------------------------------ compiler semantics

@pytest.fixture(scope="module")
def probes():
    subprocess.run(["make", "-C", ORACLE_DIR, "probes"], check=True, capture_output=True)
    return os.path.join(ORACLE_DIR, "_build")


def test_gxx_evaluates_sampler_arguments_right_to_left(probes):
    out = subprocess.run([os.path.join(probes, "probe_argorder")], capture_output=True, text=True, check=True).stdout
    # the first draw (1) lands in the last argument: u1 (lens) and u12
    assert "camera u0=2 u1=1" in out
    assert "bsdf u0=2 u12=1" in out


def test_gxx_complex_float_semantics(probes):
    out = subprocess.run([os.path.join(probes, "probe_complex")], capture_output=True, text=True, check=True).stdout
    vals = dict(line.split() for line in out.strip().splitlines())
    n = int(vals["n"])
    # the oracle's FresnelComplex uses exactly these formulas
    for k in ("div_double", "mul_float", "mul_real", "sub_real", "abs_hypot", "norm"):
        assert int(vals[k]) == n, k
    assert int(vals["div_smith"]) < n  # Smith's algorithm is not what g++ emits here


# ------------------------------------------------------------------ camera

def _pinhole(width, height, vfov=90.0, aperture=0.0, focus=1.0):
    cam = capi.BbCameraDesc()
    cam.position[:] = [0.0, 0.0, 0.0]
    q = farr(0, 0, 0, 0)
    orc_lib = __import__("oracle").load_oracle()
    orc_lib.orc_look_at(farr(0, 0, 0), farr(0, 0, -1), farr(0, 1, 0), q)
    cam.orientation[:] = list(q)
    cam.vfov_deg = vfov
    cam.aperture_radius = aperture
    cam.focus_distance = focus
    cam.width = width
    cam.height = height
    cam.filter_extent = 1.0
    return cam


def test_look_at_down_negative_z_is_half_turn_about_y(orc):
    q = farr(0, 0, 0, 0)
    orc.orc_look_at(farr(0, 0, 0), farr(0, 0, -1), farr(0, 1, 0), q)
    assert list(q) == [0.0, 1.0, 0.0, 0.0]


def test_camera_centre_ray_closed_form(orc):
    cam = _pinhole(3, 3)
    ray = capi.BbRay()
    orc.orc_camera_ray(ctypes.byref(cam), 1, 1, farr(0.5, 0.5), farr(0.5, 0.5), ctypes.byref(ray))
    assert list(ray.o) == [0.0, 0.0, 0.0]
    assert list(ray.d) == [0.0, 0.0, -1.0]
    orc.orc_camera_ray(ctypes.byref(cam), 0, 0, farr(0.5, 0.5), farr(0.5, 0.5), ctypes.byref(ray))
    expect = np.array([-2.0, -2.0, -3.0]) / np.sqrt(17.0)
    assert np.allclose(list(ray.d), expect, rtol=0, atol=1e-6)


def test_camera_filter_and_lens_offsets(orc):
    cam = _pinhole(3, 3, aperture=0.25)
    ray = capi.BbRay()
    # u1 = (1, 0.5): concentric disk -> (1, 0): origin moves by the lens radius along u
    orc.orc_camera_ray(ctypes.byref(cam), 1, 1, farr(0.5, 0.5), farr(1.0, 0.5), ctypes.byref(ray))
    assert np.allclose(list(ray.o), [0.25, 0.0, 0.0], atol=1e-7)
    # u0 = (1, 0.5): box filter offset +0.5 px in x -> pixel point x = -1 + 2 * 2 / 3
    orc.orc_camera_ray(ctypes.byref(cam), 1, 1, farr(1.0, 0.5), farr(0.5, 0.5), ctypes.byref(ray))
    p = np.array([-1 + 2 * 2.0 / 3.0, 0.0, -1.0])
    assert np.allclose(list(ray.d), p / np.linalg.norm(p), atol=1e-6)


def _filter_sample(orc, cam, u):
    out = farr(0, 0)
    orc.orc_filter_sample(ctypes.byref(cam), farr(*u), out)
    return np.array(out[:], np.float64)


def test_box_and_tent_filter_closed_forms(orc):
    cam = scenes.set_filter(_pinhole(3, 3), "box", extent=1.5)
    assert np.array_equal(_filter_sample(orc, cam, (0.0, 1.0)), [-0.75, 0.75])
    assert np.array_equal(_filter_sample(orc, cam, (0.5, 0.25)), [0.0, -0.375])
    cam = scenes.set_filter(_pinhole(3, 3), "tent")  # extent 2: half extent 1
    # x = h (sqrt(2u) - 1) below 1/2, h (1 - sqrt(1 - 2 (u - 1/2))) above
    assert np.array_equal(_filter_sample(orc, cam, (0.0, 0.5)), [-1.0, 0.0])
    assert np.array_equal(_filter_sample(orc, cam, (0.125, 0.875)), [-0.5, 0.5])
    # uniform u gives the triangle density (1 - |x|) on [-1, 1]
    u = (np.arange(20000) + 0.5) / 20000
    xs = np.array([_filter_sample(orc, cam, (v, 0.5))[0] for v in u])
    hist, edges = np.histogram(xs, bins=10, range=(-1, 1))
    c = 0.5 * (edges[:-1] + edges[1:])
    assert np.allclose(hist / len(xs), (1 - np.abs(c)) * 0.2, atol=2e-3)


def test_gaussian_filter_samples_follow_its_table(orc):
    """GaussianFilter (gaussian_filter.cpp:8-43): a 32 x 32 piecewise-constant
    Distribution2D of the separable Gaussian over [-e/2, e/2]^2.  Stratified
    u must put into each cell the probability of the truncated Gaussian's
    midpoint rule, the sampling is odd-symmetric about u = 1/2, and linear
    inside a cell."""
    sigma, e = 0.5, 3.0
    cam = scenes.set_filter(_pinhole(3, 3), "gaussian", sigma=sigma)
    n = 8192
    u = (np.arange(n) + 0.5) / n
    xs = np.array([_filter_sample(orc, cam, (v, 0.3))[0] for v in u])
    ys = np.array([_filter_sample(orc, cam, (0.3, v))[1] for v in u])
    centres = ((np.arange(32) + 0.5) / 32) * e - e / 2
    g = np.exp(-centres ** 2 / (2 * sigma ** 2))
    want = g / g.sum()
    for s_ in (xs, ys):
        assert s_.min() >= -e / 2 and s_.max() <= e / 2
        cell = np.clip(((s_ + e / 2) / e * 32).astype(int), 0, 31)
        got = np.bincount(cell, minlength=32) / n
        assert np.allclose(got, want, atol=2.0 / n)
    # odd symmetry and linearity within the central cells
    a = _filter_sample(orc, cam, (0.3, 0.3))
    b = _filter_sample(orc, cam, (0.7, 0.7))
    assert np.allclose(a, -b, atol=1e-6)
    mid = [_filter_sample(orc, cam, (v, 0.5))[0] for v in (0.49, 0.495, 0.5)]
    assert abs((mid[1] - mid[0]) - (mid[2] - mid[1])) < 1e-6
    # the sample variance is that of the truncated Gaussian (~sigma^2 less the tails)
    assert 0.9 * sigma ** 2 < np.mean(xs ** 2) < sigma ** 2


def test_concentric_disk_and_sphere_samples(orc):
    out = farr(0, 0)
    orc.orc_sample_concentric(farr(0.5, 0.5), out)
    assert list(out) == [0.0, 0.0]
    orc.orc_sample_concentric(farr(0.5, 1.0), out)
    assert np.allclose(list(out), [0.0, 1.0], atol=1e-7)
    rng = np.random.default_rng(0)
    s = farr(0, 0, 0)
    for u in rng.random((200, 2)):
        orc.orc_sample_uniform_sphere(farr(*u), s)
        assert abs(np.linalg.norm(list(s)) - 1.0) < 1e-6
        assert np.isclose(s[2], 1.0 - 2.0 * np.float32(u[0]), atol=1e-6)


# ------------------------------------------------------------- triangles

def _tri_hit(orc, o, d, t_min=1e-4, t_max=np.inf):
    ray = capi.BbRay()
    ray.o[:] = list(o)
    ray.d[:] = list(d)
    ray.t_min = t_min
    ray.t_max = t_max
    t, u, v = F(), F(), F()
    hit = orc.orc_triangle_intersect(farr(0, 0, 0), farr(1, 0, 0), farr(0, 1, 0), ctypes.byref(ray), ctypes.byref(t),
                                     ctypes.byref(u), ctypes.byref(v))
    return hit, t.value, u.value, v.value


def test_moller_trumbore_exact_cases(orc):
    assert _tri_hit(orc, (0.25, 0.25, 1.0), (0, 0, -1)) == (1, 1.0, 0.25, 0.25)
    # t is measured in units of the (unnormalised) direction
    assert _tri_hit(orc, (0.25, 0.25, 1.0), (0, 0, -2)) == (1, 0.5, 0.25, 0.25)
    # back faces are hit too (no culling)
    assert _tri_hit(orc, (0.25, 0.5, -1.0), (0, 0, 1)) == (1, 1.0, 0.25, 0.5)
    # u + v == 1 on the hypotenuse is inside, beyond it is not
    assert _tri_hit(orc, (0.5, 0.5, 1.0), (0, 0, -1))[0] == 1
    assert _tri_hit(orc, (0.5, 0.625, 1.0), (0, 0, -1))[0] == 0
    assert _tri_hit(orc, (-0.125, 0.25, 1.0), (0, 0, -1))[0] == 0
    # parallel ray
    assert _tri_hit(orc, (0.25, 0.25, 0.0), (1, 0, 0))[0] == 0
    # t range, inclusive bounds
    assert _tri_hit(orc, (0.25, 0.25, 1.0), (0, 0, -1), t_max=0.5)[0] == 0
    assert _tri_hit(orc, (0.25, 0.25, 1.0), (0, 0, -1), t_max=1.0)[0] == 1
    assert _tri_hit(orc, (0.25, 0.25, 1.0), (0, 0, -1), t_min=1.0)[0] == 1


# ------------------------------------------------------------------ BVH

def _soup(n, seed, spread=4.0, size=0.3):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 3))
    p = c[:, None, :] + rng.uniform(-size, size, (n, 3, 3))
    sc = scenes.SceneData()
    m = sc.add_material(scenes.diffuse((0.5, 0.5, 0.5)))
    pos = p.reshape(-1, 3)
    nrm = np.tile([0.0, 0.0, 1.0], (len(pos), 1))
    tan = np.tile([1.0, 0.0, 0.0], (len(pos), 1))
    sc.add_mesh(pos, nrm, tan, np.arange(3 * n).reshape(-1, 3), m)
    return sc


def _check_bvh_invariants(sc, nodes, order):
    n = len(sc.prims)
    assert sorted(order.tolist()) == list(range(n))
    covered = np.zeros(n, np.int32)
    pos = sc.positions
    tri = sc.indices
    eps = np.float32(np.finfo(np.float32).eps * 10)
    for i, nd in enumerate(nodes):
        if nd["count"] > 0:
            sl = order[nd["offset"]: nd["offset"] + nd["count"]]
            covered[nd["offset"]: nd["offset"] + nd["count"]] += 1
            prims = sc.prims[sl]
            if np.all(prims[:, 0] == capi.BB_SHAPE_TRIANGLE):
                v = pos[tri[prims[:, 1]]]
                assert np.array_equal(nd["bmin"], (v.min(axis=(0, 1)) - eps).astype(np.float32))
                assert np.array_equal(nd["bmax"], (v.max(axis=(0, 1)) + eps).astype(np.float32))
        else:
            c1, c2 = nodes[i + 1], nodes[nd["offset"]]
            assert nd["offset"] > i + 1
            assert np.array_equal(nd["bmin"], np.minimum(c1["bmin"], c2["bmin"]))
            assert np.array_equal(nd["bmax"], np.maximum(c1["bmax"], c2["bmax"]))
    assert np.all(covered == 1)
    assert len(nodes) <= 2 * n - 1


@pytest.mark.parametrize("seed", [0, 1])
def test_bvh_invariants_and_brute_force(seed):
    from oracle import OracleScene

    sc = _soup(3000, seed)
    o = OracleScene(sc)
    nodes, order = o.bvh()
    _check_bvh_invariants(sc, nodes, order)
    bmin, bmax = o.world_bounds()
    rays = scenes.random_rays(bmin, bmax, 4000, seed + 10)
    h, cnt = o.trace(rays, threads=4)
    b = o.trace_brute(rays)
    assert np.array_equal(h["prim"], b["prim"])
    assert np.array_equal(h["t"].view(np.int32), b["t"].view(np.int32))
    assert cnt.rays == 4000 and cnt.nodes >= cnt.rays
    # any-hit agrees with closest-hit on whether something is hit
    a, _ = o.trace(rays, any_hit=True, threads=4)
    assert np.array_equal(a["prim"] >= 0, h["prim"] >= 0)


def test_bvh_degenerate_centroids_make_one_leaf():
    from oracle import OracleScene

    # identical triangles: centroid extent 0 -> a single leaf in input order
    sc = scenes.SceneData()
    m = sc.add_material(scenes.diffuse((0.5, 0.5, 0.5)))
    p = np.tile([[0, 0, 0], [1, 0, 0], [0, 1, 0]], (7, 1)).astype(np.float64)
    sc.add_mesh(p, np.tile([0, 0, 1.0], (21, 1)), np.tile([1.0, 0, 0], (21, 1)), np.arange(21).reshape(-1, 3), m)
    nodes, order = OracleScene(sc).bvh()
    assert len(nodes) == 1 and nodes[0]["count"] == 7
    assert order.tolist() == list(range(7))
    # ray through all seven: the last tested one wins the tie
    rays = scenes.random_rays([0, 0, 0], [1, 1, 1], 1, 0)
    rays["o"] = [0.25, 0.25, 1.0]
    rays["d"] = [0.0, 0.0, -1.0]
    h, _ = OracleScene(sc).trace(rays, threads=1)
    assert h["prim"][0] == 6 and h["t"][0] == 1.0


def test_bvh_two_primitives_sorted_by_centroid():
    from oracle import OracleScene

    sc = scenes.SceneData()
    m = sc.add_material(scenes.diffuse((0.5, 0.5, 0.5)))
    p = np.array([[5, 0, 0], [6, 0, 0], [5, 1, 0], [0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64)
    sc.add_mesh(p, np.tile([0, 0, 1.0], (6, 1)), np.tile([1.0, 0, 0], (6, 1)), np.arange(6).reshape(-1, 3), m)
    nodes, order = OracleScene(sc).bvh()
    # nth_element on two: the smaller centroid first
    assert len(nodes) == 3 and order.tolist() == [1, 0] and nodes[0]["axis"] == 0


# --------------------------------------------------------------- Fresnel

def test_fresnel_closed_forms(orc):
    for eta in (1.33, 1.5, 2.4):
        r0 = ((eta - 1) / (eta + 1)) ** 2
        assert np.isclose(orc.orc_fresnel_dielectric(1.0, eta), r0, rtol=1e-6)
        # total internal reflection from the dense side
        assert orc.orc_fresnel_dielectric(-0.1, eta) == 1.0
        # grazing incidence reflects everything
        assert np.isclose(orc.orc_fresnel_dielectric(0.0, eta), 1.0, atol=1e-6)
        # k = 0 conductor equals the dielectric for eta > 1
        for c in (0.2, 0.5, 0.9):
            assert np.isclose(orc.orc_fresnel_complex(c, eta, 0.0), orc.orc_fresnel_dielectric(c, eta), rtol=2e-6)
    for eta, k in [(0.161, 4.08769), (0.492, 2.32625), (1.426, 1.846)]:
        expect = ((eta - 1) ** 2 + k * k) / ((eta + 1) ** 2 + k * k)
        assert np.isclose(orc.orc_fresnel_complex(1.0, eta, k), expect, rtol=2e-6)


# ----------------------------------------------------------------- BSDFs

N = farr(0, 0, 1)
T = farr(1, 0, 0)


def _eval(orc, m, wo, wi, front=1):
    f = farr(0, 0, 0)
    pdf = F()
    flags = orc.orc_bsdf_eval(ctypes.byref(m), N, T, front, farr(*wo), farr(*wi), f, ctypes.byref(pdf))
    return np.array(list(f)), pdf.value, flags


def _sample(orc, m, wo, u0, u12, front=1):
    f, wi, pdf, fl, eta = farr(0, 0, 0), farr(0, 0, 0), F(), ctypes.c_int32(), F()
    ok = orc.orc_bsdf_sample(ctypes.byref(m), N, T, front, farr(*wo), u0, farr(*u12), f, wi, ctypes.byref(pdf),
                             ctypes.byref(fl), ctypes.byref(eta))
    return ok, np.array(list(f)), np.array(list(wi)), pdf.value, fl.value, eta.value


def _dir(theta, phi):
    return (np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta))


def test_lambertian_closed_form(orc):
    m = scenes.diffuse((0.8, 0.4, 0.2))
    wo, wi = _dir(0.3, 0.1), _dir(0.9, 2.0)
    f, pdf, flags = _eval(orc, m, wo, wi)
    assert np.allclose(f, np.array([0.8, 0.4, 0.2]) / np.pi, rtol=1e-6)
    assert np.isclose(pdf, np.cos(0.9) / np.pi, rtol=1e-5)
    assert flags == 4 | 1
    f, pdf, _ = _eval(orc, m, wo, (wi[0], wi[1], -wi[2]))
    assert np.all(f == 0) and pdf == 0


# inside: also draw wo below the shading frame.  The plain DielectricBxDF
# samples the VNDF cap of wo's own hemisphere (sampling.h:206-232) and uses
# that normal unflipped (dielectric_bxdf.cpp:192-194), while PDF() flips the
# half vector up (:104-108); from below the two pdfs differ by construction
# in the reference.  The compensated variant flips the sampled normal
# (dielectric_multi_scattering_bxdf.cpp:255-259) and stays consistent.
# MetallicRoughnessBxDF's diffuse branch never flips wi below the frame
# (metallic_roughness_bxdf.cpp:133-137), so its f() is 0 there.  The
# compensated dielectric's Sample_f has no backfacing test on transmission:
# its VNDF normal can face wo from below (then Refract bends with the
# unflipped eta while eta' follows wo's side, :285-292), and its cosine-lobe
# transmission can yield a generalized half vector that f() and PDF() reject
# as backfacing (:44-47, :136-139); so it is compared for wo above the frame
# and without the cosine-lobe transmission samples (flags 6).
SKIP_NONE = ()
SKIP_DIFFUSE_TRANSMISSION = (4 | 2,)


def rough_diffuse(color, r):
    """DiffuseMaterial with roughness r > 0: the EON BxDF"""
    m = scenes.diffuse(color)
    m.u_roughness = m.v_roughness = r
    return m


@pytest.mark.parametrize("mat,pdf_consistent,inside,skip", [
    (scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.3), True, True, SKIP_NONE),
    (scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.6), True, True, SKIP_NONE),
    (scenes.dielectric(1.5, roughness=0.4), True, False, SKIP_NONE),
    # MetallicRoughnessBxDF::Sample_f weighs its lobes with the VNDF-sampled
    # half vector even when it then samples the diffuse lobe
    # (metallic_roughness_bxdf.cpp:106-151), while PDF() uses the half vector
    # of (wo, wi): the two pdfs differ by design, so only f is compared.
    (scenes.metallic_roughness((0.7, 0.5, 0.3), 0.0, 0.5), False, False, SKIP_NONE),
    (scenes.metallic_roughness((0.7, 0.5, 0.3), 1.0, 0.3), False, False, SKIP_NONE),
    (scenes.diffuse((0.5, 0.6, 0.7)), True, True, SKIP_NONE),
    (scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.5, energy_compensation=1), True, True, SKIP_NONE),
    (scenes.dielectric(1.5, roughness=0.5, energy_compensation=1), True, False, SKIP_DIFFUSE_TRANSMISSION),
    (scenes.dielectric(1.33, roughness=0.85, color=(0.9, 0.7, 0.5), energy_compensation=1), True, False,
     SKIP_DIFFUSE_TRANSMISSION),
    (rough_diffuse((0.6, 0.5, 0.4), 0.3), True, False, SKIP_NONE),
    (rough_diffuse((0.9, 0.9, 0.9), 1.0), True, False, SKIP_NONE),
], ids=["gold0.3", "gold0.6", "roughglass", "mr_dielectric", "mr_metal", "lambert", "gold0.5_ms", "roughglass0.5_ms",
        "water0.85_ms", "eon0.3", "eon1.0"])
def test_sampled_pdf_and_value_match_evaluation(orc, mat, pdf_consistent, inside, skip):
    """Sample_f's value and pdf equal f() and PDF() at the sampled direction,
    on both faces and from both sides of the shading frame."""
    rng = np.random.default_rng(3)
    n_ok = 0
    for it in range(300):
        theta = rng.uniform(0.05, 1.4)
        if inside and it % 3 == 2:
            theta = np.pi - theta  # wo below the shading frame
        front = 0 if it % 4 == 3 else 1
        wo = _dir(theta, rng.uniform(0, 2 * np.pi))
        ok, f, wi, pdf, fl, eta = _sample(orc, mat, wo, rng.random(), rng.random(2), front=front)
        if not ok or fl in skip:
            continue
        n_ok += 1
        f2, pdf2, _ = _eval(orc, mat, wo, wi, front=front)
        assert pdf > 0
        if pdf_consistent:
            assert np.isclose(pdf, pdf2, rtol=2e-4, atol=1e-7)
        assert np.allclose(f, f2, rtol=2e-4, atol=1e-7)
    assert n_ok > 100


@pytest.mark.parametrize("mat", [
    scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.4),
    scenes.dielectric(1.5, roughness=0.4),
    scenes.diffuse((0.3, 0.6, 0.9)),
    scenes.conductor(scenes.GOLD_ETA, scenes.GOLD_K, 0.4, energy_compensation=1),
    rough_diffuse((0.3, 0.6, 0.9), 0.7),
], ids=["gold", "roughglass", "lambert", "gold_ms", "eon"])
def test_bsdf_reciprocity(orc, mat):
    rng = np.random.default_rng(5)
    for _ in range(200):
        wo = _dir(rng.uniform(0.05, 1.4), rng.uniform(0, 2 * np.pi))
        wi = _dir(rng.uniform(0.05, 1.4), rng.uniform(0, 2 * np.pi))
        f1, _, _ = _eval(orc, mat, wo, wi)
        f2, _, _ = _eval(orc, mat, wi, wo)
        assert np.allclose(f1, f2, rtol=1e-4, atol=1e-7)


def test_ggx_albedo_near_one_for_perfect_reflector(orc):
    # R -> 1 conductor (F ~ 1): the directional albedo E[f cos / pdf] of a
    # smooth-ish GGX lobe approaches 1 and never exceeds it
    m = capi.BbMaterial()
    m.type = capi.BB_MAT_CONDUCTOR_R
    m.k[:] = [1.0, 1.0, 1.0]
    m.color[:] = [1.0, 1.0, 1.0]
    rng = np.random.default_rng(11)
    for rough, lo in ((0.1, 0.99), (0.5, 0.85)):
        m.u_roughness = m.v_roughness = rough
        wo = _dir(0.5, 0.2)
        acc = []
        for u in rng.random((3000, 3)):
            ok, f, wi, pdf, _, _ = _sample(orc, m, wo, u[0], u[1:])
            acc.append(f[0] * abs(wi[2]) / pdf if ok else 0.0)
        e = float(np.mean(acc))
        assert lo < e <= 1.0 + 3 * np.std(acc) / np.sqrt(len(acc)), (rough, e)


@pytest.mark.parametrize("r", [0.2, 0.6, 1.0])
def test_eon_white_furnace(orc, r):
    """EON is energy-preserving: with rho = 1 its multiple-scattering lobe
    returns exactly the energy the FON lobe loses, so the directional albedo
    E[f cos / pdf] is 1 for every wo (Portsmouth et al. 2025, eq. for
    rho_ms); a dropped or mis-weighted term moves it off 1."""
    m = rough_diffuse((1.0, 1.0, 1.0), r)
    rng = np.random.default_rng(21)
    for theta in (0.2, 0.8, 1.3):
        wo = _dir(theta, 0.4)
        acc = []
        for u in rng.random((20000, 3)):
            ok, f, wi, pdf, _, _ = _sample(orc, m, wo, u[0], u[1:])
            acc.append(f[0] * abs(wi[2]) / pdf if ok else 0.0)
        acc = np.array(acc)
        err = 4 * acc.std() / np.sqrt(len(acc))
        assert abs(acc.mean() - 1.0) < max(err, 2e-3), (r, theta, acc.mean(), err)


def test_eon_tends_to_lambert(orc):
    """With a vanishing roughness EON is the Lambertian lobe rho / pi."""
    m = rough_diffuse((0.5, 0.3, 0.8), 1e-6)
    rng = np.random.default_rng(22)
    for _ in range(50):
        wo = _dir(rng.uniform(0.05, 1.4), rng.uniform(0, 2 * np.pi))
        wi = _dir(rng.uniform(0.05, 1.4), rng.uniform(0, 2 * np.pi))
        f, _, _ = _eval(orc, m, wo, wi)
        assert np.allclose(f, np.array([0.5, 0.3, 0.8]) / np.pi, rtol=1e-4)


def test_smooth_dielectric_throughput(orc):
    m = scenes.dielectric(1.5)
    wo = _dir(0.6, 0.0)
    R = orc.orc_fresnel_dielectric(np.cos(0.6), 1.5)
    ok, f, wi, pdf, fl, eta = _sample(orc, m, wo, 0.0, (0.5, 0.5))  # u0 < R: reflection
    assert ok and fl == (16 | 1) and np.isclose(pdf, R, rtol=1e-6)
    assert np.isclose(f[0] * abs(wi[2]) / pdf, 1.0, rtol=1e-5)
    ok, f, wi, pdf, fl, eta = _sample(orc, m, wo, 0.99, (0.5, 0.5))  # transmission
    assert ok and fl == (16 | 2) and eta == np.float32(1.5)
    assert np.isclose(f[0] * abs(wi[2]) / pdf, 1.0 / 1.5 ** 2, rtol=1e-5)
    # Snell's law
    assert np.isclose(np.sqrt(wi[0] ** 2 + wi[1] ** 2), np.sin(0.6) / 1.5, rtol=1e-5)


# ---------------------------------------------------------------- lights

def test_power_light_pmf_closed_form():
    from oracle import OracleScene

    sc, _ = scenes.icosphere_scene(width=32, height=18, level=2, n_lights=16)
    o = OracleScene(sc)
    pmf = o.light_pmf()
    areas = []
    for L in sc.lights:
        v = sc.positions[sc.indices[sc.prims[L.primitive][1]]].astype(np.float64)
        areas.append(0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0])))
    power = np.array([L.spectrum[0] for L in sc.lights]) * np.array(areas) * np.pi
    assert np.isclose(pmf.sum(), 1.0, atol=1e-6)
    assert np.allclose(pmf, power / power.sum(), rtol=1e-5)


def test_mirror_reflects_with_weight_r(orc):
    """MirrorMaterial (SpecularReflectionBxDF, specular_reflection_bxdf.cpp):
    the sample is wo mirrored about the normal with f |cos| / pdf = R on
    both sides, and f and pdf of any pair are 0."""
    R = (0.9, 0.5, 0.25)
    m = scenes.mirror(R)
    for theta, phi in ((0.3, 0.2), (1.2, 2.5), (2.6, -1.0)):
        wo = _dir(theta, phi)
        ok, f, wi, pdf, flags, eta = _sample(orc, m, wo, 0.37, (0.1, 0.9))
        assert ok and pdf == 1.0 and eta == 1.0
        assert np.allclose(wi, [-wo[0], -wo[1], wo[2]], atol=1e-7)
        assert np.allclose(np.asarray(f) * abs(wi[2]) / pdf, R, rtol=1e-6)
        fe, pe, _ = _eval(orc, m, wo, wi)
        assert np.all(np.asarray(fe) == 0) and pe == 0


def test_directional_light_closed_form():
    """DirectionalLight (directional_light.cpp:29-38, a delta light): light
    arriving along w = Normalize(direction) with intensity E on a Lambertian
    plane gives rho / pi * E * cos(theta); its power is E pi R^2 with R the
    world's bounding-sphere radius."""
    from oracle import OracleScene

    sc = scenes.SceneData()
    m = sc.add_material(scenes.diffuse((0.6, 0.3, 0.15)))
    q = np.array([[-4, 0, -4], [4, 0, -4], [4, 0, 4], [-4, 0, 4]], np.float32)
    sc.add_mesh(q, [[0, 1, 0]] * 4, [[1, 0, 0]] * 4, [[0, 2, 1], [0, 3, 2]], m)
    E = np.array([2.0, 1.0, 0.5])
    sc.add_directional_light((0.0, 2.0, 2.0 * np.sqrt(3.0)), E)  # 60 degrees from the normal
    cam = scenes.make_camera((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), 0.05, 0.0, 1.0, 4, 4, up=(0.0, 0.0, -1.0))
    s, c, _ = OracleScene(sc).render(cam, scenes.render_params(16, max_bounces=1), threads=1)
    img = s / c[..., None]
    want = np.array([0.6, 0.3, 0.15]) / np.pi * E * 0.5
    assert np.allclose(img.reshape(-1, 3), want, rtol=1e-4), (img.reshape(-1, 3)[0], want)


def test_point_light_direct_lighting_closed_form():
    """A point light of intensity I at height h over a Lambertian plane of
    albedo rho, direct light only (max_bounces 1): straight below the light
    the radiance is rho / pi * I / h^2 (PointLight::Sample_Li,
    point_light.cpp:30-42; a delta light, so no MIS weight).  A dropped
    square, an MIS weight or a pdf other than 1 moves it.  The power-sampler
    PMF of a point light is 4 pi I's luminance over the total."""
    from oracle import OracleScene

    sc = scenes.SceneData()
    m = sc.add_material(scenes.diffuse((0.5, 0.25, 0.125)))
    q = np.array([[-50, 0, -50], [50, 0, -50], [50, 0, 50], [-50, 0, 50]], np.float32)
    sc.add_mesh(q, [[0, 1, 0]] * 4, [[1, 0, 0]] * 4, [[0, 2, 1], [0, 3, 2]], m)
    h, I = 2.0, np.array([3.0, 2.0, 1.0])
    sc.add_point_light((0.0, h, 0.0), I)
    sc.add_spot_light((5.0, 5.0, 5.0), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), 30.0, 20.0)
    cam = scenes.make_camera((0.0, 1.0, 0.0), (0.0, 0.0, 0.0), 0.05, 0.0, 1.0, 4, 4, up=(0.0, 0.0, -1.0))
    o = OracleScene(sc)
    pmf = o.light_pmf()
    lum = lambda c: 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]
    spot_phi = 2 * np.pi * ((1 - np.cos(np.radians(20))) + (np.cos(np.radians(20)) - np.cos(np.radians(30))) / 2)
    powers = np.array([4 * np.pi * lum(I), spot_phi * lum((1, 1, 1))])
    assert np.allclose(pmf, powers / powers.sum(), rtol=1e-5)
    # the point light alone: NEE then has no light-selection noise
    sc.lights = sc.lights[:1]
    s, c, _ = OracleScene(sc).render(cam, scenes.render_params(16, max_bounces=1), threads=1)
    img = s / c[..., None]
    want = np.array([0.5, 0.25, 0.125]) / np.pi * I / h ** 2
    assert np.allclose(img.reshape(-1, 3), want, rtol=1e-4), (img.reshape(-1, 3)[0], want)


# ------------------------------------------------------------ integrator

def test_furnace_convex_lambertian_sphere_in_uniform_light():
    """A convex Lambertian sphere lit only by a uniform environment L_env has
    radiance albedo * L_env everywhere; the background is L_env.  Pins NEE,
    BSDF sampling and their MIS weights together."""
    from oracle import OracleScene

    sc, cam = scenes.readme_scene(width=96, height=54, with_spot=False)
    o = OracleScene(sc)
    s, c, _ = o.render(cam, scenes.render_params(256), threads=os.cpu_count())
    img = s / c[..., None]
    assert np.all(c == 256)
    bg = img[:4].reshape(-1, 3)
    assert np.allclose(bg, 0.1, rtol=1e-5)  # float32 sum of 256 samples
    # pixels well inside the sphere's silhouette
    cy, cx = 54 // 2, 96 // 2
    inner = img[cy - 6:cy + 6, cx - 6:cx + 6].reshape(-1, 3)
    mean = inner.mean(axis=0)
    assert np.allclose(mean, [0.085, 0.02, 0.02], rtol=0.01), mean


def test_readme_render_is_deterministic_and_sample_ordered():
    from oracle import OracleScene

    sc, cam = scenes.readme_scene(width=48, height=27)
    o = OracleScene(sc)
    a, ca, _ = o.render(cam, scenes.render_params(4), threads=4)
    b, cb, _ = o.render(cam, scenes.render_params(4), threads=2)
    assert np.array_equal(a, b) and np.array_equal(ca, cb)
    # samples [0, 2) then [2, 4) accumulate to the same sums as [0, 4)
    film = (np.zeros_like(a), np.zeros_like(ca))
    o.render(cam, scenes.render_params(2), threads=4, film=film)
    o.render(cam, scenes.render_params(2, sample_begin=2), threads=4, film=film)
    assert np.array_equal(film[0], a) and np.array_equal(film[1], ca)


# ------------------------------------------- multiple-scattering compensation

def test_permutation_element_is_a_permutation(orc):
    """PermutationElement (hash.h:128-159, Kensler's correlated multi-jittered
    permutation) maps [0, l) onto itself one-to-one for any l and pattern."""
    for l, p in ((1024, 1), (1024, 2), (1000, 7), (37, 123456789)):
        perm = [orc.orc_permutation_element(i, l, p) for i in range(l)]
        assert sorted(perm) == list(range(l)), (l, p)
    assert [orc.orc_permutation_element(i, 1024, 1) for i in range(8)] != list(range(8))


def _tables(orc):
    E = (F * 1024)()
    E_avg = (F * 32)()
    u0 = (F * 1024)()
    u12 = (F * 2048)()
    orc.orc_reflectance_tables(E, E_avg, u0, u12)
    return (np.array(E, np.float64).reshape(32, 32), np.array(E_avg, np.float64), np.array(u0, np.float64),
            np.array(u12, np.float64).reshape(1024, 2))


def test_reflectance_table_samples_are_stratified(orc):
    """PrepareReflectanceTextures (microfacet.cpp:239-270): u0 has one sample
    in each of 1024 strata and u12 one in each cell of a 32 x 32 grid."""
    _, _, u0, u12 = _tables(orc)
    assert np.all((u0 >= 0) & (u0 < 1)) and np.all((u12 >= 0) & (u12 < 1))
    assert sorted(np.floor(u0 * 1024).astype(int)) == list(range(1024))
    cells = np.floor(u12[:, 1] * 32).astype(int) * 32 + np.floor(u12[:, 0] * 32).astype(int)
    assert sorted(cells) == list(range(1024))


def _ggx_albedo_quadrature(cos_o, alpha, n=400):
    """Single-scattering albedo of an isotropic GGX reflector with F = 1,
    E(wo) = int D(wm) G(wo, wi) / (4 cos_o) dwi, by quadrature over wi in
    float64 (independent of the oracle's Monte Carlo estimate)."""
    a2 = alpha * alpha
    sin_o = np.sqrt(1 - cos_o * cos_o)
    wo = np.array([sin_o, 0.0, cos_o])
    t = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * (2 * np.pi)
    T, P = np.meshgrid(t, ph, indexing="ij")
    wi = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    wm = wi + wo
    wm /= np.linalg.norm(wm, axis=-1, keepdims=True)
    c2 = wm[..., 2] ** 2
    D = 1.0 / (np.pi * a2 * c2 * c2 * (1 + (1 - c2) / c2 / a2) ** 2)

    def lam(w):
        c = w[..., 2] ** 2
        return (np.sqrt(1 + a2 * (1 - c) / c) - 1) / 2

    G = 1.0 / (1 + lam(wo[None, None, :] + 0 * wi) + lam(wi))
    f = D * G / (4 * cos_o * wi[..., 2])
    dw = np.sin(T) * (np.pi / 2 / n) * (2 * np.pi / (2 * n))
    return float(np.sum(f * wi[..., 2] * dw))


def test_ggx_reflectance_table_matches_quadrature(orc):
    """TrowbridgeReitzDistribution::PrepareReflectanceTexture (microfacet.cpp:
    31-75): E(cos_theta_i, alpha_j) is the albedo of GGX with F = 1 (Monte
    Carlo over the 1024 stratified samples); E_avg = 2 int E mu dmu."""
    E, E_avg, _, _ = _tables(orc)
    d = 1.0 / 32
    for j in (2, 8, 16, 28):
        a = d / 2 + d * j
        for i in (4, 12, 20, 30):
            c = d / 2 + d * i
            assert abs(E[j, i] - _ggx_albedo_quadrature(c, a)) < 0.01, (j, i)
        assert abs(E_avg[j] - 2 * np.sum(E[j] * (d / 2 + d * np.arange(32)) * d)) < 1e-5
    assert np.all(E <= 1.0 + 1e-3) and np.all(E > 0.3)
    assert np.all(np.diff(E_avg) < 1e-4)  # rougher loses more energy


def test_compensated_conductor_conserves_energy(orc):
    """ConductorMultiScatteringBxDF (conductor_multi_scattering_bxdf.cpp): for
    F -> 1 the single-scattering albedo E(wo) plus the compensation lobe
    (1 - E(wo)) is 1 at every roughness, while the plain ConductorBxDF loses
    energy as roughness grows.  Albedo by quadrature of f cos over wi."""
    n = 160
    t = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * (2 * np.pi)
    for rough in (0.3, 0.7, 0.95):
        albedo = {}
        for ec in (0, 1):
            m = capi.BbMaterial()
            m.type = capi.BB_MAT_CONDUCTOR_R
            m.k[:] = [1.0, 1.0, 1.0]
            m.color[:] = [1.0, 1.0, 1.0]
            m.u_roughness = m.v_roughness = rough
            m.energy_compensation = ec
            wo = _dir(0.6, 0.0)
            acc = 0.0
            for ti in t:
                for pj in ph[::4]:
                    f, _, _ = _eval(orc, m, wo, _dir(ti, pj))
                    acc += f[0] * np.cos(ti) * np.sin(ti)
            albedo[ec] = acc * (np.pi / 2 / n) * (2 * np.pi / (2 * n)) * 4
        assert abs(albedo[1] - 1.0) < 0.03, (rough, albedo)
        assert albedo[0] < albedo[1] - (0.02 if rough > 0.5 else 0.0), (rough, albedo)


# ------------------------------------------------ compensated rough dielectric

def _dielectric_tables(orc):
    n = 32
    bufs = [(F * (n ** 3))(), (F * (n ** 3))(), (F * (n * n))(), (F * (n * n))()]
    orc.orc_dielectric_tables(*bufs)
    E, E_inv = (np.array(b, np.float32).reshape(n, n, n) for b in bufs[:2])  # [alpha][cos][f0]
    A, A_inv = (np.array(b, np.float32).reshape(n, n) for b in bufs[2:])     # [alpha][f0]
    return E, E_inv, A, A_inv


def test_library_tables_equal_oracle_tables(orc):
    """libbbgpu computes the reflectance tables on the host with its own code
    and uploads them; they equal the oracle's bit for bit (the GPU renders
    read them), in the layouts that bbgpu_reflectance_table documents."""
    E, E_inv, A, A_inv = _dielectric_tables(orc)
    ggx_E, ggx_avg, _, _ = _tables(orc)
    pairs = [(capi.TABLE_GGX_E, ggx_E.astype(np.float32)), (capi.TABLE_GGX_E_AVG, ggx_avg.astype(np.float32)),
             (capi.TABLE_DIELECTRIC_E, E), (capi.TABLE_DIELECTRIC_E_INV, E_inv),
             (capi.TABLE_DIELECTRIC_E_AVG, A), (capi.TABLE_DIELECTRIC_E_INV_AVG, A_inv)]
    for table, ref in pairs:
        got = capi.reflectance_table(table)
        assert got.size == ref.size, table
        assert np.array_equal(got.view(np.int32), ref.reshape(-1).view(np.int32)), table


def _fresnel64(c, eta):
    """FresnelDielectric (scattering.h:51-77) in float64, c = cos_i >= 0"""
    s2t = (1 - c * c) / (eta * eta)
    if s2t >= 1:
        return 1.0
    ct = np.sqrt(1 - s2t)
    rp = (eta * c - ct) / (eta * c + ct)
    rs = (c - eta * ct) / (c + eta * ct)
    return (rp * rp + rs * rs) / 2


def _rough_dielectric_albedo_quadrature(eta, cos_o, alpha, n=700):
    """Directional albedo of the rough dielectric interface (reflection plus
    transmission, importance transport: no 1 / eta^2), by deterministic
    quadrature over the microfacet normal in float64.  A VNDF sample of wm
    has weight G(wo, wi) / G1(wo) whether it reflects (probability F) or
    refracts (1 - F), so E(wo) = int (wo.wm) D(wm) / cos_o [F G(wo, wr) +
    (1 - F) G(wo, wt)] dwm over visible normals, each term only where its
    direction lies on its side."""
    a2 = alpha * alpha
    wo = np.array([np.sqrt(1 - cos_o * cos_o), 0.0, cos_o])
    t = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * (2 * np.pi)
    T, P = np.meshgrid(t, ph, indexing="ij")
    wm = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1)
    c2 = wm[..., 2] ** 2
    D = 1.0 / (np.pi * a2 * c2 * c2 * (1 + (1 - c2) / c2 / a2) ** 2)

    def lam(w):
        c = np.clip(w[..., 2] ** 2, 1e-300, None)
        return (np.sqrt(1 + a2 * (1 - c) / c) - 1) / 2

    ci = wm @ wo  # wo . wm
    vis = ci > 0
    wr = -wo + 2 * ci[..., None] * wm
    s2t = (1 - ci * ci) / (eta * eta)
    tir = s2t >= 1
    ct = np.sqrt(np.clip(1 - s2t, 0, None))
    wt = -wo / eta + (ci / eta - ct)[..., None] * wm
    Fr = np.where(tir, 1.0, 0.0)
    ok = ~tir
    rp = (eta * ci - ct) / (eta * ci + ct)
    rs = (ci - eta * ct) / (ci + eta * ct)
    Fr = np.where(ok, (rp * rp + rs * rs) / 2, Fr)
    lo = lam(wo[None, None, :] + 0 * wm)
    Gr = 1.0 / (1 + lo + lam(wr))
    Gt = 1.0 / (1 + lo + lam(wt))
    term = Fr * Gr * (wr[..., 2] > 0) + (1 - Fr) * Gt * (ok & (wt[..., 2] < 0))
    dens = np.where(vis, ci * D / cos_o, 0.0)
    dw = np.sin(T) * (np.pi / 2 / n) * (2 * np.pi / (2 * n))
    return float(np.sum(dens * term * dw))


def test_dielectric_tables_match_quadrature(orc):
    """DielectricMultiScatteringBxDF::PrepareReflectanceTexture (microfacet.cpp:
    173-236): E(f0_k, cos_i, alpha_j) is BxDF::rho of DielectricBxDF(ior, GGX(a, a))
    in the ToCamera direction with ior = MapF0toIOR(f0) (scattering.h:188-208),
    and E_inv the same with 1 / ior; each agrees with an independent float64
    quadrature.  The nodes are asymmetric in (f0, cos, alpha), so a transposed
    layout fails; E_avg = 2 sum E mu dmu over the cos axis."""
    E, E_inv, A, A_inv = _dielectric_tables(orc)
    d = 1.0 / 32
    cen = d / 2 + d * np.arange(32)
    for k, i, j in ((14, 5, 9), (4, 20, 26), (25, 28, 14), (20, 10, 3)):
        f = cen[k]
        ior = (1 + f * f) / (1 - f * f)
        for table, eta in ((E, ior), (E_inv, 1 / ior)):
            q = _rough_dielectric_albedo_quadrature(eta, cen[i], cen[j])
            assert abs(table[j, i, k] - q) < 0.015, (k, i, j, eta, table[j, i, k], q)
    for table, avg in ((E, A), (E_inv, A_inv)):
        assert np.allclose(avg, 2 * np.einsum("jik,i->jk", table.astype(np.float64), cen * d), atol=2e-5)
    assert np.all(E <= 1.0 + 1e-3) and np.all(E_inv <= 1.0 + 1e-3) and np.all(E_inv > 0.05)


def _fresnel_average64(eta, n=20000):
    mu = (np.arange(n) + 0.5) / n
    return float(2 * np.sum([_fresnel64(m, eta) * m for m in mu]) / n)


def test_compensated_dielectric_lobes(orc):
    """DielectricMultiScatteringBxDF::f (dielectric_multi_scattering_bxdf.cpp:
    9-89): f minus the plain DielectricBxDF's f is the compensation lobe.
    (1) On wo's side it holds ratio (1 - E(wo)) of energy (quadrature over wi;
    E at a table node), which measures ComputeScatteringRatio (bxdfs.h:
    222-242) from outside and from inside.  (2) The two ratios satisfy the
    balance that defines them, (1 - ratio_out) / (1 - Favg(eta)) +
    (1 - ratio_in) / (1 - Favg(1 / eta)) = 1, with Favg = 2 int F mu dmu by
    quadrature (the reference uses the Imageworks fit, scattering.h:154-172).
    (3) Its transmission part obeys ft(o -> i) = eta^2 ft(i -> o) for wo
    outside (radiance transport, no 1 / eta'^2 on this lobe): the tables and
    the fit cancel, so this holds to rounding.  On the transmitted side f()
    also drops directions whose generalized half vector is backfacing
    (:44-47), so that side's energy is not (1 - ratio)(1 - E) in the
    reference and is not checked."""
    E, E_inv, _, _ = _dielectric_tables(orc)
    d = 1.0 / 32
    cen = d / 2 + d * np.arange(32)
    k, i, j = 14, 18, 24
    f0 = np.float32(cen[k])
    ior = float((1 + f0 * f0) / (1 - f0 * f0))
    rough = float(np.sqrt(cen[j]))
    mats = {ec: scenes.dielectric(ior, roughness=rough, energy_compensation=ec) for ec in (0, 1)}

    def lobe(wo, wi):
        f1, _, _ = _eval(orc, mats[1], wo, wi)
        f0_, _, _ = _eval(orc, mats[0], wo, wi)
        return f1[0] - f0_[0]

    n = 90
    t = (np.arange(n) + 0.5) / n * (np.pi / 2)
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * (2 * np.pi)
    dw = (np.pi / 2 / n) * (2 * np.pi / (2 * n))
    ratios = {}
    for side, table in ((1, E), (-1, E_inv)):
        wo = (np.sqrt(1 - cen[i] ** 2), 0.0, side * cen[i])
        acc = 0.0
        for ti in t:
            for pj in ph:
                wi = _dir(ti, pj)
                acc += lobe(wo, (wi[0], wi[1], side * wi[2])) * np.cos(ti) * np.sin(ti)
        ratios[side] = acc * dw / (1.0 - float(table[j, i, k]))
        assert 0.0 < ratios[side] < 1.0, ratios
    balance = (1 - ratios[1]) / (1 - _fresnel_average64(ior)) + (1 - ratios[-1]) / (1 - _fresnel_average64(1 / ior))
    assert abs(balance - 1.0) < 0.02, (ratios, balance)
    rng = np.random.default_rng(17)
    checked = 0
    for _ in range(60):
        wo = _dir(rng.uniform(0.1, 1.4), rng.uniform(0, 2 * np.pi))
        wi = _dir(np.pi - rng.uniform(0.1, 1.4), rng.uniform(0, 2 * np.pi))
        fo, fi = lobe(wo, wi), lobe(wi, wo)
        if fo == 0.0 or fi == 0.0:  # a backfacing generalized half vector
            continue
        checked += 1
        assert np.isclose(fo, ior * ior * fi, rtol=2e-3), (wo, wi, fo, fi)
    assert checked > 20
