"""Pins of the oracle's spot light, diffuse area light (Triangle::Sample and
Triangle::PDF) and MetallicRoughnessBxDF::f against closed forms of the
mathematics.  CPU only.

Each check_* takes an oracle library (oracle.load_oracle() or a build of
oracle.c with a planted mistake) and raises AssertionError when the oracle
disagrees with the closed form; tests/test_oracle_planted.py shows that each
check fails for the mistakes it is meant to catch."""
import ctypes

import numpy as np

from bulbit_amd import scenes

F = ctypes.c_float
RHO = np.array([0.5, 0.25, 0.125])


def farr(*v):
    return (F * len(v))(*v)


def _plane(sc, rho=RHO, half=50.0):
    m = sc.add_material(scenes.diffuse(tuple(rho)))
    q = np.array([[-half, 0, -half], [half, 0, -half], [half, 0, half], [-half, 0, half]], np.float32)
    sc.add_mesh(q, [[0, 1, 0]] * 4, [[1, 0, 0]] * 4, [[0, 2, 1], [0, 3, 2]], m)
    return m


def _look_down_at(p, pixels=4):
    """a pinhole 1 unit above p looking straight down with a 0.05 degree field
    of view: every camera ray hits the plane within 1e-3 of p"""
    p = np.asarray(p, np.float64)
    return scenes.make_camera(tuple(p + [0.0, 1.0, 0.0]), tuple(p), 0.05, 0.0, 1.0, pixels, pixels, up=(0.0, 0.0, -1.0))


def _direct(lib, sc, p, spp=16):
    """mean radiance leaving the Lambertian plane at p towards the camera with
    direct light only (max_bounces 1: NEE, then a BSDF sample whose next
    vertex is not lit)"""
    from oracle import OracleScene

    o = OracleScene(sc, lib=lib)
    s, c, _ = o.render(_look_down_at(p), scenes.render_params(spp, max_bounces=1), threads=4)
    assert np.all(c == spp)
    return (s / c[..., None]).reshape(-1, 3).mean(axis=0)


# ----------------------------------------------------------------- spot light

def _smooth_step(a, b, x):
    t = np.clip((x - a) / (b - a), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def check_spot_light(lib):
    """SpotLight::Sample_Li (spot_light.cpp:46-68): a delta light at L whose
    intensity I falls off with SmoothStep(cos theta_max, cos theta_min,
    cos alpha) (math.h:75-84), alpha the angle between the direction to the
    light and Normalize(direction); cos theta_max = cos(angle_max) (the outer
    cone), cos theta_min = cos(angle_falloff_start).  On a Lambertian plane of
    albedo rho the direct radiance is rho / pi * I * SmoothStep * cos(theta_s)
    / d^2 with theta_s the angle of incidence.  Points inside the inner cone,
    at a quarter, half and three quarters of the falloff band (SmoothStep
    0.15625, 0.5, 0.84375, not the linear 0.25, 0.5, 0.75) and outside the
    outer cone.  Catches swapped cone angles, a linear falloff, the frame's
    sign, a missing cosine or square."""
    Lp = np.array([0.4, 2.5, -0.3])
    direction = np.array([0.35, 1.0, -0.2])  # Normalize() is the light's job
    dn = direction / np.linalg.norm(direction)
    I = np.array([3.0, 2.0, 1.0])
    amax, afall = 35.0, 20.0
    cmax, cmin = np.cos(np.radians(amax)), np.cos(np.radians(afall))
    # the cone's axis meets the plane at p0; walk along a line in the plane
    p0 = Lp - dn * (Lp[1] / dn[1])
    walk = np.array([1.0, 0.0, 0.4]) / np.linalg.norm([1.0, 0.0, 0.4])

    def cos_alpha(s):
        p = p0 + walk * s
        w = Lp - p
        return float(np.dot(w / np.linalg.norm(w), dn))

    def point_at(c):  # bisection for the point along the walk with cos alpha = c
        lo, hi = 0.0, 20.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if cos_alpha(mid) > c else (lo, mid)
        return p0 + walk * (0.5 * (lo + hi))

    targets = [1.0, np.cos(np.radians(12.0))] + [cmax + f * (cmin - cmax) for f in (0.25, 0.5, 0.75)] + \
        [np.cos(np.radians(40.0))]
    sc = scenes.SceneData()
    _plane(sc)
    sc.add_spot_light(tuple(Lp), tuple(direction), tuple(I), amax, afall)
    for c in targets:
        p = p0 if c == 1.0 else point_at(c)
        w = Lp - p
        d = np.linalg.norm(w)
        cos_s = w[1] / d
        want = RHO / np.pi * I * _smooth_step(cmax, cmin, cos_alpha_at(Lp, p, dn)) * cos_s / d ** 2
        got = _direct(lib, sc, p)
        if np.all(want == 0):
            assert np.all(got == 0), (c, got)
        else:
            assert np.allclose(got, want, rtol=2e-3), (c, got, want)


def cos_alpha_at(Lp, p, dn):
    w = Lp - p
    return float(np.dot(w / np.linalg.norm(w), dn))


def test_spot_light_falloff_closed_form(orc):
    check_spot_light(orc)


# ----------------------------------------------------------- triangle light

TRI = np.array([[-0.6, 2.0, -0.4], [0.7, 2.2, -0.3], [0.1, 1.9, 0.8]], np.float32)  # cross(e1, e2) points down (-y)
EMIT = np.array([4.0, 2.0, 1.0])


def _light_scene(two_sided=False, with_plane=True):
    sc = scenes.SceneData()
    if with_plane:
        _plane(sc)
    m = sc.add_material(scenes.diffuse((0.0, 0.0, 0.0)))
    n = np.cross(TRI[1] - TRI[0], TRI[2] - TRI[0])
    n = n / np.linalg.norm(n)
    t = (TRI[1] - TRI[0]) / np.linalg.norm(TRI[1] - TRI[0])
    prim = sc.add_mesh(TRI, [n] * 3, [t] * 3, [[0, 1, 2]], m)
    sc.add_area_light(prim[0], tuple(EMIT), two_sided=two_sided)
    return sc, len(sc.lights) - 1


def solid_angle(p, tri):
    """Van Oosterom and Strackee (1983): the solid angle of a triangle seen from p"""
    a, b, c = (np.asarray(v, np.float64) - p for v in tri)
    la, lb, lc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(c)
    num = abs(np.dot(a, np.cross(b, c)))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2.0 * np.arctan2(num, den)


def projected_solid_angle(p, n, tri):
    """Lambert's polygon formula: the integral of cos(theta) over the solid
    angle of the polygon, 1/2 sum_i angle(u_i, u_i+1) (n . unit(u_i x u_i+1)),
    with u_i the unit vectors from p to the vertices"""
    u = [(np.asarray(v, np.float64) - p) / np.linalg.norm(np.asarray(v, np.float64) - p) for v in tri]
    s = 0.0
    for i in range(len(u)):
        a, b = u[i], u[(i + 1) % len(u)]
        g = np.cross(a, b)
        s += np.arccos(np.clip(np.dot(a, b), -1.0, 1.0)) * np.dot(n, g / np.linalg.norm(g))
    return abs(0.5 * s)


def _grid(n):
    g = (np.arange(n) + 0.5) / n
    u, v = np.meshgrid(g, g, indexing="ij")
    return np.stack([u.ravel(), v.ravel()], 1).astype(np.float32)


def check_area_light_sampling(lib):
    """DiffuseAreaLight::Sample_Li over Triangle::Sample (diffuse_area_light.cpp:
    42-72, triangle.cpp:159-229): the sampled points are uniform on the
    triangle (barycentric means 1/3 and second moments 1/6) and lie on it; the
    solid-angle pdf makes E[1 / pdf] the triangle's solid angle (Van Oosterom-
    Strackee) and E[cos theta / pdf] its projected solid angle (Lambert's
    formula); Li is the emission seen from the front, 0 from the back of a
    one-sided light and the emission from either side of a two-sided one."""
    from oracle import OracleScene

    p = np.array([0.2, 0.0, 0.3])
    sc, li = _light_scene()
    o = OracleScene(sc, lib=lib)
    u = _grid(400)
    s = o.light_sample(li, p, u)
    assert np.all(s["ok"] == 1)
    pdf = s["pdf"].astype(np.float64)
    wi = s["wi"].astype(np.float64)
    assert np.all(pdf > 0)
    assert np.allclose(np.linalg.norm(wi, axis=1), 1.0, atol=1e-6)
    assert np.allclose(s["Li"], EMIT)
    # sampled points: p + wi * (visibility + Ray::epsilon)
    pts = p + wi * (s["visibility"].astype(np.float64) + 1e-4)[:, None]
    t0, t1, t2 = (TRI[k].astype(np.float64) for k in range(3))
    e1, e2 = t1 - t0, t2 - t0
    nrm = np.cross(e1, e2)
    assert np.max(np.abs((pts - t0) @ nrm) / np.linalg.norm(nrm)) < 1e-5
    # barycentrics by least squares in the triangle's plane
    M = np.stack([e1, e2], 1)
    bc = np.linalg.lstsq(M, (pts - t0).T, rcond=None)[0].T
    b = np.column_stack([1.0 - bc.sum(1), bc])
    assert b.min() > -1e-4 and b.max() < 1 + 1e-4
    assert np.allclose(b.mean(0), 1 / 3, atol=2e-3), b.mean(0)
    assert np.allclose((b * b).mean(0), 1 / 6, atol=2e-3), (b * b).mean(0)
    omega = solid_angle(p, TRI)
    assert np.isclose(np.mean(1.0 / pdf), omega, rtol=2e-3), (np.mean(1.0 / pdf), omega)
    proj = projected_solid_angle(p, np.array([0.0, 1.0, 0.0]), TRI)
    assert np.isclose(np.mean(wi[:, 1] / pdf), proj, rtol=2e-3), (np.mean(wi[:, 1] / pdf), proj)
    # from above, the back of the one-sided light: no emission
    above = np.array([0.1, 4.0, 0.0])
    s = o.light_sample(li, above, _grid(8))
    assert np.all(s["ok"] == 1) and np.all(s["Li"] == 0)
    sc2, li2 = _light_scene(two_sided=True)
    s = OracleScene(sc2, lib=lib).light_sample(li2, above, _grid(8))
    assert np.allclose(s["Li"], EMIT)


def test_area_light_sampling_closed_forms(orc):
    check_area_light_sampling(orc)


def check_area_light_hit_pdf(lib):
    """Triangle::PDF(isect, ray) (triangle.cpp:247-253), the density an
    emitter hit's MIS weight uses: over the sphere of directions it integrates
    to 1 (stratified directions in a cone holding the triangle, traced through
    the oracle's BVH), and along sampled directions it equals the pdf that
    Sample_Li returned."""
    from oracle import OracleScene

    p = np.array([0.2, 0.0, 0.3])
    sc, li = _light_scene()
    o = OracleScene(sc, lib=lib)
    c = TRI.astype(np.float64).mean(0) - p
    axis = c / np.linalg.norm(c)
    vs = [(v - p) / np.linalg.norm(v - p) for v in TRI.astype(np.float64)]
    gamma = max(np.arccos(np.clip(np.dot(axis, v), -1, 1)) for v in vs) * 1.05
    cos_g = np.cos(gamma)
    g = _grid(700).astype(np.float64)
    cos_t = 1.0 - g[:, 0] * (1.0 - cos_g)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t ** 2))
    phi = 2 * np.pi * g[:, 1]
    tx = np.cross(axis, [1.0, 0.0, 0.0])
    tx /= np.linalg.norm(tx)
    ty = np.cross(axis, tx)
    d = (np.cos(phi) * sin_t)[:, None] * tx + (np.sin(phi) * sin_t)[:, None] * ty + cos_t[:, None] * axis
    pdf = o.light_hit_pdf(li, p, d).astype(np.float64)
    cone = 2 * np.pi * (1.0 - cos_g)
    integral = cone * pdf.mean()
    assert np.isclose(integral, 1.0, rtol=3e-3), integral
    # interior points of the triangle (a ray to an edge point may miss it by rounding)
    s = o.light_sample(li, p, 0.05 + 0.4 * _grid(40))
    hp = o.light_hit_pdf(li, p, s["wi"])
    assert np.all(hp > 0)
    assert np.allclose(hp, s["pdf"], rtol=2e-4), np.max(np.abs(hp / s["pdf"] - 1))


def test_area_light_hit_pdf_integrates_to_one(orc):
    check_area_light_hit_pdf(orc)


BIG_TRI = np.array([[-1.5, 0.5, -1.2], [1.6, 0.6, -0.9], [0.1, 0.45, 1.7]], np.float32)  # faces down, 0.5 above


def check_area_light_direct(lib):
    """The rendered direct light of one large emissive triangle close above
    a Lambertian plane (max_bounces 1: NEE and the BSDF-sampled emitter hit,
    combined by the power heuristic) is rho / pi * L_e * (projected solid
    angle) in expectation.  The triangle covers 3.5 sr, so both strategies
    carry weight; their weights only sum to one when the NEE density and the
    emitter-hit density agree, so a wrong Triangle::PDF at the hit or a wrong
    sampling pdf biases the estimate (by 3 to 44 % for the planted cases)."""
    from oracle import OracleScene

    p = np.array([0.2, 0.0, 0.3])
    sc = scenes.SceneData()
    _plane(sc)
    m = sc.add_material(scenes.diffuse((0.0, 0.0, 0.0)))
    n = np.cross(BIG_TRI[1] - BIG_TRI[0], BIG_TRI[2] - BIG_TRI[0])
    t = BIG_TRI[1] - BIG_TRI[0]
    prim = sc.add_mesh(BIG_TRI, [n / np.linalg.norm(n)] * 3, [t / np.linalg.norm(t)] * 3, [[0, 1, 2]], m)
    sc.add_area_light(prim[0], tuple(EMIT))
    # the camera looks down from below the light
    cam = scenes.make_camera(tuple(p + [0.0, 0.2, 0.0]), tuple(p), 0.05, 0.0, 1.0, 4, 4, up=(0.0, 0.0, -1.0))
    spp = 8192
    s, c, _ = OracleScene(sc, lib=lib).render(cam, scenes.render_params(spp, max_bounces=1), threads=8)
    assert np.all(c == spp)
    got = (s / c[..., None]).reshape(-1, 3).mean(axis=0)
    want = RHO / np.pi * EMIT * projected_solid_angle(p, np.array([0.0, 1.0, 0.0]), BIG_TRI)
    assert np.allclose(got, want, rtol=1.2e-2), (got, want)


def test_area_light_direct_lighting_closed_form(orc):
    check_area_light_direct(orc)


# -------------------------------------------------- metallic-roughness f

def _mr_eval(lib, color, metallic, roughness, wo, wi):
    m = scenes.metallic_roughness(tuple(color), metallic, roughness)
    f, pdf = farr(0, 0, 0), F()
    lib.orc_bsdf_eval(ctypes.byref(m), farr(0, 0, 1), farr(1, 0, 0), 1, farr(*wo), farr(*wi), f, ctypes.byref(pdf))
    return np.array(list(f), np.float64)


def _mr_model(color, metallic, roughness, wo, wi):
    """The metallic-roughness model of MetallicRoughnessBxDF::f
    (metallic_roughness_bxdf.cpp:8-38, bxdfs.h:475-489) in float64 for
    isotropic roughness, written from its published parts: f0 = lerp(0.04,
    color, metallic); Schlick F = f0 + (1 - f0) (1 - wi.wm)^5; GGX D(wm) =
    1 / (pi alpha^2 cos^4 (1 + tan^2 / alpha^2)^2); height-correlated Smith
    G = 1 / (1 + Lambda(wo) + Lambda(wi)), Lambda = (sqrt(1 + alpha^2
    tan^2) - 1) / 2; alpha = max(roughness^2, 1e-3); f = (1 - metallic)
    (1 - F) color / pi + F D G / (4 cos_i cos_o).  wo, wi: [n, 3] above the
    surface; returns [n, 3]"""
    color = np.asarray(color, np.float64)
    wo = np.atleast_2d(np.asarray(wo, np.float64))
    wi = np.atleast_2d(np.asarray(wi, np.float64))
    alpha = max(roughness ** 2, 1e-3)
    wm = wo + wi
    wm /= np.linalg.norm(wm, axis=1, keepdims=True)
    f0 = 0.04 * (1.0 - metallic) + color * metallic
    F_ = f0 + (1.0 - f0) * ((1.0 - np.sum(wi * wm, 1)) ** 5)[:, None]
    cm = wm[:, 2]
    tan2m = (1.0 - cm * cm) / (cm * cm)
    D = 1.0 / (np.pi * alpha * alpha * cm ** 4 * (1.0 + tan2m / (alpha * alpha)) ** 2)

    def lam(w):
        c2 = w[:, 2] ** 2
        return (np.sqrt(1.0 + alpha * alpha * (1.0 - c2) / c2) - 1.0) / 2.0

    G = 1.0 / (1.0 + lam(wo) + lam(wi))
    spec = F_ * (D * G / (4.0 * wi[:, 2] * wo[:, 2]))[:, None]
    return (1.0 - metallic) * (1.0 - F_) * color / np.pi + spec


def check_metallic_roughness_f(lib):
    """MetallicRoughnessBxDF::f against the model in float64: at normal
    incidence (Schlick's F is f0 itself: metallic 0 gives 0.04, metallic 1
    the basecolor), at mirrored pairs (half vector = normal) and at random
    pairs, for metallic 0, 1 and a mix.  The white directional albedo
    (midpoint quadrature of f cos over the hemisphere) is at most 1 up to
    theta_o = 0.6; the model itself exceeds 1 towards grazing view (1.005 at
    theta_o = 0.9 and 1.09 at 1.3 for roughness 0.4: its diffuse weight
    (1 - F(wi.wm)) is not the specular lobe's complement), and there the
    oracle's albedo equals the float64 model's.  Catches a reversed f0 lerp,
    a diffuse weight without (1 - F) or (1 - metallic), a wrong Schlick
    exponent, D or G."""
    color = (0.9, 0.5, 0.2)
    # normal incidence: D(n) = 1 / (pi alpha^2), G = 1, F = f0
    for metallic, f0 in ((0.0, np.full(3, 0.04)), (1.0, np.asarray(color))):
        a = 0.5 ** 2
        got = _mr_eval(lib, color, metallic, 0.5, (0, 0, 1), (0, 0, 1))
        want = (1.0 - metallic) * (1.0 - f0) * np.asarray(color) / np.pi + f0 / (4 * np.pi * a * a)
        assert np.allclose(got, want, rtol=2e-5), (metallic, got, want)
    rng = np.random.default_rng(11)
    for metallic in (0.0, 1.0, 0.35):
        for roughness in (0.3, 0.7):
            for theta in (0.5, 1.1):  # mirrored pairs
                wo = (np.sin(theta), 0.0, np.cos(theta))
                wi = (-np.sin(theta), 0.0, np.cos(theta))
                got = _mr_eval(lib, color, metallic, roughness, wo, wi)
                want = _mr_model(color, metallic, roughness, wo, wi)[0]
                assert np.allclose(got, want, rtol=2e-5), (metallic, roughness, theta, got, want)
            for _ in range(60):  # random pairs above the surface
                to, ti = rng.uniform(0.02, 1.45, 2)
                po, pi_ = rng.uniform(0, 2 * np.pi, 2)
                wo = (np.sin(to) * np.cos(po), np.sin(to) * np.sin(po), np.cos(to))
                wi = (np.sin(ti) * np.cos(pi_), np.sin(ti) * np.sin(pi_), np.cos(ti))
                got = _mr_eval(lib, color, metallic, roughness, wo, wi)
                want = _mr_model(color, metallic, roughness, wo, wi)[0]
                assert np.allclose(got, want, rtol=2e-4, atol=1e-7), (metallic, roughness, wo, wi, got, want)
    # white directional albedo
    nt, npf = 90, 180
    th = (np.arange(nt) + 0.5) / nt * (np.pi / 2)
    ph = (np.arange(npf) + 0.5) / npf * (2 * np.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    dw = ((np.pi / 2 / nt) * (2 * np.pi / npf) * np.sin(T)).reshape(-1)
    wis = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    for roughness in (0.4, 0.7):
        for th_o in (0.2, 0.6, 1.3):
            wo = (np.sin(th_o), 0.0, np.cos(th_o))
            f = np.array([_mr_eval(lib, (1, 1, 1), 0.0, roughness, wo, w)[0] for w in wis])
            E = float(np.sum(f * wis[:, 2] * dw))
            E64 = float(np.sum(_mr_model((1, 1, 1), 0.0, roughness, np.tile(wo, (len(wis), 1)), wis)[:, 0] *
                               wis[:, 2] * dw))
            assert np.isclose(E, E64, rtol=1e-4), (roughness, th_o, E, E64)
            if th_o <= 0.6:
                assert 0.5 < E <= 1.0, (roughness, th_o, E)


def test_metallic_roughness_f_closed_form(orc):
    check_metallic_roughness_f(orc)
