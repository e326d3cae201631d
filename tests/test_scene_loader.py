"""Mitsuba 3 XML scenes through LoadScene (cli/scene_loader.cpp).

CPU tests load a scene with `describe_scene`, which prints what the C++ API
built, and check it against values computed independently here: transforms
in float64 (rotation matrices, not quaternions), the lookat orientation
against the library's own LookAt, and the reference loader's element rules
(its checkerboard colour order, sqrt of the roughness `alpha`, the ior ratio,
defaults, the OBJ transform).  The GPU test renders the scene the way the
CLI does (`render_scene`: LoadScene, the Create factories, Render) and
compares the film with the described scene rendered through the C ABI
(bit-equal) and by the oracle.
"""

import json
import os
import subprocess

import numpy as np
import pytest

from bulbit_amd import build, capi, scenes

from conftest import pixel_agreement
from test_cpp_api import read_pfm

F = np.float32

XML = """<?xml version="1.0" encoding="utf-8"?>
<!-- every element kind the GPU path renders -->
<scene version="3.0.0">
    <default name="spp" value="16"/>
    <default name="res" value="96"/>
    <integrator type="path">
        <integer name="max_depth" value="6"/>
        <integer name="rr_depth" value="2"/>
    </integrator>
    <sensor type="perspective">
        <float name="fov" value="45"/>
        <transform name="to_world">
            <lookat origin="0, 1.5, 5" target="0, 0.5, 0" up="0, 1, 0"/>
        </transform>
        <sampler type="independent">
            <integer name="sample_count" value="$spp"/>
        </sampler>
        <film type="hdrfilm">
            <integer name="width" value="$res"/>
            <integer name="height" value="64"/>
            <rfilter type="gaussian"><float name="stddev" value="0.4"/></rfilter>
        </film>
    </sensor>

    <bsdf type="twosided" id="floor">
        <bsdf type="diffuse">
            <texture type="checkerboard" name="reflectance">
                <rgb name="color0" value="0.8 0.1 0.1"/>
                <rgb name="color1" value="0.1, 0.1, 0.8"/>
                <transform name="to_uv"><scale x="4" y="3"/></transform>
            </texture>
        </bsdf>
    </bsdf>
    <bsdf type="roughconductor" id="metal">
        <rgb name="eta" value="0.2 0.9 1.1"/>
        <rgb name="k" value="3.9 2.4 2.2"/>
        <float name="alpha" value="0.09"/>
    </bsdf>
    <bsdf type="roughdielectric" id="glass">
        <float name="int_ior" value="1.5"/>
        <float name="ext_ior" value="1.0"/>
        <float name="alpha_u" value="0.04"/>
        <float name="alpha_v" value="0.16"/>
    </bsdf>
    <bsdf type="diffuse" id="rough">
        <rgb name="reflectance" value="0.7 0.6 0.3"/>
        <float name="roughness" value="0.5"/>
    </bsdf>
    <bsdf type="normalmap" id="bumpy">
        <texture type="bitmap" name="normalmap"><string name="filename" value="normal.ppm"/></texture>
        <bsdf type="diffuse"><srgb name="reflectance" value="0.5 0.25 0.75"/></bsdf>
    </bsdf>
    <bsdf type="mask" id="cutout">
        <texture type="checkerboard" name="opacity">
            <rgb name="color0" value="1"/>
            <rgb name="color1" value="0.3"/>
        </texture>
        <bsdf type="diffuse"><rgb name="reflectance" value="0.9"/></bsdf>
    </bsdf>

    <shape type="rectangle">
        <transform name="to_world"><scale value="3"/><rotate x="1" angle="-90"/></transform>
        <ref id="floor"/>
    </shape>
    <shape type="cube">
        <transform name="to_world"><scale value="0.4"/><translate x="-1" y="0.4" z="0.2"/></transform>
        <ref id="metal"/>
    </shape>
    <shape type="sphere">
        <point name="center" x="0.9" y="0.5" z="0.1"/>
        <float name="radius" value="0.5"/>
        <ref id="glass"/>
    </shape>
    <shape type="obj">
        <string name="filename" value="tri.obj"/>
        <transform name="to_world"><rotate y="1" angle="30"/><translate x="0.2" y="0.05" z="1.2"/></transform>
        <ref id="rough"/>
    </shape>
    <shape type="cube">
        <transform name="to_world">
            <matrix value="0.25 0 0 0.3  0 0.3 0.1 0.35  0 -0.1 0.3 -0.8  0 0 0 1"/>
        </transform>
        <ref id="bumpy"/>
    </shape>
    <shape type="rectangle">
        <transform name="to_world"><scale x="0.6" y="0.4"/><translate x="-0.2" y="0.6" z="-1.2"/></transform>
        <ref id="cutout"/>
    </shape>
    <shape type="rectangle">
        <transform name="to_world"><scale value="0.5"/><rotate x="1" angle="90"/><translate y="2.5"/></transform>
        <bsdf type="diffuse"><rgb name="reflectance" value="0.5"/></bsdf>
        <emitter type="area"><rgb name="radiance" value="10 9 8"/></emitter>
    </shape>
    <shape type="sphere">
        <point name="center" x="-0.3" y="0.25" z="0.8"/>
        <float name="radius" value="0.25"/>
        <bsdf type="conductor"><string name="material" value="none"/><rgb name="eta" value="1"/></bsdf>
    </shape>
    <emitter type="spot">
        <rgb name="intensity" value="4"/>
        <float name="cutoff_angle" value="30"/>
        <transform name="to_world"><lookat origin="2 3 2" target="0 0 0" up="0 1 0"/></transform>
    </emitter>
    <emitter type="point">
        <point name="position" x="0.5" y="1.5" z="1"/>
        <rgb name="intensity" value="2 1.5 1"/>
        <transform name="to_world"><translate x="0.25"/></transform>
    </emitter>
    <emitter type="directional">
        <vector name="direction" value="0, 0, 1"/>
        <rgb name="irradiance" value="0.5 0.4 0.3"/>
        <transform name="to_world"><rotate x="1" angle="-60"/></transform>
    </emitter>
    <emitter type="constant"><rgb name="radiance" value="0.1"/></emitter>
</scene>
"""

OBJ = """v -0.3 0 0
v 0.3 0 0
v 0 0.6 0
f 1 2 3
"""

NORMAL8 = np.array([[[128, 128, 255], [160, 128, 230]], [[128, 100, 240], [110, 140, 235]]], np.uint8)


def _exe(path):
    if not os.path.exists(path):
        build.build_cpp()
    return path


def _write_files(tmp_path, xml=XML):
    (tmp_path / "scene.xml").write_text(xml)
    (tmp_path / "tri.obj").write_text(OBJ)
    with open(tmp_path / "normal.ppm", "wb") as f:
        f.write(b"P6\n2 2\n255\n" + NORMAL8.tobytes())
    return tmp_path / "scene.xml"


def describe(path):
    r = subprocess.run([_exe(build.CPP_DESCRIBE_SCENE), str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def load_fails(tmp_path, xml):
    path = _write_files(tmp_path, xml)
    r = subprocess.run([_exe(build.CPP_DESCRIBE_SCENE), str(path)], capture_output=True, text=True, timeout=60)
    return r.returncode, r.stderr


@pytest.fixture(scope="module")
def desc(tmp_path_factory):
    return describe(_write_files(tmp_path_factory.mktemp("xml")))


@pytest.mark.parametrize("face", ["f 1//2 2//1 3//1", "f 1/4 2/1 3/1", "f 1/1/1 2/1/1 3/1/0", "f 1 2 -4", "f 0 2 3"])
def test_obj_faces_naming_missing_vertices_are_refused(tmp_path, desc, face):
    # a texture coordinate, normal or position the file does not have fails
    # the model's read instead of reading past its arrays
    path = _write_files(tmp_path)
    (tmp_path / "tri.obj").write_text(OBJ.replace("f 1 2 3", "vt 0 0\nvn 0 0 1\n" + face))
    d = describe(path)
    assert len(d["primitives"]) == len(desc["primitives"]) - 1  # the OBJ's one triangle is missing
    (tmp_path / "tri.obj").write_text(OBJ.replace("f 1 2 3", "vt 0 0\nvn 0 0 1\nf 1/1/1 2/1/1 -1/-1/-1"))
    assert len(describe(path)["primitives"]) == len(desc["primitives"])


def _rot(axis, deg):
    """rotation matrix about a unit axis (Rodrigues), float64"""
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    t = np.radians(deg)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(t) * K + (1 - np.cos(t)) * K @ K


def _tri_positions(d, prim):
    return np.array([c["p"] for c in d["primitives"][prim]["triangle"]], float)


# corners of the unit quad and box faces as the builders lay them out
# (cli/primitive_builder.cpp:105-258): triangles (0, 1, 2), (0, 2, 3) per face
QUAD = np.array([[-0.5, -0.5, 0], [0.5, -0.5, 0], [0.5, 0.5, 0], [-0.5, 0.5, 0]])
BOX_FACES = [  # corner positions, normal
    ([[-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], [0, 0, 1]),
    ([[1, -1, 1], [1, -1, -1], [1, 1, -1], [1, 1, 1]], [1, 0, 0]),
    ([[1, -1, -1], [-1, -1, -1], [-1, 1, -1], [1, 1, -1]], [0, 0, -1]),
    ([[-1, -1, -1], [-1, -1, 1], [-1, 1, 1], [-1, 1, -1]], [-1, 0, 0]),
    ([[-1, 1, 1], [1, 1, 1], [1, 1, -1], [-1, 1, -1]], [0, 1, 0]),
    ([[1, -1, 1], [-1, -1, 1], [-1, -1, -1], [1, -1, -1]], [0, -1, 0]),
]


def _quad_tris(corners):
    c = np.asarray(corners, float)
    return [c[[0, 1, 2]], c[[0, 2, 3]]]


def test_camera_film_sampler_and_integrator(desc):
    cam = desc["camera"]
    assert cam["transform"]["p"] == [0, 1.5, 5] and cam["transform"]["s"] == [1, 1, 1]
    # <lookat> is Transform(origin, Quat(Normalize(target - origin), up)): the LookAt orientation
    assert np.array_equal(F(cam["transform"]["q"]), F(scenes.look_at((0, 1.5, 5), (0, 0.5, 0), (0, 1, 0))))
    # the horizontal fov of 45 degrees at aspect 96/64, as a vertical fov
    assert cam["fov"] == pytest.approx(np.degrees(2 * np.arctan(np.tan(np.radians(22.5)) / 1.5)), abs=2e-5)
    assert cam["resolution"] == [96, 64] and cam["spp"] == 16
    assert cam["filter"] == {"type": "gaussian", "extent": 3, "stddev": pytest.approx(0.4)}
    assert cam["aperture_radius"] == 0 and cam["focus_distance"] == 1
    assert desc["integrator"] == {"max_bounces": 6, "rr_min_bounces": 2}


def test_materials_follow_the_reference_rules(desc):
    m = desc["materials"]
    assert [x["type"] for x in m] == ["diffuse", "conductor", "dielectric", "diffuse", "diffuse", "diffuse", "diffuse",
                                      "mirror"]
    floor = m[0]["reflectance"]
    # color0 is the checker's second value, color1 its first; resolution = 2 x to_uv's scale
    assert floor["checker"][0]["constant"] == pytest.approx([0.1, 0.1, 0.8])
    assert floor["checker"][1]["constant"] == pytest.approx([0.8, 0.1, 0.1])
    assert floor["resolution"] == [8, 6]
    metal = m[1]
    assert metal["eta"]["constant"] == pytest.approx([0.2, 0.9, 1.1])
    assert metal["k"]["constant"] == pytest.approx([3.9, 2.4, 2.2])
    assert metal["u_roughness"]["constant"] == pytest.approx(0.3)  # sqrt(alpha)
    assert metal["v_roughness"] == metal["u_roughness"] and metal["energy_compensation"] and metal["reflectance"] is None
    glass = m[2]
    assert glass["eta"] == pytest.approx(1.5) and glass["energy_compensation"]
    assert glass["u_roughness"]["constant"] == pytest.approx(0.2) and glass["v_roughness"]["constant"] == pytest.approx(0.4)
    assert glass["reflectance"]["constant"] == [1, 1, 1]
    rough = m[3]
    assert rough["roughness"]["constant"] == 0.5 and rough["reflectance"]["constant"] == pytest.approx([0.7, 0.6, 0.3])
    bumpy = m[4]
    # an srgb triple decoded by RGB_from_sRGB (include/bulbit/color.h:10-29)
    srgb = np.array([0.5, 0.25, 0.75])
    lin = np.where(srgb <= 0.04045, srgb / 12.92, ((srgb + 0.055) / 1.055) ** 2.4)
    assert bumpy["reflectance"]["constant"] == pytest.approx(lin, rel=1e-6)
    # the normal map as non-colour data, bottom row first
    assert bumpy["normal"]["image"] == [2, 2]
    texels = np.array(bumpy["normal"]["texels"]).reshape(2, 2, 3)
    assert np.array_equal(F(texels), (NORMAL8[::-1].astype(F) / F(255)))
    cutout = m[5]
    assert cutout["alpha"]["checker"][0]["constant"] == pytest.approx(0.3)
    assert cutout["alpha"]["checker"][1]["constant"] == 1
    lamp = m[6]
    assert lamp["reflectance"]["constant"] == [0.5, 0.5, 0.5]
    # material="none": a mirror; the parameters after it are not read
    assert m[7] == {"type": "mirror", "reflectance": {"constant": [1, 1, 1]}, "normal": None, "alpha": None}


def test_shapes_are_placed_by_their_transforms(desc):
    p = desc["primitives"]
    assert len(p) == 2 + 12 + 1 + 1 + 12 + 2 + 2 + 1 and p[32]["material"] == 7
    # floor: Rx(-90) * scale 3 * (scale 2 of the unit quad)
    R = _rot((1, 0, 0), -90)
    for i, tri in enumerate(_quad_tris(QUAD * 6)):
        assert np.allclose(_tri_positions(desc, i), tri @ R.T, atol=2e-6)
        assert np.allclose([c["n"] for c in p[i]["triangle"]], [[0, 1, 0]] * 3, atol=1e-6)
        assert p[i]["material"] == 0
    assert [c["uv"] for c in p[0]["triangle"]] == [[0, 0], [1, 0], [1, 1]]
    assert [c["uv"] for c in p[1]["triangle"]] == [[0, 0], [1, 1], [0, 1]]
    # metal cube: scale 0.4 of the 2-unit box, then translate
    k = 2
    for corners, n in BOX_FACES:
        for tri in _quad_tris(np.asarray(corners) * 0.4 + [-1, 0.4, 0.2]):
            assert np.allclose(_tri_positions(desc, k), tri, atol=1e-6)
            assert np.allclose([c["n"] for c in p[k]["triangle"]], [n] * 3, atol=1e-6)
            assert p[k]["material"] == 1
            k += 1
    sphere = p[14]
    assert sphere["sphere"] == {"p": pytest.approx([0.9, 0.5, 0.1]), "q": [0, 0, 0, 1], "s": [1, 1, 1]}
    assert sphere["radius"] == 0.5 and sphere["material"] == 2
    # the OBJ triangle: Ry(30) then translate, through the Mat4 -> Transform round trip of LoadModel
    obj = np.array([[-0.3, 0, 0], [0.3, 0, 0], [0, 0.6, 0]]) @ _rot((0, 1, 0), 30).T + [0.2, 0.05, 1.2]
    assert np.allclose(_tri_positions(desc, 15), obj, atol=2e-6) and p[15]["material"] == 3
    # the matrix cube: rows of the matrix, applied to the 2-unit box
    M = np.array([[0.25, 0, 0, 0.3], [0, 0.3, 0.1, 0.35], [0, -0.1, 0.3, -0.8], [0, 0, 0, 1]])
    k = 16
    for corners, n in BOX_FACES:
        for tri in _quad_tris(corners):
            want = tri @ M[:3, :3].T + M[:3, 3]
            assert np.allclose(_tri_positions(desc, k), want, atol=2e-6)
            nn = M[:3, :3] @ n
            assert np.allclose([c["n"] for c in p[k]["triangle"]], [nn / np.linalg.norm(nn)] * 3, atol=2e-6)
            k += 1
    # the cutout: a non-uniform scale
    for i, tri in enumerate(_quad_tris(QUAD * [1.2, 0.8, 0] + [-0.2, 0.6, -1.2])):
        assert np.allclose(_tri_positions(desc, 28 + i), tri, atol=1e-6)
    # the lamp faces down: Rx(90) * scale 0.5 * 2, lifted to y = 2.5
    for i, tri in enumerate(_quad_tris(QUAD @ _rot((1, 0, 0), 90).T + [0, 2.5, 0])):
        assert np.allclose(_tri_positions(desc, 30 + i), tri, atol=1e-6)
        assert np.allclose([c["n"] for c in p[30 + i]["triangle"]], [[0, -1, 0]] * 3, atol=1e-6)


def test_lights(desc):
    L = desc["lights"]
    assert [list(x)[0] for x in L] == ["area", "area", "spot", "point", "directional", "infinite"]
    assert [x["area"]["primitive"] for x in L[:2]] == [30, 31]
    assert all(x["area"]["emission"]["constant"] == [10, 9, 8] and not x["area"]["two_sided"] for x in L[:2])
    spot = L[2]["spot"]
    d = -np.array([2.0, 3.0, 2.0]) / np.linalg.norm([2, 3, 2])
    assert spot["position"] == [2, 3, 2] and np.allclose(spot["direction"], d, atol=1e-6)
    # beam_width defaults to 3/4 of the default cutoff (20), not of the given one
    assert spot["angle_max"] == 30 and spot["angle_falloff_start"] == 15 and spot["intensity"] == [4, 4, 4]
    # the point's position moved by to_world
    assert L[3]["point"] == {"position": [0.75, 1.5, 1], "intensity": [2, 1.5, 1]}
    # the direction rotated by to_world's rotation
    assert np.allclose(L[4]["directional"]["direction"], _rot((1, 0, 0), -60) @ [0, 0, 1], atol=1e-6)
    assert L[4]["directional"]["intensity"] == pytest.approx([0.5, 0.4, 0.3])
    assert L[5]["infinite"] == {"l": pytest.approx([0.1, 0.1, 0.1]), "scale": 1}


def test_transform_math_against_rotation_matrices(tmp_path):
    """rotate / scale / translate chains and a look-at orientation, against
    float64 rotation matrices: catches a swapped quaternion product, a wrong
    half angle, or composition applied on the wrong side."""
    chain = ('<transform name="to_world"><scale x="0.5" y="2" z="1.5"/><rotate x="1" y="0" z="0" angle="35"/>'
             '<rotate x="0" y="0" z="1" angle="-70"/><translate x="0.3" y="-0.2" z="0.9"/>'
             '<rotate y="1" angle="120"/></transform>')
    xml = f"""<scene><shape type="cube">{chain}<bsdf type="diffuse"/></shape>
    <emitter type="spot"><transform name="to_world"><lookat origin="1 -2 3" target="-0.5 0.25 1" up="0.2 1 0.1"/></transform></emitter>
    </scene>"""
    d = describe(_write_files(tmp_path, xml))
    A = _rot((0, 1, 0), 120) @ _rot((0, 0, 1), -70) @ _rot((1, 0, 0), 35)
    S = np.diag([0.5, 2, 1.5])
    t = _rot((0, 1, 0), 120) @ np.array([0.3, -0.2, 0.9])
    for k, (corners, n) in enumerate(BOX_FACES):
        for j, tri in enumerate(_quad_tris(corners)):
            want = tri @ (A @ S).T + t
            assert np.allclose(_tri_positions(d, 2 * k + j), want, atol=3e-6)
    spot = d["lights"][0]["spot"]
    w = np.array([-1.5, 2.25, -2.0])
    assert np.allclose(spot["direction"], w / np.linalg.norm(w), atol=1e-6)
    # a diffuse BSDF with no reflectance is the reference's magenta default
    assert d["materials"][0]["reflectance"]["constant"] == [1, 0, 1]


def test_filters_and_defaults(tmp_path):
    def film(rfilter):
        xml = f"""<scene><sensor type="perspective"><film type="hdrfilm">{rfilter}</film></sensor>
        <shape type="sphere"><point name="center" x="0" y="0" z="0"/><bsdf type="diffuse"/></shape></scene>"""
        return describe(_write_files(tmp_path, xml))["camera"]

    c = film("")
    assert c["filter"] == {"type": "gaussian", "extent": 3, "stddev": 0.5}
    assert c["resolution"] == [1280, 720] and c["spp"] == 64
    assert c["fov"] == pytest.approx(np.degrees(2 * np.arctan(np.tan(np.radians(17.5)) * 720 / 1280)), abs=2e-5)
    assert film('<rfilter type="box"/>')["filter"]["extent"] == 1
    assert film('<rfilter type="tent"><float name="radius" value="0.75"/></rfilter>')["filter"] == {
        "type": "tent", "extent": 1.5, "stddev": 0.5}
    assert film('<rfilter type="tent"/>')["filter"]["extent"] == 2


@pytest.mark.parametrize("xml, what", [
    ('<emitter type="envmap"><string name="filename" value="sky.hdr"/></emitter>', "envmap emitter"),
    ('<bsdf type="principled" id="p"/>', "principled BSDF"),
    ('<bsdf type="plastic" id="p"/>', "plastic BSDF"),
    ('<medium type="homogeneous" id="fog"/>', "participating media"),
    ('<integrator type="bdpt"/>', "bdpt integrator"),
    ('<sensor type="perspective"><sampler type="stratified"/></sensor>', "stratified sampler"),
    ('<shape type="gltf"><string name="filename" value="m.gltf"/><boolean name="use_model_material" '
     'value="true"/></shape>', "glTF model materials"),
])
def test_features_outside_the_path_fail_loudly(tmp_path, xml, what):
    rc, err = load_fails(tmp_path, f"<scene>{xml}</scene>")
    assert rc == 1 and what in err and "not supported on the GPU path" in err, err


@pytest.mark.parametrize("xml, what", [
    ("<scene><shape type='sphere'></scene>", "XML line 1"),
    ("<scene><shape type=sphere/></scene>", "expected a quoted value"),
    ("<scene><bsdf type='diffuse'><rgb name='reflectance' value='0.5 x 1'/></bsdf></scene>", "not a number list"),
    ("<scene><shape type='sphere'><ref id='nope'/></shape></scene>", "Material not found by id: nope"),
    ("<nothing/>", "no <scene> element"),
])
def test_malformed_scenes_are_reported(tmp_path, xml, what):
    rc, err = load_fails(tmp_path, xml)
    assert rc == 1 and what in err, err


def test_elements_the_reference_skips_are_skipped(tmp_path):
    xml = """<scene><shape type="disk"/><emitter type="nope"/><bsdf type="nope" id="x"/>
    <shape type="sphere"><point name="center" x="0" y="0" z="0"/><bsdf type="diffuse"><float name="weird" value="1"/></bsdf></shape>
    </scene>"""
    path = _write_files(tmp_path, xml)
    r = subprocess.run([_exe(build.CPP_DESCRIBE_SCENE), str(path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    for msg in ("Shape type not supported: disk", "Light not supported: nope", "BSDF not supported: nope",
                "Unsupported parameter for diffuse material: weird"):
        assert msg in r.stderr
    assert len(json.loads(r.stdout)["primitives"]) == 1


# ------------------------------------------------------------------ twin


def _texture(sc, t, channels):
    if "constant" in t:
        return sc.add_texture(scenes.constant_texture(t["constant"]))
    if "checker" in t:
        a, b = (_texture(sc, x, channels) for x in t["checker"])
        return sc.add_texture(scenes.checker_texture(a, b, t["resolution"], channels))
    w, h = t["image"]
    texels = np.array(t["texels"], F).reshape((h, w) if channels == 1 else (h, w, 3))
    return sc.add_texture(scenes.image_texture(texels))


def _param(sc, t, fallback, channels):
    """(constant value, 0) or (fallback, slot): the exporter's rule for a material parameter"""
    if t is None:
        return fallback, 0
    if "constant" in t:
        return t["constant"], 0
    return fallback, _texture(sc, t, channels)


def scene_from_description(d):
    """SceneData + camera + render parameters of a describe_scene result"""
    sc = scenes.SceneData()
    mats = []
    for m in d["materials"]:
        b = capi.BbMaterial()
        kind = m["type"]
        if kind == "diffuse":
            b.type = capi.BB_MAT_DIFFUSE
            c, b.tex_color = _param(sc, m["reflectance"], [0, 0, 0], 3)
            b.color[:] = c
            r, b.tex_u_roughness = _param(sc, m["roughness"], 0.0, 1)
            b.u_roughness = b.v_roughness = r
        elif kind == "dielectric":
            b.type = capi.BB_MAT_DIELECTRIC
            b.energy_compensation = int(m["energy_compensation"])
            b.eta[:] = [m["eta"]] * 3
            c, b.tex_color = _param(sc, m["reflectance"], [1, 1, 1], 3)
            b.color[:] = c
            b.u_roughness, b.tex_u_roughness = _param(sc, m["u_roughness"], 0.0, 1)
            b.v_roughness, b.tex_v_roughness = _param(sc, m["v_roughness"], 0.0, 1)
        elif kind == "mirror":
            b.type = capi.BB_MAT_MIRROR
            c, b.tex_color = _param(sc, m["reflectance"], [1, 1, 1], 3)
            b.color[:] = c
        elif kind == "conductor":
            b.type = capi.BB_MAT_CONDUCTOR
            b.energy_compensation = int(m["energy_compensation"])
            b.eta[:] = m["eta"]["constant"]
            b.k[:] = m["k"]["constant"]
            c, b.tex_color = _param(sc, m["reflectance"], [1, 1, 1], 3)
            b.color[:] = c
            b.u_roughness, b.tex_u_roughness = _param(sc, m["u_roughness"], 0.0, 1)
            b.v_roughness, b.tex_v_roughness = _param(sc, m["v_roughness"], 0.0, 1)
        else:
            raise AssertionError(kind)
        if m.get("normal"):
            b.tex_normal = _texture(sc, m["normal"], 3)
        if m.get("alpha"):
            b.tex_alpha = _texture(sc, m["alpha"], 1)
        mats.append(sc.add_material(b))
    for p in d["primitives"]:
        if "sphere" in p:
            i = sc.add_sphere(p["sphere"]["p"], p["radius"], mats[p["material"]])
            s = sc.spheres[-1]
            s.q[:] = p["sphere"]["q"]
            s.s[:] = p["sphere"]["s"]
        else:
            c = p["triangle"]
            v0 = len(sc.positions)
            sc.add_mesh([x["p"] for x in c], [x["n"] for x in c], [x["t"] for x in c], [[0, 1, 2]],
                        mats[p["material"]], uvs=[x["uv"] for x in c])
            # the described normals and tangents are already the Mesh's (normalized once)
            sc.normals[v0:] = np.array([x["n"] for x in c], F)
            sc.tangents[v0:] = np.array([x["t"] for x in c], F)
    for L in d["lights"]:
        (kind, v), = L.items()
        if kind == "area":
            sc.add_area_light(v["primitive"], v["emission"]["constant"], v["two_sided"])
        elif kind == "directional":
            sc.add_directional_light(v["direction"], v["intensity"])
        elif kind == "point":
            sc.add_point_light(v["position"], v["intensity"])
        elif kind == "spot":
            sc.add_spot_light(v["position"], v["direction"], v["intensity"], v["angle_max"], v["angle_falloff_start"])
        else:
            sc.add_infinite_light(v["l"], v["scale"])
    ci = d["camera"]
    cam = capi.BbCameraDesc()
    cam.position[:] = ci["transform"]["p"]
    cam.orientation[:] = ci["transform"]["q"]
    cam.vfov_deg = ci["fov"]
    cam.aperture_radius = ci["aperture_radius"]
    cam.focus_distance = ci["focus_distance"]
    cam.width, cam.height = ci["resolution"]
    f = ci["filter"]
    scenes.set_filter(cam, f["type"], f["extent"], f["stddev"])
    params = scenes.render_params(ci["spp"], max_bounces=d["integrator"]["max_bounces"],
                                  rr_min_bounces=d["integrator"]["rr_min_bounces"])
    return sc, cam, params


@pytest.mark.gpu
def test_xml_scene_renders_like_its_description(tmp_path):
    """render_scene (LoadScene + the Create factories + Render) equals the
    described scene rendered through the C ABI, bit for bit, and the oracle
    within 1e-3 (alpha cut-outs: 99.5 %, see the textured scene tests)."""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    path = _write_files(tmp_path)
    d = describe(path)
    out = tmp_path / "xml.pfm"
    print("[xml] render_scene 96x64 x 16 spp", flush=True)
    r = subprocess.run([_exe(build.CPP_RENDER_SCENE), str(path), str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    img = read_pfm(out)
    sc, cam, params = scene_from_description(d)
    s, c, _ = GpuScene(sc).render(cam, params)
    ref = s / c[..., None]
    assert np.array_equal(img.view(np.int32), ref.astype(np.float32).view(np.int32))
    os_, oc, _ = OracleScene(sc).render(cam, params, threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    frac = pixel_agreement(img, os_ / oc[..., None])
    print(f"[xml] XML scene vs oracle: {frac:.6f} within 1e-3, mean {img.mean():.4f}", flush=True)
    assert frac >= 0.995 and img.mean() > 0.01


# ------------------------------------------------------------------ glTF

def _gltf_model():
    """A glTF 2.0 model (cli/model_loader.cpp:405-740 reads it): node 0 (a
    matrix, no mesh) with child node 1 (a matrix, mesh 0: a quad with
    tangents of both handedness and 16-bit indices, a triangle without
    tangents and 8-bit indices), node 2 (no matrix, mesh 1: a triangle with
    32-bit indices and a line primitive, which is skipped).  Returns the JSON
    document (buffer 0 without uri) and the binary buffer."""
    f = np.float32
    quad_p = np.array([[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1]], f)
    quad_n = np.array([[0, 1, 0], [0, 1, 0], [0.6, 0.8, 0], [0, 0.8, 0.6]], f)
    quad_uv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], f)
    quad_t = np.array([[1, 0, 0, 1], [1, 0, 0, -1], [0, 0, 1, 1], [1, 0, 0, -1]], f)
    quad_i = np.array([0, 1, 2, 0, 2, 3], np.uint16)
    tri_p = np.array([[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0.25]], f)
    tri_n = np.array([[0, 0, 1], [0, 0.6, 0.8], [0, 0, -1]], f)
    tri_uv = np.array([[0.25, 0.5], [0.75, 0.5], [0.5, 0.125]], f)
    tri_i = np.array([0, 1, 2], np.uint8)
    t2_i = np.array([2, 1, 0], np.uint32)
    blobs, views, accessors = [], [], []

    def add(arr, type_, ct):
        off = sum(len(b) for b in blobs)
        data = arr.tobytes()
        data += b"\0" * (-len(data) % 4)
        blobs.append(data)
        views.append({"buffer": 0, "byteOffset": off, "byteLength": arr.nbytes})
        accessors.append({"bufferView": len(views) - 1, "componentType": ct, "count": len(arr), "type": type_})
        return len(accessors) - 1

    A = {"qp": add(quad_p, "VEC3", 5126), "qn": add(quad_n, "VEC3", 5126), "quv": add(quad_uv, "VEC2", 5126),
         "qt": add(quad_t, "VEC4", 5126), "qi": add(quad_i, "SCALAR", 5123), "tp": add(tri_p, "VEC3", 5126),
         "tn": add(tri_n, "VEC3", 5126), "tuv": add(tri_uv, "VEC2", 5126), "ti": add(tri_i, "SCALAR", 5121),
         "t2i": add(t2_i, "SCALAR", 5125)}
    s0 = [2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, 0, 1, 0, 0, 1]       # scale 2, then translate x 1 (column-major)
    r1 = [0, 0, -1, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0.5, 0, 1]    # rotate 90 degrees about y, translate y 0.5
    doc = {
        "asset": {"version": "2.0"}, "scene": 0, "scenes": [{"nodes": [0, 2]}],
        "nodes": [{"matrix": s0, "children": [1]}, {"matrix": r1, "mesh": 0}, {"mesh": 1, "name": "tést"}],
        "meshes": [
            {"primitives": [
                {"attributes": {"POSITION": A["qp"], "NORMAL": A["qn"], "TEXCOORD_0": A["quv"], "TANGENT": A["qt"]},
                 "indices": A["qi"]},
                {"attributes": {"POSITION": A["tp"], "NORMAL": A["tn"], "TEXCOORD_0": A["tuv"]}, "indices": A["ti"]}]},
            {"primitives": [
                {"attributes": {"POSITION": A["tp"], "NORMAL": A["tn"], "TEXCOORD_0": A["tuv"]}, "indices": A["t2i"]},
                {"attributes": {"POSITION": A["tp"], "NORMAL": A["tn"], "TEXCOORD_0": A["tuv"]}, "indices": A["t2i"],
                 "mode": 1}]}],
        "accessors": accessors, "bufferViews": views,
        "buffers": [{"byteLength": sum(len(b) for b in blobs)}],
    }
    mats = {"s0": np.array(s0, float).reshape(4, 4).T, "r1": np.array(r1, float).reshape(4, 4).T}
    parts = {"quad": (quad_p, quad_n, quad_uv, quad_t, quad_i), "tri": (tri_p, tri_n, tri_uv, None, tri_i),
             "tri2": (tri_p, tri_n, tri_uv, None, t2_i)}
    return doc, b"".join(blobs), mats, parts


def _write_gltf(tmp_path, kind):
    import base64
    import struct
    doc, blob, _, _ = _gltf_model()
    if kind == "gltf":
        doc["buffers"][0]["uri"] = "model%20data.bin"
        (tmp_path / "model data.bin").write_bytes(blob)
        (tmp_path / "m.gltf").write_text(json.dumps(doc))
        return "m.gltf"
    if kind == "datauri":
        doc["buffers"][0]["uri"] = "data:application/octet-stream;base64," + base64.b64encode(blob).decode()
        (tmp_path / "m.gltf").write_text(json.dumps(doc))
        return "m.gltf"
    js = json.dumps(doc).encode()
    js += b" " * (-len(js) % 4)
    body = struct.pack("<II", len(js), 0x4E4F534A) + js + struct.pack("<II", len(blob), 0x004E4942) + blob
    (tmp_path / "m.glb").write_bytes(struct.pack("<III", 0x46546C67, 2, 12 + len(body)) + body)
    return "m.glb"


def _expected_gltf_triangles(to_world):
    """Mesh (cli: Mesh::Mesh) of each primitive: positions through the
    node's world matrix, normals and tangents through it (w = 0) and
    normalized, texture v negated (flip_tex_coord false); tangents as
    given (w != 1 negates) or CoordinateSystem of the normal"""
    _, _, mats, parts = _gltf_model()
    node1 = to_world @ mats["s0"] @ mats["r1"]
    node2 = to_world
    out = []
    for M, name in ((node1, "quad"), (node1, "tri"), (node2, "tri2")):
        p, n, uv, t, idx = parts[name]
        if t is None:
            t = np.array([list(_coordinate_system_f64(v)) + [1] for v in n], float)
        tv = np.where(t[:, 3:4] == 1, t[:, :3], -t[:, :3])
        for k in range(0, len(idx), 3):
            corners = []
            for v in idx[k:k + 3]:
                pw = M @ np.append(p[v], 1)
                nw = M[:3, :3] @ n[v]
                tw = M[:3, :3] @ tv[v]
                corners.append((pw[:3], nw / np.linalg.norm(nw), tw / np.linalg.norm(tw), [uv[v][0], -uv[v][1]]))
            out.append(corners)
    return out


def _coordinate_system_f64(n):
    sign = np.copysign(1.0, n[2])
    a = -1 / (sign + n[2])
    b = n[0] * n[1] * a
    return np.array([1 + sign * n[0] * n[0] * a, sign * b, -sign * n[0]])


GLTF_XML = """<scene version="3.0.0">
    <shape type="gltf">
        <string name="filename" value="{name}"/>
        <transform name="to_world"><translate y="0.5" z="-0.25"/></transform>
        <bsdf type="diffuse"><rgb name="reflectance" value="0.3 0.4 0.5"/></bsdf>
    </shape>
</scene>
"""


@pytest.mark.parametrize("kind", ["gltf", "datauri", "glb"])
def test_gltf_model_geometry(tmp_path, kind):
    """LoadGLTF through the scene loader (a gltf shape with a bsdf: the
    model is drawn with it): every triangle's corners, normals, tangents and
    texture coordinates against the reference's rules, from an external
    buffer file (with a %-escaped name), a base64 data URI and a .glb"""
    name = _write_gltf(tmp_path, kind)
    (tmp_path / "scene.xml").write_text(GLTF_XML.format(name=name))
    d = describe(tmp_path / "scene.xml")
    prims = d["primitives"]
    T = np.eye(4)
    T[:3, 3] = [0, 0.5, -0.25]
    want = _expected_gltf_triangles(T)
    assert len(prims) == len(want) == 4
    for prim, corners in zip(prims, want):
        assert prim["material"] == 0
        for c, (p, n, t, uv) in zip(prim["triangle"], corners):
            assert np.allclose(c["p"], p, atol=2e-6)
            assert np.allclose(c["n"], n, atol=2e-6)
            assert np.allclose(c["t"], t, atol=2e-6)
            assert np.allclose(c["uv"], uv, atol=1e-7)
    assert len(d["materials"]) == 1 and d["materials"][0]["type"] == "diffuse"


def test_gltf_model_materials_are_refused(tmp_path):
    """A model whose primitives name their own materials (PrincipledMaterial
    in the reference) and no fallback material: reported, nothing loaded"""
    doc, blob, _, _ = _gltf_model()
    doc["meshes"][0]["primitives"][0]["material"] = 0
    doc["materials"] = [{"pbrMetallicRoughness": {"baseColorFactor": [1, 1, 1, 1]}}]
    doc["buffers"][0]["uri"] = "b.bin"
    (tmp_path / "b.bin").write_bytes(blob)
    (tmp_path / "m.gltf").write_text(json.dumps(doc))
    rc, err = load_fails(tmp_path, '<scene><shape type="gltf"><string name="filename" value="m.gltf"/>'
                                   '<boolean name="use_model_material" value="true"/></shape></scene>')
    assert rc == 1 and "glTF model materials" in err, err


GLTF_RENDER_XML = """<scene version="3.0.0">
    <integrator type="path"><integer name="max_depth" value="5"/></integrator>
    <sensor type="perspective">
        <float name="fov" value="40"/>
        <transform name="to_world"><lookat origin="2, 3.5, 6" target="1, 0.5, 0" up="0, 1, 0"/></transform>
        <sampler type="independent"><integer name="sample_count" value="8"/></sampler>
        <film type="hdrfilm"><integer name="width" value="96"/><integer name="height" value="64"/>
            <rfilter type="box"/></film>
    </sensor>
    <shape type="gltf">
        <string name="filename" value="m.glb"/>
        <transform name="to_world"><translate y="0.5" z="-0.25"/></transform>
        <bsdf type="twosided"><bsdf type="diffuse"><rgb name="reflectance" value="0.7 0.5 0.3"/></bsdf></bsdf>
    </shape>
    <shape type="rectangle">
        <transform name="to_world"><scale value="1.5"/><rotate x="1" angle="90"/><translate x="1" y="4"/></transform>
        <bsdf type="diffuse"/>
        <emitter type="area"><rgb name="radiance" value="6 6 5"/></emitter>
    </shape>
    <emitter type="constant"><rgb name="radiance" value="0.15"/></emitter>
</scene>
"""


@pytest.mark.gpu
def test_gltf_scene_renders_like_its_description(tmp_path):
    """An XML scene with a .glb model: render_scene equals its description
    rendered through the C ABI bit for bit, and the oracle within 1e-3"""
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    _write_gltf(tmp_path, "glb")
    path = tmp_path / "scene.xml"
    path.write_text(GLTF_RENDER_XML)
    d = describe(path)
    assert len(d["primitives"]) == 4 + 2 and len(d["lights"]) == 3
    out = tmp_path / "gltf.pfm"
    print("[gltf] render_scene 96x64 x 8 spp", flush=True)
    r = subprocess.run([_exe(build.CPP_RENDER_SCENE), str(path), str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    img = read_pfm(out)
    sc, cam, params = scene_from_description(d)
    s, c, _ = GpuScene(sc).render(cam, params)
    ref = s / c[..., None]
    assert np.array_equal(img.view(np.int32), ref.astype(np.float32).view(np.int32))
    os_, oc, _ = OracleScene(sc).render(cam, params, threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    frac = pixel_agreement(img, os_ / oc[..., None])
    print(f"[gltf] glTF scene vs oracle: {frac:.6f} within 1e-3, mean {img.mean():.4f}", flush=True)
    assert frac >= 0.999 and img.mean() > 0.01
