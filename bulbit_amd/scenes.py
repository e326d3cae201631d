"""Scene descriptions and the procedural workloads of the benchmark.

``SceneData`` holds a scene in the layout of ``BbSceneDesc`` (include/bbgpu.h):
one vertex pool, one triangle index list, spheres, primitives in
``Scene::GetPrimitives()`` order, materials and lights in ``Scene::GetLights()``
order.  Host-side float32 preprocessing that the reference performs before
rendering (the Mesh constructor's normalisation, ``Normalize`` of a light
direction, ``Dist`` for a focus distance) is done here in float32 with the
reference's operation order.

Workloads:
  * ``readme_scene``  — the README library example (README.md:140-190):
    one diffuse sphere, a spot light and a uniform environment light.
  * ``icosphere_scene`` — a displaced icosphere (level 8 = 1,310,720
    triangles) on a 2-triangle ground plane under 64 emissive triangles,
    40/30/30 Lambertian/conductor/dielectric triangle groups.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import capi

F32 = np.float32
EPS32 = np.finfo(np.float32).eps

# gold, as in the reference's Cornell-box sample (cli/scene/cornell_box.cpp:19)
GOLD_ETA = (0.161, 0.492, 1.426)
GOLD_K = (4.08769, 2.32625, 1.846)


def normalize_member32(v: np.ndarray) -> np.ndarray:
    """Vector3::Normalize (include/bulbit/vectors.h:121-133) in float32:
    multiply by 1/length unless length < epsilon."""
    v = np.asarray(v, dtype=F32)
    l2 = v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1]
    l2 = l2 + v[:, 2] * v[:, 2]
    length = np.sqrt(l2)
    ok = length >= EPS32
    out = v.copy()
    with np.errstate(divide="ignore"):
        inv = F32(1) / length
    out[ok] = v[ok] * inv[ok, None]
    return out


def normalize32(v) -> np.ndarray:
    """Free Normalize (include/bulbit/math.h:56-60) of one float32 vector."""
    v = np.asarray(v, dtype=F32)
    l2 = v[0] * v[0] + v[1] * v[1]
    l2 = F32(l2 + v[2] * v[2])
    inv = F32(1) / np.sqrt(l2)
    return (v * inv).astype(F32)


def dist32(a, b) -> np.float32:
    """Dist(a, b) = Length(b - a) (include/bulbit/vectors.h:455-459) in float32."""
    d = np.asarray(b, dtype=F32) - np.asarray(a, dtype=F32)
    l2 = d[0] * d[0] + d[1] * d[1]
    l2 = F32(l2 + d[2] * d[2])
    return F32(np.sqrt(l2))


def look_at(position, target, up=(0.0, 1.0, 0.0)):
    """Transform::LookAt orientation, computed by the library's host helper."""
    lib = capi.load_library()
    p = (ctypes.c_float * 3)(*position)
    t = (ctypes.c_float * 3)(*target)
    u = (ctypes.c_float * 3)(*up)
    q = (ctypes.c_float * 4)()
    capi.check(lib.bbgpu_look_at(p, t, u, q), "bbgpu_look_at", lib)
    return tuple(float(x) for x in q)


def make_camera(position, target, vfov, aperture, focus, width, height, filter_extent=1.0, up=(0.0, 1.0, 0.0)):
    q = look_at(position, target, up)
    cam = capi.BbCameraDesc()
    cam.position[:] = [float(F32(x)) for x in position]
    cam.orientation[:] = q
    cam.vfov_deg = vfov
    cam.aperture_radius = aperture
    cam.focus_distance = float(focus)
    cam.width = width
    cam.height = height
    cam.filter_extent = filter_extent
    return cam


def set_filter(cam, kind: str, extent=None, sigma=0.5):
    """The camera's filter: "box" (extent 1), "tent" (extent 2) or
    "gaussian" (sigma 0.5, extent 3), with the reference's defaults
    (include/bulbit/filters.h)."""
    types = {"box": (capi.BB_FILTER_BOX, 1.0), "tent": (capi.BB_FILTER_TENT, 2.0),
             "gaussian": (capi.BB_FILTER_GAUSSIAN, 3.0)}
    t, default_extent = types[kind]
    cam.filter_type = t
    cam.filter_extent = default_extent if extent is None else extent
    cam.filter_sigma = sigma
    return cam


def diffuse(color):
    m = capi.BbMaterial()
    m.type = capi.BB_MAT_DIFFUSE
    m.color[:] = list(color)
    return m


def dielectric(eta, roughness=0.0, color=(1.0, 1.0, 1.0), energy_compensation=0):
    m = capi.BbMaterial()
    m.type = capi.BB_MAT_DIELECTRIC
    m.energy_compensation = energy_compensation
    m.eta[:] = [eta, eta, eta]
    m.color[:] = list(color)
    m.u_roughness = roughness
    m.v_roughness = roughness
    return m


def conductor(eta, k, roughness, color=(1.0, 1.0, 1.0), energy_compensation=0):
    m = capi.BbMaterial()
    m.type = capi.BB_MAT_CONDUCTOR
    m.energy_compensation = energy_compensation
    m.eta[:] = list(eta)
    m.k[:] = list(k)
    m.color[:] = list(color)
    m.u_roughness = roughness
    m.v_roughness = roughness
    return m


def mirror(color=(1.0, 1.0, 1.0)):
    m = capi.BbMaterial()
    m.type = capi.BB_MAT_MIRROR
    m.color[:] = list(color)
    return m


def metallic_roughness(color, metallic, roughness):
    m = capi.BbMaterial()
    m.type = capi.BB_MAT_METALLIC_ROUGHNESS
    m.color[:] = list(color)
    m.metallic = metallic
    m.u_roughness = roughness
    m.v_roughness = roughness
    return m


def constant_texture(value):
    """ConstantTexture: a float (1 channel) or an RGB triple (3 channels)"""
    t = capi.BbTexture()
    t.type = capi.BB_TEX_CONSTANT
    v = np.atleast_1d(np.asarray(value, np.float32))
    t.channels = 1 if v.size == 1 else 3
    t.value[:] = [float(v[0])] * 3 if v.size == 1 else [float(x) for x in v[:3]]
    return t


def image_texture(texels, filter=capi.BB_TEXCOORD_REPEAT):
    """ImageTexture (bilinear) over texels[height][width] (1 channel) or
    [height][width][3]; texel (x, y) is texels[y][x].  The array is kept by
    the returned structure."""
    a = np.ascontiguousarray(texels, np.float32)
    t = capi.BbTexture()
    t.type = capi.BB_TEX_IMAGE
    t.channels = 1 if a.ndim == 2 else a.shape[2]
    t.height, t.width = a.shape[0], a.shape[1]
    t.filter = filter
    t.texels = a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    t._texels = a  # noqa: keeps the array alive with the structure
    return t


def checker_texture(a_slot, b_slot, resolution, channels):
    """CheckerTexture(a, b, resolution); a_slot, b_slot are add_texture values"""
    t = capi.BbTexture()
    t.type = capi.BB_TEX_CHECKER
    t.channels = channels
    t.checker[:] = [a_slot, b_slot]
    t.resolution[:] = list(resolution)
    return t


@dataclass
class SceneData:
    positions: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), F32))
    normals: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), F32))
    tangents: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), F32))
    indices: np.ndarray = field(default_factory=lambda: np.zeros((0, 3), np.int32))
    spheres: list = field(default_factory=list)
    prims: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))
    materials: list = field(default_factory=list)
    lights: list = field(default_factory=list)
    uvs: np.ndarray | None = None  # [n_vertices][2] once a mesh brings texture coordinates
    textures: list = field(default_factory=list)  # BbTexture, texel arrays kept in _texels

    def add_texture(self, tex) -> int:
        """Appends a texture (see constant_texture, image_texture,
        checker_texture); returns its material-slot value, index + 1."""
        self.textures.append(tex)
        return len(self.textures)

    def add_material(self, m) -> int:
        self.materials.append(m)
        return len(self.materials) - 1

    def add_mesh(self, positions, normals, tangents, indices, materials, uvs=None) -> np.ndarray:
        """Mesh + one primitive per triangle (cli/primitive_builder.cpp:70-101).
        Normals and tangents are normalised as the Mesh constructor does with
        an identity transform (src/geometry/mesh.cpp:8-39).  uvs: texture
        coordinates per vertex (zeros when absent).  Returns the new
        primitive indices."""
        nv_before = len(self.positions)
        n_new = len(positions)
        if uvs is not None or self.uvs is not None:
            old = self.uvs if self.uvs is not None else np.zeros((nv_before, 2), F32)
            new = np.asarray(uvs, F32).reshape(-1, 2) if uvs is not None else np.zeros((n_new, 2), F32)
            self.uvs = np.concatenate([old, new])
        positions = np.asarray(positions, F32) + F32(0)
        normals = normalize_member32(np.asarray(normals, F32) + F32(0))
        tangents = normalize_member32(np.asarray(tangents, F32) + F32(0))
        indices = np.asarray(indices, np.int32).reshape(-1, 3)
        vbase = len(self.positions)
        tbase = len(self.indices)
        self.positions = np.concatenate([self.positions, positions])
        self.normals = np.concatenate([self.normals, normals])
        self.tangents = np.concatenate([self.tangents, tangents])
        self.indices = np.concatenate([self.indices, indices + vbase])
        n = len(indices)
        mats = np.broadcast_to(np.asarray(materials, np.int32), (n,))
        p = np.zeros((n, 4), np.int32)
        p[:, 0] = capi.BB_SHAPE_TRIANGLE
        p[:, 1] = tbase + np.arange(n, dtype=np.int32)
        p[:, 2] = mats
        pbase = len(self.prims)
        self.prims = np.concatenate([self.prims, p])
        return pbase + np.arange(n, dtype=np.int32)

    def add_sphere(self, center, radius, material) -> int:
        s = capi.BbSphere()
        s.p[:] = list(center)
        s.q[:] = [0.0, 0.0, 0.0, 1.0]
        s.s[:] = [1.0, 1.0, 1.0]
        s.radius = radius
        self.spheres.append(s)
        p = np.array([[capi.BB_SHAPE_SPHERE, len(self.spheres) - 1, material, 0]], np.int32)
        self.prims = np.concatenate([self.prims, p])
        return len(self.prims) - 1

    def add_spot_light(self, position, direction, intensity, angle_max, angle_falloff_start):
        L = capi.BbLight()
        L.type = capi.BB_LIGHT_SPOT
        L.position[:] = list(position)
        L.direction[:] = [float(x) for x in direction]
        L.spectrum[:] = list(intensity)
        L.angle_max = angle_max
        L.angle_falloff_start = angle_falloff_start
        self.lights.append(L)

    def add_point_light(self, position, intensity):
        L = capi.BbLight()
        L.type = capi.BB_LIGHT_POINT
        L.position[:] = list(position)
        L.spectrum[:] = list(intensity)
        self.lights.append(L)

    def add_directional_light(self, direction, intensity):
        L = capi.BbLight()
        L.type = capi.BB_LIGHT_DIRECTIONAL
        L.direction[:] = [float(x) for x in direction]
        L.spectrum[:] = list(intensity)
        self.lights.append(L)

    def add_infinite_light(self, l, scale=1.0):
        L = capi.BbLight()
        L.type = capi.BB_LIGHT_UNIFORM_INFINITE
        L.spectrum[:] = list(l)
        L.scale = scale
        self.lights.append(L)

    def add_area_light(self, prim, emission, two_sided=False):
        L = capi.BbLight()
        L.type = capi.BB_LIGHT_DIFFUSE_AREA
        L.primitive = int(prim)
        L.spectrum[:] = list(emission)
        L.two_sided = int(two_sided)
        self.lights.append(L)

    def desc(self):
        """A BbSceneDesc pointing into this object's arrays (keep self alive)."""
        self._keep = []

        def arr(a, ctype):
            a = np.ascontiguousarray(a)
            self._keep.append(a)
            return a.ctypes.data_as(ctypes.POINTER(ctype)) if a.size else None

        d = capi.BbSceneDesc()
        d.n_vertices = len(self.positions)
        d.positions = arr(self.positions.astype(F32), ctypes.c_float)
        d.normals = arr(self.normals.astype(F32), ctypes.c_float)
        d.tangents = arr(self.tangents.astype(F32), ctypes.c_float)
        d.uvs = arr(self.uvs.astype(F32), ctypes.c_float) if self.uvs is not None else None
        d.n_triangles = len(self.indices)
        d.indices = arr(self.indices.astype(np.int32), ctypes.c_int32)
        sph = (capi.BbSphere * max(1, len(self.spheres)))(*self.spheres)
        mats = (capi.BbMaterial * max(1, len(self.materials)))(*self.materials)
        lights = (capi.BbLight * max(1, len(self.lights)))(*self.lights)
        prims = np.ascontiguousarray(self.prims.astype(np.int32))
        self._keep += [sph, mats, lights, prims]
        d.n_spheres = len(self.spheres)
        d.spheres = ctypes.cast(sph, ctypes.POINTER(capi.BbSphere))
        d.n_primitives = len(self.prims)
        d.primitives = prims.ctypes.data_as(ctypes.POINTER(capi.BbPrimitive))
        d.n_materials = len(self.materials)
        d.materials = ctypes.cast(mats, ctypes.POINTER(capi.BbMaterial))
        d.n_lights = len(self.lights)
        d.lights = ctypes.cast(lights, ctypes.POINTER(capi.BbLight))
        texs = (capi.BbTexture * max(1, len(self.textures)))(*self.textures)
        self._keep.append(texs)
        d.n_textures = len(self.textures)
        d.textures = ctypes.cast(texs, ctypes.POINTER(capi.BbTexture))
        return d


# ------------------------------------------------------------------ workloads


def readme_scene(width=640, height=360, with_spot=True, with_env=True):
    """The README example (README.md:140-190) at the given resolution."""
    sc = SceneData()
    mat = sc.add_material(diffuse((0.85, 0.2, 0.2)))
    sphere_pos = (0.0, 1.0, 0.0)
    sc.add_sphere(sphere_pos, 1.0, mat)
    light_pos = (2.5, 4.0, 2.0)
    light_dir = normalize32(np.asarray(light_pos, F32) - np.asarray(sphere_pos, F32))
    if with_spot:
        sc.add_spot_light(light_pos, light_dir, (80.0, 80.0, 80.0), 35.0, 20.0)
    if with_env:
        sc.add_infinite_light((0.1, 0.1, 0.1))
    camera_pos = (3.0, 2.0, 5.0)
    cam = make_camera(camera_pos, sphere_pos, 35.0, 0.01, dist32(camera_pos, sphere_pos), width, height)
    return sc, cam


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [
            [-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
            [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
            [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1],
        ],
        dtype=np.float64,
    )
    f = np.array(
        [
            [0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
            [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
            [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
            [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1],
        ],
        dtype=np.int64,
    )
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(level: int):
    """Unit icosphere; each subdivision replaces face i by faces 4i..4i+3, so
    a block of 4^k consecutive faces is one compact patch."""
    v, f = _icosahedron()
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (len(v) + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (len(v) + 1)
        b = uniq % (len(v) + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1).T + len(v)  # midpoints of edges (01, 12, 20)
        v = np.concatenate([v, mid])
        nf = np.empty((len(f), 4, 3), np.int64)
        nf[:, 0] = np.stack([f[:, 0], m[:, 0], m[:, 2]], 1)
        nf[:, 1] = np.stack([m[:, 0], f[:, 1], m[:, 1]], 1)
        nf[:, 2] = np.stack([m[:, 2], m[:, 1], f[:, 2]], 1)
        nf[:, 3] = np.stack([m[:, 0], m[:, 1], m[:, 2]], 1)
        f = nf.reshape(-1, 3)
    return v, f


def value_noise3(p: np.ndarray, seed: int) -> np.ndarray:
    """Seeded trilinear value noise in [-1, 1] with smoothstep weights."""
    rng = np.random.Generator(np.random.PCG64(seed))
    table = rng.uniform(-1.0, 1.0, 256)
    perm = rng.permutation(256)
    pi = np.floor(p).astype(np.int64)
    fr = p - pi
    w = fr * fr * (3.0 - 2.0 * fr)

    def lat(ix, iy, iz):
        return table[perm[(perm[(perm[ix & 255] + iy) & 255] + iz) & 255]]

    out = 0.0
    for dx in (0, 1):
        wx = w[:, 0] if dx else 1.0 - w[:, 0]
        for dy in (0, 1):
            wy = w[:, 1] if dy else 1.0 - w[:, 1]
            for dz in (0, 1):
                wz = w[:, 2] if dz else 1.0 - w[:, 2]
                out = out + wx * wy * wz * lat(pi[:, 0] + dx, pi[:, 1] + dy, pi[:, 2] + dz)
    return out


def _tangents_for(n: np.ndarray) -> np.ndarray:
    up = np.zeros_like(n)
    up[:, 1] = 1.0
    alt = np.abs(n[:, 1]) > 0.999
    up[alt] = [1.0, 0.0, 0.0]
    t = np.cross(up, n)
    return t / np.linalg.norm(t, axis=1, keepdims=True)


ICOSPHERE_DEFAULTS = dict(
    level=8,
    radius=1.0,
    amplitude=0.05,
    noise_frequency=3.0,
    group_size=1024,
    n_lights=64,
    light_radius=6.0,
    light_size=0.6,
    ground_half_extent=8.0,
    seed=2026,
)


def icosphere_scene(width=1920, height=1080, level=8, seed=2026, n_lights=64):
    """Displaced icosphere on a ground plane under emissive triangles.

    * icosphere of unit radius, centre (0, 1 + amplitude, 0), radial
      displacement r = 1 + 0.05 * noise(3 p) (seeded value noise);
    * smooth vertex normals (area-weighted face normals) and tangents
      orthogonal to them;
    * triangles grouped in blocks of 1024 consecutive faces (compact patches);
      each group is Lambertian (40%, albedo U[0.1, 0.9] per channel),
      conductor (30%, gold, GGX roughness U[0.05, 0.6]) or smooth dielectric
      (30%, IOR 1.5), drawn from a seeded generator;
    * ground: 2 triangles at y = 0 over [-8, 8]^2, Lambertian 0.5;
    * 64 one-sided emissive triangles facing the origin on the upper
      hemisphere of radius 6 (elevations 20-70 degrees, Fibonacci azimuths),
      grey radiance U[5, 20], black Lambertian material;
    * camera at (0, 2.4, 5.2) looking at (0, 0.9, 0), vfov 38, pinhole.
    Primitive order: icosphere, ground, lights.
    """
    cfg = dict(ICOSPHERE_DEFAULTS)
    rng = np.random.Generator(np.random.PCG64(seed))
    amp = cfg["amplitude"]
    v, f = icosphere(level)
    disp = 1.0 + amp * value_noise3(v * cfg["noise_frequency"], seed + 1)
    pos = v * (cfg["radius"] * disp)[:, None]
    pos[:, 1] += 1.0 + amp
    # smooth normals
    fn = np.cross(pos[f[:, 1]] - pos[f[:, 0]], pos[f[:, 2]] - pos[f[:, 0]])
    vn = np.zeros_like(pos)
    for k in range(3):
        np.add.at(vn, f[:, k], fn)
    vn /= np.linalg.norm(vn, axis=1, keepdims=True)
    vt = _tangents_for(vn)

    sc = SceneData()
    n_groups = (len(f) + cfg["group_size"] - 1) // cfg["group_size"]
    kinds = rng.choice(3, size=n_groups, p=[0.4, 0.3, 0.3])
    group_mat = np.empty(n_groups, np.int32)
    for g in range(n_groups):
        if kinds[g] == 0:
            m = diffuse(tuple(rng.uniform(0.1, 0.9, 3)))
        elif kinds[g] == 1:
            m = conductor(GOLD_ETA, GOLD_K, float(rng.uniform(0.05, 0.6)))
        else:
            m = dielectric(1.5)
        group_mat[g] = sc.add_material(m)
    tri_mat = np.repeat(group_mat, cfg["group_size"])[: len(f)]
    sc.add_mesh(pos, vn, vt, f, tri_mat)

    # ground plane
    e = cfg["ground_half_extent"]
    gpos = np.array([[-e, 0, -e], [e, 0, -e], [e, 0, e], [-e, 0, e]], np.float64)
    gn = np.tile([0.0, 1.0, 0.0], (4, 1))
    gt = np.tile([1.0, 0.0, 0.0], (4, 1))
    gmat = sc.add_material(diffuse((0.5, 0.5, 0.5)))
    sc.add_mesh(gpos, gn, gt, np.array([[0, 2, 1], [0, 3, 2]]), gmat)

    # emissive triangles facing the origin
    black = sc.add_material(diffuse((0.0, 0.0, 0.0)))
    R = cfg["light_radius"]
    s = cfg["light_size"]
    golden = np.pi * (3.0 - 5.0 ** 0.5)
    for i in range(n_lights):
        elev = np.radians(20.0 + 50.0 * (i + 0.5) / n_lights)
        az = i * golden
        c = R * np.array([np.cos(elev) * np.cos(az), np.sin(elev), np.cos(elev) * np.sin(az)])
        nrm = -c / np.linalg.norm(c)
        t1 = np.cross(nrm, [0.0, 1.0, 0.0])
        t1 /= np.linalg.norm(t1)
        t2 = np.cross(nrm, t1)
        corners = [c + s * (np.cos(a) * t1 + np.sin(a) * t2) for a in (0.0, 2 * np.pi / 3, 4 * np.pi / 3)]
        tri = np.array(corners)
        # wind so that Cross(e1, e2) points along nrm (towards the origin)
        if np.dot(np.cross(tri[1] - tri[0], tri[2] - tri[0]), nrm) < 0:
            tri = tri[[0, 2, 1]]
        prim = sc.add_mesh(tri, np.tile(nrm, (3, 1)), np.tile(t1, (3, 1)), np.array([[0, 1, 2]]), black)
        radiance = float(rng.uniform(5.0, 20.0))
        sc.add_area_light(prim[0], (radiance, radiance, radiance))

    cam_pos = (0.0, 2.4, 5.2)
    target = (0.0, 0.9, 0.0)
    cam = make_camera(cam_pos, target, 38.0, 0.0, dist32(cam_pos, target), width, height)
    return sc, cam


BATCH_RESOLUTIONS = [(256, 256), (512, 512), (1024, 576), (1024, 1024)]


def batch_workload_specs(n_jobs=32, seed=303):
    """The batched workload: n_jobs independent seeded scenes of one call.
    Each is the displaced-icosphere scene with a triangle count drawn
    log-uniform in [1e4, 2e6] (realized as the nearest subdivision level,
    5..8, of 20 * 4^level faces), a resolution from BATCH_RESOLUTIONS, 16 spp,
    max_bounces uniform in 4..12 and 4 to 256 area lights.  Returns a list of
    dicts: level, width, height, n_lights, max_bounces, scene_seed, seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for i in range(n_jobs):
        tris = float(np.exp(rng.uniform(np.log(1e4), np.log(2e6))))
        level = int(np.clip(np.round(np.log(tris / 20.0) / np.log(4.0)), 5, 8))
        w, h = BATCH_RESOLUTIONS[int(rng.integers(0, 4))]
        n_lights = int(rng.integers(4, 257))
        max_bounces = int(rng.integers(4, 13))
        out.append(dict(level=level, width=w, height=h, n_lights=n_lights, max_bounces=max_bounces,
                        scene_seed=1000 + i, seed=i))
    return out


def batch_workload_job(spec, spp=16):
    """(SceneData, camera, BbRenderParams) of one batch_workload_specs entry"""
    sc, cam = icosphere_scene(width=spec["width"], height=spec["height"], level=spec["level"],
                              seed=spec["scene_seed"], n_lights=spec["n_lights"])
    return sc, cam, render_params(spp, max_bounces=spec["max_bounces"], seed=spec["seed"])


def poisson_disk(n, half_extent, min_dist, seed, k=30):
    """n points in [-half_extent, half_extent]^2, pairwise at least min_dist
    apart: Bridson's dart throwing (annulus candidates around active points,
    grid of cell size min_dist / sqrt(2)), seeded.  Raises if the square is
    full before n points."""
    rng = np.random.Generator(np.random.PCG64(seed))
    size = 2.0 * half_extent
    cell = min_dist / np.sqrt(2.0)
    gn = int(np.ceil(size / cell))
    grid = -np.ones((gn, gn), np.int64)
    pts = []

    def cell_of(p):
        return min(int((p[0] + half_extent) / cell), gn - 1), min(int((p[1] + half_extent) / cell), gn - 1)

    def fits(p):
        if not (-half_extent <= p[0] < half_extent and -half_extent <= p[1] < half_extent):
            return False
        cx, cy = cell_of(p)
        for ix in range(max(cx - 2, 0), min(cx + 3, gn)):
            for iy in range(max(cy - 2, 0), min(cy + 3, gn)):
                j = grid[ix, iy]
                if j >= 0 and (pts[j][0] - p[0]) ** 2 + (pts[j][1] - p[1]) ** 2 < min_dist * min_dist:
                    return False
        return True

    def add(p):
        grid[cell_of(p)] = len(pts)
        pts.append(p)
        active.append(len(pts) - 1)

    active = []
    add(rng.uniform(-half_extent, half_extent, 2))
    while active and len(pts) < n:
        a = int(rng.integers(0, len(active)))
        base = pts[active[a]]
        for _ in range(k):
            r = min_dist * np.sqrt(rng.uniform(1.0, 4.0))
            t = rng.uniform(0.0, 2.0 * np.pi)
            p = base + r * np.array([np.cos(t), np.sin(t)])
            if fits(p):
                add(p)
                break
        else:
            active[a] = active[-1]
            active.pop()
    if len(pts) < n:
        raise ValueError(f"only {len(pts)} points of spacing {min_dist} fit in the square")
    return np.array(pts[:n])


LARGE_DEFAULTS = dict(
    n_spheres=4096,
    level=5,
    half_extent=50.0,
    min_dist=1.2,
    radius=(0.3, 0.55),
    amplitude=0.05,
    noise_frequency=3.0,
    n_lights=1024,
    light_height=(8.0, 14.0),
    light_size=0.8,
    power_sigma=1.5,
    radiance_scale=8.0,
    seed=5150,
)


def large_scene(width=3840, height=2160, n_spheres=4096, level=5, n_lights=1024, seed=5150, batch=64):
    """Many displaced icospheres on a plane under many emissive triangles.

    * n_spheres icospheres of subdivision `level` (4096 at level 5: 83,886,080
      triangles), centres on the plane y = 0 placed by seeded Poisson-disk
      sampling over [-50, 50]^2 with spacing 1.2, radius U[0.3, 0.55], each
      displaced radially by 1 + 0.05 * noise(3 p + offset) (value noise, a
      random offset per sphere) and raised by radius * 1.05;
    * smooth vertex normals (area-weighted face normals), tangents orthogonal
      to them;
    * one metallic-roughness material per sphere: metallic 0 or 1 (even odds),
      roughness U[0.05, 1], base colour U[0.2, 0.9] per channel;
    * ground: 2 triangles at y = 0 over [-55, 55]^2, metallic-roughness
      (grey 0.5, metallic 0, roughness 0.9);
    * n_lights one-sided emissive triangles facing down, centres uniform over
      the plane at heights U[8, 14], circumradius 0.8; power log-normal with
      sigma 1.5 (equal areas, so radiance = 8 exp(N(0, 1.5^2)));
    * camera at (0, 10, 58) looking at (0, 0, 10), vfov 40, pinhole.
    Primitive order: spheres (sphere by sphere), ground, lights.  The
    spheres' vertex arrays are built in batches of `batch` spheres.
    """
    import scipy.sparse

    cfg = dict(LARGE_DEFAULTS)
    rng = np.random.Generator(np.random.PCG64(seed))
    amp = cfg["amplitude"]
    v, f = icosphere(level)
    V, F = len(v), len(f)
    centres = poisson_disk(n_spheres, cfg["half_extent"], cfg["min_dist"], seed + 1)
    radii = rng.uniform(*cfg["radius"], n_spheres)
    offsets = rng.uniform(0.0, 100.0, (n_spheres, 3))
    metallic = rng.integers(0, 2, n_spheres).astype(np.float64)
    rough = rng.uniform(0.05, 1.0, n_spheres)
    colours = rng.uniform(0.2, 0.9, (n_spheres, 3))
    # vertex-face incidence: smooth normal = sum of the adjacent face normals
    inc = scipy.sparse.csr_matrix((np.ones(3 * F), (f.T.ravel(), np.tile(np.arange(F), 3))), shape=(V, F))

    pos = np.empty((n_spheres * V, 3), F32)
    nrm = np.empty((n_spheres * V, 3), F32)
    tan = np.empty((n_spheres * V, 3), F32)

    def cross(a, b):
        return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1], a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                         a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], -1)

    def build(s0):
        s1 = min(s0 + batch, n_spheres)
        S = s1 - s0
        p = (v[None] * cfg["noise_frequency"] + offsets[s0:s1, None, :]).reshape(-1, 3)
        disp = (1.0 + amp * value_noise3(p, seed + 2)).reshape(S, V)
        P = v[None] * (radii[s0:s1, None, None] * disp[..., None])
        P[..., 0] += centres[s0:s1, 0, None]
        P[..., 1] += (radii[s0:s1] * (1.0 + amp))[:, None]
        P[..., 2] += centres[s0:s1, 1, None]
        fn = cross(P[:, f[:, 1]] - P[:, f[:, 0]], P[:, f[:, 2]] - P[:, f[:, 0]])  # (S, F, 3)
        vn = (inc @ fn.transpose(1, 0, 2).reshape(F, 3 * S)).reshape(V, S, 3).transpose(1, 0, 2).reshape(-1, 3)
        vn /= np.linalg.norm(vn, axis=1, keepdims=True)
        pos[s0 * V:s1 * V] = P.reshape(-1, 3)
        nrm[s0 * V:s1 * V] = vn
        tan[s0 * V:s1 * V] = _tangents_for(vn)

    # numpy releases the GIL in the array work: batches run on threads
    import concurrent.futures
    import os

    with concurrent.futures.ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(build, range(0, n_spheres, batch)))
    idx = (f[None].astype(np.int64) + (np.arange(n_spheres, dtype=np.int64) * V)[:, None, None]).reshape(-1, 3)

    sc = SceneData()
    mats = np.array([sc.add_material(metallic_roughness(tuple(colours[i]), float(metallic[i]), float(rough[i])))
                     for i in range(n_spheres)], np.int32)
    sc.add_mesh(pos, nrm, tan, idx, np.repeat(mats, F))
    del pos, nrm, tan, idx

    e = cfg["half_extent"] + 5.0
    gpos = np.array([[-e, 0, -e], [e, 0, -e], [e, 0, e], [-e, 0, e]], np.float64)
    gmat = sc.add_material(metallic_roughness((0.5, 0.5, 0.5), 0.0, 0.9))
    sc.add_mesh(gpos, np.tile([0.0, 1.0, 0.0], (4, 1)), np.tile([1.0, 0.0, 0.0], (4, 1)),
                np.array([[0, 2, 1], [0, 3, 2]]), gmat)

    black = sc.add_material(diffuse((0.0, 0.0, 0.0)))
    lc = np.stack([rng.uniform(-cfg["half_extent"], cfg["half_extent"], n_lights),
                   rng.uniform(*cfg["light_height"], n_lights),
                   rng.uniform(-cfg["half_extent"], cfg["half_extent"], n_lights)], 1)
    ang = rng.uniform(0.0, 2.0 * np.pi, n_lights)
    power = np.exp(rng.normal(0.0, cfg["power_sigma"], n_lights))
    s = cfg["light_size"]
    tri = np.empty((n_lights, 3, 3))
    for c in range(3):
        a = ang + c * 2.0 * np.pi / 3.0
        # wound so that Cross(e1, e2) points down (-y)
        tri[:, c] = lc + s * np.stack([np.cos(a), np.zeros(n_lights), np.sin(a)], 1)
    down = np.tile([0.0, -1.0, 0.0], (3 * n_lights, 1))
    side = np.tile([1.0, 0.0, 0.0], (3 * n_lights, 1))
    prims = sc.add_mesh(tri.reshape(-1, 3), down, side, np.arange(3 * n_lights).reshape(-1, 3), black)
    for i in range(n_lights):
        L = cfg["radiance_scale"] * float(power[i])
        sc.add_area_light(prims[i], (L, L, L))

    cam_pos = (0.0, 10.0, 58.0)
    target = (0.0, 0.0, 10.0)
    cam = make_camera(cam_pos, target, 40.0, 0.0, dist32(cam_pos, target), width, height)
    return sc, cam


LIGHTS_POWER, LIGHTS_ALIAS = 0, 1  # BbRenderParams.light_sampler


def render_params(spp, max_bounces=8, seed=0, rr_min_bounces=1, sample_begin=0, wavefront_paths=0, trace_lds_stack=0,
                  trace_refill=0, shadow_refill=0, graph=0, light_sampler=LIGHTS_POWER, trace_block=0, trace_stack=0):
    p = capi.BbRenderParams()
    p.spp = spp
    p.seed = seed
    p.max_bounces = max_bounces
    p.rr_min_bounces = rr_min_bounces
    p.regularize_bsdf = 0
    p.wavefront_paths = wavefront_paths
    p.trace_lds_stack = trace_lds_stack
    p.trace_refill = trace_refill
    p.shadow_refill = shadow_refill
    p.sample_begin = sample_begin
    p.light_sampler = light_sampler
    p.trace_block = trace_block
    p.trace_stack = trace_stack
    p.graph = graph
    return p


def random_rays(bmin, bmax, n, seed, t_max=np.inf):
    """Seeded random rays: origins uniform in the box scaled by 1.2 about its
    centre, directions by SampleUniformSphere (sampling.h:64-74) on float32
    uniforms, t in [1e-4, t_max]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    bmin = np.asarray(bmin, np.float64)
    bmax = np.asarray(bmax, np.float64)
    c = 0.5 * (bmin + bmax)
    h = 0.6 * (bmax - bmin)
    o = (c + h * rng.uniform(-1.0, 1.0, (n, 3))).astype(F32)
    u = rng.random((n, 2), dtype=np.float32)
    z = F32(1) - F32(2) * u[:, 0]
    r = np.sqrt(np.maximum(F32(0), F32(1) - z * z))
    phi = F32(2 * np.float32(np.pi)) * u[:, 1]
    rays = np.zeros(n, capi.RAY_DTYPE)
    rays["o"] = o
    rays["d"][:, 0] = r * np.cos(phi).astype(F32)
    rays["d"][:, 1] = r * np.sin(phi).astype(F32)
    rays["d"][:, 2] = z
    rays["t_min"] = F32(1e-4)
    if np.isscalar(t_max):
        rays["t_max"] = F32(t_max)
    else:
        rays["t_max"] = np.asarray(t_max, F32)
    return rays


def portable_random_rays(bmin, bmax, n, seed, t_max=np.inf):
    """Seeded random rays built from IEEE-exact operations only (+, -, *, /,
    sqrt; no sin/cos), so every machine produces the same bits: origins
    uniform in the box scaled by 1.2 about its centre, directions uniform on
    the sphere by rejection from the cube [-1, 1]^3, t in [1e-4, t_max]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    bmin = np.asarray(bmin, np.float64)
    bmax = np.asarray(bmax, np.float64)
    c = 0.5 * (bmin + bmax)
    h = 0.6 * (bmax - bmin)
    o = (c + h * rng.uniform(-1.0, 1.0, (n, 3))).astype(F32)
    d = np.empty((n, 3), np.float64)
    filled = 0
    while filled < n:
        m = max(1024, int(1.95 * (n - filled)))
        v = rng.uniform(-1.0, 1.0, (m, 3))
        r2 = v[:, 0] * v[:, 0] + v[:, 1] * v[:, 1] + v[:, 2] * v[:, 2]
        v = v[(r2 > 1e-6) & (r2 <= 1.0)]
        r2 = r2[(r2 > 1e-6) & (r2 <= 1.0)]
        k = min(len(v), n - filled)
        d[filled:filled + k] = v[:k] / np.sqrt(r2[:k])[:, None]
        filled += k
    rays = np.zeros(n, capi.RAY_DTYPE)
    rays["o"] = o
    rays["d"] = d.astype(F32)
    rays["t_min"] = F32(1e-4)
    if np.isscalar(t_max):
        rays["t_max"] = F32(t_max)
    else:
        rays["t_max"] = np.asarray(t_max, F32)
    return rays


def cache_params(query_bounce=2, train_stride=16, train_extra_bounces=4, train_batch=65536, train_steps=1,
                 query_primary=False, query_direct=False, query_first=0):
    """BbCacheParams of the cached path tracer (defaults: query at the first
    rough vertex from vertex 1 on, or at vertex 2; the cache stands in for all
    the vertex reflects unless query_direct; 1/16 training paths extended by
    4 vertices; one Adam step of up to 65,536 rows per pass)."""
    c = capi.BbCacheParams()
    c.query_primary = int(query_primary)
    c.query_direct = int(query_direct)
    c.query_first = int(query_first)
    c.query_bounce = query_bounce
    c.train_stride = train_stride
    c.train_extra_bounces = train_extra_bounces
    c.train_batch = train_batch
    c.train_steps = train_steps
    return c
