"""ctypes mirror of include/bbgpu.h and the loader of libbbgpu.so.

The structures here have exactly the layout of the C declarations; the
library is loaded from the package directory (built in-tree by
``bulbit_amd.build``).  A missing library raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_float, c_int32, c_size_t, c_uint8, c_uint16, c_uint64, c_int64, c_void_p

import numpy as np

ABI_VERSION = 20  # BBGPU_ABI_VERSION of include/bbgpu.h that these structures mirror

LIB_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(LIB_DIR, "libbbgpu.so")

BB_SHAPE_TRIANGLE = 0
BB_SHAPE_SPHERE = 1

BB_MAT_DIFFUSE = 0
BB_MAT_DIELECTRIC = 1
BB_MAT_CONDUCTOR = 2
BB_MAT_CONDUCTOR_R = 3
BB_MAT_METALLIC_ROUGHNESS = 4
BB_MAT_MIRROR = 5

BB_FILTER_BOX = 0
BB_FILTER_TENT = 1
BB_FILTER_GAUSSIAN = 2

BB_LIGHT_SPOT = 0
BB_LIGHT_UNIFORM_INFINITE = 1
BB_LIGHT_DIFFUSE_AREA = 2
BB_LIGHT_POINT = 3
BB_LIGHT_DIRECTIONAL = 4


class BbPrimitive(ctypes.Structure):
    _fields_ = [("shape_kind", c_int32), ("shape_index", c_int32), ("material", c_int32), ("reserved", c_int32)]


class BbSphere(ctypes.Structure):
    _fields_ = [("p", c_float * 3), ("q", c_float * 4), ("s", c_float * 3), ("radius", c_float)]


class BbMaterial(ctypes.Structure):
    _fields_ = [
        ("type", c_int32),
        ("energy_compensation", c_int32),
        ("color", c_float * 3),
        ("eta", c_float * 3),
        ("k", c_float * 3),
        ("u_roughness", c_float),
        ("v_roughness", c_float),
        ("metallic", c_float),
        ("tex_color", c_int32),
        ("tex_u_roughness", c_int32),
        ("tex_v_roughness", c_int32),
        ("tex_metallic", c_int32),
        ("tex_normal", c_int32),
        ("tex_alpha", c_int32),
    ]


BB_TEX_CONSTANT, BB_TEX_IMAGE, BB_TEX_CHECKER = 0, 1, 2
BB_TEXCOORD_REPEAT, BB_TEXCOORD_CLAMP = 0, 1


class BbTexture(ctypes.Structure):
    _fields_ = [
        ("type", c_int32),
        ("channels", c_int32),
        ("width", c_int32),
        ("height", c_int32),
        ("filter", c_int32),
        ("checker", c_int32 * 2),
        ("resolution", c_float * 2),
        ("value", c_float * 3),
        ("texels", POINTER(c_float)),
    ]


class BbLight(ctypes.Structure):
    _fields_ = [
        ("type", c_int32),
        ("primitive", c_int32),
        ("two_sided", c_int32),
        ("reserved", c_int32),
        ("position", c_float * 3),
        ("direction", c_float * 3),
        ("spectrum", c_float * 3),
        ("scale", c_float),
        ("angle_max", c_float),
        ("angle_falloff_start", c_float),
    ]


class BbSceneDesc(ctypes.Structure):
    _fields_ = [
        ("n_vertices", c_int32),
        ("positions", POINTER(c_float)),
        ("normals", POINTER(c_float)),
        ("tangents", POINTER(c_float)),
        ("uvs", POINTER(c_float)),
        ("n_triangles", c_int32),
        ("indices", POINTER(c_int32)),
        ("n_spheres", c_int32),
        ("spheres", POINTER(BbSphere)),
        ("n_primitives", c_int32),
        ("primitives", POINTER(BbPrimitive)),
        ("n_materials", c_int32),
        ("materials", POINTER(BbMaterial)),
        ("n_lights", c_int32),
        ("lights", POINTER(BbLight)),
        ("n_textures", c_int32),
        ("textures", POINTER(BbTexture)),
    ]


class BbBvhNode(ctypes.Structure):
    _fields_ = [
        ("bmin", c_float * 3),
        ("bmax", c_float * 3),
        ("offset", c_int32),
        ("count", c_uint16),
        ("axis", c_uint8),
        ("pad", c_uint8),
    ]


BVH_NODE_DTYPE = np.dtype(
    [("bmin", "<f4", 3), ("bmax", "<f4", 3), ("offset", "<i4"), ("count", "<u2"), ("axis", "u1"), ("pad", "u1")]
)
assert BVH_NODE_DTYPE.itemsize == 32


class BbRay(ctypes.Structure):
    _fields_ = [("o", c_float * 3), ("t_min", c_float), ("d", c_float * 3), ("t_max", c_float)]


class BbHit(ctypes.Structure):
    _fields_ = [("prim", c_int32), ("t", c_float), ("u", c_float), ("v", c_float)]


class BbSceneInfo(ctypes.Structure):
    _fields_ = [
        ("n_nodes", c_int32),
        ("n_primitives", c_int32),
        ("n_lights", c_int32),
        ("n_infinite_lights", c_int32),
        ("world_min", c_float * 3),
        ("world_max", c_float * 3),
        ("build_seconds", c_double),
        ("device_bytes", c_size_t),
    ]


class BbTraceCounters(ctypes.Structure):
    _fields_ = [("rays", c_uint64), ("nodes", c_uint64), ("prims", c_uint64), ("accepts", c_uint64)]

    def as_dict(self):
        return {"rays": self.rays, "nodes": self.nodes, "prims": self.prims, "accepts": self.accepts}


class BbCameraDesc(ctypes.Structure):
    _fields_ = [
        ("position", c_float * 3),
        ("orientation", c_float * 4),
        ("vfov_deg", c_float),
        ("aperture_radius", c_float),
        ("focus_distance", c_float),
        ("width", c_int32),
        ("height", c_int32),
        ("filter_extent", c_float),
        ("filter_type", c_int32),
        ("filter_sigma", c_float),
    ]


class BbRenderParams(ctypes.Structure):
    _fields_ = [
        ("spp", c_int32),
        ("seed", c_int32),
        ("max_bounces", c_int32),
        ("rr_min_bounces", c_int32),
        ("regularize_bsdf", c_int32),
        ("wavefront_paths", c_int32),
        ("trace_lds_stack", c_int32),
        ("sample_begin", c_int32),
        ("trace_refill", c_int32),
        ("shadow_refill", c_int32),
        ("graph", c_int32),
        ("trace_block", c_int32),
        ("trace_stack", c_int32),
        ("light_sampler", c_int32),
    ]


class BbRenderStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", c_double),
        ("ms_trace_closest", c_double),
        ("ms_trace_shadow", c_double),
        ("ms_shade", c_double),
        ("launches_closest", c_int64),
        ("launches_shadow", c_int64),
        ("rays_closest", c_int64),
        ("rays_shadow", c_int64),
        ("paths", c_int64),
        ("closest", BbTraceCounters),
        ("shadow", BbTraceCounters),
    ]


class BbRenderJob(ctypes.Structure):
    _fields_ = [
        ("scene", c_void_p),
        ("camera", BbCameraDesc),
        ("params", BbRenderParams),
        ("film_sum_dev", c_void_p),
        ("film_count_dev", c_void_p),
    ]


class BbBatchStats(ctypes.Structure):
    _fields_ = [("ms_total", c_double), ("samples", c_int64), ("streams", c_int32), ("job_ms", POINTER(c_double))]


class BbCacheParams(ctypes.Structure):
    _fields_ = [
        ("query_bounce", c_int32),
        ("train_stride", c_int32),
        ("train_extra_bounces", c_int32),
        ("train_batch", c_int32),
        ("train_steps", c_int32),
        ("query_primary", c_int32),
        ("query_first", c_int32),
        ("query_direct", c_int32),
    ]


class BbCacheStats(ctypes.Structure):
    _fields_ = [
        ("ms_total", c_double),
        ("ms_query", c_double),
        ("ms_train", c_double),
        ("queries", c_int64),
        ("train_rows", c_int64),
        ("loss", c_float),
        ("diverged", c_int32),
        ("passes_redone", c_int32),
        ("ms_rows", c_double),
    ]


RAY_DTYPE = np.dtype([("o", "<f4", 3), ("t_min", "<f4"), ("d", "<f4", 3), ("t_max", "<f4")])
HIT_DTYPE = np.dtype([("prim", "<i4"), ("t", "<f4"), ("u", "<f4"), ("v", "<f4")])

# entry points declared in include/bbgpu.h (checked by tests/test_capi_exports.py)
EXPORTS = [
    "bbgpu_abi_version",
    "bbgpu_last_error",
    "bbgpu_device_count",
    "bbgpu_alias_table",
    "bbgpu_build_bvh",
    "bbgpu_reflectance_table",
    "bbgpu_scene_create",
    "bbgpu_scene_destroy",
    "bbgpu_scene_info",
    "bbgpu_trace",
    "bbgpu_trace_noalloc",
    "bbgpu_render",
    "bbgpu_look_at",
    "bbgpu_nrc_param_count",
    "bbgpu_nrc_create",
    "bbgpu_nrc_destroy",
    "bbgpu_nrc_get_weights",
    "bbgpu_nrc_set_weights",
    "bbgpu_nrc_query",
    "bbgpu_nrc_train",
    "bbgpu_nrc_diverged",
    "bbgpu_render_cached",
    "bbgpu_render_batch",
    "bbgpu_autotune",
]


class BbError(RuntimeError):
    pass


_lib = None


def load_library(path: str | None = None):
    """Loads libbbgpu.so (once) and declares the signatures."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise BbError(f"{p} is missing: build it with bulbit_amd.build.build() (no CPU fallback exists)")
    # torch ships its own libamdhip64.so / libhsa-runtime64.so under the same
    # sonames as /opt/rocm's.  Loading torch first makes libbbgpu.so bind to
    # the runtime torch already holds, so one process never carries two HIP
    # runtimes (a second one sees no device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(p)
    lib.bbgpu_abi_version.restype = c_int32
    lib.bbgpu_last_error.argtypes = [c_char_p, c_size_t]
    lib.bbgpu_device_count.argtypes = [POINTER(c_int32)]
    lib.bbgpu_alias_table.argtypes = [c_void_p, c_int32, c_void_p, c_void_p]
    lib.bbgpu_build_bvh.argtypes = [POINTER(BbSceneDesc), c_void_p, c_void_p, POINTER(c_int32)]
    lib.bbgpu_reflectance_table.argtypes = [c_int32, c_void_p, c_int32, POINTER(c_int32)]
    lib.bbgpu_scene_create.argtypes = [POINTER(BbSceneDesc), POINTER(c_void_p)]
    lib.bbgpu_scene_destroy.argtypes = [c_void_p]
    lib.bbgpu_scene_info.argtypes = [c_void_p, POINTER(BbSceneInfo)]
    lib.bbgpu_trace.argtypes = [c_void_p, c_void_p, c_int32, c_int32, c_void_p, POINTER(BbTraceCounters), c_void_p]
    lib.bbgpu_trace_noalloc.argtypes = lib.bbgpu_trace.argtypes
    lib.bbgpu_render.argtypes = [
        c_void_p,
        POINTER(BbCameraDesc),
        POINTER(BbRenderParams),
        c_void_p,
        c_void_p,
        c_void_p,
        POINTER(BbRenderStats),
        c_int32,
    ]
    lib.bbgpu_look_at.argtypes = [POINTER(c_float), POINTER(c_float), POINTER(c_float), POINTER(c_float)]
    lib.bbgpu_render_batch.argtypes = [POINTER(BbRenderJob), c_int32, c_int32, POINTER(BbBatchStats)]
    lib.bbgpu_autotune.argtypes = [c_void_p, POINTER(BbCameraDesc), POINTER(BbRenderParams), c_int32,
                                   POINTER(BbRenderParams)]
    lib.bbgpu_render_cached.argtypes = [c_void_p, c_void_p, POINTER(BbCameraDesc), POINTER(BbRenderParams),
                                        POINTER(BbCacheParams), c_void_p, c_void_p, c_void_p, POINTER(BbCacheStats)]
    for name in EXPORTS:
        getattr(lib, name).restype = c_int32
    if lib.bbgpu_abi_version() != ABI_VERSION:
        raise BbError(f"{p} has ABI version {lib.bbgpu_abi_version()}, expected {ABI_VERSION}: rebuild it")
    if path is None:
        _lib = lib
    return lib


TABLE_GGX_E, TABLE_GGX_E_AVG = 0, 1
TABLE_DIELECTRIC_E, TABLE_DIELECTRIC_E_INV, TABLE_DIELECTRIC_E_AVG, TABLE_DIELECTRIC_E_INV_AVG = 2, 3, 4, 5


def reflectance_table(table: int, lib=None):
    """bbgpu_reflectance_table as a float32 numpy array (host computation)."""
    import numpy as np

    lib = lib or load_library()
    n = c_int32(0)
    check(lib.bbgpu_reflectance_table(int(table), None, 0, ctypes.byref(n)), "bbgpu_reflectance_table", lib)
    out = np.zeros(n.value, np.float32)
    check(lib.bbgpu_reflectance_table(int(table), out.ctypes.data, n.value, ctypes.byref(n)),
          "bbgpu_reflectance_table", lib)
    return out


def last_error(lib=None) -> str:
    lib = lib or load_library()
    buf = ctypes.create_string_buffer(1024)
    lib.bbgpu_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(status: int, what: str, lib=None):
    if status != 0:
        raise BbError(f"{what} failed with status {status}: {last_error(lib)}")
