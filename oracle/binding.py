"""ctypes binding of liboracle.so (test infrastructure, see oracle.h)."""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_char_p, c_double, c_float, c_int32, c_uint32, c_uint64, c_void_p

import numpy as np

from bulbit_amd import capi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def build_oracle(force: bool = False) -> str:
    """Compiles liboracle.so with the committed Makefile."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True, capture_output=True)
    return LIB


def load_oracle():
    """The committed oracle, liboracle.so (built with the Makefile when missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build_oracle()
    _lib = load_oracle_from(LIB)
    return _lib


def load_oracle_from(path):
    """An oracle library at `path` with its entry points typed (tests load
    builds of oracle.c with a planted mistake this way)."""
    lib = ctypes.CDLL(path)
    lib.orc_murmur64a.restype = c_uint64
    lib.orc_murmur64a.argtypes = [c_void_p, c_int32, c_uint64]
    lib.orc_pixel_sample_hash.restype = c_uint64
    lib.orc_pixel_sample_hash.argtypes = [c_int32, c_int32, c_int32, c_int32]
    lib.orc_pcg32_outputs.argtypes = [c_uint64, c_uint64, c_int32, POINTER(c_uint32)]
    lib.orc_sampler_floats.argtypes = [c_int32, c_int32, c_int32, c_int32, c_int32, POINTER(c_float)]
    lib.orc_sample_concentric.argtypes = [POINTER(c_float), POINTER(c_float)]
    lib.orc_sample_uniform_sphere.argtypes = [POINTER(c_float), POINTER(c_float)]
    lib.orc_triangle_intersect.restype = c_int32
    lib.orc_triangle_intersect.argtypes = [POINTER(c_float)] * 3 + [c_void_p, POINTER(c_float), POINTER(c_float), POINTER(c_float)]
    lib.orc_fresnel_dielectric.restype = c_float
    lib.orc_fresnel_dielectric.argtypes = [c_float, c_float]
    lib.orc_fresnel_complex.restype = c_float
    lib.orc_fresnel_complex.argtypes = [c_float, c_float, c_float]
    lib.orc_look_at.argtypes = [POINTER(c_float)] * 4
    lib.orc_permutation_element.restype = c_int32
    lib.orc_permutation_element.argtypes = [c_uint32, c_uint32, c_uint32]
    lib.orc_reflectance_tables.argtypes = [POINTER(c_float)] * 4
    lib.orc_dielectric_tables.argtypes = [POINTER(c_float)] * 4
    lib.orc_ggx_E.restype = c_float
    lib.orc_ggx_E.argtypes = [c_float, c_float, POINTER(c_float)]
    lib.orc_ggx_E_avg.restype = c_float
    lib.orc_ggx_E_avg.argtypes = [c_float, c_float]
    lib.orc_filter_sample.argtypes = [POINTER(capi.BbCameraDesc), POINTER(c_float), POINTER(c_float)]
    lib.orc_camera_ray.argtypes = [POINTER(capi.BbCameraDesc), c_int32, c_int32, POINTER(c_float), POINTER(c_float), c_void_p]
    lib.orc_bsdf_eval.restype = c_int32
    lib.orc_bsdf_eval.argtypes = [POINTER(capi.BbMaterial), POINTER(c_float), POINTER(c_float), c_int32,
                                  POINTER(c_float), POINTER(c_float), POINTER(c_float), POINTER(c_float)]
    lib.orc_bsdf_sample.restype = c_int32
    lib.orc_bsdf_sample.argtypes = [POINTER(capi.BbMaterial), POINTER(c_float), POINTER(c_float), c_int32,
                                    POINTER(c_float), c_float, POINTER(c_float), POINTER(c_float), POINTER(c_float),
                                    POINTER(c_float), POINTER(c_int32), POINTER(c_float)]
    lib.orc_scene_create.restype = c_void_p
    lib.orc_scene_create.argtypes = [POINTER(capi.BbSceneDesc), c_char_p, c_int32]
    lib.orc_scene_destroy.argtypes = [c_void_p]
    lib.orc_scene_node_count.restype = c_int32
    lib.orc_scene_node_count.argtypes = [c_void_p]
    lib.orc_scene_bvh.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.orc_scene_world_bounds.argtypes = [c_void_p, POINTER(c_float), POINTER(c_float)]
    lib.orc_scene_light_pmf.argtypes = [c_void_p, POINTER(c_float)]
    lib.orc_scene_build_seconds.restype = c_double
    lib.orc_scene_build_seconds.argtypes = [c_void_p]
    lib.orc_trace.argtypes = [c_void_p, c_void_p, c_int32, c_int32, c_void_p, c_int32, POINTER(capi.BbTraceCounters)]
    lib.orc_trace_brute.argtypes = [c_void_p, c_void_p, c_int32, c_void_p]
    lib.orc_texture_eval.argtypes = [c_void_p, c_int32, POINTER(c_float), POINTER(c_float)]
    lib.orc_set_cr_libm.argtypes = [c_int32]
    lib.orc_render.restype = c_double
    lib.orc_render.argtypes = [c_void_p, POINTER(capi.BbCameraDesc), POINTER(capi.BbRenderParams), c_int32, c_int32,
                               c_int32, c_int32, c_void_p, c_void_p, c_int32]
    lib.orc_sample_li.argtypes = [c_void_p, POINTER(capi.BbCameraDesc), POINTER(capi.BbRenderParams), c_int32, c_int32,
                                  c_int32, POINTER(c_float), POINTER(c_int32)]
    lib.orc_light_sample.argtypes = [c_void_p, c_int32, POINTER(c_float), c_int32, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p]
    lib.orc_light_hit_pdf.argtypes = [c_void_p, c_int32, POINTER(c_float), c_int32, c_void_p, c_void_p]
    return lib


def fptr(a):
    return np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(POINTER(c_float))


def correctly_rounded_libm(on=True):
    """Switch the oracle's sampling transcendentals (sin, cos, acos, atan2,
    pow) between glibc's float functions (the reference's, the default) and
    double evaluation rounded to float (the GPU's).  For tests that separate
    the libm difference from everything else."""
    load_oracle().orc_set_cr_libm(1 if on else 0)


class OracleScene:
    """A scene built by the oracle (its own BVH and light set-up)."""

    def __init__(self, scene_data, lib=None):
        self.lib = lib if lib is not None else load_oracle()
        self.data = scene_data
        desc = scene_data.desc()
        err = ctypes.create_string_buffer(512)
        self.h = self.lib.orc_scene_create(ctypes.byref(desc), err, 512)
        if not self.h:
            raise ValueError("oracle scene: " + err.value.decode())
        self.n_prims = len(scene_data.prims)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            self.lib.orc_scene_destroy(h)
            self.h = None

    @property
    def n_nodes(self):
        return self.lib.orc_scene_node_count(self.h)

    @property
    def build_seconds(self):
        return self.lib.orc_scene_build_seconds(self.h)

    def texture(self, index, uv):
        """Texture<T>::Evaluate of texture `index` (0-based) at uv -> 3 floats"""
        out = (c_float * 3)()
        self.lib.orc_texture_eval(self.h, int(index), (c_float * 2)(*uv), out)
        return np.array(out, np.float32)

    def bvh(self):
        nodes = np.zeros(self.n_nodes, capi.BVH_NODE_DTYPE)
        order = np.zeros(self.n_prims, np.int32)
        self.lib.orc_scene_bvh(self.h, nodes.ctypes.data, order.ctypes.data)
        return nodes, order

    def world_bounds(self):
        a = (c_float * 3)()
        b = (c_float * 3)()
        self.lib.orc_scene_world_bounds(self.h, a, b)
        return np.array(a[:], np.float32), np.array(b[:], np.float32)

    def light_pmf(self):
        n = len(self.data.lights)
        out = np.zeros(max(n, 1), np.float32)
        self.lib.orc_scene_light_pmf(self.h, out.ctypes.data_as(POINTER(c_float)))
        return out[:n]

    def trace(self, rays, any_hit=False, threads=None):
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), capi.HIT_DTYPE)
        cnt = capi.BbTraceCounters()
        self.lib.orc_trace(self.h, rays.ctypes.data, len(rays), int(any_hit), hits.ctypes.data,
                           threads or os.cpu_count() or 1, ctypes.byref(cnt))
        return hits, cnt

    def trace_brute(self, rays):
        rays = np.ascontiguousarray(rays)
        hits = np.zeros(len(rays), capi.HIT_DTYPE)
        self.lib.orc_trace_brute(self.h, rays.ctypes.data, len(rays), hits.ctypes.data)
        return hits

    def render(self, camera, params, rect=None, threads=None, film=None):
        """Returns (sum[h, w, 3], count[h, w], seconds)."""
        w, h = camera.width, camera.height
        if film is None:
            s = np.zeros((h, w, 3), np.float32)
            c = np.zeros((h, w), np.int32)
        else:
            s, c = film
        x0, y0, x1, y1 = rect or (0, 0, w, h)
        secs = self.lib.orc_render(self.h, ctypes.byref(camera), ctypes.byref(params), x0, y0, x1, y1,
                                   s.ctypes.data, c.ctypes.data, threads or os.cpu_count() or 1)
        return s, c, secs

    def light_sample(self, light, ref, u):
        """Light::Sample_Li of light `light` at the point ref for the sample
        pairs u[n, 2] -> dict of wi[n, 3], Li[n, 3], pdf[n], visibility[n], ok[n]"""
        u = np.ascontiguousarray(u, np.float32).reshape(-1, 2)
        n = len(u)
        out = {"wi": np.zeros((n, 3), np.float32), "Li": np.zeros((n, 3), np.float32), "pdf": np.zeros(n, np.float32),
               "visibility": np.zeros(n, np.float32), "ok": np.zeros(n, np.int32)}
        self.lib.orc_light_sample(self.h, int(light), (c_float * 3)(*ref), n, u.ctypes.data, out["wi"].ctypes.data,
                                  out["Li"].ctypes.data, out["pdf"].ctypes.data, out["visibility"].ctypes.data,
                                  out["ok"].ctypes.data)
        return out

    def light_hit_pdf(self, light, origin, d):
        """Triangle::PDF(isect, ray) for rays from origin along d[n, 3] whose
        closest hit is area light `light`'s triangle, 0 for the others"""
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        pdf = np.zeros(len(d), np.float32)
        self.lib.orc_light_hit_pdf(self.h, int(light), (c_float * 3)(*origin), len(d), d.ctypes.data, pdf.ctypes.data)
        return pdf

    def sample_li(self, camera, params, px, py, sample):
        L = (c_float * 3)()
        n = c_int32()
        self.lib.orc_sample_li(self.h, ctypes.byref(camera), ctypes.byref(params), px, py, sample, L, ctypes.byref(n))
        return np.array(L[:], np.float32), n.value
