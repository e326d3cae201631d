"""Python host mirror of the drop-in boundary.

``GpuScene`` wraps a ``BbScene`` handle: ``trace`` mirrors BVH::Intersect /
IntersectAny over a ray batch and ``render`` mirrors Integrator::Render
(accumulating a Film's sums and counts).  Device memory comes from torch (HIP);
every computation runs in libbbgpu.so.  Without a HIP device or without the
library these calls raise: there is no CPU path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise capi.BbError("no HIP device visible: the GPU path tracer has no CPU fallback")
    return torch


class GpuScene:
    def __init__(self, scene_data):
        self.lib = capi.load_library()
        self.torch = _torch()
        self.data = scene_data
        desc = scene_data.desc()
        h = ctypes.c_void_p()
        capi.check(self.lib.bbgpu_scene_create(ctypes.byref(desc), ctypes.byref(h)), "bbgpu_scene_create", self.lib)
        self.h = h
        info = capi.BbSceneInfo()
        capi.check(self.lib.bbgpu_scene_info(self.h, ctypes.byref(info)), "bbgpu_scene_info", self.lib)
        self.info = info

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.lib.bbgpu_scene_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def trace_device(self, rays_dev, n, hits_dev, any_hit=False, count=False, stream=None):
        cnt = capi.BbTraceCounters() if count else None
        capi.check(
            self.lib.bbgpu_trace(self.h, ctypes.c_void_p(rays_dev), int(n), int(any_hit), ctypes.c_void_p(hits_dev),
                                 ctypes.byref(cnt) if cnt is not None else None, stream),
            "bbgpu_trace",
            self.lib,
        )
        return cnt

    def trace_noalloc_device(self, rays_dev, n, hits_dev, any_hit=False, count=False, stream=None):
        """bbgpu_trace_noalloc: the same batch traced with the library's static scratch array."""
        cnt = capi.BbTraceCounters() if count else None
        capi.check(
            self.lib.bbgpu_trace_noalloc(self.h, ctypes.c_void_p(rays_dev), int(n), int(any_hit),
                                         ctypes.c_void_p(hits_dev), ctypes.byref(cnt) if cnt is not None else None,
                                         stream),
            "bbgpu_trace_noalloc",
            self.lib,
        )
        return cnt

    def trace(self, rays: np.ndarray, any_hit=False, count=False):
        """rays: RAY_DTYPE array -> (hits HIT_DTYPE array, counters or None)."""
        torch = self.torch
        rays = np.ascontiguousarray(rays)
        n = len(rays)
        r = torch.from_numpy(rays.view(np.float32).reshape(-1)).to("cuda")
        h = torch.empty(4 * max(n, 1), dtype=torch.int32, device="cuda")
        cnt = self.trace_device(r.data_ptr(), n, h.data_ptr(), any_hit, count)
        hits = h.cpu().numpy()[: 4 * n].view(capi.HIT_DTYPE).copy()
        return hits, cnt

    def render_device(self, camera, params, sum_dev, count_dev, stats=False, count_traversal=False, stream=None):
        st = capi.BbRenderStats() if stats else None
        capi.check(
            self.lib.bbgpu_render(self.h, ctypes.byref(camera), ctypes.byref(params), ctypes.c_void_p(sum_dev),
                                  ctypes.c_void_p(count_dev), stream, ctypes.byref(st) if st is not None else None,
                                  int(count_traversal)),
            "bbgpu_render",
            self.lib,
        )
        return st

    def render_cached_device(self, nrc, camera, params, cache_params, sum_dev, count_dev, stream=None):
        """bbgpu_render_cached into device film buffers; returns BbCacheStats."""
        st = capi.BbCacheStats()
        capi.check(
            self.lib.bbgpu_render_cached(self.h, nrc.h, ctypes.byref(camera), ctypes.byref(params),
                                         ctypes.byref(cache_params), ctypes.c_void_p(sum_dev),
                                         ctypes.c_void_p(count_dev), stream, ctypes.byref(st)),
            "bbgpu_render_cached",
            self.lib,
        )
        return st

    def render_cached(self, nrc, camera, params, cache_params, film=None, stream=None):
        """bbgpu_render_cached: (sum[h, w, 3], count[h, w], BbCacheStats) as numpy;
        film (device tensors from new_film) accumulates across calls."""
        s, c = film if film is not None else self.new_film(camera)
        st = self.render_cached_device(nrc, camera, params, cache_params, s.data_ptr(), c.data_ptr(), stream)
        h, w = camera.height, camera.width
        return s.cpu().numpy().reshape(h, w, 3), c.cpu().numpy().reshape(h, w), st

    def new_film(self, camera):
        torch = self.torch
        npix = camera.width * camera.height
        return (torch.zeros(3 * npix, dtype=torch.float32, device="cuda"),
                torch.zeros(npix, dtype=torch.int32, device="cuda"))

    def render(self, camera, params, stats=False, count_traversal=False, film=None):
        """Returns (sum[h, w, 3], count[h, w], stats or None) as numpy."""
        s, c = film if film is not None else self.new_film(camera)
        st = self.render_device(camera, params, s.data_ptr(), c.data_ptr(), stats, count_traversal)
        h, w = camera.height, camera.width
        return s.cpu().numpy().reshape(h, w, 3), c.cpu().numpy().reshape(h, w), st


def film_image(film_sum: np.ndarray, film_count: np.ndarray) -> np.ndarray:
    """Film::GetRenderedImage (src/reconstruction/film.cpp:80-100): sum / count."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return (film_sum / film_count[..., None].astype(np.float32)).astype(np.float32)


def render_batch(jobs, streams=0):
    """bbgpu_render_batch over [(GpuScene, camera, params, (sum_dev, count_dev)), ...]
    (films from GpuScene.new_film); returns BbBatchStats."""
    lib = capi.load_library()
    arr = (capi.BbRenderJob * max(1, len(jobs)))()
    for i, (g, cam, params, film) in enumerate(jobs):
        arr[i].scene = g.h.value
        arr[i].camera = cam
        arr[i].params = params
        arr[i].film_sum_dev = film[0].data_ptr()
        arr[i].film_count_dev = film[1].data_ptr()
    st = capi.BbBatchStats()
    job_ms = (ctypes.c_double * max(1, len(jobs)))()
    st.job_ms = job_ms
    capi.check(lib.bbgpu_render_batch(arr, len(jobs), streams, ctypes.byref(st)), "bbgpu_render_batch", lib)
    st.job_ms_list = list(job_ms[:len(jobs)])
    return st


def autotune(g, camera, params, probe_spp=2):
    """bbgpu_autotune: params with the fastest refill / LDS stack / wavefront size"""
    out = capi.BbRenderParams()
    capi.check(g.lib.bbgpu_autotune(g.h, ctypes.byref(camera), ctypes.byref(params), probe_spp, ctypes.byref(out)),
               "bbgpu_autotune", g.lib)
    return out
