"""bbgpu_trace_noalloc, the entry point that allocates no device memory,
against bbgpu_trace on the same ray batches: hit records bit for bit and the
same traversal counters, on two calls in a row with the second one smaller
(the static scratch keeps the last call's contents, so the second call shows
that what must start at zero is cleared)."""
import ctypes

import numpy as np
import pytest

from bulbit_amd import capi, scenes

pytestmark = pytest.mark.gpu


def _trace_both(g, rays, any_hit, count):
    import torch

    n = len(rays)
    r = torch.from_numpy(np.ascontiguousarray(rays).view(np.float32).reshape(-1)).to("cuda")
    h_ref = torch.full((4 * max(n, 1),), -7, dtype=torch.int32, device="cuda")
    h_new = torch.full((4 * max(n, 1),), -9, dtype=torch.int32, device="cuda")
    c_ref = g.trace_device(r.data_ptr(), n, h_ref.data_ptr(), any_hit, count)
    c_new = g.trace_noalloc_device(r.data_ptr(), n, h_new.data_ptr(), any_hit, count)
    torch.cuda.synchronize()
    a = h_ref.cpu().numpy()[: 4 * n].view(capi.HIT_DTYPE)
    b = h_new.cpu().numpy()[: 4 * n].view(capi.HIT_DTYPE)
    return a, b, c_ref, c_new


def _world(g):
    return np.array(g.info.world_min, np.float32), np.array(g.info.world_max, np.float32)


@pytest.mark.parametrize("scene", ["readme", "ico6"])
@pytest.mark.parametrize("any_hit", [False, True])
@pytest.mark.parametrize("count", [False, True])
def test_noalloc_trace_equals_trace_on_two_calls(scene, any_hit, count):
    from bulbit_amd.gpu import GpuScene

    sc = scenes.readme_scene(64, 36)[0] if scene == "readme" else scenes.icosphere_scene(64, 36, level=6)[0]
    g = GpuScene(sc)
    try:
        bmin, bmax = _world(g)
        rng = np.random.default_rng(11)
        for n, seed in ((300_000, 21), (40_000, 22)):
            t_max = rng.uniform(0.0, 10.0, n) if any_hit else np.inf
            rays = scenes.random_rays(bmin, bmax, n, seed=seed, t_max=t_max)
            print(f"[no-alloc] {scene} any_hit={any_hit} count={count} n={n}", flush=True)
            a, b, c_ref, c_new = _trace_both(g, rays, any_hit, count)
            assert np.array_equal(a.view(np.int32), b.view(np.int32))
            assert 0.05 < np.mean(a["prim"] >= 0) < 0.95
            if count:
                assert c_ref.as_dict() == c_new.as_dict()
                assert c_new.rays == n and c_new.nodes > 0
    finally:
        g.close()


def test_noalloc_trace_empty_batch_and_host_buffers_rejected():
    from bulbit_amd.gpu import GpuScene

    g = GpuScene(scenes.readme_scene(32, 18)[0])
    try:
        cnt = capi.BbTraceCounters()
        st = g.lib.bbgpu_trace_noalloc(g.h, None, 0, 0, None, ctypes.byref(cnt), None)
        assert st == 0 and cnt.rays == 0 and cnt.nodes == 0
        rays = scenes.random_rays(*_world(g), 64, seed=3)
        hits = np.zeros(64, capi.HIT_DTYPE)
        st = g.lib.bbgpu_trace_noalloc(g.h, ctypes.c_void_p(rays.ctypes.data), 64, 0, ctypes.c_void_p(hits.ctypes.data),
                                       None, None)
        assert st == 1
        assert "device memory" in capi.last_error(g.lib)
    finally:
        g.close()
