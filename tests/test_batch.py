"""Batched rendering (bbgpu_render_batch) and the autotuner (bbgpu_autotune).

Jobs of several scenes and cameras, some sharing a scene, run concurrently on
a stream pool; every job's film must equal, bit for bit, the film of the same
render alone.  The autotuner only changes speed knobs: images are unchanged.
"""
import copy

import numpy as np
import pytest

from bulbit_amd import capi, scenes

pytestmark = pytest.mark.gpu


def _jobs():
    from bulbit_amd.gpu import GpuScene

    out = []
    sc1, cam1 = scenes.icosphere_scene(width=256, height=256, level=5, n_lights=8)
    sc2, cam2 = scenes.icosphere_scene(width=320, height=180, level=6, n_lights=32, seed=7)
    sc3, cam3 = scenes.readme_scene(width=200, height=120)
    g1, g2, g3 = GpuScene(sc1), GpuScene(sc2), GpuScene(sc3)
    cam1b = copy.copy(cam1)
    scenes.set_filter(cam1b, "gaussian")
    specs = [(g1, cam1, scenes.render_params(4, max_bounces=4)), (g1, cam1b, scenes.render_params(2, max_bounces=12)),
             (g2, cam2, scenes.render_params(3, max_bounces=8, seed=5)), (g3, cam3, scenes.render_params(8)),
             (g2, cam2, scenes.render_params(2, max_bounces=6, sample_begin=3)), (g3, cam3, scenes.render_params(1))]
    for g, cam, p in specs:
        out.append((g, cam, p, g.new_film(cam)))
    return out, (g1, g2, g3)


def test_batch_equals_individual_renders():
    from bulbit_amd.gpu import render_batch

    jobs, keep = _jobs()
    st = render_batch(jobs, streams=4)
    assert st.streams == 4 and st.samples == sum(c.width * c.height * p.spp for _, c, p, _ in jobs)
    for i, (g, cam, p, film) in enumerate(jobs):
        s, c, _ = g.render(cam, p)
        bs = film[0].cpu().numpy().reshape(s.shape)
        bc = film[1].cpu().numpy().reshape(c.shape)
        assert np.array_equal(bc, c), i
        assert np.array_equal(bs.view(np.int32), s.view(np.int32)), i
    print(f"[batch] {len(jobs)} jobs, {st.samples} samples in {st.ms_total:.1f} ms", flush=True)


def test_autotune_keeps_the_image():
    from bulbit_amd.gpu import GpuScene, autotune

    sc, cam = scenes.icosphere_scene(width=320, height=180, level=6, n_lights=16)
    g = GpuScene(sc)
    p = scenes.render_params(4)
    t = autotune(g, cam, p)
    assert t.spp == p.spp and t.max_bounces == p.max_bounces and t.seed == p.seed
    assert t.trace_refill in (0, 16, 32, 48) and t.trace_lds_stack in (0, 8, 16, 24)
    assert t.trace_block in (0, 64, 128, 256) and t.trace_stack in (0, 1)
    s0, c0, _ = g.render(cam, p)
    s1, c1, _ = g.render(cam, t)
    assert np.array_equal(c0, c1) and np.array_equal(s0.view(np.int32), s1.view(np.int32))
    print(f"[batch] tuned refill {t.trace_refill}, block {t.trace_block}, stack {t.trace_stack}/{t.trace_lds_stack}, "
          f"paths {t.wavefront_paths}", flush=True)


@pytest.mark.parametrize("block", [64, 128, 256])
@pytest.mark.parametrize("stack", ["lds8", "lds32", "scratch"])
def test_traversal_launch_variants_give_the_same_image(block, stack):
    """Block size and stack placement of the traversal kernels (the
    autotuner's knobs) change nothing in the image, bit for bit, on a scene
    with all the material kinds of the icosphere generator and on the
    README scene (a sphere primitive)."""
    from bulbit_amd.gpu import GpuScene

    for sc, cam in (scenes.icosphere_scene(width=160, height=90, level=6, n_lights=16), scenes.readme_scene(96, 54)):
        g = GpuScene(sc)
        p = scenes.render_params(4)
        s0, c0, _ = g.render(cam, p)
        q = scenes.render_params(4, trace_block=block, trace_stack=1 if stack == "scratch" else 0,
                                 trace_lds_stack=0 if stack == "scratch" else int(stack[3:]))
        s1, c1, _ = g.render(cam, q)
        assert np.array_equal(c0, c1) and np.array_equal(s0.view(np.int32), s1.view(np.int32))
        g.close()


def test_graph_renders_equal_eager_renders():
    """graph = 1 captures a render's launches once and replays the graph for
    calls that differ only in sample_begin (k_raygen reads it from device
    memory).  Every replay equals the eager render bit for bit, including a
    partial last pass, with eager renders of other cameras on the same scene
    between the replays."""
    from bulbit_amd.gpu import GpuScene

    sc, cam = scenes.icosphere_scene(width=192, height=108, level=5, n_lights=8)
    cam2 = copy.copy(cam)
    scenes.set_filter(cam2, "gaussian")
    g = GpuScene(sc)
    knobs = dict(max_bounces=6, wavefront_paths=192 * 108 * 2)  # 2 samples per pass: passes of 2, 2, 1
    s, c = g.new_film(cam)
    for i, sb in enumerate([0, 5, 17, 5, 40]):
        s.zero_()
        c.zero_()
        g.render_device(cam, scenes.render_params(5, sample_begin=sb, graph=1, **knobs), s.data_ptr(), c.data_ptr())
        gs = s.cpu().numpy().reshape(108, 192, 3)
        gc = c.cpu().numpy().reshape(108, 192)
        es, ec, _ = g.render(cam, scenes.render_params(5, sample_begin=sb, **knobs))
        assert np.array_equal(gc, ec), (i, sb)
        assert np.array_equal(gs.view(np.int32), es.view(np.int32)), (i, sb)
        g.render(cam2, scenes.render_params(1, sample_begin=i))  # other eager work in between


def test_batch_with_graphs_equals_individual_renders():
    """The batched form with every job launched as a captured graph: the first
    call captures, the second replays; both equal the jobs rendered alone."""
    from bulbit_amd.gpu import render_batch

    jobs, keep = _jobs()
    for _, _, p, _ in jobs:
        p.graph = 1
    for call in range(2):
        for _, _, _, film in jobs:
            film[0].zero_()
            film[1].zero_()
        render_batch(jobs, streams=4)
        for i, (g, cam, p, film) in enumerate(jobs):
            eager = capi.BbRenderParams.from_buffer_copy(p)
            eager.graph = 0
            s, c, _ = g.render(cam, eager)
            bs = film[0].cpu().numpy().reshape(s.shape)
            bc = film[1].cpu().numpy().reshape(c.shape)
            assert np.array_equal(bc, c), (call, i)
            assert np.array_equal(bs.view(np.int32), s.view(np.int32)), (call, i)


def test_shadow_rays_beside_closest_hits_keep_the_film():
    """On scenes of 8 primitives or more each depth's shadow rays run on
    a second stream beside the next depth's closest hits; the next shading
    waits for them, so the film equals, bit for bit, that of a counting
    render, which keeps every launch on one stream.  Several passes per
    call and timing statistics (events on both streams) included."""
    from bulbit_amd.gpu import GpuScene

    sc, cam = scenes.icosphere_scene(width=240, height=135, level=6, n_lights=32, seed=3)
    assert len(sc.prims) >= 8
    g = GpuScene(sc)
    p = scenes.render_params(6, max_bounces=10, wavefront_paths=240 * 135 * 2)  # passes of 2 samples
    s0, c0, _ = g.render(cam, p, count_traversal=True)
    s1, c1, st = g.render(cam, p, stats=True)
    s2, c2, _ = g.render(cam, p)
    for s, c in ((s1, c1), (s2, c2)):
        assert np.array_equal(c, c0) and np.array_equal(s.view(np.int32), s0.view(np.int32))
    assert st.rays_shadow > 0 and st.ms_trace_shadow > 0 and st.ms_trace_closest > 0
    print(f"[batch] overlapped render: closest {st.ms_trace_closest:.2f} ms, shadow {st.ms_trace_shadow:.2f} ms, "
          f"total {st.ms_total:.2f} ms", flush=True)
    g.close()
