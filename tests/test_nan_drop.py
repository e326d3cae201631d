"""Samples whose radiance is NaN or Inf are dropped and not counted.

UniDirectionalRayIntegrator::Render tests every sample with IsNullish and
skips Film::AddSample for it (src/integrator/integrator.cpp:193-196,
src/reconstruction/film.cpp:26-37), so a pixel's sum holds only finite
samples and its count says how many.  The README scene's sphere gets an
albedo whose red channel is NaN: every path whose BSDF sample or NEE
contribution passes through the sphere carries NaN in red, while the
background and the paths that end on the sphere with a black contribution
stay finite.  The oracle must drop exactly those samples; the GPU must drop
the same ones (equal counts) and agree with the oracle on what it kept.
"""
import os

import numpy as np
import pytest

from bulbit_amd import scenes

from conftest import pixel_agreement

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
SPP = 8


def _poisoned_scene(width, height, red=float("nan")):
    sc, cam = scenes.readme_scene(width, height)
    sc.materials[0] = scenes.diffuse((red, 0.2, 0.2))
    return sc, cam


def _sphere_block(cam):
    """pixels inside the sphere's image: the camera looks at the sphere's
    centre from sqrt(35) away, so the sphere (radius 1, angular radius 9.7
    degrees) covers a disk of 0.55 half-heights around the image centre
    (vfov 35 degrees); the block of +-0.2 half-heights lies inside it"""
    cx, cy, r = cam.width // 2, cam.height // 2, max(1, int(0.2 * cam.height / 2))
    return (slice(cy - r, cy + r), slice(cx - r, cx + r))


@pytest.mark.parametrize("red", [float("nan"), float("inf")])
def test_oracle_drops_nullish_samples(red):
    from oracle import OracleScene

    sc, cam = _poisoned_scene(96, 54, red)
    o = OracleScene(sc)
    s, c, _ = o.render(cam, scenes.render_params(SPP), threads=THREADS)
    clean, cc, _ = OracleScene(scenes.readme_scene(96, 54)[0]).render(cam, scenes.render_params(SPP), threads=THREADS)
    assert np.all(cc == SPP)
    assert np.all(np.isfinite(s)), "a dropped sample reached the film sum"
    assert np.all((c >= 0) & (c <= SPP))
    # sphere pixels lose the samples that went through the poisoned albedo
    assert c[_sphere_block(cam)].mean() < 0.5 * SPP
    # the left and right image borders see only the environment: their
    # paths never touch the sphere and keep every sample, with the clean
    # scene's values
    assert np.all(c[:, :8] == SPP) and np.all(c[:, -8:] == SPP)
    assert np.array_equal(s[:, :8], clean[:, :8]) and np.array_equal(s[:, -8:], clean[:, -8:])
    assert int(c.sum()) < int(cc.sum())


@pytest.mark.gpu
@pytest.mark.parametrize("red", [float("nan"), float("inf")])
def test_gpu_drops_the_same_samples(red):
    from bulbit_amd.gpu import GpuScene
    from oracle import OracleScene

    sc, cam = _poisoned_scene(160, 90, red)
    g, o = GpuScene(sc), OracleScene(sc)
    p = scenes.render_params(SPP)
    gs, gc, _ = g.render(cam, p)
    os_, oc, _ = o.render(cam, p, threads=THREADS)
    g.close()
    assert np.all(np.isfinite(gs))
    assert np.array_equal(gc, oc), "the GPU dropped different samples than the oracle"
    assert int(gc.sum()) < SPP * cam.width * cam.height
    kept = gc > 0
    gi = gs[kept] / gc[kept][:, None]
    oi = os_[kept] / oc[kept][:, None]
    frac = pixel_agreement(gi, oi)
    print(f"poisoned albedo {red}: {int(gc.sum())} of {SPP * cam.width * cam.height} samples kept, "
          f"{frac:.6f} of kept pixels within 1e-3", flush=True)
    assert frac >= 0.999
