"""Committed golden vectors (tests/golden/, written by make_golden.py).

CPU: the oracle still produces the frozen vectors (sampler streams, the first
65,536 closest hits of the 10M-ray batch on the 1.31M-primitive scene, the
README film).  GPU: the HIP path reproduces the frozen hits bit for bit over
all 10M rays (SHA-256 of the hit array and the traversal counters) and the
README film within the render tolerance.

The scene and ray generators use numpy; if a machine's numpy produces
different scene bits (the digest in the manifest guards this), the frozen
vectors do not apply there and the comparison is skipped with that reason.
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import pytest

from bulbit_amd import capi, scenes

from conftest import pixel_agreement

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
THREADS = max(1, min(16, len(os.sched_getaffinity(0))))


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def scene_digest(sc):
    # the constant material fields (the texture slots that follow them are
    # zero in these scenes)
    const = capi.BbMaterial.tex_color.offset
    mats = np.frombuffer(b"".join(bytes(m)[:const] for m in sc.materials), np.uint8)
    lights = np.frombuffer(b"".join(bytes(l) for l in sc.lights), np.uint8)
    return sha(sc.positions.astype(np.float32), sc.normals.astype(np.float32), sc.tangents.astype(np.float32),
               sc.indices.astype(np.int32), sc.prims.astype(np.int32), mats, lights)


@pytest.fixture(scope="module")
def c2_batch():
    m = manifest()["c2"]
    sc, cam = scenes.icosphere_scene()
    if scene_digest(sc) != m["scene_sha256"]:
        pytest.skip("this machine's numpy generates different scene bits; the frozen vectors do not apply")
    rays = scenes.portable_random_rays(np.array(m["world_min"], np.float32), np.array(m["world_max"], np.float32),
                                       m["rays"], seed=m["ray_seed"])
    assert sha(rays) == m["rays_sha256"]
    gold = np.load(os.path.join(GOLDEN, "c2_rays.npz"))
    return sc, rays, gold, m


def test_sampler_streams_match_golden(orc):
    with open(os.path.join(GOLDEN, "sampler_kat.json")) as f:
        kat = json.load(f)
    for k in kat:
        out = (ctypes.c_float * 8)()
        orc.orc_sampler_floats(k["px"], k["py"], k["seed"], k["sample"], 8, out)
        bits = [int(np.float32(v).view(np.uint32)) for v in out]
        assert bits == k["floats_bits"], k
        assert int(orc.orc_pixel_sample_hash(k["px"], k["py"], k["seed"], k["sample"])) == k["hash"]


def test_oracle_reproduces_golden_c2_hits(c2_batch):
    from oracle import OracleScene

    sc, rays, gold, m = c2_batch
    o = OracleScene(sc)
    lo, hi = o.world_bounds()
    assert [float(x) for x in lo] == m["world_min"] and [float(x) for x in hi] == m["world_max"]
    n = m["subset"]
    hits, _ = o.trace(rays[:n], threads=THREADS)
    assert np.array_equal(hits.view(np.int32), gold["hits"].view(np.int32))


def test_oracle_reproduces_golden_readme_film():
    from oracle import OracleScene

    m = manifest()["readme"]
    sc, cam = scenes.readme_scene(width=64, height=36)
    assert scene_digest(sc) == m["scene_sha256"]
    gold = np.load(os.path.join(GOLDEN, "readme_64x36.npz"))
    s, c, _ = OracleScene(sc).render(cam, scenes.render_params(m["spp"]), threads=THREADS)
    assert np.array_equal(c, gold["count"])
    assert np.array_equal(s.view(np.int32), gold["sum"].view(np.int32))


@pytest.mark.gpu
def test_gpu_reproduces_golden_c2_hits_over_10m_rays(c2_batch):
    from bulbit_amd.gpu import GpuScene

    sc, rays, gold, m = c2_batch
    print(f"[golden] tracing {len(rays)} rays on the GPU", flush=True)
    g = GpuScene(sc)
    hits, cnt = g.trace(rays, count=True)
    n = m["subset"]
    assert np.array_equal(hits[:n].view(np.int32), gold["hits"].view(np.int32))
    assert sha(hits) == m["hits_sha256"]
    assert cnt.as_dict() == m["counters"]


@pytest.mark.gpu
def test_gpu_readme_film_matches_golden():
    from bulbit_amd.gpu import GpuScene

    m = manifest()["readme"]
    sc, cam = scenes.readme_scene(width=64, height=36)
    gold = np.load(os.path.join(GOLDEN, "readme_64x36.npz"))
    s, c, _ = GpuScene(sc).render(cam, scenes.render_params(m["spp"]))
    assert np.array_equal(c, gold["count"])
    frac = pixel_agreement(s / c[..., None], gold["sum"] / gold["count"][..., None])
    assert frac >= 0.999
