"""Writes the golden fixtures of tests/golden/ from the oracle.

The reference's own tests hold no vectors for this path and the reference
cannot be built in this image, so the vectors are the oracle's outputs,
frozen: they pin the oracle against later edits and give the GPU tests a
committed target that does not need the oracle at run time.

    python tests/golden/make_golden.py        # rewrites the fixtures

Fixtures:
  sampler_kat.json   first NextFloat values of (pixel, seed, sample) streams
                     and MurmurHash64A keys (independent_sampler.cpp:8-32)
  c2_rays.npz        closest-hit results of the first 65,536 of 10,000,000
                     seeded random rays on the 1,310,786-primitive scene,
                     plus SHA-256 digests of the scene, of the rays and of
                     the full 10M hit array (prim, t, u, v), and the counters
  readme_64x36.npz   README scene film (sum and count) at 64x36, 4 spp
  manifest.json      seeds, sizes and digests
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bulbit_amd import capi, scenes  # noqa: E402

C2_RAY_SEED = 20261202
C2_RAYS = 10_000_000
C2_SUBSET = 65_536
KAT_KEYS = [(0, 0, 0, 0), (1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 7, 0), (0, 0, 0, 1), (639, 359, 0, 15),
            (1919, 1079, 3, 63), (123, 456, -5, 1000)]


def sha(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def scene_digest(sc) -> str:
    # the constant material fields (the texture slots that follow them are
    # zero in these scenes)
    const = capi.BbMaterial.tex_color.offset
    mats = np.frombuffer(b"".join(bytes(m)[:const] for m in sc.materials), np.uint8)
    lights = np.frombuffer(b"".join(bytes(l) for l in sc.lights), np.uint8)
    return sha(sc.positions.astype(np.float32), sc.normals.astype(np.float32), sc.tangents.astype(np.float32),
               sc.indices.astype(np.int32), sc.prims.astype(np.int32), mats, lights)


def c2_rays(bmin, bmax):
    return scenes.portable_random_rays(bmin, bmax, C2_RAYS, seed=C2_RAY_SEED)


def main():
    from oracle import OracleScene, load_oracle

    orc = load_oracle()
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    manifest = {"generator": "tests/golden/make_golden.py", "oracle": "oracle/oracle.c"}

    # sampler streams
    import ctypes

    kat = []
    for px, py, seed, sample in KAT_KEYS:
        out = (ctypes.c_float * 8)()
        orc.orc_sampler_floats(px, py, seed, sample, 8, out)
        kat.append({"px": px, "py": py, "seed": seed, "sample": sample,
                    "hash": int(orc.orc_pixel_sample_hash(px, py, seed, sample)),
                    "floats_bits": [int(np.float32(v).view(np.uint32)) for v in out]})
    with open(os.path.join(HERE, "sampler_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)

    # C2 ray batch
    t0 = time.time()
    sc, cam = scenes.icosphere_scene()
    o = OracleScene(sc)
    bmin, bmax = o.world_bounds()
    rays = c2_rays(bmin, bmax)
    hits, cnt = o.trace(rays, threads=threads)
    print(f"traced {C2_RAYS} rays in {time.time() - t0:.1f}s: {cnt.as_dict()}", flush=True)
    np.savez_compressed(os.path.join(HERE, "c2_rays.npz"), hits=hits[:C2_SUBSET],
                        counters=np.array([cnt.rays, cnt.nodes, cnt.prims, cnt.accepts], np.uint64))
    manifest["c2"] = {
        "scene": "bulbit_amd.scenes.icosphere_scene() (level 8, seed 2026, 1920x1080)",
        "primitives": int(len(sc.prims)),
        "scene_sha256": scene_digest(sc),
        "ray_seed": C2_RAY_SEED,
        "rays": C2_RAYS,
        "subset": C2_SUBSET,
        "rays_sha256": sha(rays),
        "hits_sha256": sha(hits),
        "world_min": [float(x) for x in bmin],
        "world_max": [float(x) for x in bmax],
        "counters": cnt.as_dict(),
    }
    del o

    # README scene film
    sc, cam = scenes.readme_scene(width=64, height=36)
    o = OracleScene(sc)
    s, c, _ = o.render(cam, scenes.render_params(4), threads=threads)
    np.savez_compressed(os.path.join(HERE, "readme_64x36.npz"), sum=s, count=c)
    manifest["readme"] = {"scene": "bulbit_amd.scenes.readme_scene(64, 36)", "spp": 4, "seed": 0, "max_bounces": 8,
                          "scene_sha256": scene_digest(sc)}

    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)
    print(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    main()
