"""Benchmark: Msamples/s of the wavefront path tracer at 1080p, 8 bounces.

Workload (one step = one full render): the displaced-icosphere scene at
1920x1080, 64 samples per pixel, max_bounces 8, 1,310,786 primitives
(icosphere level 8 + ground + 64 emissive triangles), PowerLightSampler,
float32.  The scene and its BVH are built and resident in HBM before timing.

Printed on stdout: one JSON line with the metric, a roofline entry for the
dominant kernel (closest-hit traversal) from HIP events recorded around every
launch inside the timed region, and the CPU baseline: the oracle (a CPU
restatement of the reference path tracer) on a bounded, strided sample of the
same frame on the host cores.  Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Msamples/s at 1080p, 8 bounces; relMSE at equal time, cache vs plain PT"
HBM_PEAK_GBS = 8000.0
# algorithmic bytes per closest-hit ray: 32 B per node box tested, 36 B per
# triangle tested (p0, p1, p2), 4 B per accepted candidate, 28 B ray in
# (o, d, t_max), 16 B hit out
BYTES_NODE, BYTES_PRIM, BYTES_ACCEPT, BYTES_RAY = 32, 36, 4, 28 + 16
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "trace_closest_c2_traffic.json")


T_START = time.time()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--max-bounces", type=int, default=8)
    ap.add_argument("--level", type=int, default=8)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU work of the baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--wavefront-paths", type=int, default=0)
    ap.add_argument("--ref-spp", type=int, default=4096, help="spp of the relMSE reference (0: skip the quality check)")
    ap.add_argument("--batch", action="store_true",
                    help="only the batched workload: 32 seeded scenes of uneven size in one bbgpu_render_batch call")
    ap.add_argument("--cache-frames", type=int, default=256,
                    help="1-spp frames of the equal-time radiance-cache comparison (0: skip it)")
    ap.add_argument("--large", action="store_true",
                    help="the large-scene workload instead (4096 displaced level-5 icospheres, 3840x2160, 4 spp per "
                         "frame, max_bounces 12, the 128 x 5 radiance cache)")
    ap.add_argument("--large-spheres", type=int, default=4096)
    ap.add_argument("--large-frames", type=int, default=64)
    ap.add_argument("--no-extra", action="store_true",
                    help="leave out the batched and large-scene workloads of the default run")
    ap.add_argument("--large-after", type=float, default=240.0,
                    help="skip the large-scene workload of the default run when more seconds than this are spent")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        sys.exit("bench.py runs on exactly one GPU: the path tracer is a single-GPU build "
                 "(multi-GPU sharding is a design note only); rerun with --gpus 1")

    import numpy as np
    import torch

    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene

    if not torch.cuda.is_available():
        sys.exit("no HIP device visible")
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return
    if args.batch:
        print(json.dumps(batch_workload(args)), flush=True)
        return
    if args.large:
        print(json.dumps(large_workload(args)), flush=True)
        return

    log(f"[bench] generating scene level {args.level} at {args.width}x{args.height}")
    t0 = time.time()
    sc, cam = scenes.icosphere_scene(width=args.width, height=args.height, level=args.level)
    log(f"[bench] scene generated in {time.time() - t0:.2f}s: {len(sc.prims)} primitives")
    g = GpuScene(sc)
    log(f"[bench] BVH {g.info.n_nodes} nodes built in {g.info.build_seconds:.2f}s, {g.info.device_bytes / 1e6:.1f} MB on device")
    params = scenes.render_params(args.spp, max_bounces=args.max_bounces, wavefront_paths=args.wavefront_paths)
    s_dev, c_dev = g.new_film(cam)

    for i in range(args.warmup):
        s_dev.zero_()
        c_dev.zero_()
        g.render_device(cam, params, s_dev.data_ptr(), c_dev.data_ptr())
        log(f"[bench] warmup {i + 1}/{args.warmup} done")

    # counted run (untimed): traversal work of exactly the rays the timed steps trace
    s_dev.zero_()
    c_dev.zero_()
    cst = g.render_device(cam, params, s_dev.data_ptr(), c_dev.data_ptr(), stats=True, count_traversal=True)
    cl = cst.closest
    log(f"[bench] counted run: {cl.rays} closest rays, {cl.nodes / max(cl.rays, 1):.1f} nodes/ray, "
        f"{cl.prims / max(cl.rays, 1):.2f} prims/ray")

    torch.cuda.synchronize()
    stats = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        s_dev.zero_()
        c_dev.zero_()
        st = g.render_device(cam, params, s_dev.data_ptr(), c_dev.data_ptr(), stats=True)
        stats.append(st)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t_start
    for i, st in enumerate(stats):
        log(f"[bench] step {i + 1}: {st.ms_total:.1f} ms (closest {st.ms_trace_closest:.1f}, shadow "
            f"{st.ms_trace_shadow:.1f}, shade {st.ms_shade:.1f})")

    samples_per_step = args.width * args.height * args.spp
    value = samples_per_step * args.steps / elapsed / 1e6
    ms_closest = float(np.mean([s.ms_trace_closest for s in stats]))
    launches = int(stats[0].launches_closest)
    bytes_step = (BYTES_NODE * cl.nodes + BYTES_PRIM * cl.prims + BYTES_ACCEPT * cl.accepts + BYTES_RAY * cl.rays)
    achieved = bytes_step / (ms_closest / 1e3) / 1e9
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        with open(PROFILE_TRAFFIC) as f:
            traffic = json.load(f).get("bytes_per_launch")
    rays_total = stats[0].rays_closest + stats[0].rays_shadow
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "displaced icosphere (level %d, %d primitives, 64 area lights), %dx%d, %d spp, max_bounces %d, "
                        "plain wavefront path tracer" % (args.level, len(sc.prims), args.width, args.height, args.spp,
                                                         args.max_bounces),
            "width": args.width,
            "height": args.height,
            "spp": args.spp,
            "max_bounces": args.max_bounces,
            "primitives": len(sc.prims),
        },
        "roofline": {
            "kernel": "k_trace_closest (closest-hit traversal of the reference BVH, two levels per fetch)",
            "bound": "hbm",
            # what the counters show (profiles/traversal_c2_counters.md): the
            # kernel is below the HBM ceiling because of latency and lane
            # divergence, with measured HBM traffic well under the algorithmic bytes
            "limiter": "memory latency and lane divergence: 30 of 64 lanes active per VALU instruction, waves "
                       "waiting 45% of their cycles, L2 hit rate 48%; measured HBM traffic 0.20x the algorithmic "
                       "bytes (1.89 TB/s), the rest served from L2 and the Infinity Cache, so the algorithmic "
                       "rate can pass the HBM peak; each launch shares the GPU with the previous depth's shadow "
                       "traversal (profiles/traversal_c2_counters.md, trace_closest_c2_traffic.md)",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "launches_per_step": launches,
            "avg_launch_ms": round(ms_closest / max(launches, 1), 4),
            "algorithmic_bytes_per_launch": int(bytes_step / max(launches, 1)),
        },
        "rays_per_s": round(rays_total * args.steps / elapsed, 1),
        "relmse_equal_time_cache_vs_plain": None,
    }

    ref_c2 = None
    if args.ref_spp > 0:
        test = s_dev.cpu().numpy().reshape(args.height, args.width, 3) / \
            c_dev.cpu().numpy().reshape(args.height, args.width, 1).astype(np.float32)
        out["quality"], ref_c2 = quality(g, sc, cam, args, test)
    if args.cache_frames > 0:
        cq = cache_comparisons(args, g, cam, ref_c2)
        # the cache configuration as stated (README scene, 256 one-sample
        # frames under a static camera): the frames accumulated at equal
        # time is the natural reading; single frames at equal frame time
        # (the real-time use) beside it, with equal standing
        out["relmse_equal_time_cache_vs_plain"] = cq["readme"]["ratio"]
        out["relmse_equal_time_cache_vs_plain_accumulated"] = cq["readme"]["ratio"]
        out["relmse_equal_time_cache_vs_plain_per_frame"] = cq["readme"]["per_frame"]["ratio_equal_time"]
        out["radiance_cache"] = cq
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(sc, cam, params, args, samples_per_step)
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    g.close()
    del s_dev, c_dev
    torch.cuda.empty_cache()
    # the other workloads of the configuration set, as extra keys: the batched
    # form (32 scenes per call) and the large scene with the 128 x 5 cache
    if not args.no_extra:
        out["batched"] = batch_workload(args)
        spent = time.time() - T_START
        if spent < args.large_after:
            out["large_scene"] = large_workload(args, nested=True)
        else:
            out["large_scene"] = {"skipped": f"{spent:.0f}s spent before it (limit {args.large_after:.0f}s)"}
    print(json.dumps(out), flush=True)


RELMSE_EPS = 1e-2


def relmse(x, r):
    """mean over pixels and channels of (x - r)^2 / (r^2 + eps), eps = 1e-2"""
    import numpy as np
    x = x.astype(np.float64)
    r = r.astype(np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + RELMSE_EPS)))


def quality(g, sc, cam, args, test):
    """relMSE of the timed render (samples [0, spp)) against references with
    disjoint samples [spp, spp + ref_spp): the whole frame rendered on the GPU,
    and a crop rendered by the oracle on the CPU."""
    import numpy as np
    from bulbit_amd import scenes

    t0 = time.time()
    p = scenes.render_params(args.ref_spp, max_bounces=args.max_bounces, sample_begin=args.spp)
    s, c, _ = g.render(cam, p)
    ref = s / c[..., None].astype(np.float32)
    gpu_secs = time.time() - t0
    out = {"epsilon": RELMSE_EPS, "ref_spp": args.ref_spp, "ref_samples": [args.spp, args.spp + args.ref_spp],
           "relmse_vs_gpu_reference": round(relmse(test, ref), 6), "gpu_reference_seconds": round(gpu_secs, 2)}
    log(f"[bench] relMSE of the {args.spp}-spp frame vs a {args.ref_spp}-spp GPU frame: {out['relmse_vs_gpu_reference']}")
    if not args.no_cpu:
        from oracle import OracleScene
        w, h = 96, 54
        x0, y0 = cam.width // 2 - w // 2, cam.height // 2 - h // 2
        rect = (x0, y0, x0 + w, y0 + h)
        threads = host_cores()["threads"]
        o = OracleScene(sc)
        s2, c2, secs = o.render(cam, p, rect=rect, threads=threads)
        cpu_ref = (s2 / np.maximum(c2, 1)[..., None])[y0:y0 + h, x0:x0 + w]
        out["cpu_crop"] = list(rect)
        out["relmse_vs_cpu_reference_crop"] = round(relmse(test[y0:y0 + h, x0:x0 + w], cpu_ref), 6)
        out["relmse_gpu_vs_cpu_reference_crop"] = round(relmse(ref[y0:y0 + h, x0:x0 + w], cpu_ref), 8)
        out["cpu_reference_seconds"] = round(secs, 2)
        log(f"[bench] crop {rect}: relMSE vs CPU {args.ref_spp}-spp {out['relmse_vs_cpu_reference_crop']}, "
            f"GPU vs CPU reference {out['relmse_gpu_vs_cpu_reference_crop']}")
    return out, ref


def mlp_flops_query(width=64, layers=4):
    """FLOPs of one query: 74 inputs, `layers` hidden layers of `width`, 3 outputs"""
    return 2 * (74 * width + (layers - 1) * width * width + width * 3)


MLP_FLOPS_QUERY = mlp_flops_query()
# rows per Adam step of the cached renders: up to 65,536, taking a frame's
# training rows batch by batch (the README scene makes about 68K per frame,
# so its four steps take 65,536, the rest, 65,536, the rest)
CACHE_TRAIN_BATCH = 65536
BF16_PEAK_TFLOPS = 2500.0


def cache_comparison(args, g, cam, ref, name, query_primary, train_stride=16, learning_rate=1e-3, train_steps=4,
                     spp=1, max_bounces=8, shape=(64, 4), frames=None, keep_every=16, query_direct=False,
                     query_first=0, train_batch=None):
    """Radiance cache vs plain path tracing at equal wall-clock time, `spp`
    samples per frame, against `ref` (a plain render with disjoint samples).
    Plain PT renders `frames` (default `cache_frames`) frames; the cached
    tracer (a network of shape[1] hidden layers of width shape[0], trained
    online on 1/train_stride of its paths, train_steps Adam steps per frame)
    renders frames until it has used the same time.
    Two measures: the relMSE of the frames accumulated over the whole time
    (progressive rendering), and the mean relMSE of single frames from the
    second half of the run (real-time rendering), with plain PT's per-frame
    error scaled to the cached tracer's frame time (variance ~ 1 / samples)."""
    import numpy as np
    import torch
    from bulbit_amd import scenes
    from bulbit_amd.nrc import Nrc

    w, h = cam.width, cam.height
    base = 1 << 16  # frames use samples base + f, disjoint from the references

    def run(step, frames, budget=None):
        acc = g.new_film(cam)
        frame = g.new_film(cam)
        keep = []
        torch.cuda.synchronize()
        t = time.perf_counter()
        done = 0
        for f in range(frames):
            frame[0].zero_()
            frame[1].zero_()
            step(f, frame)
            acc[0].add_(frame[0])
            acc[1].add_(frame[1])
            if f % keep_every == keep_every - 1:
                keep.append((f, frame[0].clone(), frame[1].clone()))
            torch.cuda.synchronize()
            done = f + 1
            if budget is not None and time.perf_counter() - t >= budget:
                break
        secs = time.perf_counter() - t

        def err(sd, cd):
            sn, cn = sd.cpu().numpy().reshape(h, w, 3), cd.cpu().numpy().reshape(h, w)
            return relmse(sn / np.maximum(cn, 1)[..., None], ref)

        per_frame = [err(sd, cd) for f, sd, cd in keep if f >= done // 2]
        return err(acc[0], acc[1]), done, secs, float(np.mean(per_frame)) if per_frame else float("nan")

    if train_batch is None:
        train_batch = CACHE_TRAIN_BATCH
    cp = scenes.cache_params(train_stride=train_stride, train_batch=train_batch, train_steps=train_steps,
                             query_primary=query_primary, query_direct=query_direct, query_first=query_first)
    frames = args.cache_frames if frames is None else frames
    flops_q = mlp_flops_query(*shape)

    def params(f):
        return scenes.render_params(spp, max_bounces=max_bounces, sample_begin=base + f * spp)

    # warm both paths (kernel loading) outside the timed runs
    warm = scenes.render_params(spp, max_bounces=max_bounces, sample_begin=1 << 24)
    g.render_device(cam, warm, *(t.data_ptr() for t in g.new_film(cam)))
    net = Nrc(seed=1, learning_rate=learning_rate, width=shape[0], layers=shape[1])
    g.render_cached_device(net, cam, warm, cp, *(t.data_ptr() for t in g.new_film(cam)))
    net.close()

    e_plain, n_plain, t_plain, f_plain = run(
        lambda f, film: g.render_device(cam, params(f), film[0].data_ptr(), film[1].data_ptr()), frames)
    net = Nrc(seed=1, learning_rate=learning_rate, width=shape[0], layers=shape[1])
    stats = []

    def cached(f, film):
        # device film, no host copy: the same as the plain leg
        stats.append(g.render_cached_device(net, cam, params(f), cp, film[0].data_ptr(), film[1].data_ptr()))

    e_cache, n_cache, t_cache, f_cache = run(cached, 100000, budget=t_plain)
    net.close()
    q = sum(st.queries for st in stats)
    rows = sum(st.train_rows for st in stats)
    ms_q = sum(st.ms_query for st in stats)
    ms_t = sum(st.ms_train for st in stats)
    ms_r = sum(st.ms_rows for st in stats)
    tflops = (flops_q * q + 3 * flops_q * rows) / max(ms_q + ms_t, 1e-9) / 1e9
    frame_plain = t_plain / n_plain
    frame_cache = t_cache / n_cache
    f_plain_equal_time = f_plain * frame_plain / frame_cache
    log(f"[bench] {name} (train 1/{train_stride}, lr {learning_rate:g}, {train_steps} steps, direct light at the "
        f"query vertex: {bool(query_direct)}): equal time {t_plain:.2f}s: plain {n_plain} frames relMSE {e_plain:.6f}; cached {n_cache} "
        f"frames relMSE {e_cache:.6f}; per frame: plain {f_plain:.5f} ({frame_plain * 1e3:.2f} ms), cached "
        f"{f_cache:.5f} ({frame_cache * 1e3:.2f} ms); {q} queries, {rows} training rows")
    return {
        "workload": name,
        "network": {"hidden_width": shape[0], "hidden_layers": shape[1]},
        "spp_per_frame": spp,
        "query_at_primary_hit": bool(query_primary),
        "direct_light_at_query_vertex": bool(query_direct),
        "first_query_vertex": 0 if query_primary else max(1, query_first),
        "diverged": any(st.diverged for st in stats),
        "train_stride": train_stride,
        "learning_rate": learning_rate,
        "train_steps_per_frame": train_steps,
        "train_batch": train_batch,
        "seconds": round(t_plain, 3),
        "plain_frames": n_plain,
        "plain_relmse": round(e_plain, 7),
        "cache_frames": n_cache,
        "cache_relmse": round(e_cache, 7),
        "ratio": round(e_cache / e_plain, 4),
        "per_frame": {"plain_relmse": round(f_plain, 6), "plain_ms": round(frame_plain * 1e3, 3),
                      "cache_relmse": round(f_cache, 6), "cache_ms": round(frame_cache * 1e3, 3),
                      "ratio_equal_time": round(f_cache / f_plain_equal_time, 4)},
        "cache_ms_query": round(ms_q, 1),
        "cache_ms_train": round(ms_t, 1),
        "cache_ms_rows": round(ms_r, 1),
        "mlp": {"bound": "mfma", "achieved": round(tflops, 3), "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(tflops / BF16_PEAK_TFLOPS, 6),
                "flops_per_query": flops_q, "flops_per_training_row": 3 * flops_q},
    }


def cache_comparisons(args, g_c2, cam_c2, ref_c2):
    """The README scene as the cache configuration states it (query at the
    first rough hit, which on a lone convex sphere is the primary hit; 1/16
    of the paths train, Adam lr 1e-3), the same with lr 1e-2, and the 1080p
    icosphere scene as stated, where indirect light is what the cache stands
    in for (query from vertex 1 on)."""
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene

    sc, cam = scenes.readme_scene(width=1920, height=1080)
    g = GpuScene(sc)
    t0 = time.time()
    rs, rc, _ = g.render(cam, scenes.render_params(max(args.ref_spp, 1), sample_begin=1 << 20))
    log(f"[bench] README {args.ref_spp}-spp reference in {time.time() - t0:.1f}s")
    name = "README scene 1920x1080, 1 spp per frame, max_bounces 8"
    ref = rs / rc[..., None]
    # as the cache configuration states it: 1/16 of the paths train, each 4
    # bounces past its query vertex, Adam lr 1e-3; 4 steps per frame
    out = {"readme": cache_comparison(args, g, cam, ref, name, True, train_stride=16, learning_rate=1e-3),
           "readme_lr_1e-2": cache_comparison(args, g, cam, ref, name, True, train_stride=16, learning_rate=1e-2),
           # the query vertex takes its direct light; the cache stands in for
           # the indirect rest only (no bias on the direct light)
           "readme_direct": cache_comparison(args, g, cam, ref, name, True, train_stride=16, learning_rate=1e-2,
                                             query_direct=True)}
    g.close()
    if ref_c2 is not None:
        ico = "displaced icosphere 1920x1080, 1 spp per frame, max_bounces 8"
        out["icosphere"] = cache_comparison(args, g_c2, cam_c2, ref_c2, ico, False, train_stride=16,
                                            learning_rate=1e-3)
        out["icosphere_direct"] = cache_comparison(args, g_c2, cam_c2, ref_c2, ico, False, train_stride=16,
                                                   learning_rate=1e-3, query_direct=True)
    return out


def large_workload(args, nested=False):
    """The large scene: `--large-spheres` displaced level-5 icospheres (4096:
    83,886,080 triangles) Poisson-disk placed on a 100 x 100 plane under 1024
    emissive triangles of log-normal power, metallic-roughness materials;
    3840x2160, 4 spp per frame, max_bounces 12.  Reports plain-PT Msamples/s,
    rays/s and the closest-hit kernel's algorithmic bytes/s (with the
    traversal counters of one frame), then the 128 x 5 radiance cache against
    plain PT at equal time over `--large-frames` frames, against a
    `--ref-spp` plain render with disjoint samples."""
    import numpy as np
    import torch
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene

    W, H, SPP, MB = 3840, 2160, 4, 12
    # inside the default run: a few frames of its own, whatever --steps says
    steps, warmup = (3, 1) if nested else (args.steps, args.warmup)
    t0 = time.time()
    sc, cam = scenes.large_scene(width=W, height=H, n_spheres=args.large_spheres)
    gen_s = time.time() - t0
    log(f"[bench] large scene generated in {gen_s:.1f}s: {len(sc.prims)} primitives")
    g = GpuScene(sc)
    log(f"[bench] BVH {g.info.n_nodes} nodes built in {g.info.build_seconds:.1f}s, "
        f"{g.info.device_bytes / 1e9:.2f} GB on device")
    n_prims = len(sc.prims)
    del sc
    params = scenes.render_params(SPP, max_bounces=MB)
    s_dev, c_dev = g.new_film(cam)
    for i in range(warmup):
        g.render_device(cam, params, s_dev.data_ptr(), c_dev.data_ptr())
        log(f"[bench] warmup {i + 1}/{warmup} done")
    cst = g.render_device(cam, params, s_dev.data_ptr(), c_dev.data_ptr(), stats=True, count_traversal=True)
    cl = cst.closest
    log(f"[bench] counted frame: {cl.rays} closest rays, {cl.nodes / max(cl.rays, 1):.1f} nodes/ray, "
        f"{cl.prims / max(cl.rays, 1):.2f} prims/ray")
    torch.cuda.synchronize()
    stats = []
    t = time.perf_counter()
    for i in range(steps):
        s_dev.zero_()
        c_dev.zero_()
        stats.append(g.render_device(cam, params, s_dev.data_ptr(), c_dev.data_ptr(), stats=True))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t
    value = W * H * SPP * steps / elapsed / 1e6
    ms_closest = float(np.mean([s.ms_trace_closest for s in stats]))
    bytes_frame = BYTES_NODE * cl.nodes + BYTES_PRIM * cl.prims + BYTES_ACCEPT * cl.accepts + BYTES_RAY * cl.rays
    achieved = bytes_frame / (ms_closest / 1e3) / 1e9
    rays = stats[0].rays_closest + stats[0].rays_shadow
    log(f"[bench] plain: {value:.1f} Msamples/s, {elapsed / steps * 1e3:.1f} ms per frame (closest "
        f"{ms_closest:.1f} ms, {achieved:.0f} GB/s algorithmic)")
    out = {
        "metric": "Msamples/s at 2160p, 12 bounces, large scene; relMSE at equal time, cache vs plain PT",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(elapsed / steps * 1e3, 3),
        "higher_is_better": True,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%d displaced level-5 icospheres on a plane, 1024 area lights (%d primitives), "
                               "%dx%d, %d spp, max_bounces %d" % (args.large_spheres, n_prims, W, H, SPP, MB),
                   "primitives": n_prims, "scene_seconds": round(gen_s, 1),
                   "bvh_seconds": round(g.info.build_seconds, 1), "device_bytes": int(g.info.device_bytes)},
        "traversal": {"kernel": "k_trace_closest", "algorithmic_gbs": round(achieved, 1),
                      "frac_of_hbm_peak": round(achieved / HBM_PEAK_GBS, 4), "ms_per_frame": round(ms_closest, 2),
                      "nodes_per_ray": round(cl.nodes / max(cl.rays, 1), 2),
                      "prims_per_ray": round(cl.prims / max(cl.rays, 1), 3)},
        "rays_per_s": round(rays * steps / elapsed, 1),
    }
    if args.ref_spp > 0:
        # the reference in chunks of 64 spp, samples disjoint from every frame
        rs, rc = g.new_film(cam)
        t0 = time.time()
        chunk = 64
        for c0 in range(0, args.ref_spp, chunk):
            p = scenes.render_params(min(chunk, args.ref_spp - c0), max_bounces=MB, sample_begin=(1 << 20) + c0)
            g.render_device(cam, p, rs.data_ptr(), rc.data_ptr())
            torch.cuda.synchronize()
            log(f"[bench] reference {c0 + p.spp}/{args.ref_spp} spp ({time.time() - t0:.0f}s)")
        ref = rs.cpu().numpy().reshape(H, W, 3) / rc.cpu().numpy().reshape(H, W, 1).astype(np.float32)
        del rs, rc
        test = s_dev.cpu().numpy().reshape(H, W, 3) / c_dev.cpu().numpy().reshape(H, W, 1).astype(np.float32)
        out["quality"] = {"epsilon": RELMSE_EPS, "ref_spp": args.ref_spp, "ref_seconds": round(time.time() - t0, 1),
                          "relmse_4spp_frame_vs_reference": round(relmse(test, ref), 6)}
        if args.large_frames > 0:
            name = "large scene %dx%d, %d spp per frame, max_bounces %d" % (W, H, SPP, MB)
            # the query vertex takes its direct light and the cache stands in
            # for the indirect rest (the lower accumulated error); beside it the
            # cache for everything the vertex reflects (the lower per-frame error)
            cq = cache_comparison(args, g, cam, ref, name, False, train_stride=16, learning_rate=1e-3,
                                  train_steps=4, spp=SPP, max_bounces=MB, shape=(128, 5), frames=args.large_frames,
                                  keep_every=8, query_direct=True)
            ca = cache_comparison(args, g, cam, ref, name, False, train_stride=16, learning_rate=1e-3,
                                  train_steps=4, spp=SPP, max_bounces=MB, shape=(128, 5), frames=args.large_frames,
                                  keep_every=8)
            out["relmse_equal_time_cache_vs_plain"] = cq["ratio"]
            out["relmse_equal_time_cache_vs_plain_accumulated"] = cq["ratio"]
            out["relmse_equal_time_cache_vs_plain_per_frame"] = cq["per_frame"]["ratio_equal_time"]
            out["radiance_cache"] = cq
            out["radiance_cache_all"] = ca
    g.close()
    return out


def batch_workload(args):
    """32 independent seeded scenes per call: displaced icospheres with
    triangle counts log-uniform in [1e4, 2e6] (subdivision level 5..8),
    resolutions from {256x256, 512x512, 1024x576, 1024x1024}, 16 spp,
    max_bounces uniform in 4..12, 4 to 256 area lights; launch parameters
    from the autotuner (one probe per resolution); all jobs in one
    bbgpu_render_batch call on 16 streams, against the same jobs one by one."""
    import numpy as np
    import torch
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene, autotune, render_batch

    jobs, tuned = [], {}
    t0 = time.time()
    for i, spec in enumerate(scenes.batch_workload_specs(32, seed=303)):
        sc, cam, p = scenes.batch_workload_job(spec)
        w, h, level, n_lights = spec["width"], spec["height"], spec["level"], spec["n_lights"]
        g = GpuScene(sc)
        if (w, h) not in tuned:
            tuned[(w, h)] = autotune(g, cam, p)
        t = tuned[(w, h)]
        for knob in ("trace_refill", "shadow_refill", "trace_block", "trace_stack", "trace_lds_stack", "wavefront_paths"):
            setattr(p, knob, getattr(t, knob))
        jobs.append((g, cam, p, g.new_film(cam)))
        log(f"[bench] job {i}: level {level} ({len(sc.prims)} prims), {w}x{h}, max_bounces {p.max_bounces}, "
            f"{n_lights} lights")
    log(f"[bench] 32 scenes built and tuned in {time.time() - t0:.1f}s")
    render_batch(jobs, streams=16)  # warm-up (workspaces)
    for j in jobs:
        j[3][0].zero_()
        j[3][1].zero_()
    torch.cuda.synchronize()
    st = render_batch(jobs, streams=16)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for g, cam, p, film in jobs:
        g.render_device(cam, p, film[0].data_ptr(), film[1].data_ptr())
    torch.cuda.synchronize()
    serial_ms = (time.perf_counter() - t1) * 1e3
    jm = np.array(st.job_ms_list)
    out = {
        "metric": "batched Msamples/s (32 scenes per call)",
        "value": round(st.samples / st.ms_total / 1e3, 3),
        "unit": "Msamples/s",
        "samples": int(st.samples),
        "ms_total": round(st.ms_total, 2),
        "streams": int(st.streams),
        "job_ms_p50": round(float(np.percentile(jm, 50)), 2),
        "job_ms_p99": round(float(np.percentile(jm, 99)), 2),
        "serial_ms": round(serial_ms, 2),
        "serial_msamples_per_s": round(st.samples / serial_ms / 1e3, 3),
        "tuned": {f"{w}x{h}": {"refill": t.trace_refill, "block": t.trace_block,
                               "stack": "scratch" if t.trace_stack else f"lds {t.trace_lds_stack or 'auto'}",
                               "wavefront_paths": t.wavefront_paths} for (w, h), t in tuned.items()},
    }
    log(f"[bench] batch: {out['value']} Msamples/s in {st.ms_total:.1f} ms; one by one {serial_ms:.1f} ms")
    return out


def copy_calibration(mib):
    """A plain copy of known size: what the HBM byte counters read for it."""
    import torch
    a = torch.ones(mib * 1024 * 1024 // 4, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] copied 3 x {mib} MiB")


def host_cores():
    """The cores this process may use: its CPU affinity, lowered to the
    cgroup's CPU quota (cpu.max) when one is set.  Every one of the oracle's
    threads renders tiles (none only polls, unlike the reference's pool of
    n - 1 workers), so `threads` = that core count occupies every core."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(float(q) / float(period) + 0.5))
    except (OSError, ValueError):
        pass
    threads = max(1, min(affinity, quota) if quota else affinity)
    return {"threads": threads, "os_cpu_count": os.cpu_count(), "affinity": affinity, "cgroup_quota": quota}


def cpu_baseline(sc, cam, params, args, samples_full):
    """The oracle on the host cores: full frames of the same scene, fewer spp."""
    from oracle import OracleScene
    from bulbit_amd import scenes

    hc = host_cores()
    threads = hc["threads"]
    t0 = time.time()
    o = OracleScene(sc)
    log(f"[bench] oracle scene built in {time.time() - t0:.2f}s; calibrating the CPU sample on {threads} threads")
    _, _, secs = o.render(cam, scenes.render_params(1, max_bounces=args.max_bounces), threads=threads)
    npix = cam.width * cam.height
    rate = npix / max(secs, 1e-9)
    spp = int(max(1, min(args.spp, round(args.cpu_seconds * rate / npix))))
    p = scenes.render_params(spp, max_bounces=args.max_bounces, sample_begin=1)
    _, _, secs = o.render(cam, p, threads=threads)
    n = npix * spp
    value = n / secs / 1e6
    log(f"[bench] CPU baseline: {n} samples in {secs:.2f}s = {value:.3f} Msamples/s")
    return {
        "value": round(value, 4),
        "unit": "Msamples/s",
        "cores": threads,
        "os_cpu_count": hc["os_cpu_count"],
        "affinity": hc["affinity"],
        "cgroup_quota": hc["cgroup_quota"],
        "kind": "port",
        "sample": f"oracle (CPU restatement of PathIntegrator) rendering the same {cam.width}x{cam.height} frame at "
                  f"{spp} spp (samples 1..{spp}), {n} samples in {secs:.1f}s on {threads} threads (every core the process may use: "
                  f"affinity {hc['affinity']}, cgroup quota {hc['cgroup_quota']}, os.cpu_count {hc['os_cpu_count']}); "
                  f"BVH build excluded",
    }


if __name__ == "__main__":
    main()
