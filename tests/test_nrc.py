"""The radiance-cache MLP: the CPU restatement (oracle/nrc_ref.py) pinned by
float64 finite differences, and the bf16 MFMA kernels (bulbit_amd/csrc/nrc.hip)
checked against it: forward values, the loss gradient over a batch, one Adam
step, and that training fits a smooth radiance field."""
import numpy as np
import pytest

from oracle import nrc_ref as ref


def _batch(n, seed):
    rng = np.random.default_rng(seed)
    raw = np.zeros((n, 16), np.float32)
    raw[:, 0:3] = rng.random((n, 3))
    d = rng.normal(size=(n, 3))
    raw[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    nrm = rng.normal(size=(n, 3))
    raw[:, 6:9] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    raw[:, 9] = rng.random(n)
    raw[:, 10:13] = rng.random((n, 3))
    return raw


def _field(raw):
    """a smooth radiance field of position and normal"""
    x, y, z = raw[:, 0], raw[:, 1], raw[:, 2]
    base = 0.6 + 0.3 * np.sin(2 * np.pi * x) * np.cos(np.pi * y) + 0.2 * z
    return np.stack([base, base * 0.8 + 0.1 * raw[:, 8], base * 0.5 + 0.2 * raw[:, 12]], 1).astype(np.float32)


SHAPES = [pytest.param(s, id=f"{s[0]}x{s[1]}") for s in ref.SHAPES]


def _init_weights(seed, shape=(ref.WIDTH, ref.HIDDEN)):
    rng = np.random.default_rng(seed)
    hidden = shape[1]
    w = []
    for l, (k, n) in enumerate(ref.layer_shapes(*shape)):
        a = np.sqrt(6.0 / ((74 if l == 0 else k) + (3 if l == hidden else n)))
        m = rng.uniform(-a, a, (k, n))
        if l == 0:
            m[74:] = 0
        if l == hidden:
            m[:, 3:] = 0
        w.append(m.ravel())
    return np.concatenate(w).astype(np.float32)


# ------------------------------------------------------------------ oracle

def test_bf16_rounds_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3], np.float32)
    out = ref.bf16(x)
    assert out[0] == 1.0
    assert out[1] == 1.0  # tie -> even
    assert out[2] == np.float32(1.0 + 4 * 2 ** -8)  # tie -> even (up)
    assert out[3] == -2.5
    assert abs(out[4] - 1e-3) < 1e-3 * 2 ** -8


def test_encoding_layout():
    raw = _batch(5, 1).astype(np.float64)
    e = ref.encode(raw)
    assert e.shape == (5, 80)
    assert np.allclose(e[:, 0], np.sin(np.pi * raw[:, 0])) and np.allclose(e[:, 1], np.cos(np.pi * raw[:, 0]))
    assert np.allclose(e[:, 6], np.sin(2 * np.pi * raw[:, 0]))  # frequency-major: f = 1 starts at 6
    assert np.allclose(e[:, 62], np.sin(1024 * np.pi * raw[:, 1]))
    assert np.allclose(e[:, 64:74], raw[:, 3:13]) and np.all(e[:, 74:] == 0)


@pytest.mark.parametrize("shape", SHAPES)
def test_oracle_gradient_matches_finite_differences(shape):
    """The backward formulas (float64, relative-loss denominator held fixed
    as NRC's stop-gradient) against central differences of the loss."""
    raw = _batch(40, 2)
    t = _field(raw)
    w = _init_weights(3, shape).astype(np.float64)
    g = ref.grad64(w, raw, t)
    y0 = ref.forward64(w, raw)

    def frozen_loss(wv):
        y = ref.forward64(wv, raw)
        return float(np.mean((y - t) ** 2 / (y0 * y0 + 0.01)))

    rng = np.random.default_rng(4)
    offs = np.cumsum([0] + [k * n for k, n in ref.layer_shapes(*shape)])
    for l, (k, n) in enumerate(ref.layer_shapes(*shape)):
        cols = 3 if l == shape[1] else n
        rows = 74 if l == 0 else k
        for _ in range(12):
            i = offs[l] + rng.integers(0, rows) * n + rng.integers(0, cols)
            e = 1e-6
            wp, wm = w.copy(), w.copy()
            wp[i] += e
            wm[i] -= e
            fd = (frozen_loss(wp) - frozen_loss(wm)) / (2 * e)
            assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (l, i, fd, g[i])


@pytest.mark.parametrize("shape", SHAPES)
def test_bf16_oracle_tracks_float64(shape):
    raw = _batch(256, 5)
    t = _field(raw)
    w = _init_weights(6, shape)
    y, _ = ref.forward(w, raw)
    y64 = ref.forward64(w.astype(np.float64), raw)
    assert np.max(np.abs(y - y64)) < 0.05 * max(1.0, np.max(np.abs(y64)))
    _, g = ref.loss_and_grad(w, raw, t)
    g64 = ref.grad64(w.astype(np.float64), raw, t)
    cos = g @ g64 / np.linalg.norm(g) / np.linalg.norm(g64)
    assert cos > 0.99


# --------------------------------------------------------------------- GPU

@pytest.fixture(scope="module", params=SHAPES)
def nrc(request):
    from bulbit_amd.nrc import Nrc

    net = Nrc(seed=7, width=request.param[0], layers=request.param[1])
    yield net
    net.close()


@pytest.mark.gpu
def test_nrc_forward_matches_oracle(nrc):
    w = _init_weights(8, (nrc.width, nrc.layers))
    nrc.set_weights(w)
    raw = _batch(1000, 9)  # not a multiple of the 64-row tile
    y = nrc.query(raw)
    yr, _ = ref.forward(w, raw)
    rel = np.abs(y - yr) / np.maximum(np.abs(yr), 0.05)
    print(f"[nrc] forward: median rel {np.median(rel):.2e}, p99.5 {np.percentile(rel, 99.5):.2e}", flush=True)
    assert np.median(rel) < 1e-3
    assert np.percentile(rel, 99.5) < 2e-2


@pytest.mark.gpu
def test_nrc_gradient_matches_oracle(nrc):
    w = _init_weights(10, (nrc.width, nrc.layers))
    nrc.set_weights(w)
    raw = _batch(3000, 11)
    t = _field(raw)
    loss, g = nrc.train(raw, t, apply=False, want_grad=True)
    lr, gr = ref.loss_and_grad(w, raw, t)
    print(f"[nrc] loss gpu {loss:.6f} oracle {lr:.6f}", flush=True)
    assert abs(loss - lr) <= 2e-3 * lr
    cos = g @ gr / np.linalg.norm(g) / np.linalg.norm(gr)
    ratio = np.linalg.norm(g) / np.linalg.norm(gr)
    print(f"[nrc] gradient cosine {cos:.6f}, norm ratio {ratio:.5f}", flush=True)
    assert cos > 0.999 and abs(ratio - 1) < 0.01
    # the padded output columns get no gradient
    last = np.asarray(g[-nrc.width * 16:]).reshape(nrc.width, 16)
    assert np.all(last[:, 3:] == 0)


@pytest.mark.gpu
def test_nrc_gradient_over_many_passes_matches_oracle(nrc):
    """40,000 rows: more row passes than the fused step has blocks (256 x
    128 rows), so blocks accumulate dW over several passes; the three-kernel
    step of the wide network splits them into many row ranges."""
    w = _init_weights(14, (nrc.width, nrc.layers))
    nrc.set_weights(w)
    raw = _batch(40000, 15)
    t = _field(raw)
    loss, g = nrc.train(raw, t, apply=False, want_grad=True)
    lr, gr = ref.loss_and_grad(w, raw, t)
    cos = g @ gr / np.linalg.norm(g) / np.linalg.norm(gr)
    ratio = np.linalg.norm(g) / np.linalg.norm(gr)
    print(f"[nrc] 40000 rows: loss gpu {loss:.6f} oracle {lr:.6f}, cosine {cos:.6f}, norm ratio {ratio:.5f}",
          flush=True)
    assert abs(loss - lr) <= 2e-3 * lr
    assert cos > 0.999 and abs(ratio - 1) < 0.01


@pytest.mark.gpu
def test_nrc_gradient_is_deterministic(nrc):
    """The partial sums of dW and of the loss are added in a fixed order (no
    atomics): the same batch gives the same bits."""
    w = _init_weights(16, (nrc.width, nrc.layers))
    nrc.set_weights(w)
    raw = _batch(20000, 17)
    t = _field(raw)
    l1, g1 = nrc.train(raw, t, apply=False, want_grad=True)
    l2, g2 = nrc.train(raw, t, apply=False, want_grad=True)
    assert np.array_equal(g1.view(np.uint32), g2.view(np.uint32))
    assert np.float32(l1).view(np.uint32) == np.float32(l2).view(np.uint32)


@pytest.mark.gpu
def test_nrc_adam_step_matches_oracle(nrc):
    w = _init_weights(12, (nrc.width, nrc.layers))
    nrc.set_weights(w)
    raw = _batch(2048, 13)
    t = _field(raw)
    _, g = nrc.train(raw, t, apply=False, want_grad=True)
    nrc.set_weights(w)  # resets nothing of Adam's moments: they are still zero
    from bulbit_amd.nrc import Nrc

    fresh = Nrc(seed=1, learning_rate=1e-3, width=nrc.width, layers=nrc.layers)
    fresh.set_weights(w)
    fresh.train(raw, t, apply=True)
    w1 = fresh.weights()
    expect, _, _ = ref.adam(w.astype(np.float64), 0.0, 0.0, g.astype(np.float64), 1, 1e-3)
    mask = np.ones_like(w, bool)
    mask[-nrc.width * 16:] = (np.arange(nrc.width * 16) % 16) < 3
    assert np.allclose(w1[mask], expect[mask], rtol=1e-5, atol=1e-6)
    assert np.all(w1[~mask] == 0)
    fresh.close()


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES)
def test_nrc_learns_a_smooth_field(shape):
    from bulbit_amd.nrc import Nrc

    net = Nrc(seed=21, learning_rate=1e-2, width=shape[0], layers=shape[1])
    test_raw = _batch(4096, 99)
    test_t = _field(test_raw)

    def test_loss():
        y = net.query(test_raw)
        return float(np.mean((y - test_t) ** 2 / (y * y + 0.01)))

    before = test_loss()
    for step in range(300):
        raw = _batch(8192, 1000 + step)
        net.train(raw, _field(raw))
    after = test_loss()
    print(f"[nrc] held-out relative L2 {before:.4f} -> {after:.5f}", flush=True)
    assert after < before / 10
    net.close()


# ------------------------------------------------------ cached path tracer

def _relmse(x, r):
    return float(np.mean((x.astype(np.float64) - r) ** 2 / (r.astype(np.float64) ** 2 + 1e-2)))


@pytest.mark.gpu
def test_cached_render_with_all_paths_training_equals_plain():
    """Training paths that never reach their stop vertex are plain paths: the
    cache logic draws no random numbers, so with every path a training path
    and 8 extra vertices the image is the plain render, bit for bit."""
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene
    from bulbit_amd.nrc import Nrc

    sc, cam = scenes.icosphere_scene(width=160, height=90, level=4, n_lights=8)
    g = GpuScene(sc)
    p = scenes.render_params(4)
    s0, c0, _ = g.render(cam, p)
    net = Nrc(seed=3)
    # one Adam step per pass: no path is ever queried, so training cannot
    # change the image, and train_rows counts the rows it trained on
    s1, c1, st = g.render_cached(net, cam, p, scenes.cache_params(train_stride=1, train_extra_bounces=8,
                                                                   train_steps=1))
    assert np.array_equal(c0, c1)
    assert np.array_equal(s0.view(np.int32), s1.view(np.int32))
    assert st.queries == 0 and st.train_rows > 0
    net.close()


@pytest.mark.gpu
def test_training_rows_beyond_the_steps_are_not_built():
    """A pass that makes far more rows than its steps read (every path
    training, 2 steps of 512 rows): the row kernel writes only the rows the
    steps take and stops early, and the steps still take full batches."""
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene
    from bulbit_amd.nrc import Nrc

    sc, cam = scenes.icosphere_scene(width=160, height=90, level=4, n_lights=8)
    g = GpuScene(sc)
    p = scenes.render_params(4)
    s0, c0, _ = g.render(cam, p)
    net = Nrc(seed=3)
    cp = scenes.cache_params(train_stride=1, train_extra_bounces=8, train_batch=512, train_steps=2)
    s1, c1, st = g.render_cached(net, cam, p, cp)
    assert np.array_equal(s0.view(np.int32), s1.view(np.int32)) and np.array_equal(c0, c1)
    assert st.queries == 0
    assert st.train_rows == 2 * 512  # two full batches out of tens of thousands of rows
    assert np.isfinite(st.loss) and st.loss > 0
    net.close()
    g.close()


@pytest.mark.gpu
def test_cached_render_with_a_zero_cache_keeps_the_path_prefix():
    """With a cache that predicts 0, a rendering path contributes exactly what
    the plain path gathered before its query vertex, so every pixel is at
    most the plain render's and the query vertices are where they should be
    (after the first non-specular bounce)."""
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene
    from bulbit_amd.nrc import Nrc

    sc, cam = scenes.readme_scene(width=160, height=90)
    g = GpuScene(sc)
    p = scenes.render_params(4)
    s0, c0, _ = g.render(cam, p)
    net = Nrc(seed=3)
    net.set_weights(np.zeros(net.n_params, np.float32))
    n_paths = 160 * 90 * 4
    s1, c1, st = g.render_cached(net, cam, p, scenes.cache_params(train_stride=n_paths + 1, train_steps=0,
                                                                   query_primary=True))
    assert np.array_equal(c0, c1)
    assert np.all(s1 <= s0 * (1 + 1e-6) + 1e-7)
    # the diffuse sphere's primary hits are rough: they all query at vertex 0
    assert 0 < st.queries <= n_paths
    assert s1.mean() < s0.mean()
    net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("direct", [False, True], ids=["cache-all", "direct+cache"])
def test_cached_render_beats_plain_at_equal_samples_after_training(direct):
    """Displaced icosphere under area lights (indirect light matters): after
    training online over 48 one-sample frames, 16 more cached frames have
    lower relMSE against a 1024-spp plain reference than 16 plain frames (the
    cache replaces the noisy tails of the paths), whether the cache stands in
    for all the query vertex reflects or only for its indirect part."""
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene
    from bulbit_amd.nrc import Nrc

    sc, cam = scenes.icosphere_scene(width=240, height=135, level=5, n_lights=16)
    g = GpuScene(sc)
    rs, rc, _ = g.render(cam, scenes.render_params(1024, sample_begin=100000))
    ref = rs / rc[..., None]
    net = Nrc(seed=5, learning_rate=1e-2)
    cp = scenes.cache_params(train_stride=4, train_batch=16384, train_steps=4, query_direct=direct)
    for f in range(48):
        _, _, st = g.render_cached(net, cam, scenes.render_params(1, sample_begin=f), cp)
    print(f"[nrc] training loss after 48 frames {st.loss:.4f}, rows/frame {st.train_rows}", flush=True)
    film = g.new_film(cam)
    for f in range(48, 64):
        s1, c1, _ = g.render_cached(net, cam, scenes.render_params(1, sample_begin=f), cp, film=film)
    s0, c0, _ = g.render(cam, scenes.render_params(16, sample_begin=48))
    e_cache = _relmse(s1 / c1[..., None], ref)
    e_plain = _relmse(s0 / c0[..., None], ref)
    print(f"[nrc] relMSE at 16 spp (direct light at the query vertex: {direct}): cached {e_cache:.5f}, "
          f"plain {e_plain:.5f}", flush=True)
    assert e_cache < e_plain
    net.close()


# ------------------------------------------------ divergence and the fallback

def _readme_small():
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene

    sc, cam = scenes.readme_scene(width=160, height=90)
    return GpuScene(sc), cam


@pytest.mark.gpu
def test_diverged_cache_falls_back_to_the_plain_tracer():
    """Non-finite weights (set from the host) mark the network diverged at
    once; a cached render is then the plain render, bit for bit, with no
    query, and says so.  New finite weights clear the flag."""
    from bulbit_amd import scenes
    from bulbit_amd.nrc import Nrc

    g, cam = _readme_small()
    p = scenes.render_params(4)
    s0, c0, _ = g.render(cam, p)
    net = Nrc(seed=3)
    w = net.weights()
    w[5] = np.nan
    net.set_weights(w)
    assert net.diverged()
    cp = scenes.cache_params(train_stride=16, train_steps=2, query_primary=True)
    s1, c1, st = g.render_cached(net, cam, p, cp)
    assert st.diverged == 1 and st.passes_redone == 0 and st.queries == 0
    assert np.array_equal(c0, c1)
    assert np.array_equal(s0.view(np.int32), s1.view(np.int32))
    net.set_weights(np.zeros(net.n_params, np.float32))
    assert not net.diverged()
    _, _, st = g.render_cached(net, cam, p, cp)
    assert st.diverged == 0 and st.queries > 0
    net.close()
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("passes", [1, 4])
def test_non_finite_predictions_drop_and_redo_the_pass(passes):
    """Weights of 1e20 are finite, so the network starts out healthy, but its
    predictions overflow.  The pass that met them is dropped from the film
    and rendered again by the plain tracer, the network is marked diverged,
    and the rest of the call and later calls run plain.  With one pass per
    call the film is the plain render bit for bit; with four, the same
    samples are added in another order (counts equal, sums within rounding)."""
    from bulbit_amd import scenes
    from bulbit_amd.nrc import Nrc

    g, cam = _readme_small()
    npix = 160 * 90
    p = scenes.render_params(4, wavefront_paths=npix if passes == 4 else 0)
    s0, c0, _ = g.render(cam, p)
    net = Nrc(seed=3)
    net.set_weights(np.full(net.n_params, 1e20, np.float32))
    assert not net.diverged()
    cp = scenes.cache_params(train_stride=4 * npix + 1, train_steps=0, query_primary=True)
    print(f"[nrc] divergence with {passes} pass(es) per call", flush=True)
    s1, c1, st = g.render_cached(net, cam, p, cp)
    assert st.diverged == 1 and st.passes_redone == 1
    assert net.diverged()
    assert np.array_equal(c0, c1)
    if passes == 1:
        assert np.array_equal(s0.view(np.int32), s1.view(np.int32))
    else:
        assert np.allclose(s1, s0, rtol=1e-5, atol=1e-6)
    s2, c2, st2 = g.render_cached(net, cam, p, cp)
    assert st2.diverged == 1 and st2.queries == 0 and st2.passes_redone == 0
    assert np.array_equal(c0, c2)
    assert np.array_equal(s0.view(np.int32), s2.view(np.int32))
    net.close()
    g.close()


@pytest.mark.gpu
def test_direct_light_at_the_query_vertex():
    """query_direct: the query vertex takes its direct light as the plain
    tracer does and the cache stands in for the indirect rest.  With every
    path training the image is the plain render bit for bit; with a zero
    cache every pixel lies between the render without the query vertex's
    direct light and the plain render."""
    from bulbit_amd import scenes
    from bulbit_amd.nrc import Nrc

    g, cam = _readme_small()
    p = scenes.render_params(4)
    s0, c0, _ = g.render(cam, p)
    net = Nrc(seed=3)
    s1, c1, st = g.render_cached(net, cam, p, scenes.cache_params(train_stride=1, train_extra_bounces=8,
                                                                   train_steps=1, query_direct=True))
    assert np.array_equal(s0.view(np.int32), s1.view(np.int32)) and st.queries == 0 and st.train_rows > 0
    net.set_weights(np.zeros(net.n_params, np.float32))
    n_paths = 160 * 90 * 4
    cp = dict(train_stride=n_paths + 1, train_steps=0, query_primary=True)
    s_all, c_all, _ = g.render_cached(net, cam, p, scenes.cache_params(**cp))
    s_dir, c_dir, st = g.render_cached(net, cam, p, scenes.cache_params(query_direct=True, **cp))
    assert np.array_equal(c0, c_dir) and st.queries > 0
    assert np.all(s_dir <= s0 * (1 + 1e-6) + 1e-7)
    assert np.all(s_dir >= s_all * (1 - 1e-6) - 1e-7)
    # the spot light reaches the sphere: direct light is most of what the sphere reflects
    assert (s_dir - s_all).sum() > 0.5 * (s0 - s_all).sum()
    net.close()
    g.close()


@pytest.mark.gpu
def test_a_later_first_query_vertex_keeps_more_of_the_path():
    """query_first: with a zero cache, a rendering path keeps what the plain
    path gathered before its query vertex, so letting the first query come
    one vertex later gives every pixel at least as much, and never more than
    the plain render (displaced icosphere, where paths meet second surfaces)."""
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene
    from bulbit_amd.nrc import Nrc

    sc, cam = scenes.icosphere_scene(width=160, height=90, level=5, n_lights=16)
    g = GpuScene(sc)
    p = scenes.render_params(4)
    s0, c0, _ = g.render(cam, p)
    net = Nrc(seed=3)
    net.set_weights(np.zeros(net.n_params, np.float32))
    n_paths = 160 * 90 * 4
    cp = dict(train_stride=n_paths + 1, train_steps=0, query_bounce=8)
    s1, c1, st1 = g.render_cached(net, cam, p, scenes.cache_params(query_first=1, **cp))
    s2, c2, st2 = g.render_cached(net, cam, p, scenes.cache_params(query_first=2, **cp))
    assert np.array_equal(c0, c1) and np.array_equal(c0, c2)
    assert st2.queries < st1.queries
    assert np.all(s2 >= s1 * (1 - 1e-6) - 1e-7)
    assert np.all(s2 <= s0 * (1 + 1e-6) + 1e-7)
    assert s1.sum() < s2.sum() < s0.sum()
    net.close()
    g.close()
