"""Alias-table light selection (BB_LIGHTS_ALIAS), the opt-in fast mode beside
PowerLightSampler's CDF search (src/light_sampler/power_light_sampler.cpp:8-47).

CPU: the table built by bbgpu_alias_table reproduces the power PMF exactly
(the probability of each light summed over its own cell and the cells that
alias to it), and the kernel's draw-to-light map, restated in float32,
selects with those frequencies (chi-square).  GPU: a render in alias mode
agrees in expectation with the CDF render."""
import numpy as np
import pytest


def _table(bblib, w):
    w = np.ascontiguousarray(w, np.float32)
    n = len(w)
    prob = np.zeros(n, np.float32)
    alias = np.zeros(n, np.int32)
    assert bblib.bbgpu_alias_table(w.ctypes.data, n, prob.ctypes.data, alias.ctypes.data) == 0
    return prob, alias


def _implied(prob, alias):
    n = len(prob)
    p = prob.astype(np.float64).copy()
    np.add.at(p, alias, 1.0 - prob.astype(np.float64))
    return p / n


WEIGHTS = {
    "one": np.array([3.0]),
    "uniform": np.ones(17),
    "zeros_inside": np.array([0.0, 2.0, 0.0, 5.0, 1.0, 0.0]),
    "dominant": np.array([1e6, 1.0, 1.0, 1.0]),
    "skewed": np.array([1e3, 1.0, 2.0, 1.0, 0.5]),
    "lognormal": np.exp(np.random.default_rng(5).normal(0.0, 1.5, 1024)),
    "all_zero": np.zeros(5),
}


@pytest.mark.parametrize("name", list(WEIGHTS))
def test_alias_table_reproduces_the_power_pmf(bblib, name):
    w = WEIGHTS[name].astype(np.float32)
    prob, alias = _table(bblib, w)
    assert np.all((prob >= 0) & (prob <= 1)) and np.all((alias >= 0) & (alias < len(w)))
    want = np.full(len(w), 1.0 / len(w)) if w.sum() == 0 else w.astype(np.float64) / w.astype(np.float64).sum()
    got = _implied(prob, alias)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-9)
    assert np.all(got[want == 0] == 0)


def _select(u, prob, alias):
    """the kernel's map (light_select_alias) in float32"""
    n = len(prob)
    x = (u * np.float32(n)).astype(np.float32)
    i = np.clip(x.astype(np.int32), 0, n - 1)
    keep = (x - i.astype(np.float32)) < prob[i]
    return np.where(keep, i, alias[i])


@pytest.mark.parametrize("name", ["zeros_inside", "skewed", "lognormal"])
def test_alias_draws_select_with_the_pmf(bblib, name):
    from scipy.stats import chisquare

    w = WEIGHTS[name].astype(np.float32)
    prob, alias = _table(bblib, w)
    rng = np.random.default_rng(9)
    # float32 uniforms as the sampler makes them: 24-bit multiples of 2^-24
    u = (rng.integers(0, 1 << 24, 4_000_000) * np.float32(2.0 ** -24)).astype(np.float32)
    counts = np.bincount(_select(u, prob, alias), minlength=len(w)).astype(np.float64)
    p = w.astype(np.float64) / w.astype(np.float64).sum()
    assert counts[p == 0].sum() == 0
    keep = p * len(u) >= 5
    exp = p[keep] * len(u)
    obs = counts[keep]
    stat, pval = chisquare(obs, exp * obs.sum() / exp.sum())
    print(f"[alias] {name}: chi-square {stat:.1f} over {keep.sum()} cells, p = {pval:.3f}")
    assert pval > 1e-3


@pytest.mark.parametrize("name", ["zeros_inside", "skewed", "lognormal", "uniform"])
def test_alias_map_over_every_float_draw(bblib, name):
    """Every one of the 2^24 float32 uniforms the sampler can produce, put
    through the kernel's map: each light's share of them is its PMF to
    within the grid (a few 2^-24 steps per boundary)."""
    w = WEIGHTS[name].astype(np.float32)
    prob, alias = _table(bblib, w)
    u = (np.arange(1 << 24, dtype=np.int64) * (2.0 ** -24)).astype(np.float32)
    share = np.bincount(_select(u, prob, alias), minlength=len(w)) / float(1 << 24)
    p = w.astype(np.float64) / w.astype(np.float64).sum()
    assert np.all(share[p == 0] == 0)
    # a light's share is made of its own cell's kept part and the alias parts
    # of the cells that point to it: a few grid steps of rounding per part
    parts = 1 + np.bincount(alias[prob < 1], minlength=len(w))
    assert np.all(np.abs(share - p) <= 4.0 * parts * 2.0 ** -24 + 1e-7 * p)


def test_alias_table_rejects_bad_arguments(bblib):
    assert bblib.bbgpu_alias_table(None, 3, None, None) == 1  # BBGPU_ERR_INVALID
    assert bblib.bbgpu_alias_table(None, 0, None, None) == 0


@pytest.mark.gpu
def test_alias_render_agrees_in_expectation():
    """Displaced icosphere under 64 area lights of uneven power at 160x90:
    256 spp with alias selection and 256 spp with the CDF (disjoint
    samples) against a 4096-spp CDF reference: same mean and relMSE within
    noise, every sample counted."""
    from bulbit_amd import scenes
    from bulbit_amd.gpu import GpuScene

    sc, cam = scenes.icosphere_scene(width=160, height=90, level=5, n_lights=64)
    g = GpuScene(sc)
    rs, rc, _ = g.render(cam, scenes.render_params(4096, sample_begin=100000))
    ref = rs / rc[..., None]
    s0, c0, _ = g.render(cam, scenes.render_params(256, sample_begin=0))
    s1, c1, _ = g.render(cam, scenes.render_params(256, sample_begin=5000, light_sampler=scenes.LIGHTS_ALIAS))
    assert np.all(c0 == 256) and np.all(c1 == 256)
    a, b = s0 / c0[..., None], s1 / c1[..., None]

    def relmse(x):
        return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))

    e0, e1 = relmse(a), relmse(b)
    print(f"[alias] relMSE at 256 spp: CDF {e0:.6f}, alias {e1:.6f}; means {a.mean():.5f} {b.mean():.5f} "
          f"ref {ref.mean():.5f}", flush=True)
    assert abs(b.mean() - ref.mean()) < 0.01 * ref.mean()
    assert 0.6 * e0 < e1 < 1.6 * e0
    g.close()
