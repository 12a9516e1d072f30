"""Learned bucket router: gradients of the GPU training kernels against a
float64 NumPy gradient (itself pinned by central finite differences), the
nibble-table forward pass used by classification against a NumPy forward
pass, bit-exact builds with every router, and what training gains."""
import numpy as np
import pytest

import caps_sa
import oracle as O
from caps_sa import device as dev
from caps_sa import gen
from mlp_ref import np_forward, np_grad, np_loss

pytestmark = pytest.mark.gpu
IN, HID = dev.MLP_IN, dev.MLP_HID


def sample(seed, count=1000):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(dev.MLP_PARAMS) * 0.2).astype(np.float32)
    keys = rng.integers(0, 2 ** 63, count, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, count).astype(np.uint64)
    targets = rng.random(count).astype(np.float32)
    return params, keys, targets


@pytest.mark.parametrize("seed", [3, 4])
def test_gpu_gradient_matches_numpy(cuda, seed):
    params, keys, targets = sample(seed, 1000)
    g_gpu, loss_gpu = dev.mlp_grad(params, keys, targets)
    g_ref = np_grad(params.astype(np.float64), keys, targets.astype(np.float64))
    loss_ref = np_loss(params.astype(np.float64), keys, targets.astype(np.float64))
    # float32 accumulation over 1000 samples: relative tolerance 1e-4 of the scale
    scale = np.max(np.abs(g_ref))
    assert np.max(np.abs(g_gpu - g_ref)) <= 1e-4 * scale + 1e-7
    assert abs(loss_gpu - loss_ref) <= 1e-4 * abs(loss_ref)


def test_nibble_forward_matches_numpy(cuda):
    params, keys, _ = sample(5, 3000)
    got = dev.mlp_forward(params, keys)
    want, _, _ = np_forward(params.astype(np.float64), keys)
    assert np.max(np.abs(got - want)) <= 1e-4 * (1 + np.max(np.abs(want)))


def _build(text, router, buckets, flags=0):
    d_text = dev.text_to_device(text)
    sa, _ = dev.build_device(d_text, len(text), 1, idx_bytes=4, router=router, gpu_buckets=buckets, flags=flags)
    st = caps_sa.last_stats()
    return sa.cpu().numpy().view(np.uint32).astype(np.uint64), st


@pytest.mark.parametrize("buckets", [0, 4096])
def test_router_builds_are_bit_exact(cuda, buckets):
    text = gen.markov_repeats(1 << 20, seed=0xC3, mean_len=2000.0)
    ctx = O.Ctx(text)
    want = ctx.build_in_memory(threads=8)
    for router in (caps_sa.Router.BSEARCH, caps_sa.Router.PWL, caps_sa.Router.MLP, caps_sa.Router.SPLINE):
        got, st = _build(text, int(router), buckets)
        print(f"router={router.name} buckets={st['buckets']} err_sum={st['router_err_sum']} "
              f"err_max={st['router_err_max']} fallbacks={st['router_fallbacks']} count={st['router_count']}")
        assert np.array_equal(got, want), router
        if router != caps_sa.Router.BSEARCH:
            assert st["router_count"] == len(text)
        if router == caps_sa.Router.SPLINE:
            # every suffix's key fell inside the window the bound guarantees
            assert st["router_fallbacks"] == 0
            assert st["router_err_max"] <= max(-st["router_bound_lo"], st["router_bound_hi"])


def test_spline_without_bound_falls_back_exactly(cuda):
    """The spline router with a window of one position (no error bound):
    keys it mispredicts land on the widened window's ends, are counted as
    fallbacks, and the host classifies the input again by binary search;
    the suffix array is unchanged."""
    text = gen.markov_repeats(1 << 20, seed=0xC5, mean_len=2000.0)
    want = O.Ctx(text).build_in_memory(threads=8)
    for buckets in (0, 4096):
        got, st = _build(text, int(caps_sa.Router.SPLINE), buckets, caps_sa.FLAG_ROUTER_UNTRAINED)
        assert np.array_equal(got, want)
        assert st["router_fallbacks"] > 0
        got, st = _build(text, int(caps_sa.Router.SPLINE), buckets)
        assert np.array_equal(got, want) and st["router_fallbacks"] == 0


def test_training_reduces_misprediction(cuda):
    """What the trained network gains over the same network untrained: the
    mean distance between predicted and true bucket."""
    text = gen.markov_repeats(1 << 21, seed=0xC4, mean_len=2000.0)
    res = {}
    for name, router, flags in (("mlp", 2, 0), ("mlp-untrained", 2, caps_sa.FLAG_ROUTER_UNTRAINED), ("pwl", 1, 0)):
        _, st = _build(text, router, 4096, flags)
        res[name] = st["router_err_sum"] / st["router_count"]
        print(name, res[name], st["router_err_max"], st["router_fallbacks"])
    assert res["mlp"] < 0.5 * res["mlp-untrained"]
    assert res["pwl"] < res["mlp-untrained"]


def test_pwl_fit_matches_least_squares(cuda):
    """The GPU piecewise-linear fit against numpy.polyfit (degree 1) on each
    segment's own points, in float64: slopes and intercepts agree to the
    float32 rounding of x; a segment whose labels fall is flattened to its
    mean (the monotone clamp); a segment of one label is flat; the knots and
    label bounds are the segment's first sample key and labels."""
    rng = np.random.default_rng(12)
    s, ns, nb = 40_000, 64, 4096
    keys = np.sort(rng.integers(0, 1 << 63, s, dtype=np.uint64) * np.uint64(2))
    keys[1000:1600] = keys[1000]  # a run of equal keys
    # labels: a monotone step function of the key (bucket of a splitter table)
    spl = np.sort(rng.choice(keys, nb - 1, replace=False))
    labels = np.searchsorted(spl, keys, side="right").astype(np.uint32)
    got = dev.pwl_fit(keys, labels, ns, nb)
    for j in range(ns):
        i0, i1 = j * s // ns, (j + 1) * s // ns
        k0 = int(keys[i0])
        k1 = int(keys[i1]) if j + 1 < ns else int(keys[s - 1])
        span = float(k1 - k0) if k1 > k0 else 1.0
        inv = np.float32(1.0 / span)
        x = ((keys[i0:i1] - np.uint64(k0)).astype(np.float32) * inv).astype(np.float64)
        y = labels[i0:i1].astype(np.float64)
        assert got["knots"][j] == k0
        assert got["lo"][j] == labels[i0] and got["hi"][j] == (labels[i1] if j + 1 < ns else nb - 1)
        assert got["inv"][j] == inv
        if np.ptp(x) > 0:
            a, b = np.polyfit(x, y, 1)
            if a < 0:
                a, b = 0.0, y.mean()
        else:
            a, b = 0.0, y.mean()
        scale = max(1.0, abs(y).max())
        assert abs(got["a"][j] - a) <= 2e-3 * scale + 1e-3 * abs(a), (j, got["a"][j], a)
        assert abs(got["b"][j] - b) <= 2e-3 * scale, (j, got["b"][j], b)
    assert got["knots"][ns] == np.uint64(0xFFFFFFFFFFFFFFFF)


def _spline_ref(spl):
    """Knots by numpy.searchsorted, and the prediction of the spline."""
    spl = np.asarray(spl, dtype=np.uint64)
    knots = np.concatenate([np.searchsorted(spl, np.arange(4096, dtype=np.uint64) << np.uint64(52), side="left"),
                            [len(spl)]]).astype(np.int64)

    def predict(keys):
        t = (keys >> np.uint64(52)).astype(np.int64)
        f = ((keys >> np.uint64(36)) & np.uint64(0xFFFF)).astype(np.int64)
        return knots[t] + ((f * (knots[t + 1] - knots[t])) >> 16)

    return knots, predict


@pytest.mark.parametrize("case", ["uniform", "clustered", "duplicates", "extremes", "one", "empty"])
def test_spline_fit_and_bound(cuda, case):
    """The spline's knots against numpy.searchsorted, and its error bound
    against the errors of a million random keys plus every splitter, the
    keys next to them and every cell edge: no key's error leaves the bound,
    and some key reaches each end of it."""
    rng = np.random.default_rng(21)
    top = np.uint64(0xFFFFFFFFFFFFFFFF)
    if case == "uniform":
        spl = rng.integers(0, 2 ** 63, 4095, dtype=np.uint64) * np.uint64(2)
    elif case == "clustered":  # most splitters inside a few cells (a skewed key distribution)
        base = rng.integers(0, 4096, 6, dtype=np.uint64) << np.uint64(52)
        spl = (base[rng.integers(0, 6, 4000)] + rng.integers(0, 1 << 52, 4000, dtype=np.uint64))
        spl = np.concatenate([spl, rng.integers(0, 2 ** 63, 95, dtype=np.uint64) * np.uint64(2)])
    elif case == "duplicates":
        spl = np.repeat(rng.integers(0, 2 ** 63, 300, dtype=np.uint64) * np.uint64(2), 7)
    elif case == "extremes":
        spl = np.array([0, 0, 1, (1 << 52) - 1, 1 << 52, int(top) - 1, int(top), int(top)], dtype=np.uint64)
    elif case == "one":
        spl = np.array([12345 << 40], dtype=np.uint64)
    else:
        spl = np.zeros(0, dtype=np.uint64)
    spl = np.sort(spl)
    knots, (lo, hi) = dev.spline_fit(spl)
    want_knots, predict = _spline_ref(spl)
    assert np.array_equal(knots.astype(np.int64), want_knots)
    probes = [rng.integers(0, 2 ** 63, 1 << 20, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, 1 << 20).astype(np.uint64),
              spl, spl - np.uint64(1), spl + np.uint64(1),
              np.arange(4096, dtype=np.uint64) << np.uint64(52),
              (np.arange(4096, dtype=np.uint64) << np.uint64(52)) + np.uint64((1 << 52) - 1)]
    keys = np.concatenate(probes)  # wrap-around of spl - 1 / spl + 1 at the ends is still a valid key
    err = predict(keys) - np.searchsorted(spl, keys, side="right").astype(np.int64)
    assert lo <= 0 <= hi or len(spl) == 0
    assert err.min() >= lo and err.max() <= hi
    assert err.min() == lo and err.max() == hi
