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
    for router in (caps_sa.Router.BSEARCH, caps_sa.Router.PWL, caps_sa.Router.MLP):
        got, st = _build(text, int(router), buckets)
        print(f"router={router.name} buckets={st['buckets']} err_sum={st['router_err_sum']} "
              f"err_max={st['router_err_max']} fallbacks={st['router_fallbacks']} count={st['router_count']}")
        assert np.array_equal(got, want), router
        if router != caps_sa.Router.BSEARCH:
            assert st["router_count"] == len(text)


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
