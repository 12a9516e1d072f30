"""CPU pin of the router MLP reference: its analytic float64 gradient agrees
with central finite differences of its loss."""
import numpy as np

from mlp_ref import HID, IN, N_PARAMS, np_forward, np_grad, np_loss


def sample(seed, count=1000):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(N_PARAMS) * 0.2).astype(np.float32)
    keys = rng.integers(0, 2 ** 63, count, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, count).astype(np.uint64)
    targets = rng.random(count).astype(np.float32)
    return params, keys, targets


def test_numpy_gradient_matches_finite_differences():
    params, keys, targets = sample(1, 200)
    g = np_grad(params.astype(np.float64), keys, targets.astype(np.float64))
    _, pre, _ = np_forward(params.astype(np.float64), keys)
    rng = np.random.default_rng(2)
    eps = 1e-6
    checked = 0
    for idx in rng.choice(N_PARAMS, 60, replace=False):
        if idx < IN * HID + HID:  # skip first-layer entries whose units sit on a relu kink
            j = idx % HID if idx < IN * HID else idx - IN * HID
            if np.min(np.abs(pre[:, j])) < 1e-3:
                continue
        p1 = params.astype(np.float64).copy()
        p2 = p1.copy()
        p1[idx] += eps
        p2[idx] -= eps
        fd = (np_loss(p1, keys, targets) - np_loss(p2, keys, targets)) / (2 * eps)
        assert abs(fd - g[idx]) <= 1e-5 + 1e-4 * abs(fd), (idx, fd, g[idx])
        checked += 1
    assert checked > 20


