"""Float64 NumPy reference of the router MLP (64 key bits -> 32 ReLU -> 1,
mean squared error), used to check the GPU training kernels."""
import numpy as np

IN, HID = 64, 32
N_PARAMS = IN * HID + HID + HID + 1


def unpack(params):
    p = params.astype(np.float64)
    W1 = p[: IN * HID].reshape(IN, HID)
    b1 = p[IN * HID: IN * HID + HID]
    w2 = p[IN * HID + HID: IN * HID + 2 * HID]
    b2 = p[-1]
    return W1, b1, w2, b2


def bits_of(keys):
    # input i = key bit 63 - i
    k = keys.astype(np.uint64)
    return ((k[:, None] >> (63 - np.arange(64, dtype=np.uint64))[None, :]) & np.uint64(1)).astype(np.float64)


def np_forward(params, keys):
    W1, b1, w2, b2 = unpack(params)
    X = bits_of(keys)
    pre = X @ W1 + b1
    return np.maximum(pre, 0) @ w2 + b2, pre, X


def np_loss(params, keys, targets):
    out, _, _ = np_forward(params, keys)
    return np.mean((out - targets) ** 2)


def np_grad(params, keys, targets):
    W1, b1, w2, b2 = unpack(params)
    out, pre, X = np_forward(params, keys)
    d = 2.0 * (out - targets) / len(keys)
    h = np.maximum(pre, 0)
    dpre = (d[:, None] * w2[None, :]) * (pre > 0)
    gW1 = X.T @ dpre
    gb1 = dpre.sum(0)
    gw2 = h.T @ d
    gb2 = d.sum()
    return np.concatenate([gW1.ravel(), gb1, gw2, [gb2]])


