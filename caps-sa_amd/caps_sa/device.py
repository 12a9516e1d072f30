"""Device-pointer entry points of the C ABI over torch tensors.

torch is used for device memory and streams only.  u32/u64 index arrays
live in int32/int64 tensors (same bits); u16 text lives in an int16 tensor.
Calls run on torch's current stream.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import (CAPS_OK, MAX_CONTEXT, CapsError, LimitProvider, PlainText, TuneReport, _c_opts, _check, lib)

_IDX = {4: torch.int32, 8: torch.int64}


def _stream_ptr() -> int:
    return torch.cuda.current_stream().cuda_stream


def text_to_device(text: np.ndarray, device="cuda:0") -> torch.Tensor:
    text = np.ascontiguousarray(text)
    if text.dtype == np.uint8:
        return torch.from_numpy(text).to(device)
    if text.dtype == np.uint16:
        return torch.from_numpy(text.view(np.int16)).to(device)
    raise CapsError(4, f"symbol type {text.dtype} is outside the GPU build")


def build_device(d_text: torch.Tensor, n: int, sym_bytes: int, d_positions=None, n_pos=None, idx_bytes=4,
                 lp: LimitProvider = None, max_context=MAX_CONTEXT, router=0, gpu_buckets=0, key_geometry=0,
                 d_out=None, d_lcp=None, want_lcp=False, flags=0, tile_keys=0, batch_records=0):
    """caps_build_device: all buffers already in HBM.  Returns (sa, lcp)."""
    n_pos = n if d_positions is None else (n_pos if n_pos is not None else d_positions.numel())
    if d_out is None:
        d_out = torch.empty(n_pos, dtype=_IDX[idx_bytes], device=d_text.device)
    if want_lcp and d_lcp is None:
        d_lcp = torch.empty(n_pos, dtype=_IDX[idx_bytes], device=d_text.device)
    lim = (lp or PlainText(n))._c()
    o = _c_opts(max_context, router=router, gpu_buckets=gpu_buckets, key_geometry=key_geometry,
                want_lcp=int(d_lcp is not None), flags=flags, tile_keys=tile_keys, batch_records=batch_records)
    rc = lib().caps_build_device(d_text.data_ptr(), n, sym_bytes,
                                 d_positions.data_ptr() if d_positions is not None else None, n_pos, idx_bytes,
                                 ctypes.byref(lim), ctypes.byref(o), d_out.data_ptr(),
                                 d_lcp.data_ptr() if d_lcp is not None else None, _stream_ptr())
    _check(rc)
    return d_out, d_lcp


def build_device_noalloc(d_text: torch.Tensor, n: int, sym_bytes: int, d_positions=None, n_pos=None, idx_bytes=4,
                         lp: LimitProvider = None, max_context=MAX_CONTEXT, router=0, d_out=None, d_lcp=None,
                         flags=0):
    """caps_build_device_noalloc: build_device with every scratch buffer in
    the library's fixed device array (no device allocation by the library;
    the outputs are the caller's).  Returns (sa, lcp)."""
    n_pos = n if d_positions is None else (n_pos if n_pos is not None else d_positions.numel())
    if d_out is None:
        d_out = torch.empty(max(n_pos, 1), dtype=_IDX[idx_bytes], device=d_text.device)
    lim = (lp or PlainText(n))._c()
    o = _c_opts(max_context, router=router, want_lcp=int(d_lcp is not None), flags=flags)
    rc = lib().caps_build_device_noalloc(d_text.data_ptr(), n, sym_bytes,
                                         d_positions.data_ptr() if d_positions is not None else None, n_pos,
                                         idx_bytes, ctypes.byref(lim), ctypes.byref(o), d_out.data_ptr(),
                                         d_lcp.data_ptr() if d_lcp is not None else None, _stream_ptr())
    _check(rc)
    return d_out, d_lcp


def build_device_filtered(d_text: torch.Tensor, n: int, sym_bytes: int, d_keep: torch.Tensor, capacity: int,
                          idx_bytes=8, lp: LimitProvider = None, max_context=MAX_CONTEXT, router=0, d_out=None,
                          d_lcp=None, want_lcp=False, flags=0, tile_keys=0, batch_records=0, gpu_buckets=0):
    """caps_build_device_filtered: text and kept-position bitmap ((n + 63) // 64
    words, as keep_bitmap builds) in HBM.  Returns (sa, lcp, kept count); the
    outputs hold `capacity` entries, of which the first `kept` are filled."""
    if d_out is None:
        d_out = torch.empty(max(capacity, 1), dtype=_IDX[idx_bytes], device=d_text.device)
    if want_lcp and d_lcp is None:
        d_lcp = torch.empty(max(capacity, 1), dtype=_IDX[idx_bytes], device=d_text.device)
    lim = (lp or PlainText(n))._c()
    o = _c_opts(max_context, router=router, want_lcp=int(d_lcp is not None), flags=flags, tile_keys=tile_keys,
                batch_records=batch_records, gpu_buckets=gpu_buckets)
    kept = ctypes.c_uint64(0)
    rc = lib().caps_build_device_filtered(d_text.data_ptr(), n, sym_bytes, d_keep.data_ptr(), idx_bytes,
                                          ctypes.byref(lim), ctypes.byref(o), d_out.data_ptr(), capacity,
                                          ctypes.byref(kept), d_lcp.data_ptr() if d_lcp is not None else None,
                                          _stream_ptr())
    _check(rc)
    return d_out, d_lcp, kept.value


def lcp_pairs(d_text: torch.Tensor, n: int, sym_bytes: int, pairs: np.ndarray, max_context=MAX_CONTEXT):
    d_pairs = torch.from_numpy(np.ascontiguousarray(pairs, dtype=np.uint64).view(np.int64)).to(d_text.device)
    count = len(pairs) // 2
    out = torch.empty(max(count, 1), dtype=torch.int64, device=d_text.device)
    _check(lib().caps_lcp_pairs_device(d_text.data_ptr(), n, sym_bytes, d_pairs.data_ptr(), count,
                                       min(max_context, MAX_CONTEXT), out.data_ptr(), _stream_ptr()))
    return out.cpu().numpy().view(np.uint64)[:count]


def suffix_cmp_pairs(d_text: torch.Tensor, n: int, sym_bytes: int, pairs: np.ndarray, lp=None,
                     max_context=MAX_CONTEXT):
    d_pairs = torch.from_numpy(np.ascontiguousarray(pairs, dtype=np.uint64).view(np.int64)).to(d_text.device)
    count = len(pairs) // 2
    out = torch.empty(max(count, 1), dtype=torch.int32, device=d_text.device)
    lim = (lp or PlainText(n))._c()
    _check(lib().caps_suffix_cmp_pairs_device(d_text.data_ptr(), n, sym_bytes, ctypes.byref(lim),
                                              d_pairs.data_ptr(), count, min(max_context, MAX_CONTEXT),
                                              out.data_ptr(), _stream_ptr()))
    return out.cpu().numpy()[:count]


def seed_keys(d_text: torch.Tensor, n: int, sym_bytes: int, positions: np.ndarray, lp=None, geometry=1, rank=0):
    d_pos = torch.from_numpy(np.ascontiguousarray(positions, dtype=np.uint64).view(np.int64)).to(d_text.device)
    count = len(positions)
    keys = torch.empty(max(count, 1), dtype=torch.int64, device=d_text.device)
    geo = (ctypes.c_uint32 * 2)()
    lim = (lp or PlainText(n))._c()
    _check(lib().caps_seed_keys_device(d_text.data_ptr(), n, sym_bytes, ctypes.byref(lim), geometry, rank,
                                       d_pos.data_ptr(), count, keys.data_ptr(), geo, _stream_ptr()))
    return keys.cpu().numpy().view(np.uint64)[:count], (geo[0], geo[1])


def seed_params(d_text: torch.Tensor, n: int, sym_bytes: int, policy=0, remap_budget=0, max_context=MAX_CONTEXT,
                lp=None):
    """caps_seed_params_device: the reference's packed-seed decision
    (ext_mem.rs:1495-1513, pack.rs:217-234) as a dict."""
    from . import SeedParams

    out = SeedParams()
    lim = (lp or PlainText(n))._c()
    o = _c_opts(max_context, packed_prefix_seed=policy, remap_budget=remap_budget)
    _check(lib().caps_seed_params_device(d_text.data_ptr(), n, sym_bytes, ctypes.byref(lim), ctypes.byref(o),
                                         ctypes.byref(out), _stream_ptr()))
    return {name: getattr(out, name) for name, _ in out._fields_}


def pwl_fit(keys: np.ndarray, labels: np.ndarray, ns: int, n_buckets: int, device="cuda:0"):
    """caps_router_pwl_fit_device over sorted keys and their labels: dict of
    knots, a, b, inv, lo, hi (numpy arrays)."""
    k = torch.from_numpy(np.ascontiguousarray(keys, dtype=np.uint64).view(np.int64)).to(device)
    lab = torch.from_numpy(np.ascontiguousarray(labels, dtype=np.uint32).view(np.int32)).to(device)
    out = {"knots": np.zeros(ns + 1, dtype=np.uint64), "a": np.zeros(ns, dtype=np.float32),
           "b": np.zeros(ns, dtype=np.float32), "inv": np.zeros(ns, dtype=np.float32),
           "lo": np.zeros(ns, dtype=np.uint32), "hi": np.zeros(ns, dtype=np.uint32)}
    _check(lib().caps_router_pwl_fit_device(k.data_ptr(), lab.data_ptr(), len(keys), ns, n_buckets,
                                            *(out[x].ctypes.data for x in ("knots", "a", "b", "inv", "lo", "hi")),
                                            _stream_ptr()))
    return out


def spline_fit(splitters: np.ndarray, device="cuda:0"):
    """caps_router_spline_fit_device over a sorted splitter table: (knots
    (4,097 u16), (min, max) of predicted - true position over every key)."""
    n = len(splitters)
    k = torch.from_numpy(np.ascontiguousarray(splitters, dtype=np.uint64).view(np.int64)).to(device) if n else None
    knots = np.zeros(SPLINE_CELLS + 1, dtype=np.uint16)
    bound = np.zeros(2, dtype=np.int32)
    _check(lib().caps_router_spline_fit_device(k.data_ptr() if n else None, n, knots.ctypes.data,
                                               bound.ctypes.data, _stream_ptr()))
    return knots, (int(bound[0]), int(bound[1]))


SPLINE_CELLS = 4096

MLP_IN, MLP_HID = 64, 32
MLP_PARAMS = MLP_IN * MLP_HID + MLP_HID + MLP_HID + 1


def mlp_grad(params: np.ndarray, keys: np.ndarray, targets: np.ndarray, device="cuda:0"):
    """Router MLP mean-squared-error gradient on the GPU (caps_router_mlp_grad_device)."""
    p = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).to(device)
    k = torch.from_numpy(np.ascontiguousarray(keys, dtype=np.uint64).view(np.int64)).to(device)
    y = torch.from_numpy(np.ascontiguousarray(targets, dtype=np.float32)).to(device)
    g = torch.empty(MLP_PARAMS, dtype=torch.float32, device=device)
    loss = torch.empty(1, dtype=torch.float32, device=device)
    _check(lib().caps_router_mlp_grad_device(p.data_ptr(), k.data_ptr(), y.data_ptr(), len(keys), g.data_ptr(),
                                             loss.data_ptr(), _stream_ptr()))
    return g.cpu().numpy(), float(loss.cpu().numpy()[0])


def mlp_forward(params: np.ndarray, keys: np.ndarray, device="cuda:0"):
    """Router MLP output through the nibble tables used by classification."""
    p = torch.from_numpy(np.ascontiguousarray(params, dtype=np.float32)).to(device)
    k = torch.from_numpy(np.ascontiguousarray(keys, dtype=np.uint64).view(np.int64)).to(device)
    out = torch.empty(max(len(keys), 1), dtype=torch.float32, device=device)
    _check(lib().caps_router_mlp_forward_device(p.data_ptr(), k.data_ptr(), len(keys), out.data_ptr(),
                                                _stream_ptr()))
    return out.cpu().numpy()[: len(keys)]


def build_batch_device(d_text: torch.Tensor, n: int, sym_bytes: int, ends: np.ndarray, idx_bytes=4, d_out=None,
                       router=0, gpu_buckets=0, tile_keys=0, flags=0):
    """caps_build_batch_device: texts back to back in HBM, `ends` their
    cumulative ends.  Returns the global positions, text after text."""
    ends = np.ascontiguousarray(ends, dtype=np.uint64)
    if d_out is None:
        d_out = torch.empty(n, dtype=_IDX[idx_bytes], device=d_text.device)
    o = _c_opts(router=router, gpu_buckets=gpu_buckets, tile_keys=tile_keys, flags=flags)
    _check(lib().caps_build_batch_device(d_text.data_ptr(), n, sym_bytes, ends.ctypes.data, len(ends), idx_bytes,
                                         ctypes.byref(o), d_out.data_ptr(), _stream_ptr()))
    return d_out


def autotune(d_text: torch.Tensor, n: int, sym_bytes: int, lp: LimitProvider = None, idx_bytes=4, reps=2,
             router=0):
    """caps_autotune_device: returns ({"tile_keys", "flags"}, report dict)."""
    scratch = torch.empty(n, dtype=_IDX[idx_bytes], device=d_text.device)
    lim = (lp or PlainText(n))._c()
    o = _c_opts(router=router)
    rep = TuneReport()
    _check(lib().caps_autotune_device(d_text.data_ptr(), n, sym_bytes, ctypes.byref(lim), idx_bytes,
                                      scratch.data_ptr(), reps, ctypes.byref(o), _stream_ptr(), ctypes.byref(rep)))
    return {"tile_keys": o.tile_keys, "flags": o.flags}, rep.as_dict()


def autotune_batch(d_text: torch.Tensor, n: int, sym_bytes: int, ends: np.ndarray, idx_bytes=4, reps=2, router=0):
    """caps_autotune_batch_device: returns ({"tile_keys", "flags"}, report dict)."""
    ends = np.ascontiguousarray(ends, dtype=np.uint64)
    scratch = torch.empty(n, dtype=_IDX[idx_bytes], device=d_text.device)
    o = _c_opts(router=router)
    rep = TuneReport()
    _check(lib().caps_autotune_batch_device(d_text.data_ptr(), n, sym_bytes, ends.ctypes.data, len(ends), idx_bytes,
                                            scratch.data_ptr(), reps, ctypes.byref(o), _stream_ptr(),
                                            ctypes.byref(rep)))
    return {"tile_keys": o.tile_keys, "flags": o.flags}, rep.as_dict()
