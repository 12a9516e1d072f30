"""On-disk formats of caps-sa 0.7.0 (numpy views, no GPU needed).

* Suffix-array files: packed little-endian ``u32[]`` or ``u64[]`` with no
  header (examples/caps_sa.rs:117-190).  The command-line tool writes u32 when
  the in-memory path runs on a text of at most 2^32 symbols and u64 for the
  sample-sort paths (``--ext-mem`` / ``--in-mem-ss``): :func:`sa_width`.
* SaLcp records: ``(pos, lcp)``, each little-endian, 8 bytes for u32 and 16
  bytes for u64, pos first (ext_bucket.rs:65-106).  A sorted run is stored
  with its first record's lcp set to 0 (add_soa_reset_first_lcp,
  ext_bucket.rs:203-218).
"""
from __future__ import annotations

import numpy as np

SALCP_U32 = np.dtype([("pos", "<u4"), ("lcp", "<u4")])
SALCP_U64 = np.dtype([("pos", "<u8"), ("lcp", "<u8")])


def _le(idx_bytes: int) -> np.dtype:
    if idx_bytes not in (4, 8):
        raise ValueError("index width must be 4 or 8 bytes")
    return np.dtype("<u4") if idx_bytes == 4 else np.dtype("<u8")


def sa_width(n: int, sample_sort: bool) -> int:
    """Bytes per entry of the command-line tool's output
    (examples/caps_sa.rs:117): 4 for the in-memory path on n <= 2^32, else 8."""
    return 4 if n <= (1 << 32) and not sample_sort else 8


def write_sa(path, sa, idx_bytes: int) -> None:
    np.ascontiguousarray(sa).astype(_le(idx_bytes), copy=False).tofile(path)


def read_sa(path, idx_bytes: int) -> np.ndarray:
    dt = _le(idx_bytes)
    raw = np.fromfile(path, dtype=np.uint8)
    if len(raw) % dt.itemsize:
        raise ValueError(f"{path}: size is not a multiple of the index width")
    return raw.view(dt)


def salcp_dtype(idx_bytes: int) -> np.dtype:
    return SALCP_U32 if _le(idx_bytes).itemsize == 4 else SALCP_U64


def salcp_run(pos, lcp, idx_bytes: int) -> np.ndarray:
    """Records of one sorted run, first lcp stored as 0."""
    pos = np.asarray(pos)
    lcp = np.asarray(lcp)
    if pos.shape != lcp.shape:
        raise ValueError("pos and lcp differ in length")
    rec = np.empty(len(pos), dtype=salcp_dtype(idx_bytes))
    rec["pos"] = pos
    rec["lcp"] = lcp
    if len(rec):
        rec["lcp"][0] = 0
    return rec


def write_salcp(path, runs, idx_bytes: int) -> list:
    """Writes runs [(pos, lcp), ...] back to back; returns the run boundaries
    (record offsets, first 0, last the total) as the crate's bucket keeps them
    (ext_bucket.rs:1180-1197)."""
    bounds = [0]
    with open(path, "wb") as f:
        for pos, lcp in runs:
            rec = salcp_run(pos, lcp, idx_bytes)
            f.write(rec.tobytes())
            bounds.append(bounds[-1] + len(rec))
    return bounds


def read_salcp(path, idx_bytes: int):
    """(positions, lcps) of every record (load_all_soa, ext_bucket.rs:135)."""
    dt = salcp_dtype(idx_bytes)
    raw = np.fromfile(path, dtype=np.uint8)
    if len(raw) % dt.itemsize:
        raise ValueError(f"{path}: size is not a multiple of the record size")
    rec = raw.view(dt)
    return rec["pos"].copy(), rec["lcp"].copy()
