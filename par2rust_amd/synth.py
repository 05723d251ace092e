"""Synthetic workload bytes: xorshift64 (shifts 13, 7, 17), each step's state
emitted as 8 little-endian bytes, starting from the given seed.

Not part of the reference; this is the documented generator of the
benchmark and test workloads.  The recurrence is linear over GF(2), so the
byte stream is produced in parallel lanes whose start states are jumped ahead
with powers of the 64x64 step matrix; the output equals the sequential loop.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _step(s: np.ndarray) -> np.ndarray:
    s = s ^ ((s << np.uint64(13)) & _M64)
    s = s ^ (s >> np.uint64(7))
    s = s ^ ((s << np.uint64(17)) & _M64)
    return s


def _apply(cols: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Linear map given by its 64 column images, applied to every element of v."""
    out = np.zeros_like(v)
    for b in range(64):
        bit = (v >> np.uint64(b)) & np.uint64(1)
        out ^= cols[b] * bit
    return out


def _compose(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Columns of a o b."""
    return _apply(a, b)


def _power(cols: np.ndarray, e: int) -> np.ndarray:
    result = np.array([np.uint64(1) << np.uint64(b) for b in range(64)], dtype=np.uint64)
    base = cols.copy()
    while e:
        if e & 1:
            result = _compose(base, result)
        base = _compose(base, base)
        e >>= 1
    return result


def xorshift64_bytes(n: int, seed: int, lanes: int = 1 << 14) -> bytes:
    words = (n + 7) // 8
    if words == 0:
        return b""
    lanes = max(1, min(lanes, words))
    per_lane = (words + lanes - 1) // lanes
    ident = np.array([np.uint64(1) << np.uint64(b) for b in range(64)], dtype=np.uint64)
    step_cols = _step(ident)
    jump = _power(step_cols, per_lane)
    # lane start states: jump^j (seed), built by doubling
    starts = np.array([seed & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
    j = jump
    while starts.size < lanes:
        starts = np.concatenate([starts, _apply(j, starts)])
        j = _compose(j, j)
    starts = starts[:lanes]
    out = np.empty((per_lane, lanes), dtype=np.uint64)
    s = starts
    for k in range(per_lane):
        s = _step(s)
        out[k] = s
    stream = out.T.reshape(-1)[:words]  # lane-major: lane j holds words [j*L, (j+1)*L)
    return stream.astype("<u8").tobytes()[:n]
