"""Synthetic workload bytes: xorshift64 (shifts 13, 7, 17), each step's state
emitted as 8 little-endian bytes, starting from the given seed.

Not part of the reference; this is the documented generator of the
benchmark and test workloads.  The recurrence is linear over GF(2), so the
byte stream is produced in parallel lanes whose start states are jumped ahead
with powers of the 64x64 step matrix; the output equals the sequential loop.
"""
from __future__ import annotations

import math

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


class _Uniform:
    """Uniform doubles in [0, 1) from the xorshift64 stream: the top 53 bits
    of each new state."""

    def __init__(self, seed: int):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def __call__(self) -> float:
        s = self.s
        s ^= (s << 13) & 0xFFFFFFFFFFFFFFFF
        s ^= s >> 7
        s ^= (s << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = s
        return (s >> 11) / float(1 << 53)


BATCH_SLICE_SIZES = (4096, 65536, 786432)


def batch_mix(n_jobs: int = 256, seed: int = 3001, max_inputs: int = 32768):
    """The batched-create job mix: per job a slice size drawn from {4 KiB,
    64 KiB, 768 KiB}, a total source size log-uniform in [1 MiB, 1 GiB] split
    into 1-16 files (each at least one byte, sizes proportional to weights
    drawn from U[0.5, 1.5)), and recovery count ceil(U[1%, 10%] * slices),
    at least 1.  A draw with more than max_inputs slices (the PAR2 limit;
    4 KiB slices of more than 128 MiB) is drawn again.  File f of job j is
    the xorshift64 stream seeded with seed + 64 j + f + 1.

    Returns a list of dicts {slice_size, files: [(length, seed)], recovery}.
    """
    u = _Uniform(seed)
    jobs = []
    while len(jobs) < n_jobs:
        S = BATCH_SLICE_SIZES[min(2, int(3 * u()))]
        total = int(2 ** (20 + 10 * u()))
        n = -(-total // S)
        nfiles = 1 + min(15, int(16 * u()))
        frac = u()
        if n > max_inputs:
            continue
        w = [0.5 + u() for _ in range(nfiles)]
        sizes = [max(1, int(total * x / sum(w))) for x in w[:-1]]
        sizes.append(max(1, total - sum(sizes)))
        j = len(jobs)
        n = sum(-(-s // S) for s in sizes)
        if n > max_inputs:
            continue
        R = max(1, math.ceil((0.01 + 0.09 * frac) * n))
        jobs.append({"slice_size": S, "recovery": R,
                     "files": [(sz, seed + 64 * j + f + 1) for f, sz in enumerate(sizes)]})
    return jobs


def config2_files(seed: int = 2001):
    """The many-files workload: 64 files named f00.bin .. f63.bin, sizes
    floor(2^(10 + 16 u)) (log-uniform in [1 KiB, 64 MiB]); the files at the
    first 4 places of a shuffle (64 more draws, sorted) are zero-length and
    the next 8 all-zero bytes; file k otherwise holds the xorshift64 stream
    seeded with seed + 100 + k.  Returns [(name, size, kind, stream seed)]
    with kind 'random', 'zero' or 'empty'."""
    u = _Uniform(seed)
    sizes = [int(2 ** (10 + 16 * u())) for _ in range(64)]
    keys = [u() for _ in range(64)]
    order = sorted(range(64), key=lambda k: keys[k])
    empty, zero = set(order[:4]), set(order[4:12])
    out = []
    for k in range(64):
        kind = "empty" if k in empty else "zero" if k in zero else "random"
        out.append((f"f{k:02d}.bin", 0 if kind == "empty" else sizes[k], kind, seed + 100 + k))
    return out


def config2_bytes(entry) -> bytes:
    name, size, kind, sd = entry
    return xorshift64_bytes(size, sd) if kind == "random" else bytes(size)
