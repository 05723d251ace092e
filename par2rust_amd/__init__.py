"""ctypes binding of libpar2gpu.so, the MI355X PAR2 create engine.

Mirrors the reference's library surface (javi11/par2rust src/lib.rs:42-61):
``CreateOptions``, ``VolumeScheme``, ``run_create`` and ``SourceFile``-style
scanning, each a call through the C ABI declared in include/par2gpu.h.  The
library must have been built (python par2rust_amd/build.py); importing this
package without it raises, and there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import struct
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpar2gpu.so")

# status codes, include/par2gpu.h
P2G_OK = 0
ERRORS = {
    1: "Io",
    2: "InvalidSliceSize",
    3: "InvalidFileName",
    4: "NoInputFiles",
    5: "EmptyFile",
    6: "TooManyFiles",
    7: "TooManyRecoveryBlocks",
    8: "TooManyInputBlocks",
    9: "InvalidVolumeScheme",
    10: "InvalidArgument",
    11: "Device",
    12: "OutOfMemory",
}

SCHEME_SINGLE, SCHEME_EXPONENTIAL, SCHEME_UNIFORM, SCHEME_EXPLICIT = 0, 1, 2, 3

EXPORTED_SYMBOLS = [
    "p2g_open", "p2g_close", "p2g_last_error", "p2g_version",
    "p2g_slice_checksums", "p2g_encode_begin", "p2g_encode_push", "p2g_encode_finish",
    "p2g_encode_free", "p2g_recovery_packet_md5", "p2g_dev_encode",
    "p2g_dev_slice_checksums", "p2g_dev_recovery_packet_md5", "p2g_dev_create_pass",
    "p2g_pass_timing", "p2g_synchronize", "p2g_run_create",
    "p2g_batch_plan", "p2g_batch_launch", "p2g_batch_wait", "p2g_batch_job_ms",
    "p2g_batch_free", "p2g_dev_create_batch", "p2g_dev_fill_xorshift64",
    "p2g_server_open", "p2g_server_submit", "p2g_server_poll", "p2g_server_result",
    "p2g_server_release", "p2g_server_stats", "p2g_server_close", "p2g_server_set_bypass",
    "p2g_tune_bucket", "p2g_autotune", "p2g_tune_get", "p2g_tune_save", "p2g_tune_load",
    "p2g_plan_roles", "p2g_dev_create_pass_noalloc", "p2g_tune_get_encode",
]

VARIANT_MFMA, VARIANT_BITSLICE, VARIANT_NIBBLE = 0, 1, 2


class Par2Error(RuntimeError):
    """A failed p2g call; ``kind`` names the Par2Error variant (src/error.rs:7-41)."""

    def __init__(self, code: int, message: str = ""):
        self.code = code
        self.kind = ERRORS.get(code, f"Unknown({code})")
        super().__init__(f"{self.kind}: {message}" if message else self.kind)


class _DevFile(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("len", ctypes.c_uint64)]


class _BatchJob(ctypes.Structure):
    """p2g_batch_job."""
    _fields_ = [
        ("files", ctypes.c_void_p),
        ("slice_size", ctypes.c_uint64),
        ("rec", ctypes.c_void_p),
        ("slice_md5", ctypes.c_void_p),
        ("slice_crc", ctypes.c_void_p),
        ("pkt_md5", ctypes.c_void_p),
        ("nfiles", ctypes.c_uint32),
        ("recovery_count", ctypes.c_uint32),
        ("set_id", ctypes.c_uint8 * 16),
    ]


class _Request(ctypes.Structure):
    """p2g_request."""
    _fields_ = [
        ("file_lens", ctypes.POINTER(ctypes.c_uint64)),
        ("host_data", ctypes.POINTER(ctypes.c_void_p)),
        ("fill_seeds", ctypes.POINTER(ctypes.c_uint64)),
        ("slice_size", ctypes.c_uint64),
        ("nfiles", ctypes.c_uint32),
        ("recovery_count", ctypes.c_uint32),
        ("set_id", ctypes.c_uint8 * 16),
    ]


class _Scheme(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("uniform_count", ctypes.c_uint32),
        ("limited", ctypes.c_int),
        ("limit_cap", ctypes.c_uint32),
        ("explicit_sizes", ctypes.POINTER(ctypes.c_uint32)),
        ("n_explicit", ctypes.c_uint32),
        ("outer_caps", ctypes.POINTER(ctypes.c_uint32)),
        ("n_outer_caps", ctypes.c_uint32),
    ]


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing; build it with `python par2rust_amd/build.py`")
    lib = ctypes.CDLL(LIB_PATH)
    P, u8p, u32p = ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint32)
    sig = {
        "p2g_open": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(P)]),
        "p2g_close": (None, [P]),
        "p2g_last_error": (ctypes.c_char_p, [P]),
        "p2g_version": (ctypes.c_char_p, []),
        "p2g_slice_checksums": (ctypes.c_int, [P, P, ctypes.c_uint64, ctypes.c_uint64, P, P]),
        "p2g_encode_begin": (ctypes.c_int, [P, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.POINTER(P)]),
        "p2g_encode_push": (ctypes.c_int, [P, ctypes.c_uint32, P, ctypes.c_uint64]),
        "p2g_encode_finish": (ctypes.c_int, [P, P, P, P]),
        "p2g_encode_free": (None, [P]),
        "p2g_recovery_packet_md5": (ctypes.c_int, [P, P, ctypes.c_uint32, ctypes.c_uint64,
                                                   ctypes.c_uint32, P, P]),
        "p2g_dev_encode": (ctypes.c_int, [P, P, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_uint32, P, ctypes.c_int, P]),
        "p2g_dev_slice_checksums": (ctypes.c_int, [P, P, ctypes.c_uint32, ctypes.c_uint64, P, P,
                                                   P]),
        "p2g_dev_recovery_packet_md5": (ctypes.c_int, [P, P, ctypes.c_uint32, ctypes.c_uint64,
                                                       ctypes.c_uint32, P, P, P]),
        "p2g_dev_create_pass": (ctypes.c_int, [P, P, ctypes.c_uint32, ctypes.c_uint64,
                                               ctypes.c_uint32, P, P, P, P, P, P]),
        "p2g_dev_create_pass_noalloc": (ctypes.c_int, [P, P, ctypes.c_uint32, ctypes.c_uint64,
                                                       ctypes.c_uint32, P, P, P, P, P, P]),
        "p2g_pass_timing": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_double)]),
        "p2g_synchronize": (ctypes.c_int, [P]),
        "p2g_batch_plan": (ctypes.c_int, [P, P, ctypes.c_uint32, ctypes.c_int, ctypes.POINTER(P)]),
        "p2g_batch_launch": (ctypes.c_int, [P, P]),
        "p2g_batch_wait": (ctypes.c_int, [P]),
        "p2g_batch_job_ms": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_double)]),
        "p2g_batch_free": (None, [P]),
        "p2g_dev_create_batch": (ctypes.c_int, [P, P, ctypes.c_uint32, P]),
        "p2g_dev_fill_xorshift64": (ctypes.c_int, [P, P, ctypes.c_uint64, ctypes.c_uint64, P]),
        "p2g_server_open": (ctypes.c_int, [P, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(P)]),
        "p2g_server_submit": (ctypes.c_int, [P, P, ctypes.POINTER(ctypes.c_uint64)]),
        "p2g_server_set_bypass": (ctypes.c_int, [P, ctypes.c_double]),
        "p2g_server_poll": (ctypes.c_int, [P, ctypes.c_double, ctypes.POINTER(ctypes.c_uint64),
                                           ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
        "p2g_server_result": (ctypes.c_int, [P, ctypes.c_uint64, P, ctypes.POINTER(ctypes.c_double)]),
        "p2g_server_release": (ctypes.c_int, [P, ctypes.c_uint64]),
        "p2g_server_stats": (ctypes.c_int, [P, ctypes.POINTER(ctypes.c_uint32),
                                            ctypes.POINTER(ctypes.c_uint32),
                                            ctypes.POINTER(ctypes.c_uint64)]),
        "p2g_server_close": (None, [P]),
        "p2g_tune_bucket": (ctypes.c_uint32, [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]),
        "p2g_plan_roles": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                          ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, u32p,
                                          ctypes.c_uint32, u32p]),
        "p2g_autotune": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_uint32),
                                        ctypes.POINTER(ctypes.c_double)]),
        "p2g_tune_get": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_uint32)]),
        "p2g_tune_get_encode": (ctypes.c_int, [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                               ctypes.POINTER(ctypes.c_int),
                                               ctypes.POINTER(ctypes.c_double)]),
        "p2g_tune_save": (ctypes.c_int, [P, ctypes.c_char_p]),
        "p2g_tune_load": (ctypes.c_int, [P, ctypes.c_char_p]),
        "p2g_run_create": (ctypes.c_int, [P, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint32,
                                          ctypes.POINTER(_Scheme), ctypes.c_uint32,
                                          ctypes.POINTER(ctypes.c_char_p),
                                          ctypes.POINTER(ctypes.c_char_p), u32p,
                                          ctypes.c_char_p, ctypes.c_size_t]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


@dataclass
class VolumeScheme:
    """VolumeScheme, src/creator.rs:110-139.  Limited wraps the inner kind
    (limit_cap), and further Limited wrappers nest around it (outer_caps,
    innermost first)."""

    kind: int = SCHEME_SINGLE
    uniform_count: int = 0
    limited: bool = False
    limit_cap: int = 0
    explicit_sizes: List[int] = field(default_factory=list)
    outer_caps: List[int] = field(default_factory=list)

    @staticmethod
    def single() -> "VolumeScheme":
        return VolumeScheme(SCHEME_SINGLE)

    @staticmethod
    def exponential() -> "VolumeScheme":
        return VolumeScheme(SCHEME_EXPONENTIAL)

    @staticmethod
    def uniform(count: int) -> "VolumeScheme":
        return VolumeScheme(SCHEME_UNIFORM, uniform_count=count)

    @staticmethod
    def explicit(sizes: Sequence[int]) -> "VolumeScheme":
        return VolumeScheme(SCHEME_EXPLICIT, explicit_sizes=list(sizes))

    def capped(self, cap: int) -> "VolumeScheme":
        """Limited { max_blocks_per_volume: cap, inner: self }."""
        if self.limited:
            return VolumeScheme(self.kind, self.uniform_count, True, self.limit_cap,
                                list(self.explicit_sizes), self.outer_caps + [cap])
        return VolumeScheme(self.kind, self.uniform_count, True, cap, list(self.explicit_sizes))

    def _c(self):
        arr = (ctypes.c_uint32 * max(1, len(self.explicit_sizes)))(*self.explicit_sizes)
        caps = (ctypes.c_uint32 * max(1, len(self.outer_caps)))(*self.outer_caps)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        s = _Scheme(self.kind, self.uniform_count, int(self.limited), self.limit_cap,
                    ctypes.cast(arr, u32p) if self.explicit_sizes else None,
                    len(self.explicit_sizes),
                    ctypes.cast(caps, u32p) if self.outer_caps else None, len(self.outer_caps))
        return s, (arr, caps)


@dataclass
class CreateOptions:
    """CreateOptions, src/creator.rs:142-173 (library default scheme: Single)."""

    output: str
    slice_size: int = 4096
    recovery_block_count: int = 0
    volume_scheme: VolumeScheme = field(default_factory=VolumeScheme.single)


def _buf(data) -> Tuple[ctypes.c_void_p, int, object]:
    """(pointer, length, keep-alive) for bytes / bytearray / numpy arrays."""
    if isinstance(data, (bytes, bytearray)):
        b = data if isinstance(data, bytearray) else bytearray(data)
        c = (ctypes.c_char * max(1, len(b))).from_buffer(b) if len(b) else None
        return (ctypes.cast(c, ctypes.c_void_p) if c is not None else None), len(b), (b, c)
    import numpy as np
    a = np.ascontiguousarray(data, dtype=np.uint8)
    return ctypes.c_void_p(a.ctypes.data), a.nbytes, a


def _dev_ptr(x) -> int:
    """Device address of a DeviceBuffer (par2rust_amd.hipmem), a tensor or an int."""
    if isinstance(x, int):
        return x
    if hasattr(x, "addr"):
        return x.addr
    return x.data_ptr()


def _dev_len(x) -> int:
    if hasattr(x, "nbytes") and not hasattr(x, "data_ptr"):
        return x.nbytes
    return x.numel() * x.element_size()


def _dev_files(files):
    """p2g_dev_file array from DeviceBuffers / tensors / (address, length) pairs."""
    items = []
    for f in files:
        if isinstance(f, tuple):
            items.append(_DevFile(f[0], f[1]))
        else:
            items.append(_DevFile(_dev_ptr(f), _dev_len(f)))
    return (_DevFile * max(1, len(items)))(*items), items


@dataclass
class BatchJob:
    """One create job of a batch: the arguments of dev_create_pass.  Buffers
    are DeviceBuffers, tensors or device addresses; files as in dev_encode."""

    files: Sequence
    slice_size: int
    recovery_count: int
    set_id: bytes
    rec: object
    slice_md5: object
    slice_crc: object
    pkt_md5: object


def _batch_jobs(jobs: Sequence[BatchJob]):
    arr = (_BatchJob * max(1, len(jobs)))()
    keep = []
    for k, j in enumerate(jobs):
        files, items = _dev_files(j.files)
        keep.append(files)
        b = arr[k]
        b.files = ctypes.cast(files, ctypes.c_void_p).value
        b.nfiles = len(items)
        b.slice_size = j.slice_size
        b.recovery_count = j.recovery_count
        b.set_id[:] = list(bytes(j.set_id))
        b.rec = _dev_ptr(j.rec) if j.recovery_count else None
        b.slice_md5 = _dev_ptr(j.slice_md5)
        b.slice_crc = _dev_ptr(j.slice_crc)
        b.pkt_md5 = _dev_ptr(j.pkt_md5) if j.recovery_count else None
    return arr, keep


class Batch:
    """A planned batch of create jobs (p2g_batch_plan): launch it any number
    of times, wait, read per-job completion times."""

    def __init__(self, ctx: "Context", jobs: Sequence[BatchJob], use_graph: bool = True):
        self.ctx, self.njobs = ctx, len(jobs)
        arr, self._keep = _batch_jobs(jobs)
        h = ctypes.c_void_p()
        ctx._check(lib.p2g_batch_plan(ctx.h, arr, len(jobs), int(use_graph), ctypes.byref(h)))
        self.h = h

    def launch(self, stream: Optional[int] = None):
        self.ctx._check(lib.p2g_batch_launch(self.h, ctypes.c_void_p(stream) if stream else None))

    def wait(self):
        self.ctx._check(lib.p2g_batch_wait(self.h))

    def job_ms(self) -> List[float]:
        out = (ctypes.c_double * max(1, self.njobs))()
        self.ctx._check(lib.p2g_batch_job_ms(self.h, out))
        return list(out[:self.njobs])

    def free(self):
        if getattr(self, "h", None):
            lib.p2g_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """The serving loop (p2g_server_*): submit requests, poll for finished
    ones, read their outputs, release them."""

    def __init__(self, ctx: "Context", hbm_budget: int, step_ms: float = 0.0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx._check(lib.p2g_server_open(ctx.h, hbm_budget, step_ms, ctypes.byref(h)))
        self.h = h
        self._keep = {}

    def set_bypass(self, bypass_ms: float):
        """How long a request waiting for memory may be overtaken by later
        requests that fit (p2g_server_set_bypass; 0: arrival order)."""
        self.ctx._check(lib.p2g_server_set_bypass(self.h, bypass_ms))

    def submit(self, file_lens: Sequence[int], slice_size: int, recovery_count: int,
               set_id: bytes, host_data: Optional[Sequence[bytes]] = None,
               seeds: Optional[Sequence[int]] = None) -> int:
        n = len(file_lens)
        r = _Request()
        lens = (ctypes.c_uint64 * max(1, n))(*file_lens)
        keep = [lens]
        r.file_lens = lens
        if host_data is not None:
            bufs = [ctypes.create_string_buffer(bytes(d), max(1, len(d))) for d in host_data]
            ptrs = (ctypes.c_void_p * max(1, n))(*[ctypes.addressof(b) for b in bufs])
            keep += [bufs, ptrs]
            r.host_data = ptrs
        if seeds is not None:
            sd = (ctypes.c_uint64 * max(1, n))(*seeds)
            keep.append(sd)
            r.fill_seeds = sd
        r.slice_size, r.nfiles, r.recovery_count = slice_size, n, recovery_count
        r.set_id[:] = list(bytes(set_id))
        rid = ctypes.c_uint64()
        self.ctx._check(lib.p2g_server_submit(self.h, ctypes.byref(r), ctypes.byref(rid)))
        self._keep[rid.value] = keep  # host bytes stay valid until release
        return rid.value

    def poll(self, wait_ms: float = 0.0, cap: int = 4096) -> List[int]:
        ids = (ctypes.c_uint64 * cap)()
        n = ctypes.c_uint32()
        self.ctx._check(lib.p2g_server_poll(self.h, wait_ms, ids, cap, ctypes.byref(n)))
        return list(ids[:n.value])

    def result(self, rid: int):
        """(p2g_batch_job with device pointers, (submit, admit, done) ms)."""
        job = _BatchJob()
        t = (ctypes.c_double * 3)()
        self.ctx._check(lib.p2g_server_result(self.h, rid, ctypes.byref(job), t))
        return job, tuple(t)

    def download(self, rid: int):
        """(recovery blocks, slice MD5s, slice CRCs, packet MD5s) of a done request."""
        from . import hipmem

        job, _ = self.result(rid)
        S, R = job.slice_size, job.recovery_count
        files = ctypes.cast(job.files, ctypes.POINTER(_DevFile))
        n = sum((files[f].len + S - 1) // S for f in range(job.nfiles))
        rec = hipmem.download_ptr(job.rec, R * S) if R else b""
        md5 = hipmem.download_ptr(job.slice_md5, 16 * n)
        crc = list(struct.unpack(f"<{n}I", hipmem.download_ptr(job.slice_crc, 4 * n)))
        pkt = hipmem.download_ptr(job.pkt_md5, 16 * R) if R else b""
        return rec, md5, crc, pkt

    def release(self, rid: int):
        self.ctx._check(lib.p2g_server_release(self.h, rid))
        self._keep.pop(rid, None)

    def stats(self):
        q, a, u = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint64()
        self.ctx._check(lib.p2g_server_stats(self.h, ctypes.byref(q), ctypes.byref(a),
                                             ctypes.byref(u)))
        return {"queued": q.value, "active": a.value, "hbm_used": u.value}

    def close(self):
        if getattr(self, "h", None):
            lib.p2g_server_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One p2g context on one GPU (p2g_open / p2g_close)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib.p2g_open(device, ctypes.byref(h))
        if rc != P2G_OK:
            raise Par2Error(rc, "p2g_open failed")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.p2g_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != P2G_OK:
            raise Par2Error(rc, lib.p2g_last_error(self.h).decode(errors="replace"))

    # --- host-buffer entry points -------------------------------------
    def slice_checksums(self, data, slice_size: int) -> Tuple[List[bytes], List[int]]:
        p, n, keep = _buf(data)
        count = (n + slice_size - 1) // slice_size if slice_size else 0
        md5 = ctypes.create_string_buffer(16 * max(1, count))
        crc = (ctypes.c_uint32 * max(1, count))()
        self._check(lib.p2g_slice_checksums(self.h, p, n, slice_size, md5, crc))
        return [md5.raw[16 * i:16 * i + 16] for i in range(count)], [crc[i] for i in range(count)]

    def encode(self, files: Sequence, slice_size: int, recovery_count: int,
               set_id: Optional[bytes] = None, push_bytes: int = 64 << 20):
        """Recovery blocks (R * S bytes, exponent-major) of the slices of
        `files` in order, through p2g_encode_begin / push / finish.  With
        set_id, also returns the R recovery-packet MD5s."""
        total = sum((len(f) + slice_size - 1) // slice_size for f in files)
        job = ctypes.c_void_p()
        self._check(lib.p2g_encode_begin(self.h, slice_size, total, recovery_count,
                                         ctypes.byref(job)))
        try:
            idx = 0
            chunk = max(1, push_bytes // slice_size) * slice_size
            for f in files:
                p, n, keep = _buf(f)
                off = 0
                while off < n:
                    take = min(chunk, n - off)
                    self._check(lib.p2g_encode_push(job, idx, ctypes.c_void_p(p.value + off),
                                                    take))
                    idx += (take + slice_size - 1) // slice_size
                    off += take
            rec = ctypes.create_string_buffer(max(1, recovery_count * slice_size))
            md5 = ctypes.create_string_buffer(16 * max(1, recovery_count))
            sid = ctypes.create_string_buffer(bytes(set_id), 16) if set_id is not None else None
            self._check(lib.p2g_encode_finish(job, rec, sid, md5 if set_id is not None else None))
        finally:
            lib.p2g_encode_free(job)
        out = rec.raw[:recovery_count * slice_size]
        if set_id is None:
            return out
        return out, [md5.raw[16 * i:16 * i + 16] for i in range(recovery_count)]

    def recovery_packet_md5(self, rec, count: int, slice_size: int, first_exponent: int,
                            set_id: bytes) -> List[bytes]:
        p, n, keep = _buf(rec)
        out = ctypes.create_string_buffer(16 * max(1, count))
        sid = ctypes.create_string_buffer(bytes(set_id), 16)
        self._check(lib.p2g_recovery_packet_md5(self.h, p, count, slice_size, first_exponent,
                                                sid, out))
        return [out.raw[16 * i:16 * i + 16] for i in range(count)]

    # --- device-buffer entry points ----------------------------------------
    # Buffers: par2rust_amd.hipmem.DeviceBuffer, tensors, or (address, length).
    def dev_encode(self, files, slice_size: int, recovery_count: int, rec,
                   first_input: int = 0, total_inputs: Optional[int] = None,
                   accumulate: bool = False, stream: Optional[int] = None):
        arr, items = _dev_files(files)
        if total_inputs is None:
            total_inputs = first_input + sum((it.len + slice_size - 1) // slice_size
                                             for it in items)
        self._check(lib.p2g_dev_encode(self.h, arr, len(items), slice_size, first_input,
                                       total_inputs, recovery_count,
                                       ctypes.c_void_p(_dev_ptr(rec)), int(accumulate),
                                       ctypes.c_void_p(stream) if stream else None))

    def dev_slice_checksums(self, files, slice_size: int, md5_out, crc_out,
                            stream: Optional[int] = None):
        arr, items = _dev_files(files)
        self._check(lib.p2g_dev_slice_checksums(self.h, arr, len(items), slice_size,
                                                ctypes.c_void_p(_dev_ptr(md5_out)),
                                                ctypes.c_void_p(_dev_ptr(crc_out)),
                                                ctypes.c_void_p(stream) if stream else None))

    def dev_recovery_packet_md5(self, rec, count: int, slice_size: int, first_exponent: int,
                                set_id: bytes, md5_out, stream: Optional[int] = None):
        sid = ctypes.create_string_buffer(bytes(set_id), 16)
        self._check(lib.p2g_dev_recovery_packet_md5(self.h, ctypes.c_void_p(_dev_ptr(rec)),
                                                    count, slice_size, first_exponent, sid,
                                                    ctypes.c_void_p(_dev_ptr(md5_out)),
                                                    ctypes.c_void_p(stream) if stream else None))

    def dev_create_pass(self, files, slice_size: int, recovery_count: int, set_id: bytes, rec,
                        slice_md5_out, slice_crc_out, pkt_md5_out,
                        stream: Optional[int] = None):
        """Encode + slice MD5/CRC-32 + packet MD5 of a device-resident set in
        one pass (p2g_dev_create_pass); results land in the device buffers."""
        arr, items = _dev_files(files)
        sid = ctypes.create_string_buffer(bytes(set_id), 16)
        self._check(lib.p2g_dev_create_pass(self.h, arr, len(items), slice_size, recovery_count,
                                            sid, ctypes.c_void_p(_dev_ptr(rec)),
                                            ctypes.c_void_p(_dev_ptr(slice_md5_out)),
                                            ctypes.c_void_p(_dev_ptr(slice_crc_out)),
                                            ctypes.c_void_p(_dev_ptr(pkt_md5_out)),
                                            ctypes.c_void_p(stream) if stream else None))

    def dev_create_pass_noalloc(self, files, slice_size: int, recovery_count: int, set_id: bytes,
                                rec, slice_md5_out, slice_crc_out, pkt_md5_out,
                                stream: Optional[int] = None):
        """dev_create_pass without any allocation (p2g_dev_create_pass_noalloc):
        scratch in the library's static device array; at most 64 files,
        2048 slices and 64 MiB of scratch."""
        arr, items = _dev_files(files)
        sid = ctypes.create_string_buffer(bytes(set_id), 16)
        self._check(lib.p2g_dev_create_pass_noalloc(
            self.h, arr, len(items), slice_size, recovery_count, sid,
            ctypes.c_void_p(_dev_ptr(rec)), ctypes.c_void_p(_dev_ptr(slice_md5_out)),
            ctypes.c_void_p(_dev_ptr(slice_crc_out)), ctypes.c_void_p(_dev_ptr(pkt_md5_out)),
            ctypes.c_void_p(stream) if stream else None))

    def dev_create_batch(self, jobs: Sequence[BatchJob], stream: Optional[int] = None):
        """Every job's dev_create_pass work in shared persistent launches
        (p2g_dev_create_batch); synchronize() waits for it."""
        arr, keep = _batch_jobs(jobs)
        self._check(lib.p2g_dev_create_batch(self.h, arr, len(jobs),
                                             ctypes.c_void_p(stream) if stream else None))

    def dev_fill_xorshift64(self, dst, nbytes: int, seed: int, stream: Optional[int] = None):
        """The synthetic workload stream (par2rust_amd.synth) written on the device."""
        self._check(lib.p2g_dev_fill_xorshift64(self.h, ctypes.c_void_p(_dev_ptr(dst)), nbytes,
                                                seed, ctypes.c_void_p(stream) if stream else None))

    def autotune(self, slices: int, recovery_count: int, slice_size: int) -> dict:
        """Tunes the bucket of the shape on synthetic input (p2g_autotune)."""
        v, b, ms = ctypes.c_int(), ctypes.c_uint32(), ctypes.c_double()
        self._check(lib.p2g_autotune(self.h, slices, recovery_count, slice_size, ctypes.byref(v),
                                     ctypes.byref(b), ctypes.byref(ms)))
        return {"variant": v.value, "band_cols": b.value, "ms": ms.value}

    def tune_get(self, slices: int, recovery_count: int, slice_size: int) -> dict:
        v, b = ctypes.c_int(), ctypes.c_uint32()
        self._check(lib.p2g_tune_get(self.h, slices, recovery_count, slice_size, ctypes.byref(v),
                                     ctypes.byref(b)))
        return {"variant": v.value, "band_cols": b.value}

    def tune_get_encode(self, slices: int, recovery_count: int, slice_size: int) -> dict:
        """The encode kernel dev_encode / encode_push use for the shape now."""
        v, ms = ctypes.c_int(), ctypes.c_double()
        self._check(lib.p2g_tune_get_encode(self.h, slices, recovery_count, slice_size,
                                            ctypes.byref(v), ctypes.byref(ms)))
        return {"variant": v.value, "ms": ms.value}

    def tune_save(self, path: str):
        self._check(lib.p2g_tune_save(self.h, os.fsencode(path)))

    def tune_load(self, path: str):
        self._check(lib.p2g_tune_load(self.h, os.fsencode(path)))

    def server(self, hbm_budget: int, step_ms: float = 0.0) -> Server:
        return Server(self, hbm_budget, step_ms)

    def batch_plan(self, jobs: Sequence[BatchJob], use_graph: bool = True) -> Batch:
        return Batch(self, jobs, use_graph)

    def pass_timing(self) -> dict:
        """(start, end) in ms of each role of the last fused pass."""
        out = (ctypes.c_double * 8)()
        self._check(lib.p2g_pass_timing(self.h, out))
        return {role: (round(out[2 * k], 4), round(out[2 * k + 1], 4))
                for k, role in enumerate(("packet_md5", "slice_md5", "encode", "crc"))}

    def synchronize(self):
        self._check(lib.p2g_synchronize(self.h))

    # --- whole pipeline --------------------------------------------------
    def run_create(self, opts: CreateOptions, inputs: Sequence[Tuple[str, bytes]]) -> List[str]:
        """SourceFile::scan of each (path, display name), then run_create
        (src/creator.rs:182).  Returns the paths written, index first."""
        scheme, keep = opts.volume_scheme._c()
        n = len(inputs)
        paths = (ctypes.c_char_p * max(1, n))(*[os.fsencode(p) for p, _ in inputs])
        names = (ctypes.c_char_p * max(1, n))(*[bytes(nm) for _, nm in inputs])
        lens = (ctypes.c_uint32 * max(1, n))(*[len(nm) for _, nm in inputs])
        cap = 1 << 20
        out = ctypes.create_string_buffer(cap)
        rc = lib.p2g_run_create(self.h, os.fsencode(opts.output), opts.slice_size,
                                opts.recovery_block_count, ctypes.byref(scheme), n, paths, names,
                                lens, out, cap)
        if rc != P2G_OK:
            raise Par2Error(rc, lib.p2g_last_error(self.h).decode(errors="replace"))
        return [s.decode() for s in out.raw.split(b"\0") if s][: 1 + 65536]


def tune_bucket(slices: int, recovery_count: int, slice_size: int) -> int:
    """The autotuner's bucket id of a shape (p2g_tune_bucket)."""
    return lib.p2g_tune_bucket(slices, recovery_count, slice_size)


ROLE_PACKET, ROLE_SLICE, ROLE_ENCODE = 0, 1, 2


def plan_roles(kind: str, num_cus: int, max_resident: int, slices: int, recovery_count: int,
               slice_size: int) -> List[int]:
    """Role of each workgroup of the persistent launch the planner makes
    (p2g_plan_roles; host only): kind "pass" or "batch"."""
    cap = 4096
    roles = (ctypes.c_uint32 * cap)()
    grid = ctypes.c_uint32()
    rc = lib.p2g_plan_roles({"pass": 0, "batch": 1}[kind], num_cus, max_resident, slices,
                            recovery_count, slice_size, roles, cap, ctypes.byref(grid))
    if rc != P2G_OK:
        raise Par2Error(rc, "p2g_plan_roles")
    return list(roles)[:grid.value]


def version() -> str:
    return lib.p2g_version().decode()
