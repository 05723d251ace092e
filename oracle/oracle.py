"""ctypes wrapper of the CPU oracle (oracle/libpar2_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg as the checker.  The product never imports it.
Build: ``make -C oracle``.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Sequence, Tuple

HERE = os.path.dirname(os.path.abspath(__file__))
# PAR2_ORACLE_LIB: another build of the same sources (the sanitizer test
# loads an AddressSanitizer / UBSan build through it)
LIB_PATH = os.environ.get("PAR2_ORACLE_LIB") or os.path.join(HERE, "libpar2_oracle.so")
REF_MD5_PATH = os.path.join(HERE, "_ref", "libmd5ref.so")

SCHEME_SINGLE, SCHEME_EXPONENTIAL, SCHEME_UNIFORM, SCHEME_EXPLICIT = 0, 1, 2, 3


def build() -> None:
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


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


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u64, u32, u16 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint16
    sig = {
        "or_gf_tables": (None, [P, P]),
        "or_gf_mul": (u16, [u16, u16]),
        "or_gf_pow": (u16, [u16, u32]),
        "or_gf_inv": (u16, [u16]),
        "or_input_bases": (ctypes.c_int, [u32, P]),
        "or_gf_mul_xor_scalar": (None, [u16, P, P, ctypes.c_size_t]),
        "or_md5": (None, [P, ctypes.c_size_t, P]),
        "or_crc32": (u32, [P, ctypes.c_size_t]),
        "or_slice_count": (u64, [u64, u64]),
        "or_slice_checksums": (None, [P, u64, u64, P, P]),
        "or_file_hashes": (None, [P, u64, P, P]),
        "or_file_id": (None, [P, u64, P, ctypes.c_size_t, P]),
        "or_encode": (ctypes.c_int, [P, P, u32, u64, u32, P]),
        "or_encode_range": (ctypes.c_int, [P, P, u32, u64, u32, u32, P]),
        "or_encode_simd": (ctypes.c_int, [P, P, u32, u64, u32, P, u32, ctypes.c_int]),
        "or_simd_detect": (ctypes.c_int, []),
        "or_simd_dispatch_name": (ctypes.c_char_p, [ctypes.c_int]),
        "or_build_packet": (ctypes.c_size_t, [P, P, P, ctypes.c_size_t, P]),
        "or_main_packet": (ctypes.c_size_t, [u64, P, u32, P, P]),
        "or_recovery_packet": (ctypes.c_size_t, [P, u16, P, ctypes.c_size_t, P]),
        "or_creator_packet": (ctypes.c_size_t, [P, P]),
        "or_resolve_volume_sizes": (ctypes.c_int, [ctypes.POINTER(_Scheme), u32, P, u32]),
        "or_run_create": (ctypes.c_int, [ctypes.c_char_p, u64, u32, ctypes.POINTER(_Scheme), u32,
                                         ctypes.POINTER(ctypes.c_char_p),
                                         ctypes.POINTER(ctypes.c_char_p), P,
                                         ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p,
                                         ctypes.c_size_t]),
        "or_xorshift64_fill": (u64, [P, ctypes.c_size_t, u64]),
        "or_slice_checksums_mt": (None, [P, u64, u64, P, P, u32]),
        "or_recovery_md5_mt": (None, [P, u32, u64, u32, P, P, u32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


L = _load()


def _ptr(b):
    if isinstance(b, bytes):
        return ctypes.c_char_p(b)
    if isinstance(b, bytearray):
        return (ctypes.c_char * max(1, len(b))).from_buffer(b) if len(b) else None
    import numpy as np
    a = b if isinstance(b, np.ndarray) else np.frombuffer(b, dtype=np.uint8)
    return ctypes.c_void_p(a.ctypes.data)


def gf_tables() -> Tuple[List[int], List[int]]:
    log = (ctypes.c_uint16 * 65536)()
    alog = (ctypes.c_uint16 * 65536)()
    L.or_gf_tables(log, alog)
    return list(log), list(alog)


def gf_mul(a: int, b: int) -> int:
    return L.or_gf_mul(a, b)


def gf_pow(a: int, e: int) -> int:
    return L.or_gf_pow(a, e)


def gf_inv(a: int) -> int:
    return L.or_gf_inv(a)


def input_bases(n: int) -> List[int]:
    out = (ctypes.c_uint16 * max(1, n))()
    if L.or_input_bases(n, out) != 0:
        raise ValueError("bases exhausted")
    return list(out)[:n]


def gf_mul_xor_scalar(coeff: int, inp: bytes, out: bytearray) -> None:
    L.or_gf_mul_xor_scalar(coeff, _ptr(inp), _ptr(out), len(inp))


def md5(data: bytes) -> bytes:
    out = ctypes.create_string_buffer(16)
    L.or_md5(_ptr(data), len(data), out)
    return out.raw


def crc32(data: bytes) -> int:
    return L.or_crc32(_ptr(data), len(data))


def slice_checksums(data: bytes, S: int) -> Tuple[List[bytes], List[int]]:
    n = L.or_slice_count(len(data), S)
    md5s = ctypes.create_string_buffer(16 * max(1, n))
    crcs = (ctypes.c_uint32 * max(1, n))()
    L.or_slice_checksums(_ptr(data), len(data), S, md5s, crcs)
    return [md5s.raw[16 * i:16 * i + 16] for i in range(n)], list(crcs)[:n]


def slice_checksums_mt(data: bytes, S: int, threads: int = 0) -> Tuple[List[bytes], List[int]]:
    n = L.or_slice_count(len(data), S)
    md5s = ctypes.create_string_buffer(16 * max(1, n))
    crcs = (ctypes.c_uint32 * max(1, n))()
    L.or_slice_checksums_mt(_ptr(data), len(data), S, md5s, crcs, threads)
    return [md5s.raw[16 * i:16 * i + 16] for i in range(n)], list(crcs)[:n]


def recovery_md5_mt(rec: bytes, count: int, S: int, first_exp: int, set_id: bytes,
                    threads: int = 0) -> List[bytes]:
    out = ctypes.create_string_buffer(16 * max(1, count))
    L.or_recovery_md5_mt(_ptr(rec), count, S, first_exp, set_id, out, threads)
    return [out.raw[16 * i:16 * i + 16] for i in range(count)]


def file_hashes(data: bytes) -> Tuple[bytes, bytes]:
    a = ctypes.create_string_buffer(16)
    b = ctypes.create_string_buffer(16)
    L.or_file_hashes(_ptr(data), len(data), a, b)
    return a.raw, b.raw


def file_id(hash16k: bytes, length: int, name: bytes) -> bytes:
    out = ctypes.create_string_buffer(16)
    L.or_file_id(hash16k, length, name, len(name), out)
    return out.raw


def _files(files: Sequence):
    n = len(files)
    bufs = [bytes(f) for f in files]  # kept alive by the caller's reference
    ptrs = (ctypes.c_char_p * max(1, n))(*bufs)
    lens = (ctypes.c_uint64 * max(1, n))(*[len(b) for b in bufs])
    return ptrs, lens, bufs


def encode(files: Sequence[bytes], S: int, R: int) -> bytes:
    ptrs, lens, keep = _files(files)
    out = ctypes.create_string_buffer(max(1, R * S))
    rc = L.or_encode(ptrs, lens, len(files), S, R, out)
    if rc != 0:
        raise RuntimeError(f"or_encode rc={rc}")
    return out.raw[:R * S]


def encode_rows(files: Sequence[bytes], S: int, first_exp: int, count: int) -> bytes:
    """Recovery blocks of exponents first_exp .. first_exp + count - 1 only."""
    ptrs, lens, keep = _files(files)
    out = ctypes.create_string_buffer(max(1, count * S))
    rc = L.or_encode_range(ptrs, lens, len(files), S, first_exp, count, out)
    if rc != 0:
        raise RuntimeError(f"or_encode_range rc={rc}")
    return out.raw[:count * S]


def encode_simd(files: Sequence[bytes], S: int, R: int, threads: int = 0,
                dispatch: int = -1) -> bytes:
    ptrs, lens, keep = _files(files)
    out = ctypes.create_string_buffer(max(1, R * S))
    rc = L.or_encode_simd(ptrs, lens, len(files), S, R, out, threads, dispatch)
    if rc != 0:
        raise RuntimeError(f"or_encode_simd rc={rc}")
    return out.raw[:R * S]


def simd_dispatch() -> str:
    return L.or_simd_dispatch_name(L.or_simd_detect()).decode()


def build_packet(set_id: bytes, type_tag: bytes, body: bytes) -> bytes:
    out = ctypes.create_string_buffer(64 + len(body))
    n = L.or_build_packet(set_id, type_tag, _ptr(body), len(body), out)
    return out.raw[:n]


def main_packet(S: int, file_ids: Sequence[bytes]) -> Tuple[bytes, bytes]:
    ids = b"".join(file_ids)
    out = ctypes.create_string_buffer(64 + 12 + len(ids))
    sid = ctypes.create_string_buffer(16)
    n = L.or_main_packet(S, ids, len(file_ids), out, sid)
    return out.raw[:n], sid.raw


def recovery_packet(set_id: bytes, exponent: int, data: bytes) -> bytes:
    out = ctypes.create_string_buffer(68 + len(data))
    n = L.or_recovery_packet(set_id, exponent, data, len(data), out)
    return out.raw[:n]


def creator_packet(set_id: bytes) -> bytes:
    out = ctypes.create_string_buffer(128)
    n = L.or_creator_packet(set_id, out)
    return out.raw[:n]


def _scheme(kind: int, uniform_count: int = 0, limited: bool = False, limit_cap: int = 0,
            explicit: Sequence[int] = (), outer_caps: Sequence[int] = ()):
    arr = (ctypes.c_uint32 * max(1, len(explicit)))(*explicit)
    caps = (ctypes.c_uint32 * max(1, len(outer_caps)))(*outer_caps)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    s = _Scheme(kind, uniform_count, int(limited), limit_cap,
                ctypes.cast(arr, u32p) if explicit else None, len(explicit),
                ctypes.cast(caps, u32p) if outer_caps else None, len(outer_caps))
    return s, (arr, caps)


class SchemeError(ValueError):
    pass


def resolve_volume_sizes(total: int, kind: int, uniform_count: int = 0, limited: bool = False,
                         limit_cap: int = 0, explicit: Sequence[int] = (),
                         outer_caps: Sequence[int] = ()) -> List[int]:
    s, keep = _scheme(kind, uniform_count, limited, limit_cap, explicit, outer_caps)
    cap = 70000
    out = (ctypes.c_uint32 * cap)()
    n = L.or_resolve_volume_sizes(ctypes.byref(s), total, out, cap)
    if n < 0:
        raise SchemeError(f"InvalidVolumeScheme ({-n})")
    return list(out)[:n]


class OracleError(RuntimeError):
    def __init__(self, code, msg=""):
        self.code = code
        super().__init__(f"oracle rc={code} {msg}")


def run_create(output: str, S: int, R: int, inputs: Sequence[Tuple[str, bytes]],
               kind: int = SCHEME_SINGLE, uniform_count: int = 0, limited: bool = False,
               limit_cap: int = 0, explicit: Sequence[int] = (),
               outer_caps: Sequence[int] = ()) -> List[str]:
    s, keep = _scheme(kind, uniform_count, limited, limit_cap, explicit, outer_caps)
    n = len(inputs)
    paths = (ctypes.c_char_p * max(1, n))(*[os.fsencode(p) for p, _ in inputs])
    names = (ctypes.c_char_p * max(1, n))(*[bytes(nm) for _, nm in inputs])
    lens = (ctypes.c_uint32 * max(1, n))(*[len(nm) for _, nm in inputs])
    cap = 1 << 20
    out = ctypes.create_string_buffer(cap)
    msg = ctypes.create_string_buffer(512)
    rc = L.or_run_create(os.fsencode(output), S, R, ctypes.byref(s), n, paths, names, lens, out,
                         cap, msg, 512)
    if rc != 0:
        raise OracleError(rc, msg.value.decode(errors="replace"))
    return [x.decode() for x in out.raw.split(b"\0") if x]


def xorshift64_bytes(n: int, seed: int) -> bytes:
    buf = ctypes.create_string_buffer(max(1, n))
    L.or_xorshift64_fill(buf, n, seed)
    return buf.raw[:n]


def ref_md5(data: bytes):
    """MD5 by the reference tree's vendored RFC 1321 code (oracle/_ref), or
    None when that library was not built (no reference tree present)."""
    if not os.path.exists(REF_MD5_PATH):
        return None
    R = ctypes.CDLL(REF_MD5_PATH)
    R.md5_ref.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
    R.md5_ref.restype = None
    dig = (ctypes.c_uint32 * 4)()
    b = ctypes.create_string_buffer(bytes(data), max(1, len(data)))
    R.md5_ref(b, dig, len(data))
    return b"".join(int(d).to_bytes(4, "little") for d in dig)
