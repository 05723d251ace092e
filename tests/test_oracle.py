"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Three kinds of pins, none of which retypes the oracle's own formula:
  * the reference's unit tests restated (file:line cited on each test),
  * library routines that compute the same function (hashlib.md5,
    zlib.crc32) and the reference tree's vendored md5_ref.c (oracle/_ref),
  * mathematics: a textbook carry-less GF(2)[x] multiply reduced by
    0x1100B, phi(65535) = 32768, linearity and the R = 1 parity invariant.
"""
import hashlib
import math
import os
import random
import struct
import zlib

import pytest

from oracle import oracle as O

FIELD = 0x1100B


def clmul_mod(a: int, b: int) -> int:
    """Textbook GF(2^16) product: carry-less multiply, reduce by 0x1100B."""
    p = 0
    for i in range(16):
        if (b >> i) & 1:
            p ^= a << i
    for bit in range(31, 15, -1):
        if (p >> bit) & 1:
            p ^= FIELD << (bit - 16)
    return p


def scalar_mul_xor_bytes(coeff: int, data: bytes) -> bytes:
    out = bytearray(len(data))
    for k in range(0, len(data), 2):
        s = data[k] | (data[k + 1] << 8)
        p = clmul_mod(coeff, s)
        out[k] = p & 0xFF
        out[k + 1] = p >> 8
    return bytes(out)


# ---------------------------------------------------------------- GF(2^16)

def test_tables_sentinels_and_generator():
    # src/galois.rs:105-111
    log, alog = O.gf_tables()
    assert log[0] == 65535
    assert alog[0] == 1 and alog[1] == 2
    assert alog[65535] == 0
    # the antilog walk is a permutation of the non-zero field elements
    assert sorted(alog[:65535]) == list(range(1, 65536))
    for x in range(1, 65536, 97):
        assert alog[log[x]] == x


def test_gf_mul_equals_carryless_multiply():
    rnd = random.Random(7)
    pairs = [(rnd.randrange(65536), rnd.randrange(65536)) for _ in range(3000)]
    pairs += [(0, 5), (5, 0), (1, 0xFFFF), (0x8000, 2), (0xFFFF, 0xFFFF)]
    for a, b in pairs:
        assert O.gf_mul(a, b) == clmul_mod(a, b), (a, b)


def test_gf_mul_commutative_zero_identity():
    # src/galois.rs:113-140
    state = 0xCAFEBABE
    for _ in range(2000):
        state = (state * 2654435761 + 1) & 0xFFFFFFFF
        a, b = state & 0xFFFF, state >> 16
        assert O.gf_mul(a, b) == O.gf_mul(b, a)
    for x in (1, 2, 0xABCD, 0xFFFF):
        assert O.gf_mul(0, x) == 0 and O.gf_mul(x, 0) == 0
    for x in (1, 2, 3, 0x100, 0x1234, 0xFFFF):
        assert O.gf_mul(1, x) == x and O.gf_mul(x, 1) == x


def test_gf_inverse_and_pow():
    # src/galois.rs:142-183
    for x in range(1, 0x10000, 7):
        assert O.gf_mul(x, O.gf_inv(x)) == 1
    for base in (0, 1, 2, 0x1234, 0xFFFF):
        assert O.gf_pow(base, 0) == 1
    for e in (1, 2, 100, 65535):
        assert O.gf_pow(0, e) == 0
    _, alog = O.gf_tables()
    for k in (0, 1, 2, 100, 1000, 65534):
        assert O.gf_pow(2, k) == alog[k % 65535]
    for a in (1, 3, 0x1111, 0xFFFE):
        a2 = O.gf_mul(a, a)
        assert O.gf_mul(O.gf_mul(a2, a2), a2) == O.gf_pow(a, 6)
    # repeated multiplication, independent of the log tables
    for a in (3, 0x1234, 0xBEEF):
        acc = 1
        for e in range(1, 40):
            acc = clmul_mod(acc, a)
            assert O.gf_pow(a, e) == acc


# --------------------------------------------------------- RS input bases

def test_input_bases_known_answers():
    # src/reedsolomon.rs:165-200
    _, alog = O.gf_tables()
    log, _ = O.gf_tables()
    assert O.input_bases(3)[0] == 2
    assert O.input_bases(2) == [alog[1], alog[2]]
    logs = [log[b] for b in O.input_bases(20)]
    for skip in (0, 3, 5, 6, 9, 10, 15, 17):
        assert skip not in logs
    b = O.input_bases(1000)
    assert len(set(b)) == 1000


def test_input_bases_are_the_logs_coprime_to_65535():
    log, _ = O.gf_tables()
    bases = O.input_bases(32768)
    logs = [log[b] for b in bases]
    expect = [l for l in range(65535) if math.gcd(65535, l) == 1]
    assert len(expect) == 32768  # phi(65535) = 2 * 4 * 16 * 256
    assert logs == expect
    with pytest.raises(ValueError):
        O.input_bases(32769)


# ------------------------------------------------------------- mul-xor

def test_mul_xor_zero_one_and_example():
    # src/reedsolomon.rs:222-259
    out = bytearray(8)
    out[0], out[7] = 0x11, 0x77
    O.gf_mul_xor_scalar(0, b"\xff" * 8, out)
    assert out == bytearray([0x11, 0, 0, 0, 0, 0, 0, 0x77])
    out = bytearray([0xAA, 0xBB, 0xCC, 0xDD])
    O.gf_mul_xor_scalar(1, bytes([0x12, 0x34, 0x56, 0x78]), out)
    assert out == bytearray([0xAA ^ 0x12, 0xBB ^ 0x34, 0xCC ^ 0x56, 0xDD ^ 0x78])
    inp = bytes([0x11, 0x22, 0xFF, 0x00, 0x99, 0xAB])
    out = bytearray([0x10, 0x20, 0x30, 0x40, 0x50, 0x60])
    exp = bytearray(out)
    for k in range(3):
        p = clmul_mod(0x1234, inp[2 * k] | (inp[2 * k + 1] << 8))
        exp[2 * k] ^= p & 0xFF
        exp[2 * k + 1] ^= p >> 8
    O.gf_mul_xor_scalar(0x1234, inp, out)
    assert out == exp


def test_mul_xor_brute_force_all_symbols():
    # every 16-bit symbol for a handful of coefficients (galois_simd.rs:849-874)
    data = b"".join(struct.pack("<H", s) for s in range(65536))
    for c in (2, 3, 0x100B, 0x8000, 0xFFFF):
        out = bytearray(len(data))
        O.gf_mul_xor_scalar(c, data, out)
        for s in range(0, 65536, 251):
            assert struct.unpack_from("<H", out, 2 * s)[0] == clmul_mod(c, s)


def deterministic(seed: int, n: int) -> bytes:
    # LCG of src/galois_simd.rs:772-782 (reference test data generator)
    out = bytearray(n)
    s = seed
    for i in range(n):
        s = (s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        out[i] = (s >> 33) & 0xFF
    return bytes(out)


# ------------------------------------------------------------- encode

def test_encode_two_inputs_one_recovery_is_xor():
    # src/reedsolomon.rs:261-276
    in0, in1 = bytes(range(16)), bytes(range(16, 32))
    rec = O.encode([in0 + in1], 16, 1)
    assert rec == bytes(a ^ b for a, b in zip(in0, in1))


def test_encode_matches_textbook_definition():
    S, R = 64, 5
    files = [deterministic(1, 150), deterministic(2, 64), deterministic(3, 3)]
    slices = []
    for f in files:
        for off in range(0, len(f), S):
            slices.append(f[off:off + S].ljust(S, b"\0"))
    bases = O.input_bases(len(slices))
    expect = []
    for r in range(R):
        acc = bytearray(S)
        for i, sl in enumerate(slices):
            c = 1
            for _ in range(r):
                c = clmul_mod(c, bases[i])
            prod = scalar_mul_xor_bytes(c, sl)
            acc = bytearray(a ^ b for a, b in zip(acc, prod))
        expect.append(bytes(acc))
    assert O.encode(files, S, R) == b"".join(expect)


def test_encode_rows_are_rows_of_the_full_encode():
    S = 64
    files = [deterministic(21, 300), deterministic(22, 77)]
    full = O.encode(files, S, 12)
    for first, count in [(0, 12), (0, 1), (5, 3), (11, 1)]:
        assert O.encode_rows(files, S, first, count) == full[first * S:(first + count) * S]


def test_encode_rows_at_a_high_exponent_matches_textbook():
    # exponent 999 (the last block of a 1000-block create), by square-and-multiply
    S = 32
    files = [deterministic(31, 70)]
    slices = [files[0][o:o + S].ljust(S, b"\0") for o in range(0, 70, S)]
    bases = O.input_bases(len(slices))
    acc = bytearray(S)
    for i, sl in enumerate(slices):
        c, b, e = 1, bases[i], 999
        while e:
            if e & 1:
                c = clmul_mod(c, b)
            b = clmul_mod(b, b)
            e >>= 1
        acc = bytearray(x ^ y for x, y in zip(acc, scalar_mul_xor_bytes(c, sl)))
    assert O.encode_rows(files, S, 999, 1) == bytes(acc)


def test_encode_is_linear_and_row0_is_parity():
    S, R = 256, 6
    a = deterministic(11, 1000)
    b = deterministic(12, 1000)
    ab = bytes(x ^ y for x, y in zip(a, b))
    ea, eb, eab = O.encode([a], S, R), O.encode([b], S, R), O.encode([ab], S, R)
    assert eab == bytes(x ^ y for x, y in zip(ea, eb))
    # row 0 has every coefficient 1: the XOR of all zero-padded slices
    par = bytearray(S)
    for off in range(0, 1000, S):
        for k, v in enumerate(a[off:off + S]):
            par[k] ^= v
    assert ea[:S] == bytes(par)


def test_simd_baseline_equals_scalar_for_every_dispatch():
    files = [deterministic(5, 70000), deterministic(6, 4100)]
    ref = O.encode(files, 4096, 9)
    for disp in range(4):
        for threads in (1, 3, 8):
            assert O.encode_simd(files, 4096, 9, threads, disp) == ref, (disp, threads)
    # odd length tails and slice sizes that are not multiples of 64
    files = [deterministic(7, 777)]
    for S in (4, 12, 100, 132):
        ref = O.encode(files, S, 4)
        for disp in range(4):
            assert O.encode_simd(files, S, 4, 2, disp) == ref, (S, disp)


# -------------------------------------------------------------- hashes

RFC1321 = {
    b"": "d41d8cd98f00b204e9800998ecf8427e",
    b"a": "0cc175b9c0f1b6a831c399e269772661",
    b"abc": "900150983cd24fb0d6963f7d28e17f72",
    b"message digest": "f96b697d7cb7938d525a2f31aaf161d0",
    b"abcdefghijklmnopqrstuvwxyz": "c3fcd3d76192e4007dfb496cca67e13b",
    b"ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789":
        "d174ab98d277d9f5a5611c2c9f419d9f",
    b"1234567890" * 8: "57edf4a22be3c955ac49da2e2107b67a",
}


def test_md5_rfc1321_vectors():
    # src/md5_impl.rs:97-115, src/format.rs:125-134
    for msg, hexd in RFC1321.items():
        assert O.md5(msg).hex() == hexd


def test_md5_against_hashlib_and_reference_md5_ref():
    rnd = random.Random(3)
    ref_built = O.ref_md5(b"") is not None
    for n in list(range(0, 200)) + [1000, 4095, 4096, 4097, 65536 + 36, 524288]:
        d = bytes(rnd.randrange(256) for _ in range(min(n, 5000))) * (n // 5000 + 1)
        d = d[:n]
        assert O.md5(d) == hashlib.md5(d).digest(), n
        if ref_built:
            assert O.ref_md5(d) == hashlib.md5(d).digest(), n


def test_crc32_check_value_and_zlib():
    assert O.crc32(b"123456789") == 0xCBF43926
    assert O.crc32(b"") == 0
    rnd = random.Random(4)
    for n in (1, 3, 4, 17, 1000, 65536):
        d = bytes(rnd.randrange(256) for _ in range(n))
        assert O.crc32(d) == zlib.crc32(d)


# ---------------------------------------------------------------- scan

def test_scan_zero_pads_last_slice():
    # src/source.rs:374-392
    md5s, crcs = O.slice_checksums(b"ABCDE", 4)
    assert md5s == [hashlib.md5(b"ABCD").digest(), hashlib.md5(b"E\0\0\0").digest()]
    assert crcs == [zlib.crc32(b"ABCD"), zlib.crc32(b"E\0\0\0")]
    full, h16 = O.file_hashes(b"ABCDE")
    assert full == h16 == hashlib.md5(b"ABCDE").digest()


def test_scan_hash16k_truncates():
    # src/source.rs:394-404
    data = bytes(i & 0xFF for i in range(16384 + 100))
    full, h16 = O.file_hashes(data)
    assert h16 == hashlib.md5(data[:16384]).digest()
    assert full == hashlib.md5(data).digest()


def test_scan_slice_count_and_many_slices():
    # src/source.rs:428-469
    data = bytes(range(17))
    md5s, _ = O.slice_checksums(data, 8)
    assert len(md5s) == 3
    assert md5s[2] == hashlib.md5(bytes([16]) + b"\0" * 7).digest()
    data = bytes(((i * 0x9E37) >> 5) & 0xFF for i in range(4096 * 16))
    md5s, crcs = O.slice_checksums(data, 16)
    assert len(md5s) == 4096
    for i in (0, 1, 7, 1023, 4095):
        assert md5s[i] == hashlib.md5(data[16 * i:16 * i + 16]).digest()
        assert crcs[i] == zlib.crc32(data[16 * i:16 * i + 16])


def _slices_padded(data: bytes, S: int):
    """The zero-padded S-byte slices of data (src/source.rs:168-183)."""
    for o in range(0, len(data), S):
        yield data[o:o + S].ljust(S, b"\0")


@pytest.mark.parametrize("S", [4, 12, 100, 4096, 65536, 786432, 1048576])
@pytest.mark.parametrize("threads", [1, 3, 16])
def test_scan_mt_equals_hashlib_and_zlib(S, threads):
    # The multi-threaded scan (its own slicing-by-8 CRC-32) against library
    # MD5 and CRC-32 of each zero-padded slice, src/source.rs:168-183
    # (md5_impl::digest and crc32fast::hash of the padded slice buffer).
    import numpy as np

    rng = np.random.default_rng(S * 31 + threads)
    # whole slices plus a partial tail that is not a multiple of 4, 16 or 64
    n_whole = max(1, min(40, (3 << 20) // S))
    tail = max(1, (S * 5) // 7) if S > 4 else 3
    data = rng.integers(0, 256, n_whole * S + tail, dtype=np.uint8).tobytes()
    md5s, crcs = O.slice_checksums_mt(data, S, threads)
    want = list(_slices_padded(data, S))
    assert len(md5s) == len(crcs) == len(want) == n_whole + 1
    for i, sl in enumerate(want):
        assert md5s[i] == hashlib.md5(sl).digest(), (S, threads, i)
        assert crcs[i] == zlib.crc32(sl), (S, threads, i)
    # the serial scan agrees too
    assert O.slice_checksums(data, S) == (md5s, crcs)


def test_scan_mt_edge_lengths():
    # single short slice, exact multiples, one byte over
    rnd = random.Random(12)
    for S, L in ((4, 1), (4, 4), (8, 9), (4096, 4096), (4096, 4097), (65536, 3), (786432, 786431)):
        data = bytes(rnd.randrange(256) for _ in range(L))
        md5s, crcs = O.slice_checksums_mt(data, S, 3)
        want = list(_slices_padded(data, S))
        assert md5s == [hashlib.md5(s).digest() for s in want], (S, L)
        assert crcs == [zlib.crc32(s) for s in want], (S, L)


@pytest.mark.parametrize("S,first", [(4, 0), (100, 65530), (4096, 65533), (65540, 7), (524288, 65400)])
@pytest.mark.parametrize("threads", [1, 3, 16])
def test_recovery_md5_mt_equals_hashlib(S, first, threads):
    # build_packet's hash, src/format.rs:40-44, over the recovery body of
    # src/packet/recovery.rs:11-19: MD5(set_id || "PAR 2.0\0RecvSlic" ||
    # le32(exponent) || block).  S % 64 != 0 for 4, 100 and 65540.
    import numpy as np

    count = min(9, 65536 - first)
    rng = np.random.default_rng(S + first + threads)
    rec = rng.integers(0, 256, count * S, dtype=np.uint8).tobytes()
    sid = rng.integers(0, 256, 16, dtype=np.uint8).tobytes()
    got = O.recovery_md5_mt(rec, count, S, first, sid, threads)
    for r in range(count):
        msg = sid + b"PAR 2.0\0RecvSlic" + struct.pack("<I", first + r) + rec[r * S:(r + 1) * S]
        assert got[r] == hashlib.md5(msg).digest(), (S, first, threads, r)
        # and the whole packet the serial oracle builds carries the same hash
        if S <= 4096:
            assert O.recovery_packet(sid, first + r, rec[r * S:(r + 1) * S])[16:32] == got[r]


def test_file_id_derivation():
    # src/source.rs:414-426
    h16 = b"\x11" * 16
    length = 0x12345678
    name = b"example.dat"
    assert O.file_id(h16, length, name) == hashlib.md5(
        h16 + struct.pack("<Q", length) + name).digest()


# ------------------------------------------------------------- packets

TYPE_MAIN = b"PAR 2.0\0Main\0\0\0\0"
TYPE_FILE_DESC = b"PAR 2.0\0FileDesc"
TYPE_IFSC = b"PAR 2.0\0IFSC\0\0\0\0"
TYPE_RECV = b"PAR 2.0\0RecvSlic"
TYPE_CREATOR = b"PAR 2.0\0Creator\0"


def test_build_packet_layout_and_hash():
    # src/format.rs:90-123
    sid = b"\xAA" * 16
    body = bytes(8)
    pkt = O.build_packet(sid, TYPE_CREATOR, body)
    assert len(pkt) == 72
    assert pkt[0:8] == b"PAR2\0PKT"
    assert struct.unpack("<Q", pkt[8:16])[0] == 72
    assert pkt[32:48] == sid and pkt[48:64] == TYPE_CREATOR and pkt[64:] == body
    sid = b"\x11" * 16
    body = b"hello, par2!\0\0\0\0"
    pkt = O.build_packet(sid, TYPE_CREATOR, body)
    assert pkt[16:32] == hashlib.md5(sid + TYPE_CREATOR + body).digest()


def test_main_packet_sorts_ids_and_set_id_is_md5_of_body():
    # src/packet/main_packet.rs:49-102
    ids = [b"\x20" * 16, b"\x10" * 16, b"\x30" * 16]
    pkt, sid = O.main_packet(4096, ids)
    body = pkt[64:]
    assert body[12:28] == b"\x10" * 16 and body[28:44] == b"\x20" * 16 and body[44:60] == b"\x30" * 16
    pkt, sid = O.main_packet(4096, [b"\xAA" * 16])
    expect_body = struct.pack("<QI", 4096, 1) + b"\xAA" * 16
    assert sid == hashlib.md5(expect_body).digest()
    assert pkt[32:48] == sid
    pkt, _ = O.main_packet(0x123456789ABCDEF0, [])
    assert struct.unpack("<Q", pkt[64:72])[0] == 0x123456789ABCDEF0
    for n in range(8):
        pkt, _ = O.main_packet(4096, [bytes([i]) * 16 for i in range(n)])
        assert len(pkt) % 4 == 0
        assert struct.unpack("<I", pkt[72:76])[0] == n


def test_recovery_and_creator_packets():
    # src/packet/recovery.rs:26-58, src/packet/creator.rs:21-27
    pkt = O.recovery_packet(bytes(16), 0x1234, bytes([0x11, 0x22, 0x33, 0x44]))
    assert len(pkt) == 72
    assert struct.unpack("<I", pkt[64:68])[0] == 0x1234 and pkt[68:] == bytes([0x11, 0x22, 0x33, 0x44])
    pkt = O.recovery_packet(bytes(16), 0xFFFF, bytes(4))
    assert struct.unpack("<I", pkt[64:68])[0] == 0xFFFF
    sl = bytes((x ^ 0xAA) & 0xFF for x in range(256))
    pkt = O.recovery_packet(b"\xCD" * 16, 7, sl)
    assert pkt[68:] == sl and len(pkt) == 64 + 4 + 256
    assert pkt[16:32] == hashlib.md5(b"\xCD" * 16 + TYPE_RECV + struct.pack("<I", 7) + sl).digest()
    pkt = O.creator_packet(bytes(16))
    assert pkt[64:].startswith(b"Created by par2rust v0.1.0") and len(pkt) % 4 == 0
    assert len(pkt) == 64 + 28


# ------------------------------------------------------- volume splits

def test_exponential_split():
    # src/creator.rs:1347-1375
    E = O.SCHEME_EXPONENTIAL
    assert O.resolve_volume_sizes(1, E) == [1]
    assert O.resolve_volume_sizes(2, E) == [1, 1]
    assert O.resolve_volume_sizes(4, E) == [1, 1, 2]
    assert O.resolve_volume_sizes(10, E) == [1, 1, 2, 4, 2]
    assert O.resolve_volume_sizes(200, E) == [1, 1, 2, 4, 8, 16, 32, 64, 72]
    for total in list(range(1, 300)) + [1000, 3000, 65535]:
        s = O.resolve_volume_sizes(total, E)
        assert sum(s) == total and all(v > 0 for v in s)


def test_uniform_split_and_errors():
    # src/creator.rs:1493-1529
    U = O.SCHEME_UNIFORM
    assert O.resolve_volume_sizes(10, U, 5) == [2] * 5
    assert O.resolve_volume_sizes(13, U, 4) == [3, 3, 3, 4]
    assert O.resolve_volume_sizes(42, U, 1) == [42]
    assert O.resolve_volume_sizes(1000, U, 15) == [66] * 14 + [76]
    with pytest.raises(O.SchemeError):
        O.resolve_volume_sizes(10, U, 0)
    with pytest.raises(O.SchemeError):
        O.resolve_volume_sizes(3, U, 4)
    for total in range(1, 60):
        for count in range(1, total + 1):
            s = O.resolve_volume_sizes(total, U, count)
            assert len(s) == count and sum(s) == total


def test_limited_and_explicit():
    # src/creator.rs:1531-1578, 1422-1485
    E, U, X = O.SCHEME_EXPONENTIAL, O.SCHEME_UNIFORM, O.SCHEME_EXPLICIT
    s = O.resolve_volume_sizes(20, E, limited=True, limit_cap=3)
    assert all(0 < v <= 3 for v in s) and sum(s) == 20
    assert O.resolve_volume_sizes(20, U, 2, limited=True, limit_cap=4) == [4, 4, 2, 4, 4, 2]
    with pytest.raises(O.SchemeError):
        O.resolve_volume_sizes(20, E, limited=True, limit_cap=0)
    assert O.resolve_volume_sizes(6, X, explicit=[3, 3]) == [3, 3]
    with pytest.raises(O.SchemeError):
        O.resolve_volume_sizes(6, X, explicit=[3, 2])
    with pytest.raises(O.SchemeError):
        O.resolve_volume_sizes(6, X, explicit=[6, 0])
    # cap_volume_sizes cases, src/creator.rs:1531-1545
    assert O.resolve_volume_sizes(6, X, explicit=[1, 2, 3], limited=True, limit_cap=4) == [1, 2, 3]
    assert O.resolve_volume_sizes(13, X, explicit=[2, 10, 1], limited=True, limit_cap=3) == \
        [2, 3, 3, 3, 1, 1]
    assert O.resolve_volume_sizes(8, X, explicit=[8], limited=True, limit_cap=4) == [4, 4]


def test_nested_limited():
    # Limited { inner: Box<VolumeScheme> } nests (src/creator.rs:131-134,
    # :330-344): the inner Limited resolves first, the outer one caps its
    # sizes.  Limited(4, Limited(6, Explicit([10]))): [6, 4] -> [4, 2, 4],
    # which differs from one cap of 4 ([4, 4, 2]).
    E, U, X = O.SCHEME_EXPONENTIAL, O.SCHEME_UNIFORM, O.SCHEME_EXPLICIT
    assert O.resolve_volume_sizes(10, X, explicit=[10], limited=True, limit_cap=6,
                                  outer_caps=[4]) == [4, 2, 4]
    assert O.resolve_volume_sizes(10, X, explicit=[10], limited=True, limit_cap=4) == [4, 4, 2]
    # each level equals a single-level Limited over the level below, taken
    # as an explicit scheme
    for kind, cnt, total, caps in ((E, 0, 200, [50, 7, 3]), (U, 3, 100, [40, 9]),
                                   (E, 0, 1000, [5, 300])):
        below = O.resolve_volume_sizes(total, kind, cnt)
        for k, cap in enumerate(caps):
            below = O.resolve_volume_sizes(total, X, explicit=below, limited=True, limit_cap=cap)
            assert O.resolve_volume_sizes(total, kind, cnt, limited=True, limit_cap=caps[0],
                                          outer_caps=caps[1:k + 1]) == below
    # a zero cap at any level is an error, checked before the inner scheme's
    # own errors (Uniform count above the total would be one)
    for caps in ([0, 4], [4, 0], [3, 5, 0]):
        with pytest.raises(O.SchemeError):
            O.resolve_volume_sizes(10, U, 99, limited=True, limit_cap=caps[0], outer_caps=caps[1:])
    with pytest.raises(O.SchemeError):  # outer wrappers need the first one
        O.resolve_volume_sizes(10, E, outer_caps=[4])
