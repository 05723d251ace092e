"""GPU parity: the HIP kernels through the C ABI against the oracle.

Bit-exact on every case (all of this path is integer / byte work).  Sizes
are small enough for the scalar oracle where it is the checker; the SIMD
oracle (itself pinned to the scalar one in test_oracle.py) checks the larger
shapes.
"""
import hashlib
import os
import struct
import zlib

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def data(n: int, seed: int) -> bytes:
    return O.xorshift64_bytes(n, seed)


ENCODE_CASES = [
    # (file lengths, S, R)
    ([150, 64, 3], 64, 5),          # ragged tails, multiple files
    ([777], 12, 4),                 # S % 16 != 0 (word path)
    ([1000], 100, 3),               # S % 16 != 0, tail inside a word
    ([999], 4, 2),                  # smallest slice size
    ([5000], 80, 9),                # S % 64 != 0 with vector loads
    ([70000, 4100], 4096, 1),       # R = 1: parity row only
    ([70000, 4100], 4096, 7),
    ([70000, 4100], 4096, 8),       # one full recovery tile
    ([70000, 4100], 4096, 17),      # tile remainder
    ([1 << 20], 16384, 33),
    ([40000, 1, 123457], 65536, 12),
]


@pytest.mark.parametrize("lens,S,R", ENCODE_CASES)
def test_encode_job_matches_oracle(ctx, lens, S, R):
    files = [data(n, 100 + i) for i, n in enumerate(lens)]
    want = O.encode(files, S, R) if sum(lens) * R < (4 << 20) else O.encode_simd(files, S, R)
    got = ctx.encode(files, S, R, push_bytes=3 * S)  # several pushes per file
    assert got == want


def test_encode_all_zero_and_single_byte_files(ctx):
    files = [bytes(10000), b"\x01", data(4097, 9)]
    assert ctx.encode(files, 4096, 6) == O.encode(files, 4096, 6)


def test_encode_maximum_input_count(ctx):
    # N = 32768 inputs uses every base coprime to 65535 (src/creator.rs:103-106)
    f = data(4 * 32768, 3)
    want = O.encode_simd([f], 4, 3)
    assert ctx.encode([f], 4, 3) == want


def test_encode_config0_shape(ctx):
    f = data(4 << 20, 1)
    assert ctx.encode([f], 4096, 10) == O.encode_simd([f], 4096, 10)


def test_dev_encode_with_resident_inputs(ctx):
    from par2rust_amd.hipmem import DeviceBuffer, synchronize

    lens, S, R = [123457, 65536, 9], 4096, 11
    files = [data(n, 40 + i) for i, n in enumerate(lens)]
    dev = [DeviceBuffer.from_bytes(f) for f in files]
    rec = DeviceBuffer(R * S)
    ctx.dev_encode(dev, S, R, rec)
    ctx.synchronize()
    assert rec.download() == O.encode_simd(files, S, R)
    # accumulate: a second pass XORs the same contribution in again -> zero
    ctx.dev_encode(dev, S, R, rec, accumulate=True)
    ctx.synchronize()
    assert rec.download() == bytes(R * S)


def test_dev_encode_split_over_calls_equals_one_call(ctx):
    from par2rust_amd.hipmem import DeviceBuffer

    S, R = 4096, 5
    a, b = data(3 * S + 100, 50), data(2 * S, 51)
    da, db = DeviceBuffer.from_bytes(a), DeviceBuffer.from_bytes(b)
    rec = DeviceBuffer(R * S)
    rec.zero()
    # second file first: pushes may come in any order
    ctx.dev_encode([db], S, R, rec, first_input=4, total_inputs=6, accumulate=True)
    ctx.dev_encode([da], S, R, rec, first_input=0, total_inputs=6, accumulate=True)
    ctx.synchronize()
    assert rec.download() == O.encode([a, b], S, R)


def test_dev_checksums_and_packet_md5(ctx):
    from par2rust_amd.hipmem import DeviceBuffer

    S, R = 65536, 3
    files = [data(200001, 60), data(S, 61)]
    dev = [DeviceBuffer.from_bytes(f) for f in files]
    n = sum((len(f) + S - 1) // S for f in files)
    md5, crc = DeviceBuffer(16 * n), DeviceBuffer(4 * n)
    ctx.dev_slice_checksums(dev, S, md5, crc)
    rec = DeviceBuffer(R * S)
    ctx.dev_encode(dev, S, R, rec)
    sid = data(16, 62)
    pkt = DeviceBuffer(16 * R)
    ctx.dev_recovery_packet_md5(rec, R, S, 0, sid, pkt)
    ctx.synchronize()
    want_md5, want_crc = [], []
    for f in files:
        m, c = O.slice_checksums(f, S)
        want_md5 += m
        want_crc += c
    assert md5.download() == b"".join(want_md5)
    assert list(struct.unpack(f"<{n}I", crc.download())) == want_crc
    r = rec.download()
    assert r == O.encode(files, S, R)
    assert pkt.download() == b"".join(O.recovery_packet(sid, k, r[k * S:(k + 1) * S])[16:32]
                                      for k in range(R))


CHECKSUM_CASES = [(5, 4), (17, 8), (4096 * 16, 16), (1000, 100), (100001, 4096),
                  (524288 * 2 + 7, 524288), (65536 * 3, 65536), (12345, 12), (63, 64)]


@pytest.mark.parametrize("n,S", CHECKSUM_CASES)
def test_slice_checksums_match_oracle(ctx, n, S):
    d = data(n, 7)
    md5s, crcs = ctx.slice_checksums(d, S)
    w_md5, w_crc = O.slice_checksums(d, S)
    assert md5s == w_md5
    assert crcs == w_crc


def test_slice_checksums_library_routines(ctx):
    # ABCDE at S = 4, src/source.rs:374-392, against hashlib / zlib directly
    md5s, crcs = ctx.slice_checksums(b"ABCDE", 4)
    assert md5s == [hashlib.md5(b"ABCD").digest(), hashlib.md5(b"E\0\0\0").digest()]
    assert crcs == [zlib.crc32(b"ABCD"), zlib.crc32(b"E\0\0\0")]


@pytest.mark.parametrize("S,count,first", [(4, 3, 0), (64, 5, 100), (100, 2, 65533),
                                           (4096, 70, 0), (524288, 3, 7)])
def test_recovery_packet_md5(ctx, S, count, first):
    rec = data(S * count, 21)
    sid = data(16, 22)
    got = ctx.recovery_packet_md5(rec, count, S, first, sid)
    for r in range(count):
        pkt = O.recovery_packet(sid, first + r, rec[r * S:(r + 1) * S])
        assert got[r] == pkt[16:32]
        assert got[r] == hashlib.md5(sid + b"PAR 2.0\0RecvSlic" + struct.pack("<I", first + r)
                                     + rec[r * S:(r + 1) * S]).digest()


def test_encode_finish_packet_md5(ctx):
    S, R = 4096, 12
    f = data(50000, 31)
    sid = data(16, 32)
    rec, md5s = ctx.encode([f], S, R, set_id=sid)
    assert rec == O.encode([f], S, R)
    for r in range(R):
        assert md5s[r] == O.recovery_packet(sid, r, rec[r * S:(r + 1) * S])[16:32]
