"""GPU parity of the split-nibble encode (k_encode_nibble: 3-bit chunks of
each symbol looked up with v_perm_b32 in per-coefficient byte tables), the
third multiply variant, against the oracle: R = 1..16 (the small-R shapes
where the matrix-core form pads rows to 16), ragged last slices, slice sizes
not a multiple of 16 (word loads), several files, pushes in any order,
accumulate mode, device-resident inputs and the create pass running it."""
import struct

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def data(n: int, seed: int) -> bytes:
    return O.xorshift64_bytes(n, seed)


@pytest.fixture(scope="module")
def nib():
    import os

    import par2rust_amd as p2

    old = os.environ.get("P2G_ENCODE")
    os.environ["P2G_ENCODE"] = "nibble"
    try:
        c = p2.Context(0)
    finally:
        if old is None:
            del os.environ["P2G_ENCODE"]
        else:
            os.environ["P2G_ENCODE"] = old
    yield c
    c.close()


@pytest.mark.parametrize("R", list(range(1, 17)))
def test_nibble_small_r_ragged(nib, R):
    # two files with ragged tails (one ending inside a 16-byte chunk), S % 16 == 0
    lens, S = [70000, 4100 + 2 * R], 4096
    files = [data(n, 300 + 7 * R + i) for i, n in enumerate(lens)]
    assert nib.encode(files, S, R, push_bytes=3 * S) == O.encode(files, S, R)


@pytest.mark.parametrize("lens,S,R", [
    ([150, 64, 3], 64, 5),
    ([777], 12, 4),            # S % 16 != 0: word loads, word stores
    ([1000], 100, 3),
    ([999], 4, 2),             # the smallest slice
    ([5000], 80, 9),
    ([40000, 1, 123457], 65536, 12),
    ([1 << 20], 16384, 33),    # several row blocks of 8
])
def test_nibble_shapes(nib, lens, S, R):
    files = [data(n, 100 + i) for i, n in enumerate(lens)]
    want = O.encode(files, S, R) if sum(lens) * R < (4 << 20) else O.encode_simd(files, S, R)
    assert nib.encode(files, S, R, push_bytes=5 * S) == want


def test_nibble_every_symbol_value(nib):
    # one slice holding all 65536 symbols, R = 5: every chunk value of every
    # chunk position meets every coefficient row
    f = b"".join(struct.pack("<H", s) for s in range(65536))
    g = data(len(f), 5)
    assert nib.encode([f, g], len(f), 5) == O.encode([f, g], len(f), 5)


def test_nibble_dev_encode_accumulate(nib):
    from par2rust_amd.hipmem import DeviceBuffer

    lens, S, R = [123457, 65536, 9], 4096, 11
    files = [data(n, 40 + i) for i, n in enumerate(lens)]
    dev = [DeviceBuffer.from_bytes(f) for f in files]
    rec = DeviceBuffer(R * S)
    nib.dev_encode(dev, S, R, rec)
    nib.synchronize()
    assert rec.download() == O.encode_simd(files, S, R)
    nib.dev_encode(dev, S, R, rec, accumulate=True)
    nib.synchronize()
    assert rec.download() == bytes(R * S)


def test_nibble_create_pass_forced_by_the_cache(tmp_path):
    import os

    import par2rust_amd as p2
    from par2rust_amd.hipmem import DeviceBuffer

    files, S, R = [data(300000, 61), data(777, 62)], 8192, 3
    sid = data(16, 63)
    n = sum(-(-len(f) // S) for f in files)
    cache = tmp_path / "nibble.txt"
    cache.write_text(f"{p2.tune_bucket(n, R, S)} {p2.VARIANT_NIBBLE} 0 1.0 {p2.VARIANT_NIBBLE} 1.0\n")
    os.environ["P2G_TUNE_CACHE"] = str(cache)
    try:
        c = p2.Context(0)
    finally:
        del os.environ["P2G_TUNE_CACHE"]
    try:
        assert c.tune_get(n, R, S)["variant"] == p2.VARIANT_NIBBLE
        assert c.tune_get_encode(n, R, S)["variant"] == p2.VARIANT_NIBBLE
        dev = [DeviceBuffer.from_bytes(f) for f in files]
        outs = [DeviceBuffer(R * S), DeviceBuffer(16 * n), DeviceBuffer(4 * n), DeviceBuffer(16 * R)]
        c.dev_create_pass([(d.addr, len(f)) for d, f in zip(dev, files)], S, R, sid, *outs)
        c.synchronize()
        rec = O.encode(files, S, R)
        md5s, crcs = [], []
        for f in files:
            m, cr = O.slice_checksums(f, S)
            md5s += m
            crcs += cr
        assert outs[0].download() == rec
        assert outs[1].download() == b"".join(md5s)
        assert list(struct.unpack(f"<{n}I", outs[2].download())) == crcs
        assert outs[3].download() == b"".join(
            O.recovery_packet(sid, r, rec[r * S:(r + 1) * S])[16:32] for r in range(R))
        # the device encode takes the bucket's tuned kernel
        rec_dev = DeviceBuffer(R * S)
        c.dev_encode(dev, S, R, rec_dev)
        c.synchronize()
        assert rec_dev.download() == rec
    finally:
        c.close()
