"""GPU parity of the fused create pass (p2g_dev_create_pass): recovery blocks,
slice MD5 / CRC-32 and recovery packet MD5s of one launch, bit-exact against
the oracle, across band widths, ragged and unaligned slices, several packet
and slice hashing workgroups, sets above one descriptor table (a one-job
batched launch) and the bit-sliced variant's separate kernels."""
import os
import struct

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def data(n: int, seed: int) -> bytes:
    return O.xorshift64_bytes(n, seed)


def run_pass(c, files, S, R, set_id):
    from par2rust_amd.hipmem import DeviceBuffer

    dev = [DeviceBuffer.from_bytes(f) if f else DeviceBuffer(16) for f in files]
    items = [(d.addr, len(f)) for d, f in zip(dev, files)]
    n = sum((len(f) + S - 1) // S for f in files)
    rec = DeviceBuffer(max(1, R * S))
    md5 = DeviceBuffer(16 * max(1, n))
    crc = DeviceBuffer(4 * max(1, n))
    pkt = DeviceBuffer(16 * max(1, R))
    c.dev_create_pass(items, S, R, set_id, rec, md5, crc, pkt)
    c.synchronize()
    return (rec.download(R * S), md5.download(16 * n),
            list(struct.unpack(f"<{n}I", crc.download(4 * n))), pkt.download(16 * R))


def expected(files, S, R, set_id):
    big = sum(len(f) for f in files) * R >= (4 << 20)
    rec = O.encode_simd(files, S, R) if big else O.encode(files, S, R)
    md5s, crcs = [], []
    for f in files:
        m, cr = O.slice_checksums(f, S)
        md5s += m
        crcs += cr
    pkt = b"".join(O.recovery_packet(set_id, r, rec[r * S:(r + 1) * S])[16:32] for r in range(R))
    return rec, b"".join(md5s), crcs, pkt


@pytest.fixture
def band_ctx():
    """A context whose fused pass uses bands of P2G_BAND_COLS column groups."""
    import par2rust_amd as p2

    made = []

    def open_with(band_cols):
        old = os.environ.get("P2G_BAND_COLS")
        os.environ["P2G_BAND_COLS"] = str(band_cols)
        try:
            c = p2.Context(0)
        finally:
            if old is None:
                del os.environ["P2G_BAND_COLS"]
            else:
                os.environ["P2G_BAND_COLS"] = old
        made.append(c)
        return c

    yield open_with
    for c in made:
        c.close()


PASS_CASES = [
    # (file lengths, S, R)
    ([70000, 4100], 4096, 7),        # two files, ragged last slices
    ([777], 12, 4),                  # S % 16 != 0, one column group
    ([1000], 100, 3),                # S % 16 != 0, tail inside a word
    ([1 << 20], 16384, 33),          # three row blocks
    ([300000], 65536, 300),          # R = 300: five packet-hash workgroups
    ([4 * 600], 4, 5),               # 600 slices: ten slice-hash workgroups
    ([20000], 64, 600),              # R = 600: eight packet-hash workgroups, lanes loop
    ([150, 64, 3], 64, 1),           # R = 1
]


@pytest.mark.parametrize("lens,S,R", PASS_CASES)
def test_create_pass_matches_oracle(ctx, lens, S, R):
    files = [data(n, 300 + i) for i, n in enumerate(lens)]
    sid = data(16, 77)
    assert run_pass(ctx, files, S, R, sid) == expected(files, S, R, sid)


@pytest.mark.parametrize("band_cols,lens,S,R", [
    (4, [1 << 20], 65536, 20),       # 32 bands of 4 column groups
    (3, [23000, 7], 5000, 9),        # S not a multiple of 512: last band one group wide
    (1, [40000], 2048, 18),          # bands of one column group, block boundaries inside
])
def test_create_pass_many_bands(band_ctx, band_cols, lens, S, R):
    c = band_ctx(band_cols)
    files = [data(n, 400 + i) for i, n in enumerate(lens)]
    sid = data(16, 78)
    assert run_pass(c, files, S, R, sid) == expected(files, S, R, sid)


@pytest.mark.parametrize("lens,S,R", [
    ([3000 * 4], 4, 6),                          # 3000 slices of 4 bytes
    ([2500 * 64 + 10, 700 * 64 + 3], 64, 40),    # two files, ragged tails, 3202 slices
    ([5000 * 100 + 52], 100, 17),                # S % 16 != 0, 5001 slices
    ([9000 * 16], 16, 300),                      # 9000 slices: 3 chunks of 4096 per unit
])
def test_create_pass_beyond_one_descriptor_table(ctx, lens, S, R):
    # more than 2048 slices: the entry point runs the set as a one-job batched launch
    files = [data(n, 90 + i) for i, n in enumerate(lens)]
    sid = data(16, 91)
    assert run_pass(ctx, files, S, R, sid) == expected(files, S, R, sid)
    assert ctx.pass_timing()["encode"] == (0.0, 0.0)  # no fused pass ran


def test_create_pass_large_sets_queued(ctx):
    # three large sets queued before one synchronize: at most two plans stay
    # pending, the oldest is retired by the next call; every output is whole
    from par2rust_amd.hipmem import DeviceBuffer

    S, R = 32, 9
    sets = [[data(2100 * S + 8 * k, 600 + k)] for k in range(3)]
    sid = data(16, 95)
    outs = []
    for files in sets:
        dev = [DeviceBuffer.from_bytes(f) for f in files]
        n = sum((len(f) + S - 1) // S for f in files)
        o = (DeviceBuffer(R * S), DeviceBuffer(16 * n), DeviceBuffer(4 * n), DeviceBuffer(16 * R))
        ctx.dev_create_pass([(d.addr, len(f)) for d, f in zip(dev, files)], S, R, sid, *o)
        outs.append((dev, n, o))
    ctx.synchronize()
    for files, (_, n, (rec, md5, crc, pkt)) in zip(sets, outs):
        got = (rec.download(R * S), md5.download(16 * n),
               list(struct.unpack(f"<{n}I", crc.download(4 * n))), pkt.download(16 * R))
        assert got == expected(files, S, R, sid)


def test_create_pass_bitslice_beyond_one_launch(monkeypatch):
    # the bit-sliced variant runs the separate kernels at any size
    import par2rust_amd as p2

    monkeypatch.setenv("P2G_ENCODE", "bitslice")
    c = p2.Context(0)
    try:
        S, R = 4, 6
        files = [data(3000 * S, 96)]
        sid = data(16, 97)
        assert run_pass(c, files, S, R, sid) == expected(files, S, R, sid)
    finally:
        c.close()


def test_create_pass_twice_reuses_state(ctx):
    # queue, band counts and fragments are reset per pass; a smaller second pass
    # after a larger one must not see the first one's state
    sid = data(16, 92)
    a = [data(200000, 93)]
    b = [data(50000, 94)]
    assert run_pass(ctx, a, 8192, 12, sid) == expected(a, 8192, 12, sid)
    assert run_pass(ctx, b, 4096, 5, sid) == expected(b, 4096, 5, sid)


@pytest.mark.parametrize("lens,S,R", [([70000, 4100], 4096, 7), ([777], 12, 4), ([1 << 20, 9], 16384, 33),
                                      ([4096 * 10 + 1000], 4096, 5), ([65536 * 70 + 40000], 65536, 3)])
def test_create_pass_partial_slices_in_place(ctx, monkeypatch, lens, S, R):
    # no staging budget: partial slices are read in place (the general
    # instantiation of the pass, with tail words and masked word addresses).
    # The last two put a partial slice with whole 64-byte blocks beside whole
    # slices in one wave: its MD5 streams the wave's common whole prefix, then
    # selects the tail block and zero blocks per block.
    monkeypatch.setenv("P2G_STAGE_BYTES", "0")
    files = [data(n, 500 + i) for i, n in enumerate(lens)]
    sid = data(16, 79)
    assert run_pass(ctx, files, S, R, sid) == expected(files, S, R, sid)
