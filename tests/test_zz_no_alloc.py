"""p2g_dev_create_pass_noalloc (the create pass whose scratch lives in the
library's static device array, for callers that cannot allocate) against
p2g_dev_create_pass on the same device bytes: recovery blocks, slice MD5 /
CRC-32 and packet MD5s bit-exact, on two calls in a row, the second with a
smaller set (the array keeps the first call's contents), plus the limits it
refuses with a clear error.  Calls go through the C ABI (ctypes)."""
import struct

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _outs(n, S, R):
    from par2rust_amd.hipmem import DeviceBuffer

    return [DeviceBuffer(R * S), DeviceBuffer(16 * n), DeviceBuffer(4 * n), DeviceBuffer(16 * R)]


def _get(outs, n, S, R):
    return (outs[0].download(R * S), outs[1].download(16 * n),
            list(struct.unpack(f"<{n}I", outs[2].download(4 * n))), outs[3].download(16 * R))


# (file lengths, slice size, recovery count): the first is the larger set
CASES = [
    ([538_218_411], 524_288, 200),               # the headline set: 1027 slices, one partial
    ([70_000, 4_100, 9, 130_000], 4096, 37),     # several files, ragged tails, small R
]


def test_noalloc_equals_the_allocating_pass_twice():
    import par2rust_amd as p2
    from par2rust_amd.hipmem import DeviceBuffer

    ctx = p2.Context(0)
    try:
        for k, (lens, S, R) in enumerate(CASES):
            srcs = [DeviceBuffer(max(16, n)) for n in lens]
            for j, (b, n) in enumerate(zip(srcs, lens)):
                ctx.dev_fill_xorshift64(b.addr, n, 500 + 10 * k + j)
            files = [(b.addr, n) for b, n in zip(srcs, lens)]
            n = sum(-(-x // S) for x in lens)
            sid = bytes((17 * k + i) & 0xFF for i in range(16))
            want, got = _outs(n, S, R), _outs(n, S, R)
            ctx.dev_create_pass(files, S, R, sid, *want)
            ctx.synchronize()
            print(f"noalloc case {k}: {len(lens)} files, S={S}, n={n}, R={R}", flush=True)
            ctx.dev_create_pass_noalloc(files, S, R, sid, *got)
            ctx.synchronize()
            assert _get(got, n, S, R) == _get(want, n, S, R), f"case {k}"
            if k == 1:
                # and the oracle for the small set
                host = [O.xorshift64_bytes(x, 500 + 10 * k + j) for j, x in enumerate(lens)]
                rec = O.encode(host, S, R)
                assert got[0].download() == rec
    finally:
        ctx.close()


def test_noalloc_limits_are_refused_with_a_message():
    import par2rust_amd as p2
    from par2rust_amd.hipmem import DeviceBuffer

    ctx = p2.Context(0)
    try:
        src = DeviceBuffer(1 << 20)
        outs = _outs(4096, 256, 1)
        with pytest.raises(p2.Par2Error) as e:   # more slices than one fused pass holds
            ctx.dev_create_pass_noalloc([(src.addr, 1 << 20)], 256, 1, bytes(16), *outs)
        assert "2048" in str(e.value)
        big = DeviceBuffer(8 << 20)
        outs = _outs(2048, 4096, 2048)
        with pytest.raises(p2.Par2Error) as e:   # fragments above the 64 MiB array
            ctx.dev_create_pass_noalloc([(big.addr, 8 << 20)], 4096, 2048, bytes(16), *outs)
        assert "64 MiB" in str(e.value)
        with pytest.raises(p2.Par2Error) as e:   # more than 64 files
            ctx.dev_create_pass_noalloc([(src.addr, 64)] * 65, 64, 1, bytes(16), *_outs(65, 64, 1))
        assert "64 files" in str(e.value)
    finally:
        ctx.close()
