"""The batched-create workload at full size against the oracle.

The 256-job mix of par2rust_amd.synth.batch_mix (31.8 GB of sources: slice
sizes 4 KiB, 64 KiB and 768 KiB, sources log-uniform in [1 MiB, 1 GiB] over
1-16 files, recovery counts 1-10% of the slice count) is written on the
device, every (N, R, S) bucket is tuned first (p2g_autotune) so that the plan
takes its band widths from the autotuner cache, and the whole mix runs as one
batched plan.  Every job is then checked against the oracle computing that
job alone, as run_create would (src/creator.rs:883-1018 for the recovery
blocks, src/source.rs:168-183 for the slice checksums, src/format.rs:40-44
for the packet hashes): all recovery blocks (the oracle's GFNI restatement on
the host threads), every slice MD5 / CRC-32 and every packet MD5, bit-exact.
The mix holds the shapes that stress the batched kernel: a 4 KiB-slice job
of 32,695 slices (K chunks of 4096 inputs), 64 KiB and 768 KiB jobs of about
1 GB (row pieces, the longest serial hashes) and ragged last slices.
"""
import os
import struct
import time

import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_config3_mix_every_job_equals_the_oracle():
    import par2rust_amd as p2
    from par2rust_amd import hipmem, synth

    mix = synth.batch_mix(256)
    # the mix holds the shapes named above
    n_of = lambda j: sum(-(-sz // j["slice_size"]) for sz, _ in j["files"])
    size_of = lambda j: sum(sz for sz, _ in j["files"])
    assert max(n_of(j) for j in mix if j["slice_size"] == 4096) > 30000
    assert max(size_of(j) for j in mix if j["slice_size"] == 786432) > 900e6
    assert max(size_of(j) for j in mix if j["slice_size"] == 65536) > 1e9
    assert any(len(j["files"]) == 16 for j in mix)

    threads = max(1, min(16, os.cpu_count() or 1))
    ctx = p2.Context(0)
    try:
        t0 = time.time()
        jobs = []
        for k, j in enumerate(mix):
            S, R = j["slice_size"], j["recovery"]
            offs, o = [], 0
            for sz, _ in j["files"]:
                offs.append(o)
                o += (sz + 255) // 256 * 256
            arena = hipmem.DeviceBuffer(o)
            for (sz, seed), off in zip(j["files"], offs):
                ctx.dev_fill_xorshift64(arena.addr + off, sz, seed)
            n = n_of(j)
            outs = [hipmem.DeviceBuffer(R * S), hipmem.DeviceBuffer(16 * n),
                    hipmem.DeviceBuffer(4 * n), hipmem.DeviceBuffer(16 * R)]
            sid = bytes((k * 7 + i) & 0xFF for i in range(16))
            files = [(arena.addr + off, sz) for (sz, _), off in zip(j["files"], offs)]
            jobs.append((j, n, files, arena, outs, sid))
        ctx.synchronize()
        # tune every bucket on its largest job's shape, as the bench does
        shapes = {}
        for j, n, *_ in jobs:
            b = p2.tune_bucket(n, j["recovery"], j["slice_size"])
            shapes[b] = max(shapes.get(b, (0, 0, 0)), (n, j["recovery"], j["slice_size"]))
        picks = [ctx.autotune(*s) for s in shapes.values()]
        print(f"config3: {len(mix)} jobs generated, {len(picks)} buckets tuned "
              f"in {time.time() - t0:.1f}s", flush=True)
        plan = ctx.batch_plan([p2.BatchJob(files, j["slice_size"], j["recovery"], sid, *outs)
                               for j, n, files, arena, outs, sid in jobs], use_graph=False)
        plan.launch()
        plan.wait()
        plan.free()
        print(f"config3: batched run done at {time.time() - t0:.1f}s", flush=True)
        for k, (j, n, files, arena, outs, sid) in enumerate(jobs):
            S, R = j["slice_size"], j["recovery"]
            host = [hipmem.download_ptr(addr, sz) for addr, sz in files]
            rec = O.encode_simd(host, S, R, threads)
            assert outs[0].download() == rec, f"job {k} (S={S} n={n} R={R}): recovery blocks"
            md5s, crcs = [], []
            for f in host:
                m, c = O.slice_checksums_mt(f, S, threads)
                md5s += m
                crcs += c
            assert outs[1].download() == b"".join(md5s), f"job {k}: slice MD5"
            assert list(struct.unpack(f"<{n}I", outs[2].download())) == crcs, f"job {k}: slice CRC-32"
            assert outs[3].download() == b"".join(O.recovery_md5_mt(rec, R, S, 0, sid, threads)), \
                f"job {k}: packet MD5"
            del host, rec
            for b in outs:
                b.free()
            arena.free()
        print(f"config3: all {len(jobs)} jobs equal the oracle at {time.time() - t0:.1f}s", flush=True)
    finally:
        ctx.close()
