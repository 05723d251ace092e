"""Benchmark of the PAR2 create hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload headline]

Besides the headline line it reports, in the same JSON object: the whole
create's wall time on the headline file (end_to_end_create), the device pass
over the many-files workload (64 files, S = 64 KiB, R = 1000: many_files,
with its whole create from files),
the batched job mix against the same jobs run one at a time (batched), and a
serving run with Poisson arrivals and a 32 GiB job (serving); flags skip
each.

A step is one pass of the device hot path over the workload resident in
HBM: the GF(2^16) recovery encode (K1), the per-slice MD5 + CRC-32 of the IFSC
packet (K2) and the recovery-packet MD5s (K3).  By default a step is one
p2g_dev_create_pass call (all three in one persistent launch, the packet MD5
following the encode band by band); --mode separate runs K1, K2 and K3 as
their own launches (K2 on a second stream).  `value` is source bytes per
second over the K timed steps.

The headline workload is one 538,218,411-byte file of xorshift64 bytes
(seed 1) at slice size 524288 (1027 slices, the last one zero padded) with
200 recovery blocks.  The CPU baseline times the oracle's restatement of the
reference's rayon SIMD path (GFNI-AVX512 where the host has it) on the same
bytes, once, on at most 16 host threads, and checks the GPU output against it.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import struct
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "source GB/s at 1027×512KiB c=200; batched vs plain jobs/s + p99, bit-exact"

WORKLOADS = {
    # name: (file bytes, slice size, recovery count, seed)
    "headline": (538_218_411, 524_288, 200, 1),
    "cpu-case": (4 << 20, 4096, 10, 1),
}

HBM_PEAK_GBPS = 8000.0
FP4_DENSE_PEAK_TFLOPS = 10100.0  # MI355X dense MXFP4 matrix peak
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # 32-bit lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz
KERNEL_NAME = {"mfma": "k_encode_mfma (K1)", "bitslice": "k_encode_bits (K1)"}
PASS_KERNEL = "k_create_pass"
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", f"{k}_headline_traffic.md")
                 for k in (PASS_KERNEL, "k_encode_mfma")]


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE, each scaled by a plain copy's reading), or
    None when the summary has no figure for it."""
    for path in TRAFFIC_FILES:
        try:
            for line in open(path):
                if line.startswith("traffic_bytes_per_launch") and f" {kernel} " in line:
                    return int(line.split("=")[1].split()[0])
        except OSError:
            pass
    return None


def measured_counters(kernel: str):
    """MFMA-pipe busy and VALU-issue fractions of `kernel` from the committed
    counter summary (SQ_VALU_MFMA_BUSY_CYCLES and SQ_INSTS_VALU against
    GRBM_GUI_ACTIVE), or None when the summary has none."""
    import re

    path = os.path.join(ROOT, "profiles", f"{kernel}_headline_traffic.md")
    out = {}
    try:
        for line in open(path):
            m = re.search(r"MFMA pipe busy ([0-9.]+)%", line)
            if m:
                out["mfma_pipe_busy"] = round(float(m.group(1)) / 100, 3)
            m = re.search(r"VALU issue ([0-9.]+)% of cycles", line)
            if m:
                out["valu_issue"] = round(float(m.group(1)) / 100, 3)
    except OSError:
        return None
    if not out:
        return None
    out["source"] = f"profiles/{kernel}_headline_traffic.md (rocprofv3 --pmc, one dispatch)"
    return out


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_batch_mix(ctx, n_jobs: int, reps: int, tune: bool = True) -> dict:
    """The batched-create comparison: the job mix of synth.batch_mix, every
    file generated on the device, run (1) one job at a time through the plain
    device path (p2g_dev_create_pass, then p2g_synchronize, per job) and (2)
    as one batch (p2g_batch_plan + launch + wait, planning included), then
    (3) as replays of a planned batch (HIP graph).  Jobs/s = jobs / wall time
    of the whole run; latency of a job = its completion time from the start
    of the run, all jobs being submitted together at the start.  Every job's
    outputs of (2) are compared with those of (1).

    With `tune`, every (N, R, S) bucket of the mix is first tuned on
    synthetic input of its largest job's shape (p2g_autotune); both runs then
    take their choices from that cache: the plain path its variant and band
    width, the batched planner its band widths (the batched kernel is
    MFMA-only)."""
    import numpy as np

    import par2rust_amd as p2
    from par2rust_amd import hipmem, synth

    mix = synth.batch_mix(n_jobs)
    src_bytes = sum(sz for j in mix for sz, _ in j["files"])
    t = time.perf_counter()
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
        n = sum((sz + S - 1) // S for sz, _ in j["files"])
        files = [(arena.addr + off, sz) for (sz, _), off in zip(j["files"], offs)]
        outs = {m: [hipmem.DeviceBuffer(R * S), hipmem.DeviceBuffer(16 * n),
                    hipmem.DeviceBuffer(4 * n), hipmem.DeviceBuffer(16 * R)]
                for m in ("plain", "batch")}
        sid = bytes((k * 7 + i) & 0xFF for i in range(16))
        jobs.append({"S": S, "R": R, "n": n, "files": files, "arena": arena, "outs": outs,
                     "sid": sid})
    ctx.synchronize()
    log(f"bench: batch mix of {n_jobs} jobs, {src_bytes / 1e9:.2f} GB, generated on the device "
        f"in {time.perf_counter() - t:.1f}s")

    tuned = None
    if tune:
        t = time.perf_counter()
        shapes = {}
        for j in jobs:
            b = p2.tune_bucket(j["n"], j["R"], j["S"])
            shapes[b] = max(shapes.get(b, (0, 0, 0)), (j["n"], j["R"], j["S"]))
        picks = [ctx.autotune(*shape) for shape in shapes.values()]
        t_tune = time.perf_counter() - t
        bands = {}
        for c in picks:
            if c["variant"] == p2.VARIANT_MFMA:
                bands[str(c["band_cols"])] = bands.get(str(c["band_cols"]), 0) + 1
        names = {p2.VARIANT_MFMA: "mfma", p2.VARIANT_BITSLICE: "bitslice",
                 p2.VARIANT_NIBBLE: "nibble"}
        encode_kernels = {}
        for shape in shapes.values():
            k = names.get(ctx.tune_get_encode(*shape)["variant"], "other")
            encode_kernels[k] = encode_kernels.get(k, 0) + 1
        tuned = {"buckets": len(picks), "seconds": round(t_tune, 2),
                 "mfma_pass": sum(1 for c in picks if c["variant"] == p2.VARIANT_MFMA),
                 "bitslice": sum(1 for c in picks if c["variant"] == p2.VARIANT_BITSLICE),
                 "nibble": sum(1 for c in picks if c["variant"] == p2.VARIANT_NIBBLE),
                 "mfma_band_cols": bands, "encode_kernel": encode_kernels}
        log(f"bench: autotuned {len(picks)} buckets in {t_tune:.1f}s: {tuned}")

    def plain_run():
        lat = []
        t0 = time.perf_counter()
        for j in jobs:
            o = j["outs"]["plain"]
            ctx.dev_create_pass(j["files"], j["S"], j["R"], j["sid"], *o)
            ctx.synchronize()
            lat.append(time.perf_counter() - t0)
        return time.perf_counter() - t0, lat

    bjobs = [p2.BatchJob(j["files"], j["S"], j["R"], j["sid"], *j["outs"]["batch"]) for j in jobs]

    def batch_run():
        t0 = time.perf_counter()
        b = ctx.batch_plan(bjobs, use_graph=False)
        t_plan = time.perf_counter() - t0
        b.launch()
        b.wait()
        wall = time.perf_counter() - t0
        ms = np.array(b.job_ms()) / 1e3
        b.free()
        # device completion times, placed so that the last one is the host's end
        return wall, list(wall - (ms.max() - ms)), t_plan

    plain_run()  # warm-up: scratch buffers reach their sizes
    plain_wall, plain_lat = plain_run()
    log(f"bench: plain one job at a time {plain_wall:.3f}s")
    batch_run()
    runs = [batch_run() for _ in range(reps)]
    batch_wall, batch_lat, t_plan = min(runs, key=lambda r: r[0])
    log(f"bench: batched {batch_wall:.3f}s (planning {t_plan * 1e3:.1f} ms)")

    same = True
    for j in jobs:
        a, b = j["outs"]["plain"], j["outs"]["batch"]
        same &= all(x.download() == y.download() for x, y in zip(a, b))
    plan = ctx.batch_plan(bjobs, use_graph=True)
    plan.launch()
    plan.wait()
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.launch()
    plan.wait()
    replay_wall = (time.perf_counter() - t0) / reps
    groups_ms = plan.job_ms()
    plan.free()
    log(f"bench: graph replay {replay_wall:.3f}s per batch; batch == plain: {same}")

    pct = lambda v, q: round(float(np.percentile(np.array(v) * 1e3, q)), 2)
    fp4_ops = sum(256 * j["n"] * j["R"] * j["S"] for j in jobs)
    return {
        "jobs": n_jobs, "source_bytes": src_bytes, "autotune": tuned,
        "slice_sizes": {str(S): sum(1 for j in jobs if j["S"] == S) for S in synth.BATCH_SLICE_SIZES},
        "plain_jobs_per_s": round(n_jobs / plain_wall, 2),
        "batched_jobs_per_s": round(n_jobs / batch_wall, 2),
        "speedup": round(plain_wall / batch_wall, 3),
        "replay_jobs_per_s": round(n_jobs / replay_wall, 2),
        "plain_latency_ms": {"p50": pct(plain_lat, 50), "p99": pct(plain_lat, 99)},
        "batched_latency_ms": {"p50": pct(batch_lat, 50), "p99": pct(batch_lat, 99)},
        "batched_planning_ms": round(t_plan * 1e3, 2),
        "batched_source_GBps": round(src_bytes / batch_wall / 1e9, 2),
        "plain_source_GBps": round(src_bytes / plain_wall / 1e9, 2),
        "replay_last_job_ms": round(max(groups_ms), 2),
        # fp4 work of the GF(2) matrix products over the last replay's device
        # time, from the first batched launch's start to the last job's
        # completion stamp (the first group's fragment launch, ~15 ms with
        # the 64 GiB budget, and the staging copies come before it), and over
        # the replays' host wall time (every launch included)
        "fp4_ops": fp4_ops,
        "fp4_TFLOPs_device": round(fp4_ops / (max(groups_ms) * 1e-3) / 1e12, 1),
        "fp4_frac_device": round(fp4_ops / (max(groups_ms) * 1e-3) / 1e12 / FP4_DENSE_PEAK_TFLOPS, 4),
        "fp4_frac_replay": round(fp4_ops / replay_wall / 1e12 / FP4_DENSE_PEAK_TFLOPS, 4),
        "bit_exact_vs_plain": bool(same),
        "latency": "completion time from a common submission at t = 0",
    }


def run_serving(ctx, seconds: float, rate: float, budget: int, verify: int = 4,
                step_ms: float = 0.0) -> dict:
    """The serving run: requests of the batched job mix (synth.batch_mix,
    seed 4001) arrive as a Poisson process of `rate` jobs/s for `seconds`,
    plus one 32 GiB job (one file, slice size 1 MiB: 32768 slices, 3000
    recovery blocks, seed 4000) at 0.5 s; every file is written on the device
    by the server (seeded xorshift64).  The server admits requests by
    footprint within `budget` bytes of HBM and runs continuous batching steps.
    Latency of a request = time from its arrival to the poll that reports it
    done.  A sample of `verify` mix jobs is recomputed alone
    (p2g_dev_create_pass on the same device bytes) and compared."""
    import numpy as np

    from par2rust_amd import synth

    n_max = int(rate * seconds * 1.5) + 16
    mix = synth.batch_mix(n_max, seed=4001)
    u = synth._Uniform(4002)
    arrivals, t = [], 0.0
    while True:
        t += -math.log(1.0 - u()) / rate
        if t >= seconds or len(arrivals) >= len(mix):
            break
        arrivals.append(t)
    big = {"slice_size": 1 << 20, "recovery": 3000, "files": [(32 << 30, 4000)]}
    reqs = [(a, m) for a, m in zip(arrivals, mix)]
    reqs.append((0.5, big))
    reqs.sort(key=lambda r: r[0])
    t_open = time.perf_counter()
    srv = ctx.server(budget, step_ms)
    log(f"bench: serving {len(reqs)} requests over {seconds}s (rate {rate:.1f}/s, HBM budget "
        f"{budget / 1e9:.0f} GB, pool ready in {time.perf_counter() - t_open:.1f}s)")
    sample = set(range(1, len(reqs), max(1, len(reqs) // max(1, verify))))
    info, lat, keep, checked, same = {}, {}, [], 0, True
    max_q, max_used, k = 0, 0, 0
    src_bytes = sum(sz for _, m in reqs for sz, _ in m["files"])
    t0 = time.perf_counter()
    while len(lat) < len(reqs):
        now = time.perf_counter() - t0
        while k < len(reqs) and reqs[k][0] <= now:
            arr, m = reqs[k]
            sid = bytes((k * 13 + i) & 0xFF for i in range(16))
            rid = srv.submit([sz for sz, _ in m["files"]], m["slice_size"], m["recovery"], sid,
                             seeds=[sd for _, sd in m["files"]])
            info[rid] = (k, arr, m, sid)
            k += 1
        wait = 5.0 if k >= len(reqs) else max(0.0, min(5.0, (reqs[k][0] - now) * 1e3))
        for rid in srv.poll(wait):
            kk, arr, m, sid = info[rid]
            lat[rid] = time.perf_counter() - t0 - arr
            if kk in sample and m is not big:
                keep.append(rid)  # checked after the run, so the check does not share the GPU
            else:
                srv.release(rid)
        st = srv.stats()
        max_q, max_used = max(max_q, st["queued"]), max(max_used, st["hbm_used"])
        if time.perf_counter() - t0 > seconds + 120:
            raise RuntimeError("serving run did not finish")
    wall = time.perf_counter() - t0
    for rid in keep:
        kk, arr, m, sid = info[rid]
        same &= _same_alone(ctx, srv, rid, m, sid)
        checked += 1
        srv.release(rid)
    srv.close()
    big_rid = next(r for r, v in info.items() if v[2] is big)
    small = [v for r, v in lat.items() if r != big_rid]
    pct = lambda v, q: round(float(np.percentile(np.array(v) * 1e3, q)), 2)
    out = {
        "requests": len(reqs), "arrival_rate_jobs_per_s": round(rate, 1), "arrival_seconds": seconds,
        "hbm_budget_bytes": budget, "step_target_ms": step_ms or 35.0, "source_bytes": src_bytes,
        "source_GBps": round(src_bytes / wall / 1e9, 2), "wall_s": round(wall, 3),
        "latency_ms": {"p50": pct(small, 50), "p99": pct(small, 99), "max": pct(small, 100)},
        "big_job": {"bytes": 32 << 30, "slice_size": 1 << 20, "slices": 32768, "recovery": 3000,
                    "arrival_s": 0.5, "latency_ms": round(lat[big_rid] * 1e3, 1)},
        "max_queued": max_q, "max_hbm_used_bytes": max_used,
        "bit_exact_vs_alone": {"jobs_checked": checked, "equal": bool(same)},
        "latency": "arrival to completion reported by the server loop",
    }
    log(f"bench: serving done in {wall:.2f}s: {out['source_GBps']} GB/s, p50 "
        f"{out['latency_ms']['p50']} ms, p99 {out['latency_ms']['p99']} ms, big job "
        f"{out['big_job']['latency_ms']} ms, checked {checked} equal {same}")
    return out


def _same_alone(ctx, srv, rid, m, sid) -> bool:
    """The served outputs of request rid equal p2g_dev_create_pass of the
    same device bytes."""
    from par2rust_amd import hipmem

    import ctypes

    import par2rust_amd as p2

    job, _ = srv.result(rid)
    S, R = m["slice_size"], m["recovery"]
    df = ctypes.cast(job.files, ctypes.POINTER(p2._DevFile))
    files = [(df[f].data, df[f].len) for f in range(job.nfiles)]
    n = sum((L + S - 1) // S for _, L in files)
    outs = [hipmem.DeviceBuffer(R * S), hipmem.DeviceBuffer(16 * n), hipmem.DeviceBuffer(4 * n),
            hipmem.DeviceBuffer(16 * R)]
    ctx.dev_create_pass(files, S, R, sid, *outs)
    ctx.synchronize()
    served = srv.download(rid)
    alone = (outs[0].download(), outs[1].download(),
             list(struct.unpack(f"<{n}I", outs[2].download())), outs[3].download())
    return served == alone


def run_many_files(ctx, reps: int = 3) -> dict:
    """The many-files workload (synth.config2_files: 64 files log-uniform in
    [1 KiB, 64 MiB], 8 all-zero, 4 zero-length and left out as the
    reference's scan rejects them; S = 64 KiB, R = 1000), written on the
    device and run through p2g_dev_create_pass: more than 2048 slices, so a
    one-job batched launch.  Bytes against the oracle are checked by
    tests/test_gpu_fullsize.py."""
    from par2rust_amd import hipmem, synth

    S, R = 65536, 1000
    ents = [e for e in synth.config2_files() if e[2] != "empty"]
    offs, o = [], 0
    for _, sz, _, _ in ents:
        offs.append(o)
        o += (sz + 255) // 256 * 256
    arena = hipmem.DeviceBuffer(o)
    arena.zero()
    for (_, sz, kind, seed), off in zip(ents, offs):
        if kind == "random":
            ctx.dev_fill_xorshift64(arena.addr + off, sz, seed)
    files = [(arena.addr + off, sz) for (_, sz, _, _), off in zip(ents, offs)]
    n = sum(-(-sz // S) for _, sz, _, _ in ents)
    src = sum(sz for _, sz, _, _ in ents)
    outs = [hipmem.DeviceBuffer(R * S), hipmem.DeviceBuffer(16 * n), hipmem.DeviceBuffer(4 * n),
            hipmem.DeviceBuffer(16 * R)]
    sid = bytes(range(16))
    ctx.dev_create_pass(files, S, R, sid, *outs)
    ctx.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        ctx.dev_create_pass(files, S, R, sid, *outs)
        ctx.synchronize()
    dt = (time.perf_counter() - t) / reps
    ops = 256.0 * n * R * S
    log(f"bench: many files {len(files)} files {src / 1e6:.1f} MB N={n} R={R}: {dt * 1e3:.2f} ms")
    return {"files": len(files), "source_bytes": src, "slice_size": S, "slices": n, "recovery": R,
            "ms_per_pass": round(dt * 1e3, 3), "source_GBps": round(src / dt / 1e9, 2),
            "gf_mul_xor_per_s": round(n * R * (S / 2) / dt, 1),
            "fp4_TFLOPs": round(ops / dt / 1e12, 1), "fp4_frac": round(ops / dt / 1e12 / FP4_DENSE_PEAK_TFLOPS, 3),
            "path": "p2g_dev_create_pass (one-job batched launch: N > 2048)",
            "timing": "host wall time per call, synchronized, inputs resident in HBM"}


def run_many_files_end_to_end(ctx) -> dict:
    """Wall time of the whole create (p2g_run_create, every file read once
    into HBM) of the many-files workload (the 60 non-empty files of
    synth.config2_files, S = 64 KiB, R = 1000, uniform split into 15
    volumes) from the page cache of a temporary directory.  The bytes are
    checked against the committed manifest by tests/test_gpu_fullsize.py."""
    import shutil
    import tempfile

    import par2rust_amd as p2
    from par2rust_amd import synth

    tmp = tempfile.mkdtemp(prefix="p2g_e2e2_")
    try:
        inputs, total = [], 0
        for e in synth.config2_files():
            if e[2] == "empty":
                continue
            path = os.path.join(tmp, e[0])
            with open(path, "wb") as f:
                f.write(synth.config2_bytes(e))
            inputs.append((path, e[0].encode()))
            total += e[1]
        os.makedirs(os.path.join(tmp, "out"))
        opts = p2.CreateOptions(os.path.join(tmp, "out", "recovery.par2"), 65536, 1000,
                                p2.VolumeScheme.uniform(15))
        best = 1e9
        for _ in range(2):
            t = time.perf_counter()
            ctx.run_create(opts, inputs)
            best = min(best, time.perf_counter() - t)
        out = {"files": len(inputs), "file_bytes": total, "slice_size": 65536, "recovery": 1000,
               "gpu_resident_s": round(best, 3), "gpu_resident_GBps": round(total / best / 1e9, 3),
               "source": "page-cached files in a temporary directory"}
        log(f"bench: many files end to end {out}")
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def run_end_to_end(ctx, host: bytes, S: int, R: int, cpu: bool) -> dict:
    """Wall time of the whole create (SourceFile::scan of the file, then
    run_create writing the index and the volumes) on the headline file, read
    from the page cache of a temporary directory: p2g_run_create with every
    file read once into HBM, p2g_run_create streaming (scan, then a second
    read for the encode), and, with cpu, the oracle's run_create (the
    reference's CPU pipeline restated: SIMD encode on all host threads)."""
    import shutil
    import tempfile

    import par2rust_amd as p2

    tmp = tempfile.mkdtemp(prefix="p2g_e2e_")
    try:
        src = os.path.join(tmp, "input.bin")
        with open(src, "wb") as f:
            f.write(host)
        opts = p2.CreateOptions(os.path.join(tmp, "out", "recovery.par2"), S, R,
                                p2.VolumeScheme.exponential())
        os.makedirs(os.path.join(tmp, "out"))
        out = {"file_bytes": len(host), "slice_size": S, "recovery": R,
               "source": "page-cached file in a temporary directory"}
        for mode in ("resident", "stream"):
            if mode == "stream":
                os.environ["P2G_CREATE_STREAM"] = "1"
            best = 1e9
            for _ in range(2):
                t = time.perf_counter()
                ctx.run_create(opts, [(src, b"input.bin")])
                best = min(best, time.perf_counter() - t)
            os.environ.pop("P2G_CREATE_STREAM", None)
            out[f"gpu_{mode}_s"] = round(best, 3)
            out[f"gpu_{mode}_GBps"] = round(len(host) / best / 1e9, 3)
        if cpu:
            from oracle import oracle as O

            t = time.perf_counter()
            O.run_create(os.path.join(tmp, "out", "cpu.par2"), S, R, [(src, b"input.bin")],
                         O.SCHEME_EXPONENTIAL)
            ct = time.perf_counter() - t
            out["cpu_reference_port_s"] = round(ct, 3)
            out["cpu_threads"] = os.cpu_count()
            same = all(open(os.path.join(tmp, "out", a), "rb").read() ==
                       open(os.path.join(tmp, "out", a.replace("recovery", "cpu")), "rb").read()
                       for a in os.listdir(os.path.join(tmp, "out")) if a.startswith("recovery"))
            out["bit_exact_vs_cpu"] = bool(same)
        log(f"bench: end to end {out}")
        return out
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


ENCODE_VARIANT_SHAPES = [  # (slices, slice size, recovery count)
    (1024, 4096, 1), (1024, 4096, 2), (1024, 4096, 4), (1024, 4096, 8), (1024, 4096, 16),
    (1024, 4096, 32), (1027, 524288, 1), (1027, 524288, 4), (1027, 524288, 16), (1027, 524288, 64),
    (1027, 524288, 200), (16000, 65536, 1), (16000, 65536, 4), (16000, 65536, 16),
    (16000, 65536, 160), (2, 786432, 1), (20, 786432, 1), (200, 786432, 2), (1226, 786432, 13),
]


def run_encode_variants(reps: int) -> dict:
    """The encode kernels alone (p2g_dev_encode with P2G_ENCODE = mfma,
    bitslice, nibble) on one resident file of each shape: best wall time of
    `reps` synchronized calls after a warm-up call, and whether the three
    outputs are identical."""
    import par2rust_amd as p2
    from par2rust_amd import hipmem

    ctxs = {}
    for name in ("mfma", "bitslice", "nibble"):
        os.environ["P2G_ENCODE"] = name
        ctxs[name] = p2.Context(0)
    os.environ.pop("P2G_ENCODE", None)
    out = []
    for N, S, R in ENCODE_VARIANT_SHAPES:
        src = hipmem.DeviceBuffer(N * S)
        ctxs["mfma"].dev_fill_xorshift64(src.addr, N * S, 7)
        rec = hipmem.DeviceBuffer(R * S)
        ms, blocks = {}, {}
        for name, c in ctxs.items():
            c.dev_encode([src], S, R, rec)
            c.synchronize()
            best = 1e9
            for _ in range(reps):
                t = time.perf_counter()
                c.dev_encode([src], S, R, rec)
                c.synchronize()
                best = min(best, time.perf_counter() - t)
            ms[name] = round(best * 1e3, 4)
            blocks[name] = rec.download()
        same = blocks["mfma"] == blocks["bitslice"] == blocks["nibble"]
        out.append({"slices": N, "slice_size": S, "recovery": R, "ms": ms,
                    "fastest": min(ms, key=ms.get), "equal": bool(same)})
        log(f"bench: encode variants N={N} S={S} R={R}: {ms} equal {same}")
        del src, rec
    for c in ctxs.values():
        c.close()
    return {"shapes": out, "timing": "host wall time per synchronized p2g_dev_encode call, "
                                      "inputs resident in HBM"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="headline", choices=sorted(WORKLOADS))
    ap.add_argument("--mode", default="pass", choices=["pass", "separate"],
                    help="pass: one fused create pass per step; separate: K1, K2, K3 launches")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-jobs", type=int, default=256,
                    help="jobs of the batched-create comparison (0 skips it)")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the batch mix without tuning its buckets first")
    ap.add_argument("--serve-seconds", type=float, default=10.0,
                    help="arrival window of the serving run (0 skips it)")
    ap.add_argument("--serve-load", type=float, default=0.5,
                    help="arrival rate as a fraction of the measured batched jobs/s")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-step-ms", type=float, default=0.0,
                    help="the server's step target (0: its default, 35 ms)")
    ap.add_argument("--no-many-files", action="store_true",
                    help="skip the many-files (64 files, R = 1000) pass timing")
    ap.add_argument("--no-end-to-end", action="store_true",
                    help="skip the whole-create wall-time comparison")
    ap.add_argument("--encode-variants", action="store_true",
                    help="only time the three encode kernels alone over small- and large-R "
                         "shapes (JSON line) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run one 512 MiB device-to-device copy (the reference reading "
                         "for the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        log("bench.py: only --gpus 1 is supported (single-GPU build; see DESIGN.md on sharding)")
        return 2

    import numpy as np

    import par2rust_amd as p2
    from par2rust_amd import hipmem, synth

    if args.encode_variants:
        print(json.dumps({"encode_variants": run_encode_variants(max(1, args.steps))}), flush=True)
        return 0

    if args.copy_calibration:
        n = 512 << 20
        a, b = hipmem.DeviceBuffer(n), hipmem.DeviceBuffer(n)
        a.zero()
        hipmem.copy_device(b, a, n)
        hipmem.synchronize()
        print(json.dumps({"copy_calibration_bytes": n}), flush=True)
        return 0

    nbytes, S, R, seed = WORKLOADS[args.workload]
    N = (nbytes + S - 1) // S
    log(f"bench: workload={args.workload} bytes={nbytes} S={S} N={N} R={R}")

    t = time.time()
    host = synth.xorshift64_bytes(nbytes, seed)
    log(f"bench: generated input in {time.time() - t:.1f}s")

    ctx = p2.Context(0)
    src = hipmem.DeviceBuffer.from_bytes(host)
    rec = hipmem.DeviceBuffer(R * S)
    slice_md5 = hipmem.DeviceBuffer(N * 16)
    slice_crc = hipmem.DeviceBuffer(N * 4)
    pkt_md5 = hipmem.DeviceBuffer(R * 16)
    set_id = bytes(range(16))

    s_enc, s_sum = hipmem.Stream(), hipmem.Stream()
    kinds = ("pass",) if args.mode == "pass" else ("k1", "k2", "k3")
    ev = {k: [(hipmem.Event(), hipmem.Event()) for _ in range(args.steps)] for k in kinds}

    def step_pass(i=None):
        if i is not None:
            ev["pass"][i][0].record(s_enc)
        ctx.dev_create_pass([src], S, R, set_id, rec, slice_md5, slice_crc, pkt_md5,
                            stream=s_enc.addr)
        if i is not None:
            ev["pass"][i][1].record(s_enc)

    def step_separate(i=None):
        if i is not None:
            ev["k2"][i][0].record(s_sum)
        ctx.dev_slice_checksums([src], S, slice_md5, slice_crc, stream=s_sum.addr)
        if i is not None:
            ev["k2"][i][1].record(s_sum)
            ev["k1"][i][0].record(s_enc)
        ctx.dev_encode([src], S, R, rec, stream=s_enc.addr)
        if i is not None:
            ev["k1"][i][1].record(s_enc)
            ev["k3"][i][0].record(s_enc)
        ctx.dev_recovery_packet_md5(rec, R, S, 0, set_id, pkt_md5, stream=s_enc.addr)
        if i is not None:
            ev["k3"][i][1].record(s_enc)

    step = step_pass if args.mode == "pass" else step_separate
    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    log(f"bench: {args.warmup} warmup steps done")

    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    ctx.synchronize()  # also reports a pass that could not complete
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_step = 1000.0 * elapsed / args.steps
    log(f"bench: {args.steps} timed steps, {ms_step:.3f} ms/step")

    k_ms = {k: sum(a.elapsed_ms(b) for a, b in v) / args.steps for k, v in ev.items()}
    # the launch the roofline is about: the fused pass, or K1 alone
    k1_ms = k_ms["pass"] if args.mode == "pass" else k_ms["k1"]
    value = nbytes * args.steps / elapsed / 1e9
    gf_ops = N * R * (S // 2)
    alg_bytes = N * S + R * S  # source read once, recovery written once
    # The encode as a GF(2) matrix product (rows 16 R, k 16 N, columns S / 2):
    # 2 * 16R * 16N * S/2 = 256 N R S operations, run as dense fp4 MFMA.
    mfma_ops = 256 * N * R * S
    kname = KERNEL_NAME[os.environ.get("P2G_ENCODE", "mfma")]
    traffic_kernel = "k_encode_mfma"
    if args.mode == "pass" and kname.startswith("k_encode_mfma"):
        kname = (f"{PASS_KERNEL} (encode on MFMA + slice MD5/CRC-32 + packet MD5; timed with "
                 "its descriptor, fragment and reset launches)")
        traffic_kernel = PASS_KERNEL
    if "mfma" in kname or PASS_KERNEL in kname:
        roof = {"bound": "mfma", "kernel": kname,
                "achieved": round(mfma_ops / (k1_ms / 1e3) / 1e12, 2),
                "peak": FP4_DENSE_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(mfma_ops / (k1_ms / 1e3) / 1e12 / FP4_DENSE_PEAK_TFLOPS, 4),
                "traffic": measured_traffic(traffic_kernel),
                "counters": measured_counters(traffic_kernel),
                "alg_ops_per_launch": mfma_ops, "alg_bytes_per_launch": alg_bytes,
                "hbm_frac": round(alg_bytes / (k1_ms / 1e3) / 1e9 / HBM_PEAK_GBPS, 5)}
    else:
        # bit-sliced form: 16 x 16 masked XORs (one v_bitop3 each) per 32 symbols
        valu_ops = N * R * (S // 2) // 32 * 256
        roof = {"bound": "valu", "kernel": kname,
                "achieved": round(valu_ops / (k1_ms / 1e3) / 1e12, 2),
                "peak": round(VALU_PEAK_TOPS, 1), "unit": "Tlane-op/s",
                "frac": round(valu_ops / (k1_ms / 1e3) / 1e12 / VALU_PEAK_TOPS, 4),
                "traffic": None, "alg_bytes_per_launch": alg_bytes}

    gpu_rec = rec.download()
    gpu_md5 = slice_md5.download()
    gpu_crc = np.frombuffer(slice_crc.download(), dtype="<u4").tolist()
    gpu_pkt = pkt_md5.download()

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": f"synthetic: xorshift64 seed {seed}, {nbytes} bytes, resident in HBM",
        "config": {"workload": f"{args.workload}: 1 file x {nbytes} B, -s {S} "
                               f"({N} slices), -c {R}; K1 encode + K2 slice md5/crc32 + "
                               f"K3 packet md5 ({args.mode})",
                   "file_bytes": nbytes, "slice_size": S, "slices": N, "recovery": R,
                   "parallelism": "1 GPU"},
        "gf_mul_xor_per_s": round(gf_ops * args.steps / elapsed, 1),
        "kernel_ms": ({"create_pass": round(k_ms["pass"], 4),
                       "pass_roles_start_end_ms": ctx.pass_timing()} if args.mode == "pass" else
                      {"k1_encode": round(k_ms["k1"], 4), "k2_slice_md5_crc": round(k_ms["k2"], 4),
                       "k3_packet_md5": round(k_ms["k3"], 4)}),
        "roofline": roof,
    }
    if args.mode == "pass" and "mfma" in roof.get("bound", ""):
        # the encode role alone: from the pass's first stamp to the last encode
        # unit's (235 of 256 CUs encode; the hash roles end after it)
        enc = out["kernel_ms"]["pass_roles_start_end_ms"].get("encode", (0.0, 0.0))
        if enc[1] > 0:
            roof["encode_role"] = {
                "ms": round(enc[1], 4),
                "achieved": round(mfma_ops / (enc[1] / 1e3) / 1e12, 2),
                "frac": round(mfma_ops / (enc[1] / 1e3) / 1e12 / FP4_DENSE_PEAK_TFLOPS, 4),
                "note": "fp4 operations over the encode role's end in the last timed pass"}

    if not args.no_end_to_end:
        out["end_to_end_create"] = run_end_to_end(ctx, host, S, R, not args.no_cpu_baseline)
    if not args.no_many_files and args.workload == "headline":
        out["many_files"] = run_many_files(ctx)
        if not args.no_end_to_end:
            out["many_files"]["end_to_end_create"] = run_many_files_end_to_end(ctx)
    if args.batch_jobs > 0:
        out["batched"] = run_batch_mix(ctx, args.batch_jobs, args.batch_reps,
                                       tune=not args.no_autotune)
    if args.serve_seconds > 0:
        cap = out["batched"]["batched_jobs_per_s"] if "batched" in out else 400.0
        out["serving"] = run_serving(ctx, args.serve_seconds, args.serve_load * cap,
                                     int(args.serve_budget_gb * 1e9), step_ms=args.serve_step_ms)

    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = max(1, min(16, os.cpu_count() or 1))
        # three runs on all threads (a run is well under a second, so one
        # reading varies with the host's other load); the median is reported
        cts = []
        for _ in range(3):
            t = time.perf_counter()
            cpu_rec = O.encode_simd([host], S, R, threads)
            cpu_md5, cpu_crc = O.slice_checksums_mt(host, S, threads)
            cpu_pkt = O.recovery_md5_mt(cpu_rec, R, S, 0, set_id, threads)
            cts.append(time.perf_counter() - t)
        ct = sorted(cts)[1]
        # the same step on one core (the encode is the whole step's bulk and
        # scales with the thread count; the check below still uses the
        # multi-threaded outputs)
        t = time.perf_counter()
        O.encode_simd([host], S, R, 1)
        O.slice_checksums_mt(host, S, 1)
        O.recovery_md5_mt(cpu_rec, R, S, 0, set_id, 1)
        ct1 = time.perf_counter() - t
        out["cpu_baseline"] = {
            "value": round(nbytes / ct / 1e9, 4), "unit": "GB/s", "cores": threads,
            "kind": "port",
            "sample": f"the full {args.workload} step (encode {O.simd_dispatch()} + "
                      f"slice md5/crc32 + packet md5) on {threads} threads, median of 3 runs "
                      f"({min(cts):.2f}-{max(cts):.2f}s), then once on one core",
            "single_core": {"value": round(nbytes / ct1 / 1e9, 4), "unit": "GB/s",
                            "seconds": round(ct1, 2)},
            "host": {"cpu_model": _cpu_model(), "logical_cpus": os.cpu_count()},
        }
        out["parity"] = {
            "recovery_blocks": gpu_rec == cpu_rec,
            "slice_md5": gpu_md5 == b"".join(cpu_md5),
            "slice_crc32": gpu_crc == cpu_crc,
            "packet_md5": gpu_pkt == b"".join(cpu_pkt),
        }
        log(f"bench: cpu baseline {ct:.2f}s, parity {out['parity']}")

    ctx.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
