"""Benchmark of the PAR2 create hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload headline]

A step is one pass of the device hot path over the workload resident in
HBM: the GF(2^16) recovery encode (K1), the per-slice MD5 + CRC-32 of the IFSC
packet (K2, on a second stream, overlapping K1) and the recovery-packet MD5s
(K3, after K1).  `value` is source bytes per second over the K timed steps.

The headline workload is one 538,218,411-byte file of xorshift64 bytes
(seed 1) at slice size 524288 (1027 slices, the last one zero padded) with
200 recovery blocks.  The CPU baseline times the oracle's restatement of the
reference's rayon SIMD path (GFNI-AVX512 where the host has it) on the same
bytes, once, on at most 16 host threads, and checks the GPU output against it.
"""
from __future__ import annotations

import argparse
import json
import os
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="headline", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        log("bench.py: only --gpus 1 is supported (single-GPU build; see DESIGN.md on sharding)")
        return 2

    import numpy as np
    import torch

    import par2rust_amd as p2
    from par2rust_amd import synth

    nbytes, S, R, seed = WORKLOADS[args.workload]
    N = (nbytes + S - 1) // S
    log(f"bench: workload={args.workload} bytes={nbytes} S={S} N={N} R={R}")

    t = time.time()
    host = synth.xorshift64_bytes(nbytes, seed)
    log(f"bench: generated input in {time.time() - t:.1f}s")

    dev = torch.cuda.current_device()
    src = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(f"cuda:{dev}")
    rec = torch.empty(R * S, dtype=torch.uint8, device=src.device)
    slice_md5 = torch.empty(N * 16, dtype=torch.uint8, device=src.device)
    slice_crc = torch.empty(N, dtype=torch.int32, device=src.device)
    pkt_md5 = torch.empty(R * 16, dtype=torch.uint8, device=src.device)
    set_id = bytes(range(16))

    ctx = p2.Context(dev)
    s_enc = torch.cuda.Stream(device=src.device)
    s_sum = torch.cuda.Stream(device=src.device)
    ev_k1 = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
             for _ in range(args.steps)]
    ev_k2 = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
             for _ in range(args.steps)]
    ev_k3 = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
             for _ in range(args.steps)]

    def step(i=None):
        if i is not None:
            ev_k2[i][0].record(s_sum)
        ctx.dev_slice_checksums([src], S, slice_md5, slice_crc, stream=s_sum.cuda_stream)
        if i is not None:
            ev_k2[i][1].record(s_sum)
            ev_k1[i][0].record(s_enc)
        ctx.dev_encode([src], S, R, rec, stream=s_enc.cuda_stream)
        if i is not None:
            ev_k1[i][1].record(s_enc)
            ev_k3[i][0].record(s_enc)
        ctx.dev_recovery_packet_md5(rec, R, S, 0, set_id, pkt_md5, stream=s_enc.cuda_stream)
        if i is not None:
            ev_k3[i][1].record(s_enc)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"bench: {args.warmup} warmup steps done")

    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_step = 1000.0 * elapsed / args.steps
    log(f"bench: {args.steps} timed steps, {ms_step:.3f} ms/step")

    k1_ms = sum(a.elapsed_time(b) for a, b in ev_k1) / args.steps
    k2_ms = sum(a.elapsed_time(b) for a, b in ev_k2) / args.steps
    k3_ms = sum(a.elapsed_time(b) for a, b in ev_k3) / args.steps
    value = nbytes * args.steps / elapsed / 1e9
    gf_ops = N * R * (S // 2)
    alg_bytes = N * S + R * S  # source read once, recovery written once
    achieved = alg_bytes / (k1_ms / 1e3) / 1e9

    gpu_rec = rec.cpu().numpy().tobytes()
    gpu_md5 = slice_md5.cpu().numpy().tobytes()
    gpu_crc = slice_crc.cpu().numpy().astype(np.uint32).tolist()
    gpu_pkt = pkt_md5.cpu().numpy().tobytes()

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
                               "K3 packet md5",
                   "file_bytes": nbytes, "slice_size": S, "slices": N, "recovery": R,
                   "parallelism": "1 GPU"},
        "gf_mul_xor_per_s": round(gf_ops * args.steps / elapsed, 1),
        "kernel_ms": {"k1_encode": round(k1_ms, 4), "k2_slice_md5_crc": round(k2_ms, 4),
                      "k3_packet_md5": round(k3_ms, 4)},
        "roofline": {"bound": "hbm", "kernel": "k_encode_bits (K1)",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": None,
                     "alg_bytes_per_launch": alg_bytes},
    }

    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = max(1, min(16, os.cpu_count() or 1))
        t = time.perf_counter()
        cpu_rec = O.encode_simd([host], S, R, threads)
        cpu_md5, cpu_crc = O.slice_checksums_mt(host, S, threads)
        cpu_pkt = O.recovery_md5_mt(cpu_rec, R, S, 0, set_id, threads)
        ct = time.perf_counter() - t
        out["cpu_baseline"] = {
            "value": round(nbytes / ct / 1e9, 4), "unit": "GB/s", "cores": threads,
            "kind": "port",
            "sample": f"the full {args.workload} step once (encode {O.simd_dispatch()} + "
                      f"slice md5/crc32 + packet md5) on {threads} threads, {ct:.2f}s",
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
