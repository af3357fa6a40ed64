#!/usr/bin/env python3
"""Benchmark of the hot path: fused threshold + 2x2x2 pattern histogram
(PBM+BPFD) of a 1024^3 float32 Gaussian random field on one MI355X.

One step = one mkfFusedBPFD launch over the whole field (resident in HBM)
plus the 2 KiB histogram clear, launched on torch's current stream through
the C ABI. Prints ONE JSON line on stdout; progress goes to stderr.

  value            whole-job throughput, Gvoxels/s over the timed steps
  roofline         HBM roofline of the fused kernel: algorithmic bytes per
                   launch (voxels x element size) / mean launch time, from HIP
                   events recorded around each launch on the launch stream
  cpu_baseline     the oracle's C restatement of the reference CPU path
                   (threshold + 64-bit chunk walk) on the same field bytes,
                   timed on this host with `cores` threads
"""

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gvox/s fused PBM+BPFD 1024³ f32, % HBM roofline; served vs plain, p50/p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--sigma", type=float, default=4.0)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--threshold", type=float, default=0.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the oracle check")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(16, host cores)")
    ap.add_argument("--tiling", type=str, default="",
                    help="planePairs,waves,rowPairs[,voxelsPerLane] (default: library choice)")
    ap.add_argument("--pmc-workload", action="store_true",
                    help="for counter runs: 3 fused launches, then 3 copies of a 1 GiB buffer (the calibration "
                         "kernel for FETCH_SIZE / WRITE_SIZE); prints no JSON")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): "
              "this build has no multi-GPU path; see DESIGN.md for how the field would shard.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import mkf_amd as M
    from mkf_amd import synth

    lib = M.load()
    dev = torch.device("cuda:0")
    S = args.size
    log(f"[bench] generating {S}^3 GRF sigma={args.sigma} seed={args.seed} ({args.dtype})")
    t0 = time.perf_counter()
    field = synth.grf_field((S, S, S), args.sigma, args.seed, device=dev, dtype=torch.float64)
    if args.dtype == "f32":
        field = field.to(torch.float32)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log(f"[bench] field ready in {time.perf_counter() - t0:.1f}s")

    enc = M.ENC_F32 if args.dtype == "f32" else M.ENC_F64
    eb = 4 if args.dtype == "f32" else 8
    fi = M.FieldInfo(field.data_ptr(), (S, S, S), enc)
    bm = M.bin_map(">=", args.threshold)
    tiling = None
    if args.tiling:
        tiling = M.MkfTiling(*[int(v) for v in args.tiling.split(",")])
    else:
        tiling = lib.mkfDefaultTiling(M.dims((S, S, S)), eb)
    stream = torch.cuda.current_stream()
    sptr = C.c_void_p(stream.cuda_stream)
    hist = torch.zeros(256, dtype=torch.int64, device=dev)

    def launch():
        rc = lib.mkfFusedBPFD(None, hist.data_ptr(), None, fi.ref, C.byref(bm), C.byref(tiling), sptr)
        if rc != 0:
            raise RuntimeError(f"mkfFusedBPFD failed: {M.last_error()}")

    if args.pmc_workload:
        for _ in range(3):
            launch()
        torch.cuda.synchronize()
        del field
        torch.cuda.empty_cache()
        a = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("[bench] counter workload done: 3 fused launches, 3 copies of 1 GiB")
        return

    # Warm-up: at least W launches and at least 0.5 s of launches, so the
    # GPU has left its idle clocks before the timed region.
    t_w = time.perf_counter()
    n_w = 0
    while n_w < args.warmup or time.perf_counter() - t_w < 0.5:
        hist.zero_()
        launch()
        n_w += 1
        if n_w % 16 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()

    K = args.steps
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(K)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(K):
        hist.zero_()
        starts[i].record(stream)
        launch()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    gpu_hist = hist.cpu().numpy().view(np.uint64).copy()

    nvox = S * S * S
    ms_per_step = wall * 1e3 / K
    value = nvox / (ms_per_step * 1e-3) / 1e9
    kern_avg = sum(kern_ms) / len(kern_ms)
    bytes_per_launch = nvox * eb
    achieved = bytes_per_launch / (kern_avg * 1e-3) / 1e9
    log(f"[bench] {K} steps: {ms_per_step:.3f} ms/step wall, kernel {kern_avg:.3f} ms "
        f"(min {min(kern_ms):.3f}), {achieved:.0f} GB/s = {achieved / HBM_PEAK_GBS:.1%} of HBM peak")

    workload = f"fused PBM+BPFD, {S}x{S}x{S} {'float32' if eb == 4 else 'float64'} GRF (sigma {args.sigma:g}), >= {args.threshold:g}"
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "fused_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tab = json.load(fh)
        traffic = tab.get(f"{S}_{args.dtype}")

    result = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Gvox/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": f"synthetic: seeded GRF (N(0,1) noise, periodic Gaussian blur sigma={args.sigma:g}, standardised), seed {args.seed}",
        "config": {"workload": workload, "dims": [S, S, S], "elem_bytes": eb, "relation": ">=",
                   "threshold": args.threshold,
                   "tiling": {"planePairs": tiling.planePairs, "waves": tiling.waves, "rowPairs": tiling.rowPairs}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel_ms_mean": round(kern_avg, 4), "kernel_ms_min": round(min(kern_ms), 4),
                     "bytes_per_launch": bytes_per_launch},
        "cpu_baseline": None,
    }

    if not args.no_cpu:
        from oracle import oracle as orc
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        log(f"[bench] copying field to host for the CPU baseline ({threads} threads)")
        f_np = field.cpu().numpy()
        orc.set_threads(threads)
        t0 = time.perf_counter()
        cpu_hist, _ = orc.bpfd_field(f_np, bm.t[0], bm.m)
        cpu_s = time.perf_counter() - t0
        parity = bool(np.array_equal(cpu_hist, gpu_hist))
        log(f"[bench] CPU oracle: {cpu_s:.2f}s = {nvox / cpu_s / 1e9:.3f} Gvox/s; GPU == CPU: {parity}")
        result["cpu_baseline"] = {
            "value": round(nvox / cpu_s / 1e9, 4), "unit": "Gvox/s", "cores": threads, "kind": "port",
            "sample": f"the whole {S}^3 field (same bytes as the GPU run), one pass, "
                      f"{cpu_s:.2f}s; C restatement of mkfAccGetBPFDSimple (gcc -O3, pthreads over plane pairs)",
        }
        result["parity"] = "bit-exact vs oracle" if parity else "MISMATCH vs oracle"
        if not parity:
            print(json.dumps(result))
            sys.exit(1)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
