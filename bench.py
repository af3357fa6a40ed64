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
  roofline.measured_roof
                   the practical read roof: a streaming-read kernel
                   (mkfReadRoof) over the field's own bytes, best of a small
                   sweep of load kinds and workgroups per CU
  cpu_baseline     the oracle's C restatement of the reference CPU path
                   (threshold + 64-bit chunk walk) on the same field bytes,
                   timed on this host in two modes: 1 thread (the reference
                   built with gcc) and the host's CPU share (its OpenACC
                   multicore target), 3 passes each, with the CPU model
  f64              the same field in float64 (8 GiB): its own autotuned
                   tiling, timing, roofline and oracle parity

The default run also reports the drop-in pair (binMapCUDA + mkfCUDAGetBPFD),
phase moments, the host-field stream (PCIe-inclusive), the batched form on
the uneven 64-field mix, and the serving form on two request streams.
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


def bench_batch(args, dev, stream):
    """The uneven batch: 64 GRF fields (each axis dim uniform in [32, 384],
    half float32 / half float64, sigma uniform in [1, 6]) x 8 levels >= t,
    t = linspace(-1.5, 1.5, 8) passed as float; one graph replay of the
    persistent batch kernel per step, against plain per-field launches."""
    import numpy as np
    import torch

    import mkf_amd as M
    from mkf_amd import synth
    from mkf_amd.batch import Batch

    lib = M.load()
    sptr = C.c_void_p(stream.cuda_stream)
    specs = synth.batch_mix(seed=args.batch_seed)
    t0 = time.perf_counter()
    fields = synth.make_fields(specs, dev)
    torch.cuda.synchronize()
    log(f"[batch] {len(fields)} fields generated in {time.perf_counter() - t0:.1f}s")
    levels = [float(v) for v in np.linspace(-1.5, 1.5, 8)]
    L = len(levels)
    lv = [[(">=", t) for t in levels] for _ in fields]
    nvox = sum(f.numel() for f in fields)
    nbytes = sum(f.numel() * f.element_size() for f in fields)
    work = nvox * L

    B = Batch(fields, lv, zero=True, stream=stream)
    if args.pmc_workload:
        # For counter runs: three launches of the default shape, nothing else.
        for _ in range(3):
            B.run(stream)
        torch.cuda.synchronize()
        log("[batch] counter workload done: 3 batched launches")
        B.close()
        return {}
    cands = [(pp, sg) for pp in (16, 32, 64, 128) for sg in (8, 2, 1)]
    t0 = time.perf_counter()
    best = B.autotune(cands, reps=3)
    info = B.info()
    log(f"[batch] autotune over {len(cands)} shapes in {time.perf_counter() - t0:.2f}s -> planePairs {best[0]}, "
        f"segments {best[1]}, {info['items']} items, grid {info['grid']}")

    def timed(fn, reps, min_s=1.0):
        t_w = time.perf_counter()
        n = 0
        while n < 2 or time.perf_counter() - t_w < min_s:
            fn()
            n += 1
            if n % 8 == 0:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
        t1 = time.perf_counter()
        for a, b in ev:
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t1) / reps
        ms = [a.elapsed_time(b) for a, b in ev]
        return wall * 1e3, sum(ms) / len(ms), min(ms)

    steps = max(5, args.steps)
    b_wall, b_ms, b_min = timed(lambda: B.run(stream), steps)
    batch_hist = B.out.cpu().numpy().view(np.uint64).copy()
    d_wall, d_ms, _ = timed(lambda: B.run(stream, graph=False), steps)
    log(f"[batch] graph replay {b_ms:.3f} ms (min {b_min:.3f}), direct launch {d_ms:.3f} ms; "
        f"{work / (b_wall * 1e-3) / 1e9:.0f} Gvox-levels/s, {nbytes / (b_ms * 1e-3) / 1e9:.0f} GB/s")

    # Plain 1: one mkfFusedBPFD launch per field and level (the field is read
    # once per level), asynchronous on the same stream.
    fused_out = torch.zeros((len(fields) * L, 256), dtype=torch.int64, device=dev)
    infos = [M.FieldInfo(f.data_ptr(), tuple(reversed(f.shape)), M.ENC_F32 if f.dtype == torch.float32 else M.ENC_F64)
             for f in fields]
    maps = [M.bin_map(">=", t) for t in levels]

    def plain_fused():
        fused_out.zero_()
        for i, fi in enumerate(infos):
            for j, bm in enumerate(maps):
                rc = lib.mkfFusedBPFD(None, fused_out[i * L + j].data_ptr(), None, fi.ref, C.byref(bm), None, sptr)
                if rc:
                    raise RuntimeError(M.last_error())

    f_wall, f_ms, _ = timed(plain_fused, max(3, steps // 2))
    fused_same = bool(np.array_equal(fused_out.cpu().numpy().view(np.uint64), batch_hist))
    log(f"[batch] plain fused launches: {f_wall:.3f} ms wall ({f_ms:.3f} ms GPU), same histograms: {fused_same}")

    # Plain 2: the reference's drop-in sequence per field and level:
    # binMapCUDA into a PBM buffer, then mkfCUDAGetBPFD (both blocking).
    drop_out = torch.zeros((len(fields) * L, 256), dtype=torch.int64, device=dev)
    maxw = max(((f.shape[2] + 31) // 32) * f.shape[1] * f.shape[0] for f in fields)
    pbm = torch.empty(maxw, dtype=torch.int32, device=dev)

    def plain_dropin():
        drop_out.zero_()
        torch.cuda.synchronize()
        for i, fi in enumerate(infos):
            for j, bm in enumerate(maps):
                org = M.BMOrg()
                if not lib.binMapCUDA(None, pbm.data_ptr(), C.byref(org), fi.ref, C.byref(bm)):
                    raise RuntimeError(M.last_error())
                lib.mkfCUDAGetBPFD(None, drop_out[i * L + j].data_ptr(), C.byref(org), pbm.data_ptr(), 3)

    plain_dropin()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    reps_d = 2
    for _ in range(reps_d):
        plain_dropin()
    torch.cuda.synchronize()
    dr_wall = (time.perf_counter() - t1) * 1e3 / reps_d
    drop_same = bool(np.array_equal(drop_out.cpu().numpy().view(np.uint64), batch_hist))
    log(f"[batch] plain drop-in calls: {dr_wall:.3f} ms wall, same histograms: {drop_same}")

    res = {
        "workload": f"{len(fields)} GRF fields, dims U[32,384] per axis, 32 f32 + 32 f64, sigma U[1,6], "
                    f"{L} levels >= linspace(-1.5,1.5) as float; batch seed {args.batch_seed}",
        "voxels": nvox, "field_bytes": nbytes, "levels": L,
        "value": round(work / (b_wall * 1e-3) / 1e9, 2), "unit": "Gvox-levels/s",
        "ms_per_batch": round(b_wall, 4), "kernel_ms_mean": round(b_ms, 4), "kernel_ms_min": round(b_min, 4),
        "direct_launch_ms": round(d_ms, 4),
        "hbm_GBps": round(nbytes / (b_ms * 1e-3) / 1e9, 1),
        "hbm_frac": round(nbytes / (b_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "shape": {"planePairs": best[0], "segments": best[1]}, "items": info["items"], "grid": info["grid"],
        "plain_fused_ms": round(f_wall, 4), "plain_dropin_ms": round(dr_wall, 4),
        "speedup_vs_plain_fused": round(f_wall / b_wall, 2), "speedup_vs_plain_dropin": round(dr_wall / b_wall, 2),
        "same_as_plain_fused": fused_same, "same_as_plain_dropin": drop_same,
    }
    if not args.no_cpu:
        from oracle import oracle as orc
        orc.set_threads(args.cpu_threads or cpu_share())
        t1 = time.perf_counter()
        ok = True
        for i, f in enumerate(fields):
            f_np = f.cpu().numpy()
            for j, bm in enumerate(maps):
                h, _ = orc.bpfd_field(f_np, bm.t[0], bm.m)
                ok &= bool(np.array_equal(h, batch_hist[i * L + j]))
        res["parity"] = "bit-exact vs oracle (all 512 histograms)" if ok else "MISMATCH vs oracle"
        log(f"[batch] oracle check of every histogram in {time.perf_counter() - t1:.1f}s: {ok}")
    B.close()
    del fields, fused_out, drop_out, pbm
    torch.cuda.empty_cache()
    return res


def serve_sweep(lib, dev, fields, reqs, rates, seconds, seed, tag):
    """One Poisson sweep through the server against plain per-request
    launches of the same requests in arrival order. Served: submissions at
    their arrival times, results polled as they come. Plain: per request,
    copy in if the field is on the host, one mkfFusedBPFD per level, copy
    the histograms out, synchronise. Latency is completion minus arrival;
    throughput is voxel-levels from the first arrival to the last
    completion."""
    import numpy as np
    import torch

    import mkf_amd as M
    from mkf_amd.serve import Server, poisson_arrivals

    NMAX = len(reqs)
    work = [fields[i].numel() * len(r["levels"]) for i, r in enumerate(reqs)]
    maps = [[M.bin_map(*lv) for lv in r["levels"]] for r in reqs]
    stream = torch.cuda.current_stream()
    sptr = C.c_void_p(stream.cuda_stream)
    S = Server(budget=64 << 30, max_in_flight=64, max_batch=32)
    for i in range(4):  # warm-up
        S.submit(fields[i], reqs[i]["levels"], tag=10 ** 9 + i)
    while S.outstanding:
        S.poll(wait_us=10000)

    def served(n, arr):
        done, t_sub = {}, S.clock()
        st0 = S.stats()
        i = 0
        while len(done) < n:
            now = S.clock() - t_sub
            while i < n and arr[i] <= now:
                S.submit(fields[i], reqs[i]["levels"], tag=i)
                i += 1
            wait = min(0.002, arr[i] - now) if i < n else 0.002
            for r in S.poll(wait_us=max(0, int(wait * 1e6))):
                done[r["tag"]] = r
        st1 = S.stats()
        lat = np.array([done[k]["done"] - t_sub - arr[k] for k in range(n)])
        span = max(done[k]["done"] for k in range(n)) - t_sub - arr[0]
        parts = {"queue_ms": [done[k]["admit"] - t_sub - arr[k] for k in range(n)],
                 "copy_ms": [done[k]["launch"] - done[k]["admit"] for k in range(n)],
                 "run_ms": [done[k]["done"] - done[k]["launch"] for k in range(n)]}
        batches = np.array([done[k]["batch"] for k in range(n)])
        launches = st1["launches"] - st0["launches"]
        return {"lat": lat, "span": span, "hist": [done[k]["hist"] for k in range(n)],
                "batch": float(np.mean(batches)),
                "batch_dist": {"p50": float(np.percentile(batches, 50)), "p90": float(np.percentile(batches, 90)),
                               "max": int(batches.max()), "launches": int(launches),
                               "requests_per_launch": round(n / max(launches, 1), 2)},
                "gpu_busy_frac": round((st1["gpuBusy"] - st0["gpuBusy"]) / span, 3),
                "bytes_in": st1["bytesIn"] - st0["bytesIn"],
                "parts": {k: round(float(np.mean(v)) * 1e3, 3) for k, v in parts.items()}}

    def plain(n, arr):
        hist = torch.zeros((16, 256), dtype=torch.int64, device=dev)
        out, lat = [], []
        torch.cuda.synchronize()
        t_sub = time.perf_counter()
        bytes_in = 0
        for i in range(n):
            now = time.perf_counter() - t_sub
            if now < arr[i]:
                time.sleep(arr[i] - now)
            f = fields[i]
            if not f.is_cuda:
                f = f.to(dev)
                bytes_in += f.numel() * f.element_size()
            Z, Y, X = f.shape
            fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), M.ENC_F32 if f.dtype == torch.float32 else M.ENC_F64)
            L = len(maps[i])
            hist[:L].zero_()
            for j, bm in enumerate(maps[i]):
                if lib.mkfFusedBPFD(None, hist[j].data_ptr(), None, fi.ref, C.byref(bm), None, sptr):
                    raise RuntimeError(M.last_error())
            out.append(hist[:L].cpu().numpy().view(np.uint64).copy())
            lat.append(time.perf_counter() - t_sub - arr[i])
        span = time.perf_counter() - t_sub - arr[0]
        return {"lat": np.array(lat), "span": span, "hist": out, "bytes_in": bytes_in}

    rows = []
    for rate in rates:
        n = int(min(NMAX, max(40, rate * seconds)))
        arr = poisson_arrivals(n, rate, seed=seed + int(rate))
        arr = arr - arr[0]
        sv = served(n, arr)
        pl = plain(n, arr)
        w = sum(work[:n])
        same = all(np.array_equal(a, b) for a, b in zip(sv["hist"], pl["hist"]))
        on_dev = np.array([reqs[k]["sim"] is not None for k in range(n)])

        def pct(lat, sel):
            v = lat[sel]
            return {"p50_ms": round(float(np.percentile(v, 50)) * 1e3, 3),
                    "p99_ms": round(float(np.percentile(v, 99)) * 1e3, 3)} if v.size else None
        row = {"rate": rate, "requests": n,
               "served": {"Gvox_levels_s": round(w / sv["span"] / 1e9, 2),
                          "p50_ms": round(float(np.percentile(sv["lat"], 50)) * 1e3, 3),
                          "p99_ms": round(float(np.percentile(sv["lat"], 99)) * 1e3, 3),
                          "pcie_GBps": round(sv["bytes_in"] / sv["span"] / 1e9, 2), "mean_batch": round(sv["batch"], 2),
                          "batch_dist": sv["batch_dist"], "gpu_busy_frac": sv["gpu_busy_frac"]},
               "plain": {"Gvox_levels_s": round(w / pl["span"] / 1e9, 2),
                         "p50_ms": round(float(np.percentile(pl["lat"], 50)) * 1e3, 3),
                         "p99_ms": round(float(np.percentile(pl["lat"], 99)) * 1e3, 3),
                         "pcie_GBps": round(pl["bytes_in"] / pl["span"] / 1e9, 2)},
               "same_histograms": same,
               "served_by_origin": {"device": pct(sv["lat"], on_dev), "host": pct(sv["lat"], ~on_dev)},
               "plain_by_origin": {"device": pct(pl["lat"], on_dev), "host": pct(pl["lat"], ~on_dev)},
               "served_mean_parts_ms": sv["parts"]}
        row["speedup"] = round(row["served"]["Gvox_levels_s"] / row["plain"]["Gvox_levels_s"], 2)
        log(f"[serve:{tag}] {rate:6.0f}/s x{n}: served {row['served']['Gvox_levels_s']:8.1f} Gvox-levels/s "
            f"p50 {row['served']['p50_ms']:9.2f} p99 {row['served']['p99_ms']:9.2f} ms (batch {sv['batch']:.1f}, "
            f"{sv['batch_dist']['launches']} launches, GPU busy {sv['gpu_busy_frac']:.0%}, "
            f"PCIe {row['served']['pcie_GBps']:.1f} GB/s) | plain {row['plain']['Gvox_levels_s']:8.1f} "
            f"p50 {row['plain']['p50_ms']:9.2f} p99 {row['plain']['p99_ms']:9.2f} ms | x{row['speedup']} same={same}")
        log(f"[serve:{tag}]          device-resident p50/p99 served {row['served_by_origin']['device']} plain "
            f"{row['plain_by_origin']['device']}; served mean parts {sv['parts']}")
        rows.append(row)
    S.close()
    return rows


def snapshot_rings(sim_spec, dev, seed, base_seed):
    """Per simulation, a ring of 3 device snapshots f(k+1) = f(k) + 0.01
    noise of a GRF (sigma 3)."""
    import torch
    from mkf_amd import synth
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    sims = []
    for i, (side, dt) in enumerate(sim_spec):
        ring = [synth.grf_field_direct((side, side, side), 3.0, seed=base_seed + i, device=dev, dtype=dt)]
        for _ in range(2):
            ring.append(ring[-1] + 0.01 * torch.randn(ring[-1].shape, generator=g, device=dev, dtype=dt))
        sims.append(ring)
    return sims


def bench_serve(args, dev):
    """Two request streams through the server (continuous batching, 64 GB
    budget), each at several Poisson rates beside plain per-request launches:
    * the configured stream: cube side log-uniform 64-768, f32:f64 3:1, 1-16
      levels, 30 % device-resident snapshots, host fields pinned (PCIe-bound);
    * a device-resident stream: the same distribution, every request an
      in-situ snapshot of one of 16 simulations (no PCIe), where batching
      has the GPU to itself."""
    import torch

    from mkf_amd import synth
    from mkf_amd.serve import request_stream

    lib = __import__("mkf_amd").load()
    NMAX = args.serve_requests
    pool_spec, sim_spec, reqs = request_stream(NMAX, seed=args.serve_seed)
    t0 = time.perf_counter()
    pool = []
    for i, (side, dt) in enumerate(pool_spec):
        f = synth.grf_field_direct((side, side, side), 3.0, seed=7000 + i, device=dev, dtype=dt)
        h = torch.empty(f.shape, dtype=dt, pin_memory=True)
        h.copy_(f)
        pool.append(h)
        del f
    sims = snapshot_rings(sim_spec, dev, args.serve_seed, 8000)
    fields, turn = [], [0] * len(sims)
    for r in reqs:
        if r["sim"] is None:
            fields.append(pool[r["pool"]])
        else:
            s = r["sim"]
            fields.append(sims[s][turn[s] % 3])
            turn[s] += 1
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log(f"[serve] {NMAX} requests over {len(pool)} pinned host fields and {len(sims)} simulations "
        f"ready in {time.perf_counter() - t0:.1f}s")
    rows = serve_sweep(lib, dev, fields, reqs, args.serve_rates, args.serve_seconds, args.serve_seed, "mixed")
    del fields, sims, pool
    torch.cuda.empty_cache()

    # Device-resident stream: 16 simulations, every request a snapshot.
    dseed = args.serve_seed + 100
    _, dsim_spec, dreqs = request_stream(NMAX, seed=dseed, device_share=1.0, n_sims=16)
    t0 = time.perf_counter()
    dsims = snapshot_rings(dsim_spec, dev, dseed, 9000)
    dfields, turn = [], [0] * len(dsims)
    for r in dreqs:
        s = r["sim"]
        dfields.append(dsims[s][turn[s] % 3])
        turn[s] += 1
    torch.cuda.synchronize()
    log(f"[serve] device-resident stream: {NMAX} requests over {len(dsims)} simulations "
        f"ready in {time.perf_counter() - t0:.1f}s")
    drows = serve_sweep(lib, dev, dfields, dreqs, args.serve_device_rates, args.serve_seconds, dseed, "device")
    del dfields, dsims
    torch.cuda.empty_cache()
    return {"workload": f"Poisson arrivals; cube side log-uniform 64-768, f32:f64 3:1, 1-16 levels >= in [-1.5,1.5], "
                        f"30% device-resident snapshots of {len(sim_spec)} simulations (f += 0.01 noise), host fields "
                        f"pinned from a pool of {len(pool_spec)}; budget 64 GB; seed {args.serve_seed}",
            "rates": rows,
            "device_resident": {
                "workload": f"the same distribution, every request a device-resident snapshot of one of "
                            f"{len(dsim_spec)} simulations; seed {dseed}",
                "rates": drows}}


def cpu_share():
    """Worker threads for host-side work: the CPUs this process may run on,
    at most 16 (one GPU's share of its host)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def pick_tiling(lib, fi, bm, spec):
    """The tiling of the fused kernel: --tiling's, else the library's
    autotuner's choice on this device and field."""
    import mkf_amd as M
    if spec:
        return M.MkfTiling(*[int(v) for v in spec.split(",")])
    tiling = M.MkfTiling()
    M.check(lib.mkfFusedAutotune(fi.ref, C.byref(bm), C.byref(tiling), C.c_void_p(0)), "mkfFusedAutotune")
    log(f"[bench] autotuned tiling: planePairs {tiling.planePairs}, waves {tiling.waves}, "
        f"rowPairs {tiling.rowPairs}, voxelsPerLane {tiling.voxelsPerLane}")
    return tiling


def fused_leg(lib, fi, bm, tiling, hist, stream, K, W, label, warm_s=2.0):
    """Warm-up (at least W launches and warm_s seconds), then K timed steps of
    the histogram clear plus one mkfFusedBPFD launch, with HIP events around
    each launch on the launch stream. Returns the per-launch kernel times, the
    wall time per step and the last step's histogram. The GPU reaches its
    steady clocks only after about a second of load: with 0.5 s of warm-up the
    first timed window of a process ran 4 % slower (0.690 against 0.664 ms on
    the 1024^3 f32 field, same box)."""
    import numpy as np
    import torch
    import mkf_amd as M
    sptr = C.c_void_p(stream.cuda_stream)

    def launch():
        rc = lib.mkfFusedBPFD(None, hist.data_ptr(), None, fi.ref, C.byref(bm), C.byref(tiling), sptr)
        if rc != 0:
            raise RuntimeError(f"mkfFusedBPFD failed: {M.last_error()}")

    t_w = time.perf_counter()
    n_w = 0
    while n_w < W or time.perf_counter() - t_w < warm_s:
        hist.zero_()
        launch()
        n_w += 1
        if n_w % 16 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
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
    log(f"[bench] {label}: {K} steps: {wall * 1e3 / K:.3f} ms/step wall, kernel {sum(kern_ms) / K:.3f} ms "
        f"(min {min(kern_ms):.3f})")
    return {"kern_ms": kern_ms, "wall_ms": wall * 1e3 / K, "hist": hist.cpu().numpy().view(np.uint64).copy()}


def traffic_of(S, dt, tiling):
    """HBM bytes per launch from the rocprof FETCH_SIZE / WRITE_SIZE runs in
    profiles/, per field shape, dtype and tiling (None if not measured)."""
    tpath = os.path.join(ROOT, "profiles", "fused_traffic.json")
    if not os.path.exists(tpath):
        return None
    with open(tpath) as fh:
        ent = json.load(fh).get(f"{S}_{dt}")
    if isinstance(ent, dict):
        ent = ent.get(f"{tiling.planePairs},{tiling.waves},{tiling.rowPairs},{tiling.voxelsPerLane}")
    return ent


def read_roof(lib, ptr, nbytes, stream):
    """Practical HBM read roof: mkfReadRoof over nbytes of device memory,
    non-temporal and ordinary 16-byte loads, waves taking blocks round robin
    or each streaming its own contiguous run, 8 or 16 KiB in flight per
    wave, 1, 2 or 4 workgroups of 4 waves per CU; the best mean of 10
    launches (after 3 untimed ones) per configuration."""
    import torch
    import mkf_amd as M
    sptr = C.c_void_p(stream.cuda_stream)
    out = torch.zeros(1, dtype=torch.int32, device="cuda")
    best = None
    rows = []
    for flags in (1, 3, 0, 2):
        for u in (8, 16):
            for bpc in (1, 2, 4):
                args = (C.c_void_p(ptr), nbytes, flags, u, bpc, C.c_void_p(out.data_ptr()), sptr)
                for _ in range(3):
                    M.check(lib.mkfReadRoof(*args), "mkfReadRoof")
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for _ in range(10):
                    M.check(lib.mkfReadRoof(*args), "mkfReadRoof")
                e1.record(stream)
                torch.cuda.synchronize()
                ms = e0.elapsed_time(e1) / 10
                gbps = nbytes / (ms * 1e-3) / 1e9
                rows.append({"nontemporal": bool(flags & 1), "chunked": bool(flags & 2), "kib_in_flight": u,
                             "blocks_per_cu": bpc, "ms": round(ms, 4), "GBps": round(gbps, 1)})
                if best is None or gbps > best["GBps"]:
                    best = rows[-1]
    log(f"[bench] read roof: {best['GBps']} GB/s ({'non-temporal' if best['nontemporal'] else 'ordinary'} loads, "
        f"{'chunked' if best['chunked'] else 'round robin'}, {best['kib_in_flight']} KiB per wave in flight, "
        f"{best['blocks_per_cu']} workgroups/CU) = {best['GBps'] / HBM_PEAK_GBS:.1%} of the 8 TB/s peak")
    return {"GBps": best["GBps"], "frac_of_peak": round(best["GBps"] / HBM_PEAK_GBS, 4),
            "kernel": "mkfReadRoof (streaming 16-byte loads, XOR checksum)", "bytes": nbytes,
            "best": best, "sweep": rows}


def cpu_baseline(f_np, bm, threads):
    """The oracle's C restatement of the reference CPU path
    (mkfAccGetBPFDSimple: threshold + 64-bit chunk walk) on the same field
    bytes, in two modes: 1 thread, as the reference runs when built with gcc
    (its OpenACC pragmas inert), on a 129-plane slab of the field; and the
    host's CPU share (plane-pair ranges on pthreads, private histograms),
    which models the OpenACC multicore target, on the whole field. Each mode
    is timed over 3 passes; the best pass gives its rate. Returns the bench
    line's cpu_baseline object and the whole field's histogram."""
    from oracle import oracle as orc
    Z, Y, X = f_np.shape
    sample = f_np[: min(Z, 129)]

    def passes(arr, nthreads):
        orc.set_threads(nthreads)
        ts, h = [], None
        for _ in range(3):
            t1 = time.perf_counter()
            h, _ = orc.bpfd_field(arr, bm.t[0], bm.m)
            ts.append(time.perf_counter() - t1)
        return ts, h

    t1s, _ = passes(sample, 1)
    tNs, hist = passes(f_np, threads)
    one = sample.size / min(t1s) / 1e9
    alln = f_np.size / min(tNs) / 1e9
    log(f"[bench] CPU oracle: 1 thread {one:.3f} Gvox/s ({sample.shape[0]} planes), {threads} threads "
        f"{alln:.3f} Gvox/s (whole field); host: {os.cpu_count()} CPUs, {cpu_model()}")
    return {
        "value": round(alln, 4), "unit": "Gvox/s", "cores": threads, "kind": "port",
        "sample": f"all-cores mode: the whole {X}x{Y}x{Z} field (same bytes as the GPU run), best of 3 passes "
                  f"({', '.join(f'{t:.2f}' for t in tNs)} s); C restatement of mkfAccGetBPFDSimple (gcc -O3 -march=x86-64-v3, "
                  f"pthreads over plane pairs, private histograms)",
        "modes": {
            "1_thread": {"value": round(one, 4), "unit": "Gvox/s", "threads": 1,
                         "sample": f"planes 0..{sample.shape[0] - 1} of the field ({sample.size} voxels), "
                                   f"best of 3 passes ({', '.join(f'{t:.2f}' for t in t1s)} s)"},
            "all_cores": {"value": round(alln, 4), "unit": "Gvox/s", "threads": threads,
                          "sample": f"the whole field ({f_np.size} voxels), best of 3 passes"},
        },
        "host_cpus": os.cpu_count(), "cpus_available": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
        "cpu_model": cpu_model(),
    }, hist


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
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = the CPUs available, at most 16")
    ap.add_argument("--no-f64", action="store_true", help="skip the float64 leg of the default run")
    ap.add_argument("--tiling", type=str, default="",
                    help="planePairs,waves,rowPairs[,voxelsPerLane] (default: library choice)")
    ap.add_argument("--mode", choices=["all", "fused", "batch", "serve"], default="all",
                    help="fused: the 1024^3 kernel only; batch: the uneven batch only; serve: the request "
                         "streams only; all: everything (default)")
    ap.add_argument("--serve-requests", type=int, default=300)
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--serve-seconds", type=float, default=2.0, help="arrival window per rate (capped by requests)")
    ap.add_argument("--serve-rates", type=float, nargs="+", default=[50, 100, 200, 500, 1000, 2000])
    ap.add_argument("--serve-device-rates", type=float, nargs="+", default=[500, 2000, 5000, 10000],
                    help="rates of the device-resident stream")
    ap.add_argument("--batch-seed", type=int, default=4)
    ap.add_argument("--pmc-workload", action="store_true",
                    help="for counter runs: 3 fused launches, then 3 copies of a 1 GiB buffer (the calibration "
                         "kernel for FETCH_SIZE / WRITE_SIZE); with --mode batch, 3 batched launches; prints no JSON")
    ap.add_argument("--dims", type=str, default="",
                    help="X,Y,Z of the --pmc-workload field (default: --size cubed)")
    ap.add_argument("--pmc-target", choices=("fused", "bpfd", "pbm"), default="fused",
                    help="with --pmc-workload: 'bpfd' makes the PBM once and runs 3 histograms of it instead "
                         "(with --tiling's tiling, else the library's default); 'pbm' runs 3 map passes (then the copies)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): "
              "there is no multi-GPU path (a field would shard along z, one slab and one shared plane per GPU, "
              "the 256-bin histograms summed).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import mkf_amd as M
    from mkf_amd import synth

    lib = M.load()
    dev = torch.device("cuda:0")
    if args.mode == "serve":
        res = bench_serve(args, dev)
        print(json.dumps({"metric": "served vs plain per-request launches, p50/p99 ms", "serve": res}), flush=True)
        return
    if args.mode == "batch":
        res = bench_batch(args, dev, torch.cuda.current_stream())
        if args.pmc_workload:
            return
        out = {"metric": "Gvox-levels/s batched uneven mix (64 fields x 8 levels)", "value": res["value"],
               "unit": res["unit"], "n_gpus": 1, "steps": max(5, args.steps), "warmup": args.warmup,
               "ms_per_step": res["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic: seeded GRF mix",
               "config": {"workload": res["workload"]}, "batched": res}
        print(json.dumps(out), flush=True)
        return
    S = args.size
    if args.dims and not args.pmc_workload:
        raise SystemExit("--dims applies to --pmc-workload only")
    FD = tuple(int(v) for v in args.dims.split(",")) if args.dims else (S, S, S)  # X, Y, Z
    log(f"[bench] generating {FD[0]}x{FD[1]}x{FD[2]} GRF sigma={args.sigma} seed={args.seed} ({args.dtype})")
    t0 = time.perf_counter()
    field64 = synth.grf_field(FD[::-1], args.sigma, args.seed, device=dev, dtype=torch.float64)
    field = field64.to(torch.float32) if args.dtype == "f32" else field64
    # The float64 leg runs on the same field before its rounding (default run only).
    f64_leg = args.dtype == "f32" and not args.no_f64 and not args.pmc_workload
    if not f64_leg and args.dtype == "f32":
        del field64
        field64 = None
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log(f"[bench] field ready in {time.perf_counter() - t0:.1f}s")

    enc = M.ENC_F32 if args.dtype == "f32" else M.ENC_F64
    eb = 4 if args.dtype == "f32" else 8
    fi = M.FieldInfo(field.data_ptr(), FD, enc)
    bm = M.bin_map(">=", args.threshold)
    stream = torch.cuda.current_stream()
    sptr = C.c_void_p(stream.cuda_stream)
    tiling = pick_tiling(lib, fi, bm, args.tiling)
    hist = torch.zeros(256, dtype=torch.int64, device=dev)

    def launch():
        rc = lib.mkfFusedBPFD(None, hist.data_ptr(), None, fi.ref, C.byref(bm), C.byref(tiling), sptr)
        if rc != 0:
            raise RuntimeError(f"mkfFusedBPFD failed: {M.last_error()}")

    if args.pmc_workload and args.pmc_target == "bpfd":
        # The PBM of the field once, then 3 histogram launches of the PBM.
        org = M.BMOrg()
        nwords = lib.setBMO(C.byref(org), M.dims(FD), b"\x00")
        pbm = torch.empty(nwords, dtype=torch.int32, device=dev)
        M.check(lib.mkfPBMAsync(pbm.data_ptr(), C.byref(org), fi.ref, C.byref(bm), sptr), "mkfPBMAsync")
        del field
        torch.cuda.synchronize()
        # --tiling selects the histogram's tiling; without it, the library's default
        bt = C.byref(tiling) if args.tiling else None
        for _ in range(3):
            M.check(lib.mkfBPFDAsync(hist.data_ptr(), C.byref(org), pbm.data_ptr(), bt, sptr), "mkfBPFDAsync")
        torch.cuda.synchronize()
        log("[bench] counter workload done: 1 PBM, 3 histograms of the PBM")
        return
    if args.pmc_workload:
        if args.pmc_target == "pbm":
            # 3 map passes of the field into a caller buffer.
            org = M.BMOrg()
            nwords = lib.setBMO(C.byref(org), M.dims(FD), b"\x00")
            pbm = torch.empty(nwords, dtype=torch.int32, device=dev)
            for _ in range(3):
                M.check(lib.mkfPBMAsync(pbm.data_ptr(), C.byref(org), fi.ref, C.byref(bm), sptr), "mkfPBMAsync")
            torch.cuda.synchronize()
            del pbm
        else:
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
        log(f"[bench] counter workload done: 3 {'map passes' if args.pmc_target == 'pbm' else 'fused launches'}, "
            "3 copies of 1 GiB")
        return

    leg = fused_leg(lib, fi, bm, tiling, hist, stream, args.steps, args.warmup, f"{S}^3 {args.dtype}")
    gpu_hist = leg["hist"]
    kern_ms = leg["kern_ms"]
    K = args.steps
    nvox = S * S * S
    ms_per_step = leg["wall_ms"]
    value = nvox / (ms_per_step * 1e-3) / 1e9
    kern_avg = sum(kern_ms) / len(kern_ms)
    bytes_per_launch = nvox * eb
    achieved = bytes_per_launch / (kern_avg * 1e-3) / 1e9

    workload = f"fused PBM+BPFD, {S}x{S}x{S} {'float32' if eb == 4 else 'float64'} GRF (sigma {args.sigma:g}), >= {args.threshold:g}"
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
                   "tiling": {"planePairs": tiling.planePairs, "waves": tiling.waves, "rowPairs": tiling.rowPairs,
                              "voxelsPerLane": tiling.voxelsPerLane, "autotuned": not args.tiling}},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_of(S, args.dtype, tiling),
                     "kernel_ms_mean": round(kern_avg, 4), "kernel_ms_min": round(min(kern_ms), 4),
                     "kernel_ms_median": round(float(np.median(kern_ms)), 4),
                     "bytes_per_launch": bytes_per_launch},
        "cpu_baseline": None,
    }

    # The practical read roof of this device: the streaming read kernel over
    # the field's own 4 GiB, best configuration of a small sweep.
    roof = read_roof(lib, field.data_ptr(), nvox * eb, stream)
    result["roofline"]["measured_roof"] = roof
    result["roofline"]["frac_of_measured_roof"] = round(achieved / roof["GBps"], 4)

    # The reference's two-kernel drop-in path on the same field: binMapCUDA
    # (PBM into a caller buffer) then mkfCUDAGetBPFD (histogram of the PBM),
    # timed with HIP events per kernel on the launch stream.
    org = M.BMOrg()
    nwords = lib.setBMO(C.byref(org), M.dims((S, S, S)), b"\x00")
    pbm = torch.empty(nwords, dtype=torch.int32, device=dev)
    h2 = torch.zeros(256, dtype=torch.int64, device=dev)
    pbm_ms, bpfd_ms = [], []
    for i in range(max(3, K // 4) + 2):
        e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        h2.zero_()
        e[0].record(stream)
        M.check(lib.mkfPBMAsync(pbm.data_ptr(), C.byref(org), fi.ref, C.byref(bm), sptr), "mkfPBMAsync")
        e[1].record(stream)
        M.check(lib.mkfBPFDAsync(h2.data_ptr(), C.byref(org), pbm.data_ptr(), None, sptr), "mkfBPFDAsync")
        e[2].record(stream)
        torch.cuda.synchronize()
        if i >= 2:
            pbm_ms.append(e[0].elapsed_time(e[1]))
            bpfd_ms.append(e[1].elapsed_time(e[2]))
    dropin_ok = bool(np.array_equal(h2.cpu().numpy().view(np.uint64), gpu_hist))
    pbm_bytes = nvox * eb + nwords * 4
    result["dropin"] = {
        "pbm_ms": round(sum(pbm_ms) / len(pbm_ms), 4), "bpfd_ms": round(sum(bpfd_ms) / len(bpfd_ms), 4),
        "pbm_GBps": round(pbm_bytes / (sum(pbm_ms) / len(pbm_ms) * 1e-3) / 1e9, 1),
        "bpfd_GBps": round(nwords * 4 / (sum(bpfd_ms) / len(bpfd_ms) * 1e-3) / 1e9, 1),
        "same_histogram_as_fused": dropin_ok,
    }
    log(f"[bench] drop-in path: PBM {result['dropin']['pbm_ms']:.3f} ms ({result['dropin']['pbm_GBps']} GB/s), "
        f"BPFD {result['dropin']['bpfd_ms']:.3f} ms, same histogram: {dropin_ok}")
    del pbm

    # Phase moments of the same field: one fused pass (histogram + moments)
    # against the histogram-only pass above plus a moments-only map pass.
    # One call of each from zero gives the results; the times are means over
    # back-to-back calls (whose outputs keep accumulating), so that the host's
    # launch work overlaps the GPU's as in the fused leg.
    hm = torch.zeros(256, dtype=torch.int64, device=dev)
    mom_f = torch.zeros(M.MKF_MOM_N, dtype=torch.float64, device=dev)
    mom_m = torch.zeros(M.MKF_MOM_N, dtype=torch.float64, device=dev)

    def mom_fused():
        M.check(lib.mkfFusedBPFDMoments(hm.data_ptr(), mom_f.data_ptr(), None, fi.ref, C.byref(bm), None, sptr),
                "mkfFusedBPFDMoments")

    def mom_map():
        M.check(lib.mkfPBMMomentsAsync(None, mom_m.data_ptr(), C.byref(org), fi.ref, C.byref(bm), sptr),
                "mkfPBMMomentsAsync")

    mom_fused()
    mom_map()
    torch.cuda.synchronize()
    mf, mm = mom_f.cpu().numpy(), mom_m.cpu().numpy()
    mom_hist = hm.cpu().numpy().view(np.uint64).copy()

    def back_to_back(fn, n):
        t_w = time.perf_counter()
        while time.perf_counter() - t_w < 0.3:
            fn()
        torch.cuda.synchronize()
        e = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        e[0].record(stream)
        for _ in range(n):
            fn()
        e[1].record(stream)
        torch.cuda.synchronize()
        return e[0].elapsed_time(e[1]) / n

    n_mom = max(10, K)
    fm, mo = back_to_back(mom_fused, n_mom), back_to_back(mom_map, n_mom)
    result["phase_moments"] = {
        "fused_ms": round(fm, 4), "moments_only_ms": round(mo, 4), "two_pass_ms": round(kern_avg + mo, 4),
        "same_histogram_as_fused": bool(np.array_equal(mom_hist, gpu_hist)),
        "moments_agree": bool(mf[4] == mm[4] and mf[5] == mm[5] and np.allclose(mf[:4], mm[:4], rtol=1e-12)),
        "phase_means": [round(float(mf[0] / max(mf[4], 1)), 6), round(float(mf[2] / max(mf[5], 1)), 6)],
    }
    log(f"[bench] phase moments: fused histogram+moments {fm:.3f} ms vs two passes {kern_avg + mo:.3f} ms "
        f"(moments-only map pass {mo:.3f} ms); same histogram {result['phase_moments']['same_histogram_as_fused']}, "
        f"moments agree {result['phase_moments']['moments_agree']}")
    del hm, mom_f, mom_m

    # The host-field path (PCIe-inclusive): the same field from pinned host
    # memory, streamed through the fused kernel in z-slabs.
    hostf = torch.empty(field.shape, dtype=field.dtype, pin_memory=True)
    hostf.copy_(field)
    hh = np.zeros(256, dtype=np.uint64)
    hs_ms = []
    for i in range(3):
        hh[:] = 0
        t1 = time.perf_counter()
        M.check(lib.mkfHostStreamBPFD(hh.ctypes.data_as(C.POINTER(C.c_size_t)), hostf.data_ptr(), enc,
                                      M.dims((S, S, S)), C.byref(bm), 0), "mkfHostStreamBPFD")
        if i > 0:
            hs_ms.append((time.perf_counter() - t1) * 1e3)
    hs = sum(hs_ms) / len(hs_ms)
    result["host_stream"] = {"ms": round(hs, 2), "Gvox_s": round(nvox / (hs * 1e-3) / 1e9, 2),
                             "pcie_GBps": round(nvox * eb / (hs * 1e-3) / 1e9, 1),
                             "same_histogram_as_fused": bool(np.array_equal(hh, gpu_hist))}
    log(f"[bench] host-field path (pinned, slab-streamed): {hs:.1f} ms = {result['host_stream']['Gvox_s']} Gvox/s, "
        f"{result['host_stream']['pcie_GBps']} GB/s over PCIe, same histogram: "
        f"{result['host_stream']['same_histogram_as_fused']}")
    del hostf

    if f64_leg:
        # The same field in float64 (8 GiB): its own tiling, timing and roofline.
        fi64 = M.FieldInfo(field64.data_ptr(), (S, S, S), M.ENC_F64)
        t64 = pick_tiling(lib, fi64, bm, "")
        h64 = torch.zeros(256, dtype=torch.int64, device=dev)
        leg64 = fused_leg(lib, fi64, bm, t64, h64, stream, K, args.warmup, f"{S}^3 f64")
        k64 = sum(leg64["kern_ms"]) / len(leg64["kern_ms"])
        a64 = nvox * 8 / (k64 * 1e-3) / 1e9
        roof64 = read_roof(lib, field64.data_ptr(), nvox * 8, stream)
        result["f64"] = {
            "workload": f"fused PBM+BPFD, {S}x{S}x{S} float64 GRF (the same field before rounding), >= {args.threshold:g}",
            "value": round(nvox / (leg64["wall_ms"] * 1e-3) / 1e9, 2), "unit": "Gvox/s",
            "ms_per_step": round(leg64["wall_ms"], 4),
            "tiling": {"planePairs": t64.planePairs, "waves": t64.waves, "rowPairs": t64.rowPairs,
                       "voxelsPerLane": t64.voxelsPerLane, "autotuned": True},
            "roofline": {"bound": "hbm", "achieved": round(a64, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(a64 / HBM_PEAK_GBS, 4), "traffic": traffic_of(S, "f64", t64),
                         "measured_roof": roof64, "frac_of_measured_roof": round(a64 / roof64["GBps"], 4),
                         "kernel_ms_mean": round(k64, 4), "kernel_ms_min": round(min(leg64["kern_ms"]), 4),
                         "kernel_ms_median": round(float(np.median(leg64["kern_ms"])), 4),
                         "bytes_per_launch": nvox * 8},
        }
        if not args.no_cpu:
            from oracle import oracle as orc
            orc.set_threads(cpu_share())
            ref64, _ = orc.bpfd_field(field64.cpu().numpy(), bm.t[0], bm.m)
            ok64 = bool(np.array_equal(ref64, leg64["hist"]))
            result["f64"]["parity"] = "bit-exact vs oracle" if ok64 else "MISMATCH vs oracle"
            log(f"[bench] f64 leg: GPU == CPU: {ok64}")
        del field64, h64
        torch.cuda.empty_cache()

    if not args.no_cpu:
        f_np = field.cpu().numpy()
        result["cpu_baseline"], cpu_hist = cpu_baseline(f_np, bm, args.cpu_threads or cpu_share())
        parity = bool(np.array_equal(cpu_hist, gpu_hist))
        log(f"[bench] GPU == CPU: {parity}")
        result["parity"] = "bit-exact vs oracle" if parity else "MISMATCH vs oracle"
        if not parity:
            print(json.dumps(result))
            sys.exit(1)
    if args.mode == "all":
        del field
        torch.cuda.empty_cache()
        result["batched"] = bench_batch(args, dev, stream)
        result["serve"] = bench_serve(args, dev)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
