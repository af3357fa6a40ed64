"""Host-side wrapper of the serving entry points (mkfServeCreate & co.) and the
seeded request stream used to exercise them.

Server.submit() takes a (Z, Y, X) float32/float64 tensor: a CUDA tensor is
read in place (the request is device-resident), a CPU tensor is copied in by
the server (pin it for the copy to overlap). The wrapper keeps the field, the
threshold maps and the output array alive until the result is polled. Each
request travels under an id of its own, so a caller's tag may repeat while
earlier requests with that tag are outstanding."""

import ctypes as C
import math

import numpy as np
import torch

from . import (ENC_F32, ENC_F64, MKF_SERVE_AFTER_STREAM, MKF_SERVE_ON_DEVICE, BinMapF64, MkfBatchShape,
               MkfServeConfig, MkfServeRequest, MkfServeResult, MkfServeStats, bin_map, check, last_error, load)


class Server:
    def __init__(self, budget=0, max_in_flight=0, max_batch=0, plane_pairs=0, segments=0):
        self.lib = load()
        cfg = MkfServeConfig()
        cfg.hbmBudget = int(budget)
        cfg.maxInFlight = int(max_in_flight)
        cfg.maxBatch = int(max_batch)
        cfg.shape = MkfBatchShape(int(plane_pairs), int(segments))
        self.handle = self.lib.mkfServeCreate(C.byref(cfg))
        if not self.handle:
            raise RuntimeError(f"mkfServeCreate failed: {last_error()}")
        self._live = {}  # request id -> (caller's tag, field, maps, output)
        self._next = 0
        self._buf = (MkfServeResult * 256)()

    def submit(self, field, levels, tag=None, stream=None):
        """Queues one request; returns its tag. levels: (relop, t) pairs or
        BinMapF64. A CUDA field is read after the work queued so far on
        `stream` (default: torch's current stream), so it needs no host sync."""
        if field.dim() != 3 or not field.is_contiguous() or field.dtype not in (torch.float32, torch.float64):
            raise ValueError("need a contiguous 3-D float32/float64 tensor")
        maps = (BinMapF64 * len(levels))()
        for j, lv in enumerate(levels):
            maps[j] = lv if isinstance(lv, BinMapF64) else bin_map(*lv)
        out = np.zeros((len(levels), 256), dtype=np.uint64)
        rid = self._next
        self._next += 1
        tag = rid if tag is None else int(tag)
        Z, Y, X = field.shape
        r = MkfServeRequest()
        r.pF = field.data_ptr()
        r.flags = 0
        if field.is_cuda:
            st = stream if stream is not None else torch.cuda.current_stream(field.device)
            r.flags = MKF_SERVE_ON_DEVICE | MKF_SERVE_AFTER_STREAM
            r.stream = st.cuda_stream if hasattr(st, "cuda_stream") else int(st)
        r.enc = ENC_F32 if field.dtype == torch.float32 else ENC_F64
        r.def_[0], r.def_[1], r.def_[2] = X, Y, Z
        r.nLevels = len(levels)
        r.pM = C.cast(maps, C.POINTER(BinMapF64))
        r.pBPFD = out.ctypes.data
        r.tag = rid
        self._live[rid] = (tag, field, maps, out)
        rc = self.lib.mkfServeSubmit(self.handle, C.byref(r))
        if rc != 0:
            del self._live[rid]
            raise RuntimeError(f"mkfServeSubmit failed ({rc}): {last_error()}")
        return tag

    def poll(self, wait_us=0):
        """Reported results: a list of dicts (tag, status, hist (levels, 256)
        uint64, batch, and the server-clock times)."""
        n = self.lib.mkfServePoll(self.handle, self._buf, len(self._buf), int(wait_us))
        if n < 0:
            raise RuntimeError(f"mkfServePoll failed: {last_error()}")
        res = []
        for i in range(n):
            r = self._buf[i]
            tag, _, _, out = self._live.pop(int(r.tag))
            res.append({"tag": tag, "status": int(r.status), "hist": out, "batch": int(r.batchSize),
                        "submit": r.tSubmit, "admit": r.tAdmit, "launch": r.tLaunch, "done": r.tDone})
        return res

    @property
    def outstanding(self):
        return len(self._live)

    def clock(self):
        return self.lib.mkfServeClock(self.handle)

    def stats(self):
        s = MkfServeStats()
        check(self.lib.mkfServeGetStats(self.handle, C.byref(s)), "stats")
        return {k: getattr(s, k) for k, _ in MkfServeStats._fields_}

    def close(self):
        if self.handle:
            self.lib.mkfServeDestroy(self.handle)
            self.handle = None
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def request_stream(n, seed=5, side_range=(64, 768), f64_share=0.25, levels_range=(1, 16), device_share=0.3,
                   n_sims=4, host_pool=24):
    """n request specs. Every field is a cube of side log-uniform in
    side_range, float64 with probability f64_share; each request asks for a
    level count uniform in levels_range, levels evenly spaced in [-1.5, 1.5]
    (">="). A device_share of requests are in-situ snapshots of one of n_sims
    simulations (device-resident); the others are host fields drawn from a
    pool of host_pool such fields (so host memory stays bounded).
    Returns (pool specs, sim specs, request dicts with levels, sim, pool)."""
    rng = np.random.default_rng(seed)

    def cube():
        side = int(round(math.exp(rng.uniform(math.log(side_range[0]), math.log(side_range[1])))))
        return side, torch.float64 if rng.random() < f64_share else torch.float32

    pool = [cube() for _ in range(host_pool)]
    sims = [cube() for _ in range(n_sims)]
    reqs = []
    for _ in range(n):
        nl = int(rng.integers(levels_range[0], levels_range[1] + 1))
        lv = [(">=", float(t)) for t in np.linspace(-1.5, 1.5, nl)] if nl > 1 else [(">=", 0.0)]
        if rng.random() < device_share:
            reqs.append({"levels": lv, "sim": int(rng.integers(0, n_sims)), "pool": None})
        else:
            reqs.append({"levels": lv, "sim": None, "pool": int(rng.integers(0, host_pool))})
    return pool, sims, reqs


def poisson_arrivals(n, rate, seed=6):
    """Arrival times (seconds from 0) of a Poisson process of the given rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
