"""The configurations the path is measured on, at their real sizes, against
the oracle on the same bytes (the reference's own differential check is
compareNZ over CPU and GPU histograms, src/mkfTest.c:64-85, 101-176):

* 1024^3 Gaussian random field (sigma 4, standardised), threshold >= 0, in
  float32 and float64: the built-in tiling, every tiling the autotuner tries,
  the tuned choice, and the drop-in pair binMapCUDA + mkfCUDAGetBPFD (map
  words and histogram), plus the measures;
* the uneven batch: 64 fields (dims uniform in [32, 384], half float32, half
  float64, sigma uniform in [1, 6]) x 8 levels >= linspace(-1.5, 1.5): all
  512 histograms from the graph replay and from the direct launch;
* the serving stream: 40 Poisson-stream requests (cube sides log-uniform in
  64..768 with three sides above 600, float32:float64 3:1, 1-16 levels, 30 %
  device-resident snapshots of evolving simulations), each against the
  oracle, and one request alone against the same request inside the mix.
Integer results must be bit-exact; the tolerance is 0.
"""

import ctypes as C

import numpy as np
import pytest
import torch

import mkf_amd as M
from mkf_amd import synth
from mkf_amd.batch import Batch
from mkf_amd.serve import Server, request_stream

pytestmark = pytest.mark.gpu

S = 1024


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def grf1024(cuda):
    """The 1024^3 field in float64 (8 GiB) and float32 (its rounding), on the
    device and on the host, seed 2024 as the benchmark uses."""
    say("generating the 1024^3 GRF (sigma 4, seed 2024)")
    f64 = synth.grf_field((S, S, S), 4.0, 2024, device=cuda, dtype=torch.float64)
    f32 = f64.to(torch.float32)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    h64, h32 = f64.cpu().numpy(), f32.cpu().numpy()
    yield {"f64": (f64, h64), "f32": (f32, h32)}
    del f64, f32
    torch.cuda.empty_cache()


def fused(lib, field, bm, tiling=None):
    Z, Y, X = field.shape
    h = torch.zeros(256, dtype=torch.int64, device=field.device)
    fi = M.FieldInfo(field.data_ptr(), (X, Y, Z), M.ENC_F32 if field.dtype == torch.float32 else M.ENC_F64)
    t = M.MkfTiling(*tiling) if tiling is not None else None
    M.check(lib.mkfFusedBPFD(None, h.data_ptr(), None, fi.ref, C.byref(bm), C.byref(t) if t else None, None),
            f"mkfFusedBPFD {tiling}")
    torch.cuda.synchronize()
    return h.cpu().numpy().view(np.uint64).copy()


TILINGS = {
    # the built-in choice (None) and each candidate mkfFusedAutotune times
    # (171 and 341 plane pairs: the persistent kernels spread over three
    # quarters of their slots)
    "f32": [None, (128, 16, 4, 16), (128, 16, 4, 8), (64, 16, 2, 16), (256, 16, 4, 4), (128, 16, 4, 1),
            (171, 16, 4, 16), (341, 16, 4, 8)],
    "f64": [None, (512, 16, 4, 2), (512, 16, 8, 1), (256, 16, 4, 2)],
}


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_grf_1024_fused_every_tiling(lib, orc, grf1024, dt):
    field, host = grf1024[dt]
    bm = M.bin_map(">=", 0.0)
    say(f"1024^3 {dt}: oracle")
    ref, _ = orc.bpfd_field(host, bm.t[0], bm.m)
    assert int(ref.sum()) == (S - 1) ** 3
    for tl in TILINGS[dt]:
        say(f"1024^3 {dt}: fused, tiling {tl or 'built-in'}")
        np.testing.assert_array_equal(fused(lib, field, bm, tl), ref, err_msg=f"tiling {tl}")
    # the autotuner's choice, then a call without a tiling (which uses it)
    fi = M.FieldInfo(field.data_ptr(), (S, S, S), M.ENC_F32 if dt == "f32" else M.ENC_F64)
    chosen = M.MkfTiling()
    M.check(lib.mkfFusedAutotune(fi.ref, C.byref(bm), C.byref(chosen), None), "mkfFusedAutotune")
    say(f"1024^3 {dt}: autotuned {chosen.planePairs},{chosen.waves},{chosen.rowPairs},{chosen.voxelsPerLane}")
    np.testing.assert_array_equal(fused(lib, field, bm), ref, err_msg="autotuned tiling")
    # measures of the histogram: bit-equal to the oracle's restatement
    m = (C.c_float * 4)()
    sym = C.create_string_buffer(16)
    h = np.ascontiguousarray(ref)
    lib.mkfSelectMeasuresBPFD(m, sym, h.ctypes.data_as(C.POINTER(C.c_size_t)), C.c_float(3.0 / (3 * S)), 2)
    om, _ = orc.measures(ref, 3.0 / (3 * S))
    assert np.array_equal(np.array(list(m), dtype=np.float32), np.asarray(om, dtype=np.float32))


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_grf_1024_dropin_pair(lib, orc, grf1024, dt, cuda):
    """binMapCUDA into a caller buffer, then mkfCUDAGetBPFD (profile FAST):
    every map word (tail bits are zero: X is a multiple of 32) and the
    histogram equal the oracle's."""
    field, host = grf1024[dt]
    bm = M.bin_map(">=", 0.0)
    say(f"1024^3 {dt}: drop-in pair")
    org = M.BMOrg()
    nw = lib.setBMO(C.byref(org), M.dims((S, S, S)), b"\x00")
    words = torch.full((nw,), -1, dtype=torch.int32, device=cuda)
    fi = M.FieldInfo(field.data_ptr(), (S, S, S), M.ENC_F32 if dt == "f32" else M.ENC_F64)
    ki = M.KernInfo()
    assert lib.binMapCUDA(C.byref(ki), words.data_ptr(), C.byref(org), fi.ref, C.byref(bm)) == words.data_ptr()
    h = torch.zeros(256, dtype=torch.int64, device=cuda)
    assert lib.mkfCUDAGetBPFD(C.byref(ki), h.data_ptr(), C.byref(org), words.data_ptr(), 3) == h.data_ptr()
    torch.cuda.synchronize()
    assert ki.dtms[0] > 0 and ki.dtms[1] > 0
    ref_words = orc.pbm_c(host, bm.t[0], bm.m)
    got_words = words.cpu().numpy().view(np.uint32).reshape(ref_words.shape)
    assert np.array_equal(got_words, ref_words), "map words differ"
    ref = orc.bpfd_chunk(ref_words, S)
    np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64), ref)
    del words


def test_batch_mix_512_histograms(lib, orc, cuda):
    """The uneven batch at its real size: 512 histograms, graph replay and
    direct launch, against the oracle."""
    specs = synth.batch_mix(seed=4)
    say(f"batch mix: generating {len(specs)} fields")
    fields = synth.make_fields(specs, cuda)
    torch.cuda.synchronize()
    levels = [float(v) for v in np.linspace(-1.5, 1.5, 8)]
    lv = [[(">=", t) for t in levels] for _ in fields]
    B = Batch(fields, lv, zero=True)
    try:
        info = B.info()
        say(f"batch mix: {info['items']} items, grid {info['grid']}")
        B.run()
        torch.cuda.synchronize()
        graph = B.out.cpu().numpy().view(np.uint64).copy()
        B.out.fill_(5)
        B.run(graph=False)
        torch.cuda.synchronize()
        direct = B.out.cpu().numpy().view(np.uint64).copy()
        say("batch mix: oracle")
        maps = [M.bin_map(">=", t) for t in levels]
        for i, f in enumerate(fields):
            h = f.cpu().numpy()
            for j, bm in enumerate(maps):
                ref, _ = orc.bpfd_field(h, bm.t[0], bm.m)
                k = i * len(levels) + j
                np.testing.assert_array_equal(graph[k], ref, err_msg=f"graph: field {i} {specs[i][:2]} level {j}")
                np.testing.assert_array_equal(direct[k], ref, err_msg=f"direct: field {i} level {j}")
    finally:
        B.close()


def _pinned(t):
    p = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    p.copy_(t)
    return p


def test_serving_stream_at_real_sizes(lib, orc, cuda):
    """40 requests of the serving stream (seed 14: sides 66..726, three above
    600, 12 device-resident snapshots), submitted at once so they batch;
    each result against the oracle, and the largest host request alone
    against the same request in the mix."""
    n = 40
    pool_spec, sim_spec, reqs = request_stream(n, seed=14)
    say(f"serving stream: pool {[s for s, _ in pool_spec]}, simulations {[s for s, _ in sim_spec]}")
    pool = [None] * len(pool_spec)
    used = {r["pool"] for r in reqs if r["sim"] is None}
    for i in sorted(used):
        side, dt = pool_spec[i]
        pool[i] = _pinned(synth.grf_field_direct((side, side, side), 3.0, seed=7000 + i, device=cuda, dtype=dt))
    g = torch.Generator(device=cuda)
    g.manual_seed(14)
    sims = []
    for i, (side, dt) in enumerate(sim_spec):
        ring = [synth.grf_field_direct((side, side, side), 3.0, seed=8000 + i, device=cuda, dtype=dt)]
        for _ in range(2):
            ring.append(ring[-1] + 0.01 * torch.randn(ring[-1].shape, generator=g, device=cuda, dtype=dt))
        sims.append(ring)
    fields, turn = [], [0] * len(sims)
    for r in reqs:
        if r["sim"] is None:
            fields.append(pool[r["pool"]])
        else:
            fields.append(sims[r["sim"]][turn[r["sim"]] % 3])
            turn[r["sim"]] += 1
    torch.cuda.synchronize()
    sides = [f.shape[0] for f in fields]
    assert sum(s > 600 for s in sides) >= 3 and sum(r["sim"] is not None for r in reqs) == 12
    probe = max((i for i, r in enumerate(reqs) if r["sim"] is None), key=lambda i: sides[i])
    srv = Server(budget=48 << 30, max_in_flight=64, max_batch=32)
    try:
        say(f"serving stream: request {probe} (side {sides[probe]}) alone")
        srv.submit(fields[probe], reqs[probe]["levels"], tag=-1)
        alone = None
        while alone is None:
            for res in srv.poll(wait_us=20000):
                alone = res
        assert alone["status"] == 0
        say(f"serving stream: {n} requests at once")
        for i in range(n):
            srv.submit(fields[i], reqs[i]["levels"], tag=i)
        got = {}
        while len(got) < n:
            for res in srv.poll(wait_us=20000):
                got[res["tag"]] = res
        assert all(got[i]["status"] == 0 for i in range(n))
        assert max(got[i]["batch"] for i in range(n)) > 1, "no two requests shared a launch"
        np.testing.assert_array_equal(got[probe]["hist"], alone["hist"])
        say("serving stream: oracle")
        for i in range(n):
            h = fields[i].cpu().numpy()
            ref = np.stack([orc.bpfd_field(h, bm.t[0], bm.m)[0]
                            for bm in (M.bin_map(*lv) for lv in reqs[i]["levels"])])
            np.testing.assert_array_equal(got[i]["hist"], ref, err_msg=f"request {i} side {sides[i]}")
    finally:
        srv.close()
