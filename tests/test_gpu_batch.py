"""Batched form: many fields x many threshold levels in one persistent,
graph-captured launch. Every histogram equals the oracle's for that field and
level, bit for bit, and equals the single-field fused kernel's."""

import ctypes as C

import numpy as np
import pytest
import torch

import mkf_amd as M
from mkf_amd import synth

pytestmark = pytest.mark.gpu

RELOPS = [">=", "<", "<=", ">", "=", "!=", "", "!"]


def make_batch(cuda, specs, seed=0):
    """specs: list of ((Z, Y, X), dtype, levels[(relop, t)]) -> fields, maps, outputs."""
    fields, maps, outs, arr = [], [], [], (M.MkfBatchField * len(specs))()
    for i, (shape, dtype, levels) in enumerate(specs):
        f = synth.grf_field(shape, sigma=1.5, seed=seed + i, device=cuda, dtype=dtype)
        fields.append(f)
        bms = (M.BinMapF64 * len(levels))()
        for j, (rel, t) in enumerate(levels):
            bms[j] = M.bin_map(rel, t)
        maps.append(bms)
        out = torch.zeros(len(levels) * 256, dtype=torch.int64, device=cuda)
        outs.append(out)
        Z, Y, X = shape
        arr[i].pF = f.data_ptr()
        arr[i].enc = M.ENC_F32 if dtype == torch.float32 else M.ENC_F64
        arr[i].def_[0], arr[i].def_[1], arr[i].def_[2] = X, Y, Z
        arr[i].nLevels = len(levels)
        arr[i].pM = C.cast(bms, C.POINTER(M.BinMapF64))
        arr[i].pBPFD = out.data_ptr()
    return fields, maps, outs, arr


def oracle_hists(orc, fields, maps):
    res = []
    for f, bms in zip(fields, maps):
        f_np = f.cpu().numpy()
        hs = []
        for bm in bms:
            words = orc.pbm_c(f_np, bm.t[0], bm.m)
            X = f_np.shape[2]
            hs.append(orc.bpfd_chunk(words, X) if X >= 32 else orc.bpfd_direct_c(words, X))
        res.append(np.stack(hs))
    return res


def got(out, n):
    return out.cpu().numpy().view(np.uint64).reshape(n, 256)


SPECS = [
    ((9, 11, 2), torch.float32, [(">=", 0.0)]),
    ((7, 5, 33), torch.float64, [(">=", -0.5), ("<", 0.3), ("!=", 0.0)]),
    ((12, 10, 65), torch.float32, [(r, 0.1 * k) for k, r in enumerate(RELOPS)]),
    ((20, 17, 129), torch.float64, [(">=", t) for t in np.linspace(-1.5, 1.5, 8)]),
    ((6, 40, 300), torch.float32, [("<=", t) for t in np.linspace(-1.0, 1.0, 16)]),
    ((30, 9, 512), torch.float64, [(">", 0.2), ("=", 0.0)]),
    ((2, 2, 64), torch.float32, [(">=", 0.0)] * 2),
    ((41, 3, 100), torch.float32, [(">=", 0.25)]),
    ((5, 7, 1030), torch.float32, [(">=", -0.3), ("<", 0.3), ("!=", 0.1)]),
    ((11, 13, 700), torch.float64, [("<=", 0.0)]),
    ((1, 6, 50), torch.float32, [(">=", 0.0)]),
]


def bshape(pp, segs):
    s = M.MkfBatchShape()
    s.planePairs, s.segments = pp, segs
    return s


@pytest.mark.parametrize("depth,segs", [(1, 8), (3, 2), (32, 1), (7, 4), (0, 0)])
def test_batch_matches_oracle_and_graph_replays(lib, orc, cuda, depth, segs):
    fields, maps, outs, arr = make_batch(cuda, SPECS)
    ref = oracle_hists(orc, fields, maps)
    b = lib.mkfBatchCreate(arr, len(SPECS), C.byref(bshape(depth, segs)), M.MKF_BATCH_ZERO_OUTPUT, None)
    assert b, M.last_error()
    try:
        for rep in range(3):
            # eager work between replays, on the same stream
            for o in outs:
                o.add_(7)
            M.check(lib.mkfBatchRun(b, None), "mkfBatchRun")
            torch.cuda.synchronize()
            for i, (o, bms) in enumerate(zip(outs, maps)):
                np.testing.assert_array_equal(got(o, len(bms)), ref[i], err_msg=f"replay {rep} field {i}")
        M.check(lib.mkfBatchRunDirect(b, None), "direct")
        torch.cuda.synchronize()
        for i, (o, bms) in enumerate(zip(outs, maps)):
            np.testing.assert_array_equal(got(o, len(bms)), ref[i])
    finally:
        lib.mkfBatchDestroy(b)


def test_batch_accumulates_without_zero_flag(lib, orc, cuda):
    fields, maps, outs, arr = make_batch(cuda, SPECS[1:5], seed=10)
    ref = oracle_hists(orc, fields, maps)
    b = lib.mkfBatchCreate(arr, 4, C.byref(bshape(8, 8)), 0, None)
    assert b
    try:
        for _ in range(2):
            M.check(lib.mkfBatchRun(b, None))
        torch.cuda.synchronize()
        for i, (o, bms) in enumerate(zip(outs, maps)):
            np.testing.assert_array_equal(got(o, len(bms)), 2 * ref[i])
    finally:
        lib.mkfBatchDestroy(b)


def test_batch_equals_fused_per_field(lib, cuda):
    specs = [((24, 30, 200), torch.float32, [(">=", t) for t in (-0.5, 0.0, 0.5)]),
             ((16, 20, 384), torch.float64, [("<", t) for t in (-1.0, 1.0)])]
    fields, maps, outs, arr = make_batch(cuda, specs, seed=20)
    b = lib.mkfBatchCreate(arr, 2, None, M.MKF_BATCH_ZERO_OUTPUT, None)
    assert b
    try:
        grid = [(4, 8), (16, 2), (64, 1), (16, 0)]
        cands = (M.MkfBatchShape * len(grid))(*[bshape(*g) for g in grid])
        best = M.MkfBatchShape()
        M.check(lib.mkfBatchAutotune(b, cands, len(grid), 2, C.byref(best), None), "autotune")
        assert (best.planePairs, best.segments) in [(4, 8), (16, 2), (64, 1), (16, 8)]
        n = C.c_int(0)
        info = M.MkfBatchShape()
        g = C.c_int(0)
        M.check(lib.mkfBatchInfo(b, C.byref(n), C.byref(info), C.byref(g)))
        assert (info.planePairs, info.segments) == (best.planePairs, best.segments)
        assert n.value > 0 and g.value > 0
        M.check(lib.mkfBatchRun(b, None))
        torch.cuda.synchronize()
        for f, bms, o in zip(fields, maps, outs):
            Z, Y, X = f.shape
            for j in range(len(bms)):
                h = torch.zeros(256, dtype=torch.int64, device=cuda)
                fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), M.ENC_F32 if f.dtype == torch.float32 else M.ENC_F64)
                M.check(lib.mkfFusedBPFD(None, h.data_ptr(), None, fi.ref, C.byref(bms[j]), None, None))
                torch.cuda.synchronize()
                np.testing.assert_array_equal(got(o, len(bms))[j], h.cpu().numpy().view(np.uint64))
    finally:
        lib.mkfBatchDestroy(b)


def test_batch_rejects_bad_arguments(lib, cuda):
    _, _, _, arr = make_batch(cuda, [((4, 4, 40), torch.float32, [(">=", 0.0)])])
    for bad in [(8, 3), (5000, 8), (-1, 8)]:
        assert not lib.mkfBatchCreate(arr, 1, C.byref(bshape(*bad)), 0, None)
        assert "shape" in M.last_error()
    arr[0].nLevels = 17
    assert not lib.mkfBatchCreate(arr, 1, None, 0, None)
    assert "field 0" in M.last_error()
    arr[0].nLevels = 1
    assert not lib.mkfBatchCreate(arr, 0, None, 0, None)
    arr[0].enc = 9
    assert not lib.mkfBatchCreate(arr, 1, None, 0, None)


def _special_field(shape, dtype, seed, levels):
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(shape).astype(np.float64)
    flat = f.reshape(-1)
    n = flat.size
    for vals, frac in (([np.nan], 0.04), ([np.inf, -np.inf], 0.02), ([0.0, -0.0], 0.03)):
        idx = rng.choice(n, int(frac * n), replace=False)
        flat[idx] = rng.choice(vals, idx.size)
    idx = rng.choice(n, int(0.05 * n), replace=False)
    flat[idx] = rng.choice([t for _, t in levels], idx.size)  # exactly on a threshold
    if dtype == torch.float32:
        # thresholds are passed as float, so values equal to the float threshold
        f = f.astype(np.float32)
    return torch.from_numpy(f)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batch_special_values_every_relation(lib, orc, cuda, dtype):
    levels = [(r, t) for r, t in zip(RELOPS + ["<", ">="], [0.0, 0.5, -0.5, 0.0, 0.25, 0.0, 1.0, -1.0, 0.1, 0.1])]
    specs = [((9, 7, 130), dtype, levels), ((5, 6, 64), dtype, levels[:5]), ((4, 3, 70), dtype, levels[3:])]
    fields, maps, outs, arr = make_batch(cuda, specs, seed=30)
    for i, (shape, _, lv) in enumerate(specs):
        fields[i].copy_(_special_field(shape, dtype, 40 + i, [(r, float(np.float32(t))) for r, t in lv]).to(cuda))
    ref = oracle_hists(orc, fields, maps)
    for segs in (1, 8):
        for o in outs:
            o.fill_(3)
        b = lib.mkfBatchCreate(arr, len(specs), C.byref(bshape(5, segs)), M.MKF_BATCH_ZERO_OUTPUT, None)
        assert b, M.last_error()
        try:
            M.check(lib.mkfBatchRun(b, None))
            torch.cuda.synchronize()
            for i, (o, bms) in enumerate(zip(outs, maps)):
                np.testing.assert_array_equal(got(o, len(bms)), ref[i], err_msg=f"segments {segs} field {i}")
        finally:
            lib.mkfBatchDestroy(b)


def test_batch_autotune_failure_restores_or_refuses(lib, orc, cuda):
    """A re-plan that fails part way (forced from the host through the
    library's fault points) leaves the previous plan in place, so runs still
    give the oracle's histograms; if restoring the plan fails too, the batch
    refuses to run instead of reading a stale work queue."""
    fields, maps, outs, arr = make_batch(cuda, SPECS[2:6], seed=50)
    ref = oracle_hists(orc, fields, maps)
    b = lib.mkfBatchCreate(arr, 4, C.byref(bshape(8, 8)), M.MKF_BATCH_ZERO_OUTPUT, None)
    assert b, M.last_error()
    cands = (M.MkfBatchShape * 3)(bshape(4, 2), bshape(16, 1), bshape(2, 4))
    try:
        M.check(lib.mkfDebugInjectFault(b"batch.upload", 2))  # the second candidate's upload
        rc = lib.mkfBatchAutotune(b, cands, 3, 1, None, None)
        assert rc == M.MKF_ERR_MEM and "injected" in M.last_error()
        n, shp, g = C.c_int(0), M.MkfBatchShape(), C.c_int(0)
        M.check(lib.mkfBatchInfo(b, C.byref(n), C.byref(shp), C.byref(g)))
        assert (shp.planePairs, shp.segments) == (8, 8)
        for graph in (True, False):
            for o in outs:
                o.fill_(9)
            M.check((lib.mkfBatchRun if graph else lib.mkfBatchRunDirect)(b, None), "run after failed tune")
            torch.cuda.synchronize()
            for i, (o, bms) in enumerate(zip(outs, maps)):
                np.testing.assert_array_equal(got(o, len(bms)), ref[i], err_msg=f"graph={graph} field {i}")
        # first candidate's upload fails, then the restoring capture fails
        M.check(lib.mkfDebugInjectFault(b"batch.upload", 1))
        M.check(lib.mkfDebugInjectFault(b"batch.capture", 1))
        assert lib.mkfBatchAutotune(b, cands, 3, 1, None, None) != M.MKF_OK
        assert lib.mkfBatchRun(b, None) == M.MKF_ERR_ARG and "unusable" in M.last_error()
        assert lib.mkfBatchRunDirect(b, None) == M.MKF_ERR_ARG
        assert lib.mkfBatchAutotune(b, cands, 3, 1, None, None) == M.MKF_ERR_ARG
    finally:
        lib.mkfDebugInjectFault(b"batch.upload", 0)
        lib.mkfDebugInjectFault(b"batch.capture", 0)
        lib.mkfBatchDestroy(b)


INF = float("inf")
GRAY_LEVELS = {
    # ">=" / "<" family, unsorted, with duplicates, "all" ("!") and "none" ("")
    "ge8": [(">=", -1.0), ("<", 0.5), (">=", 0.0), ("<", 0.0), (">=", 1.5), ("<", -1.5), ("!", 0.0), ("", 0.0)],
    # ">" / "<=" family, 5 levels (8-level planes, 3 padded)
    "gt5": [(">", 0.2), ("<=", -0.4), (">", -0.4), ("<=", 0.9), (">", 0.2)],
    # 3 levels: the 4-level planes
    "ge3": [(">=", 0.1), ("<", -0.1), (">=", 0.1)],
    # 4 levels at infinite and zero thresholds
    "gt4": [(">", INF), ("<=", -INF), (">", -0.0), ("<=", 0.0)],
    # 9 levels: parts of 5 and 4
    "ge9": [("<" if k % 3 == 0 else ">=", t) for k, t in enumerate(np.linspace(1.2, -1.2, 9))],
    # 16 levels: two parts of 8, thresholds unsorted, both relations
    "gt16": [("<=" if k % 2 else ">", float(t)) for k, t in
             enumerate(np.random.default_rng(3).uniform(-1.8, 1.8, 16))],
    # only "all" and "none": the default family at infinite thresholds
    "allnone": [("!", 0.3), ("", 0.3), ("!", -2.0)],
    # not the Gray-plane form: a NaN threshold, and levels of both families
    "nan": [(">=", float("nan")), ("<", 0.5), (">=", -0.5)],
    "both": [(">=", 0.1), (">", 0.1), ("<=", -0.3), ("<", 0.6)],
}


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_batch_gray_levels_match_oracle(lib, orc, cuda, dtype):
    """Fields whose levels all belong to one family (">=" and "<", or ">" and
    "<=") take the Gray-plane form: sorted levels, complemented histograms,
    parts of at most 8 levels, 63-cell segments. Every histogram equals the
    oracle's, on row lengths around the 63-cell segment edges and with NaN,
    +-inf, +-0 and exact-threshold voxels."""
    names = list(GRAY_LEVELS)
    xs = [2, 3, 62, 63, 64, 65, 126, 127, 128, 190, 253]
    specs = []
    for k, x in enumerate(xs):
        lv = GRAY_LEVELS[names[k % len(names)]]
        specs.append(((3 + k % 4, 2 + (5 * k) % 11, x), dtype, lv))
    for k, name in enumerate(names):
        specs.append(((9, 13, 70 + 17 * k), dtype, GRAY_LEVELS[name]))
    fields, maps, outs, arr = make_batch(cuda, specs, seed=60)
    for i, (shape, _, lv) in enumerate(specs):
        thr = [(r, float(np.float32(t))) for r, t in lv if np.isfinite(t)] or [(">=", 0.0)]
        fields[i].copy_(_special_field(shape, dtype, 70 + i, thr).to(cuda))
    ref = oracle_hists(orc, fields, maps)
    for depth, segs in ((1, 8), (4, 2), (64, 1)):
        for o in outs:
            o.fill_(5)
        b = lib.mkfBatchCreate(arr, len(specs), C.byref(bshape(depth, segs)), M.MKF_BATCH_ZERO_OUTPUT, None)
        assert b, M.last_error()
        try:
            print(f"gray levels {dtype}: depth {depth} segments {segs}", flush=True)
            M.check(lib.mkfBatchRun(b, None))
            torch.cuda.synchronize()
            for i, (o, bms) in enumerate(zip(outs, maps)):
                np.testing.assert_array_equal(got(o, len(bms)), ref[i],
                                              err_msg=f"depth {depth} segments {segs} field {i} {specs[i][0]}")
        finally:
            lib.mkfBatchDestroy(b)
