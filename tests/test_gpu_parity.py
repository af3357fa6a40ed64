"""GPU parity: every result of libmkf_amd.so's HIP kernels, called through
the C ABI, equals the CPU oracle bit for bit (integer/bit work: exact)."""

import ctypes as C
import math

import numpy as np
import pytest
import torch

import mkf_amd as M
from mkf_amd import synth
from tests.gpu_util import dev_bpfd_from_pbm, dev_fused, dev_pbm, enc_of, host_np

pytestmark = pytest.mark.gpu

RELOPS = ["<", "=", "<=", ">", "!=", ">=", "", "!"]


def oracle_all(orc, f_np, relop, thr):
    """(pbm words, hist) of the oracle for a host field and relation."""
    m = M.bin_map(relop, thr)
    words = orc.pbm_c(f_np, m.t[0], m.m)
    X = f_np.shape[2]
    if min(f_np.shape) < 2:
        return words, np.zeros(256, np.uint64)
    hist = orc.bpfd_chunk(words, X) if X >= 32 else orc.bpfd_direct_c(words, X)
    return words, hist


def check_all_paths(lib, orc, field, relop, thr):
    f_np = host_np(field)
    w_ref, h_ref = oracle_all(orc, f_np, relop, thr)
    words, dw, org, _ = dev_pbm(lib, field, relop, thr)
    np.testing.assert_array_equal(words, w_ref)
    for prof in (M.MKFCU_PROFILE_FAST, M.MKFCU_PROFILE_FLEX, 0):
        h, _ = dev_bpfd_from_pbm(lib, dw, org, prof)
        np.testing.assert_array_equal(h, h_ref)
    h, _ = dev_fused(lib, field, relop, thr)
    np.testing.assert_array_equal(h, h_ref)
    Z, Y, X = field.shape
    assert int(h_ref.sum()) == max(X - 1, 0) * max(Y - 1, 0) * max(Z - 1, 0)
    return h_ref


SMALL = [(2, 2, 2), (3, 3, 3), (5, 4, 31), (4, 5, 32), (3, 7, 33), (6, 3, 63), (3, 3, 64), (4, 4, 65),
         (5, 6, 127), (3, 4, 128), (4, 3, 129), (9, 17, 200), (2, 70, 1000), (5, 9, 1025), (17, 2, 70),
         (70, 33, 2)]


@pytest.mark.parametrize("shape", SMALL)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_small_shapes(lib, orc, cuda, shape, dtype):
    print(f"small shape {shape} {dtype}", flush=True)
    g = torch.Generator(device=cuda).manual_seed(hash(shape) & 0xFFFF)
    f = torch.rand(shape, generator=g, device=cuda, dtype=torch.float64).to(dtype)
    check_all_paths(lib, orc, f, ">=", 0.5)


@pytest.mark.parametrize("relop", RELOPS)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_relations_and_special_values(lib, orc, cuda, relop, dtype):
    print(f"relation {relop!r} {dtype}", flush=True)
    rng = np.random.default_rng(5)
    choices = np.array([0.0, -0.0, 1e-40, -1e-40, 5e-324, np.nan, np.inf, -np.inf, 0.25, -0.25, 1.0])
    vals = rng.choice(choices, size=(7, 9, 97)).astype(np.float64)
    f = torch.from_numpy(vals).to(dtype).to(cuda)
    for thr in (0.0, 0.25, -0.25):
        check_all_paths(lib, orc, f, relop, thr)


def test_f32_threshold_rounding(lib, orc, cuda):
    # setBinMapF64 stores float(0.1); a float field of float(0.1) is EQ.
    f = torch.full((4, 4, 40), float(np.float32(0.1)), dtype=torch.float32, device=cuda)
    words, _, _, _ = dev_pbm(lib, f, "=", 0.1)
    assert np.all(words[..., 0] == 0xFFFFFFFF) and np.all(words[..., 1] == 0xFF)
    # A double field holding the double 0.1 is not EQ to float(0.1).
    f64 = torch.full((4, 4, 40), 0.1, dtype=torch.float64, device=cuda)
    words, _, _, _ = dev_pbm(lib, f64, "=", 0.1)
    assert np.all(words == 0)


# (planePairs, waves, rowPairs): narrow workgroups split each row over
# several tiles, so the bit right of a tile's last segment is read from
# memory instead of the neighbour's LDS slot.
TILINGS = [(1, 16, 4), (3, 16, 2), (8, 2, 4), (64, 3, 8), (5, 1, 2), (128, 16, 8), (7, 4, 4), (2, 5, 8)]


@pytest.mark.parametrize("tiling", TILINGS)
@pytest.mark.parametrize("vpl", [1, 0])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("X", [300, 1030])
def test_fused_and_pbm_histogram_tilings(lib, orc, cuda, tiling, vpl, dtype, X):
    """voxelsPerLane 0 selects 16 bytes per lane (4 floats / 2 doubles)."""
    tiling = tiling + (vpl,)
    print(f"tiling {tiling} {dtype} X={X}", flush=True)
    f = synth.grf_field((37, 41, X), sigma=2.0, seed=3, device=cuda, dtype=dtype)
    w_ref, h_ref = oracle_all(orc, host_np(f), ">=", 0.0)
    h, _ = dev_fused(lib, f, ">=", 0.0, tiling=tiling)
    np.testing.assert_array_equal(h, h_ref)
    _, dw, org, _ = dev_pbm(lib, f, ">=", 0.0)
    hb = torch.zeros(256, dtype=torch.int64, device=cuda)
    t = M.MkfTiling(*tiling)
    M.check(lib.mkfBPFDAsync(hb.data_ptr(), C.byref(org), dw.data_ptr(), C.byref(t), None))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hb.cpu().numpy().view(np.uint64), h_ref)


# Maps whose rows are 1 to 8 words: the histogram's lanes walk 8, 4, 2 or 1
# row stacks side by side (257: 9 words, the per-segment kernel). Y - 1 from below one stack (the per-segment kernel) to
# many stacks, with a tile's last groups past the field (read from
# Y - 1 - rowPairs); 2, 4 and 8 row pairs; slab depths that do and do not
# divide the planes; 1 to 16 waves per workgroup, each on a tile of its own
# (some past the last tile); the default tiling.
NARROW_CASES = [(3, 5, (4, 1, 4, 4)), (5, 9, (4, 1, 4, 4)), (3, 7, (3, 1, 2, 4)), (9, 6, (2, 1, 8, 4)),
                (37, 11, (5, 3, 4, 4)), (70, 13, (32, 16, 2, 4)), (75, 10, (3, 8, 8, 4)), (41, 17, None),
                (130, 40, (7, 2, 4, 4))]


@pytest.mark.parametrize("X", [2, 3, 17, 31, 32, 33, 50, 64, 65, 96, 100, 127, 128, 129, 200, 256, 257])
def test_pbm_histogram_of_narrow_rows(lib, orc, cuda, X):
    for k, (Y, Z, tiling) in enumerate(NARROW_CASES):
        print(f"narrow PBM histogram X={X} Y={Y} Z={Z} tiling {tiling}", flush=True)
        f = synth.grf_field((Z, Y, X), sigma=1.5, seed=X * 16 + k, device=cuda, dtype=torch.float32)
        for relop, thr in ((">=", 0.0), ("<", 0.3)):
            _, h_ref = oracle_all(orc, host_np(f), relop, thr)
            _, dw, org, _ = dev_pbm(lib, f, relop, thr)
            hb = torch.zeros(256, dtype=torch.int64, device=cuda)
            t = M.MkfTiling(*tiling) if tiling else None
            M.check(lib.mkfBPFDAsync(hb.data_ptr(), C.byref(org), dw.data_ptr(), C.byref(t) if t else None, None))
            torch.cuda.synchronize()
            np.testing.assert_array_equal(hb.cpu().numpy().view(np.uint64), h_ref, err_msg=f"Y={Y} Z={Z} {relop}")


# Fused launches of more than 4096 workgroups store one histogram per
# workgroup and add them up in a second kernel: 1, 2 and 4 voxels per lane,
# whole and ragged rows, x-tiles with the halo bit from memory, the general
# (two-field) form, and accumulation into a nonzero histogram.
ROWS_CASES = [((40, 500, 64), torch.float32, (1, 16, 2, 1)), ((40, 500, 300), torch.float32, (1, 1, 2, 4)),
              ((40, 500, 200), torch.float64, (1, 16, 2, 2)), ((33, 301, 97), torch.float64, (1, 16, 2, 1))]


@pytest.mark.parametrize("case", range(len(ROWS_CASES)))
def test_fused_many_workgroups_reduce_histogram_rows(lib, orc, cuda, case):
    shape, dt, tiling = ROWS_CASES[case]
    Z, Y, X = shape
    print(f"fused histogram rows {shape} {dt} tiling {tiling}", flush=True)
    f = synth.grf_field(shape, sigma=2.0, seed=500 + case, device=cuda, dtype=dt)
    for relop, thr in ((">=", 0.0), ("<", 0.4)):
        _, h_ref = oracle_all(orc, host_np(f), relop, thr)
        h, _ = dev_fused(lib, f, relop, thr, tiling=tiling)
        np.testing.assert_array_equal(h, h_ref, err_msg=f"{relop} {thr}")
        h2, _ = dev_fused(lib, f, relop, thr, tiling=tiling, pre=h_ref)
        np.testing.assert_array_equal(h2, 2 * h_ref)
    g = synth.grf_field(shape, sigma=2.0, seed=600 + case, device=cuda, dtype=dt)
    fi = M.FieldInfo([f.data_ptr(), g.data_ptr()], (X, Y, Z), enc_of(f))
    bm = M.bin_map(">", 0.1)
    h = torch.zeros(256, dtype=torch.int64, device=cuda)
    t = M.MkfTiling(tiling[0], tiling[1], tiling[2], 1)  # the general form takes 1 voxel per lane
    M.check(lib.mkfFusedBPFD(None, h.data_ptr(), None, fi.ref, C.byref(bm), C.byref(t), None), "two fields")
    torch.cuda.synchronize()
    _, h_ref = oracle_all(orc, host_np(f) + host_np(g), ">", 0.1)
    np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64), h_ref)


# 2 voxels per lane on double rows whose last 128-voxel segment is partial:
# the kernel form with a second walk for the partial waves. Workgroups as wide
# as the row or split into x-tiles (the bit right of a tile from memory), 2, 4
# and 8 row pairs, one to many plane pairs per slab.
F64_PW_TILINGS = [(1, 16, 4, 2), (3, 16, 2, 2), (8, 2, 4, 2), (64, 3, 8, 2), (5, 1, 2, 2), (7, 4, 4, 2)]


@pytest.mark.parametrize("tiling", F64_PW_TILINGS)
@pytest.mark.parametrize("X", [129, 200, 386, 1030, 1089])  # 129: whole segments (control)
def test_f64_two_voxels_on_ragged_rows(lib, orc, cuda, tiling, X):
    print(f"f64 2 voxels per lane {tiling} X={X}", flush=True)
    f = synth.grf_field((19, 23, X), sigma=2.0, seed=X, device=cuda, dtype=torch.float64)
    for relop, thr in ((">=", 0.0), ("<", -0.4)):
        _, h_ref = oracle_all(orc, host_np(f), relop, thr)
        h, _ = dev_fused(lib, f, relop, thr, tiling=tiling)
        np.testing.assert_array_equal(h, h_ref, err_msg=f"{relop} {thr}")


# 16 voxels per lane (float fields): rows of one to three 1024-voxel
# segments, exact and ragged (the last segment gathered, the voxel right of a
# segment read when it exists), 2 and 4 row pairs, slab depths that do and do
# not divide the planes, slabs deeper than the field.
V16_CASES = [(5, (2, 1, 4)), (300, (5, 1, 4)), (512, (2, 1, 4)), (513, (6, 1, 2)), (1024, (7, 1, 4)),
             (1025, (3, 1, 2)), (1030, (1, 1, 4)), (1537, (5, 1, 4)), (2048, (4, 1, 4)), (2049, (128, 1, 2)),
             (2100, (9, 16, 4)), (3071, (2, 1, 2))]


@pytest.mark.parametrize("V", [16, 8])
@pytest.mark.parametrize("X,tiling", V16_CASES)
def test_fused_16_voxels_per_lane(lib, orc, cuda, X, tiling, V):
    """The persistent fused kernel with 16 or 8 voxels per lane; the map's
    histogram with the same tiling (16 voxels per lane, or 4 for 8)."""
    print(f"{V} voxels per lane X={X} tiling {tiling}", flush=True)
    f = synth.grf_field((23, 13, X), sigma=2.0, seed=11, device=cuda, dtype=torch.float32)
    f_np = host_np(f)
    t = M.MkfTiling(*(tiling + (V,)))
    for relop, thr in ((">=", 0.0), ("<", 0.3), ("!", 0.1), ("=", 0.0)):
        _, h_ref = oracle_all(orc, f_np, relop, thr)
        h, _ = dev_fused(lib, f, relop, thr, tiling=tiling + (V,))
        np.testing.assert_array_equal(h, h_ref)
        # the histogram of the map
        _, dw, org, _ = dev_pbm(lib, f, relop, thr)
        hb = torch.zeros(256, dtype=torch.int64, device=cuda)
        M.check(lib.mkfBPFDAsync(hb.data_ptr(), C.byref(org), dw.data_ptr(), C.byref(t), None))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(hb.cpu().numpy().view(np.uint64), h_ref)


@pytest.mark.parametrize("relop", RELOPS)
def test_fused_16_special_values(lib, orc, cuda, relop):
    """NaN, signed zeros, subnormals and infinities at every lane position
    of two segments and the gathered tail."""
    print(f"16 and 8 voxels per lane, special values, relation {relop!r}", flush=True)
    rng = np.random.default_rng(17)
    choices = np.array([0.0, -0.0, 1e-40, -1e-40, np.nan, np.inf, -np.inf, 0.25, -0.25, 1.0], np.float32)
    f = torch.from_numpy(rng.choice(choices, size=(5, 6, 2100))).to(cuda)
    for thr in (0.0, 0.25):
        _, h_ref = oracle_all(orc, host_np(f), relop, thr)
        _, dw, org, _ = dev_pbm(lib, f, relop, thr)
        for ty, v in ((2, 16), (4, 16), (2, 8), (4, 8)):
            h, _ = dev_fused(lib, f, relop, thr, tiling=(3, 1, ty, v))
            np.testing.assert_array_equal(h, h_ref)
            hb = torch.zeros(256, dtype=torch.int64, device=cuda)
            t = M.MkfTiling(3, 1, ty, v)
            M.check(lib.mkfBPFDAsync(hb.data_ptr(), C.byref(org), dw.data_ptr(), C.byref(t), None))
            torch.cuda.synchronize()
            np.testing.assert_array_equal(hb.cpu().numpy().view(np.uint64), h_ref)


def test_fused_16_padded_unaligned_rows(lib, orc, cuda):
    """A sub-volume whose rows start one float past a 16-byte boundary and
    are padded (row stride X + 3): the 16-byte loads are 4-byte aligned."""
    full = synth.grf_field((12, 11, 1031), sigma=2.0, seed=21, device=cuda, dtype=torch.float32)
    sub = full[1:11, 1:10, 1:1029]
    Z, Y, X = sub.shape
    _, h_ref = oracle_all(orc, np.ascontiguousarray(host_np(sub)), ">", 0.05)
    bm = M.bin_map(">", 0.05)
    fi = M.FieldInfo(sub.data_ptr(), (X, Y, Z), M.ENC_F32, strides=(1, 1031, 1031 * 11))
    for tiling in (None, (4, 1, 4, 16), (2, 1, 2, 16), (3, 1, 4, 8), (5, 1, 2, 8)):
        h = torch.zeros(256, dtype=torch.int64, device=cuda)
        t = M.MkfTiling(*tiling) if tiling else None
        M.check(lib.mkfFusedBPFD(None, h.data_ptr(), None, fi.ref, C.byref(bm), C.byref(t) if t else None, None))
        torch.cuda.synchronize()
        np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64), h_ref)


def test_bad_tiling_is_rejected(lib, cuda):
    f = torch.rand((8, 8, 64), device=cuda)
    bm = M.bin_map(">=", 0.5)
    h = torch.zeros(256, dtype=torch.int64, device=cuda)
    fi = M.FieldInfo(f.data_ptr(), (64, 8, 8), M.ENC_F32)
    for bad in [(0, 16, 4, 1), (8, 17, 4, 1), (8, 16, 3, 1), (8, 0, 4, 1), (8, 16, 4, 3), (8, 16, 4, 2),
                (8, 16, 8, 16), (8, 16, 8, 8), (8, 16, 4, 12), (65537, 16, 4, 1)]:
        t = M.MkfTiling(*bad)
        assert lib.mkfFusedBPFD(None, h.data_ptr(), None, fi.ref, C.byref(bm), C.byref(t), None) == -1


def test_fused_accumulates(lib, orc, cuda):
    f = synth.grf_field((30, 40, 70), sigma=1.5, seed=4, device=cuda, dtype=torch.float64)
    _, h_ref = oracle_all(orc, host_np(f), "<", 0.2)
    pre = np.arange(256, dtype=np.uint64) * 7
    h, _ = dev_fused(lib, f, "<", 0.2, pre=pre)
    np.testing.assert_array_equal(h, h_ref + pre)


def test_dims_of_one_are_compacted(lib, orc, cuda):
    f = torch.rand((40, 50, 1), device=cuda, dtype=torch.float32)  # X = 1
    Z, Y, X = f.shape
    # The reference drops X = 1 and treats the field as 50 x 40 x 1.
    bm = M.bin_map(">=", 0.5)
    org = M.BMOrg()
    words = torch.zeros(4096, dtype=torch.int32, device=cuda)
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), M.ENC_F32)
    assert lib.binMapCUDA(None, words.data_ptr(), C.byref(org), fi.ref, C.byref(bm))
    assert (org.rowElem, org.rowPairs, org.planePairs, org.rowWS) == (50, 39, 0, 2)
    w_ref = orc.pbm_c(host_np(f).reshape(1, 40, 50), bm.t[0], bm.m)
    got = words[: 2 * 40].cpu().numpy().view(np.uint32).reshape(1, 40, 2)
    np.testing.assert_array_equal(got, w_ref)


def test_strided_subvolume(lib, orc, cuda):
    full = synth.grf_field((40, 50, 130), sigma=2.0, seed=8, device=cuda, dtype=torch.float32)
    sub = full[3:33, 5:45, 7:107]  # (30, 40, 100) view
    Z, Y, X = sub.shape
    bm = M.bin_map(">=", 0.1)
    h = torch.zeros(256, dtype=torch.int64, device=cuda)
    fi = M.FieldInfo(sub.data_ptr(), (X, Y, Z), M.ENC_F32, strides=(1, 130, 130 * 50))
    M.check(lib.mkfFusedBPFD(None, h.data_ptr(), None, fi.ref, C.byref(bm), None, None))
    torch.cuda.synchronize()
    _, h_ref = oracle_all(orc, np.ascontiguousarray(host_np(sub)), ">=", 0.1)
    np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64), h_ref)


def _all_paths_on(lib, fi, X, Y, Z, relop, thr, cuda):
    """PBM (binMapCUDA), PBM histogram (mkfCUDAGetBPFD) and fused histogram of
    one BMFieldInfo -> (words, hist, fused hist)."""
    bm = M.bin_map(relop, thr)
    org = M.BMOrg()
    n = lib.setBMO(C.byref(org), M.dims((X, Y, Z)), b"\x00")
    words = torch.full((n,), -1, dtype=torch.int32, device=cuda)
    assert lib.binMapCUDA(None, words.data_ptr(), C.byref(org), fi.ref, C.byref(bm)), M.last_error()
    h = torch.zeros(256, dtype=torch.int64, device=cuda)
    assert lib.mkfCUDAGetBPFD(None, h.data_ptr(), C.byref(org), words.data_ptr(), 3)
    hf = torch.zeros(256, dtype=torch.int64, device=cuda)
    M.check(lib.mkfFusedBPFD(None, hf.data_ptr(), None, fi.ref, C.byref(bm), None, None), "fused")
    torch.cuda.synchronize()
    return (words.cpu().numpy().view(np.uint32).reshape(Z, Y, org.rowWS), h.cpu().numpy().view(np.uint64),
            hf.cpu().numpy().view(np.uint64))


@pytest.mark.parametrize("nf", [2, 3, 4])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_multi_field_sum(lib, orc, cuda, nf, dtype):
    """Several enabled fields are thresholded as their sum in table order, in
    the element type (src/binMapCUDA.cuh:196-201)."""
    print(f"multi-field {nf} {dtype}", flush=True)
    Z, Y, X = 23, 29, 135
    fs = [synth.grf_field((Z, Y, X), sigma=1.5, seed=60 + i, device=cuda, dtype=dtype) for i in range(nf)]
    s = host_np(fs[0]).copy()
    for f in fs[1:]:
        s = s + host_np(f)  # same element type, same order
    w_ref, h_ref = oracle_all(orc, s, ">=", 0.25)
    fi = M.FieldInfo([f.data_ptr() for f in fs], (X, Y, Z), enc_of(fs[0]))
    w, h, hf = _all_paths_on(lib, fi, X, Y, Z, ">=", 0.25, cuda)
    np.testing.assert_array_equal(w, w_ref)
    np.testing.assert_array_equal(h, h_ref)
    np.testing.assert_array_equal(hf, h_ref)


def test_multi_field_table_mask_and_limit(lib, orc, cuda):
    """Entries outside the mask and NULL entries are skipped; at most 4 of the
    remaining ones are summed (BMCU_FIELD_MAX)."""
    Z, Y, X = 11, 13, 70
    fs = [synth.grf_field((Z, Y, X), sigma=1.0, seed=70 + i, device=cuda, dtype=torch.float32) for i in range(6)]
    ptrs = [fs[0].data_ptr(), fs[1].data_ptr(), None, fs[2].data_ptr(), fs[3].data_ptr(), fs[4].data_ptr(),
            fs[5].data_ptr()]
    # entry 1 is valid but not enabled, entry 2 enabled but NULL
    fi = M.FieldInfo(ptrs, (X, Y, Z), M.ENC_F32, mask=0b1111101)
    used = [fs[0], fs[2], fs[3], fs[4]]  # entries 0, 3, 4, 5: the first four enabled non-NULL ones
    s = host_np(used[0]).copy()
    for f in used[1:]:
        s = s + host_np(f)
    w_ref, h_ref = oracle_all(orc, s, "<", 0.0)
    w, h, hf = _all_paths_on(lib, fi, X, Y, Z, "<", 0.0, cuda)
    np.testing.assert_array_equal(w, w_ref)
    np.testing.assert_array_equal(hf, h_ref)


@pytest.mark.parametrize("sx", [2, 3])
def test_x_strided_field(lib, orc, cuda, sx):
    """pS = (sx, sy, sz): an x-strided view (every sx-th element of a row)."""
    Z, Y, X = 17, 19, 150
    base = synth.grf_field((Z, Y, X * sx), sigma=2.0, seed=80 + sx, device=cuda, dtype=torch.float64)
    view = host_np(base)[:, :, ::sx]
    w_ref, h_ref = oracle_all(orc, np.ascontiguousarray(view), ">=", 0.1)
    fi = M.FieldInfo(base.data_ptr(), (X, Y, Z), M.ENC_F64, strides=(sx, X * sx, X * sx * Y))
    w, h, hf = _all_paths_on(lib, fi, X, Y, Z, ">=", 0.1, cuda)
    np.testing.assert_array_equal(w, w_ref)
    np.testing.assert_array_equal(h, h_ref)
    np.testing.assert_array_equal(hf, h_ref)
    # two strided fields summed
    base2 = synth.grf_field((Z, Y, X * sx), sigma=2.0, seed=90 + sx, device=cuda, dtype=torch.float64)
    s = view + host_np(base2)[:, :, ::sx]
    _, h_ref2 = oracle_all(orc, np.ascontiguousarray(s), ">=", 0.1)
    fi2 = M.FieldInfo([base.data_ptr(), base2.data_ptr()], (X, Y, Z), M.ENC_F64,
                      strides=(sx, X * sx, X * sx * Y))
    _, _, hf2 = _all_paths_on(lib, fi2, X, Y, Z, ">=", 0.1, cuda)
    np.testing.assert_array_equal(hf2, h_ref2)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("shape", [(6, 300, 32), (9, 70, 96), (5, 41, 160), (3, 13, 2080), (40, 200, 64)])
def test_whole_word_rows(lib, orc, cuda, shape, dtype):
    """Rows of whole map words (X a multiple of 32 but not of 1024): the map
    pass takes each plane as one row of X Y voxels. Maps (poisoned buffers)
    and histograms equal the oracle's."""
    print(f"whole-word rows {shape} {dtype}", flush=True)
    f = synth.grf_field(shape, sigma=2.0, seed=shape[2] + shape[1], device=cuda, dtype=dtype)
    for relop, thr in ((">=", 0.1), ("<", -0.3)):
        check_all_paths(lib, orc, f, relop, thr)


@pytest.mark.parametrize("width", [64, 50, 100, 29])
def test_narrow_window_of_a_wider_field(lib, orc, cuda, width):
    """A narrow window (row stride 130) keeps its own rows: its map equals
    the oracle's map of the window (64: whole words; 50, 100, 29: row groups
    of the map pass with a row stride of their own)."""
    full = synth.grf_field((12, 40, 130), sigma=2.0, seed=81, device=cuda, dtype=torch.float32)
    sub = full[2:10, 3:35, 5:5 + width]
    Z, Y, X = sub.shape
    fi = M.FieldInfo(sub.data_ptr(), (X, Y, Z), M.ENC_F32, strides=(1, 130, 130 * 40))
    w_ref, h_ref = oracle_all(orc, np.ascontiguousarray(host_np(sub)), ">", 0.0)
    w, h, hf = _all_paths_on(lib, fi, X, Y, Z, ">", 0.0, cuda)
    np.testing.assert_array_equal(w, w_ref)
    np.testing.assert_array_equal(h, h_ref)
    np.testing.assert_array_equal(hf, h_ref)


def test_invalid_inputs_return_null(lib, cuda):
    bm = M.bin_map(">=", 0.5)
    org = M.BMOrg()
    f = torch.rand((4, 4, 40), device=cuda)
    fi = M.FieldInfo(f.data_ptr(), (40, 4, 4), M.ENC_F32)
    fi.info.fieldTableMask = 0
    assert not lib.binMapCUDA(None, None, C.byref(org), fi.ref, C.byref(bm))
    fi = M.FieldInfo(f.data_ptr(), (40, 4, 4), 0x08)  # ENC_U8: not a float field
    assert not lib.binMapCUDA(None, None, C.byref(org), fi.ref, C.byref(bm))
    fi = M.FieldInfo(f.data_ptr(), (40, 4, 4), M.ENC_F32, strides=(0, 80, 320))
    assert not lib.binMapCUDA(None, None, C.byref(org), fi.ref, C.byref(bm))
    assert "stride" in M.last_error()
    assert lib.mkfCUDAGetBPFD(None, None, C.byref(org), None, 3) is None


def test_library_pbm_buffer_grows(lib, orc, cuda):
    bm = M.bin_map(">=", 0.5)
    for shape in [(8, 8, 64), (40, 50, 300)]:
        f = torch.rand(shape, device=cuda)
        Z, Y, X = shape
        fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), M.ENC_F32)
        org = M.BMOrg()
        p = lib.binMapCUDA(None, None, C.byref(org), fi.ref, C.byref(bm))
        assert p
        n = org.planeWS * (org.planePairs + 1)
        got = (C.c_uint32 * n)()
        torch.cuda.synchronize()
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so.7")
        assert hip.hipMemcpy(got, C.c_void_p(p), n * 4, 2) == 0
        w_ref = orc.pbm_c(host_np(f), bm.t[0], bm.m)
        np.testing.assert_array_equal(np.frombuffer(got, dtype=np.uint32).reshape(w_ref.shape), w_ref)
    lib.binMapCUDACleanup()


def test_host_result_and_host_field_entry_points(lib, orc, cuda):
    f64 = synth.grf_field((33, 41, 75), sigma=2.0, seed=6, device="cpu", dtype=torch.float64).numpy()
    f64 = np.ascontiguousarray(f64)
    bm = M.bin_map(">=", 0.0)
    w_ref, h_ref = oracle_all(orc, f64, ">=", 0.0)
    Z, Y, X = f64.shape
    d = M.dims((X, Y, Z))
    # mkfAccGetBPFDSimple: host double field -> host PBM + host BPFD
    h = np.zeros(256, dtype=np.uint64)
    w = np.zeros(w_ref.size, dtype=np.uint32)
    ok = lib.mkfAccGetBPFDSimple(h.ctypes.data_as(C.POINTER(C.c_size_t)), w.ctypes.data, f64.ctypes.data, d,
                                 C.byref(bm))
    assert ok
    np.testing.assert_array_equal(w.reshape(w_ref.shape), w_ref)
    np.testing.assert_array_equal(h, h_ref)
    # mkfAccCUDAGetBPFD with both encodings
    for arr, enc in ((f64, M.ENC_F64), (f64.astype(np.float32), M.ENC_F32)):
        _, h_e = oracle_all(orc, arr, ">=", 0.0)
        h = np.zeros(256, dtype=np.uint64)
        assert lib.mkfAccCUDAGetBPFD(h.ctypes.data_as(C.POINTER(C.c_size_t)), arr.ctypes.data, d, enc, C.byref(bm))
        np.testing.assert_array_equal(h, h_e)
    # mkfCUDAGetBPFDH: library host buffer and caller buffer
    field = torch.from_numpy(f64).to(cuda)
    _, dw, org, _ = dev_pbm(lib, field, ">=", 0.0)
    p = lib.mkfCUDAGetBPFDH(None, None, C.byref(org), dw.data_ptr(), 3)
    assert p
    got = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint64)), shape=(256,)).copy()
    np.testing.assert_array_equal(got, h_ref)
    h = np.zeros(256, dtype=np.uint64)
    ki = M.KernInfo()
    p2 = lib.mkfCUDAGetBPFDH(C.byref(ki), h.ctypes.data, C.byref(org), dw.data_ptr(), 2)
    assert p2 == h.ctypes.data
    np.testing.assert_array_equal(h, h_ref)
    assert ki.dtms[1] > 0
    lib.mkfCUDACleanup(0)


def test_auto_context(lib, orc, cuda):
    X, Y, Z = 96, 40, 30
    f = synth.uniform_field((Z, Y, X), seed=2)
    w_ref, h_ref = oracle_all(orc, f, ">=", 0.5)
    ctx = M.Context()
    ctx.nField, ctx.nElem, ctx.enc, ctx.bytesElem = 1, X * Y * Z, M.ENC_F32, 4
    ctx.bytesF = f.nbytes
    ctx.nU = w_ref.size
    ctx.bytesU = w_ref.size * 4
    ctx.nZ, ctx.bytesZ = 256, 2048
    hu = np.zeros(w_ref.size, dtype=np.uint32)
    hz = np.zeros(256, dtype=np.uint64)
    ctx.pHF, ctx.pHU, ctx.pHZ = f.ctypes.data, hu.ctypes.data, hz.ctypes.data
    dz = torch.zeros(256, dtype=torch.int64, device=cuda)
    ctx.pDZ = dz.data_ptr()
    assert lib.mkfCUDAGetBPFDautoCtx(C.byref(ctx), M.dims((X, Y, Z)), C.byref(M.bin_map(">=", 0.5)), 0) == 1
    np.testing.assert_array_equal(hu.reshape(w_ref.shape), w_ref)
    np.testing.assert_array_equal(hz, h_ref)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipFree(C.c_void_p(ctx.pDF))
    hip.hipFree(C.c_void_p(ctx.pDU))


def test_config_small_uniform_256(lib, orc, cuda):
    """256^3 U[0,1) float32, threshold >= 0.5: PBM, BPFD and measures."""
    f_np = synth.uniform_field((256, 256, 256), seed=1)
    f = torch.from_numpy(f_np).to(cuda)
    h_ref = check_all_paths(lib, orc, f, ">=", 0.5)
    s = 3.0 / (256 * 3)
    m_ref, _ = orc.measures(h_ref, s)
    m = (C.c_float * 4)()
    sym = C.create_string_buffer(16)
    h, _ = dev_fused(lib, f, ">=", 0.5)
    lib.mkfSelectMeasuresBPFD(m, sym, h.ctypes.data_as(C.POINTER(C.c_size_t)), C.c_float(s), 2)
    np.testing.assert_array_equal(np.array(list(m), np.float32), m_ref)
    assert abs(m[3] - 0.5) < 1e-3


def test_config_odd_dims_sdf(lib, orc, cuda):
    """509 x 511 x 517 float32 SDF of 32 spheres and 8 boxes, "<=" 0."""
    f = synth.sdf_field((517, 511, 509), seed=11, device=cuda, dtype=torch.float32)
    check_all_paths(lib, orc, f, "<=", 0.0)


def test_auto_context_two_fields(lib, orc, cuda):
    """A context holding two fields back to back (autoFieldPtr): the PBM is
    of their sum."""
    X, Y, Z = 70, 30, 20
    a = synth.uniform_field((Z, Y, X), seed=3)
    b = synth.uniform_field((Z, Y, X), seed=4)
    f = np.ascontiguousarray(np.stack([a, b]))  # field 0 then field 1
    w_ref, h_ref = oracle_all(orc, a + b, ">=", 1.0)
    ctx = M.Context()
    ctx.nField, ctx.nElem, ctx.enc, ctx.bytesElem = 2, X * Y * Z, M.ENC_F32, 4
    ctx.bytesF = f.nbytes
    ctx.nU = w_ref.size
    ctx.bytesU = w_ref.size * 4
    ctx.nZ, ctx.bytesZ = 256, 2048
    hu = np.zeros(w_ref.size, dtype=np.uint32)
    hz = np.zeros(256, dtype=np.uint64)
    ctx.pHF, ctx.pHU, ctx.pHZ = f.ctypes.data, hu.ctypes.data, hz.ctypes.data
    dz = torch.zeros(256, dtype=torch.int64, device=cuda)
    ctx.pDZ = dz.data_ptr()
    assert lib.mkfCUDAGetBPFDautoCtx(C.byref(ctx), M.dims((X, Y, Z)), C.byref(M.bin_map(">=", 1.0)), 0) == 1
    np.testing.assert_array_equal(hu.reshape(w_ref.shape), w_ref)
    np.testing.assert_array_equal(hz, h_ref)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipFree(C.c_void_p(ctx.pDF))
    hip.hipFree(C.c_void_p(ctx.pDU))


@pytest.mark.parametrize("slab", [1, 3, 7, 0, 1000])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_host_stream_slabs(lib, orc, cuda, slab, dtype):
    """A host field streamed through the fused kernel in z-slabs that share
    one plane: every plane pair is counted once, for any slab depth."""
    print(f"host stream slab={slab} {dtype}", flush=True)
    Z, Y, X = 37, 29, 101
    f = synth.grf_field((Z, Y, X), sigma=2.0, seed=12, device="cpu", dtype=torch.float64).numpy()
    f = np.ascontiguousarray(f.astype(dtype))
    enc = M.ENC_F32 if dtype == np.float32 else M.ENC_F64
    _, h_ref = oracle_all(orc, f, ">=", 0.0)
    h = np.full(256, 5, dtype=np.uint64)  # the histogram is added to
    bm = M.bin_map(">=", 0.0)
    M.check(lib.mkfHostStreamBPFD(h.ctypes.data_as(C.POINTER(C.c_size_t)), f.ctypes.data, enc, M.dims((X, Y, Z)),
                                  C.byref(bm), slab), "mkfHostStreamBPFD")
    np.testing.assert_array_equal(h, h_ref + 5)


def test_host_stream_pinned_and_edges(lib, orc, cuda):
    bm = M.bin_map("<", 0.3)
    f = synth.grf_field((64, 48, 300), sigma=3.0, seed=13, device=cuda, dtype=torch.float32)
    p = torch.empty(f.shape, dtype=f.dtype, pin_memory=True)
    p.copy_(f)
    _, h_ref = oracle_all(orc, p.numpy(), "<", 0.3)
    h = np.zeros(256, dtype=np.uint64)
    M.check(lib.mkfHostStreamBPFD(h.ctypes.data_as(C.POINTER(C.c_size_t)), p.data_ptr(), M.ENC_F32,
                                  M.dims((300, 48, 64)), C.byref(bm), 5))
    np.testing.assert_array_equal(h, h_ref)
    # two planes: one pair, one slab
    two = np.ascontiguousarray(p.numpy()[:2])
    _, h2 = oracle_all(orc, two, "<", 0.3)
    h = np.zeros(256, dtype=np.uint64)
    M.check(lib.mkfHostStreamBPFD(h.ctypes.data_as(C.POINTER(C.c_size_t)), two.ctypes.data, M.ENC_F32,
                                  M.dims((300, 48, 2)), C.byref(bm), 0))
    np.testing.assert_array_equal(h, h2)
    # a dim of one: no cells, nothing added
    h = np.zeros(256, dtype=np.uint64)
    M.check(lib.mkfHostStreamBPFD(h.ctypes.data_as(C.POINTER(C.c_size_t)), two.ctypes.data, M.ENC_F32,
                                  M.dims((300, 48, 1)), C.byref(bm), 0))
    assert not h.any()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fused_autotune(lib, orc, cuda, dtype):
    """The autotuner leaves the caller's histogram alone, returns a supported
    tiling, and later calls without a tiling (which use it) stay exact."""
    f = synth.grf_field((40, 50, 333), sigma=2.0, seed=14, device=cuda, dtype=dtype)
    Z, Y, X = f.shape
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), enc_of(f))
    bm = M.bin_map(">", -0.2)
    t = M.MkfTiling()
    M.check(lib.mkfFusedAutotune(fi.ref, C.byref(bm), C.byref(t), None), "autotune")
    assert t.planePairs >= 1 and 1 <= t.waves <= 16 and t.rowPairs in (2, 4, 8)
    assert t.voxelsPerLane in ((1, 4, 8, 16) if dtype == torch.float32 else (1, 2))
    _, h_ref = oracle_all(orc, host_np(f), ">", -0.2)
    h, _ = dev_fused(lib, f, ">", -0.2)
    np.testing.assert_array_equal(h, h_ref)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_crofton_measures_of_device_histogram(lib, orc, cuda, dtype):
    """A hollow ball thresholded on the GPU: the histogram equals the oracle's
    and its Crofton measures meet the closed forms (chi = 2 exactly; M and S
    within the digitisation error of radii 20 and 12, 0.2 %)."""
    print(f"crofton hollow ball {dtype}", flush=True)
    n = 52
    c = torch.tensor([n / 2 + 0.3, n / 2 - 0.2, n / 2 + 0.1], dtype=torch.float64)
    z, y, x = torch.meshgrid(*(torch.arange(n, dtype=torch.float64),) * 3, indexing="ij")
    r = torch.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    f = torch.maximum(r - 20, 12 - r).to(dtype).to(cuda)
    _, h_ref = oracle_all(orc, host_np(f), "<=", 0.0)
    h, _ = dev_fused(lib, f, "<=", 0.0)
    np.testing.assert_array_equal(h, h_ref)
    m = M.crofton_measures(h, 0.0)
    assert m[0] == pytest.approx(8 * math.pi, abs=1e-9)
    assert m[1] == pytest.approx(4 * math.pi * 8, rel=2e-3)
    assert m[2] == pytest.approx(4 * math.pi * (400 + 144), rel=2e-3)
    d = M.crofton_measures(h, 1.0 / n)
    assert d[3] == pytest.approx(volume_fraction := float(m[3]) / float(h.sum()), rel=1e-15)
    assert 0 < volume_fraction < 1
