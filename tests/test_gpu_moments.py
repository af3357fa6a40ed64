"""GPU: mkfPBMMomentsAsync, phase moments fused into the PBM pass, against the
oracle's exactly rounded sums (the categorical sums the reference declares in
CUDAImgMom, src/binMapCUDA.cuh:89-121, and never instantiates).

Counts are exact. The sums are floating point: the kernel adds values in
double per lane, then per wave, workgroup and grid in fixed orders. Each sum
must be within 1e-12 x (sum of the magnitudes of its terms) of the exactly
rounded sum; the bound of naive summation for the longest chain here (a few
hundred terms per lane plus ~25 tree levels) is below 1e-13 of that.
The map written alongside equals mkfPBMAsync's bit for bit."""

import ctypes as C
import math

import numpy as np
import pytest
import torch

import mkf_amd as M
from mkf_amd import synth
from tests.gpu_util import enc_of, host_np

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _moments(lib, fi, X, Y, Z, relop, thr, cuda, with_map=True, pre=None):
    bm = M.bin_map(relop, thr)
    org = M.BMOrg()
    n = lib.setBMO(C.byref(org), M.dims((X, Y, Z)), b"\x00")
    words = torch.full((n,), -1, dtype=torch.int32, device=cuda) if with_map else None
    mom = torch.zeros(M.MKF_MOM_N, dtype=torch.float64, device=cuda)
    if pre is not None:
        mom += torch.tensor(pre, dtype=torch.float64, device=cuda)
    rc = lib.mkfPBMMomentsAsync(words.data_ptr() if with_map else None, mom.data_ptr(), C.byref(org), fi.ref,
                                C.byref(bm), None)
    M.check(rc, "mkfPBMMomentsAsync")
    torch.cuda.synchronize()
    w = words.cpu().numpy().view(np.uint32).reshape(Z, Y, org.rowWS) if with_map else None
    return mom.cpu().numpy(), w


def _pbm(lib, fi, X, Y, Z, relop, thr, cuda):
    bm = M.bin_map(relop, thr)
    org = M.BMOrg()
    n = lib.setBMO(C.byref(org), M.dims((X, Y, Z)), b"\x00")
    words = torch.full((n,), -1, dtype=torch.int32, device=cuda)
    M.check(lib.mkfPBMAsync(words.data_ptr(), C.byref(org), fi.ref, C.byref(bm), None), "mkfPBMAsync")
    torch.cuda.synchronize()
    return words.cpu().numpy().view(np.uint32).reshape(Z, Y, org.rowWS)


def _check(got, f_np, relop, thr):
    bm = M.bin_map(relop, thr)
    want = M_oracle().phase_moments(f_np, bm.t[0], bm.m)
    assert got[4] == want[4] and got[5] == want[5]
    b = M_oracle().threshold_bits(f_np, bm.t[0], bm.m).ravel()
    v = f_np.astype(np.float64).ravel()
    for ph in (0, 1):
        sel = v[b == bool(ph)]
        for j, mag in ((2 * ph, np.abs(sel).sum()), (2 * ph + 1, (sel * sel).sum())):
            if not math.isfinite(want[j]):
                assert (math.isnan(got[j]) and math.isnan(want[j])) or got[j] == want[j], (j, got[j], want[j])
            else:
                assert abs(got[j] - want[j]) <= TOL * mag, (j, got[j], want[j])


def M_oracle():
    from oracle import oracle as O

    return O


# (9, 70, 96) and (3, 13, 2080): rows of whole map words, taken a plane at a
# time; (3, 5, 31), (20, 33, 50) and (7, 9, 100): row groups of 8, 4 and 2
SHAPES = [(2, 2, 2), (3, 5, 31), (4, 4, 1025), (9, 17, 200), (40, 50, 333), (3, 2, 2049), (9, 70, 96),
          (3, 13, 2080), (20, 33, 50), (7, 9, 100)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_moments_match_oracle(lib, cuda, shape, dtype):
    print(f"moments {shape} {dtype}", flush=True)
    Z, Y, X = shape
    f = synth.grf_field(shape, sigma=1.5, seed=sum(shape), device=cuda, dtype=dtype) * 3 + 0.5
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), enc_of(f))
    f_np = host_np(f)
    for relop, thr in ((">=", 0.25), ("<", -1.0), ("!=", 0.0), ("", 0.0)):
        got, w = _moments(lib, fi, X, Y, Z, relop, thr, cuda)
        _check(got, f_np, relop, thr)
        np.testing.assert_array_equal(w, _pbm(lib, fi, X, Y, Z, relop, thr, cuda))


def test_moments_accumulate_are_deterministic_and_map_optional(lib, cuda):
    Z, Y, X = 64, 96, 700
    f = synth.grf_field((Z, Y, X), sigma=2.0, seed=5, device=cuda, dtype=torch.float32)
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), M.ENC_F32)
    a, _ = _moments(lib, fi, X, Y, Z, ">", 0.1, cuda)
    b, _ = _moments(lib, fi, X, Y, Z, ">", 0.1, cuda, with_map=False)
    assert a.tobytes() == b.tobytes()  # same partials, same fixed order
    pre = [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]
    c, _ = _moments(lib, fi, X, Y, Z, ">", 0.1, cuda, pre=pre)
    np.testing.assert_allclose(c, a + np.array(pre), rtol=1e-15)
    _check(a, host_np(f), ">", 0.1)
    assert a[4] + a[5] == X * Y * Z
    # rows of whole map words (each plane taken as one row) alike
    g = f[:, :, :640].contiguous()
    fg = M.FieldInfo(g.data_ptr(), (640, Y, Z), M.ENC_F32)
    a, _ = _moments(lib, fg, 640, Y, Z, ">", 0.1, cuda)
    b, _ = _moments(lib, fg, 640, Y, Z, ">", 0.1, cuda, with_map=False)
    assert a.tobytes() == b.tobytes()
    _check(a, host_np(g), ">", 0.1)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_moments_special_values(lib, cuda, dtype):
    rng = np.random.default_rng(9)
    choices = np.array([0.0, -0.0, 1e-40, -1e-40, 0.25, -0.25, 1.0, 3.0, np.inf])
    vals = rng.choice(choices, size=(5, 7, 97))
    f = torch.from_numpy(vals).to(dtype).to(cuda)
    Z, Y, X = f.shape
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), enc_of(f))
    # inf lands in phase 1 for '>=' and in phase 0 for '<': its phase's sums are inf
    for relop in (">=", "<"):
        got, _ = _moments(lib, fi, X, Y, Z, relop, 0.5, cuda)
        _check(got, host_np(f), relop, 0.5)
        inf_ph = 1 if relop == ">=" else 0
        assert math.isinf(got[2 * inf_ph]) and math.isfinite(got[2 * (1 - inf_ph)])
    # NaN compares as EQ: with '<' it joins phase 0 (with the inf), whose sums become NaN
    vals[2, 3, 50] = np.nan
    f = torch.from_numpy(vals).to(dtype).to(cuda)
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), enc_of(f))
    got, _ = _moments(lib, fi, X, Y, Z, "<", 0.5, cuda)
    _check(got, host_np(f), "<", 0.5)
    assert math.isnan(got[0]) and math.isnan(got[1]) and math.isfinite(got[2]) and math.isfinite(got[3])
    assert got[4] + got[5] == X * Y * Z


def test_moments_of_summed_and_strided_fields(lib, cuda):
    Z, Y, X = 13, 21, 150
    fs = [synth.grf_field((Z, Y, X), sigma=1.5, seed=40 + i, device=cuda, dtype=torch.float64) for i in range(3)]
    fi = M.FieldInfo([g.data_ptr() for g in fs], (X, Y, Z), M.ENC_F64)
    s = host_np(fs[0]) + host_np(fs[1]) + host_np(fs[2])
    got, _ = _moments(lib, fi, X, Y, Z, ">=", 0.0, cuda)
    _check(got, s, ">=", 0.0)
    sx = 3
    base = synth.grf_field((Z, Y, X * sx), sigma=2.0, seed=44, device=cuda, dtype=torch.float32)
    fi = M.FieldInfo(base.data_ptr(), (X, Y, Z), M.ENC_F32, strides=(sx, X * sx, X * sx * Y))
    got, _ = _moments(lib, fi, X, Y, Z, "<=", -0.3, cuda)
    _check(got, np.ascontiguousarray(host_np(base)[:, :, ::sx]), "<=", -0.3)


def test_moments_uniform_256(lib, cuda):
    """256^3 uniform [0,1) f32 split at 0.5: phase means near 0.25 / 0.75,
    and the oracle's sums to the stated tolerance."""
    f = torch.rand((256, 256, 256), generator=torch.Generator(device=cuda).manual_seed(11), device=cuda)
    fi = M.FieldInfo(f.data_ptr(), (256, 256, 256), M.ENC_F32)
    got, _ = _moments(lib, fi, 256, 256, 256, ">=", 0.5, cuda)
    _check(got, host_np(f), ">=", 0.5)
    assert got[0] / got[4] == pytest.approx(0.25, abs=1e-3)
    assert got[2] / got[5] == pytest.approx(0.75, abs=1e-3)


def _fused_moments(lib, fi, X, Y, Z, relop, thr, cuda, tiling=None):
    bm = M.bin_map(relop, thr)
    h = torch.zeros(256, dtype=torch.int64, device=cuda)
    mom = torch.zeros(M.MKF_MOM_N, dtype=torch.float64, device=cuda)
    t = M.MkfTiling(*tiling) if tiling is not None else None
    rc = lib.mkfFusedBPFDMoments(h.data_ptr(), mom.data_ptr(), None, fi.ref, C.byref(bm),
                                 C.byref(t) if t is not None else None, None)
    M.check(rc, "mkfFusedBPFDMoments")
    torch.cuda.synchronize()
    return h.cpu().numpy().view(np.uint64).copy(), mom.cpu().numpy()


# Every voxel must be counted once. The last column x = X - 1 is held by a
# lane of a real segment (X = 1090, 333), by an extra wave of the workgroup
# past the last segment (X = 1089), or only read from memory as the halo of
# the last segment (X = 65, 1025, 2049 at one voxel per lane; 257 at four;
# 129 at two for doubles). Slabs of 3 plane pairs give several z-tiles.
FUSED_CASES = [
    ((5, 7, 65), torch.float32, None),
    ((5, 7, 65), torch.float32, (3, 16, 4, 1)),
    ((4, 6, 257), torch.float32, (3, 16, 4, 4)),
    ((4, 6, 257), torch.float32, (3, 16, 4, 1)),
    ((3, 4, 1025), torch.float32, (2, 16, 4, 1)),
    ((3, 3, 1089), torch.float32, (1, 16, 4, 1)),
    ((3, 3, 1090), torch.float32, (1, 16, 4, 1)),
    ((2, 2, 2049), torch.float32, None),
    ((11, 13, 333), torch.float32, (3, 16, 4, 4)),
    ((4, 5, 129), torch.float64, (3, 16, 4, 2)),
    ((4, 5, 129), torch.float64, (3, 16, 4, 1)),
    ((6, 9, 1025), torch.float64, None),
    ((17, 19, 333), torch.float64, (4, 16, 8, 2)),
    # voxelsPerLane 16 or 8: the persistent kernel's moments form, always
    # with 4 row pairs; ragged and exact rows, one to three segments, the
    # last column read only as the halo of the last segment (X = 1025 and
    # 2049 at 16 voxels per lane, 513 at 8).
    ((3, 4, 1025), torch.float32, (2, 16, 4, 16)),
    ((3, 4, 1024), torch.float32, (2, 16, 4, 16)),
    ((5, 6, 1030), torch.float32, (3, 16, 4, 16)),
    ((4, 5, 2049), torch.float32, (1, 16, 2, 16)),
    ((7, 9, 300), torch.float32, (3, 16, 4, 16)),
    ((6, 9, 1000), torch.float32, None),
    ((9, 6, 2100), torch.float32, (4, 16, 2, 16)),
    ((3, 4, 513), torch.float32, (2, 16, 4, 8)),
    ((5, 6, 700), torch.float32, (3, 16, 4, 8)),
    ((4, 7, 4), torch.float32, (2, 16, 4, 8)),
    ((6, 5, 5), torch.float32, (2, 16, 4, 16)),
    # more tiles (13 row stacks x 299 plane pairs) than resident workgroups:
    # each workgroup sums the moments of several tiles
    ((300, 50, 300), torch.float32, (1, 16, 4, 16)),
    ((300, 50, 300), torch.float32, (1, 16, 4, 8)),
]


@pytest.mark.parametrize("shape,dtype,tiling", FUSED_CASES)
def test_fused_moments_match_oracle(lib, orc, cuda, shape, dtype, tiling):
    print(f"fused moments {shape} {dtype} {tiling}", flush=True)
    Z, Y, X = shape
    f = synth.grf_field(shape, sigma=1.5, seed=X + 7 * Y + 31 * Z, device=cuda, dtype=dtype)
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), enc_of(f))
    f_np = host_np(f)
    for relop, thr in ((">=", 0.1), ("<", -0.5)):
        h, mom = _fused_moments(lib, fi, X, Y, Z, relop, thr, cuda, tiling)
        bm = M.bin_map(relop, thr)
        words = orc.pbm_c(f_np, bm.t[0], bm.m)
        h_ref = orc.bpfd_chunk(words, X) if X >= 32 else orc.bpfd_direct_c(words, X)
        np.testing.assert_array_equal(h, h_ref)
        _check(mom, f_np, relop, thr)


ALL_RELOPS = ["<", "=", "<=", ">", "!=", ">=", "", "!"]  # masks 1 .. 6, 0 and 7


@pytest.mark.parametrize("shape,tiling", [((5, 7, 1025), (3, 16, 4, 16)),   # last column as the halo
                                          ((6, 5, 600), (2, 16, 4, 8)),      # partial 512-voxel segment
                                          ((4, 9, 1030), (2, 16, 4, 16))])   # partial 1024-voxel segment
def test_fused_moments_special_values_all_relations(lib, orc, cuda, shape, tiling):
    """The persistent kernel's moments form on ±0, denormals, ±inf and NaN
    under all 8 relations: each value's phase follows the map bit (NaN
    compares as EQ), so the phase holding an inf or a NaN has that sum."""
    print(f"fused moments, special values {shape} {tiling}", flush=True)
    rng = np.random.default_rng(sum(shape))
    choices = np.array([0.0, -0.0, 1e-40, -1e-40, 0.25, -0.25, 0.5, 1.0, 3.0, np.inf, -np.inf], dtype=np.float32)
    vals = rng.choice(choices, size=shape).astype(np.float32)
    Z, Y, X = shape
    vals[1, 2, X - 1] = np.nan   # last column
    vals[Z - 1, Y - 1, 3] = np.nan
    f = torch.from_numpy(vals).to(cuda)
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), M.ENC_F32)
    for relop in ALL_RELOPS:
        h, mom = _fused_moments(lib, fi, X, Y, Z, relop, 0.5, cuda, tiling)
        bm = M.bin_map(relop, 0.5)
        np.testing.assert_array_equal(h, orc.bpfd_chunk(orc.pbm_c(vals, bm.t[0], bm.m), X), err_msg=relop)
        _check(mom, vals, relop, 0.5)
        assert mom[4] + mom[5] == X * Y * Z


def test_fused_moments_summed_strided_and_cell_free(lib, orc, cuda):
    Z, Y, X = 9, 11, 200
    fs = [synth.grf_field((Z, Y, X), sigma=1.5, seed=50 + i, device=cuda, dtype=torch.float32) for i in range(2)]
    fi = M.FieldInfo([g.data_ptr() for g in fs], (X, Y, Z), M.ENC_F32)
    s = host_np(fs[0]) + host_np(fs[1])
    h, mom = _fused_moments(lib, fi, X, Y, Z, ">", 0.0, cuda)
    bm = M.bin_map(">", 0.0)
    np.testing.assert_array_equal(h, orc.bpfd_chunk(orc.pbm_c(s, bm.t[0], bm.m), X))
    _check(mom, s, ">", 0.0)
    sx = 2
    base = synth.grf_field((Z, Y, X * sx), sigma=2.0, seed=58, device=cuda, dtype=torch.float64)
    fi = M.FieldInfo(base.data_ptr(), (X, Y, Z), M.ENC_F64, strides=(sx, X * sx, X * sx * Y))
    view = np.ascontiguousarray(host_np(base)[:, :, ::sx])
    h, mom = _fused_moments(lib, fi, X, Y, Z, "<=", 0.2, cuda)
    bm = M.bin_map("<=", 0.2)
    np.testing.assert_array_equal(h, orc.bpfd_chunk(orc.pbm_c(view, bm.t[0], bm.m), X))
    _check(mom, view, "<=", 0.2)
    # one plane: no cells, the histogram stays zero, the moments still come
    flat = synth.grf_field((1, 7, 90), sigma=1.0, seed=59, device=cuda, dtype=torch.float32)
    fi = M.FieldInfo(flat.data_ptr(), (90, 7, 1), M.ENC_F32)
    h, mom = _fused_moments(lib, fi, 90, 7, 1, ">=", 0.0, cuda)
    assert not h.any()
    _check(mom, host_np(flat), ">=", 0.0)
    assert mom[4] + mom[5] == 90 * 7


def test_fused_moments_agree_with_map_moments(lib, cuda):
    Z, Y, X = 70, 90, 1030
    f = synth.grf_field((Z, Y, X), sigma=2.5, seed=77, device=cuda, dtype=torch.float32)
    fi = M.FieldInfo(f.data_ptr(), (X, Y, Z), M.ENC_F32)
    _, a = _fused_moments(lib, fi, X, Y, Z, ">=", 0.3, cuda)
    b, _ = _moments(lib, fi, X, Y, Z, ">=", 0.3, cuda, with_map=False)
    assert a[4] == b[4] and a[5] == b[5]
    np.testing.assert_allclose(a[:4], b[:4], rtol=1e-12)
    _, a2 = _fused_moments(lib, fi, X, Y, Z, ">=", 0.3, cuda)
    assert a.tobytes() == a2.tobytes()
