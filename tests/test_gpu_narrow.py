"""Rows narrower than one PBM word (X < 32) under each narrow-row mode:
* cells (default): the X-1 cells of each row pair, as the oracle's direct
  enumeration counts them;
* reference: the reference GPU walk's 2X-1 patterns per row pair
  (src/bpfdCUDA.cuh:92-118), as the oracle's restatement of that walk counts
  them, through binMapCUDA + mkfCUDAGetBPFD[H], mkfBPFDAsync and the host
  field entry point mkfAccCUDAGetBPFD; mkfAccGetBPFDSimple (whose reference
  is the CPU walk, undefined there) refuses;
* reject: every drop-in histogram entry point refuses and leaves the
  histogram untouched.
From X = 32 on every mode gives the cell histogram."""

import ctypes as C

import numpy as np
import pytest
import torch

import mkf_amd as M
from gpu_util import dev_pbm, enc_of, host_np

pytestmark = pytest.mark.gpu


@pytest.fixture
def mode(lib):
    yield lib.mkfSetNarrowRowMode
    lib.mkfSetNarrowRowMode(M.MKF_NARROW_CELLS)


def hist_of(lib, words, org, cuda, pre=0):
    h = torch.full((256,), pre, dtype=torch.int64, device=cuda)
    r = lib.mkfCUDAGetBPFD(None, h.data_ptr(), C.byref(org), words.data_ptr(), M.MKFCU_PROFILE_FAST)
    torch.cuda.synchronize()
    return r, h.cpu().numpy().view(np.uint64).copy()


def host_entry(lib, field):
    Z, Y, X = field.shape
    f = np.ascontiguousarray(host_np(field))
    out = np.zeros(256, dtype=np.uint64)
    ok = lib.mkfAccCUDAGetBPFD(out.ctypes.data_as(C.POINTER(C.c_size_t)), f.ctypes.data, M.dims((X, Y, Z)),
                               enc_of(field), C.byref(M.bin_map(">=", 0.0)))
    return ok, out


@pytest.mark.parametrize("X", [2, 3, 5, 17, 31, 32, 33])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_narrow_rows_every_mode(lib, orc, cuda, mode, X, dtype):
    Y, Z = 23, 19
    g = torch.Generator(device=cuda).manual_seed(X)
    field = torch.randn((Z, Y, X), generator=g, device=cuda, dtype=dtype)
    words_h, words, org, _ = dev_pbm(lib, field, ">=", 0.0)
    cells = orc.bpfd_direct_c(words_h, X)
    walk = orc.bpfd_gpu_walk(words_h, X)
    if X >= 32:
        np.testing.assert_array_equal(walk, cells)
    else:
        assert int(walk.sum()) == (2 * X - 1) * (Y - 1) * (Z - 1)

    assert mode(M.MKF_NARROW_CELLS) == M.MKF_NARROW_CELLS
    r, h = hist_of(lib, words, org, cuda)
    assert r
    np.testing.assert_array_equal(h, cells)
    ok, hh = host_entry(lib, field)
    assert ok
    np.testing.assert_array_equal(hh, cells)

    mode(M.MKF_NARROW_REFERENCE)
    r, h = hist_of(lib, words, org, cuda)
    assert r
    np.testing.assert_array_equal(h, walk)
    hb = torch.zeros(256, dtype=torch.int64, device=cuda)
    M.check(lib.mkfBPFDAsync(hb.data_ptr(), C.byref(org), words.data_ptr(), None, None), "mkfBPFDAsync")
    torch.cuda.synchronize()
    np.testing.assert_array_equal(hb.cpu().numpy().view(np.uint64), walk)
    hH = np.zeros(256, dtype=np.uint64)
    assert lib.mkfCUDAGetBPFDH(None, hH.ctypes.data, C.byref(org), words.data_ptr(), M.MKFCU_PROFILE_FAST)
    np.testing.assert_array_equal(hH, walk)
    ok, hh = host_entry(lib, field)
    assert ok
    np.testing.assert_array_equal(hh, walk)

    mode(M.MKF_NARROW_REJECT)
    r, h = hist_of(lib, words, org, cuda, pre=3)
    if X >= 32:
        assert r
        np.testing.assert_array_equal(h, cells + 3)
    else:
        assert not r and "narrow-row" in M.last_error()
        assert np.all(h == 3)  # untouched
        assert lib.mkfBPFDAsync(hb.data_ptr(), C.byref(org), words.data_ptr(), None, None) == M.MKF_ERR_UNSUPPORTED
        ok, _ = host_entry(lib, field)
        assert not ok


@pytest.mark.parametrize("X", [5, 31])
def test_cpu_entry_point_refuses_outside_cells_mode(lib, orc, cuda, mode, X):
    """mkfAccGetBPFDSimple (a double field on the host) counts cells in the
    default mode and refuses in the others."""
    Y, Z = 11, 9
    f = np.random.default_rng(X).standard_normal((Z, Y, X))
    words = np.zeros(orc.pbm_shape(X, Y, Z), dtype=np.uint32)
    out = np.zeros(256, dtype=np.uint64)
    bm = M.bin_map(">=", 0.0)

    def call():
        return lib.mkfAccGetBPFDSimple(out.ctypes.data_as(C.POINTER(C.c_size_t)), words.ctypes.data, f.ctypes.data,
                                       M.dims((X, Y, Z)), C.byref(bm))
    assert call()
    np.testing.assert_array_equal(out, orc.bpfd_field(f, bm.t[0], bm.m)[0])
    for m in (M.MKF_NARROW_REFERENCE, M.MKF_NARROW_REJECT):
        mode(m)
        assert not call() and "mkfAccGetBPFDSimple" in M.last_error()
