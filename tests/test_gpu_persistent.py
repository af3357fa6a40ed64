"""The persistent histogram kernels (kFused16 with 16 and 8 voxels per lane,
kBPFD16) on grids where every workgroup walks several tiles, so each carries
its cell count, all-one tally and LDS histogram from tile to tile. Shallow
slabs (1 or 2 plane pairs) on a wide field give several times more tiles
than resident slots; the launch geometry is read back from the library and
printed. Rows of 2000 voxels end inside their last segment (ragged), and
every relation runs. One shape checks that the built-in tiling itself picks
a multi-tile persistent grid. Reference grid: the whole-volume launch of
src/mkfCUDA.cu:23-31, one thread per row pair."""

import ctypes as C

import numpy as np
import pytest
import torch

import mkf_amd as M
from mkf_amd import synth

pytestmark = pytest.mark.gpu

RELOPS = [">=", "<", "<=", ">", "=", "!=", "", "!"]


@pytest.fixture(scope="module")
def wide(cuda):
    Z, Y, X = 48, 1200, 2000
    f = synth.grf_field((Z, Y, X), 2.0, seed=77, device=cuda, dtype=torch.float32)
    # exact-threshold values so "=" and "!=" see both outcomes
    f.view(-1)[::97] = 0.25
    torch.cuda.synchronize()
    return f, f.cpu().numpy()


def info_fused(lib, field, bm, tiling):
    Z, Y, X = field.shape
    fi = M.FieldInfo(field.data_ptr(), (X, Y, Z), M.ENC_F32)
    li = M.MkfLaunchInfo()
    t = M.MkfTiling(*tiling) if tiling is not None else None
    M.check(lib.mkfFusedLaunchInfo(fi.ref, C.byref(bm), C.byref(t) if t else None, C.byref(li)), "launch info")
    return li


def fused(lib, field, bm, tiling):
    Z, Y, X = field.shape
    h = torch.zeros(256, dtype=torch.int64, device=field.device)
    fi = M.FieldInfo(field.data_ptr(), (X, Y, Z), M.ENC_F32)
    t = M.MkfTiling(*tiling) if tiling is not None else None
    M.check(lib.mkfFusedBPFD(None, h.data_ptr(), None, fi.ref, C.byref(bm), C.byref(t) if t else None, None),
            "mkfFusedBPFD")
    torch.cuda.synchronize()
    return h.cpu().numpy().view(np.uint64).copy()


@pytest.mark.parametrize("vpl,ty", [(16, 4), (16, 2), (8, 4), (8, 2)])
def test_fused16_multi_tile_every_relation(lib, orc, wide, vpl, ty):
    field, host = wide
    tiling = (1, 16, ty, vpl)
    for rel in RELOPS:
        bm = M.bin_map(rel, 0.25)
        li = info_fused(lib, field, bm, tiling)
        print(f"kFused16 V={vpl} Ty={ty} '{rel}': {li.tiles} tiles on {li.workgroups} workgroups "
              f"({li.tiles / max(li.workgroups, 1):.1f} per workgroup)", flush=True)
        assert (li.used.voxelsPerLane, li.used.rowPairs) == (vpl, ty)
        assert li.wavesPerGroup == 1 and li.tiles >= 3 * li.workgroups
        ref, _ = orc.bpfd_field(host, bm.t[0], bm.m)
        np.testing.assert_array_equal(fused(lib, field, bm, tiling), ref, err_msg=f"relation '{rel}'")


@pytest.mark.parametrize("ty,tp", [(4, 1), (2, 1), (4, 2)])
def test_bpfd16_multi_tile(lib, orc, wide, cuda, ty, tp):
    field, host = wide
    Z, Y, X = field.shape
    for rel in (">=", "=", "<"):
        bm = M.bin_map(rel, 0.25)
        org = M.BMOrg()
        nw = lib.setBMO(C.byref(org), M.dims((X, Y, Z)), b"\x00")
        words = torch.empty(nw, dtype=torch.int32, device=cuda)
        fi = M.FieldInfo(field.data_ptr(), (X, Y, Z), M.ENC_F32)
        M.check(lib.mkfPBMAsync(words.data_ptr(), C.byref(org), fi.ref, C.byref(bm), None), "mkfPBMAsync")
        t = M.MkfTiling(tp, 16, ty, 16)
        li = M.MkfLaunchInfo()
        M.check(lib.mkfBPFDLaunchInfo(C.byref(org), C.byref(t), C.byref(li)), "launch info")
        print(f"kBPFD16 Ty={ty} TP={tp} '{rel}': {li.tiles} tiles on {li.workgroups} workgroups "
              f"({li.tiles / max(li.workgroups, 1):.1f} per workgroup)", flush=True)
        assert li.used.voxelsPerLane == 16 and li.tiles >= 2 * li.workgroups
        h = torch.zeros(256, dtype=torch.int64, device=cuda)
        M.check(lib.mkfBPFDAsync(h.data_ptr(), C.byref(org), words.data_ptr(), C.byref(t), None), "mkfBPFDAsync")
        torch.cuda.synchronize()
        ref, _ = orc.bpfd_field(host, bm.t[0], bm.m)
        np.testing.assert_array_equal(h.cpu().numpy().view(np.uint64), ref, err_msg=f"relation '{rel}'")
        del words


def test_default_tiling_picks_multi_tile_grid(lib, orc, cuda):
    """Rows of 1024 voxels and more row stacks than resident slots: the
    built-in tiling is the persistent 16-voxel kernel with shallow slabs, and
    its workgroups walk several tiles."""
    Z, Y, X = 100, 8200, 1024
    field = synth.grf_field((Z, Y, X), 3.0, seed=78, device=cuda, dtype=torch.float32)
    bm = M.bin_map(">=", 0.1)
    li = info_fused(lib, field, bm, None)
    print(f"built-in tiling {li.used.planePairs},{li.used.waves},{li.used.rowPairs},{li.used.voxelsPerLane}: "
          f"{li.tiles} tiles on {li.workgroups} workgroups", flush=True)
    assert li.used.voxelsPerLane in (8, 16) and li.tiles >= 3 * li.workgroups
    got = fused(lib, field, bm, None)
    ref, _ = orc.bpfd_field(field.cpu().numpy(), bm.t[0], bm.m)
    np.testing.assert_array_equal(got, ref)
    del field
    torch.cuda.empty_cache()
