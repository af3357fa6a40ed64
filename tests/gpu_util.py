"""Helpers for GPU tests: call the C ABI with torch-owned device memory."""

import ctypes as C

import numpy as np
import torch

import mkf_amd as M


def enc_of(t):
    return M.ENC_F32 if t.dtype == torch.float32 else M.ENC_F64


def dev_pbm(lib, field, relop, thr, prof_id=0):
    """binMapCUDA on a device field (Z, Y, X) -> (words (Z, Y, rowWS) uint32
    numpy, BMOrg, KernInfo)."""
    Z, Y, X = field.shape
    bm = M.bin_map(relop, thr)
    org = M.BMOrg()
    n = M.load().setBMO(C.byref(org), M.dims((X, Y, Z)), b"\x00")
    words = torch.full((n,), -1, dtype=torch.int32, device=field.device)  # poison
    fi = M.FieldInfo(field.data_ptr(), (X, Y, Z), enc_of(field), prof_id)
    ki = M.KernInfo()
    r = lib.binMapCUDA(C.byref(ki), words.data_ptr(), C.byref(org), fi.ref, C.byref(bm))
    assert r == words.data_ptr(), M.last_error()
    host = words.cpu().numpy().view(np.uint32).reshape(Z, Y, org.rowWS)
    return host, words, org, ki


def dev_bpfd_from_pbm(lib, words, org, profile=M.MKFCU_PROFILE_FAST):
    h = torch.zeros(256, dtype=torch.int64, device=words.device)
    ki = M.KernInfo()
    r = lib.mkfCUDAGetBPFD(C.byref(ki), h.data_ptr(), C.byref(org), words.data_ptr(), profile)
    assert r == h.data_ptr()
    return h.cpu().numpy().view(np.uint64).copy(), ki


def dev_fused(lib, field, relop, thr, tiling=None, stream=None, pre=None):
    """mkfFusedBPFD (blocking, timed) -> (hist uint64 numpy, KernInfo)."""
    Z, Y, X = field.shape
    bm = M.bin_map(relop, thr)
    h = torch.zeros(256, dtype=torch.int64, device=field.device)
    if pre is not None:
        h += torch.from_numpy(pre.astype(np.int64)).to(field.device)
    fi = M.FieldInfo(field.data_ptr(), (X, Y, Z), enc_of(field))
    ki = M.KernInfo()
    t = None
    if tiling is not None:
        t = M.MkfTiling(*tiling)
    rc = lib.mkfFusedBPFD(C.byref(ki), h.data_ptr(), None, fi.ref, C.byref(bm),
                          C.byref(t) if t is not None else None, stream)
    M.check(rc, "mkfFusedBPFD")
    return h.cpu().numpy().view(np.uint64).copy(), ki


def host_np(field):
    return field.detach().cpu().numpy()
