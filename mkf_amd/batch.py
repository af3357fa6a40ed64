"""Host-side wrapper of the batched C entry points (mkfBatchCreate & co.).

A Batch owns the ctypes descriptions of its fields and one contiguous device
tensor of histograms, (sum of levels) x 256 uint64 stored as int64, whose
rows for field f are `hist(f)`. Every call goes through libmkf_amd.so."""

import ctypes as C

import torch

from . import (ENC_F32, ENC_F64, MKF_BATCH_ZERO_OUTPUT, BinMapF64, MkfBatchField, MkfBatchShape, bin_map, check,
               last_error, load)


def shape(plane_pairs=0, segments=0):
    s = MkfBatchShape()
    s.planePairs, s.segments = int(plane_pairs), int(segments)
    return s


def _maps(levels):
    if isinstance(levels, C.Array):
        return levels
    arr = (BinMapF64 * len(levels))()
    for j, lv in enumerate(levels):
        arr[j] = lv if isinstance(lv, BinMapF64) else bin_map(*lv)
    return arr


class Batch:
    """fields: contiguous (Z, Y, X) float32/float64 device tensors.
    levels: per field, a list of (relop, threshold) pairs or BinMapF64."""

    def __init__(self, fields, levels, shape_=None, zero=True, stream=None):
        self.lib = load()
        if len(fields) != len(levels):
            raise ValueError("one list of levels per field")
        self.fields = list(fields)
        self.maps = [_maps(lv) for lv in levels]
        n_lv = [len(m) for m in self.maps]
        dev = self.fields[0].device
        self.offsets = [0]
        for n in n_lv:
            self.offsets.append(self.offsets[-1] + n)
        self.out = torch.zeros((self.offsets[-1], 256), dtype=torch.int64, device=dev)
        self.desc = (MkfBatchField * len(self.fields))()
        for i, (f, m) in enumerate(zip(self.fields, self.maps)):
            if not f.is_contiguous() or f.dim() != 3 or f.dtype not in (torch.float32, torch.float64):
                raise ValueError(f"field {i}: need a contiguous 3-D float32/float64 tensor")
            Z, Y, X = f.shape
            d = self.desc[i]
            d.pF = f.data_ptr()
            d.enc = ENC_F32 if f.dtype == torch.float32 else ENC_F64
            d.def_[0], d.def_[1], d.def_[2] = X, Y, Z
            d.nLevels = len(m)
            d.pM = C.cast(m, C.POINTER(BinMapF64))
            d.pBPFD = self.out[self.offsets[i]].data_ptr() if len(m) else 0
        self.stream = stream
        sp = C.byref(shape_) if shape_ is not None else None
        self.handle = self.lib.mkfBatchCreate(self.desc, len(self.fields), sp,
                                              MKF_BATCH_ZERO_OUTPUT if zero else 0, self._s(stream))
        if not self.handle:
            raise RuntimeError(f"mkfBatchCreate failed: {last_error()}")

    @staticmethod
    def _s(stream):
        if stream is None:
            return None
        return C.c_void_p(stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream))

    def hist(self, i):
        """(levels, 256) int64 view of field i's histograms."""
        return self.out[self.offsets[i]:self.offsets[i + 1]]

    def run(self, stream=None, graph=True):
        fn = self.lib.mkfBatchRun if graph else self.lib.mkfBatchRunDirect
        check(fn(self.handle, self._s(stream if stream is not None else self.stream)), "batch run")

    def autotune(self, candidates, reps=3, stream=None):
        arr = (MkfBatchShape * len(candidates))(*[shape(*c) for c in candidates])
        best = MkfBatchShape()
        check(self.lib.mkfBatchAutotune(self.handle, arr, len(candidates), int(reps), C.byref(best),
                                        self._s(stream if stream is not None else self.stream)), "autotune")
        return best.planePairs, best.segments

    def info(self):
        n, g, s = C.c_int(0), C.c_int(0), MkfBatchShape()
        check(self.lib.mkfBatchInfo(self.handle, C.byref(n), C.byref(s), C.byref(g)), "info")
        return {"items": n.value, "planePairs": s.planePairs, "segments": s.segments, "grid": g.value}

    def close(self):
        if self.handle:
            self.lib.mkfBatchDestroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
