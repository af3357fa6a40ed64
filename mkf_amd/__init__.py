"""Python binding of libmkf_amd.so, the MI355X (gfx950) MKF library.

The C ABI (include/mkf.h) is the product; this module only declares its
structs and functions for ctypes so Python tests and benchmarks call exactly
what a C caller would. Names mirror the reference's C API
(binMapCUDA, mkfCUDAGetBPFD, mkfSelectMeasuresBPFD, ...).

Loading fails loudly when the shared library has not been built: there is no
CPU fallback for any GPU entry point.
"""

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmkf_amd.so")

MKF_BINS = 256
ENC_F32 = 0xF4
ENC_F64 = 0xF8
ENC_U1 = 0x01
BMC_LT, BMC_EQ, BMC_GT, BMC_NV, BMC_AL = 1, 2, 4, 0, 7
MKFCU_PROFILE_FAST = 3
MKFCU_PROFILE_FLEX = 2
MKF_OK = 0
MKF_ERR_ARG, MKF_ERR_MEM, MKF_ERR_HIP, MKF_ERR_UNSUPPORTED, MKF_ERR_BUDGET = -1, -2, -3, -4, -5


class BinMapF64(C.Structure):
    """src/binMapUtil.h:28-32"""
    _fields_ = [("t", C.c_double * 1), ("m", C.c_uint32)]


class BMOrg(C.Structure):
    """src/binMapUtil.h:52-57"""
    _fields_ = [("rowElem", C.c_uint32), ("rowPairs", C.c_uint32), ("planePairs", C.c_uint32),
                ("rowWS", C.c_uint32), ("planeWS", C.c_uint32)]


class ConstFieldPtr(C.Union):
    """src/binMapUtil.h:65"""
    _fields_ = [("p", C.c_void_p), ("w", C.c_size_t)]


class BMFieldInfo(C.Structure):
    """src/binMapUtil.h:69-79 (the anonymous leading struct is flattened)."""
    _fields_ = [("fieldTableMask", C.c_uint32), ("elemID", C.c_uint8), ("oprID", C.c_uint8),
                ("profID", C.c_uint8), ("flags", C.c_uint8),
                ("pD", C.POINTER(C.c_int)), ("pS", C.POINTER(C.c_long)),
                ("pFieldDevPtrTable", C.POINTER(ConstFieldPtr))]


class KernInfo(C.Structure):
    """src/binMapUtil.h:107-110"""
    _fields_ = [("dtms", C.c_float * 2)]


class Context(C.Structure):
    """src/binMapUtil.h:113-125"""
    _fields_ = [("nField", C.c_int), ("nElem", C.c_int), ("enc", C.c_uint8), ("bytesElem", C.c_uint8),
                ("pad", C.c_uint8 * 2), ("bytesF", C.c_long), ("nU", C.c_long), ("bytesU", C.c_long),
                ("nZ", C.c_long), ("bytesZ", C.c_long), ("pDF", C.c_void_p), ("pHF", C.c_void_p),
                ("pDU", C.c_void_p), ("pHU", C.c_void_p), ("pDZ", C.c_void_p), ("pHZ", C.c_void_p),
                ("bmo", BMOrg), ("ki", KernInfo)]


class MkfTiling(C.Structure):
    _fields_ = [("planePairs", C.c_int32), ("waves", C.c_int32), ("rowPairs", C.c_int32),
                ("voxelsPerLane", C.c_int32)]


class MkfLaunchInfo(C.Structure):
    _fields_ = [("tiles", C.c_int64), ("workgroups", C.c_int64), ("wavesPerGroup", C.c_int32),
                ("used", MkfTiling)]


class MkfBatchField(C.Structure):
    _fields_ = [("pF", C.c_void_p), ("enc", C.c_uint8), ("def_", C.c_int * 3), ("nLevels", C.c_int32),
                ("pM", C.POINTER(BinMapF64)), ("pBPFD", C.c_void_p)]


class MkfBatchShape(C.Structure):
    _fields_ = [("planePairs", C.c_int32), ("segments", C.c_int32)]


class MkfServeConfig(C.Structure):
    _fields_ = [("hbmBudget", C.c_size_t), ("maxInFlight", C.c_int32), ("maxBatch", C.c_int32),
                ("shape", MkfBatchShape)]


class MkfServeRequest(C.Structure):
    _fields_ = [("pF", C.c_void_p), ("flags", C.c_uint32), ("stream", C.c_void_p), ("enc", C.c_uint8),
                ("def_", C.c_int * 3),
                ("nLevels", C.c_int32), ("pM", C.POINTER(BinMapF64)), ("pBPFD", C.c_void_p), ("tag", C.c_uint64)]


class MkfServeResult(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int32), ("batchSize", C.c_int32), ("tSubmit", C.c_double),
                ("tAdmit", C.c_double), ("tLaunch", C.c_double), ("tDone", C.c_double)]


class MkfServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("launches", C.c_uint64),
                ("bytesIn", C.c_uint64), ("heldMax", C.c_uint64), ("gpuBusy", C.c_double)]


MKF_SERVE_ON_DEVICE = 1
MKF_SERVE_AFTER_STREAM = 2
MKF_BATCH_MAX_LEVELS = 16
MKF_MOM_N = 6
MKF_NOALLOC_SCRATCH_BYTES = 4 << 20
MKF_NARROW_CELLS, MKF_NARROW_REFERENCE, MKF_NARROW_REJECT = 0, 1, 2
MKF_BATCH_ZERO_OUTPUT = 1

_P = C.c_void_p
_SZP = C.POINTER(C.c_size_t)

# name: (restype, argtypes)
_SIGNATURES = {
    "encSizeN": (C.c_size_t, [C.POINTER(C.c_int), C.c_size_t, C.c_uint8]),
    "encAlignment": (C.c_int, [C.c_uint8]),
    "setBinMapF64": (C.POINTER(BinMapF64), [C.POINTER(BinMapF64), C.c_char_p, C.c_float]),
    "setBMO": (C.c_size_t, [C.POINTER(BMOrg), C.POINTER(C.c_int), C.c_char]),
    "copyNonZeroDef": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]),
    "copyValidPtrByMask": (C.c_int, [C.POINTER(ConstFieldPtr), C.c_int, C.POINTER(ConstFieldPtr), C.c_uint32]),
    "countValidPtrByMask": (C.c_int, [C.POINTER(ConstFieldPtr), C.c_uint32]),
    "genStride": (C.c_int, [C.POINTER(C.c_long), C.c_int, C.c_int, C.POINTER(C.c_int), C.c_long]),
    "copyOrGenStride": (C.c_int, [C.POINTER(C.c_long), C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_long)]),
    "asFieldTab": (C.POINTER(ConstFieldPtr), [C.POINTER(C.c_void_p), C.c_uint8]),
    "setupFields": (C.POINTER(BMFieldInfo), [C.POINTER(BMFieldInfo), C.POINTER(C.c_void_p), C.c_int,
                                             C.POINTER(C.c_int), C.c_int, C.c_int]),
    "autoFieldPtr": (C.POINTER(C.c_void_p), [C.POINTER(ConstFieldPtr), C.POINTER(Context)]),
    "binMapCUDA": (_P, [C.POINTER(KernInfo), _P, C.POINTER(BMOrg), C.POINTER(BMFieldInfo), C.POINTER(BinMapF64)]),
    "binMapCUDACleanup": (None, []),
    "mkfCUDAGetBPFD": (_P, [C.POINTER(KernInfo), _P, C.POINTER(BMOrg), _P, C.c_int]),
    "mkfCUDAGetBPFDH": (_P, [C.POINTER(KernInfo), _P, C.POINTER(BMOrg), _P, C.c_int]),
    "mkfCUDAGetBPFDautoCtx": (C.c_int, [C.POINTER(Context), C.POINTER(C.c_int), C.POINTER(BinMapF64), C.c_int]),
    "mkfCUDACleanup": (None, [C.c_int]),
    "mkfAccGetBPFDSimple": (C.c_int32, [_SZP, _P, _P, C.POINTER(C.c_int), C.POINTER(BinMapF64)]),
    "mkfAccCUDAGetBPFD": (C.c_int32, [_SZP, _P, C.POINTER(C.c_int), C.c_uint8, C.POINTER(BinMapF64)]),
    "mkfRefMeasureBPFD": (None, [C.POINTER(C.c_float), _SZP, C.c_float]),
    "mkfSelectMeasuresBPFD": (C.c_int, [C.POINTER(C.c_float), C.c_char_p, _SZP, C.c_float, C.c_int]),
    "volFrac": (C.c_float, [_SZP]),
    "mkfCroftonMeasuresBPFD": (C.c_int, [C.POINTER(C.c_double), _SZP, C.c_double]),
    "mkfKvWrapCount": (C.c_int, [_SZP]),
    "mkfLastError": (C.c_char_p, []),
    "mkfDebugInjectFault": (C.c_int, [C.c_char_p, C.c_int]),
    "mkfSetNarrowRowMode": (C.c_int, [C.c_int]),
    "mkfFusedLaunchInfo": (C.c_int, [C.POINTER(BMFieldInfo), C.POINTER(BinMapF64), C.POINTER(MkfTiling),
                                     C.POINTER(MkfLaunchInfo)]),
    "mkfBPFDLaunchInfo": (C.c_int, [C.POINTER(BMOrg), C.POINTER(MkfTiling), C.POINTER(MkfLaunchInfo)]),
    "mkfReadRoof": (C.c_int, [_P, C.c_size_t, C.c_int, C.c_int, C.c_int, _P, _P]),
    "mkfFusedBPFD": (C.c_int, [C.POINTER(KernInfo), _P, C.POINTER(BMOrg), C.POINTER(BMFieldInfo),
                               C.POINTER(BinMapF64), C.POINTER(MkfTiling), _P]),
    "mkfPBMAsync": (C.c_int, [_P, C.POINTER(BMOrg), C.POINTER(BMFieldInfo), C.POINTER(BinMapF64), _P]),
    "mkfPBMMomentsAsync": (C.c_int, [_P, _P, C.POINTER(BMOrg), C.POINTER(BMFieldInfo), C.POINTER(BinMapF64), _P]),
    "mkfFusedBPFDMoments": (C.c_int, [_P, _P, C.POINTER(BMOrg), C.POINTER(BMFieldInfo), C.POINTER(BinMapF64),
                                      C.POINTER(MkfTiling), _P]),
    "mkfFusedBPFDMomentsNoAlloc": (C.c_int, [_P, _P, C.POINTER(BMOrg), C.POINTER(BMFieldInfo),
                                             C.POINTER(BinMapF64), C.POINTER(MkfTiling), _P]),
    "mkfBPFDAsync": (C.c_int, [_P, C.POINTER(BMOrg), _P, C.POINTER(MkfTiling), _P]),
    "mkfDefaultTiling": (MkfTiling, [C.POINTER(C.c_int), C.c_int]),
    "mkfBatchCreate": (_P, [C.POINTER(MkfBatchField), C.c_int, C.POINTER(MkfBatchShape), C.c_int, _P]),
    "mkfBatchRun": (C.c_int, [_P, _P]),
    "mkfBatchRunDirect": (C.c_int, [_P, _P]),
    "mkfBatchAutotune": (C.c_int, [_P, C.POINTER(MkfBatchShape), C.c_int, C.c_int, C.POINTER(MkfBatchShape), _P]),
    "mkfBatchInfo": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(MkfBatchShape), C.POINTER(C.c_int)]),
    "mkfBatchDestroy": (None, [_P]),
    "mkfFusedAutotune": (C.c_int, [C.POINTER(BMFieldInfo), C.POINTER(BinMapF64), C.POINTER(MkfTiling), _P]),
    "mkfHostStreamBPFD": (C.c_int, [_SZP, _P, C.c_uint8, C.POINTER(C.c_int), C.POINTER(BinMapF64), C.c_int]),
    "mkfServeCreate": (_P, [C.POINTER(MkfServeConfig)]),
    "mkfServeSubmit": (C.c_int, [_P, C.POINTER(MkfServeRequest)]),
    "mkfServePoll": (C.c_int, [_P, C.POINTER(MkfServeResult), C.c_int, C.c_int]),
    "mkfServeClock": (C.c_double, [_P]),
    "mkfServeGetStats": (C.c_int, [_P, C.POINTER(MkfServeStats)]),
    "mkfServeDestroy": (None, [_P]),
    "mkfPatternGroupTable": (C.POINTER(C.c_uint8), []),
    "mkfGroupWeightKTable": (C.POINTER(C.c_int8), []),
}

# Functions bound here; the CPU test suite checks this list against the
# declarations in include/*.h and against the library's exports.
EXPORTED = sorted(_SIGNATURES)

_lib = None


def load():
    """Loads libmkf_amd.so (raises if it has not been built).

    torch, when installed, is imported first: its wheel carries its own HIP
    runtime with the same soname (libamdhip64.so.7), and loading torch first
    makes this library bind to that one runtime instead of a second copy."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python mkf_amd/build.py` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().mkfLastError().decode(errors="replace")


def check(rc, what="mkf call"):
    if rc != MKF_OK:
        raise RuntimeError(f"{what} failed ({rc}): {last_error()}")


# ---------------------------------------------------------------------------
# Small helpers that build the reference's argument structs.
# ---------------------------------------------------------------------------

def bin_map(relop: str, t: float) -> BinMapF64:
    """setBinMapF64 through the library (src/binMapUtil.c:37-58)."""
    m = BinMapF64()
    load().setBinMapF64(C.byref(m), relop.encode(), C.c_float(t))
    return m


def dims(def3):
    return (C.c_int * 3)(*[int(d) for d in def3])


class FieldInfo:
    """A BMFieldInfo plus the arrays it points to. dev_ptr is one device
    pointer or a table of them (None for empty entries); mask defaults to
    every entry of the table."""

    def __init__(self, dev_ptr, def3, enc: int, prof_id: int = 0, strides=None, mask=None):
        self.def3 = dims(def3)
        ptrs = list(dev_ptr) if isinstance(dev_ptr, (list, tuple)) else [dev_ptr]
        self.table = (ConstFieldPtr * len(ptrs))()
        for i, p in enumerate(ptrs):
            self.table[i].p = p
        self.strides = None if strides is None else (C.c_long * 3)(*strides)
        self.info = BMFieldInfo()
        self.info.fieldTableMask = (1 << len(ptrs)) - 1 if mask is None else mask
        self.info.elemID = enc
        self.info.oprID = 0
        self.info.profID = prof_id
        self.info.flags = 0
        self.info.pD = self.def3
        self.info.pS = self.strides if self.strides is not None else C.POINTER(C.c_long)()
        self.info.pFieldDevPtrTable = self.table

    @property
    def ref(self):
        return C.byref(self.info)


def enc_of(dtype_name: str) -> int:
    if dtype_name in ("float32", "torch.float32", "f32"):
        return ENC_F32
    if dtype_name in ("float64", "torch.float64", "f64"):
        return ENC_F64
    raise ValueError(f"unsupported dtype {dtype_name}")


def crofton_measures(hist, s: float = 0.0):
    """{Kv, Mv, Sv, Vv} (densities for s > 0, lattice-unit totals for s == 0)
    of a 256-bin histogram through mkfCroftonMeasuresBPFD."""
    import numpy as np

    h = np.ascontiguousarray(np.asarray(hist, dtype=np.uint64).reshape(256))
    m = (C.c_double * 4)()
    rc = load().mkfCroftonMeasuresBPFD(m, h.ctypes.data_as(_SZP), C.c_double(s))
    if rc != 4:
        raise RuntimeError(f"mkfCroftonMeasuresBPFD: {rc}")
    return np.array(list(m))
