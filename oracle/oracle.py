"""CPU ORACLE for the MKF hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only to check results (or as the timed CPU baseline).
The product library never calls it.

Two independent restatements of the reference's CPU path are provided:
  * the C library oracle/liboracle.so (mkf_oracle.c): the reference's
    threshold (src/binMapACC.c:32-93), 64-bit chunk walk
    (src/mkfACC.c:28-188) and measures (src/mkfUtil.c:64-135);
  * numpy brute force below, written from the definitions alone:
    threshold bit (m >> d) & 1 with d = 1 - (f<t) + (f>t); PBM bit x & 31
    of word (x >> 5) + y*rowWS + z*planeWS; cell code bit i =
    P(x+(i&1), y+((i>>1)&1), z+(i>>2)).

How each part is pinned (no golden vectors exist in the reference; it has no
tests, fixtures or known-answer tables, and it cannot be built here because
its un-vendored ../Common/src/util.h is absent):
  * threshold: numpy's IEEE comparisons for all 8 relation masks, NaN, +-0,
    infinities, denormals, and float rounding of the threshold;
  * PBM: numpy bit packing (np.packbits, little bit order) of the same
    threshold bits;
  * histogram: C chunk walk == numpy direct enumeration == C direct
    enumeration on random and structured volumes; closed-form invariants
    (total = (X-1)(Y-1)(Z-1), all-zero/all-one volumes, single voxel ->
    bins 1,2,4,..,128, complement maps bin b -> 255-b, axis swaps permute
    code bits, translation invariance);
  * Kv: Euler weights from the textbook cubical-complex count (V-E+F-C)
    and orbits from the 48 cube symmetries, checked against closed forms
    K = 4*pi*chi (voxel, box, two boxes, hollow box with cavity, ring) and
    against the reference tables group by group. The reference's weights
    differ from the count for two orbits (0x3d: -8 vs -9, 0x7e: -12 vs -18);
    the oracle applies the reference's two values, which only the text of
    its table (src/mkfUtil.c:33-38) pins. Vv: set lower corners over cells,
    checked by direct counting.
  * Parity unpinned: the reference's book measures Mv and Sv (refMeasures.c
    is absent from its tree; as it compiles both are -1). The library's
    non-reference Crofton estimates are checked by oracle/crofton.py.
  * phase moments: the categorical sums of src/binMapCUDA.cuh:89-121
    (declared, never instantiated in the reference) restated with exactly
    rounded sums (math.fsum), pinned by closed forms on integer ramps and
    constant fields and by brute-force loops on tiny volumes.
"""

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_U64P = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_U32P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_F32P = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_F64P = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_U8P = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        from mkf_amd import build as _b  # builds only the oracle here
        _b.build_oracle()
    lib = C.CDLL(LIB_PATH)
    I, L, D, U = C.c_int, C.c_long, C.c_double, C.c_uint32
    sig = {
        "orc_relbit_f64": (U, [D, D, U]),
        "orc_pbm_f64": (None, [_U32P, _F64P, I, I, I, D, U]),
        "orc_pbm_f32": (None, [_U32P, _F32P, I, I, I, D, U]),
        "orc_bpfd_chunk": (I, [_U64P, _U32P, I, I, I, I, L]),
        "orc_bpfd_chunk_range": (I, [_U64P, _U32P, I, I, I, I, L, I, I]),
        "orc_bpfd_direct": (None, [_U64P, _U32P, I, I, I, I, L]),
        "orc_bpfd_gpu_walk": (None, [_U64P, _U32P, I, I, I, I, L]),
        "orc_bpfd_field_f64": (I, [_U64P, _U32P, _F64P, I, I, I, D, U]),
        "orc_bpfd_field_f32": (I, [_U64P, _U32P, _F32P, I, I, I, D, U]),
        "orc_num_threads": (I, []),
        "orc_set_threads": (None, [I]),
        "orc_euler_weight": (I, [U]),
        "orc_reference_weight": (I, [U]),
        "orc_orbits": (I, [_U8P]),
        "orc_measures": (I, [C.POINTER(C.c_float), _U64P, C.c_float, C.POINTER(C.c_int)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


# ---------------------------------------------------------------------------
# Geometry of the packed map
# ---------------------------------------------------------------------------

def row_words(X):
    return (int(X) + 31) >> 5


def pbm_shape(X, Y, Z):
    """(Z, Y, rowWS) word array shape of a full-size PBM."""
    return (int(Z), int(Y), row_words(X))


# ---------------------------------------------------------------------------
# Threshold (numpy, from the definition)
# ---------------------------------------------------------------------------

def threshold_bits(field, t, m):
    """Boolean array of threshold bits; field is (Z, Y, X). A float32 field
    is compared with t rounded to float32, a float64 field with t as double."""
    f = np.asarray(field)
    tt = f.dtype.type(t)
    d = 1 - (f < tt).astype(np.int32) + (f > tt).astype(np.int32)
    return ((int(m) >> d) & 1).astype(bool)


def pack_bits(bits):
    """(Z, Y, X) booleans -> (Z, Y, rowWS) uint32 words, bit x & 31 of word
    x >> 5, tail bits zero."""
    Z, Y, X = bits.shape
    nW = row_words(X)
    padded = np.zeros((Z, Y, nW * 32), dtype=bool)
    padded[:, :, :X] = bits
    by = np.packbits(padded, axis=2, bitorder="little")  # (Z, Y, nW*4) bytes
    return np.ascontiguousarray(by).view(np.uint32).reshape(Z, Y, nW)


def unpack_bits(words, X):
    Z, Y, nW = words.shape
    by = np.ascontiguousarray(words).view(np.uint8).reshape(Z, Y, nW * 4)
    return np.unpackbits(by, axis=2, bitorder="little")[:, :, :X].astype(bool)


def pbm_numpy(field, t, m):
    return pack_bits(threshold_bits(field, t, m))


def phase_moments(field, t, m):
    """{sum0, sumsq0, sum1, sumsq1, n0, n1} of a (Z, Y, X) field split by its
    threshold bits: per phase the exactly rounded sum of the values (as
    double) and of their squares (each square rounded to double), and the
    voxel count. Layout of CUDAImgMom::addCat (src/binMapCUDA.cuh:117-121:
    category c at 2c, its square at 2c + 1), with the counts appended. A
    sum holding a NaN, or both infinities, is NaN in every order of
    addition (math.fsum raises for inf + -inf); one that holds infinities of
    one sign only is that infinity."""
    import math

    def exact_sum(a):
        if not np.all(np.isfinite(a)):
            if np.isnan(a).any() or (np.isposinf(a).any() and np.isneginf(a).any()):
                return math.nan
        return math.fsum(a)

    f = np.asarray(field)
    b = threshold_bits(f, t, m).ravel()
    v = f.astype(np.float64).ravel()
    out = []
    for ph in (False, True):
        sel = v[b == ph]
        out += [exact_sum(sel), exact_sum(sel * sel)]
    n1 = int(np.count_nonzero(b))
    return np.array(out + [float(b.size - n1), float(n1)])


# ---------------------------------------------------------------------------
# Histogram (numpy, from the definition)
# ---------------------------------------------------------------------------

def cell_codes(bits):
    """8-bit code of every cell of a (Z, Y, X) boolean volume."""
    b = bits.astype(np.uint16)
    code = np.zeros((b.shape[0] - 1, b.shape[1] - 1, b.shape[2] - 1), dtype=np.uint16)
    for i in range(8):
        dx, dy, dz = i & 1, (i >> 1) & 1, i >> 2
        code |= b[dz:dz + code.shape[0], dy:dy + code.shape[1], dx:dx + code.shape[2]] << i
    return code


def bpfd_numpy(bits):
    """256-bin histogram (uint64) of the cell codes of a boolean volume."""
    if min(bits.shape) < 2:
        return np.zeros(256, dtype=np.uint64)
    return np.bincount(cell_codes(bits).ravel(), minlength=256).astype(np.uint64)


# ---------------------------------------------------------------------------
# C restatement wrappers
# ---------------------------------------------------------------------------

def pbm_c(field, t, m):
    f = np.ascontiguousarray(field)
    Z, Y, X = f.shape
    w = np.zeros(pbm_shape(X, Y, Z), dtype=np.uint32)
    if f.dtype == np.float32:
        load().orc_pbm_f32(w, f, X, Y, Z, float(t), int(m))
    elif f.dtype == np.float64:
        load().orc_pbm_f64(w, f, X, Y, Z, float(t), int(m))
    else:
        raise TypeError(f.dtype)
    return w


def bpfd_chunk(words, X):
    """Chunk-walk histogram of a (Z, Y, rowWS) PBM (requires X >= 32)."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    Z, Y, nW = w.shape
    h = np.zeros(256, dtype=np.uint64)
    if min(X, Y, Z) < 2:
        return h
    rc = load().orc_bpfd_chunk(h, w, int(X), Y, Z, nW, nW * Y)
    if rc != 0:
        raise ValueError("the chunk walk needs X >= 32")
    return h


def bpfd_direct_c(words, X):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    Z, Y, nW = w.shape
    h = np.zeros(256, dtype=np.uint64)
    load().orc_bpfd_direct(h, w, int(X), Y, Z, nW, nW * Y)
    return h


def bpfd_gpu_walk(words, X):
    """The reference GPU row walk (src/bpfdCUDA.cuh:92-118) over a (Z, Y,
    rowWS) PBM: equal to the cell histogram for X >= 32; for X < 32 it emits
    2X - 1 patterns per row pair."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    Z, Y, nW = w.shape
    h = np.zeros(256, dtype=np.uint64)
    if min(X, Y, Z) < 2:
        return h
    load().orc_bpfd_gpu_walk(h, w, int(X), Y, Z, nW, nW * Y)
    return h


def narrow_walk_numpy(words, X):
    """Closed form of the GPU row walk for X < 32 on a map with zero tail
    bits, from the register contents alone: after the first word's X - 1
    patterns the registers hold bit X - 1 of each row; ORing the same word
    in, shifted up by one, leaves bits (X-1, 0, 1, ..., X-1), whose X
    patterns are the wrap-around cell (x = X-1 beside x = 0) and the cells
    0..X-2 again. So the walk is twice the cell histogram plus the histogram
    of the wrap-around cells."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    Z, Y, nW = w.shape
    assert X < 32 and nW == 1
    bits = ((w[:, :, 0, None] >> np.arange(X, dtype=np.uint32)) & 1).astype(np.uint8)  # (Z, Y, X)
    cells = bpfd_numpy(bits)
    wrap = np.stack([bits[:, :, X - 1], bits[:, :, 0]], axis=2)  # (Z, Y, 2): x-low = X-1, x-high = 0
    return (2 * cells + bpfd_numpy(wrap)).astype(np.uint64)


def bpfd_field(field, t, m):
    """The reference CPU path (mkfAccGetBPFDSimple restated): (hist, pbm).
    X >= 32 uses the chunk walk; smaller X the direct enumeration."""
    f = np.ascontiguousarray(field)
    Z, Y, X = f.shape
    w = pbm_c(f, t, m)
    if X >= 32:
        return bpfd_chunk(w, X), w
    return bpfd_direct_c(w, X), w


def measures(hist, s):
    """(m[4], overflow) from the oracle's derived weights and orbits."""
    h = np.ascontiguousarray(hist, dtype=np.uint64)
    m = (C.c_float * 4)()
    ovf = C.c_int(0)
    load().orc_measures(m, h, C.c_float(s), C.byref(ovf))
    return np.array(list(m), dtype=np.float32), bool(ovf.value)


def euler_weights():
    """Textbook weight 3V - 6E + 12F - 24C of every pattern."""
    lib = load()
    return np.array([lib.orc_euler_weight(p) for p in range(256)], dtype=np.int32)


def reference_weights():
    """Weight the reference applies to every pattern (orbit-wise)."""
    lib = load()
    lab, _ = orbits()
    return np.array([lib.orc_reference_weight(int(lab[p])) for p in range(256)], dtype=np.int32)


def orbits():
    lab = np.zeros(256, dtype=np.uint8)
    n = load().orc_orbits(lab)
    return lab, n


def set_threads(n):
    load().orc_set_threads(int(n))


def num_threads():
    return load().orc_num_threads()
