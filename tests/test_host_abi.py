"""CPU tests of libmkf_amd.so's host side: exports, struct layouts, the
reference's helper functions, and the measures (host code) against the
oracle. No GPU entry point is called here."""

import ctypes as C
import math
import os
import re
import subprocess

import numpy as np
import pytest

import mkf_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("mkf.h", "mkf_ext.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "defined"):
                names.add(name)
    return names


def test_every_declared_symbol_is_exported(lib):
    decl = declared_functions()
    assert len(decl) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", M.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = decl - exported
    assert not missing, missing
    assert decl == set(M.EXPORTED), decl ^ set(M.EXPORTED)
    for name in decl:
        assert hasattr(lib, name)


def test_struct_layouts():
    assert C.sizeof(M.BinMapF64) == 16
    assert C.sizeof(M.BMOrg) == 20
    assert C.sizeof(M.BMFieldInfo) == 32
    assert M.BMFieldInfo.pD.offset == 8
    assert C.sizeof(M.KernInfo) == 8
    assert M.Context.bytesF.offset == 16


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """The ctypes mirrors have the sizes and offsets gcc gives the header's structs."""
    probes = {
        "BinMapF64": ["t", "m"], "BMOrg": ["rowElem", "planePairs"], "BMFieldInfo": ["pD", "pFieldDevPtrTable"],
        "Context": ["bytesF", "pDF", "bmo", "ki"], "MkfTiling": ["planePairs", "voxelsPerLane"],
        "MkfBatchField": ["pF", "enc", "def", "nLevels", "pM", "pBPFD"],
        "MkfBatchShape": ["planePairs", "segments"],
        "MkfServeConfig": ["hbmBudget", "maxInFlight", "maxBatch", "shape"],
        "MkfServeRequest": ["pF", "flags", "stream", "enc", "def", "nLevels", "pM", "pBPFD", "tag"],
        "MkfServeResult": ["tag", "status", "batchSize", "tSubmit", "tDone"],
        "MkfServeStats": ["submitted", "bytesIn", "heldMax", "gpuBusy"],
    }
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mkf.h"', "int main(void){"]
    for s, fs in probes.items():
        lines.append(f'printf("{s} %zu\\n", sizeof({s}));')
        for f in fs:
            lines.append(f'printf("{s}.{f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    got = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                       text=True).stdout.splitlines())
    for s, fs in probes.items():
        cls = getattr(M, s)
        assert int(got[s]) == C.sizeof(cls), s
        for f in fs:
            assert int(got[f"{s}.{f}"]) == getattr(cls, "def_" if f == "def" else f).offset, (s, f)


@pytest.mark.parametrize("relop,m", [(">=", 6), ("<=", 3), ("!=", 5), ("<", 1), ("=", 2), (">", 4),
                                      ("", 0), ("!", 7), ("!<", 6), ("!>=", 1), ("<>", 5), (">=x<", 6),
                                      ("=!", 2), ("<<", 1)])
def test_setBinMapF64_parsing(lib, relop, m):
    bm = M.bin_map(relop, 0.5)
    assert bm.m == m
    assert bm.t[0] == 0.5


def test_setBinMapF64_stores_float_threshold(lib):
    bm = M.bin_map(">=", 0.1)
    assert bm.t[0] == float(np.float32(0.1))


@pytest.mark.parametrize("X,Y,Z", [(256, 256, 256), (509, 511, 517), (32, 2, 2), (33, 7, 5), (1, 1, 1)])
def test_setBMO_full(lib, X, Y, Z):
    o = M.BMOrg()
    n = lib.setBMO(C.byref(o), M.dims((X, Y, Z)), b"\x00")
    rw = (X + 31) // 32
    assert (o.rowElem, o.rowPairs, o.planePairs, o.rowWS, o.planeWS) == (X, Y - 1, Z - 1, rw, rw * Y)
    assert n == rw * Y * Z


def test_setBMO_partial_buffer(lib):
    o = M.BMOrg()
    n = lib.setBMO(C.byref(o), M.dims((100, 50, 40)), b"\x01")
    assert (o.rowPairs, o.planePairs, o.planeWS) == (1, 1, 4 * 2)
    assert n == 4 * 2 * 2


def test_copyNonZeroDef(lib):
    d = (C.c_int * 3)()
    assert lib.copyNonZeroDef(d, M.dims((1, 256, 256)), 3) == 2
    assert list(d) == [256, 256, 1]
    assert lib.copyNonZeroDef(d, M.dims((1, 1, 1)), 3) == 0
    assert list(d) == [0, 0, 0]
    assert lib.copyNonZeroDef(d, M.dims((5, 1, 7)), 3) == 2
    assert list(d) == [5, 7, 1]


def test_valid_ptr_by_mask(lib):
    tab = (M.ConstFieldPtr * 4)()
    tab[0].p, tab[1].p, tab[2].p, tab[3].p = 0x1000, None, 0x3000, 0x4000
    assert lib.countValidPtrByMask(tab, 0b1111) == 3
    assert lib.countValidPtrByMask(tab, 0b0010) == 0
    out = (M.ConstFieldPtr * 4)()
    assert lib.copyValidPtrByMask(out, 4, tab, 0b1101) == 3
    assert [out[i].w for i in range(3)] == [0x1000, 0x3000, 0x4000]
    assert lib.copyValidPtrByMask(out, 1, tab, 0b1100) == 1
    assert out[0].w == 0x3000


def test_genStride(lib):
    fs = (C.c_long * 3)()
    assert lib.genStride(fs, 3, 0, M.dims((4, 5, 6)), 1) == 3
    assert list(fs) == [1, 4, 20]
    assert lib.genStride(fs, 1, 2, M.dims((4, 5, 6)), 1) == 1
    assert fs[0] == 20
    assert lib.genStride(fs, 3, 0, M.dims((4, 5, 6)), 0) == 0
    assert lib.copyOrGenStride(fs, 2, 1, M.dims((4, 5, 6)), (C.c_long * 3)(1, 7, 70)) == 2
    assert list(fs)[:2] == [7, 70]


def test_encodings(lib):
    bits = C.c_int(0)
    assert lib.encSizeN(C.byref(bits), 10, M.ENC_F32) == 40 and bits.value == 32
    assert lib.encSizeN(C.byref(bits), 10, M.ENC_F64) == 80 and bits.value == 64
    assert lib.encSizeN(C.byref(bits), 64, M.ENC_U1) == 8 and bits.value == 1
    assert lib.encAlignment(M.ENC_F32) == 4 and lib.encAlignment(M.ENC_F64) == 8
    assert lib.encAlignment(M.ENC_U1) == 1


def test_setupFields(lib):
    fi = M.BMFieldInfo()
    ptrs = (C.c_void_p * 2)(0x1000, 0x2000)
    d = M.dims((8, 8, 8))
    assert lib.setupFields(C.byref(fi), ptrs, 2, d, 8, 3)
    assert fi.elemID == M.ENC_F64 and fi.fieldTableMask == 3 and fi.profID == 3
    assert not lib.setupFields(C.byref(fi), ptrs, 1, d, 2, 0)


# ---------------------------------------------------------------------------
# measures (host code in the product) vs the oracle's derivation
# ---------------------------------------------------------------------------

def product_measures(lib, h, s):
    m = (C.c_float * 4)()
    sym = C.create_string_buffer(16)
    a = np.ascontiguousarray(h, dtype=np.uint64)
    n = lib.mkfSelectMeasuresBPFD(m, sym, a.ctypes.data_as(C.POINTER(C.c_size_t)), C.c_float(s), 2)
    return n, np.array(list(m), dtype=np.float32), sym.raw[:12]


def test_tables_match_derivation(lib, orc):
    group = np.ctypeslib.as_array(lib.mkfPatternGroupTable(), shape=(256,)).copy()
    weight = np.ctypeslib.as_array(lib.mkfGroupWeightKTable(), shape=(22,)).copy()
    lab, n = orc.orbits()
    assert n == 22 and group.max() == 21
    # same partition of the 256 patterns
    for p in range(256):
        for q in range(256):
            if lab[p] == lab[q]:
                assert group[p] == group[q]
    assert len(set(group.tolist())) == 22
    # Euler weights: the cubical-complex count everywhere except the two
    # orbits where the reference's table holds other values.
    ew = orc.euler_weights()
    ref = weight[group].astype(np.int32)
    differ = sorted({int(lab[p]) for p in range(256) if ref[p] != ew[p]})
    assert differ == [0x3D, 0x7E]
    assert ref[0x3D] == -8 and ew[0x3D] == -9
    assert ref[0x7E] == -12 and ew[0x7E] == -18
    np.testing.assert_array_equal(ref, orc.reference_weights())


@pytest.mark.parametrize("seed", range(6))
def test_measures_match_oracle(lib, orc, seed):
    rng = np.random.default_rng(seed)
    scale = [10, 10 ** 6, 10 ** 9, 10 ** 12, 2 ** 40, 3][seed]
    h = rng.integers(0, scale, size=256, dtype=np.uint64)
    for s in (0.0, 3.0 / 768, 1.0, -1.0):
        n, m, sym = product_measures(lib, h, s)
        mo, _ = orc.measures(h, s)
        assert n == 2 and sym == b"Kv       Vv\x00"
        np.testing.assert_array_equal(m, mo)


def test_measures_int32_wrap_is_reproduced(lib, orc):
    h = np.zeros(256, dtype=np.uint64)
    h[0b11101111] = 3 * 10 ** 9  # one group, weight -3 or similar, count > 2^31
    h[1] = 2 ** 33 + 5            # truncated to 5 by the int conversion
    n, m, _ = product_measures(lib, h, 0.0)
    mo, ovf = orc.measures(h, 0.0)
    assert ovf
    np.testing.assert_array_equal(m, mo)


def test_measures_on_shapes(lib, orc):
    # box with chi = 1 inside a padded volume: Kv numerator = 4 pi
    bits = np.pad(np.ones((5, 6, 7), bool), 2)
    h = orc.bpfd_numpy(bits)
    _, m, _ = product_measures(lib, h, 0.0)
    assert m[0] == np.float32(24 * math.pi / 6)
    lower = bits[:-1, :-1, :-1]
    assert m[3] == np.float32(lower.sum()) / np.float32(lower.size)


def test_ref_measures_are_undefined(lib):
    m = (C.c_float * 4)()
    h = np.ones(256, dtype=np.uint64)
    lib.mkfRefMeasureBPFD(m, h.ctypes.data_as(C.POINTER(C.c_size_t)), C.c_float(1.0))
    assert list(m) == [-1, -1, -1, -1]


def test_volFrac(lib):
    h = np.zeros(256, dtype=np.uint64)
    h[1::2] = 3
    h[0::2] = 1
    assert lib.volFrac(h.ctypes.data_as(C.POINTER(C.c_size_t))) == np.float32(384) / np.float32(512)


@pytest.mark.parametrize("dims,eb,want", [
    ((1024, 1024, 1024), 4, (16, 4, None)),  # whole 1024-voxel segments: 16 voxels per lane
    ((1000, 1000, 1000), 4, (16, 4, None)),  # rows fill 98 % of their segment
    ((900, 1000, 1000), 4, (16, 4, None)),   # 88 %
    ((1536, 1000, 700), 4, (8, 4, None)),    # whole 512-voxel segments: 8 voxels per lane
    ((512, 1024, 2048), 4, (8, 4, None)),    # one whole 512-voxel segment
    ((300, 1000, 3000), 4, (8, 4, None)),    # rows of 257..513 voxels: 8 voxels per lane
    ((900, 20, 20), 4, (4, 4, 16)),          # a small field: waves would walk too few planes
    ((1536, 20, 20), 4, (4, 4, 16)),
    ((300, 20, 20), 4, (1, 4, 16)),          # rows of 257..513 voxels on a small field: 1 voxel per lane
    ((769, 1000, 1000), 4, (16, 4, None)),   # one segment filled 75 %: 16 voxels per lane
    ((650, 1000, 1000), 4, (4, 4, 16)),      # one segment filled 63 %: 4 voxels per lane
    ((1600, 1000, 1000), 4, (4, 4, 16)),     # two segments filled 78 %: 4 voxels per lane
    ((1030, 1000, 1000), 4, (4, 4, 16)),     # 50 % of 1024, 67 % of 512
    ((1200, 100, 100), 4, (4, 4, 16)),       # 59 % of 1024, 78 % of 512
    ((257, 20, 20), 4, (1, 4, 16)),          # rows of at most 257 voxels: 1 voxel per lane
    ((514, 20, 20), 4, (4, 4, 16)),          # rows over 513 voxels: 4 voxels per lane
    ((400, 20, 20), 4, (1, 4, 16)),          # 257..513 voxels on a small grid: 1 voxel per lane
    ((350, 20, 20), 4, (1, 4, 16)),
    ((200, 1000, 5000), 4, (1, 4, 16)),
    ((1024, 1024, 1024), 8, (2, 4, 16)),     # doubles: 2 voxels per lane
    ((1000, 1000, 1000), 8, (2, 4, 16)),     # doubles with a partial 128-voxel segment: 2, with the partial walk
    ((1030, 1000, 1000), 8, (2, 4, 16)),
])
def test_default_tiling_choice(lib, dims, eb, want):
    """The built-in fused tiling (host logic only; the slab depth depends on
    the device, so only its range is checked)."""
    t = lib.mkfDefaultTiling(M.dims(dims), eb)
    v, ty, waves = want
    assert t.voxelsPerLane == v and t.rowPairs == ty
    if waves is not None:
        assert t.waves == waves
    assert 1 <= t.planePairs <= 65536


@pytest.mark.parametrize("dims", [(1, 5, 5), (0, 5, 5), (1, 1, 1), (5, 0, 5), (5, 5, -7), (1 << 30, 2, 2)])
@pytest.mark.parametrize("eb", [4, 8])
def test_default_tiling_of_degenerate_dims(lib, dims, eb):
    """Rows without cells, empty or negative dims and very long rows still
    give a supported tiling (no division by a zero segment count)."""
    t = lib.mkfDefaultTiling(M.dims(dims), eb)
    assert t.voxelsPerLane in ((1, 4, 8, 16) if eb == 4 else (1, 2))
    assert t.rowPairs in (2, 4, 8) and 1 <= t.waves <= 16 and 1 <= t.planePairs <= 65536


def test_null_arguments_do_not_crash(lib):
    """Entry points that read their pointers unchecked in the reference
    (setBinMapF64, setBMO, the measures, the field-table helpers and the
    context wrapper) and the tiling query refuse NULL instead (no GPU work)."""
    assert not lib.setBinMapF64(None, b">=", 0.5)
    bm = M.BinMapF64()
    lib.setBinMapF64(C.byref(bm), None, 0.5)
    assert bm.m == 0 and bm.t[0] == 0.5
    org = M.BMOrg(7, 7, 7, 7, 7)
    assert lib.setBMO(C.byref(org), None, b"\x00") == 0 and org.rowWS == 0 and org.rowPairs == 0
    assert lib.mkfSelectMeasuresBPFD(None, None, None, 1.0, 0) == 0
    assert lib.volFrac(None) == -1.0
    lib.mkfRefMeasureBPFD(None, None, 1.0)
    t = lib.mkfDefaultTiling(None, 4)
    assert (t.planePairs, t.waves, t.rowPairs, t.voxelsPerLane) == (0, 0, 0, 0) and "NULL dims" in M.last_error()
    assert lib.copyNonZeroDef(None, None, 3) == 0
    assert lib.copyValidPtrByMask(None, 4, None, 0xF) == 0
    assert lib.countValidPtrByMask(None, 0xF) == 0
    assert not lib.asFieldTab(None, M.ENC_F32)
    assert not lib.autoFieldPtr(None, None)
    assert lib.mkfCUDAGetBPFDautoCtx(None, None, None, 0) == 0 and "NULL context" in M.last_error()


def test_fault_points_validate_their_arguments(lib):
    """The error-path test hook accepts only known points (no GPU work)."""
    assert lib.mkfDebugInjectFault(b"no.such.point", 1) == M.MKF_ERR_ARG
    assert "unknown point" in M.last_error()
    assert lib.mkfDebugInjectFault(b"batch.upload", -1) == M.MKF_ERR_ARG
    assert lib.mkfDebugInjectFault(b"batch.upload", 3) == M.MKF_OK
    assert lib.mkfDebugInjectFault(b"batch.upload", 0) == M.MKF_OK


def test_narrow_row_mode_setter(lib):
    """mkfSetNarrowRowMode returns the previous mode and refuses unknown ones
    (no GPU work)."""
    prev = lib.mkfSetNarrowRowMode(M.MKF_NARROW_REFERENCE)
    try:
        assert prev == M.MKF_NARROW_CELLS
        assert lib.mkfSetNarrowRowMode(M.MKF_NARROW_REJECT) == M.MKF_NARROW_REFERENCE
        assert lib.mkfSetNarrowRowMode(7) == M.MKF_ERR_ARG and "narrow-row mode" in M.last_error()
        assert lib.mkfSetNarrowRowMode(M.MKF_NARROW_CELLS) == M.MKF_NARROW_REJECT
    finally:
        lib.mkfSetNarrowRowMode(M.MKF_NARROW_CELLS)


def _kv_terms(lib, h):
    """Group sums and weights from the library's tables (checked against
    their derivation in test_tables_match_derivation)."""
    group = np.ctypeslib.as_array(lib.mkfPatternGroupTable(), shape=(256,)).astype(np.int64)
    weight = np.ctypeslib.as_array(lib.mkfGroupWeightKTable(), shape=(22,)).astype(np.int64)
    g = np.zeros(22, dtype=np.uint64)
    np.add.at(g, group, np.asarray(h, dtype=np.uint64))
    return g, weight


def _kv_op_order(lib, h, s):
    """Kv restated with numpy scalars in the reference's order and
    precisions (src/mkfUtil.c:64-69, 108-115): 32-bit wrapping products
    summed in int64, times pi/6 in double, stored as float; then times the
    float 1 / (s^3 t) whose denominator is formed in float and inverted in
    double."""
    g, w = _kv_terms(lib, h)
    k = 0
    for gi, wi in zip(g, w):
        p = (int(wi) & 0xFFFFFFFF) * (int(gi) & 0xFFFFFFFF) & 0xFFFFFFFF
        k += p - (1 << 32) if p >= 1 << 31 else p
    m0 = np.float32(np.float64(k) * np.float64(math.pi) / np.float64(6))
    s32 = np.float32(s)
    denom = (s32 * s32 * s32) * np.float32(int(np.asarray(h, dtype=np.uint64).sum()))
    return np.float32(m0 * np.float32(1.0 / np.float64(denom)))


def test_kv_density_scaling_follows_float_op_order(lib, orc):
    """s > 0: the library's Kv equals the float/double op-order restatement
    on cases chosen so that evaluating 1/(s^3 t) in double throughout rounds
    differently; a library that scaled in double would fail here."""
    rng = np.random.default_rng(12)
    found = 0
    for _ in range(4000):
        h = rng.integers(0, 10 ** 7, size=256, dtype=np.uint64)
        s = float(rng.uniform(0.001, 0.5))
        want = _kv_op_order(lib, h, s)
        g, w = _kv_terms(lib, h)
        naive = np.float32(float(int((g.astype(np.int64) * w).sum())) * math.pi / 6
                           / (float(np.float32(s)) ** 3 * float(h.sum())))
        if want == naive:
            continue
        found += 1
        _, m, _ = product_measures(lib, h, s)
        mo, _ = orc.measures(h, s)
        assert m[0] == want and mo[0] == want and m[0] != naive
        if found == 20:
            break
    assert found == 20


def test_kv_wrap_count(lib, orc):
    """mkfKvWrapCount: 0 where Kv is exact, else the number of wrapped terms,
    and nonzero exactly where the oracle reports a wrap."""
    def count(h):
        a = np.ascontiguousarray(h, dtype=np.uint64)
        return lib.mkfKvWrapCount(a.ctypes.data_as(C.POINTER(C.c_size_t)))
    rng = np.random.default_rng(3)
    h = rng.integers(0, 10 ** 6, size=256, dtype=np.uint64)
    assert count(h) == 0 and not orc.measures(h, 0.0)[1]
    group, weight = (np.ctypeslib.as_array(lib.mkfPatternGroupTable(), shape=(256,)),
                     np.ctypeslib.as_array(lib.mkfGroupWeightKTable(), shape=(22,)))
    # one pattern of each of two weighted groups with a count whose product wraps
    pats = [p for p in range(256) if weight[group[p]] != 0]
    g1 = pats[0]
    g2 = next(p for p in pats if group[p] != group[g1])
    w1 = abs(int(weight[group[g1]]))
    h = np.zeros(256, dtype=np.uint64)
    h[g1] = (2 ** 31) // w1 + 1          # |w| x n just above 2^31: wraps
    assert count(h) == 1 and orc.measures(h, 0.0)[1]
    h[g2] = 2 ** 33 + 7                  # does not fit 32 bits
    assert count(h) == 2 and orc.measures(h, 0.0)[1]
    h[0] = 2 ** 40                       # group 0 has weight 0: never counted
    assert count(h) == 2
    assert lib.mkfKvWrapCount(None) == -1
