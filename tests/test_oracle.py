"""Pins the CPU oracle before GPU results are compared with it.

The reference ships no golden vectors, fixtures or tests for this path and
cannot be built here, so every check below compares the oracle's restatement
of the reference (C, oracle/mkf_oracle.c) with something written
independently of it: numpy's IEEE comparisons and bit packing, a direct
per-cell enumeration of the cell-code definition, closed-form invariants,
and the Euler characteristic of known shapes.
"""

import math

import numpy as np
import pytest

from mkf_amd import synth

RELOPS = {"<": 1, "=": 2, "<=": 3, ">": 4, "!=": 5, ">=": 6}


def rand_bits(shape, p, seed):
    rng = np.random.default_rng(seed)
    return rng.random(shape) < p


# ---------------------------------------------------------------------------
# threshold bit
# ---------------------------------------------------------------------------

def test_relbit_matches_ieee_definition(orc):
    lib = orc.load()
    vals = [-math.inf, -1.5, -0.0, 0.0, 1e-310, 0.5, 1.0, math.inf, math.nan]
    for m in range(8):
        for t in [0.0, -0.0, 0.5, math.inf]:
            for f in vals:
                d = 1 - (f < t) + (f > t)
                assert lib.orc_relbit_f64(f, t, m) == ((m >> d) & 1), (f, t, m)


def test_relbit_nan_selects_eq(orc):
    lib = orc.load()
    for m in range(8):
        assert lib.orc_relbit_f64(math.nan, 0.0, m) == ((m >> 1) & 1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("X", [1, 7, 31, 32, 33, 64, 100, 509])
def test_pbm_c_matches_numpy_packing(orc, dtype, X):
    rng = np.random.default_rng(X)
    f = rng.standard_normal((3, 5, X)).astype(dtype)
    f[0, 0, : min(X, 4)] = [np.nan, -0.0, 0.0, 1e-40][: min(X, 4)]
    for m in range(8):
        for t in (0.0, 0.1, -0.3):
            np.testing.assert_array_equal(orc.pbm_c(f, t, m), orc.pbm_numpy(f, t, m))


def test_f32_threshold_is_rounded_to_float(orc):
    # 0.1 is not a float; a float32 field equal to float32(0.1) is EQ.
    f = np.full((2, 2, 32), np.float32(0.1), dtype=np.float32)
    w = orc.pbm_c(f, 0.1, 2)  # "=" relation
    assert np.all(w == 0xFFFFFFFF)
    w64 = orc.pbm_c(f.astype(np.float64), 0.1, 2)  # double compare: not equal
    assert np.all(w64 == 0)


def test_pbm_numpy_bit_order(orc):
    bits = np.zeros((1, 1, 70), dtype=bool)
    bits[0, 0, [0, 5, 31, 32, 63, 69]] = True
    w = orc.pack_bits(bits)
    assert w.shape == (1, 1, 3)
    assert w[0, 0, 0] == (1 | (1 << 5) | (1 << 31))
    assert w[0, 0, 1] == (1 | (1 << 31))
    assert w[0, 0, 2] == (1 << 5)
    np.testing.assert_array_equal(orc.unpack_bits(w, 70), bits)


# ---------------------------------------------------------------------------
# histogram: chunk walk vs definition
# ---------------------------------------------------------------------------

SHAPES = [(2, 2, 32), (3, 4, 33), (5, 3, 63), (4, 6, 64), (3, 5, 65), (6, 4, 100), (2, 7, 128), (5, 5, 129)]


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("p", [0.05, 0.5, 0.95])
def test_chunk_walk_equals_direct_enumeration(orc, shape, p):
    bits = rand_bits(shape, p, seed=hash((shape, p)) & 0xFFFF)
    X = shape[2]
    w = orc.pack_bits(bits)
    h_np = orc.bpfd_numpy(bits)
    np.testing.assert_array_equal(orc.bpfd_chunk(w, X), h_np)
    np.testing.assert_array_equal(orc.bpfd_direct_c(w, X), h_np)


@pytest.mark.parametrize("shape", [(2, 2, 2), (3, 3, 5), (4, 2, 17), (2, 9, 31)])
def test_direct_enumeration_small_x(orc, shape):
    bits = rand_bits(shape, 0.4, seed=7)
    np.testing.assert_array_equal(orc.bpfd_direct_c(orc.pack_bits(bits), shape[2]), orc.bpfd_numpy(bits))


def test_chunk_walk_rejects_small_x(orc):
    with pytest.raises(ValueError):
        orc.bpfd_chunk(orc.pack_bits(rand_bits((3, 3, 31), 0.5, 1)), 31)


def test_tail_garbage_is_ignored(orc):
    bits = rand_bits((4, 5, 70), 0.5, 3)
    w = orc.pack_bits(bits)
    dirty = w.copy()
    dirty[:, :, -1] |= np.uint32(0xFFFFFFC0)  # bits x >= 70 of the last word
    np.testing.assert_array_equal(orc.bpfd_chunk(dirty, 70), orc.bpfd_chunk(w, 70))


def test_cell_code_bit_order(orc):
    # Each of the 8 corners of one cell alone gives code 1 << i.
    # A 2 x 2 x 32 volume has one row of 31 cells; a voxel at x = 0 or
    # x = 31 belongs to exactly one of them.
    for i in range(8):
        bits = np.zeros((2, 2, 32), dtype=bool)
        bits[i >> 2, (i >> 1) & 1, 31 if i & 1 else 0] = True
        h = orc.bpfd_chunk(orc.pack_bits(bits), 32)
        assert h[1 << i] == 1 and h[0] == 30 and h.sum() == 31


# ---------------------------------------------------------------------------
# invariants
# ---------------------------------------------------------------------------

def test_total_and_uniform_volumes(orc):
    shape = (7, 9, 70)
    total = 6 * 8 * 69
    zero = orc.bpfd_chunk(orc.pack_bits(np.zeros(shape, bool)), 70)
    one = orc.bpfd_chunk(orc.pack_bits(np.ones(shape, bool)), 70)
    assert zero[0] == total and zero.sum() == total
    assert one[255] == total and one.sum() == total


def test_single_voxel_hits_each_corner_bin_once(orc):
    bits = np.zeros((5, 5, 40), dtype=bool)
    bits[2, 2, 20] = True
    h = orc.bpfd_chunk(orc.pack_bits(bits), 40)
    for i in range(8):
        assert h[1 << i] == 1
    assert h[0] == 4 * 4 * 39 - 8


def test_complement_reverses_codes(orc):
    bits = rand_bits((5, 6, 66), 0.3, 11)
    h = orc.bpfd_chunk(orc.pack_bits(bits), 66)
    hc = orc.bpfd_chunk(orc.pack_bits(~bits), 66)
    np.testing.assert_array_equal(hc, h[::-1])


def _permute_code_bits(perm):
    """Bin mapping for an axis permutation: new axis a = old axis perm[a]."""
    out = np.zeros(256, dtype=np.int64)
    for code in range(256):
        new = 0
        for i in range(8):
            if (code >> i) & 1:
                c = [i & 1, (i >> 1) & 1, i >> 2]  # old (x, y, z)
                d = [c[perm[0]], c[perm[1]], c[perm[2]]]
                new |= 1 << (d[0] | (d[1] << 1) | (d[2] << 2))
        out[code] = new
    return out


def test_axis_swap_permutes_bins(orc):
    bits = rand_bits((40, 33, 48), 0.5, 5)  # (Z, Y, X)
    h = orc.bpfd_chunk(orc.pack_bits(bits), 48)
    swapped = np.ascontiguousarray(np.transpose(bits, (2, 1, 0)))  # x <-> z
    hs = orc.bpfd_chunk(orc.pack_bits(swapped), 40)
    mapping = _permute_code_bits((2, 1, 0))
    expect = np.zeros(256, dtype=np.uint64)
    for code in range(256):
        expect[mapping[code]] += h[code]
    np.testing.assert_array_equal(hs, expect)


def test_translation_invariance(orc):
    obj = rand_bits((6, 6, 6), 0.5, 9)
    a = np.zeros((12, 12, 40), bool)
    b = np.zeros((12, 12, 40), bool)
    a[2:8, 3:9, 4:10] = obj
    b[5:11, 1:7, 30:36] = obj
    np.testing.assert_array_equal(orc.bpfd_chunk(orc.pack_bits(a), 40), orc.bpfd_chunk(orc.pack_bits(b), 40))


def test_whole_cpu_path_matches_numpy(orc):
    f = synth.uniform_field((9, 10, 77), seed=3)
    h, w = orc.bpfd_field(f, 0.5, 6)
    bits = orc.threshold_bits(f, 0.5, 6)
    np.testing.assert_array_equal(w, orc.pack_bits(bits))
    np.testing.assert_array_equal(h, orc.bpfd_numpy(bits))


# ---------------------------------------------------------------------------
# Euler weights, orbits, measures
# ---------------------------------------------------------------------------

def test_orbits_of_cube_group(orc):
    lab, n = orc.orbits()
    assert n == 22
    counts = np.bincount([bin(p).count("1") for p in range(256) if lab[p] == p])
    assert list(counts) == [1, 1, 3, 3, 6, 3, 3, 1, 1]


def test_euler_weights_are_orbit_invariant_and_balanced(orc):
    w = orc.euler_weights()
    lab, _ = orc.orbits()
    for p in range(256):
        assert w[p] == w[lab[p]]
    assert w[0] == 0 and w[255] == 0 and w[1] == 3


def _kv_raw(orc, bits):
    """oracle m[0] without normalisation (s = 0): (float)(k * pi / 6)."""
    h = orc.bpfd_numpy(bits)
    m, ovf = orc.measures(h, 0.0)
    assert not ovf
    return m[0]


def _pad(a, n=2):
    return np.pad(a, n)


SHAPE_CHI = {
    "voxel": (lambda: _pad(np.ones((1, 1, 1), bool)), 1),
    "box": (lambda: _pad(np.ones((4, 5, 6), bool)), 1),
    "two_boxes": (lambda: _pad(np.concatenate([np.ones((3, 3, 3), bool), np.zeros((3, 3, 2), bool),
                                                np.ones((3, 3, 3), bool)], axis=2)), 2),
    "corner_touch": (lambda: _pad(np.eye(2, dtype=bool)[:, :, None] * np.eye(2, dtype=bool)[None, :, :]), 2),
    "ring": (lambda: _pad(np.pad(np.zeros((1, 3, 3), bool), ((0, 0), (1, 1), (1, 1)), constant_values=True)), 0),
}


def _hollow_box():
    a = np.ones((7, 7, 7), bool)
    a[2:5, 2:5, 2:5] = False
    return _pad(a)


@pytest.mark.parametrize("name", list(SHAPE_CHI) + ["hollow_box", "two_rings"])
def test_euler_closed_forms(orc, name):
    if name == "hollow_box":
        bits, chi = _hollow_box(), 2
    elif name == "two_rings":
        r = np.pad(np.zeros((1, 3, 3), bool), ((0, 0), (1, 1), (1, 1)), constant_values=True)
        bits, chi = _pad(np.concatenate([r, np.zeros((1, 5, 5), bool), r], axis=0)), 0
    else:
        make, chi = SHAPE_CHI[name]
        bits = make()
    # k = sum(w * n) must equal 24 chi, so Kv = (float)(24 chi * pi / 6).
    assert _kv_raw(orc, bits) == np.float32((24 * chi) * math.pi / 6)


def test_volume_fraction_is_lower_corner_count(orc):
    bits = rand_bits((6, 7, 40), 0.37, 2)
    m, _ = orc.measures(orc.bpfd_numpy(bits), 1.0)
    lower = bits[:-1, :-1, :-1]
    assert m[3] == np.float32(lower.sum()) / np.float32(lower.size)
    assert m[1] == -1 and m[2] == -1


@pytest.mark.parametrize("X", [2, 3, 5, 8, 17, 31])
def test_gpu_walk_narrow_rows_closed_form(orc, X):
    """The reference GPU row walk restated (orc_bpfd_gpu_walk) on rows of X <
    32 equals its closed form from the register contents (twice the cells
    plus the wrap-around cell per row pair), emits 2X - 1 patterns per row
    pair, and differs from the cell histogram."""
    rng = np.random.default_rng(X)
    Z, Y = 7, 9
    bits = rng.random((Z, Y, X)) < 0.45
    words = orc.pbm_numpy(bits.astype(np.float64), 0.5, 6)  # bit = (v >= 0.5), zero tail bits
    walk = orc.bpfd_gpu_walk(words, X)
    np.testing.assert_array_equal(walk, orc.narrow_walk_numpy(words, X))
    assert int(walk.sum()) == (2 * X - 1) * (Y - 1) * (Z - 1)
    cells = orc.bpfd_numpy(bits)
    assert not np.array_equal(walk, cells)


@pytest.mark.parametrize("X", [32, 33, 63, 64, 65, 100, 129])
def test_gpu_walk_equals_cells_from_one_word(orc, X):
    """From X = 32 on, the GPU row walk counts exactly the row's cells."""
    rng = np.random.default_rng(100 + X)
    bits = rng.random((5, 6, X)) < 0.5
    words = orc.pbm_numpy(bits.astype(np.float64), 0.5, 6)
    np.testing.assert_array_equal(orc.bpfd_gpu_walk(words, X), orc.bpfd_numpy(bits))
