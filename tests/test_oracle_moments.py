"""The oracle's phase moments (categorical sums of src/binMapCUDA.cuh:89-121)
against closed forms and brute-force loops."""
import math

import numpy as np
import pytest

import mkf_amd as M
from oracle import oracle as O


def test_integer_ramp_closed_form():
    """Field i = 0..N-1 (x fastest) split at threshold k by '>=': phase 1
    holds k..N-1, phase 0 holds 0..k-1."""
    Z, Y, X = 5, 6, 7
    N = X * Y * Z
    f = np.arange(N, dtype=np.float64).reshape(Z, Y, X)
    k = 77
    bm = M.bin_map(">=", k)
    m = O.phase_moments(f, bm.t[0], bm.m)

    def s1(n):  # sum 0..n-1
        return n * (n - 1) // 2

    def s2(n):  # sum of squares 0..n-1
        return (n - 1) * n * (2 * n - 1) // 6

    assert list(m) == [s1(k), s2(k), s1(N) - s1(k), s2(N) - s2(k), k, N - k]


def test_constant_field_and_inverted_relation():
    f = np.full((3, 4, 33), 2.5, dtype=np.float32)
    bm = M.bin_map(">=", 1.0)
    n = f.size
    assert list(O.phase_moments(f, bm.t[0], bm.m)) == [0, 0, 2.5 * n, 6.25 * n, 0, n]
    bm = M.bin_map("!>=", 1.0)
    assert list(O.phase_moments(f, bm.t[0], bm.m)) == [2.5 * n, 6.25 * n, 0, 0, n, 0]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_brute_force_loops(dtype):
    rng = np.random.default_rng(3)
    f = rng.standard_normal((3, 4, 5)).astype(dtype)
    f[0, 0, 0] = 0.25  # equal to the threshold: in phase 1 for '<='
    bm = M.bin_map("<=", 0.25)
    acc = [[], [], [], []]
    for z in range(3):
        for y in range(4):
            for x in range(5):
                v = float(f[z, y, x])
                ph = 1 if v <= float(dtype(0.25)) else 0
                acc[2 * ph].append(v)
                acc[2 * ph + 1].append(v * v)
    want = [math.fsum(a) for a in acc] + [len(acc[0]), len(acc[2])]
    assert list(O.phase_moments(f, bm.t[0], bm.m)) == want


def test_nan_goes_to_eq_phase():
    """NaN compares as EQ (src/binMapCUDA.cuh:131-147): with '>=' it is in
    phase 1, whose sums become NaN; phase 0 stays finite."""
    f = np.array([[[1.0, np.nan, -1.0, 3.0]] * 2] * 2)
    bm = M.bin_map(">=", 0.0)
    m = O.phase_moments(f, bm.t[0], bm.m)
    assert m[0] == -4.0 and m[1] == 4.0 and math.isnan(m[2]) and math.isnan(m[3])
    assert m[4] == 4 and m[5] == 12


def test_infinities_in_a_phase():
    """IEEE addition in any order: infinities of one sign give that infinity,
    both signs give NaN; squares of either are +inf."""
    bm = M.bin_map(">=", 0.0)
    one = np.array([[[np.inf, 1.0, 2.0, -5.0]]])
    m = O.phase_moments(one, bm.t[0], bm.m)
    assert m[0] == -5.0 and m[1] == 25.0 and m[2] == math.inf and m[3] == math.inf
    ne = M.bin_map("!=", 0.0)
    both = np.array([[[np.inf, 1.0, -np.inf, 0.0]]])
    m = O.phase_moments(both, ne.t[0], ne.m)
    assert m[0] == 0.0 and m[1] == 0.0 and m[4] == 1
    assert math.isnan(m[2]) and m[3] == math.inf and m[5] == 3   # inf, 1 and -inf in phase 1
    with np.errstate(invalid="ignore"):
        assert np.isnan(np.float64(np.inf) + np.float64(-np.inf))
