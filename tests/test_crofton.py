"""mkfCroftonMeasuresBPFD: Crofton estimates of Kv, Mv, Sv, Vv from a BPFD.

No reference counterpart exists (the reference's book measures need
refMeasures.c, absent from its tree, src/mkfUtil.c:15-22). The estimator is
pinned three ways:
* against a brute-force evaluation of the same Crofton formulas on the voxel
  set (oracle/crofton.py: global pair counts, section Euler numbers from
  connected components, direction weights from scipy's spherical Voronoi);
* against closed forms on digitised solids (ball, hollow ball, torus);
* by the symmetries the per-pattern weights must have.
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

import mkf_amd as M
from oracle import crofton as CR
from oracle import oracle as O


def _grid(n):
    c = np.array([n / 2 + 0.3, n / 2 - 0.2, n / 2 + 0.1])
    z, y, x = np.mgrid[0:n, 0:n, 0:n].astype(np.float64)
    return x - c[0], y - c[1], z - c[2]


def _measures_of(field):
    """Set = field <= 0; returns (library totals, voxel bits)."""
    h, _ = O.bpfd_field(field, 0.0, 3)
    return M.crofton_measures(h, 0.0), field <= 0


def _blobs(shape, seed, sigma=2.0):
    from scipy import ndimage

    f = ndimage.gaussian_filter(np.random.default_rng(seed).standard_normal(shape), sigma)
    f[:2] = f[-2:] = f[:, :2] = f[:, -2:] = f[:, :, :2] = f[:, :, -2:] = 1.0
    return f


@pytest.fixture(scope="module")
def weights():
    """Per-pattern weights (K, M, S, V) from one-hot histograms."""
    eye = np.eye(256, dtype=np.uint64)
    return np.array([M.crofton_measures(eye[p], 0.0) for p in range(256)])


def test_direction_weights_from_geometry():
    w = CR.direction_weights()
    assert 6 * w[1] + 12 * w[2] + 8 * w[3] == pytest.approx(1.0, abs=1e-12)
    # independent quadrature: nearest of the 26 directions over a Fibonacci sphere
    n = 400_000
    i = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * i / n)
    th = math.pi * (1 + 5 ** 0.5) * i
    p = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    d = np.array([v for v in itertools.product((-1, 0, 1), repeat=3) if any(v)], dtype=float)
    best = np.argmax(p @ (d / np.linalg.norm(d, axis=1)[:, None]).T, axis=1)
    frac = np.bincount(best, minlength=26) / n
    k = np.abs(d).sum(1)
    for c in (1, 2, 3):
        assert frac[k == c].mean() == pytest.approx(w[c], abs=2e-5)


@pytest.mark.parametrize("case", ["ball", "blobs_cube", "blobs_box", "blobs_thin", "torus"])
def test_matches_brute_force(case):
    if case == "ball":
        x, y, z = _grid(26)
        f = np.sqrt(x * x + y * y + z * z) - 9
    elif case == "torus":
        x, y, z = _grid(34)
        f = np.sqrt((np.sqrt(x * x + y * y) - 10) ** 2 + z * z) - 4
    elif case == "blobs_cube":
        f = _blobs((36, 36, 36), 3)
    elif case == "blobs_box":
        f = _blobs((20, 31, 45), 11, 1.5)
    else:
        f = _blobs((6, 40, 33), 5, 1.0)
    m, b = _measures_of(f)
    assert m[0] == pytest.approx(4 * math.pi * CR.euler_6(b), rel=1e-12, abs=1e-9)
    assert m[1] == pytest.approx(CR.mean_curvature(b), rel=1e-10, abs=1e-9)
    assert m[2] == pytest.approx(CR.surface_area(b), rel=1e-10)
    # V: cells whose lowest corner is set = set voxels off the far faces
    assert m[3] == np.count_nonzero(b[:-1, :-1, :-1])


def test_closed_forms():
    x, y, z = _grid(52)
    r = np.sqrt(x * x + y * y + z * z)
    R = 20
    m, _ = _measures_of(r - R)
    assert m[0] == pytest.approx(4 * math.pi, abs=1e-9)  # chi = 1
    assert m[1] == pytest.approx(4 * math.pi * R, rel=2e-3)
    assert m[2] == pytest.approx(4 * math.pi * R * R, rel=2e-3)
    assert m[3] == pytest.approx(4 / 3 * math.pi * R ** 3, rel=2e-3)
    # hollow ball: chi = 2, M = 4 pi (R1 - R2), S = 4 pi (R1^2 + R2^2)
    m, _ = _measures_of(np.maximum(r - 20, 12 - r))
    assert m[0] == pytest.approx(8 * math.pi, abs=1e-9)
    assert m[1] == pytest.approx(4 * math.pi * 8, rel=2e-3)
    assert m[2] == pytest.approx(4 * math.pi * (400 + 144), rel=2e-3)
    # torus (R 16, r 6): chi = 0, M = 2 pi^2 R, S = 4 pi^2 R r
    x, y, z = _grid(48)
    m, _ = _measures_of(np.sqrt((np.sqrt(x * x + y * y) - 16) ** 2 + z * z) - 6)
    assert m[0] == pytest.approx(0.0, abs=1e-9)
    assert m[1] == pytest.approx(2 * math.pi ** 2 * 16, rel=3e-2)
    assert m[2] == pytest.approx(4 * math.pi ** 2 * 16 * 6, rel=1e-2)


def test_lattice_aligned_plane_bias():
    """A lattice-aligned plane is seen with the Crofton factor
    2 sum_v c_v |n.u_v| (not 1): every line along v crosses it once."""
    w = CR.direction_weights()
    dirs = [v for v in itertools.product((-1, 0, 1), repeat=3) if any(v)]
    factor = sum(w[sum(map(abs, v))] * abs(v[2]) / math.sqrt(sum(map(abs, v))) for v in dirs) * 2
    X, Y, Z = 40, 30, 20
    f = np.ones((Z, Y, X))
    f[:8] = -1  # set below z = 7.5: (X-1)(Y-1) cells straddle the plane
    h, _ = O.bpfd_field(f, 0.0, 3)
    assert h[0x0F] == (X - 1) * (Y - 1)
    m = M.crofton_measures(h, 0.0)
    assert m[2] == pytest.approx(factor * (X - 1) * (Y - 1), rel=1e-12)


def _cube_group():
    """The 48 symmetries of the cube as permutations of the 8 vertex bits."""
    perms = []
    for ax in itertools.permutations(range(3)):
        for flip in itertools.product((0, 1), repeat=3):
            p = []
            for v in range(8):
                c = [(v >> i) & 1 for i in range(3)]
                d = [c[ax[i]] ^ flip[i] for i in range(3)]
                p.append(d[0] | d[1] << 1 | d[2] << 2)
            perms.append(p)
    return perms


def test_pattern_weight_symmetries(weights):
    for perm in _cube_group():
        img = [sum(((p >> v) & 1) << perm[v] for v in range(8)) for p in range(256)]
        np.testing.assert_allclose(weights[img, :3], weights[:, :3], atol=1e-13)
    comp = 255 - np.arange(256)
    np.testing.assert_allclose(weights[comp, 2], weights[:, 2], atol=1e-13)   # S symmetric
    np.testing.assert_allclose(weights[comp, 1], -weights[:, 1], atol=1e-13)  # M antisymmetric
    assert np.all(weights[[0, 255], :3] == 0)
    # K: 4 pi x the textbook Euler share, the oracle's independent table / 24
    np.testing.assert_allclose(weights[:, 0], O.euler_weights() * math.pi / 6, atol=1e-12)
    assert np.all(weights[:, 3] == (np.arange(256) & 1))


def test_density_scaling_and_errors():
    f = _blobs((24, 24, 24), 9)
    h, _ = O.bpfd_field(f, 0.0, 3)
    tot = M.crofton_measures(h, 0.0)
    t = float(h.sum())
    s = 0.37
    d = M.crofton_measures(h, s)
    vol = t * s ** 3
    np.testing.assert_allclose(d, [tot[0] / vol, tot[1] * s / vol, tot[2] * s * s / vol, tot[3] / t], rtol=1e-13)
    lib = M.load()
    m = (C.c_double * 4)()
    hz = np.zeros(256, dtype=np.uint64)
    P = C.POINTER(C.c_size_t)
    assert lib.mkfCroftonMeasuresBPFD(m, h.astype(np.uint64).ctypes.data_as(P), -1.0) == M.MKF_ERR_ARG
    assert lib.mkfCroftonMeasuresBPFD(m, hz.ctypes.data_as(P), 1.0) == M.MKF_ERR_ARG
    assert lib.mkfCroftonMeasuresBPFD(m, hz.ctypes.data_as(P), 0.0) == 4
    assert list(m) == [0.0, 0.0, 0.0, 0.0]
