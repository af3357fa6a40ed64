"""Seeded synthetic scalar fields for tests and benchmarks.

Every generator is deterministic for a seed and returns a (Z, Y, X) tensor
(x fastest, the planar layout the kernels read). Fields are generated once;
the same bytes go to the GPU kernels and to the CPU oracle.

  uniform_field  i.i.d. U[0,1) float32 from a counter hash: (h >> 40) * 2^-24
  grf_field      Gaussian random field: N(0,1) white noise, periodic Gaussian
                 blur (exact Gaussian transfer function exp(-2 pi^2 s^2 k^2)
                 applied in Fourier space), standardised to zero mean and unit
                 variance in float64; float32 is the rounded float64 field.
                 On the GPU torch generates it; on the CPU numpy does (the
                 two give different fields for one seed)
  grf_field_direct  the same kind of field by direct periodic convolution with
                 a truncated Gaussian (radius ceil(4 sigma), normalised), one
                 axis at a time; used for the many odd-sized fields of the
                 batch mix, where an FFT plan per size would dominate
  sdf_field      signed distance to a union of spheres and axis-aligned boxes
"""

import math

import numpy as np
import torch

_MASK64 = (1 << 64) - 1


def _mix64(x):
    """splitmix64 finaliser on a uint64 numpy array."""
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def uniform_field(shape, seed=0):
    """float32 U[0,1) field of shape (Z, Y, X) from a counter-based hash."""
    Z, Y, X = (int(s) for s in shape)
    out = np.empty((Z, Y, X), dtype=np.float32)
    key = np.uint64((seed * 0x9E3779B97F4A7C15) & _MASK64)
    plane = np.arange(Y * X, dtype=np.uint64)
    with np.errstate(over="ignore"):
        for z in range(Z):
            c = plane + np.uint64(z * Y * X)
            h = _mix64(c ^ key)
            out[z] = ((h >> np.uint64(40)).astype(np.float32) * np.float32(2.0 ** -24)).reshape(Y, X)
    return out


def grf_field(shape, sigma, seed=0, device="cpu", dtype=torch.float64):
    """Gaussian random field (Z, Y, X) with periodic Gaussian blur sigma.

    Returns a torch tensor of `dtype` on `device`; float32 is obtained by
    rounding the float64 field."""
    Z, Y, X = (int(s) for s in shape)
    if str(device) == "cpu":
        return torch.from_numpy(_grf_numpy((Z, Y, X), sigma, seed)).to(dtype)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    f = torch.randn((Z, Y, X), generator=g, dtype=torch.float64, device=device)
    F = torch.fft.rfftn(f)
    del f
    c = -2.0 * math.pi ** 2 * float(sigma) ** 2
    kz = torch.fft.fftfreq(Z, device=device, dtype=torch.float64)
    ky = torch.fft.fftfreq(Y, device=device, dtype=torch.float64)
    kx = torch.fft.rfftfreq(X, device=device, dtype=torch.float64)
    F *= torch.exp(c * kz * kz)[:, None, None]
    F *= torch.exp(c * ky * ky)[None, :, None]
    F *= torch.exp(c * kx * kx)[None, None, :]
    f = torch.fft.irfftn(F, s=(Z, Y, X))
    del F
    mean = f.mean()
    f -= mean
    std = torch.sqrt((f * f).mean())
    f /= std
    if dtype != torch.float64:
        f = f.to(dtype)
    return f


def _grf_numpy(shape, sigma, seed):
    """CPU variant of grf_field (numpy's FFT; torch's CPU FFT is not used)."""
    Z, Y, X = shape
    rng = np.random.default_rng(int(seed))
    F = np.fft.rfftn(rng.standard_normal((Z, Y, X)))
    c = -2.0 * math.pi ** 2 * float(sigma) ** 2
    kz, ky, kx = np.fft.fftfreq(Z), np.fft.fftfreq(Y), np.fft.rfftfreq(X)
    F *= np.exp(c * kz * kz)[:, None, None]
    F *= np.exp(c * ky * ky)[None, :, None]
    F *= np.exp(c * kx * kx)[None, None, :]
    f = np.fft.irfftn(F, s=(Z, Y, X), axes=(0, 1, 2))
    f -= f.mean()
    f /= np.sqrt((f * f).mean())
    return np.ascontiguousarray(f)


def sdf_shapes(shape, seed=0, n_spheres=32, n_boxes=8, r_range=(8.0, 60.0), box_range=(10.0, 80.0)):
    """Seeded spheres (centre, radius) and boxes (lo, hi) inside a volume."""
    Z, Y, X = (int(s) for s in shape)
    rng = np.random.default_rng(seed)
    dims = np.array([X, Y, Z], dtype=np.float64)
    spheres = [(rng.uniform(0, dims), rng.uniform(*r_range)) for _ in range(n_spheres)]
    boxes = []
    for _ in range(n_boxes):
        side = rng.uniform(box_range[0], box_range[1], size=3)
        lo = rng.uniform(0, np.maximum(dims - side, 1.0))
        boxes.append((lo, lo + side))
    return spheres, boxes


def sdf_field(shape, seed=0, device="cpu", dtype=torch.float32, zchunk=32, **kw):
    """Signed distance (negative inside) to the union of seeded spheres and
    axis-aligned boxes, evaluated at voxel centres (x, y, z) = indices."""
    Z, Y, X = (int(s) for s in shape)
    spheres, boxes = sdf_shapes(shape, seed, **kw)
    out = torch.empty((Z, Y, X), dtype=dtype, device=device)
    x = torch.arange(X, device=device, dtype=torch.float64)[None, None, :]
    y = torch.arange(Y, device=device, dtype=torch.float64)[None, :, None]
    for z0 in range(0, Z, zchunk):
        z1 = min(Z, z0 + zchunk)
        z = torch.arange(z0, z1, device=device, dtype=torch.float64)[:, None, None]
        d = torch.full((z1 - z0, Y, X), float("inf"), dtype=torch.float64, device=device)
        for c, r in spheres:
            d = torch.minimum(d, torch.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) - r)
        for lo, hi in boxes:
            ctr = (lo + hi) / 2
            half = (hi - lo) / 2
            q = [torch.abs(x - ctr[0]) - half[0], torch.abs(y - ctr[1]) - half[1], torch.abs(z - ctr[2]) - half[2]]
            o = [torch.clamp(qa, min=0) for qa in q]
            outside = torch.sqrt(o[0] ** 2 + o[1] ** 2 + o[2] ** 2)
            inside = torch.clamp(torch.maximum(torch.maximum(q[0], q[1]), q[2]), max=0)
            d = torch.minimum(d, outside + inside)
        out[z0:z1] = d.to(dtype)
    return out


def batch_mix(seed=4, n=64, dim_range=(32, 384), sigma_range=(1.0, 6.0)):
    """The uneven batch: n fields, each axis dim uniform in dim_range
    (inclusive), half float32 and half float64 (a seeded permutation), GRF
    blur sigma uniform in sigma_range. Returns a list of
    (shape (Z, Y, X), torch dtype, sigma, field seed)."""
    rng = np.random.default_rng(seed)
    dims = rng.integers(dim_range[0], dim_range[1] + 1, size=(n, 3))
    sig = rng.uniform(sigma_range[0], sigma_range[1], size=n)
    f64 = np.zeros(n, dtype=bool)
    f64[rng.permutation(n)[: n // 2]] = True
    return [(tuple(int(v) for v in dims[i]), torch.float64 if f64[i] else torch.float32, float(sig[i]),
             int(seed) * 1000 + i) for i in range(n)]


def grf_field_direct(shape, sigma, seed=0, device="cpu", dtype=torch.float64):
    """Gaussian random field (Z, Y, X): seeded N(0,1) noise blurred along each
    axis by periodic convolution with exp(-d^2 / (2 sigma^2)), |d| <= ceil(4
    sigma), normalised to sum 1; then standardised in float64 (float32 is the
    rounded float64 field)."""
    Z, Y, X = (int(s) for s in shape)
    g = torch.Generator(device=device)
    g.manual_seed(int(seed))
    f = torch.randn((Z, Y, X), generator=g, dtype=torch.float64, device=device)
    R = int(math.ceil(4.0 * float(sigma)))
    d = torch.arange(-R, R + 1, dtype=torch.float64)
    k = torch.exp(-0.5 * (d / float(sigma)) ** 2)
    k = (k / k.sum()).tolist()
    for axis in range(3):
        out = torch.zeros_like(f)
        for i, w in enumerate(k):
            out.add_(torch.roll(f, shifts=i - R, dims=axis), alpha=w)
        f = out
    f -= f.mean()
    f /= torch.sqrt((f * f).mean())
    return f if dtype == torch.float64 else f.to(dtype)


def make_fields(specs, device):
    """Generates the fields of batch_mix specs (grf_field_direct)."""
    return [grf_field_direct(shape, sigma, s, device=device, dtype=dtype) for shape, dtype, sigma, s in specs]
