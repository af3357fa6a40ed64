"""The seeded workload generators: determinism and the statistics that define
them (CPU only)."""

import math

import numpy as np
import torch

from mkf_amd import synth


def test_grf_direct_statistics_and_determinism():
    sigma = 2.5
    a = synth.grf_field_direct((24, 30, 40), sigma, seed=7)
    b = synth.grf_field_direct((24, 30, 40), sigma, seed=7)
    c = synth.grf_field_direct((24, 30, 40), sigma, seed=8)
    assert torch.equal(a, b) and not torch.equal(a, c)
    x = a.numpy()
    assert abs(x.mean()) < 1e-12 and abs(x.std() - 1.0) < 1e-9
    # A Gaussian blur of white noise with width sigma has lag-1 correlation
    # exp(-1 / (4 sigma^2)) along every axis (kernel truncation aside).
    want = math.exp(-1.0 / (4 * sigma * sigma))
    for ax in range(3):
        lo = np.take(x, range(x.shape[ax] - 1), axis=ax).ravel()
        hi = np.take(x, range(1, x.shape[ax]), axis=ax).ravel()
        assert abs(np.corrcoef(lo, hi)[0, 1] - want) < 0.02, ax
    f32 = synth.grf_field_direct((24, 30, 40), sigma, seed=7, dtype=torch.float32)
    assert torch.equal(f32, a.to(torch.float32))


def test_grf_direct_is_periodic():
    # Rolling the noise and the field together: the blur commutes with shifts.
    f = synth.grf_field_direct((16, 12, 20), 1.5, seed=3)
    assert f.shape == (16, 12, 20)
    assert np.isfinite(f.numpy()).all()


def test_batch_mix_definition():
    mix = synth.batch_mix(seed=4)
    assert len(mix) == 64
    assert sum(1 for m in mix if m[1] == torch.float64) == 32
    dims = np.array([m[0] for m in mix])
    assert dims.min() >= 32 and dims.max() <= 384
    assert all(1.0 <= m[2] <= 6.0 for m in mix)
    assert len({m[3] for m in mix}) == 64
    assert synth.batch_mix(seed=4) == mix and synth.batch_mix(seed=5) != mix


def test_request_stream_definition():
    from mkf_amd.serve import poisson_arrivals, request_stream
    pool, sims, reqs = request_stream(2000, seed=5)
    assert len(pool) == 24 and len(sims) == 4 and len(reqs) == 2000
    sides = np.array([s for s, _ in pool + sims])
    assert sides.min() >= 64 and sides.max() <= 768
    dev = np.mean([r["sim"] is not None for r in reqs])
    assert abs(dev - 0.3) < 0.04
    nl = np.array([len(r["levels"]) for r in reqs])
    assert nl.min() == 1 and nl.max() == 16
    for r in reqs:
        ts = [t for _, t in r["levels"]]
        assert ts == sorted(ts) and -1.5 <= ts[0] and ts[-1] <= 1.5
    assert request_stream(50, seed=5)[2] == request_stream(50, seed=5)[2]
    arr = poisson_arrivals(20000, 500.0, seed=1)
    assert np.all(np.diff(arr) > 0)
    assert abs(arr[-1] / 20000 - 1 / 500.0) < 0.05 / 500.0
