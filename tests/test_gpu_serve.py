"""Serving form: continuous batching under a memory budget. Each result is
the same, bit for bit, whether its request is served alone or in a mix of
concurrently arriving requests, and equals the oracle's histograms."""

import threading
import time

import numpy as np
import pytest
import torch

import mkf_amd as M
from mkf_amd import synth
from mkf_amd.serve import Server

pytestmark = pytest.mark.gpu


def oracle(orc, field, levels):
    f = field.cpu().numpy()
    out = []
    for lv in levels:
        bm = M.bin_map(*lv)
        out.append(orc.bpfd_field(f, bm.t[0], bm.m)[0])
    return np.stack(out)


def pinned(t):
    p = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    p.copy_(t.cpu())
    return p


def wait_all(S, n, timeout=60.0):
    got = {}
    t0 = time.time()
    while len(got) < n:
        for r in S.poll(wait_us=20000):
            got[r["tag"]] = r
        assert time.time() - t0 < timeout, f"{len(got)} of {n} results"
    return got


def mix_specs(seed, n):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        shape = tuple(int(v) for v in rng.integers(8, 90, size=3))
        dtype = torch.float64 if rng.random() < 0.3 else torch.float32
        nl = int(rng.integers(1, 17))
        rels = [">=", "<", "<=", ">", "=", "!="]
        levels = [(rels[int(rng.integers(0, len(rels)))], float(t)) for t in rng.uniform(-1.5, 1.5, nl)]
        specs.append((shape, dtype, levels, bool(rng.random() < 0.3), 1.0 + 3 * rng.random()))
    return specs


def test_serve_alone_equals_mix(orc, cuda):
    S = Server(budget=256 << 20, max_in_flight=16, max_batch=8)
    try:
        probe = pinned(synth.grf_field_direct((41, 53, 77), 2.0, seed=11, device=cuda, dtype=torch.float32))
        levels = [(">=", t) for t in (-1.0, -0.3, 0.0, 0.4, 1.2)]
        S.submit(probe, levels, tag=1)
        alone = wait_all(S, 1)[1]
        assert alone["status"] == 0 and alone["batch"] == 1
        ref = oracle(orc, probe, levels)
        np.testing.assert_array_equal(alone["hist"], ref)

        specs = mix_specs(3, 40)
        fields, expected = {}, {}
        for i, (shape, dtype, lv, on_dev, sig) in enumerate(specs):
            f = synth.grf_field_direct(shape, sig, seed=100 + i, device=cuda, dtype=dtype)
            fields[100 + i] = (f if on_dev else pinned(f), lv)
        for i, (tag, (f, lv)) in enumerate(fields.items()):
            S.submit(f, lv, tag=tag)
            if i == 20:
                S.submit(probe, levels, tag=2)
        got = wait_all(S, len(fields) + 1)
        mixed = got[2]
        assert mixed["status"] == 0
        np.testing.assert_array_equal(mixed["hist"], alone["hist"])
        assert max(r["batch"] for r in got.values()) > 1, "no request shared a launch"
        for tag, (f, lv) in fields.items():
            assert got[tag]["status"] == 0
            np.testing.assert_array_equal(got[tag]["hist"], oracle(orc, f, lv), err_msg=f"request {tag}")
        st = S.stats()
        assert st["submitted"] == st["completed"] == len(fields) + 2
        for r in got.values():
            assert r["submit"] <= r["admit"] <= r["launch"] <= r["done"]
    finally:
        S.close()


def test_serve_budget_admission(orc, cuda):
    big = synth.grf_field_direct((60, 70, 80), 2.0, seed=21, device=cuda, dtype=torch.float64)
    nbytes = big.numel() * 8
    S = Server(budget=2 * nbytes + 4096, max_in_flight=64, max_batch=4)
    try:
        host = pinned(big)
        lv = [(">=", 0.0), ("<", 0.5)]
        for tag in range(10):
            S.submit(host, lv, tag=tag)
        got = wait_all(S, 10)
        ref = oracle(orc, big, lv)
        for r in got.values():
            assert r["status"] == 0
            np.testing.assert_array_equal(r["hist"], ref)
        st = S.stats()
        assert 0 < st["heldMax"] <= 2 * nbytes + 4096 + 512
        assert st["bytesIn"] == 10 * nbytes
        # a host field larger than the budget is served in z-slabs of at most
        # a quarter of the budget, reported once with the summed histograms
        too_big = pinned(synth.grf_field_direct((100, 100, 100), 2.0, seed=22, device=cuda, dtype=torch.float64))
        S.submit(too_big, lv, tag=98)
        r = wait_all(S, 1)[98]
        assert r["status"] == 0
        np.testing.assert_array_equal(r["hist"], oracle(orc, too_big, lv))
        assert S.stats()["heldMax"] <= 2 * nbytes + 4096 + 512
        # only a field of which two planes exceed the budget is refused
        wide = pinned(torch.zeros((3, 600, 600), dtype=torch.float64))
        with pytest.raises(RuntimeError, match="budget"):
            S.submit(wide, lv)
        # a device-resident field needs no budget
        S.submit(too_big.to(cuda), lv, tag=99)
        r = wait_all(S, 1)[99]
        np.testing.assert_array_equal(r["hist"], oracle(orc, too_big, lv))
    finally:
        S.close()


def test_serve_split_requests_in_a_mix(orc, cuda):
    """Oversized host requests (split into slabs) interleaved with ordinary
    host and device requests: every result equals the oracle."""
    S = Server(budget=3 << 20, max_in_flight=24, max_batch=6)
    try:
        fields, want = {}, {}
        specs = [((40, 50, 60), torch.float32, 1.5), ((90, 80, 70), torch.float64, 2.5),
                 ((20, 30, 40), torch.float64, 1.0), ((130, 60, 110), torch.float32, 3.0),
                 ((64, 64, 64), torch.float32, 2.0), ((300, 40, 33), torch.float64, 1.5)]
        lv = [(">=", -0.5), ("<", 0.2), (">", 1.0)]
        for tag, (shape, dt, sigma) in enumerate(specs):
            f = synth.grf_field_direct(shape, sigma, seed=200 + tag, device=cuda, dtype=dt)
            fields[tag] = pinned(f) if tag % 3 != 2 else f
            want[tag] = oracle(orc, f, lv)
        for tag, f in fields.items():
            S.submit(f, lv, tag=tag)
        got = wait_all(S, len(fields))
        for tag in fields:
            assert got[tag]["status"] == 0, tag
            np.testing.assert_array_equal(got[tag]["hist"], want[tag])
        st = S.stats()
        assert st["submitted"] == len(fields) and st["completed"] == len(fields)
        assert st["heldMax"] <= (3 << 20) + 512
    finally:
        S.close()


def test_serve_device_snapshots_and_threads(orc, cuda):
    """Snapshots of an evolving device field, f(k+1) = f(k) + eps * noise, and
    host requests submitted from two threads at once."""
    S = Server(budget=128 << 20, max_in_flight=32, max_batch=16)
    try:
        g = torch.Generator(device=cuda)
        g.manual_seed(5)
        f = synth.grf_field_direct((30, 40, 50), 2.0, seed=31, device=cuda, dtype=torch.float32)
        snaps = []
        for _ in range(6):
            f = f + 0.05 * torch.randn(f.shape, generator=g, device=cuda, dtype=f.dtype)
            snaps.append(f)
        hosts = [pinned(synth.grf_field_direct((20 + 7 * i, 25, 33), 1.5, seed=40 + i, device=cuda,
                                               dtype=torch.float64)) for i in range(8)]
        lv = [(">=", -0.5), (">=", 0.0), (">=", 0.5)]
        torch.cuda.synchronize()
        errors = []

        def submit_hosts(lo, hi):
            try:
                for i in range(lo, hi):
                    S.submit(hosts[i], lv, tag=100 + i)
            except Exception as e:  # surfaced below
                errors.append(e)

        th = [threading.Thread(target=submit_hosts, args=(0, 4)), threading.Thread(target=submit_hosts, args=(4, 8))]
        for t in th:
            t.start()
        for k, s in enumerate(snaps):
            S.submit(s, lv, tag=k)
        for t in th:
            t.join()
        assert not errors
        got = wait_all(S, 14)
        for k, s in enumerate(snaps):
            np.testing.assert_array_equal(got[k]["hist"], oracle(orc, s, lv), err_msg=f"snapshot {k}")
        for i, h in enumerate(hosts):
            np.testing.assert_array_equal(got[100 + i]["hist"], oracle(orc, h, lv), err_msg=f"host {i}")
    finally:
        S.close()


def test_serve_repeated_tags(orc, cuda):
    """A caller's tag may repeat while requests with it are outstanding:
    each request keeps its own field, maps and output."""
    S = Server(budget=64 << 20, max_in_flight=16, max_batch=8)
    try:
        fs = [pinned(synth.grf_field_direct((20 + 3 * i, 30, 40), 1.5, seed=60 + i, device=cuda,
                                            dtype=torch.float32)) for i in range(3)]
        for i, f in enumerate(fs):
            S.submit(f, [(">=", 0.1 * k) for k in range(i + 1)], tag=7)  # i + 1 levels identify request i
        res = []
        t0 = time.time()
        while len(res) < 3:
            res += S.poll(wait_us=20000)
            assert time.time() - t0 < 60
        assert S.outstanding == 0
        for r in res:
            assert r["tag"] == 7 and r["status"] == 0
            i = r["hist"].shape[0] - 1
            np.testing.assert_array_equal(r["hist"], oracle(orc, fs[i], [(">=", 0.1 * k) for k in range(i + 1)]))
    finally:
        S.close()


def test_serve_split_when_two_planes_exceed_a_quarter(orc, cuda):
    """Two planes larger than a quarter of the budget but within it: the
    field is split into slabs of as many planes as the budget holds (here 2),
    not refused."""
    budget = 2_000_000
    f = synth.grf_field_direct((12, 300, 300), 2.0, seed=71, device=cuda, dtype=torch.float64)
    plane = 300 * 300 * 8
    assert budget / 4 < 2 * plane <= budget < f.numel() * 8
    S = Server(budget=budget, max_in_flight=8, max_batch=4)
    try:
        lv = [(">=", 0.0), ("<", -0.4)]
        S.submit(pinned(f), lv, tag=5)
        r = wait_all(S, 1)[5]
        assert r["status"] == 0
        np.testing.assert_array_equal(r["hist"], oracle(orc, f, lv))
        assert S.stats()["heldMax"] <= budget + 256
    finally:
        S.close()
