"""The measurement aid mkfReadRoof reads every 16-byte word once: its XOR
checksum equals numpy's over the same bytes, for both load kinds, several
grid sizes and lengths that do not fill the last grid stride; bad arguments
are refused before any launch."""

import ctypes as C

import numpy as np
import pytest
import torch

import mkf_amd as M

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("words", [1, 255, 4097, 1000003, 4 << 20])
def test_read_roof_checksum(lib, cuda, words):
    rng = np.random.default_rng(words)
    host = rng.integers(0, 2 ** 32, size=4 * words, dtype=np.uint64).astype(np.uint32)
    want = int(np.bitwise_xor.reduce(host)) if host.size else 0
    buf = torch.from_numpy(host.view(np.int32)).to(cuda)
    for nt in (0, 1, 2, 3):  # load kind x layout flags
        for u in (4, 8, 16):
            for bpc in (1, 4, 64):
                out = torch.zeros(1, dtype=torch.int32, device=cuda)
                M.check(lib.mkfReadRoof(C.c_void_p(buf.data_ptr()), 16 * words, nt, u, bpc,
                                        C.c_void_p(out.data_ptr()), None))
                torch.cuda.synchronize()
                assert int(out.cpu().numpy().view(np.uint32)[0]) == want, (nt, u, bpc)


def test_read_roof_rejects_bad_arguments(lib, cuda):
    buf = torch.zeros(64, dtype=torch.int32, device=cuda)
    out = torch.zeros(1, dtype=torch.int32, device=cuda)
    p, o = buf.data_ptr(), C.c_void_p(out.data_ptr())
    assert lib.mkfReadRoof(C.c_void_p(p + 4), 128, 1, 8, 4, o, None) == M.MKF_ERR_ARG  # unaligned
    assert lib.mkfReadRoof(C.c_void_p(p), 100, 1, 8, 4, o, None) == M.MKF_ERR_ARG     # not a multiple of 16
    assert lib.mkfReadRoof(C.c_void_p(p), 128, 1, 8, 0, o, None) == M.MKF_ERR_ARG     # no workgroups
    assert lib.mkfReadRoof(C.c_void_p(p), 128, 1, 5, 4, o, None) == M.MKF_ERR_ARG     # loads in flight
    assert lib.mkfReadRoof(C.c_void_p(p), 128, 4, 8, 4, o, None) == M.MKF_ERR_ARG     # unknown flag
    assert lib.mkfReadRoof(None, 128, 1, 8, 4, o, None) == M.MKF_ERR_ARG
