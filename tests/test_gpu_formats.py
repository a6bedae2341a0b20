"""GPU parity of the 2015-format reorder (sdbe_reorder_2015, restating
reorderTz_smem, software/prototyping/cuda/kernel_template.py:107-164) against
the oracle's numpy form (reorder.py:284-293).  Pure data movement: bit-exact.
Edge cases: the minimum of two input frames, an odd frame count, rejected
calls (one frame, overlapping buffers).
"""
import sys

import numpy as np
import pytest

from oracle import sdbe_oracle as O

pytestmark = pytest.mark.gpu

NCH, NSN = 16384, 128


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.mark.parametrize("nf", [2, 5])
def test_reorder_2015_bit_exact(torch_cuda, nf):
    torch = torch_cuda
    from sdbe_amd.formats import reorder_2015
    rng = np.random.default_rng(100 + nf)
    x = (rng.normal(size=(NCH, nf, NSN)) + 1j * rng.normal(size=(NCH, nf, NSN))).astype(np.complex64)
    _say("[fmt] reorder_2015 n_frames=%d" % nf)
    got = reorder_2015(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    want = O.reorder_2015(x)
    assert got.shape == want.shape == (nf - 1, NSN, NCH + 1)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_reorder_2015_rejects(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.formats import reorder_2015
    x = torch.zeros((NCH, 1, NSN), dtype=torch.complex64, device="cuda")
    with pytest.raises(L.SdbeError):
        reorder_2015(x)
    lib = L.load()
    y = torch.zeros((NCH, 3, NSN), dtype=torch.complex64, device="cuda")
    import ctypes as C
    # output overlapping the input
    assert lib.sdbe_reorder_2015(C.c_void_p(y.data_ptr()), 3, C.c_void_p(y.data_ptr() + 64), None) == L.SDBE_ERR_ARG
    assert lib.sdbe_reorder_2015(None, 3, C.c_void_p(y.data_ptr()), None) == L.SDBE_ERR_ARG
