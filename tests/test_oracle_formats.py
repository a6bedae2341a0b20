"""Pins of the 2015-format reorder restatement (oracle.reorder_2015, the numpy
form of software/prototyping/cuda/reorder.py:284-293).

The reference's own check (reorder.py:291-293) compares that numpy form with
the reorderTz_smem kernel (kernel_template.py:107-164).  Here the kernel's
index arithmetic is written out as an independent scatter: for every input
(channel, frame, snapshot) it computes the output snapshot and frame the way
the kernel does (kernel_template.py:120-134) and writes there.  A mistake in
the roll direction, the shifted channel set (a-d of every 8), the split
snapshot (69) or the dropped frame fails the comparison; spot values worked by
hand pin the two forms against each other's common mistakes.
"""
import numpy as np

from oracle import sdbe_oracle as O

NCH, NSN = 16384, 128


def _scatter_form(x):
    nch, nf, nsn = x.shape
    out = np.zeros((nf - 1, nsn, nch + 1), dtype=x.dtype)
    cid = np.arange(nch)[:, None]
    sid_in = np.arange(nsn)[None, :]
    shifted = ((cid // 4) & 1) == 0
    sid_out = np.where(shifted, (sid_in - 2) & 0x7F, sid_in)
    for bid_out in range(nf - 1):
        bid_in = np.where(sid_out < 69, bid_out, bid_out + 1)
        vals = x[np.broadcast_to(cid, sid_out.shape), bid_in, np.broadcast_to(sid_in, sid_out.shape)]
        out[bid_out, sid_out, np.broadcast_to(cid, sid_out.shape)] = vals
    return out


def _input(nf, seed):
    rng = np.random.default_rng(seed)
    # distinct values: real part encodes (channel, frame, snapshot)
    c, f, s = np.meshgrid(np.arange(NCH), np.arange(nf), np.arange(NSN), indexing="ij")
    return (c * 1000 + f * 128 + s).astype(np.float32) + 1j * rng.normal(size=(NCH, nf, NSN)).astype(np.float32)


def test_numpy_form_matches_kernel_index_form():
    x = _input(3, 1).astype(np.complex64)
    a = O.reorder_2015(x)
    b = _scatter_form(x)
    assert a.shape == (2, NSN, NCH + 1)
    assert np.array_equal(a, b)
    assert np.all(a[:, :, NCH] == 0)


def test_spot_values():
    x = _input(3, 2).astype(np.complex64)
    out = O.reorder_2015(x)
    # channel 0 (a of the first 8) lags: snapshot s comes from s + 2 of the same frame ...
    assert out[0, 0, 0] == x[0, 0, 2]
    assert out[0, 66, 0] == x[0, 0, 68]
    # ... and from the next frame from snapshot 69 on, wrapping at 128
    assert out[0, 69, 0] == x[0, 1, 71]
    assert out[0, 126, 0] == x[0, 1, 0]
    assert out[1, 127, 3] == x[3, 2, 1]
    # channel 4 (e) is not shifted but also switches frame at 69
    assert out[0, 68, 4] == x[4, 0, 68]
    assert out[0, 69, 4] == x[4, 1, 69]
    # channel 8 is an a-channel again, 12 an e-channel
    assert out[1, 5, 8] == x[8, 1, 7]
    assert out[1, 5, 12] == x[12, 1, 5]
    # every input value of frames 0..n-2 (snapshots < 69 after the shift) and
    # 1..n-1 (snapshots >= 69) appears exactly once
    assert np.unique(out[:, :, :NCH].real).size == 2 * NSN * NCH
