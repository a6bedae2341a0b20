"""GPU tests of the ragged multi-stream batch (sdbe_process_blocks) against the
FP64 oracle: streams of uneven length, a spectrally flipped lower sideband,
per-stream threshold policies and latch slots, per-stream VDIF stamping, and
replays of the captured HIP graph against the eager launch sequence.

Bars: VDIF headers and stream-state advance bit-exact; 2-bit codes as in
tests/parity.py (differing codes counted in shells of 1e-6 and 1e-5 sigma from a
decision boundary, none beyond 1e-5 sigma, at most one per block-pol between);
graph replays byte-identical to the eager run.
"""
import sys

import numpy as np
import pytest

from oracle import sdbe_oracle as O
from parity import ShellTally
from sdbe_amd import synth as S

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


STATIONS = ("Sm", "Aa", "Zz")


def _streams(torch, L, data, frame_bytes, policies, flips, slots, fixed=(0.35, 0.01)):
    from sdbe_amd.stage import BatchStream
    out = []
    for i, d in enumerate(data):
        nb = d["n_blocks"]
        o = torch.empty((2, 64 * nb, frame_bytes), dtype=torch.uint8, device="cuda")
        b0 = torch.from_numpy(d["beng"][0].copy()).cuda()
        b1 = torch.from_numpy(d["beng"][1].copy()).cuda()
        out.append(BatchStream(b0, b1, o, flip=flips[i], policy=policies[i], fixed=fixed, latch_slot=slots[i],
                               station=STATIONS[i % 3], bdc=i & 1, rx=(i >> 1) & 1, secs=5000 + 7 * i,
                               df=124900 + 13 * i, ref_epoch=30, psn=100 * i + 1))
    return out


def test_batch_streams_match_oracle(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage, process_blocks
    n_blocks, flips = (2, 1, 3), (False, True, False)
    policies = (L.THRESH_PER_BLOCK, L.THRESH_REFERENCE, L.THRESH_FIXED)
    slots = (0, 5, 9)
    data = S.batch_streams(71, n_blocks=n_blocks, lsb=flips)
    st = SdbeStage(out_mode=L.OUT_FRAMES, chunk=4)          # chunks straddle the stream boundaries
    st.reset_latch(5)
    streams = _streams(torch, L, data, st.frame_bytes, policies, flips, slots)
    _say("[batch] gpu process_blocks %s" % (n_blocks,))
    process_blocks(st, streams)
    torch.cuda.synchronize()
    tally = ShellTally("batch")
    for i, (d, s) in enumerate(zip(data, streams)):
        nb = d["n_blocks"]
        out = s.out.cpu().numpy()
        hdr = out[:, :, :32].copy().view(np.uint32).reshape(2, 64 * nb, 8)
        pay = out[:, :, 32:].copy().view(np.uint32).reshape(2, 64 * nb, 2048)
        state = O.StreamState(secs=5000 + 7 * i, df=124900 + 13 * i, ref_epoch=30, psn=100 * i + 1,
                              station=O.station_from_str(STATIONS[i % 3]), bdc=i & 1, rx=(i >> 1) & 1)
        _say("[batch] oracle stream %d (%d blocks, flip %d)" % (i, nb, flips[i]))
        ref = O.process_blocks(d["beng"], list(range(nb)), policy=policies[i], latch=[None],
                               fixed=(0.35, 0.01), state=state, flip=flips[i])
        assert np.array_equal(hdr, ref.headers), i
        _, nxt = O.stamp_headers(64 * nb, state)
        assert s.get_state() == dict(secs=nxt.secs, df=nxt.df, ref_epoch=30, psn=nxt.psn)
        for blk in range(nb):
            for pol in (0, 1):
                x64 = ref.samples[(blk, pol)]
                th, off = ref.thresholds[(blk, pol)]
                codes = O.unpack_codes(pay[pol, 64 * blk:64 * (blk + 1)])
                sh = tally.add(x64, codes, th, off, ref.stats[(blk, pol)][1])
                _say("[batch] stream %d blk %d pol %d: shells %s" % (i, blk, pol, sh))
    _say(tally.report())
    tally.check()
    st.close()


def test_batch_flip_moves_the_tone(torch_cuda):
    """A tone at channel c of a lower-sideband stream comes out at the
    frequency of channel 16383 - c."""
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage, process_blocks
    rng = S.rng_for(72)
    c = 3000
    codes = [S.gaussian_codes(rng, 39, 1.0, ((c, 1.0, 0.0),)) for _ in range(2)]
    beng = np.stack([S.codes_to_beng_int8(x) for x in codes])
    data = [dict(beng=beng, n_blocks=1)]
    st = SdbeStage(out_mode=L.OUT_PAYLOAD, chunk=2)
    peaks = []
    for flip in (False, True):
        s = _streams(torch, L, data, st.frame_bytes, (L.THRESH_PER_BLOCK,), (flip,), (0,))
        process_blocks(st, s)
        pay = s[0].out.cpu().numpy().view(np.uint32).reshape(2, 64, 2048)
        x = O.unpack_codes(pay[0]).astype(np.float64) - 1.5
        spec = np.abs(np.fft.rfft(x))
        spec[0] = 0
        peaks.append(int(np.argmax(spec)))
    # channel k of 16384 at 2496 Msps -> bin 39 k of the 2,097,152-point output (+-1 from the snapshot impulse)
    assert abs(peaks[0] - 39 * c) <= 1
    assert abs(peaks[1] - 39 * (16383 - c)) <= 1
    st.close()


def test_batch_graph_replays_match_eager(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage, process_blocks
    flips = (True, False, True)
    policies = (L.THRESH_PER_BLOCK, L.THRESH_REFERENCE, L.THRESH_PER_BLOCK)
    slots = (0, 3, 0)
    data = S.batch_streams(73, n_blocks=(1, 2, 1), lsb=flips)
    eager = SdbeStage(out_mode=L.OUT_FRAMES, chunk=2)
    eager.reset_latch(3)
    ref = _streams(torch, L, data, eager.frame_bytes, policies, flips, slots)
    _say("[graph] eager run")
    process_blocks(eager, ref)
    torch.cuda.synchronize()
    want = [s.out.cpu().numpy() for s in ref]
    want_state = [s.get_state() for s in ref]
    eager.close()

    g = SdbeStage(out_mode=L.OUT_FRAMES, chunk=2, use_graph=True)
    for rep in range(3):
        g.reset_latch(3)
        streams = _streams(torch, L, data, g.frame_bytes, policies, flips, slots)
        _say("[graph] replay %d" % rep)
        process_blocks(g, streams)
        junk = torch.randn(1 << 20, device="cuda") * 2.0       # eager work between replays
        torch.cuda.synchronize()
        del junk
        for i, s in enumerate(streams):
            assert np.array_equal(s.out.cpu().numpy(), want[i]), (rep, i)
            assert s.get_state() == want_state[i]
    # a batch of another size is captured anew and still matches an eager run
    data2 = S.batch_streams(74, n_blocks=(2,), lsb=(False,))
    s2 = _streams(torch, L, data2, g.frame_bytes, (L.THRESH_PER_BLOCK,), (False,), (0,))
    _say("[graph] re-capture for a second batch shape")
    process_blocks(g, s2)
    e2 = SdbeStage(out_mode=L.OUT_FRAMES, chunk=2)
    r2 = _streams(torch, L, data2, e2.frame_bytes, (L.THRESH_PER_BLOCK,), (False,), (0,))
    process_blocks(e2, r2)
    torch.cuda.synchronize()
    assert np.array_equal(s2[0].out.cpu().numpy(), r2[0].out.cpu().numpy())
    g.close()
    e2.close()


def test_batch_argument_errors(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage, process_blocks
    data = S.batch_streams(75, n_blocks=(1,), lsb=(False,))
    st = SdbeStage(out_mode=L.OUT_PAYLOAD, chunk=2, max_batch=2)
    s = _streams(torch, L, data, st.frame_bytes, (L.THRESH_PER_BLOCK,), (False,), (0,))
    s[0].desc.latch_slot = L.LATCH_SLOTS
    with pytest.raises(L.SdbeError):
        process_blocks(st, s)
    s[0].desc.latch_slot = 0
    s[0].desc.out_pol_stride_bytes = 64 * 8192 - 1                 # too small for one block
    with pytest.raises(L.SdbeError):
        process_blocks(st, s)
    s[0].desc.out_pol_stride_bytes = 64 * 8192
    s[0].desc.thresh_policy = 7
    with pytest.raises(L.SdbeError):
        process_blocks(st, s)
    s[0].desc.thresh_policy = L.THRESH_PER_BLOCK
    process_blocks(st, s)                                          # valid again
    torch.cuda.synchronize()
    st.close()
