"""GPU tests of the batching workload at its own shape: 6 independent streams
(3 beams x 2 sidebands, every lower sideband spectrally flipped, rslib.py:124),
both polarisations, per-stream sigma ~ U[0.5, 2] through a unit-sigma 2-bit
quantiser and a random tone per stream, all block-pols of all streams in one
sdbe_process_blocks call captured as a HIP graph.

* Scaled block counts {3, 4, 5, 1, 6, 2}: codes of the first, a middle and the
  last block of every stream against the FP64 oracle (tests/parity.py shells),
  headers and stream-state advance bit-exact, a second graph replay
  byte-identical to the first.
* The full shape {17, 64, 128, 5, 200, 33} (894 block-pols): each stream's
  spectra repeat a seeded 8-block pattern, so under PER_BLOCK thresholds block
  j and block j + 8 of a stream carry byte-identical payloads; the whole batch
  as one chunk (one launch per stage) captured as a graph, replayed with eager
  work in between, is byte-identical to the eager run in chunks of 256;
  headers and state advance by 64 n_blocks frames bit-exact; block 0 of every
  stream matches the FP64 oracle.
"""
import sys

import numpy as np
import pytest

from oracle import sdbe_oracle as O
from parity import ShellTally
from sdbe_amd import synth as S

pytestmark = pytest.mark.gpu

LSB = (False, True, False, True, False, True)
STATIONS = ("Sm", "Aa", "Zz")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _state(i):
    return dict(secs=7000 + 3 * i, df=124800 + 37 * i, ref_epoch=30, psn=1000 * i + 1)   # df < 125000


def _make_streams(torch, stage, pols, n_blocks, policies, slots, outs=None):
    from sdbe_amd.stage import BatchStream
    out = []
    for i, nb in enumerate(n_blocks):
        o = outs[i] if outs is not None else torch.empty((2, 64 * nb, stage.frame_bytes), dtype=torch.uint8,
                                                        device="cuda")
        out.append(BatchStream(pols[i][0], pols[i][1], o, flip=LSB[i], policy=policies[i], latch_slot=slots[i],
                               station=STATIONS[i // 2], bdc=i & 1, rx=0, **_state(i)))
    return out


def _check_headers(stream, i, nb):
    out = stream.out.cpu().numpy()
    hdr = out[:, :, :32].copy().view(np.uint32).reshape(2, 64 * nb, 8)
    st = O.StreamState(station=O.station_from_str(STATIONS[i // 2]), bdc=i & 1, rx=0, **_state(i))
    want, nxt = O.stamp_headers(64 * nb, st)
    assert np.array_equal(hdr, want), i
    assert stream.get_state() == dict(secs=nxt.secs, df=nxt.df, ref_epoch=30, psn=nxt.psn), i
    return out


def test_config4_scaled_streams_one_graph(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage, process_blocks
    n_blocks = (3, 4, 5, 1, 6, 2)
    policies = (L.THRESH_PER_BLOCK, L.THRESH_REFERENCE, L.THRESH_PER_BLOCK, L.THRESH_PER_BLOCK,
                L.THRESH_REFERENCE, L.THRESH_PER_BLOCK)
    slots = (0, 11, 0, 0, 12, 0)
    data = S.batch_streams(401, n_blocks=n_blocks, lsb=LSB)
    pols = [[torch.from_numpy(d["beng"][p].copy()).cuda() for p in range(2)] for d in data]
    st = SdbeStage(out_mode=L.OUT_FRAMES, chunk=16, use_graph=True)   # chunks straddle streams
    for s in (11, 12):
        st.reset_latch(s)
    streams = _make_streams(torch, st, pols, n_blocks, policies, slots)
    _say("[cfg4] gpu process_blocks %s, one graph" % (n_blocks,))
    process_blocks(st, streams)
    torch.cuda.synchronize()
    first = [s.out.clone() for s in streams]
    tally = ShellTally("config 4 scaled")
    for i, (d, s) in enumerate(zip(data, streams)):
        nb = n_blocks[i]
        out = _check_headers(s, i, nb)
        pay = out[:, :, 32:].copy().view(np.uint32).reshape(2, 64 * nb, 2048)
        check = sorted({0, nb // 2, nb - 1})
        _say("[cfg4] oracle stream %d (%d blocks, sigma %.3f, flip %d): blocks %s"
             % (i, nb, d["sigma"], LSB[i], check))
        # check starts at block 0, where a REFERENCE stream latches
        ref = O.process_blocks(d["beng"], check, policy=policies[i], latch=[None], flip=LSB[i])
        for blk in check:
            for pol in (0, 1):
                th, off = ref.thresholds[(blk, pol)]
                codes = O.unpack_codes(pay[pol, 64 * blk:64 * (blk + 1)])
                sh = tally.add(ref.samples[(blk, pol)], codes, th, off, ref.stats[(blk, pol)][1])
                _say("[cfg4] stream %d blk %d pol %d: shells %s" % (i, blk, pol, sh))
    _say(tally.report())
    tally.check()
    # the second call replays the captured graph: same bytes (latches cleared first)
    for s in (11, 12):
        st.reset_latch(s)
    again = _make_streams(torch, st, pols, n_blocks, policies, slots)
    _say("[cfg4] graph replay")
    process_blocks(st, again)
    torch.cuda.synchronize()
    for i, s in enumerate(again):
        assert torch.equal(s.out, first[i]), i
    st.close()


def _pattern_streams(torch, n_blocks, seed, period=8):
    """Per stream [pol0, pol1] device spectra repeating a seeded `period`-block
    pattern, and the pattern's numpy data."""
    pattern = S.batch_streams(seed, n_blocks=[min(period, n) for n in n_blocks], lsb=LSB)
    pols = []
    for d, n in zip(pattern, n_blocks):
        reps = -(-n // d["n_blocks"])
        pols.append([torch.from_numpy(d["beng"][p]).cuda().repeat(reps, 1, 1)[:39 * n].contiguous()
                     for p in range(2)])
    return pols, pattern


def test_config4_full_shape_properties(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage, process_blocks
    n_blocks = (17, 64, 128, 5, 200, 33)
    policies = (L.THRESH_PER_BLOCK,) * 6
    slots = (0,) * 6
    _say("[cfg4-full] building the %d-block-pol batch" % (2 * sum(n_blocks)))
    pols, pattern = _pattern_streams(torch, n_blocks, 402)
    eager = SdbeStage(out_mode=L.OUT_FRAMES)
    ref = _make_streams(torch, eager, pols, n_blocks, policies, slots)
    _say("[cfg4-full] eager process_blocks")
    process_blocks(eager, ref)
    torch.cuda.synchronize()
    eager.close()
    outs = [torch.empty_like(s.out) for s in ref]
    # the whole batch as one chunk (one launch per stage), captured as a graph
    g = SdbeStage(out_mode=L.OUT_FRAMES, use_graph=True, chunk=1024)
    for rep in range(3):
        for o in outs:
            o.fill_(0x5A)
        streams = _make_streams(torch, g, pols, n_blocks, policies, slots, outs=outs)
        _say("[cfg4-full] graph call %d" % rep)
        process_blocks(g, streams)
        junk = torch.randn(1 << 22, device="cuda") * 3.0          # eager work between replays
        torch.cuda.synchronize()
        del junk
        for i, s in enumerate(streams):
            assert torch.equal(s.out, ref[i].out), (rep, i)
            assert s.get_state() == ref[i].get_state(), (rep, i)
    g.close()
    tally = ShellTally("config 4 full, block 0 of each stream")
    for i, s in enumerate(ref):
        nb = n_blocks[i]
        out = _check_headers(s, i, nb)
        pay = out[:, :, 32:]
        if nb > 8:
            # period-8 input -> period-8 payloads under PER_BLOCK thresholds
            assert np.array_equal(pay[:, 64 * 8:], pay[:, :64 * (nb - 8)]), i
        d = pattern[i]
        _say("[cfg4-full] oracle block 0 of stream %d (%d blocks, flip %d)" % (i, nb, LSB[i]))
        r = O.process_blocks(d["beng"], [0], policy=O.THRESH_PER_BLOCK, flip=LSB[i])
        words = pay.copy().view(np.uint32).reshape(2, 64 * nb, 2048)
        for pol in (0, 1):
            th, off = r.thresholds[(0, pol)]
            tally.add(r.samples[(0, pol)], O.unpack_codes(words[pol, :64]), th, off, r.stats[(0, pol)][1])
    _say(tally.report())
    tally.check()
