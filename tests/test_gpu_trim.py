"""GPU parity of the band-trimmed resampler of the aphids GPU thread
(SDBE_RATE_6_11 and SDBE_RATE_8_11; vdif_inout_gpu_thread.cu:589-709, plans
:269-333, presets sample_rates.h:21-40) against the FP64 oracle
(oracle.resample_trimmed, pinned in tests/test_oracle_trim.py).

Bars, as for the FP32 64/39 pipeline:
* float samples: RMS <= 2e-6 sigma and max <= 5e-5 sigma of the FP64 chain;
  block mean and sigma within 1e-6 (relative to sigma);
* 2-bit codes as in tests/parity.py (differing codes counted in shells of
  1e-6 and 1e-5 sigma from a decision boundary, none beyond 1e-5 sigma, at most
  one per block-pol between); headers bit-exact (frames of 8224 bytes, E per
  block-pol);
* a whole group through sdbe_process_group (shuffled packets with drops,
  REFERENCE latch): the same code bar on blocks of both polarisations;
* the group's first kernel reading the packets (either geometry) gives the
  same bytes, bitmap and counters as reading the unpacked int8 rows.
Rejected: unknown presets, a skip past the band, precision variants, ragged
batches and re-framed output in a band-trimmed handle.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import sdbe_oracle as O
from parity import ShellTally
from sdbe_amd import synth as S

pytestmark = pytest.mark.gpu
NCH = 16384


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


# MHz trimmed from DC: the reference's 150 / 102 (ctrl_aphids.py:163-202) where
# the preset's excess bandwidth allows; 10/11 (4160 Msps) has 64 MHz to spare
SKIP_A = {6: 150, 8: 150, 10: 16, 11: 150}
SKIP_B = {6: 102, 8: 102, 10: 32, 11: 102}


def _stage(frac, skip=None, **kw):
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage
    kw.setdefault("policy", L.THRESH_PER_BLOCK)
    kw.setdefault("out_mode", L.OUT_PAYLOAD)
    kw.setdefault("chunk", 2)
    return SdbeStage(rate_frac=frac, skip_mhz=SKIP_A[frac] if skip is None else skip, **kw)


def _spectra(seed, D, n):
    rng = S.rng_for(seed)
    codes = S.gaussian_codes(rng, D * n)
    return S.codes_to_beng_int8(codes).reshape(n, D, NCH, 2)


def _oracle_block(beng_block, E, skip):
    return O.resample_trimmed(O.inflate(beng_block)[:, :NCH], E, skip)


# skip 0 takes the segment spectrum from DC, where the R2C post-processing pairs
# bin 0 with itself (Z[R/2 - 0] is Z[0]); 10/11 at SKIP_B (32 MHz) ends exactly
# at the last bin R/2 (test_trim_codes_and_headers)
@pytest.mark.parametrize("frac,skip_mhz", [(6, None), (8, None), (10, None), (11, None), (6, 0), (8, 0)])
def test_trim_samples_vs_fp64(torch_cuda, frac, skip_mhz):
    torch = torch_cuda
    skip_mhz = SKIP_A[frac] if skip_mhz is None else skip_mhz
    D, E, R, P, skip = O.trim_geometry(frac, skip_mhz)
    n = 2
    beng = _spectra(600 + frac, D, n)
    st = _stage(frac, skip=skip_mhz, max_spectra=n)
    _say("[trim] %d/11 resample_f32 n=%d (D=%d E=%d skip=%d)" % (frac, n, D, E, skip))
    x, stats = st.resample_f32(torch.from_numpy(beng).cuda())
    x = x.cpu().numpy().astype(np.float64)
    stats = stats.cpu().numpy()
    assert x.shape == (n, 128 * P)
    for b in range(n):
        x64 = _oracle_block(beng[b], E, skip)
        mean, sig = O.block_stats(x64)
        err = x[b] - x64
        rms, mx = np.sqrt(np.mean(err ** 2)) / sig, np.abs(err).max() / sig
        _say("[trim] %d/11 block %d: RMS %.3e sigma, max %.3e sigma; mean %.6g sigma %.6g (oracle %.6g %.6g)"
             % (frac, b, rms, mx, stats[b, 0], stats[b, 1], mean, sig))
        assert rms <= 2e-6 and mx <= 5e-5
        assert abs(stats[b, 0] - mean) <= 1e-6 * sig and abs(stats[b, 1] - sig) <= 1e-6 * sig
    st.close()


@pytest.mark.parametrize("frac", [6, 8, 10, 11])
def test_trim_codes_and_headers(torch_cuda, frac):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    D, E, R, P, skip = O.trim_geometry(frac, SKIP_B[frac])
    n = 2
    beng = _spectra(700 + frac, D, n)
    st = _stage(frac, skip=SKIP_B[frac], out_mode=L.OUT_FRAMES, max_spectra=n, station="Sm", bdc=1, rx=1)
    st.set_state(secs=4321, df=124995, ref_epoch=31, psn=900)
    _say("[trim] %d/11 process_spectra frames n=%d" % (frac, n))
    out = st.process_spectra(torch.from_numpy(beng).cuda()).cpu().numpy()
    assert out.shape == (n, E, 8224)
    hdr = out[:, :, :32].copy().view(np.uint32).reshape(n * E, 8)
    pay = out[:, :, 32:].copy().view(np.uint32).reshape(n, -1)
    state = O.StreamState(secs=4321, df=124995, ref_epoch=31, psn=900, station=O.station_from_str("Sm"),
                          bdc=1, rx=1)
    ref_hdr, ref_next = O.stamp_headers(n * E, state)
    assert np.array_equal(hdr, ref_hdr[0])
    assert st.get_state() == dict(secs=ref_next.secs, df=ref_next.df, ref_epoch=31, psn=ref_next.psn)
    tally = ShellTally("trim %d/11 codes" % frac)
    for b in range(n):
        x64 = _oracle_block(beng[b], E, skip)
        mean, sig = O.block_stats(x64)
        codes = O.unpack_codes(pay[b])
        assert codes.size == 128 * P
        _say("[trim] %d/11 block %d: shells %s" % (frac, b, tally.add(x64, codes, sig, mean)))
    _say(tally.report())
    tally.check()
    st.close()


@pytest.mark.parametrize("frac", [6, 8, 10, 11])
def test_trim_group_reference_policy(torch_cuda, frac):
    """A group is D B-frames (BENG_FRAMES_PER_GROUP = DECIMATION_FACTOR,
    sample_rates.h:67): 128 blocks of D snapshots, E frames per block-pol."""
    from sdbe_amd import _lib as L
    D, E, R, P, skip = O.trim_geometry(frac, SKIP_A[frac])
    rng = S.rng_for(800 + frac)
    _say("[trim] %d/11 generating a group of %d B-frames" % (frac, D))
    c0 = S.gaussian_codes(rng, D * 128)
    c1 = S.gaussian_codes(rng, D * 128)
    b0 = 4242
    pk = S.packetize(c0, c1, b0)
    del c0, c1
    pk_shuf, _ = S.shuffle_and_drop(pk, rng, drop_frac=0.001)
    del pk
    st = _stage(frac, policy=L.THRESH_REFERENCE, chunk=256)
    d = st.to_device_packets(pk_shuf)
    _say("[trim] %d/11 process_group (%d packets)" % (frac, pk_shuf.shape[0]))
    out = st.process_group(d, pk_shuf.shape[0], b0).cpu().numpy()
    stats = st.stats()
    nblk = 128
    assert out.shape == (2, nblk * E, 8192)
    ref = O.vdif_to_beng(pk_shuf, b0, n_frames=D)
    assert stats["n_unique"] == ref.n_unique and stats["n_missing"] == ref.n_missing == round(0.001 * D * 2048)
    snaps = [ref.beng[p].reshape(-1, NCH, 2) for p in range(2)]
    x00 = _oracle_block(snaps[0][0:D], E, skip)
    mean0, sig0 = O.block_stats(x00)            # REFERENCE: pol 0 block 0 latches both
    tally = ShellTally("trim %d/11 group" % frac)
    for blk in (0, 1, nblk - 1):
        for p in range(2):
            x64 = x00 if (blk, p) == (0, 0) else _oracle_block(snaps[p][blk * D:(blk + 1) * D], E, skip)
            words = out[p, blk * E:(blk + 1) * E].copy().view(np.uint32).reshape(-1)
            sh = tally.add(x64, O.unpack_codes(words), sig0, mean0)
            _say("[trim] %d/11 group block %d pol %d: shells %s" % (frac, blk, p, sh))
    _say(tally.report())
    tally.check()
    st.close()


@pytest.mark.parametrize("geometry", [0, 1])
@pytest.mark.parametrize("frac", [6, 8, 10, 11])
def test_trim_group_packet_c2r_matches_rows(torch_cuda, frac, geometry):
    """From a 16-byte aligned buffer the band-trimmed group's first kernel reads
    the packets (k_trim_c2r_pk); from a buffer 8 bytes off it reads the int8
    rows the unpack writes.  Whole group, both polarisations, every block: the
    output bytes, the fid bitmap and the counters are identical."""
    torch = torch_cuda
    from sdbe_amd import _lib as L
    D, E, R, P, skip = O.trim_geometry(frac, SKIP_A[frac])
    _say("[trim-pk] %d/11 geometry %d: generating a group of %d B-frames" % (frac, geometry, D))
    groups, _ = S.make_group_packets(seed=830 + frac + 10 * geometry, n_groups=1, drop_frac=0.002, frames=D,
                                     geometry=geometry)
    pk, b0 = groups[0]
    del groups
    gk = L.GEOM_8X1024 if geometry else L.GEOM_REFERENCE
    st = _stage(frac, chunk=256, packet_geometry=gk)
    d = st.to_device_packets(pk)
    raw = torch.empty(d.numel() + 8, dtype=torch.uint8, device="cuda")
    raw[8:].copy_(d)
    du = raw[8:]
    assert d.data_ptr() % 16 == 0 and du.data_ptr() % 16 == 8
    bshape, bdt = ((D, 16), torch.int16) if geometry else ((D, 256), torch.uint8)
    bm_a = torch.full(bshape, -1, dtype=bdt, device="cuda")
    bm_b = torch.full(bshape, -1, dtype=bdt, device="cuda")
    _say("[trim-pk] %d/11 geometry %d: process_group from packets (%d)" % (frac, geometry, pk.shape[0]))
    a = st.process_group(d, pk.shape[0], b0, bitmap=bm_a).cpu().numpy()
    sa = st.stats()
    _say("[trim-pk] %d/11 geometry %d: process_group from int8 rows" % (frac, geometry))
    b = st.process_group(du, pk.shape[0], b0, bitmap=bm_b).cpu().numpy()
    sb = st.stats()
    st.close()
    assert np.array_equal(a, b)
    assert torch.equal(bm_a, bm_b)
    assert sa["n_unique"] == sb["n_unique"] and sa["n_missing"] == sb["n_missing"] and sa["n_unique"] > 0


def test_trim_rejects(torch_cuda):
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage
    for kw in (dict(rate_frac=7), dict(rate_frac=6, skip_mhz=300), dict(rate_frac=8, skip_mhz=700),
               dict(rate_frac=10, skip_mhz=33), dict(rate_frac=11, skip_mhz=241),
               dict(rate_frac=6, skip_mhz=-1), dict(rate_frac=6, fft_precision=L.FFT_FP16)):
        with pytest.raises(L.SdbeError):
            SdbeStage(**kw)
    st = _stage(6)
    lib = st.lib
    assert lib.sdbe_set_output_alignment(st.h, 1, 0, 0) == L.SDBE_ERR_STATE
    desc = L.StreamDesc()
    assert lib.sdbe_process_blocks(st.h, C.byref(desc), 1, None) == L.SDBE_ERR_STATE
    st.close()
