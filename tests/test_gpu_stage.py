"""GPU parity tests: the HIP stage through its C-ABI against the FP64 oracle.

Bars:
* unpack / reorder / zero fill / fid bitmap / counters: bit-exact;
* resampled FP32 samples: RMS error <= 2e-6 sigma and max error <= 5e-5 sigma
  of the FP64 oracle (the FP32 pipeline's rounding; the accuracy target is 1e-3 sigma);
* 2-bit codes: a differing code is counted by the distance of its FP64 sample
  from the nearest decision boundary (-th, 0, +th after the offset) in three
  shells (<= 1e-6, 1e-6..1e-5, > 1e-5 sigma); none may lie beyond 1e-5 sigma
  and at most one per block-pol between 1e-6 and 1e-5 sigma (tests/parity.py
  derives that bound from the FP32 error distribution);
* VDIF headers: bit-exact.
"""
import sys

import numpy as np
import pytest

from oracle import sdbe_oracle as O
from parity import MIDDLE_MAX_PER_BP, ShellTally, boundary_distance
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


def _stage(**kw):
    from sdbe_amd.stage import SdbeStage
    return SdbeStage(**kw)


def _spectra_codes(seed, n_bp, **kw):
    rng = S.rng_for(seed)
    codes = S.gaussian_codes(rng, 39 * n_bp, **kw)
    return S.codes_to_beng_int8(codes).reshape(n_bp, 39, 16384, 2)


# --------------------------------------------------------------------- unpack
def test_unpack_group_bit_exact(torch_cuda):
    _say("[unpack] generating one shuffled group with drops, duplicates, invalid and out-of-range packets")
    rng = S.rng_for(11)
    c0 = S.gaussian_codes(rng, 39 * 128)
    c1 = S.gaussian_codes(rng, 39 * 128)
    b0 = 5000
    pk = S.packetize(c0, c1, b0)
    pk, _ = S.shuffle_and_drop(pk, rng, drop_frac=0.001)
    # duplicates with different payload placed after the original: first wins
    dup = pk[:7].copy()
    dup[:, 8:] ^= 0x55555555
    inval = pk[7:9].copy()
    inval[:, 0] |= np.uint32(1 << 31)
    inval[:, 8:] = 0
    far = pk[9:11].copy()
    far[:, 4] += np.uint32(1)          # B-count + 256: outside the 39-frame window
    pk = np.ascontiguousarray(np.concatenate([pk, dup, inval, far]))
    _say("[unpack] oracle")
    ref = O.vdif_to_beng(pk, b0)
    st = _stage()
    d = st.to_device_packets(pk)
    _say("[unpack] gpu n_pkts=%d" % pk.shape[0])
    g0, g1, bm = st.unpack_int8(d, pk.shape[0], b0)
    stats = st.stats()
    g0 = g0.cpu().numpy()
    g1 = g1.cpu().numpy()
    assert np.array_equal(g0, ref.beng[0])
    assert np.array_equal(g1, ref.beng[1])
    assert np.array_equal(bm.cpu().numpy(), ref.bitmap)
    assert stats["n_invalid"] == ref.n_invalid == 2
    assert stats["n_out_of_range"] == ref.n_out_of_range
    assert stats["n_duplicate"] == ref.n_duplicate == 7
    assert stats["n_unique"] == ref.n_unique
    assert stats["n_missing"] == ref.n_missing == round(0.001 * 79872)
    st.close()


def test_unpack_wraparound_bcount(torch_cuda):
    """B-counts straddling 2^32 of the 24+8-bit field the kernel reads, and a
    b_zero above 2^32 (only the low 32 bits enter, vdif_inout_gpu_thread.cu:437)."""
    rng = S.rng_for(12)
    c0 = S.gaussian_codes(rng, 39 * 128, tones=())
    c1 = S.gaussian_codes(rng, 39 * 128, tones=())
    b0 = (1 << 32) - 20 + (7 << 32)
    pk = S.packetize(c0, c1, b0)
    pk, _ = S.shuffle_and_drop(pk, rng, drop_frac=0.0)
    ref = O.vdif_to_beng(pk, b0)
    assert ref.n_unique == 79872
    st = _stage()
    d = st.to_device_packets(pk)
    _say("[unpack-wrap] gpu")
    g0, g1, bm = st.unpack_int8(d, pk.shape[0], b0)
    assert np.array_equal(g0.cpu().numpy(), ref.beng[0])
    assert np.array_equal(g1.cpu().numpy(), ref.beng[1])
    st.close()


# ------------------------------------------------------------------ resampler
def test_resample_f32_vs_fp64(torch_cuda):
    torch = torch_cuda
    n_bp = 3
    beng = _spectra_codes(21, n_bp)
    st = _stage(policy=1, max_spectra=n_bp)
    _say("[resample] gpu n_bp=%d" % n_bp)
    x, stt = st.resample_f32(torch.from_numpy(beng).cuda())
    x = x.cpu().numpy().astype(np.float64)
    stt = stt.cpu().numpy()
    for b in range(n_bp):
        x64 = O.resample_block(O.inflate(beng[b]))
        mean, sig = O.block_stats(x64)
        err = x[b] - x64
        rms = np.sqrt(np.mean(err ** 2)) / sig
        mx = np.abs(err).max() / sig
        # expected code flips per shell from this error distribution: a sample
        # with error e flips its code when its distance d to a boundary is below
        # |e|; distances near a boundary are uniform at the probability density
        # rho per sigma measured here, so the expected flips with d in (a, b]
        # are rho * sum(clip(|e|, a, b) - a)
        d = boundary_distance(x64, sig, mean, sig)
        rho = np.count_nonzero(d <= 1e-2) / 1e-2 / x64.size
        ae = np.abs(err) / sig
        exp_inner = rho * np.sum(np.clip(ae, 0.0, 1e-6))
        exp_mid = rho * np.sum(np.clip(ae, 1e-6, 1e-5) - 1e-6)
        exp_out = rho * np.sum(np.clip(ae, 1e-5, None) - 1e-5)
        _say("[resample] bp %d rms %.3g sigma, max %.3g sigma; boundary density %.3f / sigma; expected "
             "code flips per block-pol: %.3g within 1e-6 sigma, %.3g in 1e-6..1e-5, %.3g beyond"
             % (b, rms, mx, rho, exp_inner, exp_mid, exp_out))
        assert exp_mid <= 0.1 * MIDDLE_MAX_PER_BP and exp_out <= 0.01
        assert rms <= 2e-6
        assert mx <= 5e-5
        assert abs(stt[b, 0] - mean) <= 1e-6 * sig
        assert abs(stt[b, 1] - sig) <= 1e-6 * sig
    st.close()


def test_resample_pure_tone_is_analytic(torch_cuda):
    """Channel k = 1 + 0j in every snapshot is a cosine continuous across the
    block, i.e. bin 39 k of the 1,277,952-point DFT, so the resampled output is
    the analytic (39/64) (2/32768) cos(2 pi 39 k t / 2^21) -- the 39/64 because
    numpy's irfft of length 2^21 divides by 2^21 while the forward transform
    spanned 39 * 32768 samples (resample.py:130-138 composes it the same way).
    Pins the composition and normalisation of the whole GPU resampler."""
    torch = torch_cuda
    k = 100
    spec = np.zeros((1, 39, 16384, 2), np.int8)
    spec[0, :, k, 1] = 1
    st = _stage(policy=1, max_spectra=1)
    _say("[tone] gpu")
    x, stt = st.resample_f32(torch.from_numpy(spec).cuda())
    x = x.cpu().numpy()[0].astype(np.float64)
    t = np.arange(O.L_C2R)
    za = (39.0 / 64.0) * (2.0 / 32768) * np.cos(2 * np.pi * k * 39 * t / O.L_C2R)
    err = np.abs(x - za).max() / np.abs(za).max()
    _say("[tone] max rel err %.3g" % err)
    assert err < 2e-5
    assert abs(stt.cpu().numpy()[0, 1] - za.std(ddof=1)) < 1e-6 * za.std()
    st.close()


def test_process_spectra_codes_and_headers(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    n_bp = 2
    beng = _spectra_codes(31, n_bp)
    st = _stage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_FRAMES, max_spectra=n_bp, station="Sm",
                bdc=1, rx=0)
    st.set_state(secs=12345, df=124990, ref_epoch=30, psn=77)
    _say("[codes] gpu process_spectra n_bp=%d" % n_bp)
    out = st.process_spectra(torch.from_numpy(beng).cuda()).cpu().numpy()
    new_state = st.get_state()
    hdr = out[:, :, :32].copy().view(np.uint32).reshape(n_bp * 64, 8)
    pay = out[:, :, 32:].copy().view(np.uint32).reshape(n_bp, -1)
    state = O.StreamState(secs=12345, df=124990, ref_epoch=30, psn=77,
                          station=O.station_from_str("Sm"), bdc=1, rx=0)
    ref_hdr, ref_next = O.stamp_headers(n_bp * 64, state)
    assert np.array_equal(hdr, ref_hdr[0])
    assert new_state == dict(secs=ref_next.secs, df=ref_next.df, ref_epoch=30, psn=ref_next.psn)
    tally = ShellTally("codes")
    for b in range(n_bp):
        x64 = O.resample_block(O.inflate(beng[b]))
        mean, sig = O.block_stats(x64)
        codes = O.unpack_codes(pay[b])
        _say("[codes] bp %d: shells %s" % (b, tally.add(x64, codes, sig, mean)))
        # histogram equals the oracle's (not the Gaussian 0.159 per outer code: the
        # (code & 3) - 2 decode bias puts an impulse at the start of every snapshot)
        frac = np.bincount(codes, minlength=4) / codes.size
        ref_frac = np.bincount(O.quantize2bit(x64, sig, mean), minlength=4) / codes.size
        assert np.abs(frac - ref_frac).max() < 1e-5
    _say(tally.report())
    tally.check()
    st.close()


# ------------------------------------------------------------------ end to end
@pytest.fixture(scope="module")
def one_group():
    rng = S.rng_for(41)
    c0 = S.gaussian_codes(rng, 39 * 128)
    c1 = S.gaussian_codes(rng, 39 * 128)
    b0 = 777
    pk = S.packetize(c0, c1, b0)
    pk_shuf, _ = S.shuffle_and_drop(pk, rng, drop_frac=0.001)
    return pk, pk_shuf, b0


def test_process_group_reference_policy(torch_cuda, one_group):
    """REFERENCE policy over a shuffled group with drops: the latch is block 0's
    pol-0 (mean, sigma), and the codes of blocks across the group -- the blocks
    whose 39 snapshots straddle two B-frames (3: frames 0-1, 16: 4-5, 19: 5-6,
    42: 12-13) among them -- match the FP64 oracle in both polarisations."""
    from sdbe_amd import _lib as L
    pk, pk_shuf, b0 = one_group
    st = _stage(policy=L.THRESH_REFERENCE, out_mode=L.OUT_PAYLOAD)
    d = st.to_device_packets(pk_shuf)
    _say("[group] gpu process_group (REFERENCE policy)")
    out = st.process_group(d, pk_shuf.shape[0], b0).cpu().numpy()
    stats = st.stats()
    assert stats["n_missing"] == round(0.001 * 79872)
    ref = O.vdif_to_beng(pk_shuf, b0)
    blocks = [0, 1, 3, 16, 19, 42, 64, 127]
    _say("[group] oracle blocks %s" % blocks)
    x00 = O.resample_block(O.block_spectra(ref.beng[0], 0))
    mean0, sig0 = O.block_stats(x00)
    assert abs(stats["latched_threshold"] - sig0) <= 2e-6 * sig0
    assert abs(stats["latched_offset"] - mean0) <= 2e-6 * sig0
    pay = out.view(np.uint32).reshape(2, 8192, 2048)
    tally = ShellTally("group")
    for blk in blocks:
        for pol in (0, 1):
            x64 = x00 if (blk, pol) == (0, 0) else O.resample_block(O.block_spectra(ref.beng[pol], blk))
            codes = O.unpack_codes(pay[pol, blk * 64:(blk + 1) * 64])
            _say("[group] blk %d pol %d: shells %s" % (blk, pol, tally.add(x64, codes, sig0, mean0)))
    _say(tally.report())
    tally.check()
    st.close()


def test_process_group_reference_relatch_after_empty_blocks(torch_cuda, one_group):
    """REFERENCE policy when the stream opens with nothing: every packet of
    B-frame 0 is missing, so blocks 0-2 (snapshots 0..116) are all zero in both
    polarisations and block 3 (snapshots 117..155) holds 28 snapshots of data.
    The reference recomputes the latch while the latched threshold is 0
    (vdif_inout_gpu_thread.cu:954-1005): blocks 0-2 latch sigma = 0, mean = 0
    and quantise every sample (x = 0 >= th = 0) to code 3; block 3's pol-0
    (mean, sigma) is latched next and serves both polarisations of every later
    block."""
    from sdbe_amd import _lib as L
    _, pk_shuf, b0 = one_group
    w5 = pk_shuf[:, 5]
    bc = ((w5 >> 24) + ((pk_shuf[:, 4] & 0xFFFFFF) << 8)).astype(np.int64) - b0
    pk_cut = np.ascontiguousarray(pk_shuf[bc != 0])
    st = _stage(policy=L.THRESH_REFERENCE, out_mode=L.OUT_PAYLOAD)
    d = st.to_device_packets(pk_cut)
    _say("[relatch] gpu process_group without B-frame 0 (n_pkts=%d)" % pk_cut.shape[0])
    out = st.process_group(d, pk_cut.shape[0], b0).cpu().numpy()
    stats = st.stats()
    pay = out.view(np.uint32).reshape(2, 8192, 2048)
    for blk in (0, 1, 2):
        for pol in (0, 1):
            codes = O.unpack_codes(pay[pol, blk * 64:(blk + 1) * 64])
            assert (codes == 3).all(), (blk, pol)
    ref = O.vdif_to_beng(pk_cut, b0)
    x30 = O.resample_block(O.block_spectra(ref.beng[0], 3))
    mean3, sig3 = O.block_stats(x30)
    assert sig3 > 0
    _say("[relatch] latched (%.6g, %.6g), block 3 pol 0 (%.6g, %.6g)"
         % (stats["latched_threshold"], stats["latched_offset"], sig3, mean3))
    assert abs(stats["latched_threshold"] - sig3) <= 2e-6 * sig3
    assert abs(stats["latched_offset"] - mean3) <= 2e-6 * sig3
    tally = ShellTally("relatch")
    for blk in (3, 4, 100):
        for pol in (0, 1):
            x64 = x30 if (blk, pol) == (3, 0) else O.resample_block(O.block_spectra(ref.beng[pol], blk))
            codes = O.unpack_codes(pay[pol, blk * 64:(blk + 1) * 64])
            _say("[relatch] blk %d pol %d: shells %s" % (blk, pol, tally.add(x64, codes, sig3, mean3)))
    # the oracle's walk over blocks 0..3 latches the same values
    walk = O.process_blocks(ref.beng, [0, 1, 2, 3], policy=O.THRESH_REFERENCE, latch=[None])
    assert walk.thresholds[(0, 1)] == (0.0, 0.0)
    assert walk.thresholds[(3, 1)] == (sig3, mean3)
    _say(tally.report())
    tally.check()
    st.close()


def test_process_group_order_invariant_and_deterministic(torch_cuda, one_group):
    """Arrival order does not change the output (index-addressed reorder), and a
    repeated call is bit-identical."""
    from sdbe_amd import _lib as L
    pk, _, b0 = one_group
    rng = S.rng_for(43)
    pa, _ = S.shuffle_and_drop(pk, rng, drop_frac=0.0)
    pb, _ = S.shuffle_and_drop(pk, rng, drop_frac=0.0)
    st = _stage(policy=L.THRESH_PER_BLOCK)
    _say("[group-inv] gpu x3")
    oa = st.process_group(st.to_device_packets(pa), pa.shape[0], b0).cpu().numpy()
    ob = st.process_group(st.to_device_packets(pb), pb.shape[0], b0).cpu().numpy()
    oc = st.process_group(st.to_device_packets(pk), pk.shape[0], b0).cpu().numpy()
    assert np.array_equal(oa, ob)
    assert np.array_equal(oa, oc)
    st.close()


@pytest.mark.parametrize("tiles,ilv,chunk", [(8, 16, 256), (3, 5, 6)])
def test_process_group_c1_tile_order_does_not_change_output(torch_cuda, one_group, tiles, ilv, chunk):
    """Stage C1's tile order (c1_tiles_per_wg x c1_interleave; the second case
    leaves a partial last group in every chunk of 6 block-pols) changes only
    which workgroup runs a tile and how the Parseval sums are grouped: the
    group's output equals the one-workgroup-per-consecutive-run order."""
    from sdbe_amd import _lib as L
    pk, _, b0 = one_group
    ref = _stage(policy=L.THRESH_PER_BLOCK, chunk=chunk, c1_tiles_per_wg=tiles, c1_interleave=1)
    alt = _stage(policy=L.THRESH_PER_BLOCK, chunk=chunk, c1_tiles_per_wg=tiles, c1_interleave=ilv)
    _say("[c1-order] tiles %d, interleave 1 vs %d, chunk %d" % (tiles, ilv, chunk))
    oa = ref.process_group(ref.to_device_packets(pk), pk.shape[0], b0).cpu().numpy()
    ob = alt.process_group(alt.to_device_packets(pk), pk.shape[0], b0).cpu().numpy()
    assert np.array_equal(oa, ob)
    ref.close()
    alt.close()


def test_process_group_packet_stage_a_matches_rows(torch_cuda, one_group):
    """sdbe_process_group's stage A decodes the packets in place (through the
    packet map, a missing packet as 0 + 0j); it must equal stage A over the
    reordered int8 rows of the same group bit for bit, for all 128 blocks of both
    polarisations (including the blocks whose 39 snapshots straddle two
    B-frames), and give the same fid bitmap and counters."""
    torch = torch_cuda
    from sdbe_amd import _lib as L
    pk, pk_shuf, b0 = one_group
    # also drop every fid of one cid in the two frames under block 3 (snapshots 117..155)
    keep = np.ones(pk_shuf.shape[0], dtype=bool)
    w5 = pk_shuf[:, 5]
    bc = ((w5 >> 24) + ((pk_shuf[:, 4] & 0xFFFFFF) << 8)).astype(np.int64) - b0
    keep &= ~(((bc == 0) | (bc == 1)) & ((w5 & 0xFF) == 77))
    pk_cut = np.ascontiguousarray(pk_shuf[keep])
    st = _stage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_PAYLOAD)
    d = st.to_device_packets(pk_cut)
    bm = torch.empty((39, 256), dtype=torch.uint8, device="cuda")
    _say("[group-pk] gpu process_group, n_pkts=%d" % pk_cut.shape[0])
    out = st.process_group(d, pk_cut.shape[0], b0, bitmap=bm).cpu().numpy()
    stats = st.stats()
    ref = _stage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_PAYLOAD)
    _say("[group-pk] gpu unpack_int8 + process_spectra per polarisation")
    g0, g1, bm2 = ref.unpack_int8(d, pk_cut.shape[0], b0)
    ref_stats = ref.stats()
    for pol, g in enumerate((g0, g1)):
        rows = g.reshape(128, 39, 16384, 2).contiguous()
        want = ref.process_spectra(rows).cpu().numpy().reshape(-1)
        got = out[pol].reshape(-1)
        nd = int(np.count_nonzero(got != want))
        _say("[group-pk] pol %d: %d bytes differ" % (pol, nd))
        assert nd == 0
    assert np.array_equal(bm.cpu().numpy(), bm2.cpu().numpy())
    oracle = O.vdif_to_beng(pk_cut, b0)
    assert np.array_equal(bm.cpu().numpy(), oracle.bitmap)
    for key in ("n_unique", "n_missing", "n_duplicate", "n_invalid"):
        assert stats[key] == ref_stats[key] == getattr(oracle, key)
    assert oracle.n_missing > round(0.001 * 79872)
    st.close()
    ref.close()


def test_process_group_missing_frames_and_empty(torch_cuda, one_group):
    """Whole B-frames missing (frames 5 and 6: blocks 16 and 19 straddle them)
    and a group with no packets at all: sdbe_process_group equals stage A over
    the int8 rows of the same input (zeros where nothing arrived), byte for byte."""
    torch = torch_cuda
    from sdbe_amd import _lib as L
    _, pk_shuf, b0 = one_group
    w5 = pk_shuf[:, 5]
    bc = ((w5 >> 24) + ((pk_shuf[:, 4] & 0xFFFFFF) << 8)).astype(np.int64) - b0
    pk_cut = np.ascontiguousarray(pk_shuf[(bc != 5) & (bc != 6)])
    st = _stage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_PAYLOAD)
    ref = _stage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_PAYLOAD)
    d = st.to_device_packets(pk_cut)
    _say("[group-miss] gpu process_group without frames 5, 6 (n_pkts=%d)" % pk_cut.shape[0])
    out = st.process_group(d, pk_cut.shape[0], b0).cpu().numpy()
    g0, g1, _ = ref.unpack_int8(d, pk_cut.shape[0], b0)
    for pol, g in enumerate((g0, g1)):
        want = ref.process_spectra(g.reshape(128, 39, 16384, 2).contiguous()).cpu().numpy().reshape(-1)
        assert np.array_equal(out[pol].reshape(-1), want)
    _say("[group-miss] gpu process_group with no packets")
    empty = torch.zeros(1056, dtype=torch.uint8, device="cuda")
    out0 = st.process_group(empty, 0, b0).cpu().numpy()
    assert st.stats()["n_unique"] == 0
    want0 = ref.process_spectra(torch.zeros((128, 39, 16384, 2), dtype=torch.int8, device="cuda")).cpu().numpy()
    for pol in (0, 1):
        assert np.array_equal(out0[pol].reshape(-1), want0.reshape(-1))
    st.close()
    ref.close()


@pytest.mark.parametrize("shift", [4, 8])
def test_process_group_unaligned_packets(torch_cuda, one_group, shift):
    """A packet buffer that is not 16-byte aligned takes the int8-row path
    (stage A's packet reads are 16-byte units); the output is the same."""
    torch = torch_cuda
    from sdbe_amd import _lib as L
    _, pk_shuf, b0 = one_group
    st = _stage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_PAYLOAD)
    d = st.to_device_packets(pk_shuf)
    raw = torch.empty(d.numel() + shift, dtype=torch.uint8, device="cuda")
    raw[shift:].copy_(d)
    du = raw[shift:]
    assert du.data_ptr() % 16 == shift
    _say("[group-align] gpu process_group aligned / offset by %d bytes" % shift)
    want = st.process_group(d, pk_shuf.shape[0], b0).cpu().numpy()
    got = st.process_group(du, pk_shuf.shape[0], b0).cpu().numpy()
    assert np.array_equal(got, want)
    st.close()


def test_process_group_frames_headers(torch_cuda, one_group):
    from sdbe_amd import _lib as L
    pk, pk_shuf, b0 = one_group
    st = _stage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_FRAMES, station="Sm", bdc=1, rx=1)
    st.set_state(secs=100, df=120000, ref_epoch=30, psn=1)
    _say("[group-frames] gpu")
    out = st.process_group(st.to_device_packets(pk_shuf), pk_shuf.shape[0], b0).cpu().numpy()
    hdr = out[:, :, :32].copy().view(np.uint32).reshape(2, 8192, 8)
    state = O.StreamState(secs=100, df=120000, ref_epoch=30, psn=1,
                          station=O.station_from_str("Sm"), bdc=1, rx=1)
    ref_hdr, _ = O.stamp_headers(8192, state)
    assert np.array_equal(hdr, ref_hdr)
    # payloads equal the PAYLOAD-mode output
    st2 = _stage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_PAYLOAD)
    o2 = st2.process_group(st2.to_device_packets(pk_shuf), pk_shuf.shape[0], b0).cpu().numpy()
    assert np.array_equal(out[:, :, 32:], o2)
    st.close()
    st2.close()


def test_reframed_group_output_matches_output_thread(torch_cuda):
    """Device-side time alignment over 3 groups (first template valid, with a
    sample skip that is not a whole frame and a skip of frames to the next
    second) equals the output thread's re-framing of the group-aligned output:
    payloads, headers and frame counts."""
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import state_from_template
    _say("[align] generating 3 groups")
    groups, _ = S.make_group_packets(seed=52, n_groups=3, drop_frac=0.001)
    tmpl = O.fill_header_template(groups[0][0][0, :8], (100, 285_990_000))
    tmpl[1] = (tmpl[1] & ~np.uint32(0xFFFFFF)) | np.uint32(124_995)          # near the end of a second
    tmpl[6] = np.uint32(20_000_000)                                         # ps to skip
    inv = tmpl.copy()
    inv[0] |= np.uint32(1 << 31)
    ref = _stage(policy=L.THRESH_REFERENCE, out_mode=L.OUT_PAYLOAD)
    st = _stage(policy=L.THRESH_REFERENCE, out_mode=L.OUT_FRAMES, station="Sm")
    info = state_from_template(tmpl)
    assert info["skip_frames"] > 0 and info["skip_int32"] % 2048 != 0
    st.set_state(secs=info["secs"], df=info["df"], ref_epoch=info["ref_epoch"], psn=info["psn"])
    st.set_output_alignment(True, info["skip_int32"], info["skip_frames"])
    al = O.OutputAligner(station=O.station_from_str("Sm"))
    for g, (pk, b0) in enumerate(groups):
        _say("[align] group %d" % g)
        d = st.to_device_packets(pk)
        want_pay = ref.process_group(d, pk.shape[0], b0).cpu().numpy().view(np.uint32).reshape(2, -1)
        frames, headers = al.push(want_pay, tmpl if g == 0 else inv)
        out = st.process_group(d, pk.shape[0], b0).cpu().numpy()
        n = st.stats()["n_frames_per_pol"]
        assert n == frames.shape[1], (g, n, frames.shape)
        got_hdr = out[:, :n, :32].copy().view(np.uint32).reshape(2, n, 8)
        got_pay = out[:, :n, 32:].copy().view(np.uint32).reshape(2, n, 2048)
        assert np.array_equal(got_hdr, headers), g
        assert np.array_equal(got_pay, frames), g
    ref.close()
    st.close()
