"""FP64 numpy restatement of the SWARM data back-end GPU stage (APHIDS).

TEST INFRASTRUCTURE ONLY: this module is the checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it; the product (``sdbe_amd``) never does.

Every function cites the reference file:line it restates (paths relative to
the Smithsonian/sdbe tree).  Where the reference has undefined or racy
behaviour the oracle fixes one explicit policy and says so:

* duplicate (B-count, cid, fid) packets: the lowest packet index wins
  (the reference kernel lets the last racing writer win,
  software/aphids/vdif_inout_gpu_thread.cu:482-494);
* packets flagged invalid, or whose B-count lies outside the group window,
  are dropped (the reference input thread never inserts them,
  software/aphids/vdif_in_databuf.c:96-124; the kernel itself would index
  ``bcount % 39`` and could write out of bounds, vdif_inout_gpu_thread.cu:443);
* an fid >= 8 is dropped (the reference kernel would write out of bounds).

Pinning: see tests/test_oracle_*.py.  Each function there is checked against
closed forms, brute force on tiny inputs, independent bit-field layouts
(ctypes) or exhaustive enumeration, never against a retyped copy of itself.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

# ---------------------------------------------------------------------------
# Geometry (software/aphids/vdif_inout_gpu_thread.h:9-24, sample_rates.h:21-29,
# vdif_in_databuf.h:73,112,125, vdif_out_databuf.h:61,101)
# ---------------------------------------------------------------------------
BENG_CHANNELS = 16384            # BENG_CHANNELS_
BENG_SNAPSHOTS = 128             # snapshots per B-frame
SWARM_N_FIDS = 8                 # FID_PER_CHAN
PARALLEL_CHAN = 8                # SWARM_XENG_PARALLEL_CHAN
CHAN_PER_BENG = 256              # c-ids per B-frame
VDIF_PER_BENG_FRAME = 2048
PKT_BYTES = 1056
PKT_WORDS = PKT_BYTES // 4       # 264
HDR_WORDS = 8
DECIMATION = 39                  # DECIMATION_FACTOR / BENG_FRAMES_PER_GROUP (6/11 rate)
EXPANSION = 64                   # 64/39 resampling of resample.py:fft_batched
FFT_A = 2 * BENG_CHANNELS        # 32768-point C2R per snapshot
M_R2C = DECIMATION * FFT_A       # 1,277,952-point R2C per block
L_C2R = EXPANSION * FFT_A        # 2,097,152-point C2R per block
OUT_PAYLOAD_BYTES = 8192         # VDIF_OUT_PKT_DATA_SIZE
OUT_HDR_BYTES = 32
OUT_FRAME_BYTES = OUT_PAYLOAD_BYTES + OUT_HDR_BYTES
OUT_FRAMES_PER_BLOCK = L_C2R * 2 // 8 // OUT_PAYLOAD_BYTES   # 64
VDIF_OUT_FRAMES_PER_SECOND = 125000   # sample_rates.h:74
BLOCKS_PER_GROUP = DECIMATION * BENG_SNAPSHOTS // DECIMATION  # 128 (RESAMPLE_BATCH_ITERATIONS)

SWARM_CLK_PER_SEC = 286000000    # beng_reform.c:79
SWARM_BFRAME_CLK = 262144        # beng_reform.c:81
BFRAME_CLK_TOL = 1024            # beng_reform.c:83
MAX_COUNTABLE_BFRAME_OFFSET = 1000   # beng_reform.c:119
SWARM_ZERO_DELAY_OFFSET_PICOSECONDS = 3715791   # vdif_out_databuf.h:11
R2DBE_RATE = 4096e6              # sample_rates.h:18

# byte q of a snapshot's 8 payload bytes holds channel offset BYTE_TO_OFFSET[q]
# (vdif_inout_gpu_thread.cu:464-494: word 0 = channels d,c,b,a; word 1 = h,g,f,e)
BYTE_TO_OFFSET = np.array([3, 2, 1, 0, 7, 6, 5, 4], dtype=np.int64)


# ---------------------------------------------------------------------------
# A1/A2: VDIF-in header (vdif_in_databuf.h:29-59) and GPU-side parse
# (vdif_inout_gpu_thread.cu:353-369,437)
# ---------------------------------------------------------------------------
def in_header_fields(words):
    """Split 8 little-endian header words into the named bit fields of
    ``vdif_in_header_t`` (vdif_in_databuf.h:29-59).  GCC allocates bit-fields
    from the least significant bit on little-endian targets."""
    w = np.asarray(words, dtype=np.uint64)
    return dict(
        secs_inre=w[..., 0] & 0x3FFFFFFF, legacy=(w[..., 0] >> 30) & 1,
        invalid=(w[..., 0] >> 31) & 1,
        df_num_insec=w[..., 1] & 0xFFFFFF, ref_epoch=(w[..., 1] >> 24) & 0x3F,
        UA=(w[..., 1] >> 30) & 3,
        df_len=w[..., 2] & 0xFFFFFF, num_channels=(w[..., 2] >> 24) & 0x1F,
        ver=(w[..., 2] >> 29) & 7,
        stationID=w[..., 3] & 0xFFFF, threadID=(w[..., 3] >> 16) & 0x3FF,
        bps=(w[..., 3] >> 26) & 0x1F, dt=(w[..., 3] >> 31) & 1,
        b_upper=w[..., 4] & 0xFFFFFFFF,
        c=w[..., 5] & 0xFF, z=(w[..., 5] >> 8) & 0xFF, f=(w[..., 5] >> 16) & 0xFF,
        b_lower=(w[..., 5] >> 24) & 0xFF,
        edh_psn=(w[..., 6] & 0xFFFFFFFF) | ((w[..., 7] & 0xFFFFFFFF) << 32),
    )


def get_packet_b_count(words):
    """40-bit B-count, or -1 for invalid packets (vdif_in_databuf.c:66-74)."""
    f = in_header_fields(words)
    b = (f["b_upper"].astype(np.int64) << 8) | f["b_lower"].astype(np.int64)
    return np.where(f["invalid"] == 1, -1, b)


def gpu_cid_fid_rel(words, b_zero):
    """cid, fid and the frame offset ``(int32)(bcount - b_zero)`` exactly as the
    GPU kernel forms them (vdif_inout_gpu_thread.cu:353-369, 437): only the low
    24 bits of word 4 enter the B-count, the result is taken modulo 2**32 and
    reinterpreted as a signed 32-bit integer."""
    w = np.asarray(words, dtype=np.uint64)
    cid = (w[..., 5] & 0xFF).astype(np.int64)
    fid = ((w[..., 5] >> 16) & 0xFF).astype(np.int64)
    bc32 = ((w[..., 5] >> 24) + ((w[..., 4] & 0xFFFFFF) << 8)) & 0xFFFFFFFF
    diff = (bc32.astype(np.int64) - (int(b_zero) & 0xFFFFFFFF)) & 0xFFFFFFFF
    rel = diff.astype(np.uint32).view(np.int32).astype(np.int64)
    return cid, fid, rel


def channel_base(cid, fid):
    """First channel carried by packet (cid, fid) (vdif_inout_gpu_thread.cu:442)."""
    cid = np.asarray(cid, dtype=np.int64)
    fid = np.asarray(fid, dtype=np.int64)
    return ((cid >> 1) * 2 * PARALLEL_CHAN + (cid % 2) + fid * 2) * SWARM_N_FIDS


# ---------------------------------------------------------------------------
# A3: 2-bit decode (vdif_inout_gpu_thread.cu:387-392, swarm.py:45-52)
# ---------------------------------------------------------------------------
def decode_byte(b):
    """One payload byte -> (pol0_re, pol0_im, pol1_re, pol1_im) as int8 values
    ``(code & 3) - 2``; LSB first: pol1 im, pol1 re, pol0 im, pol0 re."""
    b = np.asarray(b, dtype=np.int64)
    p1_im = (b & 3) - 2
    p1_re = ((b >> 2) & 3) - 2
    p0_im = ((b >> 4) & 3) - 2
    p0_re = ((b >> 6) & 3) - 2
    return p0_re, p0_im, p1_re, p1_im


@dataclass
class UnpackResult:
    beng: np.ndarray            # int8 [2 pol][n_frames][128][16384][2 = (im, re)]
    bitmap: np.ndarray          # uint8 [n_frames][256], bit f of cid set if received
    n_invalid: int = 0
    n_out_of_range: int = 0
    n_duplicate: int = 0
    n_unique: int = 0
    n_missing: int = 0


# ---------------------------------------------------------------------------
# A4/A5/A6: unpack + scatter + missing-packet zero fill + flagging
# (vdif_inout_gpu_thread.cu:404-507 with the memset at :897-898;
#  flags as vdif_in_databuf.c:141-144 sets them)
# ---------------------------------------------------------------------------
def vdif_to_beng(pkts, b_zero, n_frames=DECIMATION):
    """Reference-layout unpack of one group of B-engine-over-VDIF packets.

    ``pkts`` is (n, 264) uint32.  Returns the two int8 buffers the reference
    kernel fills (``beng_0``/``beng_1``: [frame][snapshot][channel][im, re]) and
    the per-(frame, cid) fid bitmap."""
    pkts = np.ascontiguousarray(pkts, dtype=np.uint32)
    n = pkts.shape[0]
    hdr = in_header_fields(pkts[:, :HDR_WORDS])
    cid, fid, rel = gpu_cid_fid_rel(pkts[:, :HDR_WORDS], b_zero)
    invalid = hdr["invalid"] == 1
    in_range = (rel >= 0) & (rel < n_frames) & (fid < SWARM_N_FIDS)
    ok = (~invalid) & in_range
    idx_ok = np.nonzero(ok)[0]
    key = (rel[idx_ok] * CHAN_PER_BENG + cid[idx_ok]) * SWARM_N_FIDS + fid[idx_ok]
    _, first = np.unique(key, return_index=True)   # lowest packet index per key
    sel = idx_ok[first]

    beng = np.zeros((2, n_frames, BENG_SNAPSHOTS, BENG_CHANNELS, 2), dtype=np.int8)
    bitmap = np.zeros((n_frames, CHAN_PER_BENG), dtype=np.uint8)
    if sel.size:
        payload = pkts[sel, HDR_WORDS:].view(np.uint8).reshape(sel.size, BENG_SNAPSHOTS, 8)
        p0_re, p0_im, p1_re, p1_im = decode_byte(payload)
        chans = channel_base(cid[sel], fid[sel])[:, None, None] + BYTE_TO_OFFSET[None, None, :]
        fr = rel[sel][:, None, None]
        sn = np.arange(BENG_SNAPSHOTS)[None, :, None]
        beng[0, fr, sn, chans, 0] = p0_im
        beng[0, fr, sn, chans, 1] = p0_re
        beng[1, fr, sn, chans, 0] = p1_im
        beng[1, fr, sn, chans, 1] = p1_re
        np.bitwise_or.at(bitmap, (rel[sel], cid[sel]), (1 << fid[sel]).astype(np.uint8))
    return UnpackResult(
        beng=beng, bitmap=bitmap,
        n_invalid=int(invalid.sum()),
        n_out_of_range=int(((~invalid) & (~in_range)).sum()),
        n_duplicate=int(idx_ok.size - sel.size),
        n_unique=int(sel.size),
        n_missing=int(n_frames * VDIF_PER_BENG_FRAME - sel.size),
    )


# The "8 snapshots x 1024 channels" packet geometry of the benchmark's
# configuration text (no reference counterpart; the stage's SDBE_GEOM_8X1024).
# Packet (frame, cid, fid), cid < 16, fid < 16: snapshots 8 fid .. 8 fid + 7 of
# its B-frame, channels 1024 cid .. 1024 cid + 1023; payload row sn is 1024
# bytes, byte q holding channel (q & ~3) + 3 - (q & 3) (the reference's order
# within a 32-bit word), each byte decoded by decode_byte.  Header as the
# reference's (cid, fid and the B-count in words 4-5).
G2_CIDS, G2_FIDS, G2_SNAPS, G2_CHANS = 16, 16, 8, 1024
G2_PKT_WORDS = HDR_WORDS + G2_SNAPS * G2_CHANS // 4
G2_BYTE_TO_CHAN = (np.arange(G2_CHANS) & ~3) + 3 - (np.arange(G2_CHANS) & 3)


def vdif_to_beng_8x1024(pkts, b_zero, n_frames=DECIMATION):
    """Unpack one group of 8 x 1024 packets ((n, 2056) uint32) into the same
    int8 buffers as ``vdif_to_beng``; the lowest-index packet of a slot wins,
    missing slots stay 0 + 0j.  bitmap: uint16 [frame][cid], bit fid."""
    pkts = np.ascontiguousarray(pkts, dtype=np.uint32)
    hdr = in_header_fields(pkts[:, :HDR_WORDS])
    cid, fid, rel = gpu_cid_fid_rel(pkts[:, :HDR_WORDS], b_zero)
    invalid = hdr["invalid"] == 1
    in_range = (rel >= 0) & (rel < n_frames) & (cid < G2_CIDS) & (fid < G2_FIDS)
    ok = (~invalid) & in_range
    idx_ok = np.nonzero(ok)[0]
    key = (rel[idx_ok] * G2_CIDS + cid[idx_ok]) * G2_FIDS + fid[idx_ok]
    _, first = np.unique(key, return_index=True)
    sel = idx_ok[first]
    beng = np.zeros((2, n_frames, BENG_SNAPSHOTS, BENG_CHANNELS, 2), dtype=np.int8)
    bitmap = np.zeros((n_frames, G2_CIDS), dtype=np.uint16)
    for p in sel:
        payload = pkts[p, HDR_WORDS:].view(np.uint8).reshape(G2_SNAPS, G2_CHANS)
        p0_re, p0_im, p1_re, p1_im = decode_byte(payload)
        chans = G2_CHANS * cid[p] + G2_BYTE_TO_CHAN
        rows = G2_SNAPS * fid[p] + np.arange(G2_SNAPS)
        f = rel[p]
        beng[0, f, rows[:, None], chans[None, :], 0] = p0_im
        beng[0, f, rows[:, None], chans[None, :], 1] = p0_re
        beng[1, f, rows[:, None], chans[None, :], 0] = p1_im
        beng[1, f, rows[:, None], chans[None, :], 1] = p1_re
        bitmap[f, cid[p]] |= np.uint16(1 << int(fid[p]))
    return UnpackResult(
        beng=beng, bitmap=bitmap,
        n_invalid=int(invalid.sum()),
        n_out_of_range=int(((~invalid) & (~in_range)).sum()),
        n_duplicate=int(idx_ok.size - sel.size),
        n_unique=int(sel.size),
        n_missing=int(n_frames * G2_CIDS * G2_FIDS - sel.size),
    )


def inflate(beng_pol):
    """int8 (im, re) pairs -> complex spectra with the zeroed Nyquist channel
    16384 appended (vdif_inout_gpu_thread.cu:532-556 and the memset at :925-926)."""
    beng_pol = np.asarray(beng_pol)
    X = np.zeros(beng_pol.shape[:-2] + (BENG_CHANNELS + 1,), dtype=np.complex128)
    X[..., :BENG_CHANNELS] = beng_pol[..., 1].astype(np.float64) + 1j * beng_pol[..., 0]
    return X


def flip_sideband(X):
    """Reverse the channel order of a lower-sideband stream before synthesis
    (rslib.py:124 reverses the LSB spectrum).  Channels 0..16383 are reversed;
    the appended Nyquist slot stays zero."""
    Y = np.zeros_like(X)
    Y[..., :BENG_CHANNELS] = X[..., BENG_CHANNELS - 1::-1]
    return Y


# ---------------------------------------------------------------------------
# A8-A11: the 64/39 Fourier resampler, FP64, numpy normalisation
# (resample.py:128-134, fft_test.py:118-121; stage roles as in
#  vdif_inout_gpu_thread.cu:589-709).  The reference's cuFFT path is
#  unnormalised: its samples equal these times FFT_A * L_C2R (resample.py:138).
# ---------------------------------------------------------------------------
def stage_a(X):
    """Inverse real DFT of each 16385-bin snapshot to 32768 samples.  The
    imaginary parts of bin 0 and bin 16384 are discarded (numpy irfft)."""
    return np.fft.irfft(X, n=FFT_A, axis=-1)


def stage_b(x_concat):
    """Forward real DFT of one block's concatenated time series
    (N = 39 * 32768 = 1,277,952, 638,977 bins)."""
    return np.fft.rfft(x_concat, axis=-1)


def stage_c(Y):
    """Zero-pad the 638,977 bins to 1,048,577 (old Nyquist copied as is, not
    halved) and inverse real DFT to 2,097,152 samples at 4096 Msps."""
    pad = np.zeros(Y.shape[:-1] + (L_C2R // 2 + 1,), dtype=np.complex128)
    pad[..., :Y.shape[-1]] = Y
    return np.fft.irfft(pad, n=L_C2R, axis=-1)


def resample_block(X39):
    """(39, 16385) complex spectra -> 2,097,152 real samples (FP64)."""
    X39 = np.asarray(X39)
    assert X39.shape[-2:] == (DECIMATION, BENG_CHANNELS + 1)
    xa = stage_a(X39)
    y = xa.reshape(X39.shape[:-2] + (M_R2C,))
    return stage_c(stage_b(y))


# ---------------------------------------------------------------------------
# A12: statistics (vdif_inout_gpu_thread.cu:60-137 Chan/Welford combine,
# used at :953-986): offset = mean, threshold = sqrt(M2 / (n - 1)).
# ---------------------------------------------------------------------------
def block_stats(x):
    x = np.asarray(x, dtype=np.float64)
    mean = float(x.mean())
    m2 = float(((x - mean) ** 2).sum())
    return mean, math.sqrt(m2 / (x.size - 1))


def chan_combine(a, b):
    """Parallel combine of two (n, mean, M2) triples
    (vdif_inout_gpu_thread.cu:119-136)."""
    n = a[0] + b[0]
    delta = b[1] - a[1]
    mean = a[1] + delta * b[0] / n
    m2 = a[2] + b[2] + delta * delta * a[0] * b[0] / n
    return (n, mean, m2)


# ---------------------------------------------------------------------------
# A13: 2-bit quantisation and packing (vdif_inout_gpu_thread.cu:565-583)
# ---------------------------------------------------------------------------
def quantize2bit(x, th, off):
    """code = (((|v| >= th) | 2) ^ (3 * (v < 0))) & 3 with v = x - off.
    v >= th -> 3, 0 <= v < th -> 2, -th < v < 0 -> 1, v <= -th -> 0; -0.0 and
    NaN count as non-negative."""
    v = np.asarray(x, dtype=np.float64) - off
    hi = (np.abs(v) >= th).astype(np.uint8)
    neg = (v < 0).astype(np.uint8)
    return (((hi | 2) ^ (3 * neg)) & 3).astype(np.uint8)


def quantize_to_2bit_python(xs, th):
    """The prototype's CPU quantiser (sdbe_preprocess.py:668-692), values in
    {-2,-1,0,1}.  Differs from ``quantize2bit`` only at x == -th exactly."""
    xs = np.asarray(xs)
    out = np.zeros(xs.size, dtype=np.int8)
    out[xs < -th] = -2
    out[(xs >= -th) & (xs < 0)] = -1
    out[(xs >= 0) & (xs < th)] = 0
    out[xs >= th] = 1
    return out


def pack_codes(codes):
    """16 codes per little-endian uint32, sample i at bits 2i
    (vdif_inout_gpu_thread.cu:578-580, vdif.py:224-240)."""
    c = np.asarray(codes, dtype=np.uint32).reshape(-1, 16)
    shifts = (2 * np.arange(16, dtype=np.uint32))[None, :]
    return np.bitwise_or.reduce(c << shifts, axis=1).astype(np.uint32)


def unpack_codes(words):
    w = np.asarray(words, dtype=np.uint32).reshape(-1, 1)
    return ((w >> (2 * np.arange(16, dtype=np.uint32))[None, :]) & 3).astype(np.uint8).reshape(-1)


# ---------------------------------------------------------------------------
# Threshold policies.  REFERENCE: (th, off) of the first block-pol (block 0,
# pol 0) latched for the whole stream and applied to both polarisations,
# recomputed while th == 0 (vdif_inout_gpu_thread.cu:954-1005).
# PER_BLOCK: each block-pol its own mean and ddof=1 sigma (per-block thresholds).
# ---------------------------------------------------------------------------
THRESH_REFERENCE = 0
THRESH_PER_BLOCK = 1
THRESH_FIXED = 2


# ---------------------------------------------------------------------------
# A14/A15: VDIF-out header (vdif_out_databuf.h:19-52), defaults
# (vdif_out_databuf.c:35-67) and per-packet stamping
# (vdif_out_net_thread.c:294-317)
# ---------------------------------------------------------------------------
def out_header_words(secs, df, ref_epoch, station, pol, bdc, rx, psn,
                     df_len=OUT_FRAME_BYTES // 8, bps=1, edv5=2, eud5=0,
                     thread=0, nchan=0, ver=0, legacy=0, invalid=0, ua=0, dt=0, pic=0):
    w = np.zeros(8, dtype=np.uint32)
    w[0] = (secs & 0x3FFFFFFF) | ((legacy & 1) << 30) | ((invalid & 1) << 31)
    w[1] = (df & 0xFFFFFF) | ((ref_epoch & 0x3F) << 24) | ((ua & 3) << 30)
    w[2] = (df_len & 0xFFFFFF) | ((nchan & 0x1F) << 24) | ((ver & 7) << 29)
    w[3] = (station & 0xFFFF) | ((thread & 0x3FF) << 16) | ((bps & 0x1F) << 26) | ((dt & 1) << 31)
    w[4] = (pol & 1) | ((bdc & 1) << 1) | ((rx & 1) << 2) | ((eud5 & 0x1FFFFF) << 3) | ((edv5 & 0xFF) << 24)
    w[5] = pic & 0xFFFFFFFF
    w[6] = psn & 0xFFFFFFFF
    w[7] = (psn >> 32) & 0xFFFFFFFF
    return w


def station_from_str(s):
    """Two-character station code -> 16-bit field (vdif_out_net_thread.c:98)."""
    return ((ord(s[0]) & 0xFF) << 8) | (ord(s[1]) & 0xFF)


@dataclass
class StreamState:
    secs: int = 0
    df: int = 0
    ref_epoch: int = 0
    psn: int = 1
    station: int = station_from_str("XX")
    bdc: int = 0
    rx: int = 0


def stamp_headers(n_frames, st: StreamState, bit_depth=2):
    """Headers of ``n_frames`` consecutive output frames for both
    polarisations; returns (headers [2][n][8] uint32, advanced state).  Both
    polarisations of frame ii share secs/df/psn; df wraps at 125000 with
    secs++ (vdif_out_net_thread.c:295-317)."""
    hdr = np.zeros((2, n_frames, 8), dtype=np.uint32)
    secs, df, psn = st.secs, st.df, st.psn
    for ii in range(n_frames):
        for jj in range(2):
            hdr[jj, ii] = out_header_words(secs, df, st.ref_epoch, st.station, jj,
                                           st.bdc, st.rx, psn, bps=bit_depth - 1)
        df += 1
        if df == VDIF_OUT_FRAMES_PER_SECOND:
            df = 0
            secs += 1
        psn += 1
    new = StreamState(secs=secs, df=df, ref_epoch=st.ref_epoch, psn=psn,
                      station=st.station, bdc=st.bdc, rx=st.rx)
    return hdr, new


# ---------------------------------------------------------------------------
# A7: B-engine timestamp reform (beng_reform.c:16-260) and 8pac unpack
# (vdif_8pac.c:5-56)
# ---------------------------------------------------------------------------
def read_beng_timestamp(w4, w5):
    """New-format 53-bit timestamp: clk = low 29 bits, sec above
    (beng_reform.c:16-22)."""
    b = (int(w4) << 21) | ((int(w5) >> 11) & 0x1FFFFF)
    clk = b & 0x1FFFFFFF
    sec = (b & 0x001FFFFFE0000000) >> 29
    sec = sec - (1 << 32) if sec >= (1 << 31) else sec   # int32 cast
    return [sec, clk]


def ts_compare(then, now):
    """beng_reform.c:34-61, restated branch for branch."""
    if then[0] < now[0]:
        if then[0] == now[0] - 1:
            if then[1] + BFRAME_CLK_TOL > SWARM_CLK_PER_SEC and \
                    (then[1] + BFRAME_CLK_TOL) % SWARM_CLK_PER_SEC >= now[1]:
                return 0
        return -1
    elif then[0] > now[0]:
        if now[1] + BFRAME_CLK_TOL > SWARM_CLK_PER_SEC and \
                (now[1] + BFRAME_CLK_TOL) % SWARM_CLK_PER_SEC >= then[1]:
            return 0
        return 1
    else:
        if then[1] + BFRAME_CLK_TOL < now[1]:
            return -1
        elif now[1] + BFRAME_CLK_TOL < then[1]:
            return 1
        return 0


def ts_increment(t):
    """beng_reform.c:246-252 (note the strict ``>``)."""
    t = [t[0], t[1] + SWARM_BFRAME_CLK]
    if t[1] > SWARM_CLK_PER_SEC:
        t[0] += 1
        t[1] = t[1] % SWARM_CLK_PER_SEC
    return t


def ts_decrement(t):
    """beng_reform.c:254-260."""
    t = [t[0], t[1] - SWARM_BFRAME_CLK]
    if t[1] < 0:
        t[0] -= 1
        t[1] = t[1] + SWARM_CLK_PER_SEC
    return t


def ts_offset(then, now):
    """B-frame offset of ``now`` relative to ``then``, capped just beyond
    MAX_COUNTABLE_BFRAME_OFFSET (beng_reform.c:71-104)."""
    cmp = ts_compare(then, now)
    if cmp < 0:
        start, end, step = list(then), list(now), 1
    elif cmp > 0:
        start, end, step = list(now), list(then), -1
    else:
        return 0
    off = 0
    while ts_compare(start, end) < 0:
        start = ts_increment(start)
        off += step
        if off > MAX_COUNTABLE_BFRAME_OFFSET or -off > MAX_COUNTABLE_BFRAME_OFFSET:
            break
    return off


class BengReformer:
    """Stateful restatement of beng_reform_headers (beng_reform.c:137-186):
    rewrites new-format headers (53-bit timestamp) into the old format whose
    40-bit B-count is the B-frame offset from the first valid packet (T0)."""

    def __init__(self):
        self.T0 = [-1, -1]
        self.TPREV = [-1, -1]
        self.BPREV = -1

    def beng_t0(self):
        """get_beng_t0 (beng_reform.c:113-126): T0 moved back one B-frame
        (SWARM_TIMESTAMP_OFF_BY_BENG = -1)."""
        return ts_decrement(list(self.T0))

    def reform(self, hdr_words):
        """In place on an (n, >=8) uint32 array of packet words."""
        for ii in range(hdr_words.shape[0]):
            w0 = int(hdr_words[ii, 0])
            if (w0 >> 31) & 1:
                continue
            w4, w5 = int(hdr_words[ii, 4]), int(hdr_words[ii, 5])
            if self.T0 == [-1, -1]:
                self.T0 = read_beng_timestamp(w4, w5)
                self.TPREV = list(self.T0)
                self.BPREV = 0
            t_now = read_beng_timestamp(w4, w5)
            off = ts_offset(self.TPREV, t_now)
            if off > MAX_COUNTABLE_BFRAME_OFFSET or off < -MAX_COUNTABLE_BFRAME_OFFSET:
                w0 |= 1 << 31
                hdr_words[ii, 0] = w0
                off = -1
            else:
                off += self.BPREV
                self.TPREV = list(t_now)
                self.BPREV = off
            c = w5 & 0xFF
            f = (w5 >> 8) & 0x7
            off32 = off & 0xFFFFFFFF
            # old layout: byte0 = c, byte1 (z) untouched, byte2 = f, byte3 = b_lower
            w5_new = (c & 0xFF) | (w5 & 0x0000FF00) | ((f & 0xFF) << 16) | ((off32 & 0xFF) << 24)
            hdr_words[ii, 5] = w5_new
            hdr_words[ii, 4] = (off32 & 0x7FFFFF00) >> 8
        return hdr_words


def unpack_8pac(raw8):
    """8pac aggregates -> 8 plain packets each (vdif_8pac.c:5-56).  ``raw8`` is
    a uint8 array (n, 8504): pkt1 (1056 B) then 7 x (8-byte VTP + 1056 B)."""
    raw8 = np.asarray(raw8, dtype=np.uint8)
    n = raw8.shape[0]
    out = np.zeros((n * 8, PKT_WORDS), dtype=np.uint32)
    for a in range(n):
        pk = [raw8[a, 0:PKT_BYTES].copy().view(np.uint32)]
        for i in range(7):
            s = PKT_BYTES + i * (8 + PKT_BYTES) + 8
            pk.append(raw8[a, s:s + PKT_BYTES].copy().view(np.uint32))
        w0, w1, w2 = int(pk[0][0]), int(pk[0][1]), int(pk[0][2])
        df_len = w2 & 0xFFFFFF
        q = df_len - 7
        newsize = (abs(q) // 8 * (1 if q >= 0 else -1)) & 0xFFFFFF    # C truncation
        df8 = (8 * (w1 & 0xFFFFFF)) & 0xFFFFFFFF
        psn = int(pk[0][6]) | (int(pk[0][7]) << 32)
        psn8 = (8 * psn) & 0xFFFFFFFFFFFFFFFF
        inval = (w0 >> 31) & 1
        for i in range(8):
            p = pk[i]
            p[2] = (int(p[2]) & 0xFF000000) | newsize
            dfi = (df8 + i) & 0xFFFFFF
            p[1] = (int(p[1]) & 0xFF000000) | dfi
            ps = (psn8 + i) & 0xFFFFFFFFFFFFFFFF
            p[6] = ps & 0xFFFFFFFF
            p[7] = ps >> 32
            if i > 0:
                p[0] = (int(p[0]) & 0x7FFFFFFF) | (inval << 31)
            out[a * 8 + i] = p
    return out


# ---------------------------------------------------------------------------
# A5/A6 host bookkeeping: group window, insert, completion
# (vdif_in_databuf.c:76-167); A15 template (vdif_in_databuf.c:193-220) and
# output time alignment (vdif_out_net_thread.c:203-237)
# ---------------------------------------------------------------------------
BENG_GROUPS_IN_BUFFER = 4
ERR_INVALID = -1
ERR_BEFORE_START = -2
ERR_TOO_FAR_AHEAD = -3


@dataclass
class GroupCompletion:
    b: list = field(default_factory=list)                 # bfc[i].b
    frame_count: list = field(default_factory=list)       # bfc[i].beng_frame_vdif_packet_count
    cc: np.ndarray = None                                 # uint8 [39][256]
    packet_count: int = 0
    packets: list = field(default_factory=list)           # inserted packet indices in order


def init_beng_group(b_start, n_frames=DECIMATION):
    return GroupCompletion(b=[b_start + i for i in range(n_frames)],
                           frame_count=[0] * n_frames,
                           cc=np.zeros((n_frames, CHAN_PER_BENG), dtype=np.uint8))


def group_index_offset(groups, index_ref, words):
    b = int(get_packet_b_count(words))
    invalid = (int(words[0]) >> 31) & 1
    ll = groups[index_ref].b[0]
    lu = groups[index_ref].b[-1]
    n_frames = len(groups[index_ref].b)
    if b < ll or invalid:
        return ERR_INVALID if invalid else ERR_BEFORE_START
    offset = -1
    for _ in range(BENG_GROUPS_IN_BUFFER):
        offset += 1
        if ll <= b <= lu:
            return offset
        ll += n_frames
        lu += n_frames
    if b > lu + BENG_GROUPS_IN_BUFFER * n_frames:
        return ERR_TOO_FAR_AHEAD
    return offset


def insert_packet(groups, index_ref, offset, words, pkt_index):
    b = int(get_packet_b_count(words))
    n_frames = len(groups[index_ref].b)
    count = 0
    for ii in range(offset, BENG_GROUPS_IN_BUFFER):
        g = groups[(ii + index_ref) % BENG_GROUPS_IN_BUFFER]
        if g.b[0] <= b <= g.b[-1]:
            bi = b - g.b[0]
            c = int(words[5]) & 0xFF
            f = (int(words[5]) >> 16) & 0xFF
            if g.frame_count[bi] < VDIF_PER_BENG_FRAME and \
                    g.packet_count < n_frames * VDIF_PER_BENG_FRAME:
                g.frame_count[bi] += 1
                g.cc[bi, c] |= (1 << f) & 0xFF
                g.packets.append(pkt_index)
                g.packet_count += 1
                count += 1
    return count


def group_complete(g):
    return 1 if g.packet_count == len(g.b) * VDIF_PER_BENG_FRAME else 0


def fill_header_template(ref_words, T0):
    """Header template of the first output data (vdif_in_databuf.c:193-220):
    the reference packet's header with secs from T0 (moved back then forward
    one B-frame, get_beng_t0 + beng_timestamp_increment), df = ceil of the
    frame position, and the picoseconds before that frame boundary hidden in
    edh_psn.  Returns the 8 header words."""
    w = np.array(ref_words[:8], dtype=np.uint32).copy()
    t0 = ts_increment(ts_decrement(list(T0)))
    w[0] = (int(w[0]) & 0xC0000000) | (t0[0] & 0x3FFFFFFF)
    frame_full = (VDIF_OUT_FRAMES_PER_SECOND * 1.0) * (float(t0[1]) / float(SWARM_CLK_PER_SEC))
    frame_ceil = math.ceil(frame_full)
    frame_frac = frame_ceil - frame_full
    w[1] = (int(w[1]) & 0xFF000000) | (int(frame_ceil) & 0xFFFFFF)
    psn = int(math.floor(frame_frac * (1e12 / VDIF_OUT_FRAMES_PER_SECOND)))
    w[6] = psn & 0xFFFFFFFF
    w[7] = (psn >> 32) & 0xFFFFFFFF
    return w


@dataclass
class Alignment:
    secs: int
    df: int
    ref_epoch: int
    psn: int
    skip_frames: int
    skip_int32: int
    num_usable: int
    leftover_int32: int


def output_alignment(template_words, n_words_per_chan, bit_depth=2):
    """Start-of-stream alignment the output thread derives from a valid
    template (vdif_out_net_thread.c:203-237)."""
    f = in_header_fields(np.asarray(template_words, dtype=np.uint32))
    num_usable = n_words_per_chan // (OUT_PAYLOAD_BYTES // 4)
    secs = int(f["secs_inre"])
    df = int(f["df_num_insec"])
    skip_frames = 0
    if df > VDIF_OUT_FRAMES_PER_SECOND - num_usable:
        skip_frames = VDIF_OUT_FRAMES_PER_SECOND - df
        df = 0
        secs += 1
    ref_epoch = int(f["ref_epoch"])
    skip_ps = int(f["edh_psn"]) + SWARM_ZERO_DELAY_OFFSET_PICOSECONDS
    skip_samples = int(math.floor(float(skip_ps) * R2DBE_RATE / 1e12))
    skip_int32 = int(math.floor(float(skip_samples) / (8.0 / float(bit_depth)) / 4.0))
    num_usable = (n_words_per_chan - skip_int32) // (OUT_PAYLOAD_BYTES // 4)
    num_usable -= skip_frames
    leftover = (n_words_per_chan - skip_int32 - num_usable * (OUT_PAYLOAD_BYTES // 4)
                - skip_frames * (OUT_PAYLOAD_BYTES // 4))
    return Alignment(secs=secs, df=df, ref_epoch=ref_epoch, psn=1,
                     skip_frames=skip_frames, skip_int32=skip_int32,
                     num_usable=num_usable, leftover_int32=leftover)


# ---------------------------------------------------------------------------
# Whole stage, oracle form: one group of packets -> VDIF frames per pol
# ---------------------------------------------------------------------------
def block_spectra(beng_pol, block):
    """The 39 complex spectra of block ``block`` of one polarisation; blocks
    are consecutive runs of 39 snapshots in frame-major snapshot order
    (vdif_inout_gpu_thread.cu:931-937)."""
    snaps = beng_pol.reshape(-1, BENG_CHANNELS, 2)
    return inflate(snaps[block * DECIMATION:(block + 1) * DECIMATION])


@dataclass
class GroupOutput:
    samples: dict                # (block, pol) -> FP64 samples (numpy units)
    stats: dict                  # (block, pol) -> (mean, sigma)
    thresholds: dict             # (block, pol) -> (th, off) used
    codes: dict                  # (block, pol) -> uint8 codes
    payload: np.ndarray          # uint32 [2][n_blocks*64][2048]
    headers: np.ndarray          # uint32 [2][n_blocks*64][8]


def process_blocks(beng, blocks, policy=THRESH_PER_BLOCK, latch=None, fixed=(1.0, 0.0),
                   state: StreamState | None = None, flip=False):
    """Resample, threshold, quantise and pack the given blocks of one group or
    stream (both pols).  ``latch`` is a one-element list holding the
    REFERENCE-policy (th, off) across calls.  ``flip`` reverses the channel
    order of every spectrum first (lower sideband, rslib.py:124)."""
    if latch is None:
        latch = [None]
    samples, stats, ths, codes = {}, {}, {}, {}
    nb = len(blocks)
    payload = np.zeros((2, nb * OUT_FRAMES_PER_BLOCK, OUT_PAYLOAD_BYTES // 4), dtype=np.uint32)
    for bi, blk in enumerate(blocks):
        for pol in (0, 1):
            X = block_spectra(beng[pol], blk)
            x = resample_block(flip_sideband(X) if flip else X)
            samples[(blk, pol)] = x
            stats[(blk, pol)] = block_stats(x)
        for pol in (0, 1):
            if policy == THRESH_PER_BLOCK:
                mean, sig = stats[(blk, pol)]
                th, off = sig, mean
            elif policy == THRESH_REFERENCE:
                if latch[0] is None or latch[0][0] == 0:
                    mean, sig = stats[(blk, 0)]
                    latch[0] = (sig, mean)
                th, off = latch[0]
            else:
                th, off = fixed
            ths[(blk, pol)] = (th, off)
            c = quantize2bit(samples[(blk, pol)], th, off)
            codes[(blk, pol)] = c
            payload[pol, bi * OUT_FRAMES_PER_BLOCK:(bi + 1) * OUT_FRAMES_PER_BLOCK] = \
                pack_codes(c).reshape(OUT_FRAMES_PER_BLOCK, -1)
    hdr = None
    if state is not None:
        hdr, _ = stamp_headers(nb * OUT_FRAMES_PER_BLOCK, state)
    return GroupOutput(samples=samples, stats=stats, thresholds=ths, codes=codes,
                       payload=payload, headers=hdr)


# ---------------------------------------------------------------------------
# CPU baseline chain (the reference's numpy/FFTW CPU path, vectorised;
# rslib.py:84-169, resample.py:128-134, sdbe_preprocess.py:668-692,
# vdif.py:98-242).  scipy.fft (pocketfft) stands in for FFTW, parallel over
# the block-pols of a batch.
# ---------------------------------------------------------------------------
def cpu_resample_batch(X, workers=1):
    """X: (n, 39, 16385) complex spectra -> (n, 2,097,152) FP64 samples."""
    import scipy.fft as F
    n = X.shape[0]
    xa = F.irfft(X, n=FFT_A, axis=-1, workers=workers)
    Y = F.rfft(xa.reshape(n, M_R2C), axis=-1, workers=workers)
    pad = np.zeros((n, L_C2R // 2 + 1), dtype=np.complex128)
    pad[:, :Y.shape[1]] = Y
    return F.irfft(pad, n=L_C2R, axis=-1, workers=workers)


def cpu_chain(pkts, b_zero, blocks, workers=1, state=None):
    """Packets of one group -> per-block-pol VDIF payload words and headers
    for the given blocks (PER_BLOCK thresholds).  Returns (payload, headers)."""
    un = vdif_to_beng(pkts, b_zero)
    X = np.stack([block_spectra(un.beng[pol], blk) for blk in blocks for pol in (0, 1)])
    x = cpu_resample_batch(X, workers)
    words = []
    for i in range(x.shape[0]):
        mean, sig = block_stats(x[i])
        words.append(pack_codes(quantize2bit(x[i], sig, mean)))
    hdr, _ = stamp_headers(len(blocks) * OUT_FRAMES_PER_BLOCK, state or StreamState())
    return np.stack(words), hdr


# ---------------------------------------------------------------------------
# A17: polyphase filterbank and its least-squares inverse
# (software/prototyping/cuda/pfb_inverse.cu; the inverse follows the published
# method the file names -- Richard Shaw's pfb-inverse -- whose solve the CUDA
# file leaves compiled out, :392-455).
# ---------------------------------------------------------------------------
PFB_LBLOCK = 2 * BENG_CHANNELS      # 32768 samples per block (num_freq 16385)
PFB_NTAP = 4


def pfb_window_hamming(ntap=PFB_NTAP, lblock=PFB_LBLOCK):
    """The reference window over ntap * lblock samples
    (pfb_inverse.cu:51-53, used at :193 and :271)."""
    m = ntap * lblock
    n = np.arange(m, dtype=np.float64)
    return 0.54 - 0.46 * np.cos(2.0 * np.pi * n / (m - 1.0))


def pfb_window_hann_sinc(ntap=PFB_NTAP, lblock=PFB_LBLOCK):
    """Hann-windowed sinc prototype: h[n] = w_Hann[n] sinc((n - (L-1)/2) / lblock)."""
    m = ntap * lblock
    n = np.arange(m, dtype=np.float64)
    return (0.5 - 0.5 * np.cos(2.0 * np.pi * n / (m - 1.0))) * np.sinc((n - (m - 1) / 2.0) / lblock)


def pfb_forward(x, window, lblock=PFB_LBLOCK):
    """Spectra of block i: rfft of the window times samples
    [i lblock, i lblock + ntap lblock), keeping every ntap-th bin
    (pfb_inverse.cu:172-208).  Computed here as the rfft of the ntap-fold
    sum, which equals those bins.  Returns (nblock, lblock/2 + 1) complex."""
    x = np.asarray(x, dtype=np.float64)
    ntap = window.size // lblock
    nblock = x.size // lblock - (ntap - 1)
    w = window.reshape(ntap, lblock)
    seg = np.stack([x[p * lblock:(p + nblock) * lblock].reshape(nblock, lblock) for p in range(ntap)])
    folded = np.einsum("pil,pl->il", seg, w)
    return np.fft.rfft(folded, axis=-1)


def pfb_toeplitz(window, lblock=PFB_LBLOCK):
    """a[d][n] = sum_k w[n + lblock (k + d)] w[n + lblock k]: the diagonals of
    P P^T for branch n, P[i][j] = w[n + lblock (j - i)] for 0 <= j - i < ntap
    (pfb_inverse.cu:286-309)."""
    ntap = window.size // lblock
    w = window.reshape(ntap, lblock).astype(np.float64)
    return np.stack([(w[d:] * w[:ntap - d]).sum(axis=0) for d in range(ntap)])


def pfb_inverse(spec, window, lblock=PFB_LBLOCK):
    """Least-squares (minimum-norm) inverse: yh_i = irfft(spec_i) (the
    pseudo-timestream, pfb_inverse.cu:355-358 with the 1/lblock scaling of
    :377), then per branch n solve (P P^T) z = yh[:, n] and project
    x[j lblock + n] = (P^T z)[j] = sum_p w[n + lblock p] z[j - p]
    (pfb_inverse.cu:380-400).  The banded SPD system is solved by Cholesky,
    vectorised over the branches.  Returns ((nblock + ntap - 1) * lblock,)."""
    yh = np.fft.irfft(np.asarray(spec), n=lblock, axis=-1)
    nblock = yh.shape[0]
    ntap = window.size // lblock
    a = pfb_toeplitz(window, lblock)
    w = window.reshape(ntap, lblock).astype(np.float64)
    b = ntap - 1
    L = np.zeros((nblock, ntap, lblock))         # L[i][d] = L(i, i - d)
    u = np.zeros((nblock, lblock))
    for i in range(nblock):
        for d in range(b, 0, -1):
            if i - d < 0:
                continue
            acc = a[d].copy()
            for k in range(d + 1, b + 1):
                if i - k >= 0:
                    acc -= L[i][k] * L[i - d][k - d]
            L[i][d] = acc / L[i - d][0]
        diag = a[0] - sum(L[i][d] ** 2 for d in range(1, b + 1))
        L[i][0] = np.sqrt(diag)
        r = yh[i].copy()
        for d in range(1, b + 1):
            if i - d >= 0:
                r -= L[i][d] * u[i - d]
        u[i] = r / L[i][0]
    z = np.zeros((nblock, lblock))
    for i in range(nblock - 1, -1, -1):
        r = u[i].copy()
        for d in range(1, b + 1):
            if i + d < nblock:
                r -= L[i + d][d] * z[i + d]
        z[i] = r / L[i][0]
    nts = nblock + ntap - 1
    x = np.zeros((nts, lblock))
    for p in range(ntap):
        x[p:p + nblock] += w[p] * z
    return x.reshape(-1)


class OutputAligner:
    """The output thread's re-framing of the stage's per-group payload stream
    (vdif_out_net_thread.c:198-317): on a group whose header template is
    valid, derive secs / df / psn and the sample skip (output_alignment); the
    frames sent are the first num_usable frames of [leftover of the previous
    group] + [this group's words from skip_int32 + 2048 skip_frames on]; the
    last leftover_int32 words of the group are kept for the next one, and
    leftover_int32 keeps the value of the last valid template.  Frames are
    stamped as the output thread does (df wraps at 125000 with secs + 1,
    psn + 1, pol = channel)."""

    WPF = OUT_PAYLOAD_BYTES // 4         # 2048 words per frame

    def __init__(self, station=station_from_str("XX"), bdc=0, rx=0, bit_depth=2):
        self.leftover = [None, None]
        self.leftover_int32 = 0
        self.station, self.bdc, self.rx, self.bit_depth = station, bdc, rx, bit_depth
        self.secs = self.df = self.ref_epoch = 0
        self.psn = 1

    def push(self, payload, template_words):
        """payload: uint32 [2][N] (the group's words per polarisation).
        Returns (frames uint32 [2][n][2048], headers uint32 [2][n][8])."""
        N = payload.shape[1]
        skip_int32 = skip_frames = 0
        num_usable = N // self.WPF
        f = in_header_fields(np.asarray(template_words, dtype=np.uint32))
        if not int(f["invalid"]):
            a = output_alignment(template_words, N, self.bit_depth)
            self.secs, self.df, self.ref_epoch, self.psn = a.secs, a.df, a.ref_epoch, 1
            skip_int32, skip_frames, num_usable = a.skip_int32, a.skip_frames, a.num_usable
            self.leftover_int32 = a.leftover_int32
        L = self.leftover_int32
        frames = np.zeros((2, num_usable, self.WPF), dtype=np.uint32)
        for pol in (0, 1):
            parts = []
            if self.leftover[pol] is not None:
                parts.append(self.leftover[pol][:L])
            start = skip_int32 + skip_frames * self.WPF
            parts.append(payload[pol][start:start + N - skip_int32 - L])
            stream = np.concatenate(parts)
            frames[pol] = stream[:num_usable * self.WPF].reshape(num_usable, self.WPF)
            if L > 0:
                self.leftover[pol] = payload[pol][N - L:].copy()
        headers = np.zeros((2, num_usable, 8), dtype=np.uint32)
        for ii in range(num_usable):
            for pol in (0, 1):
                headers[pol, ii] = out_header_words(self.secs, self.df, self.ref_epoch, self.station, pol,
                                                    self.bdc, self.rx, self.psn, bps=self.bit_depth - 1)
            self.df += 1
            if self.df == VDIF_OUT_FRAMES_PER_SECOND:
                self.df = 0
                self.secs += 1
            self.psn += 1
        return frames, headers


# ---------------------------------------------------------------------------
# 2015-format B-engine reorder
# ---------------------------------------------------------------------------
REORDER2015_SPLIT = 69


def reorder_2015(x):
    """2015-format reorder of complex spectra x[channel][B-frame][snapshot]
    (16384 x n_frames x 128) to out[frame][snapshot][16385], the numpy form of
    software/prototyping/cuda/reorder.py:284-293: channels 0-3 of every 8 are
    rolled back two snapshots within their B-frame, then snapshots 69..127 of
    every channel are taken from the next B-frame, the last frame is dropped,
    and the result is transposed (reorder.py:291-292 compares it with
    reorderTz_smem, kernel_template.py:107-164, whose extra channel 16384 is
    zero).
    """
    x = np.asarray(x)
    nch, nf, nsn = x.shape
    r = np.empty_like(x)
    for i in range(4):                                   # reorder.py:285-287
        r[i::8] = np.roll(x[i::8], -2, axis=-1)
        r[i + 4::8] = x[i + 4::8]
    r[:, :, REORDER2015_SPLIT:] = np.roll(r[:, :, REORDER2015_SPLIT:], -1, axis=-2)   # :290
    r = r[:, :-1, :]                                      # :293
    out = np.zeros((nf - 1, nsn, nch + 1), dtype=x.dtype)
    out[:, :, :nch] = r.transpose(1, 2, 0)
    return out


# ---------------------------------------------------------------------------
# Band-trimmed rational resampler of the aphids GPU thread at reduced SWARM
# rates (software/aphids/sample_rates.h:12-40, vdif_inout_gpu_thread.h:31-53)
# ---------------------------------------------------------------------------
SWARM_RATE_FULL = 4576e6                  # sample_rates.h:17
# SWARM_RATE_FRAC -> (DECIMATION_FACTOR, EXPANSION_FACTOR), sample_rates.h:21-59
TRIM_RATES = {6: (39, 32), 8: (13, 8), 10: (65, 64), 11: (143, 128)}
TRIM_SCALE = 256                          # FFT_SIZE_SWARM_R2C_R2DBE_C2R_SCALE, vdif_inout_gpu_thread.h:45


def trim_geometry(rate_frac, skip_mhz):
    """(D, E, R, P, skip_chan) of a rate preset: D snapshots per block
    (RESAMPLE_BATCH_SNAPSHOTS = DECIMATION_FACTOR), R = 256 D the R2C size,
    P = 256 E the C2R size, skip_chan = int(skip_mhz 1e6 / (SWARM_RATE / R))
    (vdif_inout_gpu_thread.cu:269, SWARM_RATE = 4576e6 frac / 11)."""
    D, E = TRIM_RATES[rate_frac]
    R, P = TRIM_SCALE * D, TRIM_SCALE * E
    swarm_rate = SWARM_RATE_FULL * rate_frac / 11
    skip = int(skip_mhz * 1e6 / (swarm_rate / R))
    if skip < 0 or skip + P // 2 > R // 2:
        raise ValueError("skip_chan %d leaves fewer than %d bins" % (skip, P // 2 + 1))
    return D, E, R, P, skip


def resample_trimmed(X, E, skip, scale=TRIM_SCALE):
    """Band-trimmed resampler of one block: X complex spectra [D][nchan]
    (channel nchan, the SWARM Nyquist, is 0), D snapshots, in numpy units.

    SwarmC2R (vdif_inout_gpu_thread.cu:589-625): an inverse real DFT of
    2 nchan points per snapshot (numpy irfft semantics: the imaginary parts of
    bin 0 and of the Nyquist bin are discarded); the D series concatenate.
    SwarmR2C (:631-667): real DFTs of R = scale D samples, batched over the
    concatenation (FFT_BATCHES_SWARM_R2C segments).  R2dbeC2R (:673-709, plan
    :317-333): the inverse real DFT of P = scale E points reads bins
    skip .. skip + P/2 of each segment's spectrum (input pointer + skip_chan,
    batch stride R/2 + 1).  Returns the float64 samples, segment after
    segment (P per segment), normalised as numpy's irfft/rfft/irfft.
    """
    X = np.asarray(X)
    D, nchan = X.shape
    n1 = 2 * nchan
    spec = np.concatenate([X, np.zeros((D, 1), dtype=X.dtype)], axis=1)
    t = np.fft.irfft(spec, n=n1, axis=1).reshape(-1)
    R, P = scale * D, scale * E
    nseg = t.size // R
    Xs = np.fft.rfft(t[:nseg * R].reshape(nseg, R), axis=1)
    Y = Xs[:, skip:skip + P // 2 + 1]
    return np.fft.irfft(Y, n=P, axis=1).reshape(-1)


# ---------------------------------------------------------------------------
# FX cross-correlation QA (software/verification/cross_corr.py,
# verify_scan.py:231-236); Python 3 integer division where the reference
# relies on Python 2's.
# ---------------------------------------------------------------------------
def corr_Xt(X0, X1, fft_window_size=32768):
    """cross_corr.py:23-64: auto- and cross-power averaged over the windows
    (axis 0), lag function s_0x1 = irfft(S_0x1, n) / sqrt(max s_0x0 max s_1x1)."""
    S_0x0 = (X0 * X0.conjugate()).mean(axis=0)
    s_0x0 = np.fft.irfft(S_0x0, n=fft_window_size).real
    S_1x1 = (X1 * X1.conjugate()).mean(axis=0)
    s_1x1 = np.fft.irfft(S_1x1, n=fft_window_size).real
    S_0x1 = (X0 * X1.conjugate()).mean(axis=0)
    s_0x1 = np.fft.irfft(S_0x1, n=fft_window_size).real / np.sqrt(s_0x0.max() * s_1x1.max())
    return s_0x1, S_0x1


def corr_Xt_search(X0, X1, search_range, search_avg=1, fft_window_size=32768):
    """cross_corr.py:66-116: from the centre window of X0
    (X0.shape[0] / 2 - search_avg / 2), correlate search_avg windows of X0
    with the same number of X1 starting search_range[i] windows later; the
    peak |s_0x1| of each offset."""
    c = X0.shape[0] // 2 - search_avg // 2
    s = np.zeros((len(search_range), fft_window_size))
    S = np.zeros((len(search_range), X0.shape[1]), dtype=np.complex128)
    peaks = np.zeros(len(search_range))
    for i, iw in enumerate(search_range):
        s[i], S[i] = corr_Xt(X0[c:c + search_avg], X1[c + iw:c + iw + search_avg], fft_window_size)
        peaks[i] = np.abs(s[i]).max()
    return s, S, peaks


def corr_spectra(x, fft_window_size=32768):
    """Positive half-spectra of the consecutive windows of x (cross_corr.py:150-151)."""
    n = x.size // fft_window_size
    return np.fft.fft(np.asarray(x[:n * fft_window_size], dtype=np.float64).reshape(n, fft_window_size),
                      axis=1)[:, :fft_window_size // 2]


def corr_snr(s, peaks):
    """verify_scan.py:232-235: at the offset of the highest peak, the peak over
    the standard deviation of that lag function; returns (snr, signal, index)."""
    i = int(np.argmax(peaks))
    signal = np.abs(s[i]).max()
    return signal / s[i].std(), signal, i
