"""Seeded synthetic SWARM B-engine data in the reference wire format.

Spectra are circular complex Gaussian (unit variance) plus tones, each of
re/im quantised to offset-binary 2-bit codes with thresholds +-0.9816 sigma_c
(codes 0: x < -t, 1: -t <= x < 0, 2: 0 <= x < t, 3: x >= t).  The stage
decodes a code c as c - 2 (software/aphids/vdif_inout_gpu_thread.cu:387-392).

Packets follow the B-engine-over-VDIF layout of
software/aphids/vdif_in_databuf.h:29-67: a 32-byte header whose word 4 is
b_upper and whose word 5 holds (c, z, f, b_lower) bytes, then 128 snapshots x
8 bytes; byte q of a snapshot carries channel offset (3,2,1,0,7,6,5,4)[q] of
the 8 channels, LSB first: pol1 im, pol1 re, pol0 im, pol0 re
(vdif_inout_gpu_thread.cu:464-494).
"""
from __future__ import annotations

import math

import numpy as np

N_CHAN = 16384
N_SNAP_PER_FRAME = 128
PKT_WORDS = 264
THRESH_SIGMA = 0.9816
_BYTE_TO_OFFSET = np.array([3, 2, 1, 0, 7, 6, 5, 4])


def _phi(x):
    return 0.5 * (1.0 + math.erf(x / math.sqrt(2.0)))


def rng_for(seed):
    return np.random.Generator(np.random.PCG64(seed))


def quantize_components(x, sigma_c):
    """Real array -> 2-bit codes with thresholds +-0.9816 sigma_c."""
    t = THRESH_SIGMA * sigma_c
    c = np.full(x.shape, 3, dtype=np.uint8)
    c[x < t] = 2
    c[x < 0] = 1
    c[x < -t] = 0
    return c


def gaussian_codes(rng, n_snap, sigma=1.0, tones=((5000, 0.01, 0.0371),), quant_sigma=None):
    """Codes [n_snap][16384][2 = (re, im)] for CN(0, sigma^2) noise plus tones.

    ``tones`` holds (channel, power relative to sigma^2, phase step per
    snapshot in cycles).  The quantiser thresholds are +-0.9816 quant_sigma_c
    (quant_sigma defaults to sigma: a quantiser matched to the signal).
    Noise-only channels are drawn by inverse CDF from uniforms (same
    distribution as quantising Gaussian samples); tone channels are drawn as
    Gaussians plus the coherent tone and then quantised."""
    qs = sigma if quant_sigma is None else quant_sigma
    p0 = _phi(-THRESH_SIGMA * qs / sigma)
    u = rng.random((n_snap, N_CHAN, 2), dtype=np.float32)
    codes = np.full(u.shape, 3, dtype=np.uint8)
    codes[u < 1.0 - p0] = 2
    codes[u < 0.5] = 1
    codes[u < p0] = 0
    sigma_c = sigma * math.sqrt(0.5)
    for ch, power, step in tones:
        amp = sigma * math.sqrt(power)
        phase = 2.0 * math.pi * (rng.random() + step * np.arange(n_snap))
        z = rng.normal(0.0, sigma_c, (n_snap, 2))
        z[:, 0] += amp * np.cos(phase)
        z[:, 1] += amp * np.sin(phase)
        codes[:, ch, :] = quantize_components(z, qs * math.sqrt(0.5))
    return codes


def batch_streams(seed, n_blocks=(17, 64, 128, 5, 200, 33), lsb=(False, True, False, True, False, True)):
    """Ragged multi-stream input (3 beams x 2 sidebands): per stream int8
    spectra [2 pols][39 n][16384][2 = (im, re)], sigma ~ U[0.5, 2] quantised
    by a unit-sigma 2-bit quantiser, a random tone per stream.  Returns a list
    of dicts (beng, n_blocks, lsb, sigma, tone_channel)."""
    rng = rng_for(seed)
    out = []
    for nb, lo in zip(n_blocks, lsb):
        sigma = float(rng.uniform(0.5, 2.0))
        ch = int(rng.integers(100, N_CHAN - 100))
        step = float(rng.uniform(-0.5, 0.5))
        tones = ((ch, float(rng.uniform(0.005, 0.05)), step),)
        pols = [codes_to_beng_int8(gaussian_codes(rng, 39 * nb, sigma, tones, quant_sigma=1.0))
                for _ in range(2)]
        out.append(dict(beng=np.stack(pols), n_blocks=nb, lsb=lo, sigma=sigma, tone_channel=ch))
    return out


def codes_to_beng_int8(codes):
    """Codes [..., 16384, 2=(re,im)] -> reference int8 layout [..., 16384, 2=(im,re)]
    holding code - 2 (the layout of beng_0/beng_1 in the reference stage)."""
    out = np.empty(codes.shape, dtype=np.int8)
    out[..., 0] = codes[..., 1].astype(np.int8) - 2
    out[..., 1] = codes[..., 0].astype(np.int8) - 2
    return out


def packetize(codes0, codes1, b_first, secs=1000, df0=0, ref_epoch=30,
              station=0x536D):
    """Two polarisations of codes [n_frames*128][16384][2] -> packets
    (n_frames*2048, 264) uint32, ordered (frame, cid, fid)."""
    n_snap = codes0.shape[0]
    assert n_snap % N_SNAP_PER_FRAME == 0 and codes1.shape == codes0.shape
    n_frames = n_snap // N_SNAP_PER_FRAME
    byte = (codes1[..., 1] | (codes1[..., 0] << 2) | (codes0[..., 1] << 4)
            | (codes0[..., 0] << 6)).astype(np.uint8)
    # channel = 128*h + 16*fid + 8*lo + off  with cid = 2h + lo
    byte = byte.reshape(n_frames, N_SNAP_PER_FRAME, 128, 8, 2, 8)
    byte = byte[..., _BYTE_TO_OFFSET]                       # offset axis -> byte order q
    pay = np.ascontiguousarray(byte.transpose(0, 2, 4, 3, 1, 5))   # (frame, h, lo, fid, snap, q)
    pay = pay.reshape(n_frames * 2048, 1024).view(np.uint32)
    pk = np.zeros((n_frames * 2048, PKT_WORDS), dtype=np.uint32)
    pk[:, 8:] = pay
    fr = np.repeat(np.arange(n_frames, dtype=np.int64), 2048)
    idx = np.tile(np.arange(2048, dtype=np.int64), n_frames)
    cid = idx // 8
    fid = idx % 8
    b = b_first + fr
    n = np.arange(n_frames * 2048, dtype=np.int64)
    pk[:, 0] = (secs + (df0 + n) // 125000) & 0x3FFFFFFF
    pk[:, 1] = ((df0 + n) % 125000) | (ref_epoch << 24)
    pk[:, 2] = 1056 // 8
    pk[:, 3] = station | (1 << 26)
    pk[:, 4] = (b >> 8) & 0xFFFFFFFF
    pk[:, 5] = (cid | (fid << 16) | ((b & 0xFF) << 24)).astype(np.uint32)
    pk[:, 6] = n & 0xFFFFFFFF
    pk[:, 7] = 0
    return pk


PKT_WORDS_8X1024 = 2056


def packetize_8x1024(codes0, codes1, b_first, secs=1000, df0=0, ref_epoch=30, station=0x536D):
    """The "8 snapshots x 1024 channels" packet geometry (SDBE_GEOM_8X1024):
    two polarisations of codes [n_frames*128][16384][2] -> packets
    (n_frames*256, 2056) uint32 ordered (frame, cid, fid).  Packet (frame,
    cid, fid) carries snapshots 8 fid .. 8 fid + 7 of its B-frame for channels
    1024 cid .. 1024 cid + 1023: row sn of the payload is 256 words, word w
    holding channels 4 w .. 4 w + 3 with channel 4 w + j in byte 3 - j, each
    byte the reference's code layout (pol1 im, pol1 re, pol0 im, pol0 re from
    the LSB).  Header as the reference's with cid, fid in word 5."""
    n_snap = codes0.shape[0]
    assert n_snap % N_SNAP_PER_FRAME == 0 and codes1.shape == codes0.shape
    n_frames = n_snap // N_SNAP_PER_FRAME
    byte = (codes1[..., 1] | (codes1[..., 0] << 2) | (codes0[..., 1] << 4)
            | (codes0[..., 0] << 6)).astype(np.uint8)
    # (frame, fid, sn, cid, word, channel offset j) -> byte order q = 3 - j
    byte = byte.reshape(n_frames, 16, 8, 16, 256, 4)[..., ::-1]
    pay = np.ascontiguousarray(byte.transpose(0, 3, 1, 2, 4, 5))          # (frame, cid, fid, sn, w, q)
    pay = pay.reshape(n_frames * 256, 8192).view(np.uint32)
    pk = np.zeros((n_frames * 256, PKT_WORDS_8X1024), dtype=np.uint32)
    pk[:, 8:] = pay
    fr = np.repeat(np.arange(n_frames, dtype=np.int64), 256)
    idx = np.tile(np.arange(256, dtype=np.int64), n_frames)
    cid = idx // 16
    fid = idx % 16
    b = b_first + fr
    n = np.arange(n_frames * 256, dtype=np.int64)
    pk[:, 0] = (secs + (df0 + n) // 125000) & 0x3FFFFFFF
    pk[:, 1] = ((df0 + n) % 125000) | (ref_epoch << 24)
    pk[:, 2] = 8224 // 8
    pk[:, 3] = station | (1 << 26)
    pk[:, 4] = (b >> 8) & 0xFFFFFFFF
    pk[:, 5] = (cid | (fid << 16) | ((b & 0xFF) << 24)).astype(np.uint32)
    pk[:, 6] = n & 0xFFFFFFFF
    pk[:, 7] = 0
    return pk


def shuffle_and_drop(pkts, rng, drop_frac=0.001):
    """Fisher-Yates shuffle of the packet order, then drop round(frac*n)
    seed-chosen packets.  Returns (packets, indices of dropped originals)."""
    n = pkts.shape[0]
    perm = rng.permutation(n)
    n_drop = int(round(drop_frac * n))
    drop = np.sort(rng.choice(n, size=n_drop, replace=False)) if n_drop else np.zeros(0, np.int64)
    keep = np.ones(n, dtype=bool)
    keep[drop] = False
    perm = perm[keep[perm]]
    return np.ascontiguousarray(pkts[perm]), drop


def make_group_packets(seed, n_groups=2, drop_frac=0.001, shuffle=True, tones=None,
                       b_first=1000, frames=39, geometry=0):
    """Config-2 style input: ``n_groups`` groups of ``frames`` B-frames (39 at
    the 6/11 rate; BENG_FRAMES_PER_GROUP = DECIMATION_FACTOR of the rate,
    software/aphids/sample_rates.h:67), both pols, in the reference packet
    geometry (geometry 0) or the 8 x 1024 one (geometry 1,
    ``packetize_8x1024``).  Returns a list of (packets, b_zero) per group and
    the per-pol codes."""
    rng = rng_for(seed)
    n_snap = frames * N_SNAP_PER_FRAME * n_groups
    kw = {} if tones is None else {"tones": tones}
    c0 = gaussian_codes(rng, n_snap, **kw)
    c1 = gaussian_codes(rng, n_snap, **kw)
    out = []
    per = frames * N_SNAP_PER_FRAME
    for g in range(n_groups):
        b0 = b_first + frames * g
        if geometry == 1:
            pk = packetize_8x1024(c0[g * per:(g + 1) * per], c1[g * per:(g + 1) * per], b0,
                                  df0=g * frames * 256)
        else:
            pk = packetize(c0[g * per:(g + 1) * per], c1[g * per:(g + 1) * per], b0,
                           df0=g * frames * 2048)
        if shuffle:
            pk, _ = shuffle_and_drop(pk, rng, drop_frac)
        out.append((pk, b0))
    return out, (c0, c1)


SWARM_CLK_PER_SEC = 286_000_000     # B-engine clock ticks per second (beng_reform.c)
SWARM_BFRAME_CLK = 262_144          # clock ticks per B-frame


def to_8pac_new_format(pkts, b_ref, t_ref):
    """Old-format packets -> the SWARM wire format the APHIDS input thread
    receives: new-format headers carrying the 53-bit B-engine timestamp (clock
    in the low 29 bits, seconds above; word 4 = bits 21..52, word 5 =
    c | f << 8 | bits 0..20 << 11; software/aphids/beng_reform.c:16-22,137-186)
    and 8 packets per 8pac record (packet 1, then 7 x (8-byte VTP + packet);
    the record's df / psn / df_len apply to all 8, vdif_8pac.c:5-56).
    Packet p's timestamp is t_ref = (sec, clk) plus (B-count - b_ref)
    B-frames.  Returns uint8 records [n / 8][8504]; n must be a multiple of 8."""
    pkts = np.ascontiguousarray(pkts, dtype=np.uint32)
    n = pkts.shape[0]
    assert n % 8 == 0
    w = pkts.copy()
    w4, w5 = w[:, 4].astype(np.int64), w[:, 5].astype(np.int64)
    b = (w5 >> 24) + ((w4 & 0xFFFFFF) << 8)
    clk_total = int(t_ref[1]) + (b - b_ref) * SWARM_BFRAME_CLK
    sec = int(t_ref[0]) + clk_total // SWARM_CLK_PER_SEC
    clk = clk_total % SWARM_CLK_PER_SEC
    b53 = (sec << 29) | clk
    cid, fid = w5 & 0xFF, (w5 >> 16) & 0x7
    w[:, 4] = (b53 >> 21).astype(np.uint32)
    w[:, 5] = (cid | (fid << 8) | ((b53 & 0x1FFFFF) << 11)).astype(np.uint32)
    rec = np.zeros((n // 8, 1056 + 7 * 1064), dtype=np.uint8)
    k = np.arange(n // 8, dtype=np.int64)
    first = w[0::8].copy()
    first[:, 1] = (first[:, 1] & np.uint32(0xFF000000)) | (k & 0xFFFFFF).astype(np.uint32)
    first[:, 2] = (first[:, 2] & np.uint32(0xFF000000)) | np.uint32((1056 + 7 * 1064) // 8)
    first[:, 6] = (k & 0xFFFFFFFF).astype(np.uint32)
    first[:, 7] = 0
    first[:, 0] &= np.uint32(0x7FFFFFFF)
    rec[:, :1056] = first.view(np.uint8)
    for i in range(1, 8):
        s = 1056 + (i - 1) * 1064 + 8
        rec[:, s:s + 1056] = w[i::8].view(np.uint8)
    return rec


def make_stream_packets(seed, n_groups, frames=39, drop_frac=0.001, b_first=1000):
    """A received stream for the input thread: B-frames b_first ..
    b_first + frames n_groups in order (the first one is before the start the
    thread derives and is skipped by it), packets shuffled within each B-frame,
    round(drop_frac n) seed-chosen packets dropped (never the stream's first,
    the count a multiple of 8).  Returns (packets in arrival order, the per-pol
    codes of all frames)."""
    rng = rng_for(seed)
    n_frames = 1 + frames * n_groups
    n_snap = n_frames * N_SNAP_PER_FRAME
    c0 = gaussian_codes(rng, n_snap)
    c1 = gaussian_codes(rng, n_snap)
    pk = packetize(c0, c1, b_first)
    order = np.concatenate([f * 2048 + rng.permutation(2048) for f in range(n_frames)])
    n_drop = int(round(drop_frac * pk.shape[0]))
    n_drop -= n_drop % 8
    keep = np.ones(pk.shape[0], dtype=bool)
    if n_drop:
        keep[1 + rng.choice(pk.shape[0] - 1, n_drop, replace=False)] = False
    order = order[keep[order]]
    return np.ascontiguousarray(pk[order]), (c0, c1)
