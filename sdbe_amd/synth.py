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
                       b_first=1000):
    """Config-2 style input: ``n_groups`` groups of 39 B-frames, both pols.
    Returns a list of (packets, b_zero) per group and the per-pol codes."""
    rng = rng_for(seed)
    n_snap = 39 * N_SNAP_PER_FRAME * n_groups
    kw = {} if tones is None else {"tones": tones}
    c0 = gaussian_codes(rng, n_snap, **kw)
    c1 = gaussian_codes(rng, n_snap, **kw)
    out = []
    per = 39 * N_SNAP_PER_FRAME
    for g in range(n_groups):
        b0 = b_first + 39 * g
        pk = packetize(c0[g * per:(g + 1) * per], c1[g * per:(g + 1) * per], b0,
                       df0=g * 39 * 2048)
        if shuffle:
            pk, _ = shuffle_and_drop(pk, rng, drop_frac)
        out.append((pk, b0))
    return out, (c0, c1)
