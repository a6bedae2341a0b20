"""Pins of the oracle's unpack side (A1-A6) on the CPU.

The decode is checked against a second, independent route -- the generic VDIF
2-bit sample unpack (software/prototyping/vdif.py:198-209) followed by the
DBEFrame channel/polarisation split (software/prototyping/swarm.py:41-52) --
and the scatter against a bijection property plus hand-worked placements.
"""
import numpy as np
import pytest

from oracle import sdbe_oracle as O
from sdbe_amd import synth as S


def _swarm_py_route(payload_u32):
    """vdif.py: sample n of word w at bits 2n, value = code - 2; swarm.py:
    p1 = data[1::4] + 1j data[0::4], p0 = data[3::4] + 1j data[2::4], channel
    ch at p[ch_start[ch]::8] with ch_start = [3,2,1,0,7,6,5,4].
    Returns {pol: complex array [snapshot][8 channels]}."""
    words = np.asarray(payload_u32, dtype=np.uint32)
    data = np.zeros(words.size * 16, dtype=np.int64)
    for n in range(16):
        data[n::16] = (words >> (2 * n)) & 3
    data = data - 2
    p1 = data[1::4] + 1j * data[0::4]
    p0 = data[3::4] + 1j * data[2::4]
    ch_start = [3, 2, 1, 0, 7, 6, 5, 4]
    out = {}
    for name, p in (("p0", p0), ("p1", p1)):
        out[name] = np.stack([p[ch_start[ch]::8] for ch in range(8)], axis=1)
    return out


def test_decode_matches_vdif_swarm_route():
    rng = np.random.default_rng(1)
    pk = np.zeros((1, O.PKT_WORDS), dtype=np.uint32)
    pk[0, 8:] = rng.integers(0, 2 ** 32, 256, dtype=np.uint64).astype(np.uint32)
    cid, fid, b0 = 37, 5, 400
    pk[0, 4] = b0 >> 8
    pk[0, 5] = cid | (fid << 16) | ((b0 & 0xFF) << 24)
    r = O.vdif_to_beng(pk, b0, n_frames=1)
    route = _swarm_py_route(pk[0, 8:])
    base = int(O.channel_base(cid, fid))
    for pol, name in ((0, "p0"), (1, "p1")):
        got = r.beng[pol, 0, :, base:base + 8, 1] + 1j * r.beng[pol, 0, :, base:base + 8, 0]
        assert np.array_equal(got, route[name])
    assert r.n_unique == 1 and r.n_missing == O.VDIF_PER_BENG_FRAME - 1
    assert r.bitmap[0, cid] == 1 << fid


def test_decode_all_byte_values():
    b = np.arange(256)
    p0_re, p0_im, p1_re, p1_im = O.decode_byte(b)
    vals = np.stack([p0_re, p0_im, p1_re, p1_im], axis=1)
    assert vals.min() == -2 and vals.max() == 1
    # bijective: 256 bytes <-> 256 distinct value quadruples
    assert len({tuple(v) for v in vals}) == 256
    # 16 distinct complex values per polarisation (reader.py:255-262)
    assert len(set(p0_re + 1j * p0_im)) == 16 and len(set(p1_re + 1j * p1_im)) == 16
    # byte 0b11_10_01_00: pol1 im = 0-2, pol1 re = 1-2, pol0 im = 2-2, pol0 re = 3-2
    assert O.decode_byte(0b11100100) == (1, 0, -1, -2)


def test_channel_map_is_a_bijection():
    cid = np.repeat(np.arange(256), 8)
    fid = np.tile(np.arange(8), 256)
    base = O.channel_base(cid, fid)
    chans = (base[:, None] + np.arange(8)[None, :]).ravel()
    assert np.array_equal(np.sort(chans), np.arange(16384))
    # hand-worked placements: packets of a cid pair interleave by 8 channels, fid steps by 16
    assert O.channel_base(0, 0) == 0 and O.channel_base(1, 0) == 8
    assert O.channel_base(0, 1) == 16 and O.channel_base(1, 7) == 8 * 15
    assert O.channel_base(2, 0) == 128 and O.channel_base(255, 7) == 16384 - 8


def test_gpu_bcount_parse_and_wrap():
    w = np.zeros((4, 8), dtype=np.uint32)
    # b = 0x12_345678_9A: the kernel sees only the low 24 bits of word 4
    w[0, 4] = 0x12345678
    w[0, 5] = 0x9A000000 | (3 << 16) | 201
    cid, fid, rel = O.gpu_cid_fid_rel(w, 0x3456789A - 5)
    assert (cid[0], fid[0], rel[0]) == (201, 3, 5)
    assert int(O.get_packet_b_count(w)[0]) == 0x123456789A   # the host keeps all 40 bits
    # wrap across 2^32: b_zero = 2^32 - 2, packet at 2^32 + 3 -> rel 5
    b_zero = (1 << 32) - 2
    b = (1 << 32) + 3
    w[1, 4] = b >> 8
    w[1, 5] = (b & 0xFF) << 24
    assert O.gpu_cid_fid_rel(w[1:2], b_zero)[2][0] == 5
    # earlier packet -> negative offset
    w[2, 4] = (b_zero - 1) >> 8
    w[2, 5] = ((b_zero - 1) & 0xFF) << 24
    assert O.gpu_cid_fid_rel(w[2:3], b_zero)[2][0] == -1
    # invalid flag -> -1 from get_packet_b_count
    w[3, 0] = 1 << 31
    assert O.get_packet_b_count(w[3:4])[0] == -1


def _small_group(seed, n_frames=2, drop=0.0):
    rng = S.rng_for(seed)
    c0 = S.gaussian_codes(rng, 128 * n_frames, tones=((100, 0.5, 0.1),))
    c1 = S.gaussian_codes(rng, 128 * n_frames, tones=())
    b0 = 9000
    pk = S.packetize(c0, c1, b0)
    return rng, c0, c1, b0, pk


def test_unpack_round_trip_shuffle_drop():
    rng, c0, c1, b0, pk = _small_group(2)
    n_frames = 2
    ps, dropped = S.shuffle_and_drop(pk, rng, drop_frac=0.01)
    r = O.vdif_to_beng(ps, b0, n_frames=n_frames)
    e0 = S.codes_to_beng_int8(c0).reshape(n_frames, 128, 16384, 2)
    e1 = S.codes_to_beng_int8(c1).reshape(n_frames, 128, 16384, 2)
    # dropped packets leave 0 + 0j in their 128 x 8 block, everything else arrives
    for d in dropped:
        fr, cid, fid = d // 2048, (d % 2048) // 8, d % 8
        base = int(O.channel_base(cid, fid))
        e0[fr, :, base:base + 8] = 0
        e1[fr, :, base:base + 8] = 0
        assert not (r.bitmap[fr, cid] >> fid) & 1
    assert np.array_equal(r.beng[0], e0) and np.array_equal(r.beng[1], e1)
    assert r.n_missing == dropped.size and r.n_unique == pk.shape[0] - dropped.size
    # order invariance
    r2 = O.vdif_to_beng(S.shuffle_and_drop(ps, rng, 0.0)[0], b0, n_frames=n_frames)
    assert np.array_equal(r2.beng, r.beng)


def test_unpack_policies_duplicates_invalid_range():
    _, c0, c1, b0, pk = _small_group(3, n_frames=1)
    dup = pk[10:11].copy()
    dup[:, 8:] = ~dup[:, 8:]
    bad_fid = pk[11:12].copy()
    bad_fid[:, 5] = (bad_fid[:, 5] & ~np.uint32(0xFF0000)) | np.uint32(9 << 16)
    inval = pk[12:13].copy()
    inval[:, 0] |= np.uint32(1 << 31)
    ahead = pk[13:14].copy()
    ahead[:, 5] = (ahead[:, 5] & np.uint32(0x00FFFFFF)) | np.uint32(((b0 + 1) & 0xFF) << 24)
    allp = np.concatenate([pk, dup, bad_fid, inval, ahead])
    r = O.vdif_to_beng(allp, b0, n_frames=1)
    ref = O.vdif_to_beng(pk, b0, n_frames=1)
    assert np.array_equal(r.beng, ref.beng)          # first copy of packet 10 wins
    assert (r.n_duplicate, r.n_invalid, r.n_out_of_range) == (1, 1, 2)
    # the duplicate first: it wins
    r3 = O.vdif_to_beng(np.concatenate([dup, pk]), b0, n_frames=1)
    assert not np.array_equal(r3.beng, ref.beng)


# ---------------------------------------------- the 8 x 1024 packet geometry
def test_8x1024_unpack_equals_reference_geometry_on_the_same_codes():
    """The same codes packed both ways unpack to the same int8 rows: the
    8 x 1024 decode and scatter agree with the reference geometry's (itself
    pinned above against the vdif.py / swarm.py route)."""
    rng = np.random.default_rng(11)
    n_frames, b0 = 2, 5000
    c0 = rng.integers(0, 4, (n_frames * 128, 16384, 2), dtype=np.uint8)
    c1 = rng.integers(0, 4, (n_frames * 128, 16384, 2), dtype=np.uint8)
    ref = O.vdif_to_beng(S.packetize(c0, c1, b0), b0, n_frames=n_frames)
    g2 = O.vdif_to_beng_8x1024(S.packetize_8x1024(c0, c1, b0), b0, n_frames=n_frames)
    assert np.array_equal(ref.beng, g2.beng)
    assert g2.n_unique == n_frames * 256 and g2.n_missing == 0
    assert np.all(g2.bitmap == 0xFFFF)


def test_8x1024_hand_built_packet():
    """A packet built by hand from the format's definition (not by the
    packetiser): byte q of snapshot row sn holds channel 1024 cid + (q & ~3) +
    3 - (q & 3); bits from the LSB pol1 im, pol1 re, pol0 im, pol0 re; value =
    code - 2."""
    cid, fid, b0 = 9, 13, 7777
    pk = np.zeros((1, O.G2_PKT_WORDS), dtype=np.uint32)
    pk[0, 4] = b0 >> 8
    pk[0, 5] = cid | (fid << 16) | ((b0 & 0xFF) << 24)
    pay = np.full(8 * 1024, 0xAA, dtype=np.uint8)        # every field code 2: value 0
    # row 2, word 5, byte 0 -> channel 4 * 5 + 3 = 23; fields im1=0, re1=1, im0=2, re0=3
    pay[2 * 1024 + 4 * 5 + 0] = 0b11_10_01_00
    # row 7, word 255, byte 3 -> channel 1020
    pay[7 * 1024 + 4 * 255 + 3] = 0b00_01_10_11
    pk[0, 8:] = pay.view(np.uint32)
    r = O.vdif_to_beng_8x1024(pk, b0, n_frames=1)
    ch = 1024 * cid
    assert tuple(r.beng[0, 0, 8 * fid + 2, ch + 23]) == (0, 1)       # pol 0 (im, re)
    assert tuple(r.beng[1, 0, 8 * fid + 2, ch + 23]) == (-2, -1)     # pol 1
    assert tuple(r.beng[0, 0, 8 * fid + 7, ch + 1020]) == (-1, -2)
    assert tuple(r.beng[1, 0, 8 * fid + 7, ch + 1020]) == (1, 0)
    nz = np.argwhere(r.beng != 0)
    assert len(nz) == 6                 # the 8 fields set above, two of them value 0
    assert r.bitmap[0, cid] == 1 << fid and r.bitmap.sum() == 1 << fid


def test_8x1024_drops_duplicates_invalid_range():
    """Dropped slots stay zero with their bitmap bit clear; a duplicate loses
    to the lower packet index; invalid and out-of-window packets are counted
    and ignored."""
    rng = np.random.default_rng(12)
    n_frames, b0 = 2, 300
    c0 = rng.integers(0, 4, (n_frames * 128, 16384, 2), dtype=np.uint8)
    c1 = rng.integers(0, 4, (n_frames * 128, 16384, 2), dtype=np.uint8)
    pk = S.packetize_8x1024(c0, c1, b0)
    full = O.vdif_to_beng_8x1024(pk, b0, n_frames=n_frames)
    drop = np.array([3, 77, 300, 511])
    keep = np.setdiff1d(np.arange(pk.shape[0]), drop)
    dup = pk[5].copy()
    dup[8:] ^= np.uint32(0xFFFFFFFF)                     # same slot, other payload, higher index
    inval = pk[6].copy()
    inval[0] |= np.uint32(1 << 31)
    late = pk[7].copy()
    late[5] = (late[5] & 0x00FFFFFF) | (((b0 + n_frames) & 0xFF) << 24)   # frame n_frames: out of window
    late[4] = (b0 + n_frames) >> 8
    r = O.vdif_to_beng_8x1024(np.concatenate([pk[keep], dup[None], inval[None], late[None]]), b0,
                              n_frames=n_frames)
    assert r.n_invalid == 1 and r.n_out_of_range == 1 and r.n_duplicate == 1
    assert r.n_unique == pk.shape[0] - drop.size and r.n_missing == drop.size
    want = full.beng.copy()
    for p in drop:
        f, cid, fid = p // 256, (p % 256) // 16, p % 16
        want[:, f, 8 * fid:8 * fid + 8, 1024 * cid:1024 * (cid + 1)] = 0
        assert not (r.bitmap[f, cid] >> fid) & 1
    assert np.array_equal(r.beng, want)
