"""Pins of the oracle's VDIF side (A1, A7, A14, A15) and the host bookkeeping
(A5, A6) on the CPU.

Header words are checked against ctypes structures with the reference's
bit-fields (the platform C ABI lays those out, independently of the shifts
the oracle uses); B-frame reform against an absolute clock count; stamping,
template and alignment against hand-worked values.
"""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import sdbe_oracle as O


class W0(C.LittleEndianStructure):
    _fields_ = [("secs_inre", C.c_uint32, 30), ("legacy", C.c_uint32, 1), ("invalid", C.c_uint32, 1)]


class W1(C.LittleEndianStructure):
    _fields_ = [("df_num_insec", C.c_uint32, 24), ("ref_epoch", C.c_uint32, 6), ("UA", C.c_uint32, 2)]


class W2(C.LittleEndianStructure):
    _fields_ = [("df_len", C.c_uint32, 24), ("num_channels", C.c_uint32, 5), ("ver", C.c_uint32, 3)]


class W3(C.LittleEndianStructure):
    _fields_ = [("stationID", C.c_uint32, 16), ("threadID", C.c_uint32, 10), ("bps", C.c_uint32, 5),
                ("dt", C.c_uint32, 1)]


class W4out(C.LittleEndianStructure):
    _fields_ = [("pol", C.c_uint32, 1), ("bdc_sideband", C.c_uint32, 1), ("rx_sideband", C.c_uint32, 1),
                ("eud5", C.c_uint32, 21), ("edv5", C.c_uint32, 8)]


class OutHdr(C.LittleEndianStructure):
    """vdif_out_header_t (software/aphids/vdif_out_databuf.h:19-52)."""
    _fields_ = [("w0", W0), ("w1", W1), ("w2", W2), ("w3", W3), ("w4", W4out), ("PICstatus", C.c_uint32),
                ("edh_psn", C.c_uint64)]


class BengHdr(C.LittleEndianStructure):
    _fields_ = [("b_upper", C.c_uint32), ("c", C.c_uint8), ("z", C.c_uint8), ("f", C.c_uint8),
                ("b_lower", C.c_uint8)]


class InHdr(C.LittleEndianStructure):
    """vdif_in_header_t (software/aphids/vdif_in_databuf.h:29-59)."""
    _fields_ = [("w0", W0), ("w1", W1), ("w2", W2), ("w3", W3), ("beng", BengHdr), ("edh_psn", C.c_uint64)]


def _words(struct):
    return np.frombuffer(bytes(struct), dtype=np.uint32)


def test_out_header_layout_matches_c_bitfields():
    assert C.sizeof(OutHdr) == 32
    h = OutHdr()
    h.w0.secs_inre, h.w0.legacy, h.w0.invalid = 0x2ABCDEF1, 0, 1
    h.w1.df_num_insec, h.w1.ref_epoch, h.w1.UA = 124999, 33, 2
    h.w2.df_len, h.w2.num_channels, h.w2.ver = 1028, 3, 5
    h.w3.stationID, h.w3.threadID, h.w3.bps, h.w3.dt = 0x536D, 1023, 1, 1
    h.w4.pol, h.w4.bdc_sideband, h.w4.rx_sideband, h.w4.eud5, h.w4.edv5 = 1, 0, 1, 0x12345, 2
    h.PICstatus = 0xDEADBEEF
    h.edh_psn = 0x0123456789ABCDEF
    got = O.out_header_words(0x2ABCDEF1, 124999, 33, 0x536D, 1, 0, 1, 0x0123456789ABCDEF, df_len=1028,
                             bps=1, edv5=2, eud5=0x12345, thread=1023, nchan=3, ver=5, invalid=1, ua=2,
                             dt=1, pic=0xDEADBEEF)
    assert np.array_equal(got, _words(h))


def test_in_header_fields_match_c_bitfields():
    h = InHdr()
    h.w0.secs_inre, h.w0.invalid = 12345678, 1
    h.w1.df_num_insec, h.w1.ref_epoch = 77, 30
    h.w2.df_len = 132
    h.w3.stationID, h.w3.bps = 0x4142, 1
    h.beng.b_upper, h.beng.c, h.beng.z, h.beng.f, h.beng.b_lower = 0xA1B2C3, 200, 9, 6, 0x5F
    h.edh_psn = 987654321012
    f = O.in_header_fields(_words(h))
    assert (f["secs_inre"], f["invalid"], f["df_num_insec"], f["ref_epoch"]) == (12345678, 1, 77, 30)
    assert (f["df_len"], f["stationID"], f["bps"]) == (132, 0x4142, 1)
    assert (f["b_upper"], f["c"], f["z"], f["f"], f["b_lower"]) == (0xA1B2C3, 200, 9, 6, 0x5F)
    assert f["edh_psn"] == 987654321012


def test_stamp_headers_wraps_second_and_counts_psn():
    st = O.StreamState(secs=10, df=124998, ref_epoch=30, psn=1, station=O.station_from_str("Sm"), bdc=1, rx=0)
    hdr, nxt = O.stamp_headers(4, st)
    f = O.in_header_fields(hdr[0])            # words 0-3 share the layout
    assert list(f["df_num_insec"]) == [124998, 124999, 0, 1]
    assert list(f["secs_inre"]) == [10, 10, 11, 11]
    assert list(f["edh_psn"]) == [1, 2, 3, 4]
    assert (nxt.secs, nxt.df, nxt.psn) == (11, 2, 5)
    # both polarisations: identical except w4.pol
    assert np.array_equal(hdr[0][:, [0, 1, 2, 3, 5, 6, 7]], hdr[1][:, [0, 1, 2, 3, 5, 6, 7]])
    assert list(hdr[0][:, 4] & 1) == [0] * 4 and list(hdr[1][:, 4] & 1) == [1] * 4
    # defaults of init_vdif_out / the stamping loop: df_len 1028, bps 1, edv5 2, threadID 0
    assert (hdr[0][0, 2] & 0xFFFFFF) == 1028 and ((hdr[0][0, 3] >> 26) & 0x1F) == 1
    assert (hdr[0][0, 4] >> 24) == 2 and ((hdr[0][0, 3] >> 16) & 0x3FF) == 0
    assert (hdr[0][0, 3] & 0xFFFF) == 0x536D and (hdr[0][0, 4] >> 1) & 1 == 1


def _new_format_header(sec, clk, c, f):
    b53 = (sec << 29) | clk
    w = np.zeros(8, dtype=np.uint32)
    w[4] = (b53 >> 21) & 0xFFFFFFFF
    w[5] = (c & 0xFF) | ((f & 7) << 8) | ((b53 & 0x1FFFFF) << 11)
    return w


def test_beng_reform_counts_frames_from_absolute_clock():
    """Timestamps T0 + k * 262144 clocks (with second wraps, in arbitrary packet
    order within a small window) reform to B-count k."""
    T0 = (1000, 285_000_000)
    t_abs0 = T0[0] * O.SWARM_CLK_PER_SEC + T0[1]
    ks = [0, 1, 2, 3, 5, 4, 6, 10, 9, 20]
    hdr = np.zeros((len(ks), 8), dtype=np.uint32)
    for i, k in enumerate(ks):
        sec, clk = divmod(t_abs0 + k * O.SWARM_BFRAME_CLK, O.SWARM_CLK_PER_SEC)
        hdr[i] = _new_format_header(sec, clk, c=i, f=i % 8)
    ref = O.BengReformer()
    out = ref.reform(hdr.copy())
    b = O.get_packet_b_count(out)
    assert list(b) == ks
    assert list(out[:, 5] & 0xFF) == list(range(len(ks)))          # c kept
    assert list((out[:, 5] >> 16) & 0xFF) == [i % 8 for i in range(len(ks))]
    # get_beng_t0: T0 moved back one B-frame
    assert ref.beng_t0() == [T0[0], T0[1] - O.SWARM_BFRAME_CLK]


def test_beng_reform_tolerance_and_far_jump():
    ref = O.BengReformer()
    base = (500, 1_000_000)
    h = np.stack([
        _new_format_header(base[0], base[1], 0, 0),
        _new_format_header(base[0], base[1] + 1000, 1, 0),             # within 1024 clk: same frame
        _new_format_header(base[0], base[1] + O.SWARM_BFRAME_CLK - 900, 2, 0),   # next frame
        _new_format_header(base[0] + 2, base[1], 3, 0),                # > 1000 frames ahead
    ])
    out = ref.reform(h)
    f = O.in_header_fields(out)
    assert list(O.get_packet_b_count(out)[:3]) == [0, 0, 1]
    assert f["invalid"][3] == 1
    # invalid packets are skipped by later reforms and keep the previous reference
    h2 = np.stack([_new_format_header(base[0], base[1] + 2 * O.SWARM_BFRAME_CLK, 4, 0)])
    assert O.get_packet_b_count(ref.reform(h2))[0] == 2


def test_unpack_8pac():
    raw = np.zeros((1, O.PKT_BYTES + 7 * (8 + O.PKT_BYTES)), dtype=np.uint8)
    pk = np.zeros(O.PKT_WORDS, dtype=np.uint32)
    pk[0] = 1 << 31 | 5
    pk[1] = 1000 | (30 << 24)
    pk[2] = 8 * 132 + 7
    pk[6] = 11
    pk[8:] = 0xAB
    raw[0, :O.PKT_BYTES] = pk.view(np.uint8)
    for i in range(7):
        s = O.PKT_BYTES + i * (8 + O.PKT_BYTES) + 8
        q = np.zeros(O.PKT_WORDS, dtype=np.uint32)
        q[8:] = i
        raw[0, s:s + O.PKT_BYTES] = q.view(np.uint8)
    out = O.unpack_8pac(raw)
    f = O.in_header_fields(out[:, :8])
    assert list(f["df_len"]) == [132] * 8
    assert list(f["df_num_insec"]) == [8000 + i for i in range(8)]
    assert list(f["edh_psn"]) == [88 + i for i in range(8)]
    assert list(f["invalid"]) == [1] * 8
    assert f["ref_epoch"][0] == 30 and f["secs_inre"][0] == 5
    assert np.all(out[0, 8:] == 0xAB) and np.all(out[3, 8:] == 2)


def _pkt(b, c=0, f=0, invalid=0):
    w = np.zeros(8, dtype=np.uint32)
    w[0] = invalid << 31
    w[4] = b >> 8
    w[5] = c | (f << 16) | ((b & 0xFF) << 24)
    return w


def test_group_window_insert_and_completion():
    groups = [O.init_beng_group(100 + 39 * i) for i in range(4)]
    assert O.group_index_offset(groups, 0, _pkt(99)) == O.ERR_BEFORE_START
    assert O.group_index_offset(groups, 0, _pkt(120, invalid=1)) == O.ERR_INVALID
    assert O.group_index_offset(groups, 0, _pkt(100)) == 0
    assert O.group_index_offset(groups, 0, _pkt(100 + 39 * 3 + 38)) == 3
    # the window has advanced 4 groups when the "too far" test runs: b > 294 + 4 * 39
    assert O.group_index_offset(groups, 0, _pkt(451)) == O.ERR_TOO_FAR_AHEAD
    assert O.group_index_offset(groups, 0, _pkt(450)) == 3
    # past the 4 windows but not too far: the last offset is returned
    assert O.group_index_offset(groups, 0, _pkt(100 + 39 * 5)) == 3
    assert O.insert_packet(groups, 0, 0, _pkt(105, c=7, f=3), 0) == 1
    assert O.insert_packet(groups, 0, 0, _pkt(105, c=7, f=3), 1) == 1   # duplicates count twice
    assert groups[0].cc[5, 7] == 1 << 3 and groups[0].frame_count[5] == 2
    assert O.group_complete(groups[0]) == 0
    g = O.init_beng_group(0, n_frames=1)
    ring = [g] + [O.init_beng_group(1 + i, n_frames=1) for i in range(3)]
    for i in range(O.VDIF_PER_BENG_FRAME):
        O.insert_packet(ring, 0, 0, _pkt(0, c=i // 8, f=i % 8), i)
    assert O.group_complete(g) == 1 and np.all(g.cc == 0xFF)


def test_header_template_and_output_alignment_worked_example():
    ref_hdr = np.zeros(8, dtype=np.uint32)
    ref_hdr[1] = 30 << 24
    # T0 at half a second: frame position 62500 exactly, no picoseconds to skip
    w = O.fill_header_template(ref_hdr, (100, 143_000_000))
    f = O.in_header_fields(w)
    assert (f["secs_inre"], f["df_num_insec"], f["edh_psn"], f["ref_epoch"]) == (100, 62500, 0, 30)
    a = O.output_alignment(w, 8192 * 2048)
    # skip = floor(floor(3715791 ps * 4.096 GHz) / 4 / 4) = floor(15219 / 16) = 951 words
    assert (a.skip_int32, a.skip_frames, a.num_usable, a.leftover_int32) == (951, 0, 8191, 1097)
    assert (a.secs, a.df, a.psn) == (100, 62500, 1)
    # T0 just after the start of a second: ceil to the next 8 us frame, rest in edh_psn
    w = O.fill_header_template(ref_hdr, (7, 1000))
    f = O.in_header_fields(w)
    full = 125000 * 1000 / 286e6
    assert f["df_num_insec"] == math.ceil(full)
    assert f["edh_psn"] == math.floor((math.ceil(full) - full) * 8e6)
    # a start within num_usable frames of the end of the second skips to the next second
    w = O.fill_header_template(ref_hdr, (7, 285_000_000))
    a = O.output_alignment(w, 8192 * 2048)
    assert a.skip_frames == 125000 - int(O.in_header_fields(w)["df_num_insec"]) and a.df == 0 and a.secs == 8


def test_cabi_state_from_template_matches_oracle():
    """The library's host-side derivation (no GPU needed) equals the oracle."""
    from sdbe_amd.stage import state_from_template
    ref_hdr = np.zeros(8, dtype=np.uint32)
    ref_hdr[1] = 30 << 24
    for T0 in [(100, 143_000_000), (7, 1000), (7, 285_000_000), (2000, 12_345_678)]:
        w = O.fill_header_template(ref_hdr, T0)
        a = O.output_alignment(w, 8192 * 2048)
        got = state_from_template(w)
        assert got == dict(secs=a.secs, df=a.df, ref_epoch=a.ref_epoch, psn=1, skip_int32=a.skip_int32,
                           skip_frames=a.skip_frames)


@pytest.mark.parametrize("df0,psn_ps", [(124990, 12_345_678), (5, 7), (124000, 300_000)])
def test_output_aligner_sends_the_stream_from_the_skip(df0, psn_ps):
    """The frames the output thread sends, concatenated over groups, are the
    stage's word stream from skip_int32 + 2048 skip_frames on, and their
    headers count on without a gap (df wrap with secs + 1, psn + 1)."""
    rng = np.random.default_rng(df0)
    N = 64 * 2048                                  # a short group keeps the test fast
    tm = np.zeros(8, dtype=np.uint32)
    tm[0] = 100
    tm[1] = (30 << 24) | df0
    tm[6] = psn_ps & 0xFFFFFFFF
    tm[7] = psn_ps >> 32
    a = O.output_alignment(tm, N)
    inv = tm.copy()
    inv[0] |= 1 << 31
    al = O.OutputAligner(station=O.station_from_str("Sm"))
    pays = [rng.integers(0, 2 ** 32, (2, N), dtype=np.uint64).astype(np.uint32) for _ in range(3)]
    outs = [al.push(pays[0], tm), al.push(pays[1], inv), al.push(pays[2], inv)]
    D = a.skip_int32 + 2048 * a.skip_frames
    assert outs[0][0].shape[1] == a.num_usable and outs[1][0].shape[1] == N // 2048
    for pol in (0, 1):
        stream = np.concatenate([p[pol] for p in pays])[D:]
        sent = np.concatenate([o[0][pol].reshape(-1) for o in outs])
        assert np.array_equal(stream[:sent.size], sent)
    hdr = np.concatenate([o[1][0] for o in outs])
    f = O.in_header_fields(hdr)
    t = f["secs_inre"].astype(np.int64) * 125000 + f["df_num_insec"].astype(np.int64)
    assert np.all(np.diff(t) == 1)
    psn = hdr[:, 6].astype(np.uint64) | (hdr[:, 7].astype(np.uint64) << np.uint64(32))
    assert psn[0] == 1 and np.all(np.diff(psn.astype(np.int64)) == 1)
    assert (int(f["secs_inre"][0]), int(f["df_num_insec"][0])) == (a.secs, a.df)
