"""Pins of the remaining oracle helpers on the CPU: spectrum inflation,
sideband flip, block slicing, the batched CPU resampler, and the B-engine
timestamp arithmetic of beng_reform.c.

The timestamp helpers are checked against the absolute-clock rule they encode
(sec * 286e6 + clk compared with a 1024-clock tolerance, B-frames of 262144
clocks), so a wrong branch, tolerance or wrap fails a property rather than a
retyped formula.  The two places where the reference departs from that rule
are pinned as known answers and named.
"""
import numpy as np

from oracle import sdbe_oracle as O

CLK = O.SWARM_CLK_PER_SEC
TOL = O.BFRAME_CLK_TOL
FRAME = O.SWARM_BFRAME_CLK


def _abs(t):
    return t[0] * CLK + t[1]


def test_inflate_through_the_packet_decode():
    """One hand-built packet: snapshot 5, channel 0 carries pol0 = 1 - 2j and
    pol1 = -1 + 0j (codes 3/0 and 1/2); every other byte is 0xAA (all zero).
    inflate() of the unpacked rows must give exactly those complex values,
    which fixes the (im, re) order of the int8 rows and the appended bin."""
    pkt = np.zeros(O.PKT_WORDS, dtype=np.uint32)
    b_zero = 777
    pkt[4] = b_zero >> 8
    pkt[5] = 0 | (0 << 16) | ((b_zero & 0xFF) << 24)
    pay = np.full(1024, 0xAA, dtype=np.uint8)
    # byte bits from the LSB: pol1 im, pol1 re, pol0 im, pol0 re; byte 3 of a
    # snapshot is channel offset 0
    pay[8 * 5 + 3] = 2 | (1 << 2) | (0 << 4) | (3 << 6)
    pkt[8:] = pay.view(np.uint32)
    un = O.vdif_to_beng(pkt[None, :], b_zero)
    X0 = O.inflate(un.beng[0].reshape(-1, O.BENG_CHANNELS, 2)[:39])
    X1 = O.inflate(un.beng[1].reshape(-1, O.BENG_CHANNELS, 2)[:39])
    assert X0.shape == (39, O.BENG_CHANNELS + 1)
    assert X0[5, 0] == 1 - 2j and X1[5, 0] == -1 + 0j
    others = np.ones(X0.shape, dtype=bool)
    others[5, 0] = False
    assert not X0[others].any() and not X1[others].any()


def test_flip_sideband_is_channel_reversal():
    rng = np.random.default_rng(3)
    X = rng.normal(size=(2, O.BENG_CHANNELS + 1)) + 1j * rng.normal(size=(2, O.BENG_CHANNELS + 1))
    X[:, -1] = 0
    Y = O.flip_sideband(X)
    for k in (0, 1, 150, 8191, 8192, 16382, 16383):
        assert np.array_equal(Y[:, k], X[:, 16383 - k])
    assert not Y[:, -1].any()
    assert np.array_equal(O.flip_sideband(Y), X)      # an involution


def test_block_spectra_takes_consecutive_snapshots():
    """Snapshot index = frame * 128 + snapshot-in-frame; block b is snapshots
    39b .. 39b+38.  Each snapshot's channel 0/1 encode its index."""
    n = 39 * 128
    beng = np.zeros((n, O.BENG_CHANNELS, 2), dtype=np.int8)
    s = np.arange(n)
    beng[:, 0, 1] = (s % 128).astype(np.int8)
    beng[:, 1, 1] = (s // 128).astype(np.int8)
    for b in (0, 1, 77, 127):
        X = O.block_spectra(beng, b)
        idx = X[:, 0].real.astype(int) + 128 * X[:, 1].real.astype(int)
        assert list(idx) == list(range(39 * b, 39 * b + 39))


def test_cpu_resample_batch_equals_per_block_oracle():
    rng = np.random.default_rng(4)
    X = rng.integers(-2, 2, (2, 39, O.BENG_CHANNELS, 2)).astype(np.int8)
    Xc = O.inflate(X)
    got = O.cpu_resample_batch(Xc, workers=2)
    for i in range(2):
        want = O.resample_block(Xc[i])
        assert np.abs(got[i] - want).max() <= 1e-12 * np.abs(want).max()


def test_read_beng_timestamp_inverts_the_53_bit_field():
    for sec, clk in [(0, 0), (1, CLK - 1), (123456, 285_999_999), ((1 << 24) - 1, (1 << 29) - 1)]:
        b53 = (sec << 29) | clk
        w4 = (b53 >> 21) & 0xFFFFFFFF
        w5 = ((b53 & 0x1FFFFF) << 11) | 0x7FF           # low 11 bits are other fields
        assert O.read_beng_timestamp(w4, w5) == [sec, clk]


def test_ts_compare_is_the_tolerant_clock_order():
    """For timestamps at most one second apart, ts_compare(then, now) is
    -1 / 0 / +1 as now - then is above TOL / within it / below -TOL."""
    rng = np.random.default_rng(5)
    n_checked = 0
    for _ in range(20000):
        s = int(rng.integers(10, 20))
        then = [s, int(rng.integers(0, CLK))]
        ds = int(rng.integers(-1, 2))
        near = rng.random() < 0.7
        if near:       # concentrate on the tolerance edges and second wraps
            d = int(rng.integers(-3 * TOL, 3 * TOL))
            a = _abs(then) + d
            now = [a // CLK, a % CLK]
            if abs(now[0] - then[0]) > 1:
                continue
        else:
            now = [s + ds, int(rng.integers(0, CLK))]
        d = _abs(now) - _abs(then)
        if abs(d) == TOL:
            continue
        want = -1 if d > TOL else (1 if d < -TOL else 0)
        assert O.ts_compare(then, now) == want, (then, now)
        n_checked += 1
    assert n_checked > 15000
    # edges inside one second and across a wrap
    assert O.ts_compare([5, 1000], [5, 1000 + TOL]) == 0
    assert O.ts_compare([5, 1000], [5, 1001 + TOL]) == -1
    assert O.ts_compare([5, CLK - 500], [6, 200]) == 0
    assert O.ts_compare([6, 200], [5, CLK - 500]) == 0


def test_ts_compare_reference_departures():
    """Two known answers where beng_reform.c:34-61 departs from the clock
    order: (1) exactly TOL apart across a second boundary with now.clk = 0
    reads as 'earlier', not 'same'; (2) when `then` is several seconds later
    but `now` sits within TOL of the end of its second, the wrap branch does
    not check the second difference and reads 'same'."""
    assert O.ts_compare([5, CLK - TOL], [6, 0]) == -1
    assert O.ts_compare([9, 10], [4, CLK - 100]) == 0


def test_ts_increment_decrement_step_one_frame():
    rng = np.random.default_rng(6)
    for _ in range(5000):
        t = [int(rng.integers(0, 1000)), int(rng.integers(0, CLK))]
        up = O.ts_increment(t)
        if t[1] + FRAME != CLK:
            assert _abs(up) == _abs(t) + FRAME and 0 <= up[1] < CLK
        dn = O.ts_decrement(t)
        assert _abs(dn) == _abs(t) - FRAME and 0 <= dn[1] < CLK
    # the reference wraps on '>' only: landing exactly on the second leaves clk = CLK
    assert O.ts_increment([3, CLK - FRAME]) == [3, CLK]


def test_ts_offset_counts_frames_and_caps():
    rng = np.random.default_rng(7)
    for _ in range(300):
        then = [int(rng.integers(100, 200)), int(rng.integers(0, CLK))]
        k = int(rng.integers(0, 1100))
        jitter = int(rng.integers(-TOL // 2, TOL // 2))
        sign = 1 if rng.random() < 0.5 else -1
        a = _abs(then) + sign * k * FRAME + jitter
        now = [a // CLK, a % CLK]
        off = O.ts_offset(then, now)
        want = sign * min(k, O.MAX_COUNTABLE_BFRAME_OFFSET + 1)
        if k == 0:
            want = 0
        assert off == want, (then, now, k, sign)


def test_reference_policy_relatches_while_sigma_is_zero():
    """The REFERENCE threshold walk (vdif_inout_gpu_thread.cu:954-1005): the
    latch is taken from polarisation 0 and recomputed while the latched
    threshold is 0.  Block 0 has an empty pol 0 (sigma 0) and a noisy pol 1;
    block 1 has noise in both; block 2 has noise of another scale.
    Expected, from the rule alone:
    * block 0: th = off = 0 for both pols, so pol 0 (all zero) is code 3 and
      pol 1 is 3 where x >= 0, 0 where x < 0;
    * blocks 1 and 2: numpy's mean and std(ddof=1) of block 1's pol-0 samples.
    Latching from pol 1, never recomputing, or recomputing every block fails."""
    rng = np.random.default_rng(5)
    beng = np.zeros((2, 3 * O.DECIMATION, O.BENG_CHANNELS, 2), dtype=np.int8)
    beng[1, :39] = rng.integers(-2, 2, size=(39, O.BENG_CHANNELS, 2))
    beng[:, 39:78] = rng.integers(-2, 2, size=(2, 39, O.BENG_CHANNELS, 2))
    beng[:, 78:] = 2 * rng.integers(-1, 1, size=(2, 39, O.BENG_CHANNELS, 2))
    out = O.process_blocks(beng, [0, 1, 2], policy=O.THRESH_REFERENCE, latch=[None])
    x01 = O.resample_block(O.block_spectra(beng[1], 0))
    x10 = O.resample_block(O.block_spectra(beng[0], 1))
    assert not np.any(out.samples[(0, 0)])
    assert np.array_equal(out.codes[(0, 0)], np.full(x01.size, 3, np.uint8))
    assert np.array_equal(out.codes[(0, 1)], np.where(x01 >= 0, 3, 0).astype(np.uint8))
    th, off = float(np.std(x10, ddof=1)), float(np.mean(x10))
    for blk in (1, 2):
        for pol in (0, 1):
            got_th, got_off = out.thresholds[(blk, pol)]
            assert abs(got_th - th) <= 1e-12 * th and abs(got_off - off) <= 1e-12 * th
    x21 = O.resample_block(O.block_spectra(beng[1], 2))
    v = x21 - off
    want = np.select([v >= th, v >= 0, v > -th], [3, 2, 1], 0).astype(np.uint8)
    assert np.array_equal(out.codes[(2, 1)], want)
