"""CPU check of the index arithmetic of the band-trimmed snapshot C2R
(sdbe_amd/csrc/sdbe_trim.hpp, k_trim_c2r and k_trim_c2r_pk): thread t of 1024
builds elements k = t + 1024 q (q < 16) from channel ka = k (flipped:
N - 1 - k) and its mirror kb = (N - k) mod N (flipped: N - 1 - that): the
packed complex form of the 32768-point C2R that the reference runs as a library
C2R (vdif_inout_gpu_thread.cu:598-605), with the LSB spectrum reversed as
rslib.py:124 does.
The kernels form those channel indices, and the packet kernel the staged word
(channel / 4) and bit field (byte 3 - channel mod 4), as per-thread bases plus
+-1024 q (words +-256 q), clamping only the mirror of k = 0, which pairs with
the zero Nyquist bin and whose value is not used.  This restates that base-and-
step form and checks it against the definition for every thread, q and flip,
including the flipped branch that no API call reaches today (the polarisation
only adds a constant 0 or 4 to the bit position).  It also checks the segment
kernel's post-processing bases (k_trim_seg) against the R2C bin indices they
replace, at both ends of the band, and the base-plus-offset forms of the Stockham
passes' swizzled LDS addresses against the swizzle itself, for every pass run.
Runs no GPU code."""
import numpy as np

N = 16384


def _definition(t, q, flip):
    k = t + 1024 * q
    km = (N - k) & (N - 1)
    ka = N - 1 - k if flip else k
    kb = N - 1 - km if flip else km
    return k, ka, kb


def _kernel_form(t, q, flip):
    # k_trim_c2r: channel indices; k_trim_c2r_pk: staged words and byte fields
    da = -1024 if flip else 1024
    ea = N - 1 - t if flip else t
    eb = t - 1 if flip else N - t
    ka = ea + da * q
    kb = min(max(eb, 0), N - 1) if q == 0 else eb - da * q
    dw = -256 if flip else 256
    wa = (ea >> 2) + dw * q
    wb = min(max(eb >> 2, 0), N // 4 - 1) if q == 0 else (eb >> 2) - dw * q
    return ka, kb, wa, wb, 3 - (ea & 3), 3 - (eb & 3)


def test_c2r_channel_words_match_definition():
    checked = 0
    for flip in (0, 1):
        for t in range(1024):
            for q in range(16):
                k, ka, kb = _definition(t, q, flip)
                ka2, kb2, wa, wb, ba, bb = _kernel_form(t, q, flip)
                assert 0 <= ka2 < N and 0 <= kb2 < N and 0 <= wa < N // 4 and 0 <= wb < N // 4
                assert ka2 == ka and wa == ka >> 2 and ba == 3 - (ka & 3)
                if k != 0:   # the mirror of k = 0 is the zero Nyquist bin
                    assert kb2 == kb and wb == kb >> 2 and bb == 3 - (kb & 3)
                checked += 1
    assert checked == 2 * 1024 * 16


def test_c2r_flip_reverses_channels():
    # the flipped pairing reads channel N - 1 - c where the plain one reads c
    for t in range(0, 1024, 7):
        for q in range(16):
            _, ka0, kb0 = _definition(t, q, 0)
            _, ka1, kb1 = _definition(t, q, 1)
            assert ka1 == N - 1 - ka0
            if t + 1024 * q:
                assert kb1 == N - 1 - kb0
    assert np.array_equal(np.arange(N)[::-1][np.arange(N)], N - 1 - np.arange(N))


# Segment post-processing of k_trim_seg (sdbe_trim.hpp): element j = t + NT q of
# the C2R input needs X[k] and X[k'] of the R2C, k = skip + j, k' = skip + NP - j,
# each from Z[k] and Z[(NR - k) mod NR] (Z[NR] read as Z[0]).  The kernel reads
# them at per-thread bases plus non-negative multiples of NT:
#   k = ka + NT q, NR - k = kb + NT (JT-1-q), k' = kc + NT (JT-1-q), NR - k' = kd + NT q,
# with ka = skip + t, kc = skip + NT - t, kb = NR - NT (JT-1) - ka, kd = NR - NT (JT-1) - kc,
# and maps an index equal to NR to 0 at q = 0 only.
PRESETS = {6: (39, 32, 256), 8: (13, 8, 128), 10: (65, 64, 512), 11: (143, 128, 1024)}


def test_segment_postprocessing_bases():
    for frac, (D, E, NT) in PRESETS.items():
        R, P = 256 * D, 256 * E
        NR, NP, JT = R // 2, P // 2, P // 2 // NT
        bin_hz = 4576e6 * frac / 11 / R
        for skip in sorted({0, 1, NR - NP, (NR - NP) // 2, int(150e6 / bin_hz)}):
            if skip + NP > NR:
                continue
            t = np.arange(NT)
            ka, kc = skip + t, skip + NT - t
            kb, kd = NR - NT * (JT - 1) - ka, NR - NT * (JT - 1) - kc
            assert (kb >= 0).all() and (kd >= 0).all() and (kc >= 0).all()
            for q in range(JT):
                j = t + NT * q
                k, k2 = skip + j, skip + NP - j
                ia, ib = ka + NT * q, kb + NT * (JT - 1 - q)
                ic, idd = kc + NT * (JT - 1 - q), kd + NT * q
                if q == 0:
                    ib = np.where(ib == NR, 0, ib)
                    ic = np.where(ic == NR, 0, ic)
                assert np.array_equal(ia, np.where(k == NR, 0, k))
                assert np.array_equal(ib, (NR - k) % NR)
                assert np.array_equal(ic, np.where(k2 == NR, 0, k2))
                assert np.array_equal(idd, (NR - k2) % NR)
                assert max(ia.max(), ib.max(), ic.max(), idd.max()) < NR
                # the twiddle table xr has NR + 1 entries; k' = NR uses its last
                assert k2.max() <= NR and k.max() < NR


# LDS addressing of the Stockham passes (sdbe_trim.hpp pass_reads / pass_writes):
# the swizzle swz16(e) = e ^ ((e >> 4) & 15) and the decomposed forms the passes use
# in its place, chosen by the same compile-time conditions, for every pass run.
def _swz(e):
    return e ^ ((e >> 4) & 15)


def _swz_add(sb, c):
    return (sb ^ ((c & 128) >> 4)) + c


# (N, R, Ns, plain output) of each pass of every preset: the R2C's (its first pass
# reads global memory and writes as Ns = 1; its last writes the plain layout), then
# the C2R's (its last pass, None here, reads only)
PASSES = [(4992, 16, 1, False), (4992, 8, 16, False), (4992, 13, 128, False), (4992, 3, 1664, True),
          (1664, 16, 1, False), (1664, 8, 16, False), (1664, 13, 128, True),
          (8320, 16, 1, False), (8320, 8, 16, False), (8320, 5, 128, False), (8320, 13, 640, True),
          (18304, 16, 1, False), (18304, 8, 16, False), (18304, 11, 128, False), (18304, 13, 1408, True),
          (4096, 16, 1, False), (4096, 16, 16, False), (4096, 16, 256, None),
          (1024, 16, 1, False), (1024, 8, 16, False), (1024, 8, 128, None),
          (8192, 16, 1, False), (8192, 16, 16, False), (8192, 16, 256, False), (8192, 2, 4096, None)]


def test_pass_addressing_forms():
    for N, R, NS, lin in PASSES:
        NB = N // R
        b = np.arange(NB)
        for i in range(R):   # the reads (the per-access branch compares swz16 with itself)
            e = b + NB * i
            got = _swz_add(_swz(b), NB * i) if NB % 128 == 0 else _swz(e)
            assert np.array_equal(got, _swz(e)), (N, R, NS, i)
        if lin is None:
            continue   # pass_to_regs: no LDS writes
        base = (b // NS) * NS * R + (b % NS)
        sb = _swz(base)
        for k in range(R):
            e = base + NS * k
            if lin:
                got, want = e, e
            elif NS % 128 == 0:
                got, want = _swz_add(sb, NS * k), _swz(e)
            elif NS == 16 and R & (R - 1) == 0:
                got, want = (sb ^ k) + 16 * k, _swz(e)
            elif NS == 1 and R == 16:
                got, want = sb ^ k, _swz(e)
            else:
                got, want = _swz(e), _swz(e)
            assert np.array_equal(got, want), (N, R, NS, k)
            assert got.max() < N and got.min() >= 0
        # each pass's writes are a permutation of the N elements
        if not lin:
            allw = np.concatenate([_swz(base + NS * k) for k in range(R)])
            assert np.array_equal(np.sort(allw), np.arange(N))
