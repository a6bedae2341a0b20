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
only adds a constant 0 or 4 to the bit position).
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
