"""Pins of the oracle's resampler, statistics and quantiser (A8-A13) on the CPU.

Each check is chosen so that a plausible mistake fails it: brute-force DFTs on
tiny inputs fix numpy's C2R conventions (DC / Nyquist imaginary parts dropped),
an integer-bin tone fixes the composition and its 39/64 normalisation, the
band limit and linearity fix the zero padding, Chan's combine is compared with
a direct two-pass moment, and the quantiser with the rule stated in words.
"""
import math

import numpy as np
import pytest

from oracle import sdbe_oracle as O


def _brute_irfft(X, n):
    """Direct inverse real DFT with the Hermitian extension built by hand;
    the imaginary parts of bin 0 and bin n/2 carry no information in a real
    signal and are taken as zero."""
    X = np.array(X, dtype=np.complex128)
    full = np.zeros(n, dtype=np.complex128)
    full[: n // 2 + 1] = X
    full[0] = X[0].real
    full[n // 2] = X[n // 2].real
    for k in range(1, n // 2):
        full[n - k] = np.conj(X[k])
    t = np.arange(n)
    return np.array([np.sum(full * np.exp(2j * np.pi * np.arange(n) * tt / n)).real for tt in t]) / n


def test_irfft_drops_dc_and_nyquist_imag():
    rng = np.random.default_rng(0)
    for n in (8, 12, 16):
        X = rng.normal(size=n // 2 + 1) + 1j * rng.normal(size=n // 2 + 1)
        assert np.allclose(np.fft.irfft(X, n=n), _brute_irfft(X, n), atol=1e-12)


def test_stage_shapes_and_block_layout():
    X = np.zeros((39, 16385), dtype=np.complex128)
    assert O.stage_a(X).shape == (39, 32768)
    assert O.stage_b(np.zeros(O.M_R2C)).shape == (638977,)
    assert O.stage_c(np.zeros(638977, dtype=complex)).shape == (2097152,)
    assert O.OUT_FRAMES_PER_BLOCK == 64 and O.L_C2R * 2 // 8 == 64 * 8192


def test_integer_bin_tone_resamples_to_analytic_cosine():
    """y[n] = cos(2 pi m0 n / M + phi) over one block is bin m0 of the block
    DFT; the resampled output is (M/L) cos(2 pi m0 t / L + phi) exactly."""
    m0, phi = 3 * 39 * 7 + 5, 0.3
    n = np.arange(O.M_R2C)
    y = 0.25 * np.cos(2 * np.pi * m0 * n / O.M_R2C + phi)
    X = np.fft.rfft(y.reshape(39, 32768), axis=-1)
    z = O.resample_block(X)
    t = np.arange(O.L_C2R)
    za = (39 / 64) * 0.25 * np.cos(2 * np.pi * m0 * t / O.L_C2R + phi)
    assert np.abs(z - za).max() < 1e-12
    # a mistake in the padding length or in which bins are copied moves the tone
    assert np.argmax(np.abs(np.fft.rfft(z))) == m0


def test_resampled_spectrum_band_limit_and_linearity():
    rng = np.random.default_rng(3)
    X1 = rng.integers(-2, 2, (39, 16385)) + 1j * rng.integers(-2, 2, (39, 16385))
    X2 = rng.integers(-2, 2, (39, 16385)) + 1j * rng.integers(-2, 2, (39, 16385))
    X1[:, 16384] = 0
    X2[:, 16384] = 0
    z1, z2 = O.resample_block(X1), O.resample_block(X2)
    z12 = O.resample_block(2 * X1 - 3 * X2)
    assert np.abs(z12 - (2 * z1 - 3 * z2)).max() < 1e-12
    Z = np.fft.rfft(z1)
    Y = O.stage_b(O.stage_a(X1).reshape(-1))
    # bins above the old Nyquist are empty; below it the spectrum is Y itself,
    # the old Nyquist bin copied as is (not halved)
    assert np.abs(Z[638977:]).max() < 1e-9 * np.abs(Z).max()
    assert np.allclose(Z[:638977], Y, atol=1e-9 * np.abs(Y).max())
    # Parseval on the resampled block: what the GPU uses for the variance
    Yp = np.zeros(O.L_C2R // 2 + 1, dtype=complex)
    Yp[:Y.size] = Y
    w = np.full(Yp.size, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    energy = (w[0] * Yp[0].real ** 2 + np.sum(w[1:-1] * np.abs(Yp[1:-1]) ** 2)) / O.L_C2R ** 2 * O.L_C2R
    assert abs(energy - np.sum(z1 ** 2)) < 1e-9 * np.sum(z1 ** 2)


def test_chan_combine_equals_two_pass_moments():
    rng = np.random.default_rng(4)
    x = rng.normal(3.0, 2.0, 10001)
    parts = np.split(x, [17, 500, 501, 7000])
    trip = [(p.size, p.mean(), ((p - p.mean()) ** 2).sum()) for p in parts]
    acc = trip[0]
    for tr in trip[1:]:
        acc = O.chan_combine(acc, tr)
    mean, sig = O.block_stats(x)
    assert acc[0] == x.size
    assert abs(acc[1] - mean) < 1e-12
    assert abs(math.sqrt(acc[2] / (acc[0] - 1)) - sig) < 1e-12
    assert abs(sig - x.std(ddof=1)) < 1e-12


def test_quantizer_boundaries():
    th = 1.5
    v = np.array([th, np.nextafter(th, 0), 0.0, -0.0, np.nextafter(0, -1), -np.nextafter(th, 0), -th,
                  -th - 1, np.inf, -np.inf, np.nan])
    codes = O.quantize2bit(v, th, 0.0)
    # v >= th -> 3; 0 <= v < th -> 2; -th < v < 0 -> 1; v <= -th -> 0; NaN -> 2
    assert list(codes) == [3, 2, 2, 2, 1, 1, 0, 0, 3, 0, 2]
    # the offset is subtracted before the comparison
    assert list(O.quantize2bit(np.array([1.75, 0.25, 0.2, -1.25, -1.3]), th, 0.25)) == [3, 2, 1, 0, 0]
    # the prototype's CPU quantiser differs only at v == -th (code 1 there)
    py = O.quantize_to_2bit_python(np.array([-th, -th - 1e-9, -th + 1e-9, 0.0, th]), th) + 2
    gpu = O.quantize2bit(np.array([-th, -th - 1e-9, -th + 1e-9, 0.0, th]), th, 0.0)
    assert list(py) == [1, 0, 1, 2, 3] and list(gpu) == [0, 0, 1, 2, 3]


def test_pack_codes_bit_order():
    codes = np.tile(np.array([0, 1, 2, 3], dtype=np.uint8), 4)
    assert O.pack_codes(codes)[0] == 0xE4E4E4E4      # sample i at bits 2i
    codes = np.zeros(32, dtype=np.uint8)
    codes[17] = 3
    assert list(O.pack_codes(codes)) == [0, 3 << 2]
    rng = np.random.default_rng(5)
    c = rng.integers(0, 4, 16 * 100).astype(np.uint8)
    assert np.array_equal(O.unpack_codes(O.pack_codes(c)), c)


def test_cpu_chain_equals_oracle_blocks():
    from sdbe_amd import synth as S
    rng = S.rng_for(6)
    c0 = S.gaussian_codes(rng, 39 * 128)
    c1 = S.gaussian_codes(rng, 39 * 128)
    pk = S.packetize(c0, c1, 100)
    words, _ = O.cpu_chain(pk, 100, [0, 5], workers=2)
    un = O.vdif_to_beng(pk, 100)
    ref = O.process_blocks(un.beng, [0, 5], policy=O.THRESH_PER_BLOCK)
    assert np.array_equal(words.reshape(2, 2, -1)[:, 0], ref.payload[0].reshape(2, -1))
    assert np.array_equal(words.reshape(2, 2, -1)[:, 1], ref.payload[1].reshape(2, -1))
