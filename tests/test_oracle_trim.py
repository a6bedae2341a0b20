"""Pins of the band-trimmed resampler restatement (oracle.resample_trimmed,
trim_geometry; software/aphids/vdif_inout_gpu_thread.cu:589-709, plans
:269-333, presets sample_rates.h:12-40).

* brute force: on tiny sizes every transform of the chain is written as an
  explicit DFT sum (no FFT routine), including the numpy/cuFFT C2R treatment
  of bins 0 and N/2 (real parts only) -- a wrong segment length, skip
  direction, normalisation or Hermitian rule fails it;
* identity: with E = D and skip 0 the R2C / C2R pair is the identity, so the
  output is the concatenated per-snapshot C2R series;
* tone: a real tone on the R-point grid comes out as a tone at bin k0 - skip
  of the P-point grid with amplitude R/P times the input's;
* geometry: skip_chan of the presets (250 kHz and 1 MHz R2C bins).
"""
import numpy as np
import pytest

from oracle import sdbe_oracle as O


def _irfft_bruteforce(Y, n):
    """Inverse real DFT by the explicit sum (imaginary parts of bin 0 and n/2 dropped)."""
    m = np.arange(n)
    out = np.full(n, Y[0].real / n)
    for j in range(1, n // 2):
        out += 2.0 * (Y[j] * np.exp(2j * np.pi * j * m / n)).real / n
    out += Y[n // 2].real * np.cos(np.pi * m) / n
    return out


def _rfft_bruteforce(x):
    n = x.size
    k = np.arange(n // 2 + 1)[:, None]
    return (x[None, :] * np.exp(-2j * np.pi * k * np.arange(n)[None, :] / n)).sum(axis=1)


@pytest.mark.parametrize("D,E,skip,scale,nchan", [(3, 2, 1, 2, 4), (3, 2, 0, 4, 6), (5, 3, 2, 2, 5)])
def test_bruteforce_tiny(D, E, skip, scale, nchan):
    rng = np.random.default_rng(D * 100 + E * 10 + skip)
    X = rng.normal(size=(D, nchan)) + 1j * rng.normal(size=(D, nchan))
    got = O.resample_trimmed(X, E, skip, scale=scale)
    n1 = 2 * nchan
    t = np.concatenate([_irfft_bruteforce(np.append(X[s], 0.0), n1) for s in range(D)])
    R, P = scale * D, scale * E
    want = []
    for g in range(t.size // R):
        Xs = _rfft_bruteforce(t[g * R:(g + 1) * R])
        want.append(_irfft_bruteforce(Xs[skip:skip + P // 2 + 1], P))
    want = np.concatenate(want)
    assert got.shape == want.shape
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())


def test_identity_when_expansion_equals_decimation():
    rng = np.random.default_rng(7)
    D, nchan = 13, 16384
    X = rng.normal(size=(D, nchan)) + 1j * rng.normal(size=(D, nchan))
    X[:, 0] = X[:, 0].real            # the C2R drops bin 0's imaginary part either way
    got = O.resample_trimmed(X, D, 0)
    t = np.fft.irfft(np.concatenate([X, np.zeros((D, 1))], axis=1), n=2 * nchan, axis=1).reshape(-1)
    assert np.allclose(got, t, rtol=0, atol=1e-12 * np.abs(t).max())


def test_tone_moves_down_by_skip():
    # build the spectra from a real cosine that is periodic over R and over a
    # snapshot (k0 * 2 nchan / R integer): the per-snapshot C2R returns it exactly
    D, E, skip, nchan = 39, 32, 600, 16384
    R, P = 256 * D, 256 * E
    n1 = 2 * nchan
    k0 = 39 * 40                               # R2C bin; k0 * n1 / R = 5120 is an integer
    n = np.arange(D * n1)
    t = 0.75 * np.cos(2 * np.pi * k0 * n / R + 0.3)
    X = np.fft.rfft(t.reshape(D, n1), axis=1)[:, :nchan]
    assert np.abs(np.fft.rfft(t.reshape(D, n1), axis=1)[:, nchan]).max() < 1e-6
    y = O.resample_trimmed(X, E, skip)
    m = np.arange(P)
    # every segment starts a whole number of periods in, so all share the phase
    want = (0.75 * R / P) * np.cos(2 * np.pi * (k0 - skip) * m[None, :] / P + 0.3)
    assert np.allclose(y.reshape(-1, P), want, rtol=0, atol=1e-9)


def test_preset_geometry():
    assert O.trim_geometry(6, 150) == (39, 32, 9984, 8192, 600)
    assert O.trim_geometry(6, 102) == (39, 32, 9984, 8192, 408)
    assert O.trim_geometry(8, 150) == (13, 8, 3328, 2048, 150)
    # 10/11: 4160 Msps, R = 16640 (250 kHz bins); 11/11: 4576 Msps, R = 36608 (125 kHz bins)
    assert O.trim_geometry(10, 0) == (65, 64, 16640, 16384, 0)
    assert O.trim_geometry(10, 32) == (65, 64, 16640, 16384, 128)
    assert O.trim_geometry(11, 150) == (143, 128, 36608, 32768, 1200)
    with pytest.raises(ValueError):
        O.trim_geometry(10, 33)               # 132 + 8192 > 8320 bins
    with pytest.raises(ValueError):
        O.trim_geometry(6, 300)               # 1200 + 4096 > 4992 bins
