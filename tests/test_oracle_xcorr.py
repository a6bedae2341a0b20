"""Pins of the cross-correlation restatement (oracle.corr_Xt,
corr_Xt_search, corr_snr; software/verification/cross_corr.py:23-116,
verify_scan.py:231-236).

* brute force: the lag function of two short random window sets by explicit
  sums (cross-correlation in the time domain of the positive-frequency
  filtered windows is not needed: the definition is the inverse real DFT of the
  averaged cross-power, written here as a cosine/sine sum);
* identical inputs: s is the normalised autocorrelation, 1 at lag 0;
* a delayed copy: the peak sits at the delay (sample lead within the window,
  window offset across windows), with the sign convention of verify_scan.py:244-247
  (x1 leading x0 by L samples puts the peak at lag +L);
* the search indexes windows from X0's centre (Python 2 integer division).
"""
import numpy as np

from oracle import sdbe_oracle as O


def test_bruteforce_lag_function():
    rng = np.random.default_rng(3)
    n = 16
    X0 = rng.normal(size=(3, n // 2)) + 1j * rng.normal(size=(3, n // 2))
    X1 = rng.normal(size=(3, n // 2)) + 1j * rng.normal(size=(3, n // 2))
    s, S = O.corr_Xt(X0, X1, fft_window_size=n)

    def irfft0(P):      # inverse real DFT of P[0..n/2-1] (P[n/2] = 0) by the sum
        m = np.arange(n)
        out = np.full(n, P[0].real / n)
        for k in range(1, n // 2):
            out += 2 * (P[k] * np.exp(2j * np.pi * k * m / n)).real / n
        return out
    S01 = np.mean(X0 * np.conj(X1), axis=0)
    a0 = irfft0(np.mean(np.abs(X0) ** 2, axis=0))
    a1 = irfft0(np.mean(np.abs(X1) ** 2, axis=0))
    want = irfft0(S01) / np.sqrt(a0.max() * a1.max())
    assert np.allclose(S, S01) and np.allclose(s, want, atol=1e-12)
    # the autocorrelation peaks at lag 0 (the normalisation uses its max)
    assert np.argmax(a0) == 0 and np.argmax(a1) == 0


def test_identical_streams_give_unit_peak_at_zero():
    rng = np.random.default_rng(4)
    x = rng.normal(size=40 * 256)
    X = O.corr_spectra(x, 256)
    s, S, peaks = O.corr_Xt_search(X, X, np.arange(-3, 4), search_avg=8, fft_window_size=256)
    assert np.argmax(peaks) == 3 and abs(peaks[3] - 1.0) < 1e-12 and np.argmax(np.abs(s[3])) == 0


def test_delayed_copy_peaks_at_the_delay():
    rng = np.random.default_rng(5)
    n, nw = 1024, 48
    common = rng.normal(size=(nw + 4) * n)
    lag_windows, lead = 2, 37
    x0 = common[4 * n:4 * n + nw * n] + 0.3 * rng.normal(size=nw * n)
    x1 = np.roll(common, -(lag_windows * n + lead))[4 * n:4 * n + nw * n] + 0.3 * rng.normal(size=nw * n)
    # x1 is x0's signal advanced by lag_windows windows + lead samples
    X0, X1 = O.corr_spectra(x0, n), O.corr_spectra(x1, n)
    r = np.arange(-4, 5)
    s, S, peaks = O.corr_Xt_search(X0, X1, r, search_avg=16, fft_window_size=n)
    snr, sig, i = O.corr_snr(s, peaks)
    assert r[i] == -lag_windows
    # S_0x1 = X0 conj(X1): s[tau] = sum_t x0[t + tau] x1[t] peaks where x1 leads x0
    k = int(np.argmax(np.abs(s[i])))
    assert (k - n if k > n // 2 else k) == lead
    assert snr > 20
