"""Pins of the oracle's polyphase filterbank and least-squares inverse (A17).

* windows against numpy's own hamming / hanning / sinc;
* the forward PFB against its definition (window, rfft of the whole
  ntap * lblock segment, every ntap-th bin; pfb_inverse.cu:172-208), which the
  oracle computes differently (fold, then a short rfft);
* the inverse against the Moore-Penrose pseudo-inverse of the dense linear
  map x -> irfft(pfb(x)) built column by column from unit impulses (the
  minimum-norm least-squares solution the method defines), and
* at full width (lblock 32768) consistency: pfb(inverse(s)) == s.
"""
import numpy as np
import pytest

from oracle import sdbe_oracle as O


def test_windows_match_numpy():
    m = 4 * 32768
    assert np.allclose(O.pfb_window_hamming(), np.hamming(m), rtol=0, atol=1e-15)
    n = np.arange(m)
    want = np.hanning(m) * np.sinc((n - (m - 1) / 2) / 32768)
    assert np.allclose(O.pfb_window_hann_sinc(), want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("window", ["hamming", "hann_sinc"])
def test_forward_matches_definition(window):
    rng = np.random.default_rng(1)
    lb, nt = 16, 4
    w = getattr(O, "pfb_window_" + window)(nt, lb)
    x = rng.normal(size=(9 + nt - 1) * lb)
    s = O.pfb_forward(x, w, lb)
    nb = x.size // lb - (nt - 1)
    want = np.stack([np.fft.rfft(w * x[i * lb:i * lb + nt * lb])[::nt] for i in range(nb)])
    assert s.shape == (9, lb // 2 + 1)
    assert np.abs(s - want).max() < 1e-12


@pytest.mark.parametrize("window", ["hamming", "hann_sinc"])
def test_inverse_is_the_pseudo_inverse(window):
    rng = np.random.default_rng(2)
    lb, nt, nb = 8, 4, 10
    w = getattr(O, "pfb_window_" + window)(nt, lb)
    n = (nb + nt - 1) * lb
    x = rng.normal(size=n)
    s = O.pfb_forward(x, w, lb)
    cols = [np.fft.irfft(O.pfb_forward(np.eye(n)[k], w, lb), n=lb, axis=-1).ravel() for k in range(n)]
    M = np.stack(cols, axis=1)
    yh = np.fft.irfft(s, n=lb, axis=-1).ravel()
    want = np.linalg.pinv(M) @ yh
    got = O.pfb_inverse(s, w, lb)
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()
    # the reconstruction differs from x only in the null space of the map
    assert np.abs(M @ (got - x)).max() < 1e-9


def test_full_width_consistency():
    rng = np.random.default_rng(3)
    w = O.pfb_window_hann_sinc()
    x = rng.normal(size=(6 + 3) * O.PFB_LBLOCK)
    s = O.pfb_forward(x, w)
    xh = O.pfb_inverse(s, w)
    assert xh.size == x.size
    assert np.abs(O.pfb_forward(xh, w) - s).max() < 1e-9 * np.abs(s).max()
