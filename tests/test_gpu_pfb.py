"""GPU tests of the polyphase filterbank path (sdbe_pfb_*) against the FP64
oracle (oracle.pfb_forward / oracle.pfb_inverse, the latter pinned against
the pseudo-inverse in tests/test_oracle_pfb.py).

Bars (stated as SNR = 10 log10(sum ref^2 / sum (gpu - ref)^2)):
* forward spectra vs FP64: >= 120 dB (FP32 FFT rounding);
* inverse time series vs FP64 on the same spectra: >= 90 dB (FP32 C2R into
  the FP64 banded solve, amplified by the solve's conditioning);
* GPU forward of the GPU inverse reproduces the spectra: >= 100 dB;
* config 5 (4096 spectra of noise + chirp, 2-bit per channel, Hann-sinc
  prototype): vs FP64 >= 80 dB; reconstruction of the original signal from
  the 2-bit spectra equal to FP64's within 0.01 dB (about 3.25 dB: the
  least-squares inverse's response to the quantisation noise, not a numerical
  error);
* span boundaries of the segmented solve (256, 257, 512, 513 spectra): >= 100 dB.
The oracle is given the FP32-rounded window, so only computation differs.
"""
import math
import sys

import numpy as np
import pytest

from oracle import sdbe_oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.filterwarnings("error::numpy.exceptions.ComplexWarning")]

LB = 32768


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _snr(ref, got):
    """SNR of got against ref; complex arrays are compared as complex (both
    the real and the imaginary parts enter), real ones as real."""
    ref = np.asarray(ref)
    got = np.asarray(got)
    assert ref.shape == got.shape and np.iscomplexobj(ref) == np.iscomplexobj(got)
    dt = np.complex128 if np.iscomplexobj(ref) else np.float64
    ref = ref.astype(dt).ravel()
    err = got.astype(dt).ravel() - ref
    return 10.0 * math.log10(np.sum(np.abs(ref) ** 2) / max(np.sum(np.abs(err) ** 2), 1e-300))


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _window(kind):
    w = O.pfb_window_hann_sinc() if kind == "hann_sinc" else O.pfb_window_hamming()
    return w.astype(np.float32)


@pytest.mark.parametrize("kind", ["hann_sinc", "hamming"])
def test_forward_matches_oracle(torch_cuda, kind):
    torch = torch_cuda
    from sdbe_amd.pfb import Pfb
    w32 = _window(kind)
    rng = np.random.default_rng(11)
    n = 64
    x = rng.normal(size=(n + 3) * LB).astype(np.float32)
    p = Pfb(w32, max_spectra=n)
    _say("[pfb] forward %s n=%d" % (kind, n))
    got = p.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    want = O.pfb_forward(x.astype(np.float64), w32.astype(np.float64))
    snr = _snr(want, got)
    _say("[pfb] forward %s: %.1f dB vs FP64" % (kind, snr))
    assert got.shape == (n, 16385)
    assert snr >= 120.0
    p.close()


@pytest.mark.parametrize("kind", ["hann_sinc", "hamming"])
def test_inverse_matches_oracle(torch_cuda, kind):
    torch = torch_cuda
    from sdbe_amd.pfb import Pfb
    w32 = _window(kind)
    rng = np.random.default_rng(12)
    n = 200                                             # not a multiple of the solve's 16-row chunks
    x = rng.normal(size=(n + 3) * LB)
    spec = O.pfb_forward(x, w32.astype(np.float64)).astype(np.complex64)
    p = Pfb(w32, max_spectra=256)
    _say("[pfb] inverse %s n=%d" % (kind, n))
    got = p.inverse(torch.from_numpy(spec).cuda()).cpu().numpy()
    want = O.pfb_inverse(spec.astype(np.complex128), w32.astype(np.float64))
    snr = _snr(want, got)
    _say("[pfb] inverse %s: %.1f dB vs FP64" % (kind, snr))
    assert got.shape == ((n + 3) * LB,)
    assert snr >= 90.0
    # GPU forward of the GPU inverse gives the spectra back
    back = p.forward(torch.from_numpy(got).cuda()).cpu().numpy()
    rt = _snr(spec, back)
    _say("[pfb] forward(inverse) %s: %.1f dB" % (kind, rt))
    assert rt >= 100.0
    p.close()


def test_small_batches_and_errors(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.pfb import Pfb
    w32 = _window("hann_sinc")
    rng = np.random.default_rng(13)
    p = Pfb(w32, max_spectra=8)
    for n in (1, 2, 3, 5):
        spec = (rng.normal(size=(n, 16385)) + 1j * rng.normal(size=(n, 16385))).astype(np.complex64)
        got = p.inverse(torch.from_numpy(spec).cuda()).cpu().numpy()
        want = O.pfb_inverse(spec.astype(np.complex128), w32.astype(np.float64))
        snr = _snr(want, got)
        _say("[pfb] inverse n=%d: %.1f dB" % (n, snr))
        assert snr >= 100.0
    with pytest.raises(L.SdbeError):
        p.inverse(torch.zeros((9, 16385), dtype=torch.complex64, device="cuda"))
    p.close()


def test_span_boundaries(torch_cuda):
    """The banded solve runs in 256-row spans joined by a scan: one full span,
    a full span plus one row, two full spans, and two plus one row (a last span
    of one row), on a plan whose capacity (600) is not a span multiple."""
    torch = torch_cuda
    from sdbe_amd.pfb import Pfb
    w32 = _window("hann_sinc")
    rng = np.random.default_rng(14)
    p = Pfb(w32, max_spectra=600)
    for n in (256, 257, 512, 513):
        spec = (rng.normal(size=(n, 16385)) + 1j * rng.normal(size=(n, 16385))).astype(np.complex64)
        _say("[pfb] inverse n=%d (span boundaries)" % n)
        got = p.inverse(torch.from_numpy(spec).cuda()).cpu().numpy()
        want = O.pfb_inverse(spec.astype(np.complex128), w32.astype(np.float64))
        snr = _snr(want, got)
        _say("[pfb] inverse n=%d: %.1f dB" % (n, snr))
        assert snr >= 100.0
    p.close()


def test_config5_reconstruction_snr(torch_cuda):
    """4096 spectra from a forward PFB of white Gaussian noise plus a linear
    chirp, each channel's re / im quantised to 2 bits (thresholds 0, +-0.9816
    sigma_c, levels +-0.4528 / +-1.510 sigma_c), inverted on the GPU."""
    torch = torch_cuda
    from sdbe_amd.pfb import Pfb
    w32 = _window("hann_sinc")
    n = 4096
    rng = np.random.default_rng(5005)
    ns = (n + 3) * LB
    t = np.arange(ns, dtype=np.float64)
    f0, f1 = 0.05, 0.45                                 # cycles per sample, swept over the record
    chirp = 0.5 * np.cos(2 * np.pi * (f0 * t + 0.5 * (f1 - f0) * t * t / ns))
    x = (rng.normal(size=ns) + chirp).astype(np.float32)
    p = Pfb(w32, max_spectra=n)
    _say("[pfb5] forward PFB of %d spectra" % n)
    spec = p.forward(torch.from_numpy(x).cuda()).cpu().numpy().astype(np.complex128)
    sig_c = np.sqrt(0.5 * np.mean(np.abs(spec) ** 2, axis=0))
    sig_c[sig_c == 0] = 1.0

    def q2(v):
        a = np.abs(v) / sig_c
        return np.sign(v) * np.where(a < 0.9816, 0.4528, 1.510) * sig_c

    sq = (q2(spec.real) + 1j * q2(spec.imag)).astype(np.complex64)
    _say("[pfb5] GPU inverse")
    got = p.inverse(torch.from_numpy(sq).cuda()).cpu().numpy()
    _say("[pfb5] FP64 oracle inverse")
    want = O.pfb_inverse(sq.astype(np.complex128), w32.astype(np.float64))
    snr_num = _snr(want, got)
    inner = slice(8 * LB, (n - 5) * LB)
    snr_sig = _snr(x[inner], got[inner])
    snr_sig64 = _snr(x[inner], want[inner])
    _say("[pfb5] n=%d: GPU vs FP64 %.1f dB; reconstruction of the signal from 2-bit spectra: GPU %.2f dB, "
         "FP64 %.2f dB" % (n, snr_num, snr_sig, snr_sig64))
    assert snr_num >= 80.0
    assert abs(snr_sig - snr_sig64) < 0.01
    assert snr_sig > 3.0
    p.close()
