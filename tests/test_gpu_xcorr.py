"""GPU parity of the cross-correlation checker (sdbe_xcorr_search, the GPU
form of corr_Xt_search / corr_Xt, software/verification/cross_corr.py:23-116,
reduced as verify_scan.py:231-247) against the oracle restatement
(tests/test_oracle_xcorr.py pins it).

* a delayed copy (2 windows + 37 samples, independent noise on each stream):
  every lag function within 2e-5 of the oracle's (values are normalised, peak
  about 0.9), peaks and standard deviations within 1e-4 relative, the fringe at
  window -2, lead +37, SNR within 1e-3 relative;
* QA use: the stage's 2-bit output of a block against the oracle's 2-bit
  codes of the same block, decoded to {-3, -1, 1, 3}: the fringe is at offset 0,
  lead 0, with a peak above 0.999;
* rejected: a search reaching outside the series, capacity exceeded.
"""
import sys

import numpy as np
import pytest

from oracle import sdbe_oracle as O
from sdbe_amd import synth as S

pytestmark = pytest.mark.gpu

LW = 32768


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def test_delayed_copy_matches_oracle(torch_cuda):
    torch = torch_cuda
    from sdbe_amd.qa import XCorr, fringe
    rng = np.random.default_rng(21)
    nw, lag_w, lead = 11, 2, 37
    common = rng.normal(size=(nw + 4) * LW)
    x0 = (common[4 * LW:4 * LW + nw * LW] + 0.5 * rng.normal(size=nw * LW)).astype(np.float32)
    x1 = (np.roll(common, -(lag_w * LW + lead))[4 * LW:4 * LW + nw * LW]
          + 0.5 * rng.normal(size=nw * LW)).astype(np.float32)
    first, n_off, avg = -3, 7, 4
    xc = XCorr(max_windows=16, max_offsets=8)
    _say("[xcorr] gpu search offsets %d..%d avg %d" % (first, first + n_off - 1, avg))
    s, peak, std = xc.search(torch.from_numpy(x0).cuda(), torch.from_numpy(x1).cuda(), first, n_off, avg)
    snr, pk, win, ld = fringe(s, peak, std, first)
    s, peak, std = s.cpu().numpy(), peak.cpu().numpy(), std.cpu().numpy()
    r = np.arange(first, first + n_off)
    so, So, po = O.corr_Xt_search(O.corr_spectra(x0.astype(np.float64)), O.corr_spectra(x1.astype(np.float64)),
                                  r, search_avg=avg)
    snr_o, sig_o, io = O.corr_snr(so, po)
    err = np.abs(s - so).max()
    _say("[xcorr] max |s - s_oracle| %.3e; peak %.5f (oracle %.5f); SNR %.2f (oracle %.2f); window %d lead %d"
         % (err, pk, po.max(), snr, snr_o, win, ld))
    assert err <= 2e-5
    assert np.allclose(peak, po, rtol=1e-4) and np.allclose(std, so.std(axis=1), rtol=1e-4)
    assert win == -lag_w and ld == lead and r[io] == win
    assert abs(snr / snr_o - 1) < 1e-3
    xc.close()


def test_stage_output_fringes_against_oracle_codes(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.qa import XCorr, fringe
    from sdbe_amd.stage import SdbeStage
    rng = S.rng_for(31)
    beng = S.codes_to_beng_int8(S.gaussian_codes(rng, 39)).reshape(1, 39, 16384, 2)
    st = SdbeStage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_PAYLOAD, chunk=2, max_spectra=1)
    _say("[xcorr] stage output of one block")
    out = st.process_spectra(torch.from_numpy(beng).cuda()).cpu().numpy()
    codes_gpu = O.unpack_codes(out.reshape(-1).view(np.uint32))
    x64 = O.resample_block(O.inflate(beng[0]))
    mean, sig = O.block_stats(x64)
    codes_ref = O.quantize2bit(x64, sig, mean)
    lv = np.array([-3.0, -1.0, 1.0, 3.0], dtype=np.float32)
    a, b = lv[codes_ref], lv[codes_gpu]          # 2,097,152 samples = 64 windows each
    xc = XCorr(max_windows=64, max_offsets=9)
    s, peak, std = xc.search(torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(), -4, 9, 16)
    snr, pk, win, ld = fringe(s, peak, std, -4)
    _say("[xcorr] stage vs oracle codes: peak %.6f, SNR %.1f, window %d, lead %d" % (pk, snr, win, ld))
    assert win == 0 and ld == 0 and pk > 0.999
    xc.close()
    st.close()


def test_xcorr_rejects(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.qa import XCorr
    x = torch.zeros(8 * LW, dtype=torch.float32, device="cuda")
    xc = XCorr(max_windows=8, max_offsets=4)
    with pytest.raises(L.SdbeError) as e:
        xc.search(x, x, -6, 3, 2)          # centre 3: window -3 lies before the series
    assert e.value.code == L.SDBE_ERR_ARG
    with pytest.raises(L.SdbeError) as e:
        xc.search(x, x, -1, 5, 2)          # more offsets than the capacity
    assert e.value.code == L.SDBE_ERR_CAPACITY
    xc.close()
