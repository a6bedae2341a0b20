"""Python host side of the polyphase filterbank path (include/sdbe_pfb.h).

``Pfb(window)`` owns one C-ABI plan; ``forward`` maps a float32 time series
of (n + 3) * 32768 samples to n spectra of 16385 bins, ``inverse`` maps n
spectra back to the minimum-norm least-squares time series
(software/prototyping/cuda/pfb_inverse.cu:172-488).  Device memory comes from
torch; every computation runs in libsdbe.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Pfb:
    def __init__(self, window, max_spectra=4096, device=0):
        self.lib = L.load()
        w = np.ascontiguousarray(np.asarray(window, dtype=np.float32))
        if w.size != L.PFB_NTAP * L.PFB_LBLOCK:
            raise ValueError("window must have %d coefficients" % (L.PFB_NTAP * L.PFB_LBLOCK))
        self.window = w
        self.device = device
        self.max_spectra = max_spectra
        h = C.c_void_p()
        L.check(self.lib.sdbe_pfb_create(device, L.PFB_NTAP, w.ctypes.data_as(C.c_void_p), max_spectra,
                                         C.byref(h)), "sdbe_pfb_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdbe_pfb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, where):
        if rc != L.SDBE_OK:
            raise L.SdbeError(rc, where, self.lib.sdbe_pfb_last_hip_error(self.h))

    def forward(self, x, stream=None):
        """x: float32 device tensor of (n + 3) * 32768 samples -> complex64 [n][16385]."""
        assert x.dtype == torch.float32 and x.is_contiguous() and x.numel() % L.PFB_LBLOCK == 0
        n = x.numel() // L.PFB_LBLOCK - (L.PFB_NTAP - 1)
        spec = torch.empty((n, L.PFB_NBINS), dtype=torch.complex64, device=x.device)
        self._check(self.lib.sdbe_pfb_forward(self.h, _ptr(x), n, _ptr(spec), _stream(stream)), "sdbe_pfb_forward")
        return spec

    def inverse(self, spec, stream=None):
        """spec: complex64 device tensor [n][16385] -> float32 (n + 3) * 32768 samples."""
        assert spec.dtype == torch.complex64 and spec.is_contiguous() and spec.shape[-1] == L.PFB_NBINS
        n = spec.shape[0]
        x = torch.empty(((n + L.PFB_NTAP - 1) * L.PFB_LBLOCK,), dtype=torch.float32, device=spec.device)
        self._check(self.lib.sdbe_pfb_inverse(self.h, _ptr(spec), n, _ptr(x), _stream(stream)), "sdbe_pfb_inverse")
        return x


def main():
    """Time the forward PFB and the inverse on n spectra (default 4096, the
    16384-channel config of 4096 spectra) of white noise with the Hann-windowed
    sinc prototype; prints one JSON line."""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--spectra", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args()
    n, lb = args.spectra, L.PFB_LBLOCK
    m = L.PFB_NTAP * lb
    k = np.arange(m, dtype=np.float64)
    window = (0.5 - 0.5 * np.cos(2 * np.pi * k / (m - 1))) * np.sinc((k - (m - 1) / 2) / lb)
    p = Pfb(window, max_spectra=n)
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn(((n + 3) * lb,), generator=g, device="cuda", dtype=torch.float32)
    spec = p.forward(x)
    xr = p.inverse(spec)
    torch.cuda.synchronize()

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        out = []
        for _ in range(args.reps):
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            out.append(e0.elapsed_time(e1))
        return float(np.median(out))

    t_fwd = timed(lambda: p.forward(x))
    t_inv = timed(lambda: p.inverse(spec))
    samples = (n + 3) * lb
    print(json.dumps(dict(kind="pfb", spectra=n, samples=samples, forward_ms=round(t_fwd, 3),
                          inverse_ms=round(t_inv, 3), forward_gsample_s=round(samples / t_fwd * 1e-6, 2),
                          inverse_gsample_s=round(samples / t_inv * 1e-6, 2))))
    del xr
    p.close()


if __name__ == "__main__":
    main()
