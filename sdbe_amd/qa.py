"""Python host side of the cross-correlation checker (include/sdbe_qa.h).

``XCorr.search(x0, x1, offsets, avg)`` runs corr_Xt_search
(software/verification/cross_corr.py:66-116) on two device float32 series cut
into 32768-sample windows and returns the lag functions, their peaks and
standard deviations (device tensors).  ``fringe(...)`` reduces them as
verify_scan.py:231-247 does: the offset of the highest peak, the peak over
the standard deviation there (SNR), and the sample lead of the peak within
the window.  Device memory comes from torch; the computation runs in
libsdbe.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L


class XCorr:
    def __init__(self, max_windows=512, max_offsets=257, device=0):
        self.lib = L.load()
        self.device = device
        h = C.c_void_p()
        L.check(self.lib.sdbe_xcorr_create(device, max_windows, max_offsets, C.byref(h)), "sdbe_xcorr_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.sdbe_xcorr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def search(self, x0, x1, first_offset, n_offsets, avg, stream=None):
        for x in (x0, x1):
            if x.dtype != torch.float32 or not x.is_cuda or not x.is_contiguous():
                raise ValueError("series must be contiguous float32 device tensors")
        n0, n1 = x0.numel() // L.XCORR_WINDOW, x1.numel() // L.XCORR_WINDOW
        dev = x0.device
        s = torch.empty((n_offsets, L.XCORR_WINDOW), dtype=torch.float32, device=dev)
        peak = torch.empty(n_offsets, dtype=torch.float32, device=dev)
        std = torch.empty(n_offsets, dtype=torch.float32, device=dev)
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = self.lib.sdbe_xcorr_search(self.h, C.c_void_p(x0.data_ptr()), n0, C.c_void_p(x1.data_ptr()), n1,
                                        first_offset, n_offsets, avg, C.c_void_p(s.data_ptr()),
                                        C.c_void_p(peak.data_ptr()), C.c_void_p(std.data_ptr()),
                                        C.c_void_p(st.cuda_stream))
        L.check(rc, "sdbe_xcorr_search", self.h)
        return s, peak, std


def fringe(s, peak, std, first_offset):
    """(snr, peak, window offset, sample lead) at the highest peak
    (verify_scan.py:232-247); host-side bookkeeping of the device results."""
    p = peak.cpu()
    i = int(torch.argmax(p))
    k = int(torch.argmax(s[i].abs().cpu()))
    lead = k - L.XCORR_WINDOW if k > L.XCORR_WINDOW // 2 else k
    return float(p[i] / std[i].cpu()), float(p[i]), first_offset + i, lead
