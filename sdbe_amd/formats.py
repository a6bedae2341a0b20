"""Python host side of the input-format variants (include/sdbe_formats.h).

``reorder_2015(x)`` runs the 2015-format B-engine reorder of
software/prototyping/cuda/kernel_template.py:107-164 (reorderTz_smem) on a
device tensor of complex64 spectra shaped [16384, n_frames, 128] and returns
a new device tensor [n_frames - 1, 128, 16385].  Device memory comes from
torch; the computation runs in libsdbe.so.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L


def reorder_2015(x, out=None, stream=None):
    lib = L.load()
    if x.dtype != torch.complex64 or not x.is_cuda or not x.is_contiguous():
        raise ValueError("x must be a contiguous complex64 device tensor")
    if x.dim() != 3 or x.shape[0] != L.REORDER2015_CHANNELS or x.shape[2] != L.REORDER2015_SNAPSHOTS:
        raise ValueError("x must be shaped [16384, n_frames, 128]")
    nf = x.shape[1]
    shape = (nf - 1, L.REORDER2015_SNAPSHOTS, L.REORDER2015_CHANNELS + 1)
    if out is None:
        # n_frames < 2 gives an empty output shape; the library rejects it
        out = torch.empty((max(nf - 1, 0),) + shape[1:], dtype=torch.complex64, device=x.device)
    elif tuple(out.shape) != shape or out.dtype != torch.complex64 or not out.is_contiguous():
        raise ValueError("out must be a contiguous complex64 tensor shaped %s" % (shape,))
    s = stream if stream is not None else torch.cuda.current_stream()
    L.check(lib.sdbe_reorder_2015(C.c_void_p(x.data_ptr()), nf, C.c_void_p(out.data_ptr()),
                                  C.c_void_p(s.cuda_stream)), "sdbe_reorder_2015")
    return out
