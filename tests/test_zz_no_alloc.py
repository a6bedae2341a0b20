"""GPU test of sdbe_process_spectra_noalloc, the allocation-free variant of
sdbe_process_spectra (include/sdbe_stage.h), called through the C-ABI.

Its output must equal sdbe_process_spectra's on a handle of the same
configuration and chunk (2 block-pols), byte for byte, headers included, and
it must advance the stream state the same way, on two calls in a row of which
the second is smaller (the scratch array keeps the first call's contents).
A REFERENCE-policy call latches from its own first block-pol, as a fresh
handle does.  A request above SDBE_NOALLOC_MAX_BLOCKPOLS is refused.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from sdbe_amd import synth as S

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _spectra(torch, seed, n_bp):
    codes = S.gaussian_codes(S.rng_for(seed), 39 * n_bp)
    return torch.from_numpy(S.codes_to_beng_int8(codes).reshape(n_bp, 39, 16384, 2)).cuda()


def _noalloc(torch, lib, L, cfg, state, d_beng, out):
    from sdbe_amd.stage import _ptr, _stream
    rc = lib.sdbe_process_spectra_noalloc(C.byref(cfg), C.byref(state), _ptr(d_beng), d_beng.shape[0], _ptr(out),
                                          _stream(None))
    L.check(rc, "sdbe_process_spectra_noalloc")
    torch.cuda.synchronize()


def test_noalloc_matches_process_spectra(torch_cuda):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage
    lib = L.load()
    ref = SdbeStage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_FRAMES, chunk=2, max_spectra=8, station="Sm",
                    bdc=1, rx=1)
    ref.set_state(secs=3000, df=124950, ref_epoch=30, psn=5)
    cfg = L.Config()
    C.memmove(C.byref(cfg), C.byref(ref.cfg), C.sizeof(L.Config))
    state = L.StreamState(secs=3000, df=124950, ref_epoch=30, reserved=0, psn=5)
    for call, (n_bp, seed) in enumerate(((5, 901), (3, 902))):
        d = _spectra(torch, seed, n_bp)
        _say("[noalloc] call %d: %d block-pols, handle path" % (call, n_bp))
        want = ref.process_spectra(d)
        got = torch.full_like(want, 0xA5)
        _say("[noalloc] call %d: allocation-free path" % call)
        _noalloc(torch, lib, L, cfg, state, d, got)
        nd = int((got != want).sum().item())
        _say("[noalloc] call %d: %d bytes differ" % (call, nd))
        assert nd == 0
        assert dict(secs=state.secs, df=state.df, ref_epoch=state.ref_epoch, psn=state.psn) == ref.get_state()
    ref.close()

    # REFERENCE policy, PAYLOAD output: one call latches from its first block-pol
    d = _spectra(torch, 903, 4)
    fresh = SdbeStage(policy=L.THRESH_REFERENCE, out_mode=L.OUT_PAYLOAD, chunk=2, max_spectra=8)
    _say("[noalloc] REFERENCE policy, handle path")
    want = fresh.process_spectra(d)
    C.memmove(C.byref(cfg), C.byref(fresh.cfg), C.sizeof(L.Config))
    state = L.StreamState(secs=0, df=0, ref_epoch=0, reserved=0, psn=1)
    got = torch.zeros_like(want)
    _say("[noalloc] REFERENCE policy, allocation-free path")
    _noalloc(torch, lib, L, cfg, state, d, got)
    assert torch.equal(got, want)
    fresh.close()

    # beyond the stated maximum: refused before any work
    big = torch.empty((1,), dtype=torch.int8, device="cuda")
    from sdbe_amd.stage import _ptr
    rc = lib.sdbe_process_spectra_noalloc(C.byref(cfg), C.byref(state), _ptr(big), L.NOALLOC_MAX_BLOCKPOLS + 1,
                                          _ptr(big), None)
    assert rc == L.SDBE_ERR_CAPACITY
    assert lib.sdbe_process_spectra_noalloc(C.byref(cfg), C.byref(state), _ptr(big), 0, _ptr(big), None) \
        == L.SDBE_ERR_ARG
