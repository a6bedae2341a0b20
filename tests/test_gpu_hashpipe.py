"""GPU test of the hashpipe drop-in: the vdif_inout_gpu_thread plugin run by
the in-process hashpipe API on two groups, checked against the C-ABI called
directly (same kernels, same REFERENCE threshold latch order) and for the
slot protocol of software/aphids/vdif_inout_gpu_thread.cu:823-1023."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import sdbe_oracle as O
from sdbe_amd import synth as S

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _d2h(ptr, nbytes):
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = C.c_int
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    out = np.empty(nbytes, dtype=np.uint8)
    assert hip.hipMemcpy(out.ctypes.data, ptr, nbytes, 2) == 0    # device to host
    return out


def test_plugin_two_groups_match_cabi():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from sdbe_amd import _lib as L
    from sdbe_amd import databuf as D
    from sdbe_amd.stage import SdbeStage

    _say("[plugin] generating 2 groups")
    groups, _ = S.make_group_packets(seed=61, n_groups=2, drop_frac=0.001)
    dpk = [SdbeStage.to_device_packets(pk) for pk, _ in groups]
    tmpl = O.fill_header_template(groups[0][0][0, :8], (100, 143_000_000))
    hp = D.load()
    _say("[plugin] hpc_start")
    assert hp.hpc_start(b"vdif_inout_gpu_thread") == 0
    try:
        for i, (pk, b0) in enumerate(groups):
            bgc = D.BengGroupCompletion.from_address(hp.hpc_in_block(i))
            bgc.beng_group_vdif_packet_count = pk.shape[0]
            bgc.bgv_buf_gpu = dpk[i].data_ptr()
            bgc.gpu_id = 0
            for f in range(39):
                bgc.bfc[f].b = b0 + f
            C.memmove(C.addressof(bgc.vdif_header_template), tmpl.ctypes.data, 32)
            if i:
                bgc.vdif_header_template.w0.invalid = 1
            assert hp.hpc_set_in_filled(i) == 0
        got = []
        for j in range(2):
            _say("[plugin] waiting for output slot %d" % j)
            assert hp.hpc_wait_out_filled(j, 120000) == 0
            qs = D.QuantizedStorage.from_address(hp.hpc_out_block(j))
            assert qs.bit_depth == 2 and qs.gpu_id == 0
            assert qs.N_32bit_words_per_chan == 8192 * 2048
            hdr = np.array(list(qs.vdif_header_template), dtype=np.uint32)
            if j == 0:
                assert np.array_equal(hdr, tmpl)
            else:
                assert (hdr[0] >> 31) == 1
            got.append(_d2h(qs.vdg_buf_gpu, 2 * 8192 * 8192))
            assert hp.hpc_in_status(j) == 0            # input slot released
        assert D.QuantizedStorage.from_address(hp.hpc_out_block(0)).vdg_buf_gpu != \
            D.QuantizedStorage.from_address(hp.hpc_out_block(1)).vdg_buf_gpu
    finally:
        _say("[plugin] hpc_stop")
        hp.hpc_stop()
    st = SdbeStage(policy=L.THRESH_REFERENCE, out_mode=L.OUT_PAYLOAD)
    for i, (pk, b0) in enumerate(groups):
        want = st.process_group(dpk[i], pk.shape[0], b0).cpu().numpy().reshape(-1)
        assert np.array_equal(got[i], want), i
    st.close()
