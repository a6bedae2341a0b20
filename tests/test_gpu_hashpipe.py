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


@pytest.mark.parametrize("variant,trim_mhz", [("trim6", 150), ("trim8", 102), ("trim10", 16), ("trim11", 150)])
def test_plugin_band_trimmed_build_matches_cabi(variant, trim_mhz):
    """A plugin built for the reference thread's band-trimmed path
    (-DSDBE_PLUGIN_TRIMMED=1 at SWARM_RATE_FRAC 6, 8, 10 or 11): groups of
    DECIMATION_FACTOR B-frames, the trim read from the status key TRIMDC (the
    reference's trim_from_dc, vdif_inout_gpu_thread.cu:770-784), and
    N_32bit_words_per_chan = 2 * 16384 * 128 * E / 16 (:1012).  Output equals
    the C-ABI's band-trimmed handle on the same groups, byte for byte."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from sdbe_amd import _lib as L
    from sdbe_amd import databuf as D
    from sdbe_amd.stage import SdbeStage

    _, rate, _, frames, slots = D.VARIANTS[variant]
    D_, E = L.TRIM_PRESETS[rate]
    assert D_ == frames
    _say("[plugin-%s] generating 2 groups of %d B-frames" % (variant, frames))
    groups, _ = S.make_group_packets(seed=65 + rate, n_groups=2, drop_frac=0.001, frames=frames)
    dpk = [SdbeStage.to_device_packets(pk) for pk, _ in groups]
    tmpl = O.fill_header_template(groups[0][0][0, :8], (100, 143_000_000))
    Bgc = D.group_completion_type(frames)
    hp = D.load(variant)
    hp.hpc_status_clear()
    assert hp.hpc_status_puti4(b"TRIMDC", trim_mhz) == 0
    assert hp.hpc_start(b"vdif_inout_gpu_thread") == 0
    words = 2 * 16384 * 128 * E // 16
    try:
        for i, (pk, b0) in enumerate(groups):
            bgc = Bgc.from_address(hp.hpc_in_block(i))
            bgc.beng_group_vdif_packet_count = pk.shape[0]
            bgc.bgv_buf_gpu = dpk[i].data_ptr()
            bgc.gpu_id = 0
            for f in range(frames):
                bgc.bfc[f].b = b0 + f
            C.memmove(C.addressof(bgc.vdif_header_template), tmpl.ctypes.data, 32)
            assert hp.hpc_set_in_filled(i) == 0
        got = []
        for j in range(2):
            _say("[plugin-%s] waiting for output slot %d" % (variant, j))
            assert hp.hpc_wait_out_filled(j, 120000) == 0
            qs = D.QuantizedStorage.from_address(hp.hpc_out_block(j))
            assert qs.bit_depth == 2 and qs.N_32bit_words_per_chan == words
            got.append(_d2h(qs.vdg_buf_gpu, 2 * slots * 8192).reshape(2, slots, 8192))
    finally:
        hp.hpc_stop()
    st = SdbeStage(policy=L.THRESH_REFERENCE, out_mode=L.OUT_PAYLOAD, rate_frac=rate, skip_mhz=trim_mhz)
    for i, (pk, b0) in enumerate(groups):
        want = st.process_group(dpk[i], pk.shape[0], b0).cpu().numpy()
        n = 128 * E
        assert want.shape == (2, n, 8192)
        assert np.array_equal(got[i][:, :n], want), i
        _say("[plugin-%s] group %d: %d frames per polarisation equal the C-ABI's" % (variant, i, n))
    st.close()


def test_host_to_host_chain_through_databuf_copies():
    """The input thread's side (host bookkeeping into pinned group buffers,
    HIP host-to-device copy), the plugin, and the output thread's side (HIP
    device-to-host copy into a pinned buffer): the host bytes equal the C-ABI
    output for the same packets.  Prints the PCIe copy rates."""
    import time

    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from sdbe_amd import _lib as L
    from sdbe_amd import databuf as D
    from sdbe_amd.stage import SdbeStage

    hip = C.CDLL("libamdhip64.so")
    hip.hipHostFree.argtypes = [C.c_void_p]
    hip.hipFree.argtypes = [C.c_void_p]
    _say("[chain] generating 2 groups")
    groups, _ = S.make_group_packets(seed=63, n_groups=2, drop_frac=0.001)
    tmpl = O.fill_header_template(groups[0][0][0, :8], (100, 143_000_000))
    hp = D.load()
    assert hp.hpc_start(b"vdif_inout_gpu_thread") == 0
    pinned, devbufs = [], []
    try:
        db = D.VdifInDatabuf.from_address(hp.hpc_in_databuf())
        b_first = groups[0][1]
        for i in range(2):
            cpu, gpu = C.c_void_p(), C.c_void_p()
            assert hp.get_bgv_cpu_memory_cuda(C.byref(cpu), i) == 1
            assert hp.get_bgv_gpu_memory_cuda(C.byref(gpu), i) == 1
            pinned.append(cpu.value)
            devbufs.append(gpu.value)
            hp.init_beng_group(C.byref(db.bgc[i]), cpu, gpu, b_first + 39 * i)
        _say("[chain] inserting packets through the group window")
        for i, (pk, b0) in enumerate(groups):
            assert b0 == b_first + 39 * i
            for row in pk:
                p = D.VdifInPacket.from_buffer(row)
                off = hp.get_beng_group_index_offset(C.byref(db), 0, C.byref(p))
                assert off == i
                hp.insert_vdif_in_beng_group_buffer(C.byref(db), 0, off, C.byref(p))
            assert db.bgc[i].beng_group_vdif_packet_count == pk.shape[0]
        h2d = []
        for i in range(2):
            C.memmove(C.addressof(db.bgc[i].vdif_header_template), tmpl.ctypes.data, 32)
            if i:
                db.bgc[i].vdif_header_template.w0.invalid = 1
            t0 = time.perf_counter()
            assert hp.transfer_beng_group_to_gpu_cuda(C.byref(db), i) == 1
            while True:
                rv = hp.check_transfer_beng_group_to_gpu_complete_cuda(C.byref(db), i)
                assert rv in (0, 1)
                if rv == 1:
                    break
            h2d.append((time.perf_counter() - t0, groups[i][0].nbytes))
            assert db.bgc[i].gpu_id == 0
            assert hp.hpc_set_in_filled(i) == 0
        got, d2h = [], []
        for j in range(2):
            _say("[chain] output slot %d" % j)
            assert hp.hpc_wait_out_filled(j, 120000) == 0
            qs = D.QuantizedStorage.from_address(hp.hpc_out_block(j))
            host = C.c_void_p()
            assert hp.get_vdg_cpu_memory_cuda(C.byref(host), j) == 1
            pinned.append(host.value)
            qs.vdg_buf_cpu = host.value
            t0 = time.perf_counter()
            assert hp.transfer_vdif_group_to_cpu_cuda(C.c_void_p(hp.hpc_out_databuf()), j) == 1
            assert hp.check_transfer_vdif_group_to_cpu_complete_cuda(C.c_void_p(hp.hpc_out_databuf()), j) == 1
            d2h.append(time.perf_counter() - t0)
            got.append(np.ctypeslib.as_array((C.c_uint8 * (2 * 8192 * 8192)).from_address(host.value)).copy())
        for i, (dt, nb) in enumerate(h2d):
            _say("[chain] group %d: H2D %.1f MB in %.2f ms (%.1f GB/s), D2H 134.2 MB in %.2f ms (%.1f GB/s)"
                 % (i, nb / 1e6, dt * 1e3, nb / dt / 1e9, d2h[i] * 1e3, 2 * 8192 * 8192 / d2h[i] / 1e9))
    finally:
        hp.hpc_stop()
    st = SdbeStage(policy=L.THRESH_REFERENCE, out_mode=L.OUT_PAYLOAD)
    for i, (pk, b0) in enumerate(groups):
        want = st.process_group(st.to_device_packets(pk), pk.shape[0], b0).cpu().numpy().reshape(-1)
        assert np.array_equal(got[i], want), i
    st.close()
    for ptr in pinned:
        hip.hipHostFree(ptr)
    for ptr in devbufs:
        hip.hipFree(ptr)
