"""Python host side of the stage: a thin owner of one C-ABI handle.

Device memory comes from torch (plumbing only); every computation runs in
libsdbe.so.  ``SdbeStage`` mirrors the per-slot work of the reference's
``vdif_inout_gpu_thread`` run loop (software/aphids/vdif_inout_gpu_thread.cu:
874-1023): one call per group of 39 B-frames, output either as payloads
(vdif_out_data_group_t layout) or as complete VDIF frames.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def station_code(s):
    return ((ord(s[0]) & 0xFF) << 8) | (ord(s[1]) & 0xFF)


class SdbeStage:
    def __init__(self, policy=L.THRESH_REFERENCE, out_mode=L.OUT_PAYLOAD, chunk=256, device=0,
                 station="XX", bdc=0, rx=0, max_spectra=256, fixed=(1.0, 0.0), use_graph=False,
                 c1_tiles_per_wg=0, max_batch=0, fft_precision=L.FFT_FP32, rate_frac=L.RATE_64_39,
                 skip_mhz=0, c1_interleave=0, packet_geometry=L.GEOM_REFERENCE):
        self.lib = L.load()
        cfg = L.Config()
        L.check(self.lib.sdbe_default_config(C.byref(cfg)), "sdbe_default_config")
        cfg.thresh_policy = policy
        cfg.out_mode = out_mode
        cfg.chunk_blockpols = chunk
        cfg.max_spectra_blockpols = max_spectra
        cfg.station_id = station_code(station)
        cfg.bdc_sideband = bdc
        cfg.rx_sideband = rx
        cfg.fixed_threshold = fixed[0]
        cfg.fixed_offset = fixed[1]
        cfg.use_graph = 1 if use_graph else 0
        cfg.c1_tiles_per_wg = c1_tiles_per_wg
        cfg.max_batch_blockpols = max_batch
        cfg.fft_precision = fft_precision
        cfg.rate_frac = rate_frac
        cfg.skip_mhz = skip_mhz
        cfg.c1_interleave = c1_interleave
        cfg.packet_geometry = packet_geometry
        # bytes per packet of the configured geometry
        self.pkt_bytes = 4 * L.PKT_WORDS_8X1024 if packet_geometry == L.GEOM_8X1024 else L.PKT_BYTES
        self.cfg = cfg
        # geometry of a block-pol: snapshots in, samples and frames out
        # (include/sdbe_stage.h, SDBE_RATE_*)
        # a group is 128 blocks: 39 B-frames at 64/39, D B-frames band-trimmed
        if rate_frac == L.RATE_64_39:
            self.snaps, self.samples, self.frames = 39, L.SAMPLES_PER_BLOCK, L.OUT_FRAMES_PER_BLOCK
            self.group_frames = L.FRAMES_PER_GROUP
        else:
            d, e = L.TRIM_PRESETS.get(rate_frac, (0, 0))
            self.snaps, self.samples, self.frames = d, 128 * 256 * e, e
            self.group_frames = d
        self.frames_per_group = L.BLOCKS_PER_GROUP * self.frames
        self.device = device
        self.out_mode = out_mode
        self.frame_bytes = L.OUT_PAYLOAD_BYTES + (L.OUT_HEADER_BYTES if out_mode == L.OUT_FRAMES else 0)
        h = C.c_void_p()
        L.check(self.lib.sdbe_create(C.byref(cfg), device, C.byref(h)), "sdbe_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdbe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- state
    def set_state(self, secs=0, df=0, ref_epoch=0, psn=1):
        st = L.StreamState(secs=secs, df=df, ref_epoch=ref_epoch, reserved=0, psn=psn)
        L.check(self.lib.sdbe_set_stream_state(self.h, C.byref(st)), "sdbe_set_stream_state", self.h)

    def get_state(self):
        st = L.StreamState()
        L.check(self.lib.sdbe_get_stream_state(self.h, C.byref(st)), "sdbe_get_stream_state", self.h)
        return dict(secs=st.secs, df=st.df, ref_epoch=st.ref_epoch, psn=st.psn)

    def stats(self):
        s = L.Stats()
        L.check(self.lib.sdbe_get_stats(self.h, C.byref(s)), "sdbe_get_stats", self.h)
        return {k: getattr(s, k) for k, _ in L.Stats._fields_ if k != "reserved"}

    def set_output_alignment(self, enable=True, skip_int32=0, skip_frames=0):
        L.check(self.lib.sdbe_set_output_alignment(self.h, 1 if enable else 0, skip_int32, skip_frames),
                "sdbe_set_output_alignment", self.h)

    def reset_latch(self, slot, stream=None):
        L.check(self.lib.sdbe_reset_latch(self.h, slot, _stream(stream)), "sdbe_reset_latch", self.h)

    def profile(self, enable=True):
        L.check(self.lib.sdbe_profile_enable(self.h, 1 if enable else 0), "sdbe_profile_enable", self.h)

    def profile_select(self, kernel=None):
        """Time only the launches of one kernel kind (None: every kernel)."""
        L.check(self.lib.sdbe_profile_select(self.h, kernel.encode() if kernel else None),
                "sdbe_profile_select", self.h)

    def profile_read(self):
        """{kernel name: (launches, total ms, units)} since the last read."""
        buf = (L.KernelTime * 32)()
        n = C.c_int32()
        L.check(self.lib.sdbe_profile_read(self.h, buf, 32, C.byref(n)), "sdbe_profile_read", self.h)
        return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms, buf[i].units)
                for i in range(min(n.value, 32))}

    # -------------------------------------------------------------- buffers
    def group_out(self):
        """Output buffer of one group, [2 pols][frames per group][frame bytes] uint8
        (128 blocks: 8192 frames for the 64/39 resampler, 128 E band-trimmed:
        4096 at 6/11, 1024 at 8/11, 8192 at 10/11, 16384 at 11/11)."""
        return torch.empty((2, self.frames_per_group, self.frame_bytes), dtype=torch.uint8,
                           device="cuda:%d" % self.device)

    @staticmethod
    def to_device_packets(pkts_u32, device=0):
        t = torch.from_numpy(np.ascontiguousarray(pkts_u32).view(np.uint8).reshape(-1))
        return t.to("cuda:%d" % device)

    # ----------------------------------------------------------------- calls
    def process_group(self, d_pkts, n_pkts, b_zero, out=None, bitmap=None, stream=None):
        if out is None:
            out = self.group_out()
        assert out.is_contiguous() and out.numel() >= 2 * self.frames_per_group * self.frame_bytes
        assert d_pkts.numel() >= n_pkts * self.pkt_bytes
        stride = out.numel() // 2
        rc = self.lib.sdbe_process_group(self.h, _ptr(d_pkts), n_pkts, b_zero, _ptr(out), stride,
                                         _ptr(bitmap), _stream(stream))
        L.check(rc, "sdbe_process_group", self.h)
        return out

    def unpack_int8(self, d_pkts, n_pkts, b_zero, stream=None):
        dev = "cuda:%d" % self.device
        b0 = torch.empty((self.group_frames, 128, L.CHANNELS, 2), dtype=torch.int8, device=dev)
        b1 = torch.empty_like(b0)
        # fid bitmap: uint8 [frames][256] (reference geometry) or uint16
        # [frames][16] (8 x 1024 geometry)
        if self.cfg.packet_geometry == L.GEOM_8X1024:
            bm = torch.empty((self.group_frames, 16), dtype=torch.int16, device=dev)
        else:
            bm = torch.empty((self.group_frames, 256), dtype=torch.uint8, device=dev)
        assert d_pkts.numel() >= n_pkts * self.pkt_bytes
        rc = self.lib.sdbe_unpack_group_int8(self.h, _ptr(d_pkts), n_pkts, b_zero, _ptr(b0), _ptr(b1),
                                             _ptr(bm), _stream(stream))
        L.check(rc, "sdbe_unpack_group_int8", self.h)
        return b0, b1, bm

    def process_spectra(self, d_beng, out=None, stream=None):
        """d_beng: int8 [n_bp][snapshots per block][16384][2 = (im, re)] on the device."""
        n_bp = d_beng.shape[0]
        assert d_beng.dtype == torch.int8 and d_beng.is_contiguous()
        assert tuple(d_beng.shape[1:]) == (self.snaps, L.CHANNELS, 2)
        if out is None:
            out = torch.empty((n_bp, self.frames, self.frame_bytes), dtype=torch.uint8, device=d_beng.device)
        assert out.numel() >= n_bp * self.frames * self.frame_bytes
        rc = self.lib.sdbe_process_spectra(self.h, _ptr(d_beng), n_bp, _ptr(out), _stream(stream))
        L.check(rc, "sdbe_process_spectra", self.h)
        return out

    def resample_f32(self, d_beng, stream=None):
        n_bp = d_beng.shape[0]
        assert d_beng.dtype == torch.int8 and d_beng.is_contiguous()
        assert tuple(d_beng.shape[1:]) == (self.snaps, L.CHANNELS, 2)
        x = torch.empty((n_bp, self.samples), dtype=torch.float32, device=d_beng.device)
        st = torch.empty((n_bp, 2), dtype=torch.float64, device=d_beng.device)
        rc = self.lib.sdbe_resample_spectra_f32(self.h, _ptr(d_beng), n_bp, _ptr(x), _ptr(st),
                                                _stream(stream))
        L.check(rc, "sdbe_resample_spectra_f32", self.h)
        return x, st


class BatchStream:
    """One stream of a ragged batch (sdbe_process_blocks): int8 spectra of both
    polarisations (device tensors [n_blocks * 39][16384][2 = (im, re)]), the
    lower-sideband flip, threshold policy, REFERENCE latch slot, station
    fields and VDIF time state (advanced by every call).  `out` receives
    [2 pols][64 n_blocks][frame bytes]."""

    def __init__(self, beng0, beng1, out, flip=False, policy=L.THRESH_PER_BLOCK, fixed=(1.0, 0.0),
                 latch_slot=0, station="XX", bdc=0, rx=0, secs=0, df=0, ref_epoch=0, psn=1):
        assert beng0.shape == beng1.shape and beng0.dtype == torch.int8
        assert beng0.is_contiguous() and beng1.is_contiguous() and out.is_contiguous()
        assert beng0.shape[0] % 39 == 0 and tuple(beng0.shape[1:]) == (L.CHANNELS, 2)
        self.beng = (beng0, beng1)
        self.out = out
        self.n_blocks = beng0.shape[0] // 39
        self.state = L.StreamState(secs=secs, df=df, ref_epoch=ref_epoch, reserved=0, psn=psn)
        self.desc = L.StreamDesc()
        d = self.desc
        d.d_beng[0] = _ptr(beng0)
        d.d_beng[1] = _ptr(beng1)
        d.n_blocks = self.n_blocks
        d.flip = 1 if flip else 0
        d.d_out = _ptr(out)
        d.out_pol_stride_bytes = out.numel() // 2
        d.state = C.pointer(self.state)
        d.thresh_policy = policy
        d.fixed_threshold = fixed[0]
        d.fixed_offset = fixed[1]
        d.latch_slot = latch_slot
        d.station_id = station_code(station)
        d.bdc_sideband = bdc
        d.rx_sideband = rx

    def get_state(self):
        st = self.state
        return dict(secs=st.secs, df=st.df, ref_epoch=st.ref_epoch, psn=st.psn)


def process_blocks(stage, streams, stream=None):
    """Run a ragged batch of BatchStream objects through one sdbe_process_blocks call."""
    arr = (L.StreamDesc * len(streams))(*[s.desc for s in streams])
    rc = stage.lib.sdbe_process_blocks(stage.h, arr, len(streams), _stream(stream))
    L.check(rc, "sdbe_process_blocks", stage.h)   # each desc.state points at its stream's state


def state_from_template(template_words, n_words_per_chan=L.OUT_FRAMES_PER_GROUP * 2048):
    lib = L.load()
    w = (C.c_uint32 * 8)(*[int(x) for x in template_words[:8]])
    st = L.StreamState()
    sk = C.c_int32()
    sf = C.c_int32()
    L.check(lib.sdbe_stream_state_from_template(C.cast(w, C.c_void_p), n_words_per_chan, C.byref(st),
                                                C.byref(sk), C.byref(sf)), "sdbe_stream_state_from_template")
    return dict(secs=st.secs, df=st.df, ref_epoch=st.ref_epoch, psn=st.psn, skip_int32=sk.value,
                skip_frames=sf.value)
