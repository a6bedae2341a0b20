"""ctypes binding of include/sdbe_stage.h (libsdbe.so, built in-tree).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdbe.so")

SDBE_OK = 0
SDBE_ERR_ARG = -1
SDBE_ERR_HIP = -2
SDBE_ERR_NOMEM = -3
SDBE_ERR_CAPACITY = -4
SDBE_ERR_NODEV = -5
SDBE_ERR_STATE = -6
NOALLOC_MAX_BLOCKPOLS = 64       # SDBE_NOALLOC_MAX_BLOCKPOLS
NOALLOC_SCRATCH_BYTES = 41943040  # SDBE_NOALLOC_SCRATCH_BYTES

THRESH_REFERENCE = 0
THRESH_PER_BLOCK = 1
THRESH_FIXED = 2
RATE_64_39 = 0
RATE_6_11 = 6
RATE_8_11 = 8
RATE_10_11 = 10
RATE_11_11 = 11
# SWARM rate preset -> (DECIMATION_FACTOR D, EXPANSION_FACTOR E) of the
# band-trimmed path (include/sdbe_stage.h); a group is D B-frames, 128 blocks
# of D snapshots, E frames of 8192 payload bytes per block-pol
TRIM_PRESETS = {RATE_6_11: (39, 32), RATE_8_11: (13, 8), RATE_10_11: (65, 64), RATE_11_11: (143, 128)}
GEOM_REFERENCE = 0              # 1056-byte packets, 128 snapshots x 8 channels
GEOM_8X1024 = 1                 # 8224-byte packets, 8 snapshots x 1024 channels
PKT_WORDS_8X1024 = 2056
FFT_FP32 = 0
FFT_BF16X3 = 1
FFT_BF16 = 2
FFT_FP16 = 3
OUT_PAYLOAD = 0
OUT_FRAMES = 1

CHANNELS = 16384
FRAMES_PER_GROUP = 39
BLOCKS_PER_GROUP = 128
SNAPSHOTS_PER_GROUP = FRAMES_PER_GROUP * 128     # 4992
SAMPLES_PER_BLOCK = 2097152
OUT_FRAMES_PER_BLOCK = 64
OUT_PAYLOAD_BYTES = 8192
OUT_HEADER_BYTES = 32
OUT_FRAMES_PER_GROUP = BLOCKS_PER_GROUP * OUT_FRAMES_PER_BLOCK
PKT_BYTES = 1056


class Config(C.Structure):
    _fields_ = [
        ("thresh_policy", C.c_int32),
        ("fixed_threshold", C.c_float),
        ("fixed_offset", C.c_float),
        ("out_mode", C.c_int32),
        ("chunk_blockpols", C.c_int32),
        ("max_spectra_blockpols", C.c_int32),
        ("station_id", C.c_uint16),
        ("bdc_sideband", C.c_uint8),
        ("rx_sideband", C.c_uint8),
        ("use_graph", C.c_int32),
        ("c1_tiles_per_wg", C.c_int32),
        ("max_batch_blockpols", C.c_int32),
        ("fft_precision", C.c_int32),
        ("rate_frac", C.c_int32),
        ("skip_mhz", C.c_int32),
        ("c1_interleave", C.c_int32),
        ("packet_geometry", C.c_int32),
    ]


class StreamState(C.Structure):
    _fields_ = [
        ("secs", C.c_uint32),
        ("df", C.c_uint32),
        ("ref_epoch", C.c_uint32),
        ("reserved", C.c_uint32),
        ("psn", C.c_uint64),
    ]


LATCH_SLOTS = 64


class StreamDesc(C.Structure):
    _fields_ = [
        ("d_beng", C.c_void_p * 2),
        ("n_blocks", C.c_int32),
        ("flip", C.c_int32),
        ("d_out", C.c_void_p),
        ("out_pol_stride_bytes", C.c_int64),
        ("state", C.POINTER(StreamState)),
        ("thresh_policy", C.c_int32),
        ("fixed_threshold", C.c_float),
        ("fixed_offset", C.c_float),
        ("latch_slot", C.c_int32),
        ("station_id", C.c_uint16),
        ("bdc_sideband", C.c_uint8),
        ("rx_sideband", C.c_uint8),
        ("reserved", C.c_int32 * 3),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("n_packets", C.c_int32),
        ("n_invalid", C.c_int32),
        ("n_out_of_range", C.c_int32),
        ("n_duplicate", C.c_int32),
        ("n_unique", C.c_int32),
        ("n_missing", C.c_int32),
        ("n_blockpols", C.c_int32),
        ("n_frames_per_pol", C.c_int32),
        ("latched_threshold", C.c_double),
        ("latched_offset", C.c_double),
    ]


class KernelTime(C.Structure):
    _fields_ = [
        ("name", C.c_char * 32),
        ("launches", C.c_int32),
        ("reserved", C.c_int32),
        ("total_ms", C.c_double),
        ("units", C.c_double),
    ]


PFB_LBLOCK = 32768
PFB_NBINS = 16385
PFB_NTAP = 4
REORDER2015_CHANNELS = 16384
REORDER2015_SNAPSHOTS = 128
REORDER2015_SPLIT = 69
XCORR_WINDOW = 32768

# (name, restype, argtypes) of every symbol declared in include/sdbe_stage.h,
# include/sdbe_pfb.h, include/sdbe_formats.h and include/sdbe_qa.h
PROTOTYPES = [
    ("sdbe_default_config", C.c_int, [C.POINTER(Config)]),
    ("sdbe_create", C.c_int, [C.POINTER(Config), C.c_int, C.POINTER(C.c_void_p)]),
    ("sdbe_destroy", None, [C.c_void_p]),
    ("sdbe_error_string", C.c_char_p, [C.c_int]),
    ("sdbe_last_hip_error", C.c_int, [C.c_void_p]),
    ("sdbe_set_stream_state", C.c_int, [C.c_void_p, C.POINTER(StreamState)]),
    ("sdbe_get_stream_state", C.c_int, [C.c_void_p, C.POINTER(StreamState)]),
    ("sdbe_stream_state_from_template", C.c_int,
     [C.c_void_p, C.c_int32, C.POINTER(StreamState), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("sdbe_process_group", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("sdbe_unpack_group_int8", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("sdbe_process_spectra", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("sdbe_resample_spectra_f32", C.c_int,
     [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("sdbe_process_blocks", C.c_int, [C.c_void_p, C.POINTER(StreamDesc), C.c_int32, C.c_void_p]),
    ("sdbe_reset_latch", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    ("sdbe_set_output_alignment", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    ("sdbe_get_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("sdbe_process_spectra_noalloc", C.c_int,
     [C.POINTER(Config), C.POINTER(StreamState), C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    # include/sdbe_pfb.h
    ("sdbe_pfb_create", C.c_int, [C.c_int, C.c_int32, C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    ("sdbe_pfb_destroy", None, [C.c_void_p]),
    ("sdbe_pfb_last_hip_error", C.c_int, [C.c_void_p]),
    ("sdbe_pfb_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("sdbe_pfb_inverse", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("sdbe_profile_enable", C.c_int, [C.c_void_p, C.c_int]),
    ("sdbe_profile_read", C.c_int, [C.c_void_p, C.POINTER(KernelTime), C.c_int32, C.POINTER(C.c_int32)]),
    ("sdbe_profile_select", C.c_int, [C.c_void_p, C.c_char_p]),
    # include/sdbe_formats.h
    ("sdbe_reorder_2015", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    # include/sdbe_qa.h
    ("sdbe_xcorr_create", C.c_int, [C.c_int, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]),
    ("sdbe_xcorr_destroy", None, [C.c_void_p]),
    ("sdbe_xcorr_last_hip_error", C.c_int, [C.c_void_p]),
    ("sdbe_xcorr_search", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
]

_lib = None


def load():
    """Load libsdbe.so and bind every prototype; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libsdbe.so is not built (run `python -m sdbe_amd.build`); "
                           "the stage has no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in PROTOTYPES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SdbeError(RuntimeError):
    def __init__(self, code, where, hip=0):
        lib = load()
        msg = lib.sdbe_error_string(code).decode()
        super().__init__("%s failed: %s (code %d, hip error %d)" % (where, msg, code, hip))
        self.code = code
        self.hip = hip


def check(code, where, handle=None):
    if code != SDBE_OK:
        hip = load().sdbe_last_hip_error(handle) if handle else 0
        raise SdbeError(code, where, hip)
    return code
