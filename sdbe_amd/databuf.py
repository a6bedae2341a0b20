"""ctypes mirror of include/aphids_databuf.h and a binding of the hashpipe
plugin library (libsdbe_hashpipe.so: vdif_inout_gpu_thread, the host-side
group bookkeeping, B-engine timestamp reform, 8pac unpack, and the in-process
hashpipe API used to run the plugin without hashpipe)."""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsdbe_hashpipe.so")

BENG_FRAMES_PER_GROUP = 39
VDIF_PER_BENG_FRAME = 2048
BENG_GROUPS_IN_BUFFER = 4
VDIF_OUT_BUFFER_SIZE = 4


class _W0(C.Structure):
    _fields_ = [("secs_inre", C.c_uint32, 30), ("legacy", C.c_uint32, 1), ("invalid", C.c_uint32, 1)]


class _W1(C.Structure):
    _fields_ = [("df_num_insec", C.c_uint32, 24), ("ref_epoch", C.c_uint32, 6), ("UA", C.c_uint32, 2)]


class _W2(C.Structure):
    _fields_ = [("df_len", C.c_uint32, 24), ("num_channels", C.c_uint32, 5), ("ver", C.c_uint32, 3)]


class _W3(C.Structure):
    _fields_ = [("stationID", C.c_uint32, 16), ("threadID", C.c_uint32, 10), ("bps", C.c_uint32, 5),
                ("dt", C.c_uint32, 1)]


class _Beng(C.Structure):
    _fields_ = [("b_upper", C.c_uint32), ("c", C.c_uint8), ("z", C.c_uint8), ("f", C.c_uint8),
                ("b_lower", C.c_uint8)]


class VdifInHeader(C.Structure):
    _fields_ = [("w0", _W0), ("w1", _W1), ("w2", _W2), ("w3", _W3), ("beng", _Beng), ("edh_psn", C.c_uint64)]


class VdifInPacket(C.Structure):
    _fields_ = [("header", VdifInHeader), ("data", C.c_char * 1024)]


class BengFrameCompletion(C.Structure):
    _fields_ = [("b", C.c_uint64), ("beng_frame_vdif_packet_count", C.c_uint16), ("cc", C.c_uint8 * 256)]


def group_completion_type(frames):
    """beng_group_completion_t of a build whose BENG_FRAMES_PER_GROUP is
    `frames` (DECIMATION_FACTOR of its SWARM rate preset)."""
    class _Bgc(C.Structure):
        _fields_ = [("beng_group_vdif_packet_count", C.c_int32), ("bgv_buf_cpu", C.c_void_p),
                    ("bgv_buf_gpu", C.c_void_p), ("gpu_id", C.c_int), ("memcpy_stream", C.c_void_p),
                    ("bfc", BengFrameCompletion * frames),
                    ("vdif_header_template", VdifInHeader)]
    _Bgc.__name__ = "BengGroupCompletion%d" % frames
    assert C.sizeof(_Bgc) == 72 + 272 * frames
    return _Bgc


BengGroupCompletion = group_completion_type(BENG_FRAMES_PER_GROUP)


class HashpipeDatabuf(C.Structure):
    _fields_ = [("data_type", C.c_char * 64), ("header_size", C.c_size_t), ("block_size", C.c_size_t),
                ("n_block", C.c_int), ("shmid", C.c_int), ("semid", C.c_int)]


class VdifInDatabuf(C.Structure):
    _fields_ = [("header", HashpipeDatabuf), ("bgc", BengGroupCompletion * BENG_GROUPS_IN_BUFFER)]


class QuantizedStorage(C.Structure):
    _fields_ = [("vdg_buf_cpu", C.c_void_p), ("vdg_buf_gpu", C.c_void_p), ("bit_depth", C.c_int),
                ("N_32bit_words_per_chan", C.c_int), ("gpu_id", C.c_int), ("memcpy_stream", C.c_void_p),
                ("vdif_header_template", C.c_uint32 * 8)]


class ThreadDesc(C.Structure):
    _fields_ = [("name", C.c_char_p), ("skey", C.c_char_p), ("init", C.c_void_p), ("run", C.c_void_p),
                ("ibuf_create", C.c_void_p), ("obuf_create", C.c_void_p)]


class BengTimestamp(C.Structure):
    _fields_ = [("sec", C.c_int32), ("clk", C.c_int32)]


assert C.sizeof(VdifInHeader) == 32 and C.sizeof(VdifInPacket) == 1056
assert C.sizeof(BengFrameCompletion) == 272 and C.sizeof(BengGroupCompletion) == 10680
assert C.sizeof(QuantizedStorage) == 72

# Plugin builds (sdbe_amd/build.py PLUGIN_VARIANTS): variant name ->
# (library, SWARM_RATE_FRAC, band-trimmed resampler, B-frames per group,
# output slots per polarisation VDIF_OUT_PKTS_PER_BLOCK)
VARIANTS = {
    None: (LIB_PATH, 6, False, 39, 8192),
    "trim6": (os.path.join(_HERE, "libsdbe_hashpipe_trim6.so"), 6, True, 39, 8192),
    "trim8": (os.path.join(_HERE, "libsdbe_hashpipe_trim8.so"), 8, True, 13, 2048),
    "trim10": (os.path.join(_HERE, "libsdbe_hashpipe_trim10.so"), 10, True, 65, 8192),
    "trim11": (os.path.join(_HERE, "libsdbe_hashpipe_trim11.so"), 11, True, 143, 16384),
}

_libs = {}


def load(variant=None):
    """The plugin library of a build variant (None: the default 6/11 build of
    the 39 -> 64 resampler)."""
    if variant in _libs:
        return _libs[variant]
    path = VARIANTS[variant][0]
    if not os.path.exists(path):
        raise RuntimeError("%s is not built (run `python -m sdbe_amd.build`)" % os.path.basename(path))
    lib = C.CDLL(path)
    P = C.POINTER
    sigs = {
        "find_hashpipe_thread": (P(ThreadDesc), [C.c_char_p]),
        "get_packet_b_count": (C.c_int64, [P(VdifInHeader)]),
        "init_beng_group": (None, [P(BengGroupCompletion), C.c_void_p, C.c_void_p, C.c_int64]),
        "get_beng_group_index_offset": (C.c_int, [P(VdifInDatabuf), C.c_int, P(VdifInPacket)]),
        "insert_vdif_in_beng_group_buffer": (C.c_int, [P(VdifInDatabuf), C.c_int, C.c_int, P(VdifInPacket)]),
        "check_beng_group_complete": (C.c_int, [P(VdifInDatabuf), C.c_int]),
        "fill_vdif_header_template": (None, [P(VdifInHeader), P(VdifInPacket), C.c_int]),
        "beng_reform_headers": (None, [C.c_void_p, C.c_int]),
        "beng_reform_reset": (None, []),
        "get_beng_t0": (None, [P(BengTimestamp)]),
        "unpack_8pac": (None, [C.c_void_p, C.c_void_p, C.c_int]),
        "vdif_in_databuf_create": (C.c_void_p, [C.c_int, C.c_int]),
        "vdif_out_databuf_create": (C.c_void_p, [C.c_int, C.c_int]),
        "get_bgv_cpu_memory_cuda": (C.c_int, [P(C.c_void_p), C.c_int]),
        "get_bgv_gpu_memory_cuda": (C.c_int, [P(C.c_void_p), C.c_int]),
        "transfer_beng_group_to_gpu_cuda": (C.c_int, [C.c_void_p, C.c_int]),
        "check_transfer_beng_group_to_gpu_complete_cuda": (C.c_int, [C.c_void_p, C.c_int]),
        "get_vdg_cpu_memory_cuda": (C.c_int, [P(C.c_void_p), C.c_int]),
        "transfer_vdif_group_to_cpu_cuda": (C.c_int, [C.c_void_p, C.c_int]),
        "check_transfer_vdif_group_to_cpu_complete_cuda": (C.c_int, [C.c_void_p, C.c_int]),
        "hpc_in_databuf": (C.c_void_p, []),
        "hpc_out_databuf": (C.c_void_p, []),
        "hpc_start": (C.c_int, [C.c_char_p]),
        "hpc_in_block": (C.c_void_p, [C.c_int]),
        "hpc_out_block": (C.c_void_p, [C.c_int]),
        "hpc_set_in_filled": (C.c_int, [C.c_int]),
        "hpc_in_status": (C.c_int, [C.c_int]),
        "hpc_set_out_free": (C.c_int, [C.c_int]),
        "hpc_wait_out_filled": (C.c_int, [C.c_int, C.c_int]),
        "hpc_stop": (C.c_int, []),
        "hpc_pipeline_start": (C.c_int, [C.POINTER(C.c_char_p), C.c_int]),
        "hpc_thread_done": (C.c_int, [C.c_int]),
        "hpc_status_puti4": (C.c_int, [C.c_char_p, C.c_int]),
        "hpc_status_puts": (C.c_int, [C.c_char_p, C.c_char_p]),
        "hpc_status_geti4": (C.c_int, [C.c_char_p, C.POINTER(C.c_int)]),
        "hpc_status_clear": (C.c_int, []),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _libs[variant] = lib
    return lib


def pipeline_start(lib, names):
    """Start the named plugins as a hashpipe pipeline (hpc_pipeline_start)."""
    arr = (C.c_char_p * len(names))(*[n.encode() for n in names])
    return lib.hpc_pipeline_start(arr, len(names))
