"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
include/sdbe_stage.h declares, and the ctypes structures match the C layout.
No compute calls (no GPU here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDRS = [os.path.join(ROOT, "include", h) for h in ("sdbe_stage.h", "sdbe_pfb.h", "sdbe_formats.h", "sdbe_qa.h")]


def _declared_functions():
    names = set()
    for h in HDRS:
        src = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(sdbe_[a-z0-9_]+)\s*\(", src))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from sdbe_amd import _lib
    lib = _lib.load()
    declared = _declared_functions()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    bound = {p[0] for p in _lib.PROTOTYPES}
    assert set(declared) == bound


def test_struct_layouts_match_c(tmp_path):
    """Compile a probe against the header with gcc and compare sizeof/offsetof
    with the ctypes mirrors."""
    from sdbe_amd import _lib as L
    probe = tmp_path / "probe.c"
    probe.write_text("""
#include <stdio.h>
#include <stddef.h>
#include "sdbe_stage.h"
#define P(T, F) printf(#T "." #F " %zu\\n", offsetof(T, F));
int main(void) {
  printf("sdbe_config_t %zu\\n", sizeof(sdbe_config_t));
  printf("sdbe_stream_state_t %zu\\n", sizeof(sdbe_stream_state_t));
  printf("sdbe_stats_t %zu\\n", sizeof(sdbe_stats_t));
  printf("sdbe_kernel_time_t %zu\\n", sizeof(sdbe_kernel_time_t));
  printf("sdbe_stream_desc_t %zu\\n", sizeof(sdbe_stream_desc_t));
  P(sdbe_config_t, use_graph) P(sdbe_config_t, max_batch_blockpols)
  P(sdbe_stream_desc_t, n_blocks) P(sdbe_stream_desc_t, d_out) P(sdbe_stream_desc_t, state)
  P(sdbe_stream_desc_t, thresh_policy) P(sdbe_stream_desc_t, latch_slot) P(sdbe_stream_desc_t, station_id)
  P(sdbe_stream_desc_t, rx_sideband) P(sdbe_stream_desc_t, reserved)
  P(sdbe_config_t, station_id) P(sdbe_config_t, rx_sideband) P(sdbe_config_t, c1_interleave)
  P(sdbe_config_t, packet_geometry) P(sdbe_config_t, skip_mhz)
  P(sdbe_stream_state_t, psn) P(sdbe_stats_t, latched_threshold) P(sdbe_kernel_time_t, total_ms)
  return 0;
}
""")
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(probe), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.split("\n") if l)
    assert int(out["sdbe_config_t"]) == C.sizeof(L.Config)
    assert int(out["sdbe_stream_state_t"]) == C.sizeof(L.StreamState)
    assert int(out["sdbe_stats_t"]) == C.sizeof(L.Stats)
    assert int(out["sdbe_kernel_time_t"]) == C.sizeof(L.KernelTime)
    assert int(out["sdbe_config_t.station_id"]) == L.Config.station_id.offset
    assert int(out["sdbe_config_t.rx_sideband"]) == L.Config.rx_sideband.offset
    for f in ("c1_interleave", "packet_geometry", "skip_mhz"):
        assert int(out["sdbe_config_t." + f]) == getattr(L.Config, f).offset, f
    assert int(out["sdbe_stream_state_t.psn"]) == L.StreamState.psn.offset
    assert int(out["sdbe_stats_t.latched_threshold"]) == L.Stats.latched_threshold.offset
    assert int(out["sdbe_kernel_time_t.total_ms"]) == L.KernelTime.total_ms.offset
    assert int(out["sdbe_stream_desc_t"]) == C.sizeof(L.StreamDesc)
    assert int(out["sdbe_config_t.use_graph"]) == L.Config.use_graph.offset
    assert int(out["sdbe_config_t.max_batch_blockpols"]) == L.Config.max_batch_blockpols.offset
    for f in ("n_blocks", "d_out", "state", "thresh_policy", "latch_slot", "station_id", "rx_sideband",
              "reserved"):
        assert int(out["sdbe_stream_desc_t." + f]) == getattr(L.StreamDesc, f).offset, f


def test_host_only_entry_points_without_gpu():
    """Calls that touch no device: config defaults, error strings, template
    state derivation, argument validation."""
    from sdbe_amd import _lib as L
    lib = L.load()
    cfg = L.Config()
    assert lib.sdbe_default_config(C.byref(cfg)) == 0
    assert cfg.thresh_policy == L.THRESH_REFERENCE and cfg.station_id == 0x5858
    assert lib.sdbe_default_config(None) == L.SDBE_ERR_ARG
    assert lib.sdbe_error_string(L.SDBE_ERR_CAPACITY).decode() == "request exceeds handle capacity"
    assert lib.sdbe_process_group(None, None, 0, 0, None, 0, None, None) == L.SDBE_ERR_ARG
    assert lib.sdbe_get_stats(None, None) == L.SDBE_ERR_ARG
