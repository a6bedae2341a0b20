"""GPU test of the whole hashpipe chain: vdif_in_file_thread (the input
thread's grouping loop, software/aphids/vdif_in_net_thread.c:269-500, fed from
a file of 8pac records) -> vdif_inout_gpu_thread -> vdif_out_file_thread (the
output thread's time alignment and stamping, vdif_out_net_thread.c:180-336,
writing one VDIF file per polarisation), run by the in-process hashpipe API.

The stream starts one B-frame before the first group, as the input thread
expects (its groups start at b_first + 1), and carries 0.1% dropped packets,
shuffled within each B-frame, in the wire format: new-format 53-bit
timestamps packed 8 packets per 8pac record (unpacked and reformed by the
thread).  With VINFLUSH the groups still filling at the end of the file are
handed off too.  Bars:
* the input thread's counters: every packet of the first B-frame skipped, all
  others used, one group per DECIMATION_FACTOR B-frames handed off;
* the written files equal, byte for byte, the output thread's re-framing
  (oracle.OutputAligner) of the C-ABI's output for the same groups, with the
  header template of the input thread (oracle.fill_header_template from the
  stream's first timestamp).
Prints the ingest-inclusive rate of the chain (file read, unpack, reform,
sort, grouping, H2D, stage, D2H, alignment, file write).
"""
import ctypes as C
import os
import sys
import time

import numpy as np
import pytest

from oracle import sdbe_oracle as O
from sdbe_amd import synth as S

pytestmark = pytest.mark.gpu

T_REF = (1000, 285_990_000)        # first B-frame's timestamp: frames cross a second boundary
B_FIRST = 1000


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture
def short_dir():
    """A directory with a short path (status-buffer strings hold 67 characters)."""
    import shutil
    import tempfile
    d = tempfile.mkdtemp(prefix="sdbe", dir="/tmp")
    yield d
    shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("variant,n_groups", [(None, 3), ("trim8", 5)])
def test_file_pipeline_matches_cabi_and_output_thread(variant, n_groups, short_dir):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from sdbe_amd import _lib as L
    from sdbe_amd import databuf as D
    from sdbe_amd.stage import SdbeStage

    _, rate, trimmed, frames, _ = D.VARIANTS[variant]
    _say("[pipe-%s] generating %d groups of %d B-frames" % (variant, n_groups, frames))
    pk, _ = S.make_stream_packets(71 + rate, n_groups, frames=frames)
    rec = S.to_8pac_new_format(pk, B_FIRST, T_REF)
    fin = os.path.join(short_dir, "in.8pac")
    rec.tofile(fin)
    fout = [os.path.join(short_dir, "out.pol%d.vdif" % p) for p in (0, 1)]
    hp = D.load(variant)
    hp.hpc_status_clear()
    for k, v in (("VINFILE", fin), ("VOUTPOL0", fout[0]), ("VOUTPOL1", fout[1]), ("STATION", "Sm")):
        assert hp.hpc_status_puts(k.encode(), v.encode()) == 0
    for k, v in (("VIN8PAC", 1), ("VINBATCH", 256), ("VINFLUSH", 1), ("BDCSB", 1), ("RXSB", 0), ("TRIMDC", 150)):
        assert hp.hpc_status_puti4(k.encode(), v) == 0
    hp.beng_reform_reset()
    _say("[pipe-%s] starting input -> GPU -> output threads on %d records" % (variant, rec.shape[0]))
    t0 = time.perf_counter()
    assert D.pipeline_start(hp, ["vdif_in_file_thread", "vdif_inout_gpu_thread", "vdif_out_file_thread"]) == 0
    got = C.c_int(0)
    try:
        deadline = time.time() + 240
        last = -1
        while time.time() < deadline:
            if hp.hpc_status_geti4(b"VOUTGRPS", C.byref(got)) and got.value >= n_groups:
                break
            if got.value != last:
                _say("[pipe-%s] groups written: %d" % (variant, max(got.value, 0)))
                last = got.value
            time.sleep(0.2)
        dt = time.perf_counter() - t0
    finally:
        hp.hpc_stop()
    assert got.value == n_groups
    cnt = {}
    for k in ("VINPKTS", "VINUSED", "VINSKIP", "VININVAL", "VINAHEAD", "VINGRPS", "VINDONE"):
        v = C.c_int(-1)
        assert hp.hpc_status_geti4(k.encode(), C.byref(v)) == 1, k
        cnt[k] = v.value
    _say("[pipe-%s] input thread: %s" % (variant, cnt))
    b = ((pk[:, 5] >> 24) + ((pk[:, 4] & 0xFFFFFF) << 8)).astype(np.int64) - B_FIRST
    assert cnt["VINPKTS"] == pk.shape[0] and cnt["VINDONE"] == 1 and cnt["VINGRPS"] == n_groups
    assert cnt["VINSKIP"] == int((b == 0).sum()) and cnt["VINUSED"] == int((b > 0).sum())
    assert cnt["VININVAL"] == 0 and cnt["VINAHEAD"] == 0
    gsps = (pk.shape[0] - cnt["VINSKIP"]) * 1024 * 4 / dt / 1e9
    _say("[pipe-%s] %d groups through the chain in %.2f s: %.2f Gsample/s ingest-inclusive"
         % (variant, n_groups, dt, gsps))

    # expected: the C-ABI on the same groups, re-framed by the output thread's rule
    st = SdbeStage(policy=L.THRESH_REFERENCE, out_mode=L.OUT_PAYLOAD, rate_frac=rate if trimmed else 0,
                   skip_mhz=150 if trimmed else 0)
    tmpl = O.fill_header_template(pk[0, :8], T_REF)
    inval = tmpl.copy()
    inval[0] |= np.uint32(1 << 31)
    al = O.OutputAligner(station=O.station_from_str("Sm"), bdc=1, rx=0)
    want_frames, want_hdr = [[], []], [[], []]
    for g in range(n_groups):
        sel = (b >= 1 + frames * g) & (b <= frames * (g + 1))
        pg = np.ascontiguousarray(pk[sel])
        pay = st.process_group(st.to_device_packets(pg), pg.shape[0], B_FIRST + 1 + frames * g).cpu().numpy()
        fr, hd = al.push(pay.view(np.uint32).reshape(2, -1), tmpl if g == 0 else inval)
        for p in (0, 1):
            want_frames[p].append(fr[p])
            want_hdr[p].append(hd[p])
    st.close()
    for p in (0, 1):
        data = np.fromfile(fout[p], dtype=np.uint32).reshape(-1, 8224 // 4)
        wf = np.concatenate(want_frames[p])
        wh = np.concatenate(want_hdr[p])
        _say("[pipe-%s] pol %d: %d frames written, %d expected" % (variant, p, data.shape[0], wf.shape[0]))
        assert data.shape[0] == wf.shape[0]
        assert np.array_equal(data[:, :8], wh)
        assert np.array_equal(data[:, 8:], wf)
