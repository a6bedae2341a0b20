"""CPU tests of the C host side (sdbe_amd/csrc/aphids_host.c, the plugin
registration and the databuf layouts) against the oracle.  No GPU calls."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import sdbe_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def hp():
    from sdbe_amd import databuf
    return databuf.load()


def test_layout_header_compiles_with_its_static_asserts(tmp_path):
    src = tmp_path / "l.c"
    src.write_text('#include "aphids_databuf.h"\nint main(void){return (int)sizeof(vdif_in_databuf_t) & 0;}\n')
    subprocess.run(["gcc", "-std=gnu11", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "sdbe_amd", "csrc", "hashpipe_compat"), str(src)], check=True)


def test_plugin_registered_with_reference_name(hp):
    td = hp.find_hashpipe_thread(b"vdif_inout_gpu_thread")
    assert td
    assert td.contents.name == b"vdif_inout_gpu_thread" and td.contents.skey == b"VDIFIO"
    assert td.contents.run and td.contents.ibuf_create and td.contents.obuf_create
    assert not td.contents.init


def _new_hdr(sec, clk, c, f, invalid=0):
    b53 = (sec << 29) | clk
    w = np.zeros(O.PKT_WORDS, dtype=np.uint32)
    w[0] = invalid << 31
    w[4] = (b53 >> 21) & 0xFFFFFFFF
    w[5] = (c & 0xFF) | ((f & 7) << 8) | ((b53 & 0x1FFFFF) << 11)
    return w


def test_beng_reform_matches_oracle(hp):
    rng = np.random.default_rng(7)
    T0 = (3000, 285_800_000)
    t_abs0 = T0[0] * O.SWARM_CLK_PER_SEC + T0[1]
    rows = []
    for i in range(300):
        k = int(rng.integers(0, 40)) if i % 50 else int(rng.integers(1500, 3000))   # some far jumps
        jitter = int(rng.integers(-1000, 1000))
        sec, clk = divmod(t_abs0 + k * O.SWARM_BFRAME_CLK + jitter, O.SWARM_CLK_PER_SEC)
        rows.append(_new_hdr(sec, clk, int(rng.integers(0, 256)), int(rng.integers(0, 8)),
                             invalid=int(rng.random() < 0.05)))
    pk = np.stack(rows)
    ref = O.BengReformer()
    want = ref.reform(pk[:, :8].copy())
    hp.beng_reform_reset()
    got = pk.copy()
    hp.beng_reform_headers(got.ctypes.data, got.shape[0])
    assert np.array_equal(got[:, :8], want)
    from sdbe_amd.databuf import BengTimestamp
    t = BengTimestamp()
    hp.get_beng_t0(C.byref(t))
    assert [t.sec, t.clk] == ref.beng_t0()


def test_unpack_8pac_matches_oracle(hp):
    rng = np.random.default_rng(8)
    raw = rng.integers(0, 256, (3, 8504), dtype=np.uint64).astype(np.uint8)
    want = O.unpack_8pac(raw)
    got = np.zeros((24, O.PKT_WORDS), dtype=np.uint32)
    src = raw.copy()              # unpack_8pac rewrites the source headers in place: keep it alive
    hp.unpack_8pac(got.ctypes.data, src.ctypes.data, 3)
    assert np.array_equal(got, want)


def test_group_window_and_insert_match_oracle(hp):
    from sdbe_amd import databuf as D
    rng = np.random.default_rng(9)
    db = D.VdifInDatabuf()
    cpu_bufs = [np.zeros((39 * 2048, O.PKT_WORDS), dtype=np.uint32) for _ in range(4)]
    b_first = 5000
    groups = []
    for i in range(4):
        hp.init_beng_group(C.byref(db.bgc[i]), cpu_bufs[i].ctypes.data, None, b_first + 39 * i)
        groups.append(O.init_beng_group(b_first + 39 * i))
    assert db.bgc[0].vdif_header_template.w0.invalid == 1
    n = 3000
    pk = np.zeros((n, O.PKT_WORDS), dtype=np.uint32)
    b = rng.integers(b_first - 5, b_first + 39 * 9, n)
    pk[:, 4] = b >> 8
    pk[:, 5] = rng.integers(0, 256, n) | (rng.integers(0, 8, n) << 16) | ((b & 0xFF) << 24)
    pk[:, 0] = (rng.random(n) < 0.03).astype(np.uint32) << 31
    pk[:, 8:] = rng.integers(0, 2 ** 32, (n, 256), dtype=np.uint64).astype(np.uint32)
    for i in range(n):
        p = D.VdifInPacket.from_buffer(pk[i])
        off_c = hp.get_beng_group_index_offset(C.byref(db), 1, C.byref(p))
        off_o = O.group_index_offset(groups, 1, pk[i])
        assert off_c == off_o, i
        if off_c >= 0:
            assert hp.insert_vdif_in_beng_group_buffer(C.byref(db), 1, off_c, C.byref(p)) == \
                O.insert_packet(groups, 1, off_c, pk[i], i)
    for gi in range(4):
        g, o = db.bgc[gi], groups[gi]
        assert g.beng_group_vdif_packet_count == o.packet_count
        assert [g.bfc[f].beng_frame_vdif_packet_count for f in range(39)] == o.frame_count
        cc = np.array([list(g.bfc[f].cc) for f in range(39)], dtype=np.uint8)
        assert np.array_equal(cc, o.cc)
        # packets copied into the host buffer in insertion order
        assert np.array_equal(cpu_bufs[gi][:o.packet_count], pk[o.packets])
        assert hp.check_beng_group_complete(C.byref(db), gi) == O.group_complete(o)
    assert hp.check_beng_group_complete(C.byref(db), 7) == -1


def test_fill_header_template_matches_oracle(hp):
    from sdbe_amd import databuf as D
    for T0 in [(100, 143_000_000), (7, 1000), (7, 285_000_000), (2000, 12_345_678), (9, 100)]:
        hp.beng_reform_reset()
        first = _new_hdr(T0[0], T0[1], 3, 1)
        first[1] = 30 << 24
        arr = np.stack([first])
        hp.beng_reform_headers(arr.ctypes.data, 1)       # sets T0
        ref_pkt = D.VdifInPacket.from_buffer(arr[0])
        out = D.VdifInHeader()
        hp.fill_vdif_header_template(C.byref(out), C.byref(ref_pkt), 0)
        want = O.fill_header_template(arr[0, :8], T0)
        assert np.array_equal(np.frombuffer(bytes(out), dtype=np.uint32), want)
