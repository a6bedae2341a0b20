"""GPU tests of the "8 snapshots x 1024 channels" packet geometry
(SDBE_GEOM_8X1024, the packets of the benchmark configuration's text; the
reference packs 128 snapshots x 8 channels).  Bars:
* unpack of a shuffled group with dropped, duplicate, invalid and
  out-of-window packets: int8 rows, the uint16 fid bitmap and the counters
  bit-exact against oracle.vdif_to_beng_8x1024 (pinned on the CPU against the
  reference geometry and a hand-built packet);
* with no packet dropped, sdbe_process_group's output equals, byte for byte,
  the same codes sent in the reference geometry;
* with drops, each block-pol's payload equals sdbe_process_spectra on the
  oracle's rows of that block-pol (PER_BLOCK thresholds), and the fid bitmap
  and counters of sdbe_process_group (whose stage A reads the packets) equal
  the oracle's.
"""
import sys

import numpy as np
import pytest

from oracle import sdbe_oracle as O
from sdbe_amd import synth as S

pytestmark = pytest.mark.gpu

B0 = 1000


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def codes():
    rng = S.rng_for(808)
    return S.gaussian_codes(rng, 39 * 128), S.gaussian_codes(rng, 39 * 128)


def test_unpack_8x1024_bit_exact(torch_cuda, codes):
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage
    c0, c1 = codes
    pk = S.packetize_8x1024(c0, c1, B0)
    rng = S.rng_for(809)
    pk, dropped = S.shuffle_and_drop(pk, rng, drop_frac=0.002)
    dup = pk[10].copy()
    dup[8:] ^= np.uint32(0x55555555)
    inval = pk[11].copy()
    inval[0] |= np.uint32(1 << 31)
    late = pk[12].copy()
    late[4] = (B0 + 39) >> 8
    late[5] = (late[5] & 0x00FFFFFF) | (((B0 + 39) & 0xFF) << 24)
    pk = np.ascontiguousarray(np.concatenate([pk, dup[None], inval[None], late[None]]))
    st = SdbeStage(packet_geometry=L.GEOM_8X1024)
    _say("[g2-unpack] gpu %d packets (%d dropped)" % (pk.shape[0], dropped.size))
    g0, g1, bm = st.unpack_int8(st.to_device_packets(pk), pk.shape[0], B0)
    stats = st.stats()
    want = O.vdif_to_beng_8x1024(pk, B0)
    assert np.array_equal(g0.cpu().numpy(), want.beng[0])
    assert np.array_equal(g1.cpu().numpy(), want.beng[1])
    assert np.array_equal(bm.cpu().numpy().view(np.uint16), want.bitmap)
    _say("[g2-unpack] stats %s, oracle unique %d missing %d dup %d invalid %d range %d"
         % (stats, want.n_unique, want.n_missing, want.n_duplicate, want.n_invalid, want.n_out_of_range))
    assert stats["n_unique"] == want.n_unique and stats["n_missing"] == want.n_missing
    assert stats["n_invalid"] == want.n_invalid and stats["n_duplicate"] == want.n_duplicate
    st.close()


def test_process_group_8x1024_equals_reference_geometry(torch_cuda, codes):
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage
    c0, c1 = codes
    rng = S.rng_for(810)
    pk_ref, _ = S.shuffle_and_drop(S.packetize(c0, c1, B0), rng, drop_frac=0.0)
    pk_g2, _ = S.shuffle_and_drop(S.packetize_8x1024(c0, c1, B0), rng, drop_frac=0.0)
    ref = SdbeStage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_FRAMES)
    g2 = SdbeStage(policy=L.THRESH_PER_BLOCK, out_mode=L.OUT_FRAMES, packet_geometry=L.GEOM_8X1024)
    for h in (ref, g2):
        h.set_state(secs=1000, df=0, ref_epoch=30, psn=1)
    _say("[g2-group] gpu process_group, both geometries")
    a = ref.process_group(ref.to_device_packets(pk_ref), pk_ref.shape[0], B0).cpu().numpy()
    b = g2.process_group(g2.to_device_packets(pk_g2), pk_g2.shape[0], B0).cpu().numpy()
    assert np.array_equal(a, b)
    ref.close()
    g2.close()


def test_process_group_8x1024_with_drops_equals_spectra_path(torch_cuda, codes):
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage
    c0, c1 = codes
    rng = S.rng_for(811)
    pk, dropped = S.shuffle_and_drop(S.packetize_8x1024(c0, c1, B0), rng, drop_frac=0.01)
    g2 = SdbeStage(policy=L.THRESH_PER_BLOCK, packet_geometry=L.GEOM_8X1024)
    _say("[g2-drops] gpu process_group (%d of %d packets dropped)" % (dropped.size, dropped.size + pk.shape[0]))
    bm = torch.full((39, 16), -1, dtype=torch.int16, device="cuda")
    out = g2.process_group(g2.to_device_packets(pk), pk.shape[0], B0, bitmap=bm).cpu().numpy()
    got = g2.stats()
    g2.close()
    ref = O.vdif_to_beng_8x1024(pk, B0)
    # the packet-reading stage A records the fid bitmap and the counters as the unpack does
    assert np.array_equal(bm.cpu().numpy().view(np.uint16), ref.bitmap)
    assert got["n_unique"] == ref.n_unique and got["n_missing"] == ref.n_missing
    rows = ref.beng.reshape(2, 128, 39, 16384, 2)
    blocks = (0, 1, 17, 63, 64, 100, 127)
    spec = SdbeStage(policy=L.THRESH_PER_BLOCK, max_spectra=2 * len(blocks))
    d = torch.from_numpy(np.ascontiguousarray(
        np.stack([rows[pol, blk] for blk in blocks for pol in (0, 1)]))).cuda()
    _say("[g2-drops] gpu process_spectra on the oracle's rows of blocks %s" % (blocks,))
    sp = spec.process_spectra(d).cpu().numpy()
    spec.close()
    for i, blk in enumerate(blocks):
        for pol in (0, 1):
            assert np.array_equal(out[pol, 64 * blk:64 * (blk + 1)], sp[2 * i + pol]), (blk, pol)


def test_unpack_8x1024_empty_group(torch_cuda):
    """No packet at all: every row is 0 + 0j and the bitmap is clear."""
    torch = torch_cuda
    from sdbe_amd import _lib as L
    from sdbe_amd.stage import SdbeStage
    st = SdbeStage(packet_geometry=L.GEOM_8X1024)
    _say("[g2-empty] gpu unpack of an empty group")
    dummy = torch.zeros(8224, dtype=torch.uint8, device="cuda")
    g0, g1, bm = st.unpack_int8(dummy, 0, B0)
    assert not g0.any().item() and not g1.any().item() and not bm.any().item()
    assert st.stats()["n_unique"] == 0
    st.close()
