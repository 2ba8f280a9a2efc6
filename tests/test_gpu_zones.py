"""Multi-zone helpers through the C-ABI against oracle/zone_oracle.py:
launch_ghost_sync (K11, bindings.cu:341-359), launch_extract_outgoing_spikes
(K12, bindings.cu:378-418), and two zones on one GPU where zone A's outgoing
spikes of each batch become zone B's K3 schedule for its next batch
(ring_buffer.rs:35-54), checked against the CPU oracle running both zones."""
import ctypes as C

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd._ffi import day_lib
from axicor_amd.engine import DeviceBuffer
from helpers import Pair
from oracle import zone_oracle as Z

pytestmark = pytest.mark.gpu


def _extract(heads_ptr, si: np.ndarray, gid: np.ndarray, T: int, v_seg: int, stream=None):
    lib = day_lib()
    d_si, d_gid = DeviceBuffer.from_array(si.astype(np.uint32)), DeviceBuffer.from_array(gid.astype(np.uint32))
    cap = 8 * len(si)
    d_ev, d_cnt = DeviceBuffer(max(cap, 1) * 8), DeviceBuffer(4)
    try:
        lib.launch_extract_outgoing_spikes(heads_ptr, d_si.ptr, d_gid.ptr, len(si), T, v_seg, d_ev.ptr, d_cnt.ptr,
                                           stream)
        lib.gpu_device_synchronize()
        n = int(d_cnt.download(np.uint32, 1)[0])
        assert n <= cap
        ev = d_ev.download(np.uint32, 2 * n).reshape(n, 2) if n else np.zeros((0, 2), np.uint32)
        return sorted((int(g), int(o)) for g, o in ev)
    finally:
        for b in (d_si, d_gid, d_ev, d_cnt):
            b.free()


def test_ghost_sync_matches_oracle(gpu):
    lib = day_lib()
    rng = np.random.default_rng(5)
    src = rng.integers(0, 2**32, size=(3000, 8), dtype=np.uint64).astype(np.uint32)
    dst = rng.integers(0, 2**32, size=(2000, 8), dtype=np.uint64).astype(np.uint32)
    n = 1500
    si = rng.integers(0, 3000, size=n).astype(np.uint32)
    si[::9] = 0x80000000   # dead axon
    si[4::13] = 0xFFFFFFFF  # soma without axon
    di = rng.permutation(2000)[:n].astype(np.uint32)  # distinct destinations: no write race
    d_src, d_dst = DeviceBuffer.from_array(src), DeviceBuffer.from_array(dst)
    d_si, d_di = DeviceBuffer.from_array(si), DeviceBuffer.from_array(di)
    try:
        lib.launch_ghost_sync(d_src.ptr, d_dst.ptr, d_si.ptr, d_di.ptr, n, None)
        lib.gpu_device_synchronize()
        got = d_dst.download(np.uint32, dst.size).reshape(dst.shape)
    finally:
        for b in (d_src, d_dst, d_si, d_di):
            b.free()
    want = dst.copy()
    Z.ghost_sync(src, want, si, di)
    np.testing.assert_array_equal(got, want)
    assert not np.array_equal(want, dst)


@pytest.mark.parametrize("v_seg,T", [(1, 50), (3, 50), (2, 1)])
def test_extract_outgoing_spikes_matches_oracle(gpu, v_seg, T):
    """Heads spread around the batch window and the dead-tail threshold, with
    poisoned source indices; the events as a multiset."""
    rng = np.random.default_rng(v_seg * 100 + T)
    A = 4000
    heads = rng.integers(0, 3 * T * v_seg + 8, size=(A, 8)).astype(np.uint32)
    heads[rng.random((A, 8)) < 0.3] = 0x80000000
    heads[rng.random((A, 8)) < 0.02] = 0x70000000 - rng.integers(0, 4)
    si = rng.integers(0, A, size=2500).astype(np.uint32)
    si[::17] = 0x80000000
    si[5::19] = 0xFFFFFFFF
    gid = rng.integers(0, 1 << 24, size=2500).astype(np.uint32)
    d_h = DeviceBuffer.from_array(heads)
    try:
        got = _extract(d_h.ptr, si, gid, T, v_seg)
    finally:
        d_h.free()
    want = sorted(Z.extract_outgoing_spikes(heads, si, gid, T, v_seg))
    assert len(want) > 100
    assert got == want


def test_two_zones_spike_exchange(gpu):
    """Zone A's 160 boundary axons feed zone B's 160 ghost axons (B's input
    axon range, driven only through the K3 schedule). Per 20-tick batch: A
    steps, its outgoing spikes are extracted on the GPU and by the oracle
    (equal as multisets), the oracle's events become B's schedule for its
    next batch, and B steps on the GPU and on the oracle; both zones are
    compared in full after every batch."""
    T = 20
    a = Pair(B.generate(B.spec_small(3000, 40, 100, seed=41, axon_len=(1, 24))), n_out=0)
    bb = B.generate(B.spec_small(2500, 40, 160, seed=42, axon_len=(1, 24), input_ppm=0))
    b = Pair(bb, n_out=16, max_spikes=8 * 160)
    rng = np.random.default_rng(43)
    si = rng.choice(a.b.padded_n, size=160, replace=False).astype(np.uint32)
    si[::23] = 0xFFFFFFFF  # a boundary pixel without an axon
    gid = (bb.virtual_offset + np.arange(160)).astype(np.uint32)
    ids = np.zeros((T, 8 * 160), np.uint32)
    cnt = np.zeros(T, np.uint32)
    no_input = np.zeros((T, bb.input_words), np.uint32)
    total = 0
    for k in range(4):
        # B steps on the schedule built from A's previous batch.
        b.step(T, spikes=ids, counts=cnt, inputs=no_input, use_graph=bool(k % 2))
        b.check(f"zone B batch {k}")
        a.step(T, use_graph=bool(k % 2))
        a.check(f"zone A batch {k}")
        heads_ptr = a.vram.ptrs.axon_heads
        got = _extract(heads_ptr, si, gid, T, 1, a.eng.stream)
        want = Z.extract_outgoing_spikes(a.ref_heads, si, gid, T, 1)
        assert got == sorted(want), f"batch {k}"
        total += len(want)
        ids, cnt = Z.schedule_from_events(want, T, 8 * 160)
    assert total > 50
