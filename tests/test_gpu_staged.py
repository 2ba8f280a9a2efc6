"""The staged dendrite scan (per-block presynaptic entries in LDS) against the
CPU oracle and against the direct scan, on the cases that take its other
paths: blocks whose distinct presynaptic axons exceed the LDS capacity (the
slots of the axons past it spilled to the exact path) or the build's hash set
(the unstaged scan), segments whose segment + propagation length passes 255
(the exact path from the u32 heads), filled slots after an empty one
(ignored, as the reference's `break`), and targets changed between steps
(the plan is rebuilt). Each runs tick by tick and through the batch graph, where the staged pass and GSOP use
the plan's working copy of the weights."""
import ctypes as C

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd._ffi import day_lib
from axicor_amd.layout import StateView
from helpers import Pair

pytestmark = pytest.mark.gpu


@pytest.fixture
def direct_scan():
    """Runs the test with the direct scan, restoring the staged default."""
    day_lib().axc_staging_enable(0)
    yield
    day_lib().axc_staging_enable(1)


GRAPH = pytest.mark.parametrize("use_graph", [False, True], ids=["per_tick", "graph"])


@GRAPH
def test_spilled_slots_past_the_lds_capacity(gpu, use_graph):
    """40k neurons x 128 slots on a 100x100x4 grid with targets 12 voxels
    away on average: 80% of the 256-soma blocks have more than 8,704
    distinct presynaptic axons (median 10,700). They stage the first 8,704
    and the slots of the others take the exact path."""
    spec = B.spec_small(40000, 128, 2000, seed=37, axon_len=(1, 64), grid=(100, 100, 4))
    spec.mean_dist = 12.0
    p = Pair(B.generate(spec), n_out=32)
    for k in range(3):
        out_ref, _ = p.step(12, dopamine=(120 if k % 2 else -80), use_graph=use_graph)
        np.testing.assert_array_equal(p.outputs(12), out_ref)
        p.check(f"spilled slots batch {k}")


@GRAPH
def test_unstaged_blocks_hash_overflow(gpu, use_graph):
    """65,536 neurons x 128 slots on a flat 256x256 grid with targets 60
    voxels away on average: over half of the 256-soma blocks have more
    distinct presynaptic axons than the build's 16,384-entry hash set and
    take the unstaged scan; the others stage 8,704 and spill the rest."""
    spec = B.spec_small(65536, 128, 3276, seed=37, axon_len=(1, 64), grid=(256, 256, 1))
    spec.mean_dist = 60.0
    p = Pair(B.generate(spec), n_out=16)
    for k in range(3):
        p.step(8, dopamine=(90 if k % 2 else -60), use_graph=use_graph)
        p.check(f"hash overflow batch {k}")


@GRAPH
def test_wide_segments_use_the_compact_snapshot(gpu, use_graph):
    """Axons 200-255 segments long and propagation length 60: a large share
    of slots has segment + length > 255."""
    lut = B.lut_from([B.template_variant(False, signal_propagation_length=60),
                      B.template_variant(True, signal_propagation_length=60)])
    spec = B.spec_small(6000, 64, 400, seed=23, axon_len=(200, 255), lut=lut)
    p = Pair(B.generate(spec), n_out=0)
    for k in range(5):
        p.step(40, dopamine=200, use_graph=use_graph)
        p.check(f"wide segments batch {k}")


@GRAPH
def test_holes_after_the_first_empty_slot_are_ignored(gpu, use_graph):
    """Somas with filled slots after an empty one (as sprouting leaves them
    until the next sort): the scan and GSOP stop at the first empty target."""
    b = B.generate(B.spec_small(4000, 80, 200, seed=25))
    v = StateView(b.state, b.padded_n)
    rng = np.random.default_rng(3)
    somas = rng.choice(4000, 600, replace=False)
    hole = rng.integers(5, 70, somas.size)
    v.targets[hole, somas] = 0
    p = Pair(b, n_out=0)
    for k in range(3):
        p.step(20, dopamine=300, use_graph=use_graph)
        p.check(f"holes batch {k}")


@GRAPH
def test_targets_uploaded_between_steps_rebuild_the_plan(gpu, use_graph):
    """A host upload of the targets (as the Night's sprouting upload does)
    between two batches: the next batch sees the new targets."""
    p = Pair(B.generate(B.spec_small(5000, 64, 250, seed=27)), n_out=0)
    p.step(20, use_graph=use_graph)
    p.check("before upload")
    v = StateView(p.ref_state, p.b.padded_n)
    rng = np.random.default_rng(5)
    cols = rng.choice(p.b.padded_n, 1500, replace=False)
    rows = rng.integers(0, 64, cols.size)
    v.targets[rows, cols] = (rng.integers(0, 200, cols.size).astype(np.uint32) << 24) | \
        (rng.integers(0, p.b.total_axons, cols.size).astype(np.uint32) + 1)
    tg = np.ascontiguousarray(v.targets)
    assert day_lib().gpu_memcpy_host_to_device(p.vram.ptrs.dendrite_targets, tg.ctypes.data, tg.nbytes)
    for k in range(3):
        p.step(20, use_graph=use_graph)
        p.check(f"after upload batch {k}")


@GRAPH
def test_staged_and_direct_scans_agree(gpu, use_graph):
    """The same brain stepped with the staged scan and with the direct one
    ends in the same state (and both equal the oracle)."""
    spec = B.spec_small(9000, 100, 450, seed=29, axon_len=(1, 255), lut=B.lut_c2(),
                        type_ppm=[400_000, 400_000, 100_000, 100_000])
    b = B.generate(spec)
    a = Pair(b.copy(), n_out=24)
    for k in range(3):
        a.step(30, dopamine=(200 - 150 * k), use_graph=use_graph)
        a.check(f"staged batch {k}")
    lib = day_lib()
    lib.axc_staging_enable(0)
    try:
        d = Pair(b.copy(), n_out=24)
        for k in range(3):
            d.step(30, dopamine=(200 - 150 * k), use_graph=use_graph)
            d.check(f"direct batch {k}")
    finally:
        lib.axc_staging_enable(1)
    assert np.array_equal(a.vram.download_state(), d.vram.download_state())
    assert np.array_equal(a.vram.download_axons(), d.vram.download_axons())


def test_direct_scan_per_tick(gpu, direct_scan):
    """The direct scan still passes the per-tick path against the oracle."""
    p = Pair(B.generate(B.spec_small(3000, 48, 150, seed=31)), n_out=16)
    for t in range(10):
        out_ref, _ = p.step(1, dopamine=50)
        np.testing.assert_array_equal(p.outputs(1), out_ref)
        p.check(f"direct tick {t}")


def test_graph_with_spike_schedule_and_inputs(gpu):
    """Batches with input pushes every tick and network spikes (K3) on some
    ticks: tick t's GSOP in tick t+1's pass reads the previous compact
    snapshot for every slot of a tick with network spikes and for the slots
    of pushed inputs."""
    b = B.generate(B.spec_small(6000, 80, 1200, seed=33, axon_len=(1, 120)))
    p = Pair(b, max_spikes=256, max_batch=16)
    rng = np.random.default_rng(7)
    for k in range(4):
        T = 16
        sched = rng.integers(0, b.total_axons, size=(T, 256)).astype(np.uint32)
        counts = np.where(rng.random(T) < 0.4, rng.integers(1, 256, T), 0).astype(np.uint32)
        p.step(T, spikes=sched.reshape(-1), counts=counts, dopamine=(150 if k % 2 else -60), use_graph=True)
        p.check(f"spike schedule batch {k}")


def test_graph_heads_near_the_wrap_and_sentinel(gpu):
    """Heads close to the u32 wrap, at and next to the sentinel, and large
    heads, with v_seg = 3, through the batch graph: the fused GSOP's
    distances from this tick's byte entries at segment + v_seg equal those of
    the previous post-push heads."""
    spec = B.spec_small(4000, 64, 300, seed=35, axon_len=(1, 255))
    spec.v_seg = 3
    b = B.generate(spec)
    h = b.heads.reshape(-1, 8)
    rng = np.random.default_rng(9)
    rows = rng.choice(b.total_axons, 400, replace=False)
    patterns = np.array([[0xFFFFFFF0, 0x80000001, 0x7FFFFFFF, 0, 3, 0x80000000, 0xFFFFFFFF, 17],
                         [0xFFFFFFFE, 0xFFFFFFFD, 250, 251, 252, 253, 254, 255],
                         [0xFFFFFFFC, 1, 2, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF]], np.uint32)
    h[rows] = patterns[np.arange(rows.size) % 3]
    p = Pair(b, n_out=0)
    for k in range(4):
        p.step(12, dopamine=400, v_seg=3, use_graph=True)
        p.check(f"wrap heads batch {k}")


@GRAPH
@pytest.mark.parametrize("engine_order", [False, True], ids=["caller_order", "engine_order"])
def test_baker_numbering(gpu, use_graph, engine_order):
    """Somas numbered as genesis-baker numbers a shard (neuron_placement.rs
    :60-87: z-planes in order, each plane's somas shuffled): 60k neurons x
    100 slots on a 48x48x32 grid, 1.5M-neuron brain's four types, against
    the oracle; in the caller's order and in the engine's own (the soma
    positions given: bricks)."""
    spec = B.spec_small(60000, 100, 3000, seed=41, axon_len=(1, 255), grid=(48, 48, 32), lut=B.lut_c2(),
                        type_ppm=[400_000, 400_000, 100_000, 100_000])
    spec.brick = B.BAKER_ORDER
    spec.input_ppm = 20_000
    b = B.generate(spec)
    p = Pair(b, n_out=64, positions=b.soma_pos if engine_order else None)
    for k in range(3):
        out_ref, _ = p.step(20, dopamine=(60 if k % 2 else -40), use_graph=use_graph)
        np.testing.assert_array_equal(p.outputs(20), out_ref)
        p.check(f"baker numbering batch {k}")
