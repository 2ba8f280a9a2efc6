"""Windows of ticks (axc_set_window_ticks): the batch graphs stream the
dendrite slots once per window of 2-4 ticks, predict every slot's hits over
the window from the heads at its start, and correct the slots of pushed
axons tick by tick. Every case runs against the CPU oracle bit for bit:
ragged last windows and one-tick batches, inputs, spike schedules, own
spikes, heads that wrap past 2^32 inside a window, segments past the byte
form, spilled and unstaged soma blocks, the engine's soma order, Nights (the
reverse index follows the plan), v_seg too large for windows (per-tick
fallback), the work counters, and the timed kernel-by-kernel path."""
import ctypes as C

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd._ffi import TIMED_KERNELS, day_lib
from helpers import Pair

pytestmark = pytest.mark.gpu

ORDERS = pytest.mark.parametrize("engine_order", [False, True], ids=["caller_order", "engine_order"])


def random_positions(n, seed):
    """Soma positions that put the somas in a random order of bricks."""
    rng = np.random.default_rng(seed)
    x, y, z = rng.integers(0, 64, n), rng.integers(0, 64, n), rng.integers(0, 32, n)
    return (x | (y << 10) | (z << 20) | (1 << 28)).astype(np.uint32)


@pytest.fixture(params=[2, 3, 4], ids=["win2", "win3", "win4"])
def window(request, gpu):
    lib = day_lib()
    assert lib.axc_set_window_ticks(request.param) == 0
    assert lib.axc_window_ticks() == request.param
    yield request.param
    lib.axc_set_window_ticks(1)


@pytest.fixture
def window4(gpu):
    lib = day_lib()
    assert lib.axc_set_window_ticks(4) == 0
    yield 4
    lib.axc_set_window_ticks(1)


def test_window_setting_bounds(gpu):
    """Lengths above 4 are refused and leave the setting as it was; 0 and 1
    mean the per-tick pass."""
    lib = day_lib()
    assert lib.axc_set_window_ticks(3) == 0
    assert lib.axc_set_window_ticks(5) != 0
    assert lib.axc_window_ticks() == 3
    assert lib.axc_set_window_ticks(0) == 0
    assert lib.axc_window_ticks() == 1


def test_window_c1_batches_with_night(window):
    """C1 (4,096 neurons x 64 slots, 256 inputs at p=0.05) in 50-tick graph
    batches (a ragged last window for 3 and 4) to tick 500, Night every 250
    ticks, outputs mapped, dopamine of both signs."""
    p = Pair(B.generate(B.spec_c1()), n_out=64)
    while p.tick < 500:
        out_ref, _ = p.step(50, use_graph=True, dopamine=(150 if p.tick % 100 else -90))
        np.testing.assert_array_equal(p.outputs(50), out_ref, err_msg=f"window outputs ending {p.tick}")
        p.check(f"window {window}: batch ending tick {p.tick}")
        if p.tick % 250 == 0:
            p.night(15)
            p.check(f"window {window}: night at tick {p.tick}")


def test_window_short_and_ragged_batches(window):
    """Batches of 1, 2, 3, 5, 7 and 10 ticks: windows shorter than the
    setting, and last windows of every length."""
    p = Pair(B.generate(B.spec_small(3000, 48, 200, seed=41)), n_out=24)
    for k, T in enumerate([1, 2, 3, 5, 7, 10, 4, 9]):
        out_ref, _ = p.step(T, use_graph=True, dopamine=40 * k - 100)
        np.testing.assert_array_equal(p.outputs(T), out_ref, err_msg=f"outputs, batch {k} of {T} ticks")
        p.check(f"window {window}: batch {k} of {T} ticks")


@ORDERS
def test_window_schedules_and_one_graph(window, engine_order):
    """Spike schedules (empty, full, ragged; duplicates, out-of-range ids)
    push axons inside windows; one captured graph serves all six batches."""
    b = B.generate(B.spec_small(2048, 32, 128, seed=15))
    p = Pair(b, max_spikes=48, max_batch=10, n_out=24,
             positions=random_positions(b.padded_n, 5) if engine_order else None)
    rng = np.random.default_rng(21)
    T = 10
    for k in range(6):
        sched = rng.integers(0, b.total_axons, size=(T, 48)).astype(np.uint32)
        sched[:, 3] = sched[:, 4]
        sched[::3, 7] = b.total_axons + 1
        if k == 0:
            counts = np.zeros(T, np.uint32)
        elif k == 1:
            counts = np.full(T, 48, np.uint32)
        else:
            counts = rng.integers(0, 49, size=T).astype(np.uint32)
        out_ref, _ = p.step(T, spikes=sched.reshape(-1), counts=counts, use_graph=True, dopamine=30 * k)
        np.testing.assert_array_equal(p.outputs(T), out_ref, err_msg=f"outputs, schedule {k}")
        p.check(f"window {window}: schedule {k}")
    assert p.eng.graph_captures() == 1


@ORDERS
def test_window_heads_wrapping_and_edge_values(window, engine_order):
    """Heads a few ticks below 2^32 (they wrap inside a window and hit small
    segments afterwards), heads at and past the sentinel, targets with axon
    field 0 or >= A, shared and missing axons, extreme voltages, thresholds
    and weights, a refractory timer of 255."""
    b = B.generate(B.spec_small(1024, 64, 64, seed=9))
    v = b.view()
    n, A = b.padded_n, b.total_axons
    h = b.heads.reshape(-1, 8)
    h[5] = [0xFFFFFFF0, 0x80000001, 0x7FFFFFFF, 0, 3, 0x80000000, 0xFFFFFFFF, 17]
    h[6] = [1, 2, 3, 4, 5, 6, 7, 8]
    h[7] = [0xFFFFFFFE, 0xFFFFFFFC, 0xFFFFFFF9, 0x80000000, 0x80000000, 0x80000000, 0x80000000, 0x80000000]
    h[8] = [0xFFFFFFFB, 0xFFFFFFF6, 0xFFFFFFEE, 0xFFFFFFD0, 0x80000000, 0x80000000, 0x80000000, 0x80000000]
    v.targets[0, 10] = (7 << 24) | 0
    v.targets[1, 10] = (3 << 24) | (A + 9)
    v.targets[0, 11] = (0 << 24) | (5 + 1)
    v.targets[5, 12] = 0
    # Small segments of the wrapping axons on many somas.
    for k, col in enumerate(range(100, 160)):
        v.targets[k % 8, col] = ((k % 6) << 24) | (7 + 1 + (k % 2))
        v.weights[k % 8, col] = 900 if k % 3 else -700
    v.soma_to_axon[20] = v.soma_to_axon[21]
    v.soma_to_axon[22] = 0xFFFFFFFF
    v.soma_to_axon[23] = A + 100
    v.voltage[30] = 2**31 - 5
    v.voltage[31] = -2**31 + 3
    v.threshold_offset[32] = 2**31 - 10
    v.weights[:, 40] = -32768
    v.weights[:, 41] = 32767
    v.weights[:, 42] = 0
    v.timers[50] = 255
    p = Pair(b, n_out=16, positions=random_positions(b.padded_n, 3) if engine_order else None)
    for k in range(12):
        p.step(5, dopamine=(-1000 if k % 3 == 0 else 1000), use_graph=True)
        p.check(f"window {window}: edge batch {k}")
    p.night(100)
    p.check("edge night")
    p.step(7, use_graph=True)
    p.check("edge after night")


def test_window_types_v_seg_and_wide_segments(window):
    """Four variants (two adaptive), v_seg 3, axons 1-255 segments long:
    slots with segment + length > 255 take the u32 rule."""
    spec = B.spec_small(5000, 100, 300, seed=7, axon_len=(1, 255), lut=B.lut_c2(),
                        type_ppm=[400_000, 400_000, 100_000, 100_000])
    spec.v_seg = 3
    p = Pair(B.generate(spec), n_out=0)
    for k, dop in enumerate([0, 500, -700, 32767, -32768, 90]):
        p.step(11, dopamine=dop, v_seg=3, use_graph=True)
        p.check(f"window {window}: dopamine {dop}")


def test_window_long_propagation(window):
    """Propagation length 60 on axons 200-255 segments long: a large share of
    the slots is decided from the u32 heads."""
    lut = B.lut_from([B.template_variant(False, signal_propagation_length=60),
                      B.template_variant(True, signal_propagation_length=60)])
    spec = B.spec_small(6000, 64, 400, seed=23, axon_len=(200, 255), lut=lut)
    p = Pair(B.generate(spec), n_out=0)
    for k in range(4):
        p.step(13, dopamine=(-200 if k % 2 else 300), v_seg=2, use_graph=True)
        p.check(f"window {window}: long propagation batch {k}")


def test_window_short_refractory_and_short_propagation(window):
    """Refractory periods of 0-2 ticks, so a spiker's weight changes reach
    the window's later ticks (GSOP adds them to the predicted hits there),
    and propagation lengths of 1-3 segments below (K-1) v_seg, where the
    window pass tests each tick on its own, beside types that take the
    one-pass test, mixed in the same waves."""
    lut = B.lut_from([
        B.template_variant(False, refractory_period=0, signal_propagation_length=2, threshold=12000),
        B.template_variant(False, refractory_period=1, signal_propagation_length=10, threshold=12000),
        B.template_variant(True, refractory_period=2, signal_propagation_length=1, threshold=12000),
        B.template_variant(False, refractory_period=0, signal_propagation_length=3, threshold=12000,
                           adaptive_mode=1, adaptive_leak_gain=64, adaptive_leak_max=3000),
    ])
    spec = B.spec_small(4000, 64, 300, seed=31, axon_len=(1, 40), lut=lut,
                        type_ppm=[300_000, 300_000, 200_000, 200_000])
    for v_seg in (1, 2):
        p = Pair(B.generate(spec), n_out=32)
        fired = 0
        for k, dop in enumerate([0, 120, -90, 255]):
            out_ref, _ = p.step(13, dopamine=dop, v_seg=v_seg, use_graph=True)
            np.testing.assert_array_equal(p.outputs(13), out_ref)
            fired += int(out_ref.sum())
            p.check(f"window {window}: v_seg {v_seg}, short refractory batch {k}")
        assert fired > 0


def test_window_large_v_seg(window):
    """v_seg 9,000: windows of 2 ticks stay exact in the byte form; 3 and 4
    ticks reach 2^15 and the batch runs per tick (same results)."""
    spec = B.spec_small(2000, 48, 120, seed=19)
    p = Pair(B.generate(spec), n_out=8)
    for k in range(3):
        out_ref, _ = p.step(9, v_seg=9000, use_graph=True, dopamine=77)
        np.testing.assert_array_equal(p.outputs(9), out_ref)
        p.check(f"window {window}: v_seg 9000 batch {k}")


def test_window_spilled_and_unstaged_blocks(window4):
    """Soma blocks past the LDS capacity (spilled slots) and past the build's
    hash set (the unstaged scan) with windows of 4."""
    spec = B.spec_small(40000, 128, 2000, seed=37, axon_len=(1, 64), grid=(100, 100, 4))
    spec.mean_dist = 12.0
    p = Pair(B.generate(spec), n_out=32)
    for k in range(2):
        out_ref, _ = p.step(10, dopamine=(120 if k % 2 else -80), use_graph=True)
        np.testing.assert_array_equal(p.outputs(10), out_ref)
        p.check(f"spilled slots batch {k}")
    spec = B.spec_small(65536, 128, 3276, seed=37, axon_len=(1, 64), grid=(256, 256, 1))
    spec.mean_dist = 60.0
    p = Pair(B.generate(spec), n_out=16)
    for k in range(2):
        p.step(6, dopamine=(90 if k % 2 else -60), use_graph=True)
        p.check(f"hash overflow batch {k}")


def test_window_counters_match_oracle(window4):
    """The work counters of windowed batches equal the oracle's exact
    counts (hits, slots, integrations, spikes, GSOP)."""
    p = Pair(B.generate(B.spec_small(3000, 64, 200, seed=2)))
    p.eng.stats_reset()
    total = None
    for k in range(3):
        _, st = p.step(18, use_graph=True, dopamine=60)
        total = st if total is None else {kk: total[kk] + st[kk] for kk in st}
    got = {k: v for k, v in p.eng.stats().items() if not k.endswith("_fused")}
    assert got == total, (got, total)
    p.check("window counters")


def test_window_timed_path(window4):
    """With kernel timing on, the batch launches the windowed kernels one by
    one (no graph): same results, and the window kernels are the ones
    timed."""
    lib = day_lib()
    p = Pair(B.generate(B.spec_small(3000, 48, 200, seed=43)), n_out=16)
    lib.axc_kernel_timing_read(None, None, 1)
    lib.axc_kernel_timing_enable(1)
    try:
        for k in range(3):
            out_ref, _ = p.step(10, use_graph=True, dopamine=50 * k)
            np.testing.assert_array_equal(p.outputs(10), out_ref)
            p.check(f"timed window batch {k}")
    finally:
        lib.axc_kernel_timing_enable(0)
    ms = (C.c_double * len(TIMED_KERNELS))()
    n = (C.c_uint64 * len(TIMED_KERNELS))()
    assert lib.axc_kernel_timing_read(ms, n, 1) == 0
    got = dict(zip(TIMED_KERNELS, n))
    assert got["window_pass"] == 3 * 3       # windows of 4, 4, 2 per batch
    assert got["window_end"] == 3 * 3
    assert got["window_soma"] == 30 and got["window_correct"] == 30 and got["axon_tick"] == 30
    assert got["neuron_tick"] == 0
