"""Bit-exact parity of the HIP engine (through the C-ABI) with the CPU oracle.

Every test runs the same seeded brain and inputs on both sides and compares
all eight `.state` arrays and the axon heads byte for byte (and the output
raster where outputs are mapped)."""
import ctypes as C

import numpy as np
import pytest

from axicor_amd import brain as B
from helpers import Pair

pytestmark = pytest.mark.gpu

ORDERS = pytest.mark.parametrize("engine_order", [False, True], ids=["caller_order", "engine_order"])


def random_positions(n, seed):
    """Soma positions that put the somas in a random order of bricks."""
    rng = np.random.default_rng(seed)
    x, y, z = rng.integers(0, 64, n), rng.integers(0, 64, n), rng.integers(0, 32, n)
    return (x | (y << 10) | (z << 20) | (1 << 28)).astype(np.uint32)


def test_c1_per_tick_then_batches(gpu):
    """C1 workload: 4,096 neurons x 64 slots, 256 inputs at p=0.05; every tick
    for the first 20 ticks, then 50-tick batches to tick 1,000 with the Night
    sort&prune every 250 ticks (prune 15)."""
    p = Pair(B.generate(B.spec_c1()), n_out=96)
    for t in range(20):
        out_ref, _ = p.step(1)
        np.testing.assert_array_equal(p.outputs(1), out_ref, err_msg=f"outputs tick {t}")
        p.check(f"tick {t}")
    while p.tick < 1000:
        T = min(50, 250 - p.tick % 250)
        out_ref, _ = p.step(T)
        np.testing.assert_array_equal(p.outputs(T), out_ref, err_msg=f"outputs batch ending {p.tick}")
        p.check(f"batch ending tick {p.tick}")
        if p.tick % 250 == 0:
            p.night(15)
            p.check(f"night at tick {p.tick}")


def test_graph_batch_equals_oracle(gpu):
    """The HIP-graph batch path (axc_step_day_phase_batch), replayed several
    times with the program's own synchronisations in between."""
    p = Pair(B.generate(B.spec_small(3000, 48, 200, seed=11)), n_out=40)
    for k in range(6):
        out_ref, _ = p.step(25, use_graph=True, dopamine=(-300 if k % 2 else 250))
        np.testing.assert_array_equal(p.outputs(25), out_ref, err_msg=f"graph outputs replay {k}")
        p.check(f"graph replay {k}")


def test_four_types_adaptive_dopamine(gpu):
    """Four variants (two with the adaptive leak), nonzero dopamine of both
    signs, padded neurons (n not a multiple of 64), v_seg = 2."""
    spec = B.spec_small(5000, 100, 300, seed=7, axon_len=(1, 255), lut=B.lut_c2(),
                        type_ppm=[400_000, 400_000, 100_000, 100_000])
    spec.v_seg = 2
    p = Pair(B.generate(spec), n_out=0)
    for k, dop in enumerate([0, 500, -700, 32767, -32768, 90]):
        p.step(10, dopamine=dop, v_seg=2)
        p.check(f"dopamine {dop}")


def test_spike_schedule_duplicates_and_bad_ids(gpu):
    """K3 network spikes: repeated ids push once per occurrence, ids >= A are
    ignored (physics.cu:127)."""
    b = B.generate(B.spec_small(2048, 32, 128, seed=5))
    p = Pair(b, max_spikes=64, max_batch=8)
    rng = np.random.default_rng(3)
    T = 8
    sched = np.zeros((T, 64), np.uint32)
    counts = rng.integers(0, 64, size=T).astype(np.uint32)
    for t in range(T):
        ids = rng.integers(0, b.total_axons, size=64).astype(np.uint32)
        ids[:6] = ids[6]                # duplicates
        ids[10] = b.total_axons + 3     # out of range
        ids[11] = 0xFFFFFFFF
        sched[t] = ids
    p.step(T, spikes=sched.reshape(-1), counts=counts)
    p.check("spike schedule")


@ORDERS
def test_graph_reused_across_varying_schedules(gpu, engine_order):
    """The batch graph does not depend on the spike counts: six graph batches
    whose schedules differ (empty, full slices, ragged counts, duplicates,
    out-of-range ids) replay one captured graph, each equal to the oracle;
    in the caller's soma order and in the engine's (random positions)."""
    b = B.generate(B.spec_small(2048, 32, 128, seed=15))
    p = Pair(b, max_spikes=48, max_batch=10, n_out=24,
             positions=random_positions(b.padded_n, 5) if engine_order else None)
    rng = np.random.default_rng(21)
    T = 10
    for k in range(6):
        sched = rng.integers(0, b.total_axons, size=(T, 48)).astype(np.uint32)
        sched[:, 3] = sched[:, 4]            # duplicates
        sched[::3, 7] = b.total_axons + 1    # out of range
        if k == 0:
            counts = np.zeros(T, np.uint32)
        elif k == 1:
            counts = np.full(T, 48, np.uint32)
        else:
            counts = rng.integers(0, 49, size=T).astype(np.uint32)
            counts[rng.integers(0, T)] = 0
        out_ref, _ = p.step(T, spikes=sched.reshape(-1), counts=counts, use_graph=True, dopamine=30 * k)
        np.testing.assert_array_equal(p.outputs(T), out_ref, err_msg=f"outputs, schedule {k}")
        p.check(f"graph batch with schedule {k}")
    assert p.eng.graph_captures() == 1


@ORDERS
def test_edge_state_values(gpu, engine_order):
    """Hand-placed edge cases: heads close to the u32 wrap and past the
    sentinel, targets whose axon field is 0 or >= A, two somas sharing one
    axon, somas without an axon, extreme voltages/thresholds/weights; in the
    caller's soma order and in the engine's (random positions), per tick."""
    b = B.generate(B.spec_small(1024, 64, 64, seed=9))
    v = b.view()
    n, A = b.padded_n, b.total_axons
    h = b.heads.reshape(-1, 8)
    h[5] = [0xFFFFFFF0, 0x80000001, 0x7FFFFFFF, 0, 3, 0x80000000, 0xFFFFFFFF, 17]
    h[6] = [1, 2, 3, 4, 5, 6, 7, 8]
    v.targets[0, 10] = (7 << 24) | 0              # axon field 0 -> index 0xFFFFFFFF
    v.targets[1, 10] = (3 << 24) | (A + 9)        # axon >= A
    v.targets[0, 11] = (0 << 24) | (5 + 1)        # seg 0 of the wrapped axon
    v.targets[5, 12] = 0                          # hole: later slots are ignored
    v.soma_to_axon[20] = v.soma_to_axon[21]       # shared axon
    v.soma_to_axon[22] = 0xFFFFFFFF
    v.soma_to_axon[23] = A + 100                  # out of range: ignored
    v.voltage[30] = 2**31 - 5
    v.voltage[31] = -2**31 + 3
    v.threshold_offset[32] = 2**31 - 10
    v.weights[:, 40] = -32768
    v.weights[:, 41] = 32767
    v.weights[:, 42] = 0
    v.timers[50] = 255
    p = Pair(b, n_out=16, positions=random_positions(b.padded_n, 3) if engine_order else None)
    for k in range(12):
        p.step(5, dopamine=(-1000 if k % 3 == 0 else 1000))
        p.check(f"edge step {k}")
    p.night(100)
    p.check("edge night")
    p.step(5)
    p.check("edge after night")


def test_sort_and_prune_random_columns(gpu):
    """K9 alone on columns with random fill, holes, ties and garbage in empty
    slots, at prune thresholds 0, 15, 300 and -5."""
    b = B.generate(B.spec_small(1500, 128, 0, seed=21))
    v = b.view()
    rng = np.random.default_rng(8)
    v.weights[:] = rng.integers(-600, 600, size=v.weights.shape).astype(np.int16)
    v.weights[:, ::3] = rng.choice([-20, 20, 15, -15, 0], size=v.weights[:, ::3].shape)
    holes = rng.random(v.targets.shape) < 0.3
    v.targets[holes] = 0
    v.dendrite_timers[:] = rng.integers(0, 255, size=v.dendrite_timers.shape).astype(np.uint8)
    v.flags[:] = rng.integers(0, 256, size=v.flags.shape).astype(np.uint8)
    p = Pair(b)
    for prune in (0, 15, 300, -5):
        p.night(prune)
        p.check(f"sort&prune {prune}")


def test_telemetry_ascending(gpu):
    """launch_extract_telemetry lists the spiking somas in ascending order and
    appends after the caller's count."""
    from axicor_amd.engine import DeviceBuffer
    from axicor_amd._ffi import day_lib
    p = Pair(B.generate(B.spec_small(4096, 32, 256, seed=4)))
    p.step(3)
    lib = day_lib()
    ids = DeviceBuffer(4 * p.b.padded_n)
    cnt = DeviceBuffer(4)
    cnt.upload(np.array([0], np.uint32))
    lib.launch_extract_telemetry(C.byref(p.vram.ptrs), p.b.padded_n, ids.ptr, cnt.ptr, p.eng.stream)
    p.eng.synchronize()
    n = int(cnt.download(np.uint32, 1)[0])
    got = ids.download(np.uint32, n)
    flags = p.vram.download_state()[4 * p.b.padded_n:5 * p.b.padded_n]
    want = np.nonzero(flags & 1)[0].astype(np.uint32)
    np.testing.assert_array_equal(got, want)


def test_counters_match_oracle(gpu):
    """The engine's work counters equal the oracle's exact counts."""
    p = Pair(B.generate(B.spec_small(3000, 64, 200, seed=2)))
    p.eng.stats_reset()
    total = None
    for k in range(4):
        _, st = p.step(20)
        total = st if total is None else {kk: total[kk] + st[kk] for kk in st}
    got = p.eng.stats()
    got = {k: v for k, v in got.items() if not k.endswith("_fused")}
    assert got == total, (got, total)


def test_c1_graph_batches_with_night(gpu):
    """C1 through the batch entry point (HIP graph, GSOP of tick t fused into
    tick t+1): 50-tick batches to tick 1,000, Night every 250 ticks, nonzero
    dopamine, outputs mapped."""
    p = Pair(B.generate(B.spec_c1()), n_out=64)
    while p.tick < 1000:
        out_ref, _ = p.step(50, use_graph=True, dopamine=120)
        np.testing.assert_array_equal(p.outputs(50), out_ref, err_msg=f"graph outputs ending {p.tick}")
        p.check(f"graph batch ending tick {p.tick}")
        if p.tick % 250 == 0:
            p.night(15)
            p.check(f"night at tick {p.tick}")


def test_graph_and_per_tick_paths_agree_on_counters(gpu):
    """Work counters of the fused batch path equal the oracle's."""
    p = Pair(B.generate(B.spec_small(3000, 64, 200, seed=12)))
    p.eng.stats_reset()
    total = None
    for k in range(3):
        _, st = p.step(17, use_graph=True)
        total = st if total is None else {kk: total[kk] + st[kk] for kk in st}
    got = {k: v for k, v in p.eng.stats().items() if not k.endswith("_fused")}
    assert got == total
    p.check("graph counters")


@ORDERS
def test_night_phase_with_sprouting(gpu, engine_order):
    """The Night sequence of shard_thread.rs:269-378 on the GPU shard (flags
    before pruning, GPU sort&prune, host sprouting, upload) against the CPU
    path (oracle sort&prune, the same sprouting on host arrays); the Day Phase
    continues afterwards on both (in the engine's soma order: the plan is
    rebuilt in it after each upload)."""
    from axicor_amd import night as NT
    from axicor_amd.layout import StateView
    from oracle import oracle as O
    b = B.generate(B.spec_c1(), with_paths=True)
    p = Pair(b, positions=random_positions(b.padded_n, 7) if engine_order else None)
    types = NT.sprout_types_from_lut(b.lut)
    geo_gpu, geo_cpu = NT.geometry_of(b), NT.geometry_of(b)
    for night in range(2):
        for _ in range(5):
            p.step(50, use_graph=True)
        p.check(f"before night {night}")
        res = NT.execute_night_phase(p.vram, 15, geo_gpu, types, epoch=night, master_seed=b.spec.seed)
        v = StateView(p.ref_state, b.padded_n)
        flags = v.flags.copy()
        O.sort_and_prune_blob(p.ref_state, b.padded_n, 15)
        t, w = v.targets.copy(), v.weights.copy()
        ref = NT.sprout_pass(t, w, flags, v.soma_to_axon.copy(), b.padded_n, geo_cpu, types, night, b.spec.seed)
        v.targets[:] = t
        v.weights[:] = w
        assert res.new_synapses == ref.new_synapses > 0
        p.check(f"after night {night}")
    p.step(50, use_graph=True)
    p.check("day after nights")


def test_night_through_the_shm_contract(gpu):
    """The Night as genesis-node runs it with genesis-baker-daemon: the shard's
    side copies the flags, the GPU-pruned matrices and the request into the
    shared region (axicor_amd/shm.py), the daemon's side sprouts on the
    region, the shard uploads. It ends in the same state as the twin shard
    whose Night ran through execute_night_phase, and stays equal after more
    day ticks."""
    from axicor_amd import night as NT
    from axicor_amd import shm as S
    from axicor_amd.layout import StateView
    b = B.generate(B.spec_c1(), with_paths=True)
    p, q = Pair(b.copy()), Pair(b.copy())
    types = NT.sprout_types_from_lut(b.lut)
    for _ in range(4):
        p.step(50, use_graph=True)
        q.step(50, use_graph=True)
    NT.execute_night_phase(q.vram, 15, NT.geometry_of(b), types, epoch=0, master_seed=b.spec.seed)
    region = S.ShmRegion.create(0x5151, b.padded_n, b.total_axons)
    req = S.runtime_night_begin(p.vram, region, 15, max_sprouts=4)
    reply = S.daemon_night(region, req, NT.geometry_of(b), b.view().soma_to_axon, types, b.spec.seed)
    acks = S.runtime_night_end(p.vram, region, reply)
    assert len(acks) == 0 and int(region.header["state"]) == S.ShmState.IDLE
    sp, sq = p.vram.download_state(), q.vram.download_state()
    assert np.array_equal(sp, sq)
    assert not np.array_equal(StateView(sp, b.padded_n).targets, b.view().targets)  # pruned and sprouted
    p.step(50, use_graph=True)
    q.step(50, use_graph=True)
    assert np.array_equal(p.vram.download_state(), q.vram.download_state())
    assert np.array_equal(p.vram.download_axons(), q.vram.download_axons())


def test_hot_checkpoint_resume(gpu, tmp_path):
    """save_hot_checkpoint (shard_thread.rs:229-263) after 50 ticks: the files
    hold the oracle's state and heads byte for byte, and a fresh shard
    resumed from them continues exactly as the uninterrupted one."""
    from axicor_amd import baked as K
    b = B.generate(B.spec_c1())
    p = Pair(b.copy(), n_out=32)
    p.step(50, use_graph=True, dopamine=60)
    s_path, a_path = K.save_checkpoint(p.vram, str(tmp_path))
    assert open(s_path, "rb").read() == p.ref_state.tobytes()
    assert open(a_path, "rb").read() == p.ref_heads.tobytes()
    q = Pair(b.copy(), n_out=32)
    K.load_checkpoint(q.vram, str(tmp_path))
    q.ref_state[:] = p.ref_state
    q.ref_heads[:] = p.ref_heads
    q.tick = p.tick
    for k in range(2):
        out_p, _ = p.step(50, use_graph=True, dopamine=60)
        out_q, _ = q.step(50, use_graph=True, dopamine=60)
        np.testing.assert_array_equal(p.outputs(50), q.outputs(50))
        assert np.array_equal(p.vram.download_state(), q.vram.download_state())
        assert np.array_equal(p.vram.download_axons(), q.vram.download_axons())
        q.check(f"resumed batch {k}")


@ORDERS
def test_write_slots_then_step(gpu, engine_order):
    """axc_shard_write_slots puts new synapses into empty slots (targets and
    weights) after a batch, as a host Night's upload does; a column past the
    rows is skipped. The next batches (graph and windows of 4) must read the
    new targets (the staging plan is rebuilt) and equal the oracle given the
    same slots."""
    from axicor_amd._ffi import day_lib
    from axicor_amd.engine import DeviceBuffer
    from axicor_amd.layout import StateView
    b = B.generate(B.spec_c1())
    N = b.padded_n
    p = Pair(b, max_batch=50, positions=random_positions(N, 5) if engine_order else None)
    p.step(30, use_graph=True)
    p.check("before the writes")
    v = StateView(p.ref_state, N)
    rng = np.random.default_rng(3)
    t_all = v.targets.reshape(-1)
    empty = np.flatnonzero(t_all == 0)
    cols = np.sort(rng.choice(empty, size=min(4000, len(empty)), replace=False)).astype(np.uint32)
    axons = rng.integers(0, b.total_axons, size=len(cols))
    segs = rng.integers(0, 8, size=len(cols))
    tgt = ((segs << 24) | (axons + 1)).astype(np.uint32)
    wts = rng.integers(-400, 400, size=len(cols)).astype(np.int16)
    t_all[cols] = tgt
    v.weights.reshape(-1)[cols] = wts
    junk = np.uint32(128 * N + 17)  # past the rows: skipped
    d_cols = DeviceBuffer.from_array(np.append(cols, junk))
    d_t = DeviceBuffer.from_array(np.append(tgt, np.uint32(5)))
    d_w = DeviceBuffer.from_array(np.append(wts, np.int16(9)))
    lib = day_lib()
    err = lib.axc_shard_write_slots(C.byref(p.vram.ptrs), N, d_cols.ptr, d_t.ptr, d_w.ptr, len(cols) + 1,
                                    p.eng.stream)
    assert err == 0
    p.eng.synchronize()
    p.check("after the writes")
    p.step(30, use_graph=True)
    p.check("a batch after the writes")
    lib.axc_set_window_ticks(4)
    try:
        p.step(30, use_graph=True)
    finally:
        lib.axc_set_window_ticks(1)
    p.check("windows of 4 after the writes")
