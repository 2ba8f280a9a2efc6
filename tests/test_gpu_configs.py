"""Every benchmark configuration at its own size, on the GPU, against the CPU
oracle (and, for the batched brains, against each brain run alone):

- the 1.5M-neuron brain (100 slots, 150M synapses): every soma's spike bit
  every tick for 500 ticks through the batch graph (windows of 4 ticks to
  tick 250, the per-tick pass after), the whole state and the heads at
  ticks 250 and 500, one Night (GPU sort & prune at 15, host sprouting, the
  upload, shard_thread.rs:269-378) checked against the oracle's Night, then
  50 more ticks in windows of 4;
- 128 batched brains of 8k-250k neurons with 32, 64 or 128 slots, each with
  its own parameter table, inputs and dopamine, for 100 ticks: every brain's
  state and heads equal to the same brain stepped alone, and three sampled
  brains (the smallest, the largest, one other) equal to the oracle;
- surrogate-data trials (20k neurons x 64 slots, 2,000 ticks, the search's
  own Simulator on the batched engine): the final state of each trial equal
  to the oracle's;
- the 12M-neuron scale point (128 slots, 1.5B synapses) for 100 ticks per
  tick and 50 in windows of 4: 64 mapped somas' spike bits every tick and the
  whole state after each batch of 50.
"""
import ctypes as C

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd import night as NT
from axicor_amd._ffi import day_lib
from axicor_amd.batch import BatchEngine
from axicor_amd.engine import DeviceBuffer, ShardEngine, VramState, upload_constant_memory
from axicor_amd.layout import StateView
from helpers import Pair
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_c2_500_ticks_then_a_night_with_sprouting(gpu):
    b = B.generate(B.spec_c2(), with_paths=True)
    N = b.padded_n
    p = Pair(b, n_out=N, max_batch=50)
    p.mapped = np.arange(N, dtype=np.uint32)  # every soma's spike bit, every tick
    p.d_mapped = DeviceBuffer.from_array(p.mapped)
    types = NT.sprout_types_from_lut(b.lut)
    geo_gpu, geo_cpu = NT.geometry_of(b), NT.geometry_of(b)
    lib = day_lib()
    try:
        while p.tick < 500:
            lib.axc_set_window_ticks(4 if p.tick < 250 else 1)
            out_ref, _ = p.step(50, dopamine=(40 if (p.tick // 50) % 2 else -25), use_graph=True)
            np.testing.assert_array_equal(p.outputs(50), out_ref, err_msg=f"C2 raster, ticks to {p.tick}")
            print(f"[c2] ticks to {p.tick}: raster equal, {int(out_ref.sum())} spikes", flush=True)
            if p.tick == 250:
                p.check("C2 at tick 250 (windows of 4)")
    finally:
        lib.axc_set_window_ticks(1)
    p.check("C2 before the Night at tick 500")
    res = NT.execute_night_phase(p.vram, 15, geo_gpu, types, epoch=0, master_seed=b.spec.seed)
    v = StateView(p.ref_state, N)
    flags = v.flags.copy()
    O.sort_and_prune_blob(p.ref_state, N, 15)
    t, w = v.targets.copy(), v.weights.copy()
    ref = NT.sprout_pass(t, w, flags, v.soma_to_axon.copy(), N, geo_cpu, types, 0, b.spec.seed)
    v.targets[:] = t
    v.weights[:] = w
    assert res.new_synapses == ref.new_synapses > 0
    for f in ("lengths", "paths", "tips"):
        np.testing.assert_array_equal(getattr(geo_gpu, f), getattr(geo_cpu, f), err_msg=f)
    p.check("C2 after the Night at tick 500")
    print(f"[c2] Night at 500: {ref.new_synapses} new synapses, state equal", flush=True)
    lib.axc_set_window_ticks(4)
    try:
        out_ref, _ = p.step(50, dopamine=10, use_graph=True)
    finally:
        lib.axc_set_window_ticks(1)
    np.testing.assert_array_equal(p.outputs(50), out_ref, err_msg="C2 raster after the Night")
    p.check("C2 at tick 550")


def _load_member(b, dop, inputs):
    v = VramState(b.padded_n, b.total_axons)
    v.upload_state(b.state)
    v.upload_axon_heads(b.heads)
    d_in = DeviceBuffer.from_array(inputs)
    return {"vram": v, "lut": b.lut, "d_inputs": d_in, "input_words": b.input_words, "input_ticks": inputs.shape[0],
            "v_seg": 1, "virtual_offset": b.virtual_offset, "n_inputs": b.n_inputs, "dopamine": dop}


def test_c3_128_brains_batched_equal_solo_and_oracle(gpu):
    T, batches = 50, 2
    ticks = T * batches
    rng = np.random.default_rng(20261204)
    members = []
    for k in range(128):
        spec, params = B.spec_batch_member(rng)
        b = B.generate(spec)
        members.append((b, int(params[6]), b.inputs(0, ticks)))
    sizes = [b.padded_n for b, _, _ in members]
    slots = sorted(set(b.spec.slots_filled for b, _, _ in members))
    print(f"[c3] 128 brains, {sum(sizes)} neurons ({min(sizes)}..{max(sizes)}), slots {slots}", flush=True)
    assert min(sizes) < 16_000 and max(sizes) > 150_000 and len(slots) == 3
    lib = day_lib()
    mems = [_load_member(b, dop, inp) for b, dop, inp in members]
    be = BatchEngine(mems)
    for j in range(batches):
        be.step(j * T, T)
    lib.gpu_device_synchronize()
    be.destroy()
    order = np.argsort(sizes)
    sampled = {int(order[0]), int(order[-1]), int(order[len(order) // 2])}
    for k, (b, dop, inp) in enumerate(members):
        m = _load_member(b, dop, inp)
        upload_constant_memory(b.lut)
        eng = ShardEngine(m["vram"])
        for j in range(batches):
            err = lib.axc_step_day_phase_batch(C.byref(m["vram"].ptrs), b.padded_n, b.total_axons, 1, j * T, T,
                                               m["d_inputs"].ptr + j * T * b.input_words * 4, b.input_words,
                                               b.virtual_offset, b.n_inputs, None, 0, None, None, None, 0, dop,
                                               eng.stream)
            assert err == 0
        eng.synchronize()
        sb, hb = mems[k]["vram"].download_state(), mems[k]["vram"].download_axons()
        assert np.array_equal(sb, m["vram"].download_state()), f"brain {k}: batched state != solo state"
        assert np.array_equal(hb, m["vram"].download_axons()), f"brain {k}: batched heads != solo heads"
        eng.destroy()
        m["vram"].free()
        if k in sampled:
            st, hd = b.state.copy(), b.heads.copy()
            O.step_batch(st, b.padded_n, hd, b.total_axons, b.lut, 1, 0, ticks, inp, b.virtual_offset, b.n_inputs,
                         dopamine=dop)
            assert np.array_equal(sb, st) and np.array_equal(hb, hd), f"brain {k}: batched != oracle"
            print(f"[c3] brain {k} ({b.padded_n} neurons, {b.spec.slots_filled} slots): equal to the oracle",
                  flush=True)
        mems[k]["vram"].free()


def test_c4_surrogate_data_trials_equal_oracle(gpu):
    from axicor_amd import search as S
    wl = S.Workload(n_neurons=20_000, slots=64, n_inputs=1_000, ticks=2_000, batch_brains=3, n_topologies=2,
                    n_conditions=2, seed=77)
    sim = S.Simulator(wl)
    rng = np.random.default_rng(5)
    params = B.random_params(rng, 3)
    cond, topo = np.array([0, 1, 0]), np.array([0, 1, 1])
    out = sim.run(params, cond, topo)
    assert np.isfinite(out).all()
    for k in range(3):
        b = sim.topos[int(topo[k])]
        lut = B.lut_from_params(params[k])
        bits = sim.conditions[int(cond[k])].bits(wl.n_inputs, wl.ticks, wl.seed + 7 * int(cond[k]))
        st, hd = b.state.copy(), b.heads.copy()
        O.step_batch(st, b.padded_n, hd, b.total_axons, lut, 1, 0, wl.ticks, bits, b.virtual_offset, wl.n_inputs,
                     dopamine=int(params[k][6]))
        assert np.array_equal(sim.pool[k].download_state(), st), f"trial {k}: state != oracle"
        assert np.array_equal(sim.pool[k].download_axons(), hd), f"trial {k}: heads != oracle"
        print(f"[c4] trial {k}: 2,000 ticks equal to the oracle; outcomes {np.round(out[k], 3)}", flush=True)
    sim.free()


def test_c5_full_brain_150_ticks(gpu):
    b = B.generate(B.spec_c5())
    p = Pair(b, n_out=64, max_batch=50)
    for k in range(2):
        out_ref, _ = p.step(50, dopamine=(30 if k else -20), use_graph=True)
        np.testing.assert_array_equal(p.outputs(50), out_ref, err_msg=f"C5 outputs, batch {k}")
        p.check(f"C5 after {p.tick} ticks")
        print(f"[c5] ticks to {p.tick}: state equal", flush=True)
    # 50 more ticks in windows of 4 (the window pass at 12M somas, with the
    # blocks past the LDS capacity spilling slots to the exact path).
    lib = day_lib()
    lib.axc_set_window_ticks(4)
    try:
        out_ref, _ = p.step(50, dopamine=10, use_graph=True)
    finally:
        lib.axc_set_window_ticks(1)
    np.testing.assert_array_equal(p.outputs(50), out_ref, err_msg="C5 outputs, windows of 4")
    p.check(f"C5 after {p.tick} ticks, windows of 4")
    print(f"[c5] ticks to {p.tick} in windows of 4: state equal", flush=True)
