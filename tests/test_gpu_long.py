"""Long runs of the full 1.5M-neuron brain on the GPU:

- (skipped unless AXC_LONG=1: two minutes of oracle time; the default suite
  runs this brain against the oracle for 500 ticks and a Night in
  tests/test_gpu_configs.py) the brain for 2,000 ticks through the batch graph with
  a Night every 500 ticks as shard_thread.rs:269-378 runs it (the flags
  taken before the GPU sort & prune at 15, the host sprouting, the upload):
  every soma's spike bit every tick, and the whole state and the heads
  before and after every Night and at the end, bit for bit;
- the same brain at its configured length, 10,000 ticks with a Night every
  1,000 (shard_thread.rs:667-680), twice on the GPU: in windows of 4 ticks
  and with the per-tick pass, the two runs' states and heads equal before
  and after every Night and at the end (the oracle covers the first 2,000
  ticks above; 10,000 ticks of it would take over a quarter of an hour).
  Part of the default suite (about 40 s).

Run: AXC_LONG=1 python -u -m pytest tests/test_gpu_long.py -s --timeout 1100"""
import os

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd import night as NT
from axicor_amd._ffi import day_lib
from axicor_amd.engine import DeviceBuffer, IoDeviceBuffers, ShardEngine, VramState, upload_constant_memory
from axicor_amd.layout import StateView
from helpers import Pair
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(os.environ.get("AXC_LONG") != "1", reason="long run: set AXC_LONG=1")
def test_c2_2000_ticks_with_a_night_every_500(gpu):
    b = B.generate(B.spec_c2(), with_paths=True)
    N = b.padded_n
    p = Pair(b, n_out=N, max_batch=50)
    p.mapped = np.arange(N, dtype=np.uint32)  # every soma's spike bit, every tick
    p.d_mapped = DeviceBuffer.from_array(p.mapped)
    types = NT.sprout_types_from_lut(b.lut)
    geo_gpu, geo_cpu = NT.geometry_of(b), NT.geometry_of(b)
    for k in range(40):
        out_ref, _ = p.step(50, dopamine=(40 if k % 2 else -25), use_graph=True)
        np.testing.assert_array_equal(p.outputs(50), out_ref, err_msg=f"C2 raster, batch {k}")
        print(f"[c2 long] ticks to {p.tick}: raster equal, {int(out_ref.sum())} spikes", flush=True)
        if p.tick % 500 == 0 and p.tick < 2000:
            p.check(f"C2 before the Night at tick {p.tick}")
            epoch = p.tick // 500 - 1
            res = NT.execute_night_phase(p.vram, 15, geo_gpu, types, epoch=epoch, master_seed=b.spec.seed)
            v = StateView(p.ref_state, N)
            flags = v.flags.copy()
            O.sort_and_prune_blob(p.ref_state, N, 15)
            t, w = v.targets.copy(), v.weights.copy()
            ref = NT.sprout_pass(t, w, flags, v.soma_to_axon.copy(), N, geo_cpu, types, epoch, b.spec.seed)
            v.targets[:] = t
            v.weights[:] = w
            assert res.new_synapses == ref.new_synapses
            p.check(f"C2 after the Night at tick {p.tick}")
            print(f"[c2 long] Night at {p.tick}: {ref.new_synapses} new synapses, state equal", flush=True)
    p.check("C2 after 2,000 ticks")


def test_c2_10000_ticks_windows_equal_per_tick(gpu):
    b = B.generate(B.spec_c2(), with_paths=True)
    N, A = b.padded_n, b.total_axons
    lib = day_lib()
    upload_constant_memory(b.lut)
    types = NT.sprout_types_from_lut(b.lut)
    runs = []
    for win in (4, 1):
        v = VramState(N, A)
        v.upload_state(b.state)
        v.upload_axon_heads(b.heads)
        runs.append({"win": win, "vram": v, "eng": ShardEngine(v), "geo": NT.geometry_of(b),
                     "io": IoDeviceBuffers(b.input_words, 0, 0, 50)})

    def same(where):
        for r in runs:
            r["eng"].synchronize()
        a, c = runs[0]["vram"], runs[1]["vram"]
        assert np.array_equal(a.download_state(), c.download_state()), where
        assert np.array_equal(a.download_axons(), c.download_axons()), where

    tick = 0
    try:
        while tick < 10_000:
            inputs = b.inputs(tick, 50)
            for r in runs:
                lib.axc_set_window_ticks(r["win"])
                r["eng"].step_day_phase_batch(50, r["io"], inputs, None, np.zeros(50, np.uint32), b.virtual_offset,
                                              b.n_inputs, None, b.spec.v_seg, 40 if (tick // 50) % 2 else -25, tick,
                                              use_graph=True)
            tick += 50
            if tick % 1000 == 0:
                same(f"before the Night at tick {tick}")
                epoch = tick // 1000 - 1
                news = [NT.execute_night_phase(r["vram"], 15, r["geo"], types, epoch=epoch,
                                               master_seed=b.spec.seed).new_synapses for r in runs]
                assert news[0] == news[1]
                same(f"after the Night at tick {tick}")
                print(f"[c2 10k] tick {tick}: windows of 4 and per tick equal, Night {news[0]} new synapses",
                      flush=True)
    finally:
        lib.axc_set_window_ticks(1)
    same("at tick 10,000")
