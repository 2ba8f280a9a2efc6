"""The scale point's brain structure (12M neurons x 128 slots of six types,
axons of 1-255 segments, 2% of the neurons' count as inputs at p = 0.01) on
its 200k-neuron sub-shard from the same seed, against the CPU oracle at the
Night boundaries: day batches through the batch graph, a Night (flags before
the GPU sort & prune, host sprouting, upload; shard_thread.rs:269-378), then
more day batches. At this density 256-soma blocks hold up to ~9,300
distinct presynaptic axons, so the staged scan's spill path runs in every
tick. The full 12M-neuron brain is measured by `bench.py --workload c5`.

Also the full 1.5M-neuron C2 brain: the spike raster of every soma compared
with the oracle's every tick for the first 100 ticks, the whole state at the
batch boundaries and after a Night's sort & prune."""
import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd import night as NT
from axicor_amd.engine import DeviceBuffer
from axicor_amd.layout import StateView
from helpers import Pair
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_c5_subshard_bit_exact_at_night_boundaries(gpu):
    b = B.generate(B.spec_c5(200_000), with_paths=True)
    assert b.spec.slots_filled == 128 and b.spec.grid == (64, 64, 32)
    p = Pair(b, n_out=64)
    types = NT.sprout_types_from_lut(b.lut)
    geo_gpu, geo_cpu = NT.geometry_of(b), NT.geometry_of(b)
    for k in range(3):
        out_ref, _ = p.step(50, dopamine=(60 if k % 2 else -40), use_graph=True)
        np.testing.assert_array_equal(p.outputs(50), out_ref)
    p.check("c5 sub-shard before the night")
    res = NT.execute_night_phase(p.vram, 15, geo_gpu, types, epoch=0, master_seed=b.spec.seed)
    v = StateView(p.ref_state, b.padded_n)
    flags = v.flags.copy()
    O.sort_and_prune_blob(p.ref_state, b.padded_n, 15)
    t, w = v.targets.copy(), v.weights.copy()
    ref = NT.sprout_pass(t, w, flags, v.soma_to_axon.copy(), b.padded_n, geo_cpu, types, 0, b.spec.seed)
    v.targets[:] = t
    v.weights[:] = w
    assert res.new_synapses == ref.new_synapses > 0
    p.check("c5 sub-shard after the night")
    out_ref, _ = p.step(50, dopamine=30, use_graph=True)
    np.testing.assert_array_equal(p.outputs(50), out_ref)
    p.check("c5 sub-shard day after the night")


def test_c2_every_tick_for_100_ticks_and_a_night(gpu):
    b = B.generate(B.spec_c2())
    N = b.padded_n
    p = Pair(b, n_out=N, max_batch=50)
    p.mapped = np.arange(N, dtype=np.uint32)  # every soma's spike bit, every tick
    p.d_mapped = DeviceBuffer.from_array(p.mapped)
    for k in range(2):
        out_ref, _ = p.step(50, dopamine=(40 if k else -25), use_graph=True)
        got = p.outputs(50)
        for t in range(50):
            np.testing.assert_array_equal(got[t], out_ref[t], err_msg=f"C2 raster tick {50 * k + t}")
        p.check(f"C2 tick {p.tick}")
    assert 0.002 < out_ref.mean() < 0.05  # the brain is firing (about 1% per tick)
    p.night(15)
    p.check("C2 after the Night's sort & prune")
    p.step(10, use_graph=True)
    p.check("C2 day after the Night")
