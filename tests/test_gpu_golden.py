"""The engine on the C1 golden run (tests/golden/c1_golden.py): 1,000 ticks
through the batch graph with every soma's spike bit read out each tick, and
the three Nights as the node runs them (execute_night_phase: the flags taken
before the GPU sort & prune, the host sprouting, the upload;
shard_thread.rs:269-378). Every tick's raster digest and every state digest
must equal the committed ones, which the CPU oracle wrote; nothing here runs
the oracle."""
import os
import sys

import numpy as np
import pytest

from axicor_amd import night as NT
from axicor_amd._ffi import day_lib
from axicor_amd.engine import DeviceBuffer, IoDeviceBuffers, ShardEngine, VramState, upload_constant_memory

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

import c1_golden as G  # noqa: E402

pytestmark = pytest.mark.gpu


def test_c1_engine_matches_the_golden_digests(gpu):
    want = G.load()
    b = G.brain()
    N = b.padded_n
    upload_constant_memory(b.lut)
    vram = VramState(N, b.total_axons)
    vram.upload_state(b.state)
    vram.upload_axon_heads(b.heads)
    eng = ShardEngine(vram)
    io = IoDeviceBuffers(b.input_words, 0, N, G.BATCH)
    mapped = DeviceBuffer.from_array(np.arange(N, dtype=np.uint32))
    geo, types = NT.geometry_of(b), NT.sprout_types_from_lut(b.lut)
    out = np.empty((G.BATCH, N), np.uint8)
    states = iter(want["states"])

    def digests():
        eng.synchronize()
        return G.state_digests(vram.download_state(), vram.download_axons(), N)

    assert next(states) == ["start", 0, digests()]
    for k in range(G.TICKS // G.BATCH):
        t0 = k * G.BATCH
        eng.step_day_phase_batch(G.BATCH, io, b.inputs(t0, G.BATCH), None, np.zeros(G.BATCH, np.uint32),
                                 b.virtual_offset, b.n_inputs, mapped.ptr, G.V_SEG, G.dopamine_of(k), t0,
                                 use_graph=True)
        eng.synchronize()
        assert day_lib().gpu_memcpy_device_to_host(out.ctypes.data, io.d_output_history, out.nbytes)
        for t in range(G.BATCH):
            assert int(out[t].sum()) == want["spikes"][t0 + t], f"spike count, tick {t0 + t}"
            assert G.digest(out[t]) == want["raster"][t0 + t], f"raster, tick {t0 + t}"
        tick = t0 + G.BATCH
        if tick % G.NIGHT_EVERY == 0 and tick < G.TICKS:
            assert next(states) == ["before_night", tick, digests()]
            res = NT.execute_night_phase(vram, G.PRUNE, geo, types, epoch=tick // G.NIGHT_EVERY - 1,
                                         master_seed=b.spec.seed, max_sprouts=G.MAX_SPROUTS)
            assert next(states) == ["after_night", tick, digests(), res.new_synapses]
    assert next(states) == ["end", G.TICKS, digests()]
