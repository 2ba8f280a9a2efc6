"""axc_step_day_phase_noalloc (the allocation-free tick) against
cu_step_day_phase through the C-ABI: two brains in a row, the second smaller
than the first, so the second call reuses scratch that still holds the first
brain's workspace. Outputs every tick and all eight state arrays plus the
heads must be bit-identical."""
import ctypes as C

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd._ffi import day_lib
from axicor_amd.engine import DeviceBuffer, ShardEngine, VramState, upload_constant_memory
from helpers import compare_state

pytestmark = pytest.mark.gpu


def run_ticks(entry, b, T, seed):
    """T ticks of brain b through `entry`, one call per tick, with inputs, a
    spike schedule (duplicates and out-of-range ids included), mapped outputs
    and dopamine. Returns (outputs [T, n_out], state, heads)."""
    lib = day_lib()
    vram = VramState(b.padded_n, b.total_axons)
    vram.upload_state(b.state)
    vram.upload_axon_heads(b.heads)
    eng = ShardEngine(vram)
    rng = np.random.default_rng(seed)
    inputs = b.inputs(0, T)
    n_out, max_spk = 40, 24
    mapped = rng.integers(0, b.padded_n, size=n_out).astype(np.uint32)
    mapped[::9] = 0xFFFFFFFF
    sched = rng.integers(0, b.total_axons + 8, size=(T, max_spk)).astype(np.uint32)
    sched[:, 1] = sched[:, 2]
    counts = rng.integers(0, max_spk + 1, size=T).astype(np.uint32)
    d_in = DeviceBuffer.from_array(np.ascontiguousarray(inputs))
    d_sp = DeviceBuffer.from_array(sched)
    d_map = DeviceBuffer.from_array(mapped)
    d_out = DeviceBuffer(T * n_out)
    words = inputs.shape[1]
    for t in range(T):
        rc = entry(C.byref(vram.ptrs), b.padded_n, b.total_axons, 1, t, d_in.ptr + t * words * 4, b.virtual_offset,
                   b.n_inputs, d_sp.ptr + t * max_spk * 4, int(counts[t]), d_map.ptr, d_out.ptr + t * n_out, n_out,
                   (-40, 0, 75)[t % 3], eng.stream)
        assert rc == 0, (t, rc)
    eng.synchronize()
    lib.gpu_device_synchronize()
    out = d_out.download(np.uint8, T * n_out).reshape(T, n_out)
    return out, vram.download_state(), vram.download_axons()


@pytest.mark.parametrize("order", ["noalloc_first", "reference_first"])
def test_noalloc_equals_cu_step_day_phase_two_calls_second_smaller(gpu, order):
    lib = day_lib()
    brains = [B.generate(B.spec_small(6000, 48, 300, seed=31, axon_len=(1, 60))),
              B.generate(B.spec_small(1500, 32, 100, seed=32))]
    for k, b in enumerate(brains):
        upload_constant_memory(b.lut)
        print(f"[no-alloc] brain {k}: N={b.padded_n} A={b.total_axons} ({order})", flush=True)
        runs = {}
        names = ("noalloc", "reference") if order == "noalloc_first" else ("reference", "noalloc")
        for name in names:
            entry = lib.axc_step_day_phase_noalloc if name == "noalloc" else lib.cu_step_day_phase
            runs[name] = run_ticks(entry, b, 30, seed=100 + k)
        (o1, s1, h1), (o2, s2, h2) = runs["noalloc"], runs["reference"]
        assert o2.sum() > 0
        np.testing.assert_array_equal(o1, o2, err_msg=f"outputs, brain {k}")
        compare_state(s1, h1, s2, h2, b.padded_n, f"no-alloc vs cu_step_day_phase, brain {k}")


def test_noalloc_refuses_a_shard_past_its_scratch(gpu):
    """A shard whose workspace does not fit the 16 MiB scratch is refused
    before any launch."""
    lib = day_lib()
    v = VramState(1024, 1024)
    assert lib.axc_step_day_phase_noalloc(C.byref(v.ptrs), 2_000_000, 2_000_000, 1, 0, None, 0, 0, None, 0, None,
                                          None, 0, 0, None) == 1
