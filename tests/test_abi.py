"""The C-ABI library loads and exports every symbol include/axicor_day.h
declares, and the shared layouts agree (no GPU call is made)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

from axicor_amd import build
from axicor_amd._ffi import SIGNATURES, VARIANT_DTYPE, ShardVramPtrs, day_lib
from axicor_amd.layout import calculate_state_blob_size, compute_state_offsets

HEADER = os.path.join(os.path.dirname(build.REPO + "/"), "include", "axicor_day.h")


def header_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:[A-Za-z_][\w ]*\*?\s*\**)\b([a-z_][a-z0-9_]*)\s*\(", src, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "return", "sizeof")))


def test_header_lists_the_reference_ffi_block():
    """Every extern fn of genesis-compute/src/ffi.rs:50-273 is declared."""
    ffi_rs = ["cu_allocate_shard", "cu_reset_burst_counters", "cu_upload_state_blob", "cu_upload_axons_blob",
              "cu_free_shard", "cu_step_day_phase", "cu_upload_constant_memory", "gpu_malloc", "gpu_free",
              "gpu_host_alloc", "gpu_host_free", "gpu_memcpy_host_to_device_async",
              "gpu_memcpy_device_to_host_async", "gpu_memcpy_host_to_device", "gpu_memcpy_device_to_host",
              "gpu_stream_create", "gpu_stream_destroy", "gpu_stream_synchronize", "gpu_set_device",
              "gpu_device_synchronize", "gpu_synchronize", "gpu_load_constants",
              "update_constant_memory_hot_reload", "launch_sort_and_prune", "launch_extract_outgoing_spikes",
              "launch_ghost_sync", "launch_inject_inputs", "launch_apply_spike_batch", "launch_propagate_axons",
              "launch_update_neurons", "launch_apply_gsop", "launch_record_readout", "gpu_reset_telemetry_count",
              "launch_extract_telemetry", "cu_allocate_io_buffers", "cu_dma_h2d_io", "cu_dma_d2h_io"]
    declared = set(header_functions())
    missing = [n for n in ffi_rs if n not in declared]
    assert not missing, missing


def test_library_exports_every_declared_symbol():
    lib = day_lib()
    for name in header_functions():
        assert hasattr(lib, name), name
    assert set(header_functions()) == set(SIGNATURES), set(header_functions()) ^ set(SIGNATURES)
    out = subprocess.run(["nm", "-D", "--defined-only", build.DAY_LIB], capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for name in header_functions():
        assert name in exported, name


def test_version_string():
    assert b"gfx950" in day_lib().axc_version()


def test_struct_layouts():
    assert VARIANT_DTYPE.itemsize == 64
    assert VARIANT_DTYPE.fields["inertia_curve"][1] == 32
    assert VARIANT_DTYPE.fields["adaptive_leak_max"][1] == 48
    assert VARIANT_DTYPE.fields["d2_affinity"][1] == 59
    assert C.sizeof(ShardVramPtrs) == 9 * 8


def test_state_offsets_reference_contract():
    """memory.rs:71-86 and the 910 B/neuron contract (boot.rs:36-38)."""
    pn, total = calculate_state_blob_size(4000)
    assert pn == 4032 and total == 910 * 4032
    o = compute_state_offsets(64)
    assert (o.soma_voltage, o.soma_flags, o.threshold_offset, o.timers, o.soma_to_axon, o.dendrite_targets,
            o.dendrite_weights, o.dendrite_timers, o.total_bytes) == (0, 256, 320, 576, 640, 896, 33664, 50048, 58240)


def test_host_sort_and_prune_matches_oracle():
    """axc_host_sort_and_prune (host code of the library) against the oracle."""
    from axicor_amd import brain as B
    from oracle import oracle as O
    b = B.generate(B.spec_small(700, 128, 0, seed=13))
    v = b.view()
    rng = np.random.default_rng(5)
    v.weights[:] = rng.integers(-400, 400, size=v.weights.shape)
    v.targets[rng.random(v.targets.shape) < 0.25] = 0
    v.dendrite_timers[:] = rng.integers(0, 255, size=v.dendrite_timers.shape)
    ref = b.state.copy()
    O.sort_and_prune_blob(ref, b.padded_n, 20)
    day_lib().axc_host_sort_and_prune(v.flags.ctypes.data, v.targets.ctypes.data, v.weights.ctypes.data,
                                      v.dendrite_timers.ctypes.data, b.padded_n, 20)
    np.testing.assert_array_equal(b.state, ref)


def test_surrogate_library_exports_every_declared_symbol():
    from axicor_amd._ffi import SURROGATE_SIGNATURES, surrogate_lib
    src = open(os.path.join(build.REPO, "include", "axicor_surrogate.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = set(re.findall(r"\b(axc_mlp_[a-z_]+)\s*\(", src))
    assert names == set(SURROGATE_SIGNATURES), names ^ set(SURROGATE_SIGNATURES)
    lib = surrogate_lib()
    out = subprocess.run(["nm", "-D", "--defined-only", build.SUR_LIB], capture_output=True, text=True).stdout
    exported = set(line.split()[-1] for line in out.splitlines() if line.strip())
    for name in names:
        assert hasattr(lib, name) and name in exported, name


def test_batch_and_schedule_dma_refuse_counts_past_the_slice():
    """Host-side checks, made before any GPU call: a tick's spike count above
    max_spikes_per_tick (or counts without a schedule) would read past the
    schedule buffer, so both entry points refuse it; an all-zero schedule
    copies nothing."""
    import ctypes as C
    import numpy as np
    from axicor_amd._ffi import ShardVramPtrs
    lib = day_lib()
    invalid = 1  # hipErrorInvalidValue
    v = ShardVramPtrs()
    counts = np.array([0, 3, 9, 2], np.uint32)
    fake = C.c_void_p(0x1000)  # never dereferenced: the call is refused first
    rc = lib.axc_step_day_phase_batch(C.byref(v), 64, 64, 1, 0, 4, None, 0, 0, 0, fake, 8,
                                      counts.ctypes.data, None, None, 0, 0, None)
    assert rc == invalid
    rc = lib.axc_step_day_phase_batch(C.byref(v), 64, 64, 1, 0, 4, None, 0, 0, 0, None, 8,
                                      counts.ctypes.data, None, None, 0, 0, None)
    assert rc == invalid
    assert lib.axc_dma_h2d_schedule(fake, fake, 8, counts.ctypes.data, 4, None) == invalid
    zeros = np.zeros(4, np.uint32)
    assert lib.axc_dma_h2d_schedule(None, None, 8, zeros.ctypes.data, 4, None) == 0
