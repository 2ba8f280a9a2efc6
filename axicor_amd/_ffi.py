"""ctypes bindings of ``include/axicor_day.h`` (libaxicor_day.so).

The library is the product path: there is no fallback. If it is missing or
fails to load, :func:`day_lib` raises with the build command to run.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

# ---------------------------------------------------------------------------
# Binary contracts (genesis-core/src/layout.rs:10-74; ffi.rs:25-42)
# ---------------------------------------------------------------------------
MAX_DENDRITES = 128
AXON_SENTINEL = 0x80000000
NO_AXON = 0xFFFFFFFF
STATE_BYTES_PER_NEURON = 910

# VariantParameters, 64 B (layout.rs:10-48): numpy view with the ABI offsets.
VARIANT_DTYPE = np.dtype({
    "names": ["threshold", "rest_potential", "leak_rate", "homeostasis_penalty",
              "spontaneous_firing_period_ticks", "initial_synapse_weight", "gsop_potentiation",
              "gsop_depression", "homeostasis_decay", "refractory_period",
              "synapse_refractory_period", "signal_propagation_length", "is_inhibitory",
              "inertia_curve", "adaptive_leak_max", "adaptive_leak_gain", "adaptive_mode",
              "_leak_pad", "d1_affinity", "d2_affinity", "_pad"],
    "formats": ["<i4", "<i4", "<i4", "<i4", "<u4", "<u2", "<u2", "<u2", "<u2", "u1", "u1", "u1", "u1",
                ("u1", 16), "<i4", "<u2", "u1", ("u1", 3), "u1", "u1", ("u1", 4)],
    "offsets": [0, 4, 8, 12, 16, 20, 22, 24, 26, 28, 29, 30, 31, 32, 48, 52, 54, 55, 58, 59, 60],
    "itemsize": 64,
})
assert VARIANT_DTYPE.itemsize == 64


class ShardVramPtrs(C.Structure):
    """ffi.rs:25-42 — nine device pointers in `.state` order."""
    _fields_ = [
        ("soma_voltage", C.c_void_p),
        ("soma_flags", C.c_void_p),
        ("threshold_offset", C.c_void_p),
        ("timers", C.c_void_p),
        ("soma_to_axon", C.c_void_p),
        ("dendrite_targets", C.c_void_p),
        ("dendrite_weights", C.c_void_p),
        ("dendrite_timers", C.c_void_p),
        ("axon_heads", C.c_void_p),
    ]


class AxcShardStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "ticks", "integrated", "slots_scanned", "targets_read", "hits", "spikes", "gsop_slots", "gsop_targets",
        "gsop_slots_fused", "gsop_targets_fused")]


class AxcBrainDesc(C.Structure):
    """One member of a batch (include/axicor_day.h)."""
    _fields_ = [("vram", ShardVramPtrs), ("padded_n", C.c_uint32), ("total_axons", C.c_uint32),
                ("v_seg", C.c_uint32), ("virtual_offset", C.c_uint32), ("num_virtual_axons", C.c_uint32),
                ("lut", C.c_void_p), ("d_input_bitmask", C.c_void_p), ("input_words", C.c_uint32),
                ("input_ticks", C.c_uint32), ("dopamine", C.c_int16)]


class AxcBatchTuning(C.Structure):
    _fields_ = [("axon_blocks", C.c_uint32), ("soma_blocks", C.c_uint32), ("list_blocks", C.c_uint32),
                ("stage_order", C.c_uint32)]


class AxcSynapseStats(C.Structure):
    _fields_ = [("filled", C.c_uint64), ("sum_abs_weight", C.c_uint64), ("sum_sq_weight", C.c_uint64),
                ("below_prune", C.c_uint64), ("sum_threshold_offset", C.c_int64)]


_P = C.c_void_p
_U32 = C.c_uint32
_I32 = C.c_int32
_I16 = C.c_int16
_SZ = C.c_size_t
_VP = C.POINTER(ShardVramPtrs)

# name: (restype, argtypes) — every symbol of include/axicor_day.h.
SIGNATURES = {
    "cu_allocate_shard": (_I32, [_U32, _U32, _VP]),
    "cu_upload_state_blob": (_I32, [_VP, _P, _SZ]),
    "cu_upload_axons_blob": (_I32, [_VP, _P, _SZ]),
    "cu_free_shard": (None, [_VP]),
    "cu_reset_burst_counters": (None, [_VP, _U32, _P]),
    "cu_step_day_phase": (_I32, [_VP, _U32, _U32, _U32, _U32, _P, _U32, _U32, _P, _U32, _P, _P, _U32, _I16, _P]),
    "cu_upload_constant_memory": (_I32, [_P]),
    "gpu_malloc": (_P, [_SZ]),
    "gpu_free": (None, [_P]),
    "gpu_host_alloc": (_P, [_SZ]),
    "gpu_host_free": (None, [_P]),
    "gpu_memcpy_host_to_device_async": (None, [_P, _P, _SZ, _P]),
    "gpu_memcpy_device_to_host_async": (None, [_P, _P, _SZ, _P]),
    "gpu_memcpy_host_to_device": (C.c_bool, [_P, _P, _SZ]),
    "gpu_memcpy_device_to_host": (C.c_bool, [_P, _P, _SZ]),
    "gpu_stream_create": (_I32, [C.POINTER(_P)]),
    "gpu_stream_destroy": (_I32, [_P]),
    "gpu_stream_synchronize": (None, [_P]),
    "gpu_set_device": (None, [_I32]),
    "gpu_device_synchronize": (None, []),
    "gpu_synchronize": (None, []),
    "gpu_load_constants": (None, [_P]),
    "update_constant_memory_hot_reload": (None, [_P, _P]),
    "launch_sort_and_prune": (None, [_VP, _U32, _I16]),
    "launch_extract_outgoing_spikes": (None, [_P, _P, _P, _U32, _U32, _U32, _P, _P, _P]),
    "launch_ghost_sync": (None, [_P, _P, _P, _P, _U32, _P]),
    "gpu_reset_telemetry_count": (None, [_VP, _P]),
    "launch_extract_telemetry": (None, [_VP, _U32, _P, _P, _P]),
    "cu_allocate_io_buffers": (_I32, [_U32, _U32, _U32, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)]),
    "cu_dma_h2d_io": (_I32, [_P, _P, _U32, _P, _P, _U32, _P]),
    "cu_dma_d2h_io": (_I32, [_P, _P, _U32, _P]),
    "launch_inject_inputs": (None, None),
    "launch_apply_spike_batch": (None, None),
    "launch_propagate_axons": (None, None),
    "launch_update_neurons": (None, None),
    "launch_apply_gsop": (None, None),
    "launch_record_readout": (None, None),
    "axc_shard_stats_reset": (_I32, [_VP, _P]),
    "axc_shard_stats_read": (_I32, [_VP, C.POINTER(AxcShardStats)]),
    "axc_step_day_phase_batch": (_I32, [_VP, _U32, _U32, _U32, _U32, _U32, _P, _U32, _U32, _U32, _P, _U32, _P, _P,
                                        _P, _U32, _I16, _P]),
    "axc_host_sort_and_prune": (None, [_P, _P, _P, _P, _U32, _I16]),
    "axc_kernel_timing_enable": (None, [C.c_int]),
    "axc_kernel_timing_read": (_I32, [_P, _P, C.c_int]),
    "axc_batch_create": (_I32, [C.POINTER(AxcBrainDesc), _U32, C.POINTER(_P)]),
    "axc_batch_step": (_I32, [_P, _U32, _U32, _P]),
    "axc_batch_autotune": (_I32, [_P, _U32, _P, C.POINTER(AxcBatchTuning)]),
    "axc_batch_set_tuning": (_I32, [_P, C.POINTER(AxcBatchTuning)]),
    "axc_batch_get_tuning": (_I32, [_P, C.POINTER(AxcBatchTuning)]),
    "axc_batch_stats": (_I32, [_P, _U32, C.POINTER(AxcShardStats), _P]),
    "axc_batch_synapse_stats": (_I32, [_P, _P, _P, _P]),
    "axc_batch_update": (_I32, [_P, C.POINTER(AxcBrainDesc), _U32]),
    "axc_batch_stats_all": (_I32, [_P, C.POINTER(AxcShardStats), _P]),
    "axc_batch_destroy": (None, [_P]),
    "axc_version": (C.c_char_p, []),
    "axc_staging_enable": (None, [C.c_int]),
    "axc_set_window_ticks": (_I32, [_U32]),
    "axc_window_ticks": (_U32, []),
    "axc_shard_window_stats": (_I32, [_VP, C.POINTER(C.c_uint64)]),
    "axc_shard_targets_changed": (None, [C.POINTER(ShardVramPtrs)]),
    "axc_shard_write_slots": (C.c_int32, [C.POINTER(ShardVramPtrs), _U32, _P, _P, _P, _U32, _P]),
    "axc_shard_graph_captures": (_I32, [_VP, C.POINTER(C.c_uint64)]),
    "axc_shard_set_soma_positions": (_I32, [_VP, _P, _U32]),
    "axc_dma_h2d_schedule": (_I32, [_P, _P, _U32, _P, _U32, _P]),
    "axc_step_day_phase_noalloc": (_I32, [_VP, _U32, _U32, _U32, _U32, _P, _U32, _U32, _P, _U32, _P, _P, _U32, _I16,
                                          _P]),
}

TIMED_KERNELS = ["axon_tick", "neuron_tick", "push_apply", "gsop", "readout", "spike_count", "weights_copy",
                 "window_start", "window_pass", "window_correct", "window_soma", "window_end"]

SURROGATE_SIGNATURES = {
    "axc_mlp_create": (_I32, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "axc_mlp_forward": (_I32, [_P, _P, C.c_int64, _P, _P]),
    "axc_mlp_set_fused_forward": (_I32, [_P, C.c_int]),
    "axc_mlp_train_step": (_I32, [_P, _P, _P, C.c_int, C.c_float, C.POINTER(C.c_float), _P]),
    "axc_mlp_gradients": (_I32, [_P, _P, _P, C.c_int, _P, C.POINTER(C.c_float), _P]),
    "axc_mlp_layout": (_I32, [_P, C.POINTER(C.c_int32), _P, _P, C.POINTER(C.c_int64)]),
    "axc_mlp_get_params": (_I32, [_P, _P]),
    "axc_mlp_set_params": (_I32, [_P, _P]),
    "axc_mlp_reset_optimizer": (_I32, [_P]),
    "axc_mlp_gemm": (_I32, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P, C.c_int, _P, C.c_int, _P, C.c_int,
                            _P]),
    "axc_mlp_destroy": (None, [_P]),
}

_day = None
_gen = None
_sur = None


def _load(path: str, what: str) -> C.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(f"{what} is not built ({path}); run `python -m axicor_amd.build`")
    return C.CDLL(path)


def day_lib() -> C.CDLL:
    """The Day-Phase engine library, with argtypes set for every symbol."""
    global _day
    if _day is None:
        lib = _load(_build.DAY_LIB, "libaxicor_day.so")
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            if args is not None:
                fn.argtypes = args
        _day = lib
    return _day


def surrogate_lib() -> C.CDLL:
    """The surrogate library (include/axicor_surrogate.h), argtypes set."""
    global _sur
    if _sur is None:
        lib = _load(_build.SUR_LIB, "libaxicor_surrogate.so")
        for name, (res, args) in SURROGATE_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _sur = lib
    return _sur


class GenSpec(C.Structure):
    """Mirror of GenSpec in csrc/brain_gen.cpp."""
    _fields_ = [
        ("n_neurons", C.c_uint32), ("n_inputs", C.c_uint32),
        ("gx", C.c_uint32), ("gy", C.c_uint32), ("gz", C.c_uint32),
        ("slots_filled", C.c_uint32), ("axon_len_min", C.c_uint32), ("axon_len_max", C.c_uint32),
        ("w_min", C.c_int32), ("w_max", C.c_int32), ("v_seg", C.c_uint32), ("n_types", C.c_uint32),
        ("type_ppm_cdf", C.c_uint32 * 16), ("mean_dist_milli", C.c_uint32), ("master_seed", C.c_uint64),
        ("brick", C.c_uint32),
    ]


def gen_lib() -> C.CDLL:
    global _gen
    if _gen is None:
        lib = _load(_build.GEN_LIB, "libaxicor_host.so")
        lib.axc_gen_sizes.argtypes = [C.POINTER(GenSpec), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.axc_gen_sizes.restype = None
        lib.axc_gen_brain.argtypes = [C.POINTER(GenSpec), _P, _P, _P, _P, _P]
        lib.axc_gen_brain.restype = C.c_int
        lib.axc_gen_inputs.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]
        lib.axc_gen_inputs.restype = None
        lib.axc_gen_entity_seed.argtypes = [C.c_uint64, C.c_uint32]
        lib.axc_gen_entity_seed.restype = C.c_uint64
        lib.axc_gen_paths.argtypes = [C.POINTER(GenSpec), _P, _P, _P, _P, _P]
        lib.axc_gen_paths.restype = C.c_int
        lib.axc_sprout_pass.argtypes = [_P, _P, _P, _P, _P, _U32, _U32, _P, _P, _P, _U32, _U32, _U32, _U32, _U32,
                                        _P, _U32, C.c_uint64, _P, _P, _P, C.c_uint64, _U32, C.c_uint16, _P, _P,
                                        _U32, _P, _P, _U32]
        lib.axc_sprout_pass.restype = C.c_int
        _gen = lib
    return _gen


def ptr(a: np.ndarray | None) -> int | None:
    """Host address of a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data
