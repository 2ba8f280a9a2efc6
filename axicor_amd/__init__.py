"""axicor_amd — MI355X-native Day-Phase engine for Axicor (integer spiking networks).

The hot path is HIP for gfx950 behind the C-ABI of include/axicor_day.h
(libaxicor_day.so); this package holds the host-side mirror of genesis-compute's
Rust API (engine.py), the `.state`/`.axons` layout helpers (layout.py) and the
seeded synthetic brain generator (brain.py).
"""
from .layout import (MAX_DENDRITES, StateOffsets, StateView, align_to_warp, calculate_state_blob_size,
                     compute_state_offsets, pack_dendrite_target, unpack_axon_id, unpack_segment_offset)

__all__ = ["MAX_DENDRITES", "StateOffsets", "StateView", "align_to_warp", "calculate_state_blob_size",
           "compute_state_offsets", "pack_dendrite_target", "unpack_axon_id", "unpack_segment_offset"]
