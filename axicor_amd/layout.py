"""`.state` / `.axons` layout helpers — genesis-compute/src/memory.rs:14-86.

The `.state` blob is a headerless concatenation of eight SoA arrays
(memory.rs:19-29). Slot ``s`` of neuron ``i`` in a columnar dendrite array is at
index ``s * padded_n + i`` (genesis-core/src/layout.rs:236-241).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

MAX_DENDRITES = 128  # memory.rs:15


def align_to_warp(n: int) -> int:
    """genesis-core/src/layout.rs:77-79: round up to 64."""
    return (n + 63) & ~63


def calculate_state_blob_size(neuron_count: int) -> tuple[int, int]:
    """memory.rs:32-54: (padded_n, total .state bytes) = (align64(n), 910*padded_n)."""
    padded_n = align_to_warp(neuron_count)
    return padded_n, padded_n * (4 + 1 + 4 + 1 + 4) + padded_n * MAX_DENDRITES * (4 + 2 + 1)


@dataclass(frozen=True)
class StateOffsets:
    """memory.rs:58-69: byte offset of each array inside the `.state` blob."""
    soma_voltage: int
    soma_flags: int
    threshold_offset: int
    timers: int
    soma_to_axon: int
    dendrite_targets: int
    dendrite_weights: int
    dendrite_timers: int
    total_bytes: int


def compute_state_offsets(padded_n: int) -> StateOffsets:
    """memory.rs:71-86."""
    off = 0
    out = []
    for size in (4, 1, 4, 1, 4, 4 * MAX_DENDRITES, 2 * MAX_DENDRITES, 1 * MAX_DENDRITES):
        out.append(off)
        off += padded_n * size
    return StateOffsets(*out, total_bytes=off)


class StateView:
    """Typed numpy views into a `.state` blob (no copies)."""

    def __init__(self, blob: np.ndarray, padded_n: int):
        assert blob.dtype == np.uint8 and blob.size == calculate_state_blob_size(padded_n)[1]
        o = compute_state_offsets(padded_n)
        n = padded_n
        self.padded_n = n
        self.blob = blob
        self.voltage = blob[o.soma_voltage:o.soma_flags].view(np.int32)
        self.flags = blob[o.soma_flags:o.threshold_offset]
        self.threshold_offset = blob[o.threshold_offset:o.timers].view(np.int32)
        self.timers = blob[o.timers:o.soma_to_axon]
        self.soma_to_axon = blob[o.soma_to_axon:o.dendrite_targets].view(np.uint32)
        self.targets = blob[o.dendrite_targets:o.dendrite_weights].view(np.uint32).reshape(MAX_DENDRITES, n)
        self.weights = blob[o.dendrite_weights:o.dendrite_timers].view(np.int16).reshape(MAX_DENDRITES, n)
        self.dendrite_timers = blob[o.dendrite_timers:o.total_bytes].reshape(MAX_DENDRITES, n)


def pack_dendrite_target(axon_id: int, segment_offset: int) -> int:
    """genesis-core/src/layout.rs:254-265 (raises where the reference panics)."""
    if axon_id >= 0x00FFFFFF:
        raise ValueError("Axon ID exceeds 24 bits")
    if segment_offset >= 256:
        raise ValueError("Segment offset exceeds 8 bits")
    return (segment_offset << 24) | ((axon_id + 1) & 0x00FFFFFF)


def unpack_axon_id(t: int) -> int:
    """layout.rs:268-271 (saturating -1)."""
    a = t & 0x00FFFFFF
    return a - 1 if a else 0


def unpack_segment_offset(t: int) -> int:
    """layout.rs:274-277."""
    return t >> 24
