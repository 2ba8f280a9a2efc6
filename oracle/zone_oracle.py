"""CPU restatement of the multi-zone helpers: K11 ghost sync, K12 outgoing
spike extraction, and the receiver's per-tick spike schedule.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker. The product
never imports it.

- ghost_sync: genesis-compute/src/cuda/bindings.cu:341-359.
- extract_outgoing_spikes: bindings.cu:378-418. The device appends events
  in atomic order, so callers compare them as multisets.
- schedule_from_events: SpikeSchedule::push_spike,
  genesis-node/src/network/ring_buffer.rs:35-54.

Pinned by the hand-derived known answers in tests/test_oracle_kats.py (head
thresholds, integer division by v_seg, poisoned indices, dropped offsets,
the per-tick cap); the reference holds no test for these functions.
"""
from __future__ import annotations

import numpy as np

POISON = 0x80000000          # src index >= this: sentinel or "soma without axon" (bindings.cu:350-356)
DEAD_HEAD = 0x70000000       # heads >= this are dead tails (bindings.cu:404)
MAX_SPIKES_PER_TICK = 100_000  # ring_buffer.rs:6


def ghost_sync(src_heads: np.ndarray, dst_heads: np.ndarray, src_idx, dst_idx) -> None:
    """dst[dst_idx[j]] = src[src_idx[j]] (all eight heads) for every j whose
    source index is below 0x80000000 (bindings.cu:341-359), in place."""
    s = src_heads.reshape(-1, 8)
    d = dst_heads.reshape(-1, 8)
    for a, b in zip(np.asarray(src_idx, np.uint64).tolist(), np.asarray(dst_idx, np.uint64).tolist()):
        if a < POISON:
            d[b] = s[a]


def extract_outgoing_spikes(heads: np.ndarray, src_idx, ghost_ids, T: int, v_seg: int) -> list[tuple[int, int]]:
    """Events (ghost_id, tick_offset) of bindings.cu:378-418: for entry j with
    src_idx[j] < 0x80000000, each of the axon's heads h < 0x70000000 with
    t = h // v_seg < T gives (ghost_ids[j], T - 1 - t); in (j, head) order."""
    h = heads.reshape(-1, 8)
    out = []
    for a, g in zip(np.asarray(src_idx, np.uint64).tolist(), np.asarray(ghost_ids, np.uint64).tolist()):
        if a >= POISON:
            continue
        for k in range(8):
            x = int(h[a, k])
            if x >= DEAD_HEAD:
                continue
            t = x // v_seg
            if t < T:
                out.append((int(g), T - 1 - t))
    return out


def schedule_from_events(events, T: int, max_spikes: int = MAX_SPIKES_PER_TICK) -> tuple[np.ndarray, np.ndarray]:
    """SpikeSchedule::push_spike for each event in order (ring_buffer.rs:35-54):
    offsets >= T are dropped, each tick holds at most max_spikes ids. Returns
    (ids uint32[T, max_spikes], counts uint32[T])."""
    ids = np.zeros((T, max_spikes), np.uint32)
    counts = np.zeros(T, np.uint32)
    for g, off in events:
        if off >= T:
            continue
        c = int(counts[off])
        if c < max_spikes:
            ids[off, c] = g
            counts[off] = c + 1
    return ids, counts
