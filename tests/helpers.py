"""Shared parity machinery: run a brain on the engine (through the C-ABI) and
on the CPU oracle with the same inputs, and compare every byte."""
from __future__ import annotations

import ctypes as C

import numpy as np

from axicor_amd import brain as B
from axicor_amd.engine import DeviceBuffer, IoDeviceBuffers, ShardEngine, VramState, upload_constant_memory
from axicor_amd.layout import StateView
from axicor_amd._ffi import day_lib
from oracle import oracle as O

FIELDS = ["voltage", "flags", "threshold_offset", "timers", "soma_to_axon", "targets", "weights",
          "dendrite_timers"]


def first_diff(name: str, got: np.ndarray, want: np.ndarray, where: str) -> str | None:
    if got.shape != want.shape:
        return f"{where}: {name} shape {got.shape} != {want.shape}"
    neq = np.nonzero(got.reshape(-1) != want.reshape(-1))[0]
    if neq.size == 0:
        return None
    k = int(neq[0])
    return (f"{where}: {name} differs at {neq.size} positions, first flat index {k}: "
            f"got {got.reshape(-1)[k]} want {want.reshape(-1)[k]}")


def compare_state(got_state, got_heads, want_state, want_heads, padded_n, where):
    gv, wv = StateView(got_state, padded_n), StateView(want_state, padded_n)
    for f in FIELDS:
        msg = first_diff(f, getattr(gv, f), getattr(wv, f), where)
        assert msg is None, msg
    msg = first_diff("axon_heads", got_heads.reshape(-1, 8), want_heads.reshape(-1, 8), where)
    assert msg is None, msg


class Pair:
    """A brain loaded on the engine next to its oracle copy."""

    def __init__(self, brain: B.Brain, n_out: int = 0, max_batch: int = 64, max_spikes: int = 0, seed: int = 1,
                 positions: np.ndarray | None = None):
        """positions: soma positions for the engine's own order (None: the
        brain's numbering)."""
        self.b = brain
        self.ref_state = brain.state.copy()
        self.ref_heads = brain.heads.copy()
        upload_constant_memory(brain.lut)
        self.vram = VramState(brain.padded_n, brain.total_axons)
        self.vram.upload_state(brain.state)
        self.vram.upload_axon_heads(brain.heads)
        if positions is not None:
            self.vram.set_soma_positions(positions)
        self.eng = ShardEngine(self.vram)
        self.max_batch = max_batch
        self.max_spikes = max_spikes
        self.io = IoDeviceBuffers(brain.input_words, max_spikes, n_out, max_batch)
        self.n_out = n_out
        rng = np.random.default_rng(seed)
        if n_out:
            ids = rng.integers(0, brain.padded_n, size=n_out).astype(np.uint32)
            ids[::7] = 0xFFFFFFFF          # empty pixels
            ids[3::11] = brain.padded_n + 5  # out of range reads 0
            self.mapped = ids
            self.d_mapped = DeviceBuffer.from_array(ids)
        else:
            self.mapped = None
            self.d_mapped = None
        self.tick = 0

    def step(self, T: int, dopamine: int = 0, spikes: np.ndarray | None = None, counts=None, v_seg: int = 1,
             use_graph: bool = False, inputs: np.ndarray | None = None):
        b = self.b
        if inputs is None:
            inputs = b.inputs(self.tick, T) if b.n_inputs else None
        cnts = np.zeros(T, np.uint32) if counts is None else np.asarray(counts, np.uint32)
        self.eng.step_day_phase_batch(T, self.io, inputs, spikes, cnts, b.virtual_offset, b.n_inputs,
                                      self.d_mapped.ptr if self.d_mapped else None, v_seg, dopamine, self.tick,
                                      use_graph=use_graph)
        out_ref, st_ref = O.step_batch(self.ref_state, b.padded_n, self.ref_heads, b.total_axons, b.lut, v_seg,
                                       self.tick, T, inputs, b.virtual_offset, b.n_inputs, spikes=spikes,
                                       max_spikes_per_tick=self.max_spikes, spike_counts=cnts,
                                       mapped=self.mapped, n_out=self.n_out, dopamine=dopamine)
        self.tick += T
        self.eng.synchronize()
        return out_ref, st_ref

    def outputs(self, T: int) -> np.ndarray:
        buf = np.empty(T * self.n_out, np.uint8)
        if self.n_out:
            assert day_lib().gpu_memcpy_device_to_host(buf.ctypes.data, self.io.d_output_history, buf.nbytes)
        return buf.reshape(T, self.n_out)

    def check(self, where: str) -> None:
        self.eng.synchronize()
        compare_state(self.vram.download_state(), self.vram.download_axons(), self.ref_state, self.ref_heads,
                      self.b.padded_n, where)

    def night(self, prune: int) -> None:
        day_lib().launch_sort_and_prune(C.byref(self.vram.ptrs), self.b.padded_n, prune)
        day_lib().gpu_device_synchronize()
        O.sort_and_prune_blob(self.ref_state, self.b.padded_n, prune)
