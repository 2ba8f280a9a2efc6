"""Batched brains: many independent brains stepped together (axc_batch_* of
include/axicor_day.h). Each brain keeps its own VramState and parameter
table; results are bit-identical to stepping each brain alone."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import AxcBatchTuning, AxcBrainDesc, AxcShardStats, AxcSynapseStats, VARIANT_DTYPE, day_lib, ptr
from .engine import DeviceBuffer, GpuError, VramState


def _sig():
    return day_lib()


class BatchEngine:
    """Owns nothing but the batch handle: the brains' VramStates, tables and
    input buffers stay with the caller and must outlive the batch."""

    @staticmethod
    def _descs(members: list[dict]):
        luts = [np.ascontiguousarray(m["lut"], dtype=VARIANT_DTYPE) for m in members]
        arr = (AxcBrainDesc * len(members))()
        for k, m in enumerate(members):
            d = arr[k]
            d.vram = m["vram"].ptrs
            d.padded_n = m["vram"].padded_n
            d.total_axons = m["vram"].total_axons
            d.v_seg = m.get("v_seg", 1)
            d.virtual_offset = m["virtual_offset"]
            d.num_virtual_axons = m["n_inputs"]
            d.lut = ptr(luts[k])
            d.d_input_bitmask = m["d_inputs"].ptr if m.get("d_inputs") is not None else None
            d.input_words = m.get("input_words", 0)
            d.input_ticks = m.get("input_ticks", 1)
            d.dopamine = m.get("dopamine", 0)
        return arr, luts

    def __init__(self, members: list[dict]):
        """members: dicts with keys vram (VramState), lut (VARIANT_DTYPE[16]),
        d_inputs (DeviceBuffer or None), input_words, input_ticks, v_seg,
        virtual_offset, n_inputs, dopamine."""
        lib = _sig()
        arr, self._luts = self._descs(members)
        h = C.c_void_p()
        err = lib.axc_batch_create(arr, len(members), C.byref(h))
        if err:
            raise GpuError(f"axc_batch_create failed ({err})")
        self.h = h.value
        self.n = len(members)

    def step(self, tick_base: int, ticks: int, stream=None) -> None:
        err = day_lib().axc_batch_step(self.h, tick_base, ticks, stream)
        if err:
            raise GpuError(f"axc_batch_step failed ({err})")

    def autotune(self, ticks: int, stream=None) -> dict:
        t = AxcBatchTuning()
        err = day_lib().axc_batch_autotune(self.h, ticks, stream, C.byref(t))
        if err:
            raise GpuError(f"axc_batch_autotune failed ({err})")
        return {"axon_blocks": t.axon_blocks, "soma_blocks": t.soma_blocks, "list_blocks": t.list_blocks,
                "stage_order": t.stage_order}

    def set_tuning(self, axon_blocks: int, soma_blocks: int, list_blocks: int, stage_order: int = 0) -> None:
        t = AxcBatchTuning(axon_blocks, soma_blocks, list_blocks, stage_order)
        err = day_lib().axc_batch_set_tuning(self.h, C.byref(t))
        if err:
            raise GpuError(f"axc_batch_set_tuning failed ({err})")

    def stats(self, k: int) -> tuple[dict, np.ndarray]:
        s = AxcShardStats()
        ts = np.zeros(16, np.uint64)
        err = day_lib().axc_batch_stats(self.h, k, C.byref(s), ptr(ts))
        if err:
            raise GpuError(f"axc_batch_stats failed ({err})")
        return {n: getattr(s, n) for n, _ in AxcShardStats._fields_}, ts

    def update(self, members: list[dict]) -> None:
        """New tables, dopamine and inputs for the same brains (same VramStates
        in the same order); counters restart at zero (axc_batch_update)."""
        arr, self._luts = self._descs(members)
        err = day_lib().axc_batch_update(self.h, arr, len(members))
        if err:
            raise GpuError(f"axc_batch_update failed ({err})")

    def stats_all(self) -> tuple[np.ndarray, np.ndarray]:
        """Counters of every brain: ([n] structured AxcShardStats fields as a
        dict of arrays, [n][16] spikes per variant type)."""
        arr = (AxcShardStats * self.n)()
        ts = np.zeros((self.n, 16), np.uint64)
        err = day_lib().axc_batch_stats_all(self.h, arr, ptr(ts))
        if err:
            raise GpuError(f"axc_batch_stats_all failed ({err})")
        st = {n: np.array([getattr(a, n) for a in arr], np.uint64) for n, _ in AxcShardStats._fields_}
        return st, ts

    def synapse_stats(self, prune_thresholds=None, stream=None) -> np.ndarray:
        """[n][5] int64: filled slots, sum |w|, sum w^2, slots with |w| below
        the brain's prune threshold, sum of threshold offsets."""
        arr = (AxcSynapseStats * self.n)()
        pr = None if prune_thresholds is None else np.ascontiguousarray(prune_thresholds, np.int16)
        err = day_lib().axc_batch_synapse_stats(self.h, ptr(pr), arr, stream)
        if err:
            raise GpuError(f"axc_batch_synapse_stats failed ({err})")
        return np.array([[a.filled, a.sum_abs_weight, a.sum_sq_weight, a.below_prune, a.sum_threshold_offset]
                         for a in arr], dtype=np.int64)

    def destroy(self) -> None:
        if self.h:
            day_lib().axc_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
