"""ctypes wrapper of the CPU oracle (oracle/day_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker. The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "liborc.so")


class OrcStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "ticks", "integrated", "slots_scanned", "targets_read", "hits", "spikes", "gsop_slots", "gsop_targets")]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n, _ in self._fields_}


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, U32, I16 = C.c_void_p, C.c_uint32, C.c_int16
        L.orc_step_tick.restype = U32
        L.orc_step_tick.argtypes = [P, U32, P, U32, P, U32, U32, P, U32, U32, P, U32, P, P, U32, I16,
                                    C.POINTER(OrcStats), P]
        L.orc_step_batch.restype = None
        L.orc_step_batch.argtypes = [P, U32, P, U32, P, U32, U32, U32, P, U32, U32, U32, P, U32, P, P, P, U32, I16,
                                     C.POINTER(OrcStats)]
        L.orc_sort_and_prune.restype = None
        L.orc_sort_and_prune.argtypes = [P, P, P, P, U32, I16]
        L.orc_sort_and_prune_blob.restype = None
        L.orc_sort_and_prune_blob.argtypes = [P, U32, I16]
        L.orc_gsop_update.restype = C.c_int16
        L.orc_gsop_update.argtypes = [C.c_int16, P, U32, P, C.c_int32, I16]
        L.orc_active_tail.restype = C.c_int
        L.orc_active_tail.argtypes = [U32, U32, U32]
        L.orc_seed_from_str.restype = C.c_uint64
        L.orc_seed_from_str.argtypes = [C.c_char_p, C.c_size_t]
        L.orc_entity_seed.restype = C.c_uint64
        L.orc_entity_seed.argtypes = [C.c_uint64, U32]
        L.orc_fnv1a_32.restype = U32
        L.orc_fnv1a_32.argtypes = [C.c_char_p, C.c_size_t]
        L.orc_pack_target.restype = U32
        L.orc_pack_target.argtypes = [U32, U32]
        L.orc_unpack_axon.restype = U32
        L.orc_unpack_axon.argtypes = [U32]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def step_batch(state: np.ndarray, padded_n: int, heads: np.ndarray, total_axons: int, lut: np.ndarray, v_seg: int,
               tick_base: int, T: int, bitmask: np.ndarray | None, virt_off: int, n_virt: int,
               spikes: np.ndarray | None = None, max_spikes_per_tick: int = 0, spike_counts=None,
               mapped: np.ndarray | None = None, n_out: int = 0, dopamine: int = 0):
    """Runs T ticks in place on (state, heads). Returns (outputs[T, n_out] or None, stats dict)."""
    assert state.flags["C_CONTIGUOUS"] and heads.flags["C_CONTIGUOUS"]
    words = bitmask.shape[1] if bitmask is not None and bitmask.ndim == 2 else 0
    bm = np.ascontiguousarray(bitmask, dtype=np.uint32) if bitmask is not None else None
    sp = np.ascontiguousarray(spikes, dtype=np.uint32) if spikes is not None else None
    sc = np.ascontiguousarray(spike_counts, dtype=np.uint32) if spike_counts is not None else None
    mp = np.ascontiguousarray(mapped, dtype=np.uint32) if mapped is not None else None
    out = np.zeros((T, n_out), dtype=np.uint8) if (mp is not None and n_out) else None
    st = OrcStats()
    lib().orc_step_batch(_p(state), padded_n, _p(heads), total_axons, _p(np.ascontiguousarray(lut)), v_seg,
                         tick_base, T, _p(bm), words, virt_off, n_virt if bm is not None else 0, _p(sp),
                         max_spikes_per_tick, _p(sc), _p(mp), _p(out), n_out, dopamine, C.byref(st))
    return out, st.as_dict()


def sort_and_prune_blob(state: np.ndarray, padded_n: int, prune: int) -> None:
    lib().orc_sort_and_prune_blob(_p(state), padded_n, prune)


def gsop_update(w: int, heads8, seg: int, variant: np.ndarray, burst_mult: int, dopamine: int) -> int:
    h = np.ascontiguousarray(heads8, dtype=np.uint32)
    v = np.ascontiguousarray(variant)
    return int(lib().orc_gsop_update(w, _p(h), seg, _p(v), burst_mult, dopamine))
