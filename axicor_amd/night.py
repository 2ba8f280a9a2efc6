"""Night Phase: GPU sort & prune plus the host sprouting pass.

`sprout_pass` binds axc_sprout_pass (csrc/night_host.cpp, a restatement of
genesis-baker/src/bake/sprouting.rs:119-378). `execute_night_phase` is the
sequence of genesis-node/src/node/shard_thread.rs:269-378: download the flags
before pruning, launch_sort_and_prune, download weights and targets, sprout on
the host, upload the new synapses' slots of targets and weights (timers,
voltages and heads stay)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._ffi import day_lib, gen_lib, ptr
from .engine import DeviceBuffer, GpuError, PinnedBuffer, VramState
from .layout import MAX_DENDRITES

HANDOVER_DTYPE = np.dtype([("origin_zone_hash", "<u4"), ("local_axon_id", "<u4"), ("entry_x", "<u2"),
                           ("entry_y", "<u2"), ("vector_x", "i1"), ("vector_y", "i1"), ("vector_z", "i1"),
                           ("type_mask", "u1"), ("remaining_length", "<u2"), ("entry_z", "u1"),
                           ("_padding", "u1")])
assert HANDOVER_DTYPE.itemsize == 20  # genesis-core/src/ipc.rs:45-48
ACK_DTYPE = np.dtype([("target_zone_hash", "<u4"), ("src_axon_id", "<u4"), ("dst_ghost_id", "<u4")])
PRUNE_DTYPE = np.dtype([("target_zone_hash", "<u4"), ("dst_ghost_id", "<u4")])
SPROUT_TYPE_DTYPE = np.dtype([("w_distance", "<f4"), ("w_power", "<f4"), ("w_explore", "<f4"), ("w_type", "<f4"),
                              ("initial_synapse_weight", "<u2"), ("is_inhibitory", "u1"), ("_pad", "u1")])
MAX_HANDOVERS_PER_NIGHT = 10_000  # ipc.rs:23
MAX_PRUNES_PER_NIGHT = 10_000     # ipc.rs:24


def sprout_types_from_lut(lut: np.ndarray, w_distance=0.5, w_power=0.4, w_explore=0.1, w_type=0.0) -> np.ndarray:
    """Per-type sprouting preferences (template values, _template/blueprints.toml:60-63)
    with each type's initial weight and sign from its variant."""
    t = np.zeros(16, SPROUT_TYPE_DTYPE)
    t["w_distance"], t["w_power"], t["w_explore"], t["w_type"] = w_distance, w_power, w_explore, w_type
    t["initial_synapse_weight"] = lut["initial_synapse_weight"]
    t["is_inhibitory"] = lut["is_inhibitory"]
    return t


@dataclass
class Geometry:
    """Night-Phase host data of one shard (paths/tips/dirs/lengths per axon)."""
    lengths: np.ndarray       # uint8[A]
    paths: np.ndarray         # uint32[A, 256]
    tips: np.ndarray          # uint32[A]
    dirs: np.ndarray          # uint32[A]
    soma_positions: np.ndarray  # uint32[N]
    total_ghosts: int = 0
    ghost_origins: np.ndarray | None = None
    max_x: int = 1024
    max_y: int = 1024


@dataclass
class SproutResult:
    new_synapses: int
    handovers: np.ndarray = field(default_factory=lambda: np.zeros(0, HANDOVER_DTYPE))
    acks: np.ndarray = field(default_factory=lambda: np.zeros(0, ACK_DTYPE))
    prunes: np.ndarray = field(default_factory=lambda: np.zeros(0, PRUNE_DTYPE))
    changed_cols: np.ndarray | None = None  # columns (slot * padded_n + soma) of the new synapses, if asked for


class _Res(C.Structure):
    _fields_ = [("new_synapses", C.c_uint32), ("handovers_out", C.c_uint32), ("acks", C.c_uint32),
                ("prunes", C.c_uint32)]


def sprout_pass(targets: np.ndarray, weights: np.ndarray, flags: np.ndarray, soma_to_axon: np.ndarray,
                padded_n: int, geo: Geometry, types: np.ndarray | None, epoch: int, master_seed: int,
                zone_hash: int = 0, max_sprouts: int = 4,
                incoming: np.ndarray | None = None, changed: np.ndarray | None = None) -> SproutResult:
    """run_sprouting_pass in place on host arrays (targets/weights columnar
    [128, padded_n], geometry arrays updated). With `changed` (uint32, room
    for every new synapse: padded_n * max(1, max_sprouts) always suffices),
    the result's changed_cols lists the columns the pass filled."""
    for a in (targets, weights, flags, soma_to_axon, geo.lengths, geo.paths, geo.tips, geo.dirs, geo.soma_positions):
        assert a.flags["C_CONTIGUOUS"]
    A = geo.tips.shape[0]
    n_in = 0 if incoming is None else len(incoming)
    hand = np.zeros(max(MAX_HANDOVERS_PER_NIGHT, n_in), HANDOVER_DTYPE)
    if n_in:
        hand[:n_in] = incoming
    acks = np.zeros(max(n_in, 1), ACK_DTYPE)
    prunes = np.zeros(MAX_PRUNES_PER_NIGHT, PRUNE_DTYPE)
    go = geo.ghost_origins if geo.ghost_origins is not None else np.zeros(max(geo.total_ghosts, 1), np.uint32)
    ty = None if types is None else np.ascontiguousarray(types, SPROUT_TYPE_DTYPE)
    r = _Res()
    rc = gen_lib().axc_sprout_pass(ptr(targets), ptr(weights), ptr(flags), ptr(go), ptr(hand), n_in,
                                   MAX_HANDOVERS_PER_NIGHT, ptr(geo.tips), ptr(geo.dirs), ptr(soma_to_axon),
                                   padded_n, geo.total_ghosts, A, geo.max_x, geo.max_y, ptr(ty),
                                   0 if ty is None else len(ty), epoch & 0xFFFFFFFFFFFFFFFF, ptr(geo.lengths),
                                   ptr(geo.paths), ptr(geo.soma_positions), master_seed & 0xFFFFFFFFFFFFFFFF,
                                   zone_hash, max_sprouts, ptr(acks), ptr(prunes), MAX_PRUNES_PER_NIGHT, C.byref(r),
                                   ptr(changed), 0 if changed is None else len(changed))
    if rc != 0:
        raise ValueError("axc_sprout_pass rejected its arguments")
    cols = None
    if changed is not None:
        if r.new_synapses > len(changed):
            raise ValueError(f"`changed` holds {len(changed)} columns, the pass made {r.new_synapses} synapses")
        cols = changed[:r.new_synapses]
    return SproutResult(r.new_synapses, hand[:r.handovers_out].copy(), acks[:r.acks].copy(), prunes[:r.prunes].copy(),
                        cols)


class NightStaging:
    """Page-locked host copies of one shard's flags, soma-to-axon map,
    weights and targets, for the Night's downloads and upload: allocated once
    per shard and reused every Night (pinning 1.2 GB at 1.5M neurons costs
    more than a Night's copies, which page-locked memory runs at the link's
    full rate)."""

    def __init__(self, padded_n: int):
        n = int(padded_n)
        self.padded_n = n
        self._bufs = [PinnedBuffer(n, np.uint8), PinnedBuffer(n, np.uint32),
                      PinnedBuffer(MAX_DENDRITES * n, np.int16), PinnedBuffer(MAX_DENDRITES * n, np.uint32)]
        self.flags = self._bufs[0].array
        self.soma_to_axon = self._bufs[1].array
        self.weights = self._bufs[2].array.reshape(MAX_DENDRITES, n)
        self.targets = self._bufs[3].array.reshape(MAX_DENDRITES, n)
        # The host maps page-locked pages on their first CPU access: touch
        # them here, not in the first Night's sprouting pass.
        for b in self._bufs:
            b.array.fill(0)

    def changed_buffer(self, max_sprouts: int) -> np.ndarray:
        """Room for the columns of every synapse a pass can make
        (padded_n * max(1, max_sprouts)), kept for the next Nights."""
        need = self.padded_n * max(1, int(max_sprouts))
        if getattr(self, "_changed", None) is None or len(self._changed) < need:
            self._changed = np.empty(need, np.uint32)
        return self._changed[:need]

    def upload_slots(self, vram: VramState, cols: np.ndarray) -> None:
        """The host values of the listed slots (targets and weights) written
        to the shard's reference rows (axc_shard_write_slots): the upload of a
        Night whose sprouting changed only those slots."""
        n = len(cols)
        if n == 0:
            return
        if getattr(self, "_d_slots", None) is None or self._d_slots[0] < n:
            cap = max(n, 1 << 16)
            self._d_slots = (cap, DeviceBuffer(4 * cap), DeviceBuffer(4 * cap), DeviceBuffer(2 * cap))
        _, d_cols, d_t, d_w = self._d_slots
        d_cols.upload(np.ascontiguousarray(cols, np.uint32))
        d_t.upload(self.targets.reshape(-1)[cols])
        d_w.upload(self.weights.reshape(-1)[cols])
        err = day_lib().axc_shard_write_slots(C.byref(vram.ptrs), self.padded_n, d_cols.ptr, d_t.ptr, d_w.ptr, n,
                                              None)
        if err:
            raise GpuError(f"axc_shard_write_slots failed ({err})")

    @staticmethod
    def of(vram: VramState) -> "NightStaging":
        """The shard's staging buffers, made on its first Night."""
        st = getattr(vram, "_night_staging", None)
        if st is None or st.padded_n != vram.padded_n:
            st = NightStaging(vram.padded_n)
            vram._night_staging = st
        return st


def execute_night_phase(vram: VramState, prune_threshold: int, geo: Geometry | None, types: np.ndarray | None,
                        epoch: int, master_seed: int, max_sprouts: int = 4) -> SproutResult | None:
    """shard_thread.rs:269-378 on one shard; returns the sprouting result (None
    when no geometry is given: sort & prune only, as when no daemon is up).
    The host copies live in the shard's NightStaging; the upload writes only
    the slots the sprouting filled."""
    lib = day_lib()
    N = vram.padded_n
    st = NightStaging.of(vram)
    lib.gpu_device_synchronize()
    if not lib.gpu_memcpy_device_to_host(ptr(st.flags), vram.ptrs.soma_flags, N):
        raise GpuError("flags download failed")
    lib.launch_sort_and_prune(C.byref(vram.ptrs), N, prune_threshold)
    lib.gpu_device_synchronize()
    if geo is None:
        return None
    for dst, src in ((st.weights, vram.ptrs.dendrite_weights), (st.targets, vram.ptrs.dendrite_targets),
                     (st.soma_to_axon, vram.ptrs.soma_to_axon)):
        if not lib.gpu_memcpy_device_to_host(ptr(dst), src, dst.nbytes):
            raise GpuError("night download failed")
    res = sprout_pass(st.targets, st.weights, st.flags, st.soma_to_axon, N, geo, types, epoch, master_seed,
                      max_sprouts=max_sprouts, changed=st.changed_buffer(max_sprouts))
    # The device already holds the pruned rows; only the new synapses' slots
    # differ (the reference uploads both whole arrays, shard_thread.rs:339-350).
    st.upload_slots(vram, res.changed_cols)
    lib.gpu_device_synchronize()
    return res


def geometry_of(brain) -> Geometry:
    """Night geometry of a generated brain (generate(..., with_paths=True))."""
    assert brain.paths is not None, "generate the brain with with_paths=True"
    return Geometry(brain.axon_len.copy(), brain.paths.copy(), brain.tips.copy(), brain.dirs.copy(),
                    brain.soma_pos.copy(), 0, None, brain.spec.grid[0], brain.spec.grid[1])
