"""The Night Phase's shared-memory contract with genesis-baker-daemon.

A shard and its baker daemon share one file, /dev/shm/genesis_shard_<HASH>
(genesis-core/src/ipc.rs:69-81), laid out as (ipc.rs:112-176):

    64-byte ShmHeader | weights i16[128][padded_n] | targets u32[128][padded_n]
    | handovers AxonHandoverEvent[10,000] | prunes AxonHandoverPrune[10,000]
    | flags u8[padded_n]

The Night goes (genesis-node/src/node/shard_thread.rs:269-378 and
genesis-node/src/ipc.rs:76-150 on the shard's side,
genesis-baker/src/bin/daemon.rs:314-478 on the daemon's):

1. the shard copies the soma flags into the region, runs launch_sort_and_prune
   on the GPU and copies weights and targets in, writes the incoming
   handovers and their count, and sends a 16-byte BakeRequest followed by
   its ghost-origin map (u32 per ghost);
2. the daemon validates the header, runs the sprouting pass on the region's
   arrays in place, writes the outgoing handovers' count and the prunes,
   sets the state to NightDone and answers "BKOK", the ack count and the
   acks;
3. the shard uploads targets and weights and sets the state to Idle.

`runtime_night_begin` / `runtime_night_end` are the shard's side on a
VramState, `daemon_night` is the daemon's side on the native sprouting pass.
They exchange the request and reply as bytes: the socket that carries them
(a Unix socket in the reference) is not part of this package.
"""
from __future__ import annotations

import ctypes as C
import mmap
import os
import struct
from enum import IntEnum

import numpy as np

from ._ffi import day_lib, ptr
from .engine import GpuError, VramState
from .layout import MAX_DENDRITES
from .night import (ACK_DTYPE, HANDOVER_DTYPE, MAX_HANDOVERS_PER_NIGHT, MAX_PRUNES_PER_NIGHT, PRUNE_DTYPE,
                    Geometry, sprout_pass)

SHM_MAGIC = 0x47454E53         # "GENS", ipc.rs:18
SHM_VERSION = 2                # ipc.rs:21
BAKE_MAGIC = 0x42414B45        # "BAKE", ipc.rs:227
BAKE_READY_MAGIC = 0x424B4F4B  # "BKOK", ipc.rs:228
MAX_PRUNES_WRITTEN = 1000      # per-night limit of the prunes written to the region (sprouting.rs:369)


class ShmState(IntEnum):
    """ipc.rs:181-194."""
    IDLE = 0
    NIGHT_START = 1
    SPROUTING = 2
    NIGHT_DONE = 3
    ERROR = 4


# ipc.rs:122-140 (64 bytes, little-endian, no padding).
SHM_HEADER_DTYPE = np.dtype([
    ("magic", "<u4"), ("version", "u1"), ("state", "u1"), ("_pad", "<u2"), ("padded_n", "<u4"),
    ("dendrite_slots", "<u4"), ("weights_offset", "<u4"), ("targets_offset", "<u4"), ("epoch", "<u8"),
    ("total_axons", "<u4"), ("handovers_offset", "<u4"), ("handovers_count", "<u4"), ("zone_hash", "<u4"),
    ("prunes_offset", "<u4"), ("prunes_count", "<u4"), ("incoming_prunes_count", "<u4"), ("flags_offset", "<u4")])
assert SHM_HEADER_DTYPE.itemsize == 64
# ipc.rs:233-242.
BAKE_REQUEST_DTYPE = np.dtype([("magic", "<u4"), ("zone_hash", "<u4"), ("current_tick", "<u4"),
                               ("prune_threshold", "<i2"), ("max_sprouts", "<u2")])
assert BAKE_REQUEST_DTYPE.itemsize == 16


def shm_name(zone_hash: int) -> str:
    """ipc.rs:69-71."""
    return f"genesis_shard_{zone_hash & 0xFFFFFFFF:08X}"


def shm_file_path(zone_hash: int) -> str:
    """ipc.rs:75-81 (Unix)."""
    return os.path.join("/dev/shm", shm_name(zone_hash))


def shm_size(padded_n: int) -> int:
    """ipc.rs:112-119."""
    return (64 + padded_n * MAX_DENDRITES * 2 + padded_n * MAX_DENDRITES * 4
            + MAX_HANDOVERS_PER_NIGHT * HANDOVER_DTYPE.itemsize + MAX_PRUNES_PER_NIGHT * PRUNE_DTYPE.itemsize
            + padded_n)


def new_header(zone_hash: int, padded_n: int, total_axons: int) -> np.ndarray:
    """ShmHeader::new (ipc.rs:143-176): sections in order after the header,
    state Idle, epoch and counts 0."""
    h = np.zeros(1, SHM_HEADER_DTYPE)
    w_off = 64
    t_off = w_off + padded_n * MAX_DENDRITES * 2
    h_off = t_off + padded_n * MAX_DENDRITES * 4
    p_off = h_off + MAX_HANDOVERS_PER_NIGHT * HANDOVER_DTYPE.itemsize
    f_off = p_off + MAX_PRUNES_PER_NIGHT * PRUNE_DTYPE.itemsize
    for k, v in (("magic", SHM_MAGIC), ("version", SHM_VERSION), ("state", ShmState.IDLE), ("padded_n", padded_n),
                 ("dendrite_slots", MAX_DENDRITES), ("weights_offset", w_off), ("targets_offset", t_off),
                 ("total_axons", total_axons), ("handovers_offset", h_off), ("zone_hash", zone_hash & 0xFFFFFFFF),
                 ("prunes_offset", p_off), ("flags_offset", f_off)):
        h[k] = v
    return h


def validate(header) -> str | None:
    """ShmHeader::validate (ipc.rs:178-190): None, or the reference's message."""
    if int(header["magic"]) != SHM_MAGIC:
        return "SHM magic mismatch"
    if int(header["version"]) != SHM_VERSION:
        return "SHM version mismatch"
    if int(header["dendrite_slots"]) != MAX_DENDRITES:
        return "SHM dendrite_slots != 128"
    return None


class ShmRegion:
    """One shard's region over a uint8 buffer (a mapped /dev/shm file or
    memory of this process); the arrays are views into it."""

    def __init__(self, buf, owner=None):
        self.buf = np.frombuffer(buf, dtype=np.uint8) if not isinstance(buf, np.ndarray) else buf
        self._owner = owner  # keeps the mapping alive
        h = self.header
        problem = validate(h)
        if problem:
            raise ValueError(problem)
        n = int(h["padded_n"])
        if self.buf.size < shm_size(n):
            raise ValueError(f"SHM region of {self.buf.size} B is smaller than shm_size({n})")

    @classmethod
    def create(cls, zone_hash: int, padded_n: int, total_axons: int, path: str | None = None) -> "ShmRegion":
        """A new region with ShmHeader::new, in the file `path` (truncated to
        shm_size and mapped) or, with no path, in memory of this process."""
        size = shm_size(padded_n)
        if path is None:
            buf = np.zeros(size, np.uint8)
            buf[:64] = new_header(zone_hash, padded_n, total_axons).view(np.uint8)
            return cls(buf)
        with open(path, "w+b") as f:
            f.truncate(size)
            m = mmap.mmap(f.fileno(), size)
        m[:64] = new_header(zone_hash, padded_n, total_axons).tobytes()
        return cls(np.frombuffer(m, dtype=np.uint8), owner=m)

    @classmethod
    def open(cls, path: str) -> "ShmRegion":
        f = open(path, "r+b")
        m = mmap.mmap(f.fileno(), 0)
        f.close()
        return cls(np.frombuffer(m, dtype=np.uint8), owner=m)

    @property
    def header(self) -> np.ndarray:
        return self.buf[:64].view(SHM_HEADER_DTYPE)[0]

    def _view(self, off_field: str, dtype, count: int) -> np.ndarray:
        off = int(self.header[off_field])
        return self.buf[off:off + count * np.dtype(dtype).itemsize].view(dtype)

    @property
    def padded_n(self) -> int:
        return int(self.header["padded_n"])

    @property
    def weights(self) -> np.ndarray:
        n = self.padded_n
        return self._view("weights_offset", np.int16, MAX_DENDRITES * n).reshape(MAX_DENDRITES, n)

    @property
    def targets(self) -> np.ndarray:
        n = self.padded_n
        return self._view("targets_offset", np.uint32, MAX_DENDRITES * n).reshape(MAX_DENDRITES, n)

    @property
    def handovers(self) -> np.ndarray:
        return self._view("handovers_offset", HANDOVER_DTYPE, MAX_HANDOVERS_PER_NIGHT)

    @property
    def prunes(self) -> np.ndarray:
        return self._view("prunes_offset", PRUNE_DTYPE, MAX_PRUNES_PER_NIGHT)

    @property
    def flags(self) -> np.ndarray:
        return self._view("flags_offset", np.uint8, self.padded_n)

    def set_state(self, s: ShmState) -> None:
        self.buf[5] = int(s)  # byte 5, as set_state writes it (genesis-node/src/ipc.rs:168-172)


def bake_request(zone_hash: int, prune_threshold: int, max_sprouts: int, ghost_origins=None,
                 current_tick: int = 0) -> bytes:
    """The 16-byte BakeRequest and the ghost-origin map that follows it
    (genesis-node/src/ipc.rs:101-122)."""
    r = np.zeros(1, BAKE_REQUEST_DTYPE)
    r["magic"], r["zone_hash"], r["current_tick"] = BAKE_MAGIC, zone_hash & 0xFFFFFFFF, current_tick
    r["prune_threshold"], r["max_sprouts"] = prune_threshold, max_sprouts
    go = np.zeros(0, np.uint32) if ghost_origins is None else np.ascontiguousarray(ghost_origins, np.uint32)
    return r.tobytes() + go.tobytes()


def bake_reply(acks: np.ndarray) -> bytes:
    """BKOK, the ack count and the acks (daemon.rs:462-476)."""
    a = np.ascontiguousarray(acks, ACK_DTYPE)
    return struct.pack("<II", BAKE_READY_MAGIC, len(a)) + a.tobytes()


def parse_reply(data: bytes) -> np.ndarray:
    """The acks of a reply (genesis-node/src/ipc.rs:124-148); raises on an
    error magic or a short reply."""
    if len(data) < 8:
        raise ValueError("short bake reply")
    magic, n = struct.unpack_from("<II", data, 0)
    if magic != BAKE_READY_MAGIC:
        raise ValueError(f"Baker daemon returned error magic: {magic:08X}")
    need = 8 + n * ACK_DTYPE.itemsize
    if len(data) < need:
        raise ValueError("short bake reply")
    return np.frombuffer(data[8:need], ACK_DTYPE).copy()


# ---- The shard's side ------------------------------------------------------
def runtime_night_begin(vram: VramState, region: ShmRegion, prune_threshold: int, max_sprouts: int = 4,
                        incoming: np.ndarray | None = None, ghost_origins=None, current_tick: int = 0) -> bytes:
    """Steps 0-3 of shard_thread.rs:269-378 and the first half of run_night:
    flags into the region before the prune, launch_sort_and_prune, weights and
    targets into the region, the incoming handovers and their count; returns
    the request bytes for the daemon."""
    lib = day_lib()
    n = vram.padded_n
    if region.padded_n != n:
        raise ValueError("SHM region and shard sizes differ")
    inc = np.zeros(0, HANDOVER_DTYPE) if incoming is None else np.ascontiguousarray(incoming, HANDOVER_DTYPE)
    if len(inc) > MAX_HANDOVERS_PER_NIGHT:
        raise ValueError(f"Too many handovers: {len(inc)} > {MAX_HANDOVERS_PER_NIGHT}")
    lib.gpu_device_synchronize()
    flags = np.empty(n, np.uint8)
    if not lib.gpu_memcpy_device_to_host(ptr(flags), vram.ptrs.soma_flags, n):
        raise GpuError("flags download failed")
    region.flags[:] = flags
    lib.launch_sort_and_prune(C.byref(vram.ptrs), n, prune_threshold)
    lib.gpu_device_synchronize()
    w = np.empty((MAX_DENDRITES, n), np.int16)
    t = np.empty((MAX_DENDRITES, n), np.uint32)
    for dst, src in ((w, vram.ptrs.dendrite_weights), (t, vram.ptrs.dendrite_targets)):
        if not lib.gpu_memcpy_device_to_host(ptr(dst), src, dst.nbytes):
            raise GpuError("night download failed")
    region.weights[:] = w
    region.targets[:] = t
    region.handovers[:len(inc)] = inc
    region.header["handovers_count"] = len(inc)
    return bake_request(int(region.header["zone_hash"]), prune_threshold, max_sprouts, ghost_origins, current_tick)


def runtime_night_end(vram: VramState, region: ShmRegion, reply: bytes) -> np.ndarray:
    """The rest of run_night and of the Night: check the reply, upload targets
    and weights from the region, state Idle; returns the acks. On an error
    reply the state goes to Idle and the shard keeps its pruned matrices."""
    try:
        acks = parse_reply(reply)
    except ValueError:
        region.set_state(ShmState.IDLE)
        raise
    lib = day_lib()
    t = np.ascontiguousarray(region.targets)
    w = np.ascontiguousarray(region.weights)
    for src, dst in ((t, vram.ptrs.dendrite_targets), (w, vram.ptrs.dendrite_weights)):
        if not lib.gpu_memcpy_host_to_device(dst, ptr(src), src.nbytes):
            raise GpuError("night upload failed")
    lib.gpu_device_synchronize()
    region.set_state(ShmState.IDLE)
    return acks


# ---- The daemon's side -----------------------------------------------------
def daemon_night(region: ShmRegion, request: bytes, geo: Geometry, soma_to_axon: np.ndarray, types,
                 master_seed: int) -> bytes:
    """run_night_phase of daemon.rs:314-478 on a region: parse the request and
    its ghost-origin map (geo.total_ghosts entries), validate the header, run
    the sprouting pass in place on the region's targets, weights and flags
    with the incoming handovers, write the outgoing handovers' count and the
    prunes (at most 1,000), set NightDone; returns the reply bytes."""
    if len(request) < 16:
        raise ValueError("short bake request")
    req = np.frombuffer(request[:16], BAKE_REQUEST_DTYPE)[0]
    if int(req["magic"]) != BAKE_MAGIC:
        raise ValueError(f"Invalid BAKE magic: {int(req['magic']):08X}")
    ng = geo.total_ghosts
    if len(request) < 16 + 4 * ng:
        raise ValueError("short ghost-origin map")
    origins = np.frombuffer(request[16:16 + 4 * ng], np.uint32).copy()
    h = region.header
    problem = validate(h)
    if problem:
        raise ValueError(f"SHM validation failed: {problem}")
    n = region.padded_n
    n_in = int(h["handovers_count"])
    incoming = region.handovers[:n_in].copy()
    g = Geometry(geo.lengths, geo.paths, geo.tips, geo.dirs, geo.soma_positions, ng,
                 origins if ng else None, geo.max_x, geo.max_y)
    t, w = region.targets, region.weights  # views: the pass writes the region in place
    flags = np.ascontiguousarray(region.flags)
    res = sprout_pass(t, w, flags, np.ascontiguousarray(soma_to_axon, np.uint32), n, g, types, int(h["epoch"]),
                      master_seed, zone_hash=int(req["zone_hash"]), max_sprouts=int(req["max_sprouts"]),
                      incoming=incoming)
    region.handovers[:len(res.handovers)] = res.handovers
    region.header["handovers_count"] = len(res.handovers)
    if len(res.prunes):
        k = min(len(res.prunes), MAX_PRUNES_WRITTEN)
        region.prunes[:k] = res.prunes[:k]
        region.header["prunes_count"] = k
    region.set_state(ShmState.NIGHT_DONE)
    return bake_reply(res.acks)
