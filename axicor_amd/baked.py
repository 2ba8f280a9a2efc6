"""Baked shard files and hot checkpoints, in the reference's byte layouts.

A baked shard directory holds (genesis-baker/src/bake/layout.rs:127-208,
genesis-baker/src/bin/daemon.rs:200-312):

- shard.state: the .state blob, 910 B per neuron (layout.rs:127-159);
- shard.axons: raw BurstHeads8[total_axons] (layout.rs:160-168);
- shard.paths: PathsFileHeader {magic "PATH", version 1, total_axons,
  max_segments 256}, then u8 lengths[total_axons], zero padding to a
  multiple of 64 bytes, then u32 segments[total_axons][256]
  (genesis-core/src/layout.rs:157-191, layout.rs:170-208);
- shard.geom: u32 tips[total_axons] then u32 dirs[total_axons] (daemon.rs:278-281, 393-401);
- shard.pos: u32 packed soma positions[padded_n] (daemon.rs:289-291).

The daemon maps paths and geom read-write (sprouting nudges axons and
grows paths in place) and pos read-only. `save_checkpoint` is the hot
checkpoint of genesis-node/src/node/shard_thread.rs:229-263: both device
blobs downloaded, written to .tmp files and renamed over checkpoint.state
and checkpoint.axons.
"""
from __future__ import annotations

import os

import numpy as np

from .engine import VramState
from .night import Geometry

PATHS_MAGIC = 0x50415448  # "PATH", genesis-core/src/layout.rs:157
PATHS_VERSION = 1         # layout.rs:185 (baker)
MAX_SEGMENTS = 256        # genesis-core/src/layout.rs:158
PATHS_HEADER_DTYPE = np.dtype([("magic", "<u4"), ("version", "<u4"), ("total_axons", "<u4"),
                               ("max_segments", "<u4")])


def paths_matrix_offset(total_axons: int) -> int:
    """calculate_paths_matrix_offset (genesis-core/src/layout.rs:186-191)."""
    return 16 + total_axons + (64 - (16 + total_axons) % 64) % 64


def paths_file_size(total_axons: int) -> int:
    """calculate_paths_file_size (genesis-core/src/layout.rs:173-183)."""
    return paths_matrix_offset(total_axons) + total_axons * MAX_SEGMENTS * 4


def write_paths(path: str, lengths: np.ndarray, paths: np.ndarray) -> None:
    """write_paths_blob (genesis-baker/src/bake/layout.rs:170-208)."""
    A = len(lengths)
    assert paths.shape == (A, MAX_SEGMENTS)
    blob = np.zeros(paths_file_size(A), np.uint8)
    h = np.zeros(1, PATHS_HEADER_DTYPE)
    h["magic"], h["version"], h["total_axons"], h["max_segments"] = PATHS_MAGIC, PATHS_VERSION, A, MAX_SEGMENTS
    blob[:16] = h.view(np.uint8)
    blob[16:16 + A] = lengths
    off = paths_matrix_offset(A)
    blob[off:] = np.ascontiguousarray(paths, np.uint32).view(np.uint8).reshape(-1)
    _write_atomic(path, blob)


def read_paths(path: str, writable: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """(lengths u8[A], segments u32[A, 256]) of a shard.paths file, as views
    of one memory map (read-write when `writable`, like the daemon's)."""
    m = np.memmap(path, dtype=np.uint8, mode="r+" if writable else "r")
    h = m[:16].view(PATHS_HEADER_DTYPE)[0]
    if int(h["magic"]) != PATHS_MAGIC or int(h["max_segments"]) != MAX_SEGMENTS:
        raise ValueError(f"{path}: not a paths file")
    A = int(h["total_axons"])
    if m.size < paths_file_size(A):
        raise ValueError(f"{path}: short paths file")
    off = paths_matrix_offset(A)
    return m[16:16 + A], m[off:off + A * MAX_SEGMENTS * 4].view(np.uint32).reshape(A, MAX_SEGMENTS)


def write_geom(path: str, tips: np.ndarray, dirs: np.ndarray) -> None:
    assert len(tips) == len(dirs)
    _write_atomic(path, np.concatenate([np.ascontiguousarray(tips, np.uint32),
                                        np.ascontiguousarray(dirs, np.uint32)]).view(np.uint8))


def read_geom(path: str, total_axons: int, writable: bool = False) -> tuple[np.ndarray, np.ndarray]:
    """(tips, dirs) views of a shard.geom file of total_axons entries each."""
    m = np.memmap(path, dtype=np.uint32, mode="r+" if writable else "r")
    if m.size < 2 * total_axons:
        raise ValueError(f"{path}: short geom file")
    return m[:total_axons], m[total_axons:2 * total_axons]


def write_bake(directory: str, brain) -> None:
    """shard.state, shard.axons, shard.paths, shard.geom and shard.pos of a
    generated brain (generate(..., with_paths=True))."""
    if brain.paths is None:
        raise ValueError("the brain has no geometry: generate it with with_paths=True")
    os.makedirs(directory, exist_ok=True)
    _write_atomic(os.path.join(directory, "shard.state"), brain.state)
    _write_atomic(os.path.join(directory, "shard.axons"), brain.heads.view(np.uint8))
    write_paths(os.path.join(directory, "shard.paths"), brain.axon_len, brain.paths)
    write_geom(os.path.join(directory, "shard.geom"), brain.tips, brain.dirs)
    _write_atomic(os.path.join(directory, "shard.pos"), np.ascontiguousarray(brain.soma_pos, np.uint32).view(np.uint8))


def load_geometry(directory: str, padded_n: int, total_axons: int, max_x: int, max_y: int, total_ghosts: int = 0,
                  writable: bool = True) -> Geometry:
    """The Night geometry of a baked directory, as views of the daemon's
    memory maps: the sprouting pass then updates the files in place. max_x
    and max_y are the shard's width and depth in voxels (the daemon's
    shard_cfg.dimensions.w and .d, daemon.rs:275-276): axons nudged past
    them leave as handovers."""
    lengths, paths = read_paths(os.path.join(directory, "shard.paths"), writable)
    if len(lengths) != total_axons:
        raise ValueError("shard.paths has a different axon count")
    tips, dirs = read_geom(os.path.join(directory, "shard.geom"), total_axons, writable)
    pos = np.memmap(os.path.join(directory, "shard.pos"), dtype=np.uint32, mode="r")[:padded_n]
    return Geometry(lengths, paths, tips, dirs, np.ascontiguousarray(pos), total_ghosts, None, max_x, max_y)


def save_checkpoint(vram: VramState, directory: str) -> tuple[str, str]:
    """save_hot_checkpoint (shard_thread.rs:229-263): both device blobs to
    .tmp files, then renamed over checkpoint.state and checkpoint.axons."""
    state, axons = vram.download_state(), vram.download_axons()
    os.makedirs(directory, exist_ok=True)
    s, a = os.path.join(directory, "checkpoint.state"), os.path.join(directory, "checkpoint.axons")
    for path, data in ((s + ".tmp", state), (a + ".tmp", axons)):
        with open(path, "wb") as f:
            f.write(np.ascontiguousarray(data).view(np.uint8).tobytes())
    os.replace(s + ".tmp", s)
    os.replace(a + ".tmp", a)
    return s, a


def load_checkpoint(vram: VramState, directory: str) -> None:
    """Upload checkpoint.state and checkpoint.axons into a shard of the same size."""
    state = np.fromfile(os.path.join(directory, "checkpoint.state"), dtype=np.uint8)
    axons = np.fromfile(os.path.join(directory, "checkpoint.axons"), dtype=np.uint32)
    vram.upload_state(state)
    vram.upload_axon_heads(axons)


def _write_atomic(path: str, data: np.ndarray) -> None:
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(np.ascontiguousarray(data).tobytes())
    os.replace(tmp, path)
