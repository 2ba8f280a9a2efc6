"""Host-side mirror of genesis-compute's Rust API over the C-ABI.

Same names, argument meaning and fail-fast behaviour as
genesis-compute/src/memory.rs (VramState, PinnedBuffer) and
genesis-compute/src/compute/shard.rs (IoDeviceBuffers, ShardEngine): every
nonzero return of the C-ABI raises (the Rust code asserts and panics).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._ffi import AxcShardStats, ShardVramPtrs, VARIANT_DTYPE, day_lib, ptr
from .layout import calculate_state_blob_size


class GpuError(RuntimeError):
    pass


def _check(err: int, what: str) -> None:
    if err != 0:
        raise GpuError(f"FATAL: {what} failed (hipError={err})")


class DeviceBuffer:
    """A plain device allocation made with gpu_malloc (ffi.rs:114)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = day_lib().gpu_malloc(max(self.nbytes, 1))
        if not self.ptr:
            raise GpuError(f"gpu_malloc({nbytes}) failed")

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        if a.nbytes and not day_lib().gpu_memcpy_host_to_device(self.ptr + offset, ptr(a), a.nbytes):
            raise GpuError("gpu_memcpy_host_to_device failed")

    def download(self, dtype, count: int | None = None, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes and not day_lib().gpu_memcpy_device_to_host(ptr(out), self.ptr + offset, out.nbytes):
            raise GpuError("gpu_memcpy_device_to_host failed")
        return out

    def free(self) -> None:
        if self.ptr:
            day_lib().gpu_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_hip = None


def device_copy(dst: int, src: int, nbytes: int) -> None:
    """Synchronous device-to-device copy (hipMemcpy of the HIP runtime that
    libaxicor_day.so links)."""
    global _hip
    if _hip is None:
        day_lib()  # loads the runtime first
        _hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemcpy.restype = C.c_int
    _check(_hip.hipMemcpy(dst, src, nbytes, 3), "hipMemcpy device to device")


def upload_constant_memory(lut: np.ndarray) -> None:
    """cu_upload_constant_memory (physics.cu:504-506): the 16-entry table."""
    lut = np.ascontiguousarray(lut, dtype=VARIANT_DTYPE)
    assert lut.shape == (16,)
    _check(day_lib().cu_upload_constant_memory(ptr(lut)), "cu_upload_constant_memory")


class VramState:
    """memory.rs:98-184."""

    def __init__(self, padded_n: int, total_axons: int, total_ghosts: int = 0):
        self.ptrs = ShardVramPtrs()
        _check(day_lib().cu_allocate_shard(padded_n, total_axons, C.byref(self.ptrs)), "cu_allocate_shard")
        self.padded_n = padded_n
        self.total_axons = total_axons
        self.total_ghosts = total_ghosts

    allocate = classmethod(lambda cls, n, a, g=0: cls(n, a, g))

    def virtual_offset(self) -> int:
        """memory.rs:128-130."""
        return self.padded_n + self.total_ghosts

    def upload_state(self, blob: np.ndarray) -> None:
        """memory.rs:134-149: size must match exactly."""
        _, expected = calculate_state_blob_size(self.padded_n)
        if blob.nbytes != expected:
            raise GpuError(f"FATAL: .state blob size mismatch: got {blob.nbytes} expected {expected}")
        blob = np.ascontiguousarray(blob)
        _check(day_lib().cu_upload_state_blob(C.byref(self.ptrs), ptr(blob), blob.nbytes), "cu_upload_state_blob")

    def upload_axon_heads(self, blob: np.ndarray) -> None:
        """memory.rs:152-175: larger than 32*A panics; shorter is accepted."""
        blob = np.ascontiguousarray(blob)
        expected = self.total_axons * 32
        if blob.nbytes > expected:
            raise GpuError(f"FATAL: axon_heads blob too large: got {blob.nbytes} expected max {expected}")
        if blob.nbytes == 0:
            return
        _check(day_lib().cu_upload_axons_blob(C.byref(self.ptrs), ptr(blob), blob.nbytes), "cu_upload_axons_blob")

    def set_soma_positions(self, positions: np.ndarray | None) -> None:
        """axc_shard_set_soma_positions: the steps run in the engine's own
        soma order (bricks of these positions); None returns to the caller's."""
        if positions is None:
            _check(day_lib().axc_shard_set_soma_positions(C.byref(self.ptrs), None, 0), "set_soma_positions")
            return
        pos = np.ascontiguousarray(positions, np.uint32)
        _check(day_lib().axc_shard_set_soma_positions(C.byref(self.ptrs), ptr(pos), pos.size), "set_soma_positions")

    def download_state(self) -> np.ndarray:
        """The `.state` bytes, as save_hot_checkpoint reads them (shard_thread.rs:231-265)."""
        _, nbytes = calculate_state_blob_size(self.padded_n)
        out = np.empty(nbytes, dtype=np.uint8)
        if not day_lib().gpu_memcpy_device_to_host(ptr(out), self.ptrs.soma_voltage, nbytes):
            raise GpuError("state download failed")
        return out

    def download_axons(self) -> np.ndarray:
        out = np.empty(self.total_axons * 8, dtype=np.uint32)
        if not day_lib().gpu_memcpy_device_to_host(ptr(out), self.ptrs.axon_heads, out.nbytes):
            raise GpuError("axons download failed")
        return out

    def free(self) -> None:
        if self.ptrs.soma_voltage or self.ptrs.axon_heads:
            day_lib().cu_free_shard(C.byref(self.ptrs))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedBuffer:
    """memory.rs:193-233: page-locked host memory from gpu_host_alloc."""

    def __init__(self, length: int, dtype=np.uint32):
        self.dtype = np.dtype(dtype)
        self.len = int(length)
        self._p = None
        if self.len == 0:
            self.array = np.zeros(0, dtype=self.dtype)
            return
        nbytes = self.len * self.dtype.itemsize
        self._p = day_lib().gpu_host_alloc(nbytes)
        if not self._p:
            raise GpuError(f"PinnedBuffer: gpu_host_alloc failed for {nbytes} bytes")
        buf = (C.c_uint8 * nbytes).from_address(self._p)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=self.len)

    def as_ptr(self) -> int | None:
        return self._p

    def __del__(self):
        try:
            if self._p:
                day_lib().gpu_host_free(self._p)
                self._p = None
        except Exception:
            pass


class IoDeviceBuffers:
    """compute/shard.rs:4-11, allocated with cu_allocate_io_buffers (ffi.rs:248)."""

    def __init__(self, input_words_per_tick: int, max_spikes_per_tick: int, num_outputs: int, sync_batch_ticks: int):
        self.input_words_per_tick = input_words_per_tick
        self.max_spikes_per_tick = max_spikes_per_tick
        self.num_outputs = num_outputs
        a, b, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(day_lib().cu_allocate_io_buffers(input_words_per_tick * sync_batch_ticks,
                                                max_spikes_per_tick * sync_batch_ticks,
                                                num_outputs * sync_batch_ticks,
                                                C.byref(a), C.byref(b), C.byref(c)), "cu_allocate_io_buffers")
        self.d_input_bitmask = a.value or 0
        self.d_incoming_spikes = b.value or 0
        self.d_output_history = c.value or 0


class ShardEngine:
    """compute/shard.rs:18-131: owns a VramState and a non-blocking stream."""

    def __init__(self, vram: VramState):
        self.vram = vram
        s = C.c_void_p()
        _check(day_lib().gpu_stream_create(C.byref(s)), "hipStreamCreate")
        self.stream = s.value

    def step_day_phase_batch(self, sync_batch_ticks: int, io: IoDeviceBuffers, h_input_bitmask: np.ndarray | None,
                             h_incoming_spikes: np.ndarray | None, h_spike_counts: np.ndarray | list,
                             virtual_offset: int, num_virtual_axons: int, mapped_soma_ids_device: int | None,
                             v_seg: int, dopamine: int, tick_base: int, use_graph: bool = False) -> None:
        """shard.rs:40-125. With use_graph the tick loop runs as one captured
        HIP graph (axc_step_day_phase_batch); the results are identical."""
        lib = day_lib()
        T = sync_batch_ticks
        words = io.input_words_per_tick * T if h_input_bitmask is not None else 0
        hin = np.ascontiguousarray(h_input_bitmask, dtype=np.uint32) if h_input_bitmask is not None else None
        hsp = np.ascontiguousarray(h_incoming_spikes, dtype=np.uint32) if h_incoming_spikes is not None else None
        counts = np.ascontiguousarray(np.asarray(h_spike_counts, dtype=np.uint32).reshape(-1)[:T])
        if counts.size < T:
            counts = np.concatenate([counts, np.zeros(T - counts.size, np.uint32)])
        # Inputs as the reference uploads them (cu_dma_h2d_io); the schedule
        # without its full-capacity copy: only the used part of each tick's
        # slice (axc_dma_h2d_schedule), nothing for an empty schedule.
        _check(lib.cu_dma_h2d_io(io.d_input_bitmask or None, ptr(hin), words, None, None, 0, self.stream),
               "cu_dma_h2d_io")
        if hsp is not None and io.d_incoming_spikes:
            _check(lib.axc_dma_h2d_schedule(io.d_incoming_spikes, ptr(hsp), io.max_spikes_per_tick, ptr(counts), T,
                                            self.stream), "axc_dma_h2d_schedule")
        bm = io.d_input_bitmask if (io.d_input_bitmask and hin is not None) else None
        sp = io.d_incoming_spikes if (io.d_incoming_spikes and hsp is not None) else None
        outp = io.d_output_history or None
        if use_graph:
            _check(lib.axc_step_day_phase_batch(C.byref(self.vram.ptrs), self.vram.padded_n, self.vram.total_axons,
                                                v_seg, tick_base, T, bm, io.input_words_per_tick, virtual_offset,
                                                num_virtual_axons, sp, io.max_spikes_per_tick, ptr(counts),
                                                mapped_soma_ids_device, outp, io.num_outputs, dopamine, self.stream),
                   "axc_step_day_phase_batch")
            return
        lib.cu_reset_burst_counters(C.byref(self.vram.ptrs), self.vram.padded_n, self.stream)
        for t in range(T):
            tin = bm + t * io.input_words_per_tick * 4 if bm else None
            tsp = sp + t * io.max_spikes_per_tick * 4 if sp else None
            tout = outp + t * io.num_outputs if outp else None
            err = lib.cu_step_day_phase(C.byref(self.vram.ptrs), self.vram.padded_n, self.vram.total_axons, v_seg,
                                        (tick_base + t) & 0xFFFFFFFF, tin, virtual_offset, num_virtual_axons, tsp,
                                        int(counts[t]), mapped_soma_ids_device, tout, io.num_outputs, dopamine,
                                        self.stream)
            _check(err, f"Day Phase Pipeline at tick {t}")

    def synchronize(self) -> None:
        day_lib().gpu_stream_synchronize(self.stream)

    def graph_captures(self) -> int:
        """Graphs recorded for this shard by the batch entry point so far."""
        n = C.c_uint64()
        _check(day_lib().axc_shard_graph_captures(C.byref(self.vram.ptrs), C.byref(n)), "axc_shard_graph_captures")
        return n.value

    def stats_reset(self) -> None:
        _check(day_lib().axc_shard_stats_reset(C.byref(self.vram.ptrs), self.stream), "axc_shard_stats_reset")

    def stats(self) -> dict:
        self.synchronize()
        s = AxcShardStats()
        _check(day_lib().axc_shard_stats_read(C.byref(self.vram.ptrs), C.byref(s)), "axc_shard_stats_read")
        return {n: getattr(s, n) for n, _ in AxcShardStats._fields_}

    WINDOW_STATS = ("windows", "pass_somas", "pass_slots", "correction_pieces", "corrected_entries")

    def window_stats(self) -> dict:
        """Work counters of the windowed batches since stats_reset()."""
        self.synchronize()
        out = (C.c_uint64 * len(self.WINDOW_STATS))()
        _check(day_lib().axc_shard_window_stats(C.byref(self.vram.ptrs), out), "axc_shard_window_stats")
        return dict(zip(self.WINDOW_STATS, (int(x) for x in out)))

    def destroy(self) -> None:
        if self.stream:
            day_lib().gpu_stream_destroy(self.stream)
            self.stream = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
