"""The CPU oracle against the reference's own HIP kernels, built from the
reference sources by oracle/build_ref.sh (genesis-compute/src/amd/physics.hip,
bindings.hip) and run on the GPU through the reference's C-ABI.

The reference K5 reads axon heads while spiking threads of the same launch
push them (physics.hip update kernel), so its result depends on scheduling
whenever a dendrite reads an axon whose soma spikes in that tick. These brains
take every dendrite target from the input (virtual) axons, which only K2
pushes, before K5: the reference run is then deterministic, and every other
rule (K1-K7, K9, GLIF, heartbeat, homeostasis, adaptive leak, GSOP with
dopamine and bursts, readout) is compared bit for bit with the oracle."""
import ctypes as C
import os

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd._ffi import ShardVramPtrs
from oracle import oracle as O

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                   "libgenesis_amd_ref.so")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built")]


def ref_lib():
    L = C.CDLL(REF)
    P, U32, I16, VP = C.c_void_p, C.c_uint32, C.c_int16, C.POINTER(ShardVramPtrs)
    L.cu_allocate_shard.argtypes = [U32, U32, VP]
    L.cu_allocate_shard.restype = C.c_int32
    L.cu_upload_state_blob.argtypes = [VP, P, C.c_size_t]
    L.cu_upload_state_blob.restype = C.c_int32
    L.cu_upload_axons_blob.argtypes = [VP, P, C.c_size_t]
    L.cu_upload_axons_blob.restype = C.c_int32
    L.cu_upload_constant_memory.argtypes = [P]
    L.cu_upload_constant_memory.restype = C.c_int32
    L.cu_step_day_phase.argtypes = [VP, U32, U32, U32, U32, P, U32, U32, P, U32, P, P, U32, I16, P]
    L.cu_step_day_phase.restype = C.c_int32
    L.cu_reset_burst_counters.argtypes = [VP, U32, P]
    L.cu_reset_burst_counters.restype = None
    L.launch_sort_and_prune.argtypes = [VP, U32, I16]
    L.launch_sort_and_prune.restype = None
    L.gpu_malloc.argtypes = [C.c_size_t]
    L.gpu_malloc.restype = P
    L.gpu_memcpy_host_to_device.argtypes = [P, P, C.c_size_t]
    L.gpu_memcpy_host_to_device.restype = C.c_bool
    L.gpu_memcpy_device_to_host.argtypes = [P, P, C.c_size_t]
    L.gpu_memcpy_device_to_host.restype = None  # void in the reference (bindings.hip:235)
    L.gpu_device_synchronize.argtypes = []
    L.gpu_device_synchronize.restype = None
    L.cu_free_shard.argtypes = [VP]
    L.cu_free_shard.restype = None
    return L


def input_only_brain(seed: int, n: int = 3000, slots: int = 48, n_in: int = 512, lut=None, type_ppm=None):
    b = B.generate(B.spec_small(n, slots, n_in, seed=seed, lut=lut, type_ppm=type_ppm))
    v = b.view()
    rng = np.random.default_rng(seed)
    filled = v.targets != 0
    k = int(filled.sum())
    ax = rng.integers(0, n_in, size=k)
    seg = rng.integers(0, b.axon_len[b.padded_n + ax].astype(np.int64))
    v.targets[filled] = ((seg.astype(np.uint32) << 24) | (b.padded_n + ax + 1).astype(np.uint32))
    v.weights[filled] = np.abs(v.weights[filled].astype(np.int32)).astype(np.int16)
    # Some inhibitory-looking weights as well (input axons carry any sign).
    neg = rng.random(k) < 0.25
    w = v.weights[filled].astype(np.int32)
    w[neg] = -w[neg]
    v.weights[filled] = w.astype(np.int16)
    return b


def run_ref_and_oracle(b: B.Brain, batches, dop: int, n_out: int = 32, prune_every: int = 0):
    L = ref_lib()
    assert L.cu_upload_constant_memory(b.lut.ctypes.data) == 0
    ptrs = ShardVramPtrs()
    assert L.cu_allocate_shard(b.padded_n, b.total_axons, C.byref(ptrs)) == 0
    assert L.cu_upload_state_blob(C.byref(ptrs), b.state.ctypes.data, b.state.nbytes) == 0
    assert L.cu_upload_axons_blob(C.byref(ptrs), b.heads.ctypes.data, b.heads.nbytes) == 0
    words = b.input_words
    d_mask = L.gpu_malloc(words * 4)
    rng = np.random.default_rng(1)
    mapped = rng.integers(0, b.padded_n, size=n_out).astype(np.uint32)
    mapped[::5] = 0xFFFFFFFF
    d_map = L.gpu_malloc(n_out * 4)
    assert L.gpu_memcpy_host_to_device(d_map, mapped.ctypes.data, mapped.nbytes)
    d_out = L.gpu_malloc(n_out)
    ref_state, ref_heads = b.state.copy(), b.heads.copy()
    tick = 0
    for bi, T in enumerate(batches):
        inputs = b.inputs(tick, T)
        L.cu_reset_burst_counters(C.byref(ptrs), b.padded_n, None)
        outs = np.zeros((T, n_out), np.uint8)
        for t in range(T):
            row = np.ascontiguousarray(inputs[t])
            assert L.gpu_memcpy_host_to_device(d_mask, row.ctypes.data, row.nbytes)
            assert L.cu_step_day_phase(C.byref(ptrs), b.padded_n, b.total_axons, b.spec.v_seg, tick + t, d_mask,
                                       b.virtual_offset, b.n_inputs, None, 0, d_map, d_out, n_out, dop, None) == 0
            L.gpu_device_synchronize()
            o = np.zeros(n_out, np.uint8)
            L.gpu_memcpy_device_to_host(o.ctypes.data, d_out, n_out)
            outs[t] = o
        out_ref, _ = O.step_batch(ref_state, b.padded_n, ref_heads, b.total_axons, b.lut, b.spec.v_seg, tick, T,
                                  inputs, b.virtual_offset, b.n_inputs, mapped=mapped, n_out=n_out, dopamine=dop)
        tick += T
        got_s = np.empty_like(b.state)
        got_h = np.empty_like(b.heads)
        L.gpu_memcpy_device_to_host(got_s.ctypes.data, ptrs.soma_voltage, got_s.nbytes)
        L.gpu_memcpy_device_to_host(got_h.ctypes.data, ptrs.axon_heads, got_h.nbytes)
        np.testing.assert_array_equal(outs, out_ref, err_msg=f"outputs batch {bi}")
        assert np.array_equal(got_h, ref_heads), f"heads differ after batch {bi}"
        assert np.array_equal(got_s, ref_state), f"state differs after batch {bi} " + str(
            np.nonzero(got_s != ref_state)[0][:5])
        if prune_every and (bi + 1) % prune_every == 0:
            L.launch_sort_and_prune(C.byref(ptrs), b.padded_n, 40)
            L.gpu_device_synchronize()
            O.sort_and_prune_blob(ref_state, b.padded_n, 40)
            L.gpu_memcpy_device_to_host(got_s.ctypes.data, ptrs.soma_voltage, got_s.nbytes)
            assert np.array_equal(got_s, ref_state), f"sort&prune differs after batch {bi}"
    L.cu_free_shard(C.byref(ptrs))


def test_reference_kernels_two_types(gpu):
    run_ref_and_oracle(input_only_brain(31), [5, 20, 25, 50], dop=0, prune_every=2)


def test_reference_kernels_four_types_dopamine(gpu):
    b = input_only_brain(32, n=2500, slots=100, lut=B.lut_c2(), type_ppm=[400_000, 400_000, 100_000, 100_000])
    run_ref_and_oracle(b, [10, 30, 30], dop=-350, prune_every=1)
    b = input_only_brain(33, n=2500, slots=100, lut=B.lut_c2(), type_ppm=[400_000, 400_000, 100_000, 100_000])
    run_ref_and_oracle(b, [40, 40], dop=2000)
