"""Batched brains: every brain of a batch is bit-identical to the same brain
stepped alone (axc_step_day_phase_batch), and to the CPU oracle."""
import ctypes as C

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd._ffi import day_lib
from axicor_amd.batch import BatchEngine
from axicor_amd.engine import DeviceBuffer, ShardEngine, VramState, upload_constant_memory
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make_members(seed: int, n_brains: int, ticks: int):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_brains):
        spec, params = B.spec_batch_member(rng, n_min=1500, n_max=12000)
        b = B.generate(spec)
        out.append((b, int(params[6]), b.inputs(0, ticks)))
    return out


def load(b, dop, inputs):
    v = VramState(b.padded_n, b.total_axons)
    v.upload_state(b.state)
    v.upload_axon_heads(b.heads)
    d_in = DeviceBuffer.from_array(inputs)
    return {"vram": v, "lut": b.lut, "d_inputs": d_in, "input_words": b.input_words, "input_ticks": inputs.shape[0],
            "v_seg": 1, "virtual_offset": b.virtual_offset, "n_inputs": b.n_inputs, "dopamine": dop}


def test_batch_bit_identical_to_solo_and_oracle(gpu):
    T, nb = 30, 2
    members = make_members(5, 7, T * nb)
    lib = day_lib()
    # Batched
    mem_b = [load(b, dop, inp) for b, dop, inp in members]
    be = BatchEngine(mem_b)
    tune = be.autotune(4)
    assert tune["soma_blocks"] % 8 == 0
    for k in range(nb):
        be.step(k * T, T)
    lib.gpu_device_synchronize()
    # Solo: the plain engine, one brain at a time (each with its own table)
    for k, (b, dop, inp) in enumerate(members):
        m = load(b, dop, inp)
        upload_constant_memory(b.lut)
        eng = ShardEngine(m["vram"])
        for j in range(nb):
            err = lib.axc_step_day_phase_batch(C.byref(m["vram"].ptrs), b.padded_n, b.total_axons, 1, j * T, T,
                                               m["d_inputs"].ptr + j * T * b.input_words * 4, b.input_words,
                                               b.virtual_offset, b.n_inputs, None, 0, None, None, None, 0, dop,
                                               eng.stream)
            assert err == 0
        eng.synchronize()
        sb, hb = mem_b[k]["vram"].download_state(), mem_b[k]["vram"].download_axons()
        ss, hs = m["vram"].download_state(), m["vram"].download_axons()
        assert np.array_equal(sb, ss), f"brain {k}: batched state != solo state"
        assert np.array_equal(hb, hs), f"brain {k}: batched heads != solo heads"
        if k < 2:
            st, hd = b.state.copy(), b.heads.copy()
            O.step_batch(st, b.padded_n, hd, b.total_axons, b.lut, 1, 0, T * nb, inp, b.virtual_offset, b.n_inputs,
                         dopamine=dop)
            assert np.array_equal(sb, st) and np.array_equal(hb, hd), f"brain {k}: batched != oracle"
        s, ts = be.stats(k)
        assert s["ticks"] == T * nb and ts.sum() == s["spikes"]
    be.destroy()


def test_batch_graph_replays_and_autotune_restore(gpu):
    """Autotuning leaves the state as it was; repeated graph replays keep
    matching the oracle."""
    members = make_members(9, 3, 40)
    mem = [load(b, dop, inp) for b, dop, inp in members]
    be = BatchEngine(mem)
    before = [m["vram"].download_state() for m in mem]
    be.autotune(3)
    after = [m["vram"].download_state() for m in mem]
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    for k in range(4):
        be.step(k * 10, 10)
    day_lib().gpu_device_synchronize()
    for (b, dop, inp), m in zip(members, mem):
        st, hd = b.state.copy(), b.heads.copy()
        O.step_batch(st, b.padded_n, hd, b.total_axons, b.lut, 1, 0, 40, inp, b.virtual_offset, b.n_inputs,
                     dopamine=dop)
        assert np.array_equal(m["vram"].download_state(), st)
        assert np.array_equal(m["vram"].download_axons(), hd)
    be.destroy()


@pytest.mark.parametrize("stage_order", [0, 1, 3, 64])
def test_batch_stage_orders_equal_the_oracle(gpu, stage_order):
    """Every dispatch order of the staged pass (brain order, balanced runs of
    1, 3 or 64 blocks, the XCD shares padded) gives every brain the oracle's
    state: the order only moves work between XCDs."""
    members = make_members(21, 6, 36)
    mem = [load(b, dop, inp) for b, dop, inp in members]
    be = BatchEngine(mem)
    t = be.autotune(2)
    assert t["stage_order"] in (0, 4, 32, 256)
    be.set_tuning(t["axon_blocks"], t["soma_blocks"], t["list_blocks"], stage_order)
    for k in range(3):
        be.step(k * 12, 12)
    day_lib().gpu_device_synchronize()
    for (b, dop, inp), m in zip(members, mem):
        st, hd = b.state.copy(), b.heads.copy()
        O.step_batch(st, b.padded_n, hd, b.total_axons, b.lut, 1, 0, 36, inp, b.virtual_offset, b.n_inputs,
                     dopamine=dop)
        assert np.array_equal(m["vram"].download_state(), st), f"order {stage_order}: state != oracle"
        assert np.array_equal(m["vram"].download_axons(), hd), f"order {stage_order}: heads != oracle"
    be.destroy()


def test_batch_synapse_stats_match_the_downloaded_state(gpu):
    """axc_batch_synapse_stats against the same sums over the state blob."""
    from axicor_amd.layout import compute_state_offsets
    members = make_members(9, 5, 40)
    mems = [load(b, dop, inp) for b, dop, inp in members]
    be = BatchEngine(mems)
    be.step(0, 40)
    day_lib().gpu_device_synchronize()
    prune = np.array([2, 15, 30, 100, 0], np.int16)
    got = be.synapse_stats(prune)
    be.destroy()
    for k, m in enumerate(mems):
        v = m["vram"]
        st = v.download_state()
        off = compute_state_offsets(v.padded_n)
        N = v.padded_n
        t = st[off.dendrite_targets:off.dendrite_targets + 4 * 128 * N].view(np.uint32)
        w = st[off.dendrite_weights:off.dendrite_weights + 2 * 128 * N].view(np.int16).astype(np.int64)
        thr = st[off.threshold_offset:off.threshold_offset + 4 * N].view(np.int32).astype(np.int64)
        aw = np.abs(w[t != 0])
        want = [aw.size, aw.sum(), (aw * aw).sum(), np.count_nonzero(aw < prune[k]), thr.sum()]
        assert got[k].tolist() == want, (k, got[k].tolist(), want)


def test_batch_after_a_brain_is_pruned_mid_run(gpu):
    """Sort & prune at 200 on one brain between two batch steps: the batch
    rebuilds that brain's staging plan before the next step (its fills
    shrink, its slots move) and reloads its working copy of the weights.
    Every brain equals the oracle run through the same sequence."""
    import ctypes as C
    T = 24
    members = make_members(33, 5, 2 * T)
    mem = [load(b, dop, inp) for b, dop, inp in members]
    be = BatchEngine(mem)
    be.autotune(2)
    lib = day_lib()
    be.step(0, T)
    lib.gpu_device_synchronize()
    v0 = mem[0]["vram"]
    lib.launch_sort_and_prune(C.byref(v0.ptrs), v0.padded_n, 200)
    lib.gpu_device_synchronize()
    be.step(T, T)
    lib.gpu_device_synchronize()
    for k, ((b, dop, inp), m) in enumerate(zip(members, mem)):
        st, hd = b.state.copy(), b.heads.copy()
        O.step_batch(st, b.padded_n, hd, b.total_axons, b.lut, 1, 0, T, inp, b.virtual_offset, b.n_inputs,
                     dopamine=dop)
        if k == 0:
            O.sort_and_prune_blob(st, b.padded_n, 200)
        O.step_batch(st, b.padded_n, hd, b.total_axons, b.lut, 1, T, T, inp[T:], b.virtual_offset, b.n_inputs,
                     dopamine=dop)
        assert np.array_equal(m["vram"].download_state(), st), f"brain {k}: state != oracle"
        assert np.array_equal(m["vram"].download_axons(), hd), f"brain {k}: heads != oracle"
    be.destroy()
