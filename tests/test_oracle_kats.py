"""Pins the CPU oracle against the reference's own known-answer tests and
against hand-traced tiny brains whose outcome the mathematics fixes.

Reference vectors: genesis-core/src/signal.rs:71-102, test_signal.rs:6-98,
layout.rs:388-404, hash.rs:23-28, test_seed.rs:46-66, physics.rs:272-329
(the GSOP vectors that match the kernel: dopamine 0, burst multiplier 1,
cooling shift 0, slot decay 128), test_tick.rs:134-147 (propagation)."""
import ctypes as C

import numpy as np
import pytest

from axicor_amd import brain as B
from axicor_amd._ffi import VARIANT_DTYPE
from axicor_amd.layout import StateView, calculate_state_blob_size
from oracle import oracle as O

SENT = 0x80000000


def L():
    return O.lib()


# ---------------------------------------------------------------- active tail
@pytest.mark.parametrize("head,seg,prop,want", [
    (10, 8, 3, True), (10, 10, 3, True), (10, 9, 3, True),      # signal.rs:72-79
    (10, 7, 3, False), (10, 0, 3, False),                       # signal.rs:82-87
    (SENT, 5, 3, False), (SENT, 0, 255, False), (SENT, 1000, 100, False),  # signal.rs:90-95
    (10, 10, 0, False), (10, 9, 0, False),                      # signal.rs:98-102
    (SENT, 0, 10, False), (SENT, 100, 1, False),                # test_signal.rs:7-10
])
def test_active_tail_reference_vectors(head, seg, prop, want):
    assert bool(L().orc_active_tail(head, seg, prop)) == want


def test_tail_width_equals_prop():
    """test_signal.rs:59-70: head 100, prop 7 covers exactly 7 segments."""
    assert sum(L().orc_active_tail(100, s, 7) for s in range(200)) == 7


def test_sentinel_dead_for_every_u8_prop():
    """test_signal.rs:91-98 for the kernel's u8 propagation length."""
    for p in (0, 1, 10, 255):
        assert not L().orc_active_tail(SENT, 0, p)


def test_initial_head_reaches_sentinel_mid_flight_active():
    """test_signal.rs:35-55: h = SENT - len*v_seg; after len-prop ticks the
    starting segment is in the tail."""
    length, prop = 5, 3
    head = (SENT - length) & 0xFFFFFFFF
    head = (head + (length - prop)) & 0xFFFFFFFF
    assert L().orc_active_tail(head, (SENT - length) & 0xFFFFFFFF, prop)


# ------------------------------------------------------------ target packing
def test_pack_reference_vectors():
    """layout.rs:388-404."""
    t0 = L().orc_pack_target(0, 0)
    assert t0 != 0 and L().orc_unpack_axon(t0) == 0 and t0 >> 24 == 0
    tm = L().orc_pack_target(0x00FFFFFE, 255)
    assert L().orc_unpack_axon(tm) == 0x00FFFFFE and tm >> 24 == 255
    tx = L().orc_pack_target(0x123456, 0xAB)
    assert L().orc_unpack_axon(tx) == 0x123456 and tx >> 24 == 0xAB


# -------------------------------------------------------------------- hashes
def test_fnv1a_32_reference_vector():
    """hash.rs:23-28: "SensoryCortex" -> 0x273fd103."""
    s = b"SensoryCortex"
    assert L().orc_fnv1a_32(s, len(s)) == 0x273fd103


def test_seed_from_str_matches_python_fnv64():
    for s in ["GENESIS", "", "GENESIS_DAEMON_TEST", "  x  "]:
        assert L().orc_seed_from_str(s.encode(), len(s.encode())) == B.seed_from_str(s)


def test_entity_seed_avalanche():
    """test_seed.rs:55-66: adjacent ids differ in >= 16 bits."""
    a = L().orc_entity_seed(0x1234567890ABCDEF, 10)
    b = L().orc_entity_seed(0x1234567890ABCDEF, 11)
    assert a != b and bin(a ^ b).count("1") >= 16


# ---------------------------------------------------------------------- GSOP
def _variant(**kw):
    v = np.zeros(1, dtype=VARIANT_DTYPE)
    v["gsop_potentiation"] = kw.get("pot", 80)
    v["gsop_depression"] = kw.get("dep", 40)
    v["inertia_curve"] = kw.get("curve", [128] * 16)
    v["signal_propagation_length"] = kw.get("prop", 20)
    v["d1_affinity"] = kw.get("d1", 0)
    v["d2_affinity"] = kw.get("d2", 0)
    return v


def _heads(active_dist=None, seg=40):
    h = [SENT] * 8
    if active_dist is not None:
        h[0] = seg + active_dist
    return h


@pytest.mark.parametrize("w,causal,want", [
    (100, True, 180), (100, False, 60),            # physics.rs:276-295
    (-32768, True, -32767), (-32768, False, -32728),  # physics.rs:298-314
    (-500, False, -460), (-5, False, 0),           # physics.rs:317-334
    (32767, True, 32767),                          # physics.rs:337-341
    (1000, True, 1080),                            # physics.rs:350-358 (slot decay 128)
])
def test_gsop_reference_vectors(w, causal, want):
    v = _variant()
    h = _heads(0 if causal else None)
    assert O.gsop_update(w, h, 40, v, 1, 0) == want


def test_gsop_inertia_rank_boundaries():
    """physics.rs:343-347 rank = |w|>>11 (the kernel clamps at 15, not 14)."""
    curve = list(range(1, 17))  # distinct inertia per rank
    v = _variant(pot=128, dep=0, curve=curve)
    for w, rank in [(0, 0), (2047, 0), (2048, 1), (4095, 1), (4096, 2), (30719, 14), (30720, 15), (32767, 15)]:
        # delta = ((128*inertia*1)>>7 * 128) >> 7 = inertia for cs = 0
        got = O.gsop_update(w, _heads(0), 40, v, 1, 0)
        assert got == min(32767, w + curve[rank]), (w, rank, got)


def test_gsop_cooling_shift_halves_every_16():
    """cooling shift = min_dist >> 4 is applied twice (physics.cu:351-358)."""
    v = _variant(pot=128, curve=[128] * 16, prop=255)
    # dp = 128; delta = ((128 >> cs) * 128) >> (7 + cs)
    for d, cs in [(0, 0), (15, 0), (16, 1), (31, 1), (32, 2), (200, 12)]:
        want = ((128 >> cs) * 128) >> (7 + cs)
        assert O.gsop_update(1000, _heads(d), 40, v, 1, 0) == 1000 + want, d


def test_gsop_burst_and_dopamine():
    """bm multiplies both deltas; dopamine scales pot by d1 and dep by d2,
    (dop*aff)>>7 with an arithmetic shift, depression clamped at 0."""
    v = _variant(pot=100, dep=60, d1=128, d2=64)
    # dop = 256: pot_mod = 256, dep_mod = 128 -> raw_pot 356, raw_dep -68 -> 0
    assert O.gsop_update(1000, _heads(0), 40, v, 2, 256) == 1000 + 712
    assert O.gsop_update(1000, _heads(None), 40, v, 2, 256) == 1000
    # dop = -1: pot_mod = (-128)>>7 = -1, dep_mod = (-64)>>7 = -1
    assert O.gsop_update(1000, _heads(0), 40, v, 1, -1) == 1000 + 99
    assert O.gsop_update(1000, _heads(None), 40, v, 1, -1) == 1000 - 61
    # Large negative dopamine turns LTP into a penalty (no clamp on pot).
    assert O.gsop_update(1000, _heads(0), 40, v, 1, -200) == 1000 - 100


def test_gsop_zero_weight_turns_positive():
    """sign of 0 is +1 (physics.cu:323): potentiation makes it positive."""
    v = _variant()
    assert O.gsop_update(0, _heads(0), 40, v, 1, 0) == 80


# --------------------------------------------------------- hand-traced ticks
def tiny(n_real, lut, slots=4, a_extra=0):
    pn, nbytes = calculate_state_blob_size(n_real)
    state = np.zeros(nbytes, np.uint8)
    A = pn + a_extra
    heads = np.full(A * 8, SENT, np.uint32)
    return state, StateView(state, pn), heads, pn, A


def run(state, pn, heads, A, lut, T=1, tick0=0, dop=0, bitmask=None, virt_off=0, n_virt=0):
    return O.step_batch(state, pn, heads, A, lut, 1, tick0, T, bitmask, virt_off, n_virt, dopamine=dop)


def lut1(**kw):
    v = B.template_variant(False, spontaneous_firing_period_ticks=0, **kw)
    return B.lut_from([v])


def test_leak_toward_rest_closed_form():
    """No input: v = rest + sgn(d)*max(0, |d| - k*leak) after k ticks."""
    lut = lut1()
    state, v, heads, pn, A = tiny(2, lut)
    v.voltage[0], v.voltage[1] = 10000 + 5000, 10000 - 3200
    v.soma_to_axon[:] = 0xFFFFFFFF
    for k in range(1, 6):
        run(state, pn, heads, A, lut, tick0=k)
        assert v.voltage[0] == 10000 + max(0, 5000 - 1500 * k)
        assert v.voltage[1] == 10000 - max(0, 3200 - 1500 * k)


def test_adaptive_leak_mode_1():
    """physics.cu:209-275 with adaptive_mode 1: thr = max(0, thr - decay),
    leak = min(leak + (thr*gain)>>8, leak_max), the product wrapping in i32.
    Template: rest 10000, leak 1500, decay 20. Derived by hand:
      soma 0 (gain 64): thr 5020 -> 5000, leak 1500 + 1250 = 2750, v 16000 -> 13250
      soma 1 (gain 64): thr 20020 -> 20000, leak 1500 + 5000 -> max 4000, v 5000 -> 9000
      soma 2 (gain 4): thr 2^30 + 20 -> 2^30, 2^30 * 4 wraps to 0, leak 1500, v 12000 -> 10500"""
    mk = lambda gain: B.template_variant(False, spontaneous_firing_period_ticks=0, adaptive_mode=1,  # noqa: E731
                                         adaptive_leak_gain=gain, adaptive_leak_max=4000)
    lut = B.lut_from([mk(64), mk(4)])
    state, v, heads, pn, A = tiny(3, lut)
    v.soma_to_axon[:] = 0xFFFFFFFF
    v.flags[2] = 1 << 4
    v.voltage[:3] = [16000, 5000, 12000]
    v.threshold_offset[:3] = [5020, 20020, (1 << 30) + 20]
    run(state, pn, heads, A, lut, tick0=3)
    assert list(v.threshold_offset[:3]) == [5000, 20000, 1 << 30]
    assert list(v.voltage[:3]) == [13250, 9000, 10500]
    assert not (v.flags[:3] & 1).any()


def test_glif_spike_reset_refractory_and_push():
    """A soma over threshold fires: v <- rest, thr += penalty, timer <-
    refractory, flags = type<<4 | burst 2 | spike, own axon pushed (h0=0)
    after everyone integrated; the next 15 ticks are refractory."""
    lut = lut1()
    state, v, heads, pn, A = tiny(1, lut)
    v.voltage[0] = 40000
    v.soma_to_axon[0] = 0
    heads[0:8] = [5, SENT, SENT, SENT, SENT, SENT, SENT, SENT]
    run(state, pn, heads, A, lut, tick0=0)
    assert v.voltage[0] == 10000 and v.threshold_offset[0] == 10000 and v.timers[0] == 15
    assert v.flags[0] == (0 << 4) | (2 << 1) | 1
    assert list(heads[0:8]) == [0, 6, SENT, SENT, SENT, SENT, SENT, SENT]
    for k in range(15):
        run(state, pn, heads, A, lut, tick0=1 + k)
        assert v.timers[0] == 14 - k and v.flags[0] & 1 == 0
        assert v.threshold_offset[0] == 10000  # no decay while refractory
    run(state, pn, heads, A, lut, tick0=16)
    assert v.threshold_offset[0] == 10000 - 20


def test_heartbeat_spikes_without_reset():
    """(tick + i*104729) % period == 0 fires without touching v, thr, timer."""
    lut = B.lut_from([B.template_variant(False, spontaneous_firing_period_ticks=7)])
    state, v, heads, pn, A = tiny(3, lut)
    v.voltage[:3] = 10000
    v.soma_to_axon[:] = 0xFFFFFFFF
    fired = {i: [] for i in range(3)}
    for t in range(30):
        run(state, pn, heads, A, lut, tick0=t)
        for i in range(3):
            if v.flags[i] & 1:
                fired[i].append(t)
    for i in range(3):
        assert fired[i] == [t for t in range(30) if (t + i * 104729) % 7 == 0]
        assert v.voltage[i] == 10000 and v.timers[i] == 0 and v.threshold_offset[i] == 0


def test_snapshot_rule_same_tick_push_invisible_to_integration():
    """Neuron 1 reads seg 0 of neuron 0's axon. Neuron 0 fires at tick 0: its
    push (h0=0) is not seen by tick 0's integration but is seen by GSOP."""
    lut = lut1()
    state, v, heads, pn, A = tiny(2, lut)
    v.voltage[0] = 40000
    v.voltage[1] = 10000
    v.soma_to_axon[0] = 0
    v.soma_to_axon[1] = 0xFFFFFFFF
    v.targets[0, 1] = (0 << 24) | 1   # axon 0, seg 0
    v.weights[0, 1] = 5000
    run(state, pn, heads, A, lut, tick0=0)
    assert v.voltage[1] == 10000      # no current at tick 0
    run(state, pn, heads, A, lut, tick0=1)
    assert v.voltage[1] == 10000 + 5000 - 1500   # head at 1: (1-0) < 20


@pytest.mark.parametrize("dop", [0, 100])
def test_gsop_tick_after_a_spike(dop):
    """Soma 0 fires at tick 0 (v = 40000) and its GSOP runs on the heads
    after propagation and after its own push (physics.cu:281-373, template
    variant: pot 150, dep 60, length 20, curve 128, 120, ..., d1 = d2 = 128;
    a spiker's burst count is 2). Slots, derived by hand:
      0: extra axon 0, seg 3, w 100, head 4 -> 5: d = 2, rank 0, cs 0
         dop 0: dp = (150*128*2)>>7 = 300 -> 400; dop 100: pot 250, dp 500 -> 600
      1: extra axon 1, seg 0, w -500, no head: depression
         dop 0: dd = (60*128*2)>>7 = 120 -> -380; dop 100: dep max(0, 60-100) = 0 -> -500
      2: extra axon 2, seg 10, w 3000, head 26 -> 27: d = 17, rank 1 (in 120), cs 1
         dop 0: dp = (150*120*2)>>7 = 281, 281>>1 = 140, (140*128)>>8 = 70 -> 3070
         dop 100: dp = (250*120*2)>>7 = 468, 234, (234*128)>>8 = 117 -> 3117
      3: extra axon 3, seg 5, w 0, head 4 -> 5: d = 0; sign(0) = +1 -> 300 (dop 100: 500)
      4: its own axon, seg 0, w 100, no head before the tick: the spike's push
         (h = 0) comes after integration but before GSOP: d = 0 -> 400 (600)"""
    lut = lut1()
    state, v, heads, pn, A = tiny(1, lut, a_extra=8)
    v.voltage[0] = 40000
    v.soma_to_axon[0] = 0
    x = [pn + k for k in range(4)]
    for slot, (ax, seg, w) in enumerate([(x[0], 3, 100), (x[1], 0, -500), (x[2], 10, 3000), (x[3], 5, 0),
                                         (0, 0, 100)]):
        v.targets[slot, 0] = (seg << 24) | (ax + 1)
        v.weights[slot, 0] = w
    heads[x[0] * 8] = 4
    heads[x[2] * 8] = 26
    heads[x[3] * 8] = 4
    run(state, pn, heads, A, lut, tick0=0, dop=dop)
    assert v.flags[0] == (2 << 1) | 1 and v.voltage[0] == 10000
    want = [400, -380, 3070, 300, 400] if dop == 0 else [600, -500, 3117, 500, 600]
    assert [int(v.weights[s, 0]) for s in range(5)] == want
    assert heads[0] == 0  # the push


def test_inputs_push_virtual_axons():
    """Bit i of the mask pushes axon virt_off+i before propagation."""
    lut = lut1()
    state, v, heads, pn, A = tiny(1, lut, a_extra=64)
    v.soma_to_axon[:] = 0xFFFFFFFF
    mask = np.zeros((1, 2), np.uint32)
    mask[0, 0] = 1 << 3
    mask[0, 1] = 1 << 0
    run(state, pn, heads, A, lut, bitmask=mask, virt_off=pn, n_virt=40)
    h = heads.reshape(-1, 8)
    assert h[pn + 3, 0] == 1 and h[pn + 32, 0] == 1
    assert (h[pn + 4] == SENT).all()


def test_propagation_reference_vectors():
    """test_tick.rs:134-147: heads advance by v_seg, the sentinel stays."""
    lut = lut1()
    state, v, heads, pn, A = tiny(1, lut)
    v.soma_to_axon[:] = 0xFFFFFFFF
    heads[0:3] = [0, 5, 10]
    heads[8:10] = [0, SENT]
    run(state, pn, heads, A, lut)
    assert list(heads[0:3]) == [1, 6, 11] and heads[9] == SENT and heads[8] == 1


# ----------------------------------------------------------------- sort&prune
def brute_sort_prune(t, w, m, prune):
    """Python stable sort by (filled, |w|) descending, then prune."""
    order = sorted(range(128), key=lambda s: (-(1 if t[s] else 0), -(abs(int(w[s])) if t[s] else 0), s))
    t2, w2, m2 = t[order].copy(), w[order].copy(), m[order].copy()
    for s in range(128):
        if t2[s] != 0 and abs(int(w2[s])) < prune:
            t2[s], w2[s], m2[s] = 0, 0, 0
    return t2, w2, m2


def test_sort_and_prune_brute_force():
    rng = np.random.default_rng(0)
    n = 64
    pn, nbytes = calculate_state_blob_size(n)
    state = np.zeros(nbytes, np.uint8)
    v = StateView(state, pn)
    v.targets[:] = rng.integers(1, 1 << 30, size=v.targets.shape).astype(np.uint32)
    v.targets[rng.random(v.targets.shape) < 0.4] = 0
    v.weights[:] = rng.choice([-40, -15, -14, 0, 14, 15, 16, 300, -32768, 32767], size=v.weights.shape)
    v.dendrite_timers[:] = rng.integers(0, 256, size=v.dendrite_timers.shape)
    v.flags[:] = rng.integers(0, 256, size=pn)
    want = [brute_sort_prune(v.targets[:, i].copy(), v.weights[:, i].copy(), v.dendrite_timers[:, i].copy(), 15)
            for i in range(pn)]
    flags0 = v.flags.copy()
    O.sort_and_prune_blob(state, pn, 15)
    for i in range(pn):
        np.testing.assert_array_equal(v.targets[:, i], want[i][0])
        np.testing.assert_array_equal(v.weights[:, i], want[i][1])
        np.testing.assert_array_equal(v.dendrite_timers[:, i], want[i][2])
    np.testing.assert_array_equal(v.flags, flags0 & 0xF1)


def test_sort_and_prune_invariants_and_idempotence():
    rng = np.random.default_rng(1)
    pn, nbytes = calculate_state_blob_size(200)
    state = np.zeros(nbytes, np.uint8)
    v = StateView(state, pn)
    v.targets[:] = rng.integers(1, 1 << 30, size=v.targets.shape).astype(np.uint32)
    v.targets[rng.random(v.targets.shape) < 0.5] = 0
    v.weights[:] = rng.integers(-500, 500, size=v.weights.shape)
    before = sorted(map(tuple, np.stack([v.targets, v.weights], -1).reshape(-1, 2).tolist()))
    O.sort_and_prune_blob(state, pn, 0)
    after = sorted(map(tuple, np.stack([v.targets, v.weights], -1).reshape(-1, 2).tolist()))
    assert before == after  # prune 0 removes nothing: a permutation per neuron overall
    filled = v.targets != 0
    for i in range(pn):
        f = filled[:, i]
        k = int(f.sum())
        assert f[:k].all() and not f[k:].any()
        a = np.abs(v.weights[:k, i].astype(np.int32))
        assert (a[:-1] >= a[1:]).all()
    s2 = state.copy()
    O.sort_and_prune_blob(state, pn, 0)
    np.testing.assert_array_equal(state, s2)


# ---- Multi-zone helpers (oracle/zone_oracle.py), hand-derived from
# bindings.cu:341-418 and ring_buffer.rs:35-54.
def test_zone_extract_known_answers():
    from oracle import zone_oracle as Z
    heads = np.full((3, 8), 0x80000000, np.uint32)
    heads[0] = [0, 1, 49, 50, 0x6FFFFFFF, 0x70000000, 0x80000000, 7]
    heads[1] = [2, 0x80000000, 0x80000000, 0x80000000, 0x80000000, 0x80000000, 0x80000000, 0x80000000]
    # v_seg 1, T 50: heads 0, 1, 49, 7 spiked 0, 1, 49, 7 ticks ago -> offsets 49, 48, 0, 42;
    # 50 is one tick before the batch; 0x6FFFFFFF is too old; 0x70000000 and the sentinel are dead.
    ev = Z.extract_outgoing_spikes(heads, [0, 0x80000000, 0xFFFFFFFF, 1], [900, 901, 902, 903], 50, 1)
    assert ev == [(900, 49), (900, 48), (900, 0), (900, 42), (903, 47)]
    # v_seg 3: ticks since the spike are head // 3 (0, 0, 16, 16, -, -, -, 2).
    ev3 = Z.extract_outgoing_spikes(heads, [0], [5], 50, 3)
    assert ev3 == [(5, 49), (5, 49), (5, 33), (5, 33), (5, 47)]
    # A batch of one tick keeps only the heads pushed in it.
    assert Z.extract_outgoing_spikes(heads, [0, 1], [1, 2], 1, 1) == [(1, 0)]


def test_zone_schedule_known_answers():
    from oracle import zone_oracle as Z
    ids, cnt = Z.schedule_from_events([(5, 0), (6, 0), (7, 49), (8, 50), (9, 0), (10, 3)], 50, max_spikes=2)
    assert cnt.tolist()[:4] == [2, 0, 0, 1] and cnt[49] == 1 and cnt.sum() == 4
    assert ids[0, :2].tolist() == [5, 6] and ids[49, 0] == 7 and ids[3, 0] == 10  # (9, 0) over the cap, (8, 50) past T


def test_zone_ghost_sync_skips_poisoned_sources():
    from oracle import zone_oracle as Z
    src = np.arange(4 * 8, dtype=np.uint32).reshape(4, 8)
    dst = np.full((5, 8), 7, np.uint32)
    Z.ghost_sync(src, dst, [2, 0x80000000, 0xFFFFFFFF, 0], [4, 0, 1, 2])
    assert dst[4].tolist() == list(range(16, 24)) and dst[2].tolist() == list(range(0, 8))
    assert (dst[[0, 1, 3]] == 7).all()
