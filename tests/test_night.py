"""Night-Phase sprouting (axc_sprout_pass, csrc/night_host.cpp) against the
pure-Python oracle of genesis-baker/src/bake/sprouting.rs, and the invariants
the pass must keep at C1 scale."""
import numpy as np

from axicor_amd import brain as B
from axicor_amd import night as NT
from oracle import sprout_oracle as SO


def small_brain(seed, n=200, slots=20, axon_len=(1, 10), n_in=32):
    spec = B.spec_small(n, slots, n_in, seed=seed, axon_len=axon_len)
    b = B.generate(spec, with_paths=True)
    v = b.view()
    rng = np.random.default_rng(seed)
    # Empty tails of varying length so somas have slots to fill, some holes.
    for i in range(spec.n_neurons):
        k = int(rng.integers(slots // 2, slots + 1))
        v.targets[k:, i] = 0
        v.weights[k:, i] = 0
    # Spiking at the last tick (burst 2) for a third of the somas.
    v.flags[:spec.n_neurons] = (v.flags[:spec.n_neurons] & 0xF0) | np.where(rng.random(spec.n_neurons) < 0.35, 0x05, 0)
    return b


def with_ghosts(g, n_ghosts, at):
    """Geometry with n_ghosts extra axon rows inserted at index `at` (the
    ghost range [padded_n, padded_n + ghosts) of the reference layout)."""
    ins = lambda a, fill: np.ascontiguousarray(np.insert(a, at, fill, axis=0))  # noqa: E731
    return NT.Geometry(ins(g.lengths, np.zeros(n_ghosts, np.uint8)), ins(g.paths, np.zeros((n_ghosts, 256), np.uint32)),
                       ins(g.tips, np.zeros(n_ghosts, np.uint32)), ins(g.dirs, np.zeros(n_ghosts, np.uint32)),
                       g.soma_positions.copy(), n_ghosts, None, g.max_x, g.max_y)


def run_both(b, types, epoch=7, ghosts=0, incoming=None, ghost_origins=None, max_sprouts=4, geo=None, seed=None):
    seed = b.spec.seed if seed is None else seed
    v = b.view()
    t1, w1 = v.targets.copy(), v.weights.copy()
    g1 = geo if geo is not None else NT.geometry_of(b)
    g1 = NT.Geometry(g1.lengths.copy(), g1.paths.copy(), g1.tips.copy(), g1.dirs.copy(), g1.soma_positions.copy(),
                     ghosts, ghost_origins, g1.max_x, g1.max_y)
    t0 = t1.copy()
    changed = np.empty(b.padded_n * max(1, max_sprouts), np.uint32)
    res = NT.sprout_pass(t1, w1, v.flags.copy(), v.soma_to_axon.copy(), b.padded_n, g1, types, epoch,
                         seed, zone_hash=0xABCD, max_sprouts=max_sprouts, incoming=incoming, changed=changed)
    # The listed columns are exactly the slots whose targets changed (each an
    # empty slot that got a synapse), each once.
    cols = res.changed_cols
    assert len(cols) == res.new_synapses and len(np.unique(cols)) == len(cols)
    np.testing.assert_array_equal(np.sort(cols), np.flatnonzero((t1 != t0).reshape(-1)))
    assert (t0.reshape(-1)[cols] == 0).all() and (t1.reshape(-1)[cols] != 0).all()
    t2, w2 = v.targets.copy(), v.weights.copy()
    g2 = geo if geo is not None else NT.geometry_of(b)
    g2 = NT.Geometry(g2.lengths.copy(), g2.paths.copy(), g2.tips.copy(), g2.dirs.copy(), g2.soma_positions.copy(),
                     ghosts, ghost_origins, g2.max_x, g2.max_y)
    tl = None if types is None else [dict(zip(types.dtype.names, x)) for x in types.tolist()]
    inc = [] if incoming is None else [dict(zip(incoming.dtype.names, x)) for x in incoming.tolist()]
    new, hand, acks, prunes = SO.sprout_pass(t2, w2, v.flags.copy(), v.soma_to_axon.copy(), b.padded_n, g2.lengths,
                                             g2.paths, g2.tips, g2.dirs, g2.soma_positions, tl, epoch, seed,
                                             g2.max_x, g2.max_y, ghosts, ghost_origins, inc, 0xABCD, max_sprouts)
    assert res.new_synapses == new
    np.testing.assert_array_equal(t1, t2)
    np.testing.assert_array_equal(w1, w2)
    for f in ("lengths", "paths", "tips", "dirs"):
        np.testing.assert_array_equal(getattr(g1, f), getattr(g2, f), err_msg=f)
    assert len(res.handovers) == len(hand) and len(res.acks) == len(acks) and len(res.prunes) == len(prunes)
    for got, want in zip(res.handovers, hand):
        for k, val in want.items():
            assert int(got[k]) == int(val), (k, got, want)
    for got, want in zip(res.acks, acks):
        assert tuple(int(x) for x in got) == tuple(want)
    for got, want in zip(res.prunes, prunes):
        assert tuple(int(x) for x in got) == tuple(want)
    return res, t1, w1


def test_sprout_matches_oracle_template_scores():
    b = small_brain(1)
    res, _, _ = run_both(b, NT.sprout_types_from_lut(b.lut))
    assert res.new_synapses > 0


def test_sprout_matches_oracle_without_blueprints():
    """No type table: every score is 1.0 (first candidate wins) and +74."""
    b = small_brain(2)
    res, _, _ = run_both(b, None)
    assert res.new_synapses > 0


def test_sprout_matches_oracle_type_affinity_and_limits():
    b = small_brain(3)
    types = NT.sprout_types_from_lut(b.lut, w_distance=0.2, w_power=0.3, w_explore=0.3, w_type=0.7)
    run_both(b, types, epoch=12345, max_sprouts=2)


def test_sprout_ghosts_nudge_handover_and_prune():
    """Ghost axons: absorbed handovers, nudging off the grid (handovers out),
    orphan ghosts reported for pruning."""
    b = small_brain(4, n=192, n_in=0)
    N = b.padded_n
    ghosts = 16
    geo = with_ghosts(NT.geometry_of(b), ghosts, N)
    rng = np.random.default_rng(0)
    # Half of the ghost slots hold axons already, pointing off the grid.
    for k in range(0, ghosts, 2):
        a = N + k
        geo.tips[a] = int(rng.integers(0, 7)) | (int(rng.integers(0, 7)) << 10) | (1 << 20)
        geo.dirs[a] = 0x90 | (0x10 << 8) | (0x70 << 16)
        geo.lengths[a] = 1
        geo.paths[a, 0] = geo.tips[a]
    inc = np.zeros(3, NT.HANDOVER_DTYPE)
    inc["origin_zone_hash"] = [11, 12, 13]
    inc["local_axon_id"] = [5, 6, 7]
    inc["entry_x"], inc["entry_y"], inc["entry_z"] = 2, 3, 4
    inc["vector_x"] = 100
    origins = np.arange(1, ghosts + 1, dtype=np.uint32)
    res, _, _ = run_both(b, NT.sprout_types_from_lut(b.lut), ghosts=ghosts, incoming=inc, ghost_origins=origins,
                         geo=geo)
    assert len(res.acks) == 3 and len(res.handovers) > 0 and len(res.prunes) > 0


def test_sprout_invariants_c1():
    """C1 brain after a Night sort & prune: new synapses only in empty tail
    slots of somas with flags & 0x0E, at most max_sprouts each, never the
    soma's own axon, never an axon already present, sign by presynaptic type."""
    from oracle import oracle as O
    b = B.generate(B.spec_c1(), with_paths=True)
    inp = b.inputs(0, 250)
    O.step_batch(b.state, b.padded_n, b.heads, b.total_axons, b.lut, 1, 0, 250, inp, b.virtual_offset, b.n_inputs)
    v = b.view()
    flags = v.flags.copy()
    O.sort_and_prune_blob(b.state, b.padded_n, 15)
    t0, w0 = v.targets.copy(), v.weights.copy()
    t1, w1 = t0.copy(), w0.copy()
    types = NT.sprout_types_from_lut(b.lut)
    res = NT.sprout_pass(t1, w1, flags, v.soma_to_axon.copy(), b.padded_n, NT.geometry_of(b), types, 3, b.spec.seed)
    changed = t1 != t0
    assert res.new_synapses == int(changed.sum()) > 0
    cols = np.nonzero(changed.any(0))[0]
    assert ((flags[cols] & 0x0E) != 0).all()
    for i in cols:
        new_slots = np.nonzero(changed[:, i])[0]
        assert len(new_slots) <= 4
        assert (t0[new_slots, i] == 0).all()
        old = set((((t0[t0[:, i] != 0, i]) & 0xFFFFFF) - 1).tolist())
        new = (((t1[new_slots, i]) & 0xFFFFFF) - 1).tolist()
        assert len(set(new)) == len(new) and not (set(new) & old)   # rule of uniqueness (sprouting.rs:264-273)
        assert v.soma_to_axon[i] not in set(new)                     # no self connection (262)
        for s in new_slots:
            a = int(t1[s, i] & 0xFFFFFF) - 1
            src_type = int(b.paths[a, int(t1[s, i]) >> 24]) >> 28
            want = int(types[src_type]["initial_synapse_weight"])
            assert w1[s, i] == (-want if types[src_type]["is_inhibitory"] else want)


_SPROUT_CHILD = r"""
import sys, hashlib
import numpy as np
sys.path.insert(0, sys.argv[1])
from axicor_amd import brain as B, night as NT
spec = B.spec_small(6000, 48, 300, seed=17, axon_len=(1, 40))
b = B.generate(spec, with_paths=True)
v = b.view()
rng = np.random.default_rng(17)
for i in range(spec.n_neurons):
    k = int(rng.integers(24, 49))
    v.targets[k:, i] = 0
    v.weights[k:, i] = 0
v.flags[:spec.n_neurons] = (v.flags[:spec.n_neurons] & 0xF0) | np.where(rng.random(spec.n_neurons) < 0.3, 0x05, 0)
t, w = v.targets.copy(), v.weights.copy()
res = NT.sprout_pass(t, w, v.flags.copy(), v.soma_to_axon.copy(), b.padded_n, NT.geometry_of(b),
                     NT.sprout_types_from_lut(b.lut), 5, b.spec.seed)
print(res.new_synapses, hashlib.sha256(t.tobytes() + w.tobytes()).hexdigest())
"""


def test_sprout_same_result_for_any_thread_count():
    """The segment grid is built and each slot's candidate scan is split over
    OpenMP threads; the merge keeps the sequential choice, so 1, 3 and 8
    threads give the same targets and weights (6,000 somas, about 30% of
    them bursting, with empty tails), and with 8 asked for under a limit of 2
    (OMP_THREAD_LIMIT)."""
    import os
    import subprocess
    import sys
    repo = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    outs = []
    for n, limit in ((1, None), (3, None), (8, None), (8, 2)):
        env = dict(os.environ, OMP_NUM_THREADS=str(n))
        env.pop("OMP_THREAD_LIMIT", None)
        if limit:
            env["OMP_THREAD_LIMIT"] = str(limit)
        r = subprocess.run([sys.executable, "-c", _SPROUT_CHILD, repo], env=env, capture_output=True, text=True,
                           timeout=300, check=True)
        outs.append(r.stdout.strip())
    assert int(outs[0].split()[0]) > 0
    assert len(set(outs)) == 1, outs


def test_sprout_matches_oracle_over_seeds_and_settings():
    """Random scenes (a third of the somas bursting, empty tails) with the
    power term on: the host pass, whose candidates with an earlier bursting
    owner are rescored in index order, equals the oracle's sequential scan
    for max_sprouts 0 (one sprout), 1, 3 and 7, with negative and zero
    score weights among the settings."""
    settings = [dict(w_distance=0.5, w_power=0.4, w_explore=0.1, w_type=0.0),
                dict(w_distance=0.0, w_power=1.0, w_explore=0.0, w_type=0.0),
                dict(w_distance=0.1, w_power=-0.7, w_explore=0.3, w_type=0.2),
                dict(w_distance=-2.0, w_power=3.0, w_explore=0.5, w_type=-0.5)]
    for seed in range(3):
        b = small_brain(40 + seed, n=240, slots=24)
        for k, st in enumerate(settings):
            types = NT.sprout_types_from_lut(b.lut, **st)
            run_both(b, types, epoch=seed * 7 + k, max_sprouts=(0, 1, 3, 7)[(seed + k) % 4])


def test_sprout_matches_oracle_when_the_explore_noise_varies():
    """The explore noise random_f32(seed + axon + epoch) is the top 23 bits of
    the sum, so over one shard it takes at most two values, or jumps from
    nearly 1 to 0 where the sum wraps past 2^64. Random scenes with such
    seeds and large explore weights: equal to the oracle's sequential scan.
    (test_sprout_kats.py places a scene where the cell bounds decide.)"""
    b = small_brain(77, n=240, slots=24)
    A = b.total_axons
    seeds = [2**64 - A // 2,            # the sum wraps in the middle of the axons
             5 * 2**41 - A // 3,        # two noise values, 2^-23 apart
             2**64 - 3]                 # wraps at once: nearly every axon has noise 0
    for k, seed in enumerate(seeds):
        for st in (dict(w_distance=0.5, w_power=0.4, w_explore=2.0e5, w_type=0.0),
                   dict(w_distance=0.1, w_power=0.9, w_explore=-2.0e6, w_type=0.3)):
            types = NT.sprout_types_from_lut(b.lut, **st)
            res, _, _ = run_both(b, types, epoch=k, seed=seed, max_sprouts=(1, 3, 4)[k])
            assert res.new_synapses > 0


def test_sprout_rejects_more_axons_than_a_target_holds():
    """A target holds axon + 1 in 24 bits (layout.rs:268-271): a pass over
    2^24 axons or more is refused before it reads the geometry."""
    import pytest
    A = 1 << 24
    geo = NT.Geometry(np.zeros(A, np.uint8), np.zeros((1, 256), np.uint32), np.zeros(A, np.uint32),
                      np.zeros(A, np.uint32), np.zeros(64, np.uint32), 0, None, 1024, 1024)
    with pytest.raises(ValueError):
        NT.sprout_pass(np.zeros((128, 64), np.uint32), np.zeros((128, 64), np.int16), np.zeros(64, np.uint8),
                       np.zeros(64, np.uint32), 64, geo, None, 0, 1)
