"""Known answers for the Night-Phase sprouting pass, derived by hand from
genesis-baker/src/bake/sprouting.rs:32-378 and spatial_grid.rs:99-185, run on
both the pure-Python oracle (oracle/sprout_oracle.py) and the host pass
(axc_sprout_pass, csrc/night_host.cpp).

Each scene is a handful of axons placed by hand on a tiny grid, so that the
expected new targets follow from the reference's rules alone:

- candidates are visited cell by cell, z then y then x ascending over the
  radius of 2 cells (spatial_grid.rs:168-183); inside a cell the segments
  are in reverse axon order, each axon's in segment order (104-123);
- a candidate replaces the best only on a strictly greater score
  (sprouting.rs:301);
- the soma's own axon and axons already on the soma are skipped (262-273);
- an axon with no owner soma has power 1.0 (285-290);
- slots are filled from 127 downward while empty, at most max_sprouts
  (251-255, 331-333);
- the explore noise of axon a is random_f32(master_seed + a + epoch)
  (280-282), random_f32(s) = (s >> 41) / 2^23 (genesis-core/src/seed.rs:61-64);
- nudge_axon moves a tip one voxel along each axis whose direction byte
  exceeds +-64, and hands the axon over past the shard's x/y range or z > 63
  (sprouting.rs:51-117).

Plus the reference's own vectors: random_f32's range over entity seeds
(genesis-core/src/test_seed.rs:46-52) and PackedPosition unpacking
(genesis-baker/src/bake/test_sprouting.rs:1-16, genesis-core/src/types.rs:84-105).
"""
import numpy as np
import pytest

from axicor_amd import night as NT
from oracle import oracle as O
from oracle import sprout_oracle as SO

N = 64          # somas (padded); soma i owns axon i
ZONE = 0x5A5A


def pack_pos(x, y, z, t):
    """PackedPosition::pack_raw (genesis-core/src/types.rs:22-29)."""
    return (x & 0x3FF) | ((y & 0x3FF) << 10) | ((z & 0xFF) << 20) | ((t & 0xF) << 28)


def pack_target(axon, seg):
    """pack_dendrite_target (genesis-core/src/layout.rs:254-260)."""
    return (seg << 24) | ((axon + 1) & 0xFFFFFF)


def f32_closed_form(seed):
    """random_f32 from its definition: the top 23 bits of the seed as the
    fraction of a float in [1, 2), minus 1, i.e. exactly (seed >> 41) / 2^23."""
    return np.float32(((seed & 0xFFFFFFFFFFFFFFFF) >> 41) / float(1 << 23))


class Scene:
    """A tiny shard: N somas, N local axons (+ ghosts), empty dendrites, no
    segments; tests place segments, positions and flags by hand."""

    def __init__(self, ghosts=0, max_x=64, max_y=64):
        self.ghosts = ghosts
        A = N + ghosts
        self.targets = np.zeros((128, N), np.uint32)
        self.weights = np.zeros((128, N), np.int16)
        self.flags = np.zeros(N, np.uint8)
        self.s2a = np.arange(N, dtype=np.uint32)
        self.lengths = np.zeros(A, np.uint8)
        self.paths = np.zeros((A, 256), np.uint32)
        self.tips = np.zeros(A, np.uint32)
        self.dirs = np.zeros(A, np.uint32)
        self.pos = np.zeros(N, np.uint32)
        self.origins = None
        self.max_x, self.max_y = max_x, max_y

    def segments(self, axon, cells_xyz, t=0):
        """Axon `axon` gets one segment per (x, y, z) voxel, in order."""
        for k, (x, y, z) in enumerate(cells_xyz):
            self.paths[axon, k] = pack_pos(x, y, z, t)
        self.lengths[axon] = len(cells_xyz)
        self.tips[axon] = self.paths[axon, len(cells_xyz) - 1]

    def soma(self, i, xyz, t=0, burst=2):
        self.pos[i] = pack_pos(*xyz, t)
        self.flags[i] = (t << 4) | (burst << 1)

    def run(self, types, epoch=0, seed=0, max_sprouts=4, incoming=None):
        """The host pass and the Python oracle on copies; both must agree,
        and the host result is returned for the known-answer checks."""
        out = []
        for which in ("host", "oracle"):
            t, w = self.targets.copy(), self.weights.copy()
            g = NT.Geometry(self.lengths.copy(), self.paths.copy(), self.tips.copy(), self.dirs.copy(),
                            self.pos.copy(), self.ghosts, self.origins, self.max_x, self.max_y)
            if which == "host":
                res = NT.sprout_pass(t, w, self.flags.copy(), self.s2a.copy(), N, g, types, epoch, seed,
                                     zone_hash=ZONE, max_sprouts=max_sprouts, incoming=incoming)
                out.append((res.new_synapses, t, w, g, [dict(zip(h.dtype.names, h.tolist())) for h in res.handovers],
                            [tuple(int(v) for v in a) for a in res.acks]))
            else:
                tl = None if types is None else [dict(zip(types.dtype.names, x)) for x in types.tolist()]
                inc = [] if incoming is None else [dict(zip(incoming.dtype.names, x)) for x in incoming.tolist()]
                new, hand, acks, _ = SO.sprout_pass(t, w, self.flags.copy(), self.s2a.copy(), N, g.lengths, g.paths,
                                                    g.tips, g.dirs, g.soma_positions, tl, epoch, seed, self.max_x,
                                                    self.max_y, self.ghosts, self.origins, inc, ZONE, max_sprouts)
                out.append((new, t, w, g, hand, acks))
        (n1, t1, w1, g1, h1, a1), (n2, t2, w2, g2, h2, a2) = out
        assert n1 == n2
        np.testing.assert_array_equal(t1, t2)
        np.testing.assert_array_equal(w1, w2)
        for f in ("lengths", "paths", "tips", "dirs"):
            np.testing.assert_array_equal(getattr(g1, f), getattr(g2, f), err_msg=f)
        assert len(h1) == len(h2) and a1 == [tuple(a) for a in a2]
        for x, y in zip(h1, h2):
            for k, v in y.items():
                assert int(x[k]) == int(v), (k, x, y)
        return n1, t1, w1, g1, h1, a1


def types_table(w_distance=0.0, w_power=0.0, w_explore=0.0, w_type=0.0, weights=(74, 90), inhibitory=(0, 1)):
    t = np.zeros(16, NT.SPROUT_TYPE_DTYPE)
    t["w_distance"], t["w_power"], t["w_explore"], t["w_type"] = w_distance, w_power, w_explore, w_type
    t["initial_synapse_weight"][:2] = weights
    t["is_inhibitory"][:2] = inhibitory
    return t


# --------------------------------------------------------------------------
# The reference's own vectors
# --------------------------------------------------------------------------
def test_random_f32_range_over_entity_seeds():
    """test_seed.rs:46-52: random_f32(entity_seed(42, i)) in [0, 1) for
    i < 10,000; and each value is exactly (seed >> 41) / 2^23."""
    L = O.lib()
    for i in range(10_000):
        s = L.orc_entity_seed(42, i)
        f = SO.random_f32(s)
        assert 0.0 <= f < 1.0
        assert f == f32_closed_form(s)


def test_random_f32_bit_patterns():
    """Closed form at the edges of the 23-bit fraction: seeds whose top 23
    bits are 0, 1, 2^22 and 2^23 - 1 (low bits arbitrary, they are dropped),
    and the wrapping seed sums the sprouting pass forms."""
    cases = {0: 0.0, (1 << 41) - 1: 0.0, 1 << 41: 2.0 ** -23, (1 << 41) | 0x1FFFFFFFFFF: 2.0 ** -23,
             1 << 63: 0.5, 0xFFFFFFFFFFFFFFFF: 1.0 - 2.0 ** -23}
    for s, want in cases.items():
        got = SO.random_f32(s)
        assert got == np.float32(want), (hex(s), got, want)
        assert got == f32_closed_form(s)
    # master_seed + axon + epoch wraps at 2^64 (sprouting.rs:280-282).
    s = (0xFFFFFFFFFFFFFFF0 + 0x20 + 3) & 0xFFFFFFFFFFFFFFFF
    assert SO.random_f32(0xFFFFFFFFFFFFFFF0 + 0x20 + 3) == f32_closed_form(s) == np.float32(0.0)


@pytest.mark.parametrize("x,y,z,t", [(512, 256, 100, 2), (1023, 1023, 255, 15), (0, 0, 0, 0), (123, 1000, 48, 9)])
def test_position_unpacking(x, y, z, t):
    """test_sprouting.rs:1-16 and types.rs:84-105."""
    p = pack_pos(x, y, z, t)
    if (x, y, z, t) == (1023, 1023, 255, 15):
        assert p == 0xFFFFFFFF
    if (x, y, z, t) == (0, 0, 0, 0):
        assert p == 0
    assert SO.pos_xyzt(p) == (x, y, z, t)


# --------------------------------------------------------------------------
# Candidate choice
# --------------------------------------------------------------------------
def test_tie_in_one_cell_goes_to_the_higher_axon():
    """Equal scores in one cell: the cell lists axon 9 before axon 5 (reverse
    axon order), and 5 does not replace 9 (strict >)."""
    s = Scene()
    s.soma(0, (10, 10, 10))
    s.segments(5, [(10, 10, 10)])
    s.segments(9, [(11, 11, 11)])        # same 2-voxel cell (5, 5, 5)
    n, t, w, *_ = s.run(types_table(w_distance=0.5), max_sprouts=1)
    assert n == 1
    assert t[127, 0] == pack_target(9, 0) and w[127, 0] == 74
    assert (t[:127] == 0).all()


def test_tie_across_cells_goes_to_the_first_cell_visited():
    """Equal scores in two cells: z ascending, so axon 9's cell at z = 4 is
    visited before axon 5's at z = 6 and keeps the slot."""
    s = Scene()
    s.soma(0, (10, 10, 10))
    s.segments(5, [(10, 10, 12)])        # cell z = 6
    s.segments(9, [(10, 10, 8)])         # cell z = 4
    n, t, *_ = s.run(types_table(w_distance=0.5), max_sprouts=1)
    assert n == 1 and t[127, 0] == pack_target(9, 0)
    # Same cells, axons swapped: the lower axon in the earlier cell wins.
    s = Scene()
    s.soma(0, (10, 10, 10))
    s.segments(9, [(10, 10, 12)])
    s.segments(5, [(10, 10, 8)])
    n, t, *_ = s.run(types_table(w_distance=0.5), max_sprouts=1)
    assert n == 1 and t[127, 0] == pack_target(5, 0)


def test_first_segment_of_an_axon_in_the_first_cell():
    """Axon 9's segments: 0 in cell z = 6, 1 and 2 in cell z = 4. Cell z = 4
    is visited first and lists segment 1 before 2: the new target is segment
    1."""
    s = Scene()
    s.soma(0, (10, 10, 10))
    s.segments(9, [(10, 10, 12), (10, 10, 8), (11, 10, 9)])
    n, t, *_ = s.run(types_table(w_distance=0.5), max_sprouts=1)
    assert n == 1 and t[127, 0] == pack_target(9, 1)


def test_radius_is_two_cells():
    """Soma in cell (5,5,5): a segment in cell z = 8 (3 cells away) is out of
    reach, one in z = 7 is in."""
    s = Scene()
    s.soma(0, (10, 10, 10))
    s.segments(20, [(10, 10, 16)])      # cell z = 8
    n, t, *_ = s.run(types_table(w_distance=0.5))
    assert n == 0 and (t == 0).all()
    s.segments(21, [(10, 10, 15)])      # cell z = 7
    n, t, *_ = s.run(types_table(w_distance=0.5))
    assert n == 1 and t[127, 0] == pack_target(21, 0)


def test_self_and_duplicate_guards_and_sign():
    """Soma 0's own axon 0 and axon 3 (already in slot 0) would score 1.5
    (same type); axon 7 (type 1, inhibitory) scores 0.5 and is the only one
    allowed. The next slot then has no candidate left, so the soma stops
    after one sprout. The weight is -90: type 1's initial weight, negated."""
    s = Scene()
    s.soma(0, (10, 10, 10), t=0)
    s.targets[0, 0] = pack_target(3, 0)
    s.weights[0, 0] = 200
    s.segments(0, [(10, 10, 10)], t=0)
    s.segments(3, [(10, 11, 10)], t=0)
    s.segments(7, [(12, 12, 12)], t=1)
    n, t, w, *_ = s.run(types_table(w_distance=0.5, w_type=1.0))
    assert n == 1
    assert t[127, 0] == pack_target(7, 0) and w[127, 0] == -90
    assert t[0, 0] == pack_target(3, 0) and (t[1:127, 0] == 0).all()


def test_ghost_axon_has_power_one():
    """w_power only. Local axon 5's owner has sum |w| = 128 * 32767 / 2
    (power 0.5); ghost axon 64 has no owner (power 1.0) and is visited after
    it, so it wins only through the owner-less power of 1.0."""
    s = Scene(ghosts=1)
    s.soma(0, (10, 10, 10))
    s.segments(5, [(10, 10, 8)])          # cell z = 4, first
    s.segments(N, [(10, 10, 12)])         # ghost, cell z = 6, second
    s.weights[:, 5] = 16384               # 128 * 16384 = 2^21
    s.weights[0, 5] = 16384 - 64          # sum = 2^21 - 64 = 128 * 32767 / 2
    n, t, *_ = s.run(types_table(w_power=1.0), max_sprouts=1)
    assert n == 1 and t[127, 0] == pack_target(N, 0)
    # An owner with full power (every slot at 32767) ties with the ghost and,
    # visited first, keeps the slot.
    s.weights[:, 5] = 32767
    n, t, *_ = s.run(types_table(w_power=1.0), max_sprouts=1)
    assert n == 1 and t[127, 0] == pack_target(5, 0)


def test_power_index_prefers_the_stronger_owner():
    """Axon 5's owner: sum |w| = 1000; axon 9's owner: 1001 (negative weights
    count by magnitude). 9 is visited second and wins by strict >."""
    s = Scene()
    s.soma(0, (10, 10, 10))
    s.segments(5, [(10, 10, 8)])
    s.segments(9, [(10, 10, 12)])
    s.weights[0, 5], s.weights[1, 5] = 600, -400
    s.weights[0, 9], s.weights[5, 9] = -1000, 1
    n, t, *_ = s.run(types_table(w_power=1.0), max_sprouts=1)
    assert n == 1 and t[127, 0] == pack_target(9, 0)


def test_max_sprouts_fills_from_the_tail():
    """Slots 0-9 filled, six candidates of equal score in one cell: with
    max_sprouts = 4, slots 127..124 take the four highest axons (each pick is
    the first in the reverse-order cell not yet present)."""
    s = Scene()
    s.soma(0, (10, 10, 10))
    for k in range(10):
        s.targets[k, 0] = pack_target(40 + k, 0)
        s.weights[k, 0] = 100
    for a in (11, 12, 13, 14, 15, 16):
        s.segments(a, [(10, 10, 10)])
    n, t, *_ = s.run(types_table(w_distance=0.5), max_sprouts=4)
    assert n == 4
    assert [int(t[k, 0]) for k in (127, 126, 125, 124)] == [pack_target(a, 0) for a in (16, 15, 14, 13)]
    assert (t[10:124, 0] == 0).all()


def test_only_bursting_placed_somas_sprout():
    """flags & 0x0E == 0, or position 0, means no sprouting (sprouting.rs:240-243)."""
    s = Scene()
    s.segments(9, [(10, 10, 10)])
    s.soma(0, (10, 10, 10), burst=0)
    s.soma(1, (10, 10, 10), burst=1)
    s.pos[2] = 0
    s.flags[2] = 0x04
    n, t, *_ = s.run(types_table(w_distance=0.5))
    assert n == 1 and t[127, 1] == pack_target(9, 0) and (t[:, [0, 2]] == 0).all()


def test_explore_noise_orders_the_picks():
    """w_explore only. The noise keeps the top 23 bits of seed + axon +
    epoch, so axons of one shard share at most two values: with the seed
    placed so that seed + a + epoch crosses 5 * 2^41 at a = 24, axons 20-23
    score 4 / 2^23 and axons 24-27 score 5 / 2^23. Eight cells visited in
    axon order (z, then y, then x ascending): the four picks are 24, 25, 26,
    27 (the higher value first, ties to the earlier cell); shifting the epoch
    by 2 moves the boundary to a = 22."""
    epoch = 77
    seed = (5 << 41) - 24 - epoch
    s = Scene()
    s.soma(0, (10, 10, 10))
    axons = [20, 21, 22, 23, 24, 25, 26, 27]
    for k, a in enumerate(axons):
        s.segments(a, [(8 + (k % 2) * 2, 8 + ((k // 2) % 2) * 2, 8 + (k // 4) * 2)])
    assert [float(f32_closed_form(seed + a + epoch)) * 2 ** 23 for a in axons] == [4.0] * 4 + [5.0] * 4
    n, t, *_ = s.run(types_table(w_explore=1.0), epoch=epoch, seed=seed, max_sprouts=4)
    assert n == 4
    assert [int(t[k, 0]) for k in (127, 126, 125, 124)] == [pack_target(a, 0) for a in (24, 25, 26, 27)]
    n, t, *_ = s.run(types_table(w_explore=1.0), epoch=epoch + 2, seed=seed, max_sprouts=4)
    assert [int(t[k, 0]) for k in (127, 126, 125, 124)] == [pack_target(a, 0) for a in (22, 23, 24, 25)]


@pytest.mark.parametrize("case", ["two_values", "wrap", "negative_weight"])
def test_a_later_cell_with_mixed_noise_is_scanned(case):
    """One sprout (max_sprouts 1), w_explore only. Cell (3, 3, 3), the first
    of the soma's radius in scan order, holds one axon; cell (7, 7, 7), the
    last, holds axons 23 and 30, whose noise differs. The winner sits in the
    last cell, so a pass that skips cells by a bound on their scores must
    bound that cell by its best axon, not by its least or greatest axon id:
    - two_values: seed + a + epoch crosses 5 * 2^41 at a = 24: axons 20 and 23
      score 4 / 2^23, axon 30 scores 5 / 2^23: axon 30;
    - wrap: seed + a + epoch passes 2^64 at a = 25: axon 23 scores
      (2^23 - 1) / 2^23, axons 30 and 40 score 0: axon 23 (the cell's
      greatest axon has the least noise);
    - negative_weight: w_explore = -1 with the seed of two_values and axon 35
      (5 / 2^23) in the first cell: axon 23 (least noise) wins."""
    epoch = 5
    s = Scene()
    s.soma(0, (10, 10, 10))
    if case == "wrap":
        seed = (1 << 64) - 25 - epoch
        first, w_e, want = 40, 1.0, 23
    else:
        seed = (5 << 41) - 24 - epoch
        first, w_e, want = (20, 1.0, 30) if case == "two_values" else (35, -1.0, 23)
    s.segments(first, [(6, 6, 6)])
    s.segments(23, [(14, 14, 14)])
    s.segments(30, [(14, 14, 14)])
    noise = {a: float(f32_closed_form(seed + a + epoch)) for a in (first, 23, 30)}
    best = max(noise, key=lambda a: w_e * noise[a])
    assert best == want and sorted(w_e * v for v in noise.values())[-2] < w_e * noise[want]
    n, t, *_ = s.run(types_table(w_explore=w_e), epoch=epoch, seed=seed, max_sprouts=1)
    assert n == 1 and int(t[127, 0]) == pack_target(want, 0)


@pytest.mark.parametrize("w_type,first_type,last_type", [(-1.0, 0, 1), (1.0, 1, 0)])
def test_type_affinity_decides_across_cells(w_type, first_type, last_type):
    """w_distance 0.5 and w_type, one sprout; soma 5 is of type 0. Its first
    cell holds axon 30, its last cell axon 31. Score = 0.5 + w_type * [same
    type] (sprouting.rs:292-296): with a negative weight axon 30 of the
    soma's own type scores 0.5 - 1 and axon 31 of type 1 scores 0.5; with a
    positive weight axon 30 of type 1 scores 0.5 and axon 31 of type 0 1.5.
    Either way the last cell's axon wins (strictly greater)."""
    s = Scene()
    s.soma(5, (10, 10, 10), t=0)
    s.segments(30, [(6, 6, 6)], t=first_type)
    s.segments(31, [(14, 14, 14)], t=last_type)
    n, t, *_ = s.run(types_table(w_distance=0.5, w_type=w_type), max_sprouts=1)
    assert n == 1 and int(t[127, 5]) == pack_target(31, 0)


def test_an_owner_that_sprouts_earlier_wins_with_its_new_power():
    """w_power only, one sprout per soma. Soma 40 (not sprouting) owns axon 40,
    the only axon in soma 5's first cell, with sum |w| = 1000; soma 2 owns
    axon 2, the only axon in soma 5's last cell, with sum |w| = 950. Soma 2
    bursts and sprouts first (index order) onto axon 50 near it, +74, so its
    power index is 1024 / (128 * 32767) when soma 5 scores axon 2
    (sprouting.rs:285-290, 331-333): soma 5 picks axon 2, which at the start
    of the pass scored below axon 40."""
    s = Scene()
    s.soma(2, (40, 40, 40))
    s.soma(5, (10, 10, 10))
    s.segments(50, [(40, 40, 41)])     # soma 2's candidate
    s.segments(40, [(6, 6, 6)])        # soma 5's first cell
    s.segments(2, [(14, 14, 14)])      # soma 5's last cell
    s.targets[0, 2], s.weights[0, 2] = pack_target(60, 0), 950
    s.targets[0, 40], s.weights[0, 40] = pack_target(61, 0), 1000
    n, t, w, *_ = s.run(types_table(w_power=1.0), max_sprouts=1)
    assert n == 2
    assert int(t[127, 2]) == pack_target(50, 0) and int(w[127, 2]) == 74
    assert int(t[127, 5]) == pack_target(2, 0)


# --------------------------------------------------------------------------
# nudge_axon and handovers
# --------------------------------------------------------------------------
def test_nudge_moves_bursting_somas_axon_and_appends_a_segment():
    """Soma 3 with flags & 0x02: its axon's tip (5, 6, 7) with direction
    (x +100, y -100, z +10) moves to (6, 5, 7); the segment is appended at
    the old length and the length grows by one."""
    s = Scene()
    s.segments(3, [(5, 6, 6), (5, 6, 7)], t=1)
    s.dirs[3] = (100 & 0xFF) | ((-100 & 0xFF) << 8) | (10 << 16)
    s.flags[3] = 0x02
    n, t, w, g, hand, _ = s.run(types_table(w_distance=0.5))
    assert g.tips[3] == pack_pos(6, 5, 7, 1)
    assert g.lengths[3] == 3 and g.paths[3, 2] == pack_pos(6, 5, 7, 1)
    assert hand == []


def test_nudge_length_wraps_at_256():
    """lengths are u8: a 255-segment axon writes segment 255 and its length
    wraps to 0 (sprouting.rs:112-116)."""
    s = Scene()
    s.segments(3, [(5, 6, 7)] * 255)
    s.dirs[3] = 100 << 16                  # z + 1
    s.flags[3] = 0x02
    *_, g, hand, _ = s.run(types_table(w_distance=0.5))
    assert g.lengths[3] == 0 and g.paths[3, 255] == pack_pos(5, 6, 8, 0) and g.tips[3] == pack_pos(5, 6, 8, 0)


def test_nudge_off_the_shard_is_a_handover():
    """Ghost 64 at x = max_x - 1 moving +x, and ghost 65 at z = 63 moving +z,
    both leave: one handover each (fields per sprouting.rs:87-103), tips
    zeroed, paths and lengths kept."""
    s = Scene(ghosts=2, max_x=32, max_y=40)
    s.segments(N, [(31, 7, 9)] * 3, t=2)
    s.dirs[N] = 100 | ((-30 & 0xFF) << 8) | ((-65 & 0xFF) << 16)
    s.segments(N + 1, [(4, 39, 63)], t=3)
    s.dirs[N + 1] = 90 << 16
    *_, g, hand, _ = s.run(types_table(w_distance=0.5))
    assert g.tips[N] == 0 and g.tips[N + 1] == 0
    assert g.lengths[N] == 3 and g.lengths[N + 1] == 1
    assert len(hand) == 2
    h0, h1 = hand
    assert (h0["origin_zone_hash"], h0["local_axon_id"], h0["entry_x"], h0["entry_y"], h0["entry_z"]) == \
        (ZONE, N, 32, 7, 8)
    assert (h0["vector_x"], h0["vector_y"], h0["vector_z"], h0["type_mask"], h0["remaining_length"]) == \
        (100, -30, -65, 2, 253)
    assert (h1["local_axon_id"], h1["entry_x"], h1["entry_y"], h1["entry_z"], h1["type_mask"],
            h1["remaining_length"]) == (N + 1, 4, 39, 63, 3, 255)


def test_absorbed_handover_tip_packing():
    """An incoming handover (x 3, y 4, z 5, type 2) lands in the first free
    ghost slot with tip (2 << 28) | (5 << 22) | (4 << 11) | 3, acknowledged
    with that slot. The ghost is then nudged: the tip decodes (10/10/8 bits)
    as x 3, y 8, z 20, type 2, so with direction 0 the new segment 0 is
    (3, 8, 20, 2)."""
    s = Scene(ghosts=3)
    s.tips[N] = pack_pos(1, 1, 1, 0)       # slot 64 is taken
    s.lengths[N] = 1
    s.paths[N, 0] = s.tips[N]
    inc = np.zeros(1, NT.HANDOVER_DTYPE)
    inc["origin_zone_hash"], inc["local_axon_id"] = 0x77, 12
    inc["entry_x"], inc["entry_y"], inc["entry_z"], inc["type_mask"] = 3, 4, 5, 2
    *_, g, hand, acks = s.run(types_table(w_distance=0.5), incoming=inc)
    assert acks == [(0x77, 12, N + 1)]
    raw = (2 << 28) | (5 << 22) | (4 << 11) | 3
    assert SO.pos_xyzt(raw) == (3, 8, 20, 2)
    assert g.tips[N + 1] == pack_pos(3, 8, 20, 2)
    assert g.paths[N + 1, 0] == pack_pos(3, 8, 20, 2) and g.lengths[N + 1] == 1
