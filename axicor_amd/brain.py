"""Seeded synthetic brains in the genesis-baker binary layouts.

The generator itself is native (csrc/brain_gen.cpp). This module holds the
variant tables and the named workloads:

* ``c1`` — 4,096 integer GLIF neurons (80% excitatory / 20% inhibitory), 64
  dendrite slots each, 16x16x16 grid, axons 1-32 segments, |w| in [64,512],
  256 inputs at p=0.05 per tick.
* ``c2`` — 1.5M neurons of 4 types (excitatory/inhibitory x fast/slow leak),
  100 slots each (150M synapses), 128x128x64 grid, axons 1-255 segments,
  |w| in [64,512], 5% of neurons as inputs (75,000) at p=0.02 per tick;
  somas numbered in 8x8x8-voxel bricks.
* ``c5`` — the scale point: 12M neurons of 6 types, 128 slots each (1.5B
  synapses), 256x256x128 grid, axons 1-255 segments, 2% of neurons as inputs
  (240,000) at p=0.01 per tick; ``spec_c5(200_000)`` is its CPU-sized
  sub-shard from the same seed (about 1.2% firing per tick).

Parameter values follow the reference template (_template/blueprints.toml:13-37)
and the blueprint defaults (genesis-core/src/config/blueprints.rs:131-152).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._ffi import VARIANT_DTYPE, GenSpec, gen_lib, ptr
from .layout import StateView, calculate_state_blob_size

# genesis-core/src/seed.rs:29,33-43: seed_from_str("GENESIS").
def seed_from_str(s: str) -> int:
    h = 0xcbf29ce484222325
    for b in s.encode():
        h ^= b
        h = (h * 0x00000100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


DEFAULT_MASTER_SEED = seed_from_str("GENESIS")
DEFAULT_INERTIA = [128, 120, 112, 104, 96, 88, 80, 72, 64, 56, 48, 40, 32, 24, 16, 8]  # blueprints.rs:150-152


def template_variant(inhibitory: bool = False, **over) -> np.ndarray:
    """One VariantParameters record with the _template values."""
    v = np.zeros(1, dtype=VARIANT_DTYPE)
    v["threshold"] = 30000
    v["rest_potential"] = 10000
    v["leak_rate"] = 1500
    v["homeostasis_penalty"] = 10000
    v["spontaneous_firing_period_ticks"] = 400
    v["initial_synapse_weight"] = 74
    v["gsop_potentiation"] = 150
    v["gsop_depression"] = 60
    v["homeostasis_decay"] = 20
    v["refractory_period"] = 15
    v["synapse_refractory_period"] = 15
    v["signal_propagation_length"] = 20
    v["is_inhibitory"] = 1 if inhibitory else 0
    v["inertia_curve"] = DEFAULT_INERTIA
    v["adaptive_leak_max"] = 0
    v["adaptive_leak_gain"] = 0
    v["adaptive_mode"] = 0
    v["d1_affinity"] = 128
    v["d2_affinity"] = 128
    for k, val in over.items():
        v[k] = val
    return v[0]


def lut_from(variants: list) -> np.ndarray:
    """16-entry table; unused entries are zero (as the reference uploads)."""
    lut = np.zeros(16, dtype=VARIANT_DTYPE)
    for i, v in enumerate(variants):
        lut[i] = v
    return lut


def lut_c1() -> np.ndarray:
    return lut_from([template_variant(False), template_variant(True)])


def lut_c2() -> np.ndarray:
    """Four types: excitatory/inhibitory x fast/slow leak. Fast leak: the
    template leak. Slow leak: half the leak with the adaptive mode on (the leak
    grows with the homeostatic offset up to 3000). All four use the blueprint
    default propagation length of 10 segments (blueprints.rs:131) and a
    128-tick heartbeat; with these values the 1.5M-neuron brain fires at about
    1% per tick for its first ~1,500 ticks (the template's 20 segments drive it
    to the refractory-limited ~6%)."""
    common = dict(signal_propagation_length=10, spontaneous_firing_period_ticks=128)
    slow = dict(leak_rate=750, adaptive_mode=1, adaptive_leak_gain=64, adaptive_leak_max=3000)
    return lut_from([
        template_variant(False, **common),
        template_variant(False, **common, **slow),
        template_variant(True, **common),
        template_variant(True, **common, **slow),
    ])


@dataclass
class BrainSpec:
    n_neurons: int
    n_inputs: int
    grid: tuple[int, int, int]
    slots_filled: int
    axon_len: tuple[int, int]
    w_range: tuple[int, int] = (64, 512)
    v_seg: int = 1
    type_ppm: list[int] = field(default_factory=lambda: [800_000, 200_000])
    mean_dist: float = 3.0
    seed: int = DEFAULT_MASTER_SEED
    input_ppm: int = 50_000
    lut: np.ndarray | None = None
    brick: int = 0  # 0: voxels numbered x-fastest; B: BxBxB bricks (a spatial sort of the somas);
    #                 BAKER_ORDER: z-planes in order, each plane shuffled (the baker's numbering)

    def to_c(self) -> GenSpec:
        g = GenSpec()
        g.n_neurons = self.n_neurons
        g.n_inputs = self.n_inputs
        g.gx, g.gy, g.gz = self.grid
        g.slots_filled = self.slots_filled
        g.axon_len_min, g.axon_len_max = self.axon_len
        g.w_min, g.w_max = self.w_range
        g.v_seg = self.v_seg
        g.n_types = len(self.type_ppm)
        acc = 0
        for i, p in enumerate(self.type_ppm):
            acc += p
            g.type_ppm_cdf[i] = acc
        g.mean_dist_milli = int(round(self.mean_dist * 1000))
        g.master_seed = self.seed & 0xFFFFFFFFFFFFFFFF
        g.brick = self.brick & 0xFFFFFFFF
        return g


def spec_c1(seed: int = DEFAULT_MASTER_SEED) -> BrainSpec:
    return BrainSpec(4096, 256, (16, 16, 16), 64, (1, 32), seed=seed, input_ppm=50_000, lut=lut_c1())


# genesis-baker/src/bake/neuron_placement.rs:60-87: the baker sorts its
# placed somas by z only, each plane's positions drawn from a shuffled pool,
# so a baked shard numbers its somas plane by plane in random (x, y) order.
BAKER_ORDER = -1


def spec_c2(seed: int = DEFAULT_MASTER_SEED, brick: int = 8) -> BrainSpec:
    """Somas numbered brick by brick (8x8x8 voxels) unless brick=0
    (x-fastest) or BAKER_ORDER (z-planes, shuffled within each)."""
    return BrainSpec(1_500_000, 75_000, (128, 128, 64), 100, (1, 255),
                     type_ppm=[400_000, 400_000, 100_000, 100_000], seed=seed, input_ppm=20_000, lut=lut_c2(),
                     brick=brick)


def lut_c5() -> np.ndarray:
    """Six types: excitatory/inhibitory x fast, slow and medium leak. Fast and
    slow are the 1.5M-neuron brain's; medium is three quarters of the template
    leak with a gentler adaptive leak (gain 32, up to 2250). Propagation 10
    segments and a 128-tick heartbeat, as there."""
    common = dict(signal_propagation_length=10, spontaneous_firing_period_ticks=128)
    slow = dict(leak_rate=750, adaptive_mode=1, adaptive_leak_gain=64, adaptive_leak_max=3000)
    medium = dict(leak_rate=1125, adaptive_mode=1, adaptive_leak_gain=32, adaptive_leak_max=2250)
    return lut_from([
        template_variant(False, **common),
        template_variant(False, **common, **slow),
        template_variant(False, **common, **medium),
        template_variant(True, **common),
        template_variant(True, **common, **slow),
        template_variant(True, **common, **medium),
    ])


def spec_c5(n_neurons: int = 12_000_000, seed: int = DEFAULT_MASTER_SEED, brick: int = 8) -> BrainSpec:
    """The scale point: 12M neurons x 128 slots (about 1.5B synapses) of six
    types, axons of 1-255 segments, 2% of the neurons' count as inputs at
    p = 0.01 per tick. The voxel grid keeps the 1.5M-neuron brain's density
    (about 1.43 somas per voxel, x:y:z = 2:2:1, sides rounded to whole
    bricks), so a smaller n (a sub-shard from the same seed, e.g. 200,000 on
    a 64x64x32 grid for the CPU side) has nearly the same local
    connectivity."""
    vox = n_neurons / (1_500_000 / (128 * 128 * 64))
    q = max(brick, 1)  # grid dimensions are whole bricks
    z = max(q, int(round((vox / 4) ** (1 / 3) / q)) * q)
    return BrainSpec(n_neurons, n_neurons // 50, (2 * z, 2 * z, z), 128, (1, 255),
                     type_ppm=[300_000, 300_000, 200_000, 80_000, 70_000, 50_000], seed=seed, input_ppm=10_000,
                     lut=lut_c5(), brick=brick)


def spec_small(n: int, slots: int, n_inputs: int, seed: int, axon_len=(1, 32), input_ppm=50_000,
               grid=None, lut=None, type_ppm=None) -> BrainSpec:
    """A small brain for parity tests; grid of about one neuron per voxel."""
    if grid is None:
        side = max(1, int(round(n ** (1 / 3))))
        grid = (side, side, max(1, (n + side * side - 1) // (side * side)))
    return BrainSpec(n, n_inputs, grid, slots, axon_len, seed=seed, input_ppm=input_ppm,
                     lut=lut if lut is not None else lut_c1(),
                     type_ppm=type_ppm if type_ppm is not None else [800_000, 200_000])


@dataclass
class Brain:
    spec: BrainSpec
    padded_n: int
    total_axons: int
    state: np.ndarray     # uint8[910*padded_n]
    heads: np.ndarray     # uint32[total_axons*8]
    lut: np.ndarray       # VARIANT_DTYPE[16]
    axon_len: np.ndarray  # uint8[total_axons]
    soma_pos: np.ndarray  # uint32[padded_n]
    paths: np.ndarray | None = None  # uint32[total_axons, 256] (Night-Phase geometry)
    tips: np.ndarray | None = None   # uint32[total_axons]
    dirs: np.ndarray | None = None   # uint32[total_axons]

    @property
    def virtual_offset(self) -> int:
        """memory.rs:128-130 with zero ghosts: inputs follow the local axons."""
        return self.padded_n

    @property
    def n_inputs(self) -> int:
        return self.spec.n_inputs

    @property
    def input_words(self) -> int:
        return (self.spec.n_inputs + 31) // 32

    def view(self) -> StateView:
        return StateView(self.state, self.padded_n)

    def copy(self) -> "Brain":
        cp = lambda a: None if a is None else a.copy()  # noqa: E731
        return Brain(self.spec, self.padded_n, self.total_axons, self.state.copy(), self.heads.copy(),
                     self.lut.copy(), self.axon_len.copy(), self.soma_pos.copy(), cp(self.paths), cp(self.tips),
                     cp(self.dirs))

    def inputs(self, tick0: int, ticks: int) -> np.ndarray:
        """uint32[ticks, words] Bernoulli input bitmasks for ticks tick0.."""
        return gen_inputs(self.spec.seed, self.spec.n_inputs, self.spec.input_ppm, tick0, ticks)


def sizes(spec: BrainSpec) -> tuple[int, int]:
    g = spec.to_c()
    pn, a = C.c_uint32(), C.c_uint32()
    gen_lib().axc_gen_sizes(C.byref(g), C.byref(pn), C.byref(a))
    return pn.value, a.value


def generate(spec: BrainSpec, with_paths: bool = False) -> Brain:
    g = spec.to_c()
    pn, a = sizes(spec)
    _, nbytes = calculate_state_blob_size(pn)
    state = np.empty(nbytes, dtype=np.uint8)
    heads = np.empty(a * 8, dtype=np.uint32)
    alen = np.empty(a, dtype=np.uint8)
    pos = np.empty(pn, dtype=np.uint32)
    lut = spec.lut if spec.lut is not None else lut_c1()
    lut = np.ascontiguousarray(lut, dtype=VARIANT_DTYPE)
    rc = gen_lib().axc_gen_brain(C.byref(g), ptr(lut), ptr(state), ptr(heads), ptr(alen), ptr(pos))
    if rc != 0:
        raise ValueError(f"bad brain spec: {spec}")
    b = Brain(spec, pn, a, state, heads, lut, alen, pos)
    if with_paths:
        b.paths = np.empty((a, 256), dtype=np.uint32)
        b.tips = np.empty(a, dtype=np.uint32)
        b.dirs = np.empty(a, dtype=np.uint32)
        gen_lib().axc_gen_paths(C.byref(g), ptr(pos), ptr(alen), ptr(b.paths), ptr(b.tips), ptr(b.dirs))
    return b


def gen_inputs(seed: int, n_inputs: int, ppm: int, tick0: int, ticks: int) -> np.ndarray:
    words = (n_inputs + 31) // 32
    out = np.zeros((ticks, max(words, 1)), dtype=np.uint32)
    if n_inputs > 0 and ticks > 0:
        gen_lib().axc_gen_inputs(seed & 0xFFFFFFFFFFFFFFFF, n_inputs, ppm, tick0, ticks, ptr(out))
    return out[:, :words] if words else out[:, :0]


# ---------------------------------------------------------------------------
# Per-brain parameter vectors (batched brains, parameter search)
# ---------------------------------------------------------------------------
# The twelve GLIF/GSOP parameters of one brain and their ranges. The first
# six are the reference's search space (examples/bayesian_search.py:65-77:
# leak 100-3000, homeostasis penalty 500-10000, decay 1-100, D1/D2 affinity
# 64-255, prune threshold 2-30) plus its dopamine (-255..255); the rest are
# the template value +-50% (_template/blueprints.toml:13-37), propagation
# around the blueprint default of 10 segments.
PARAM_NAMES = ["leak_rate", "homeostasis_penalty", "homeostasis_decay", "d1_affinity", "d2_affinity",
               "prune_threshold", "dopamine", "threshold", "refractory_period", "gsop_potentiation",
               "gsop_depression", "signal_propagation_length"]
PARAM_RANGES = np.array([[100, 3000], [500, 10000], [1, 100], [64, 255], [64, 255], [2, 30], [-255, 255],
                         [15000, 45000], [8, 22], [75, 225], [30, 90], [5, 15]], dtype=np.int64)


def random_params(rng: np.random.Generator, n: int = 1) -> np.ndarray:
    """Integer parameter vectors, uniform within PARAM_RANGES (inclusive)."""
    lo, hi = PARAM_RANGES[:, 0], PARAM_RANGES[:, 1]
    return rng.integers(lo, hi + 1, size=(n, len(PARAM_NAMES)))


def lut_from_params(p) -> np.ndarray:
    """Two-type table (0 excitatory, 1 inhibitory) from one parameter vector;
    the inhibitory type gets 1.5x leak and 0.8x penalty, as the reference
    search sets them (examples/bayesian_search.py:86-87)."""
    p = [int(x) for x in p]
    leak, pen, dec, d1, d2, _prune, _dop, thr, refr, pot, dep, prop = p
    common = dict(homeostasis_decay=dec, d1_affinity=d1, d2_affinity=d2, threshold=thr, refractory_period=refr,
                  gsop_potentiation=pot, gsop_depression=dep, signal_propagation_length=prop,
                  spontaneous_firing_period_ticks=128)
    return lut_from([template_variant(False, leak_rate=leak, homeostasis_penalty=pen, **common),
                     template_variant(True, leak_rate=int(leak * 1.5), homeostasis_penalty=int(pen * 0.8), **common)])


def spec_batch_member(rng: np.random.Generator, n_min: int = 8_000, n_max: int = 250_000,
                      slot_choices=(32, 64, 128), params=None, seed: int | None = None) -> tuple[BrainSpec, np.ndarray]:
    """One brain of the batched workload: neuron count log-uniform in
    [n_min, n_max], slots from slot_choices, 5% inputs at p ~ U[0.01, 0.05],
    its own parameter vector. Returns (spec, params)."""
    n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
    slots = int(rng.choice(slot_choices))
    if params is None:
        params = random_params(rng)[0]
    side = max(8, int(round((n / 1.4) ** (1 / 3) / 8)) * 8)
    gz = max(8, (n + side * side - 1) // (side * side) // 8 * 8)
    p_in = int(rng.integers(10_000, 50_001))
    spec = BrainSpec(n, max(1, n // 20), (side, side, gz), slots, (1, 255), seed=int(seed if seed is not None else rng.integers(1, 2**62)),
                     input_ppm=p_in, lut=lut_from_params(params), brick=8)
    return spec, params
