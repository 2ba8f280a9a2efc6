"""The C1 golden run: digests of the CPU oracle's run of the 4,096-neuron C1
brain, committed as c1_digests.json beside this file.

The run (4,096 neurons x 64 slots on a 16^3 grid, 80% excitatory, 256 inputs
at p = 0.05, axons of 1-32 segments): 1,000 ticks in batches of 50, GSOP on,
dopamine 0 in even batches and 60 in odd ones, v_seg 1, and a Night every 250
ticks as shard_thread.rs:269-378 runs it: the flags taken before the sort &
prune (prune threshold 15, bindings.cu:450-515), then the sprouting pass
(genesis-baker/src/bake/sprouting.rs:119-378, at most 4 new synapses per
neuron, epoch = the Night's index, template sprouting weights).

Recorded, as the first 16 hex digits of SHA-256:
- every tick's spike raster: flags & 1 of all 4,096 somas, one byte each
  (K7's readout, physics.cu:378-395, with every soma mapped);
- the spike count of every tick;
- each state array (the eight arrays of the .state blob) and the axon heads
  at the start, before and after every Night, and at the end.

tests/test_golden.py re-runs the oracle's day phase on the CPU with the
engine's host sprouting for the Nights and compares; tests/test_gpu_golden.py
runs the engine on the GPU, Nights included, and compares. Regenerate with `python tests/golden/c1_golden.py` after a
deliberate change to the oracle, the generator or the sprouting pass.

Test infrastructure only: it drives the oracle (oracle/oracle.py and
oracle/sprout_oracle.py) as the checker.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

from axicor_amd import brain as B  # noqa: E402
from axicor_amd import night as NT  # noqa: E402
from axicor_amd.layout import StateView  # noqa: E402

FIXTURE = os.path.join(HERE, "c1_digests.json")
TICKS, BATCH, NIGHT_EVERY, PRUNE, MAX_SPROUTS, V_SEG = 1000, 50, 250, 15, 4, 1
FIELDS = ("voltage", "flags", "threshold_offset", "timers", "soma_to_axon", "targets", "weights",
          "dendrite_timers")


def dopamine_of(batch: int) -> int:
    return 60 if batch % 2 else 0


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]


def state_digests(state: np.ndarray, heads: np.ndarray, padded_n: int) -> dict:
    v = StateView(state, padded_n)
    d = {f: digest(getattr(v, f)) for f in FIELDS}
    d["axon_heads"] = digest(heads)
    return d


def brain() -> B.Brain:
    return B.generate(B.spec_c1(), with_paths=True)


def run_oracle(b: B.Brain | None = None, native_sprouting: bool = False) -> dict:
    """The golden run on the CPU oracle; returns the digest record. With
    native_sprouting the Nights' sprouting is the engine's host pass
    (axicor_amd.night.sprout_pass) instead of the sprouting oracle."""
    from oracle import oracle as O
    from oracle import sprout_oracle as SO
    b = b or brain()
    N, A = b.padded_n, b.total_axons
    state, heads = b.state.copy(), b.heads.copy()
    mapped = np.arange(N, dtype=np.uint32)
    geo = NT.geometry_of(b)
    types = NT.sprout_types_from_lut(b.lut)
    tl = [dict(zip(types.dtype.names, x)) for x in types.tolist()]
    rec = {"ticks": TICKS, "raster": [], "spikes": [], "states": [["start", 0, state_digests(state, heads, N)]]}
    for k in range(TICKS // BATCH):
        t0 = k * BATCH
        out, _ = O.step_batch(state, N, heads, A, b.lut, V_SEG, t0, BATCH, b.inputs(t0, BATCH), b.virtual_offset,
                              b.n_inputs, mapped=mapped, n_out=N, dopamine=dopamine_of(k))
        rec["raster"] += [digest(out[t]) for t in range(BATCH)]
        rec["spikes"] += [int(out[t].sum()) for t in range(BATCH)]
        tick = t0 + BATCH
        if tick % NIGHT_EVERY == 0 and tick < TICKS:
            rec["states"].append(["before_night", tick, state_digests(state, heads, N)])
            v = StateView(state, N)
            flags = v.flags.copy()
            O.sort_and_prune_blob(state, N, PRUNE)
            t, w = v.targets.copy(), v.weights.copy()
            epoch = tick // NIGHT_EVERY - 1
            if native_sprouting:
                new = NT.sprout_pass(t, w, flags, v.soma_to_axon.copy(), N, geo, types, epoch, b.spec.seed,
                                     max_sprouts=MAX_SPROUTS).new_synapses
            else:
                new, _, _, _ = SO.sprout_pass(t, w, flags, v.soma_to_axon.copy(), N, geo.lengths, geo.paths,
                                              geo.tips, geo.dirs, geo.soma_positions, tl, epoch, b.spec.seed,
                                              geo.max_x, geo.max_y, max_sprouts=MAX_SPROUTS)
            v.targets[:] = t
            v.weights[:] = w
            rec["states"].append(["after_night", tick, state_digests(state, heads, N), int(new)])
    rec["states"].append(["end", TICKS, state_digests(state, heads, N)])
    return rec


def load() -> dict:
    with open(FIXTURE) as f:
        return json.load(f)


if __name__ == "__main__":
    r = run_oracle()
    with open(FIXTURE, "w") as f:
        json.dump(r, f, indent=0)
    print(f"wrote {FIXTURE}: {r['ticks']} ticks, {sum(r['spikes'])} spikes, "
          f"Nights added {[s[3] for s in r['states'] if s[0] == 'after_night']} synapses")
