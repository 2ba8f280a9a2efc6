"""Parameter search over GLIF/GSOP parameter sets, plain and surrogate-guided.

The reference tunes its brain with Optuna TPE. Each trial hot-patches the
parameter table and runs a CartPole episode on a live node
(examples/bayesian_search.py:63-187). That search space is the first seven
parameters of brain.PARAM_NAMES. The objective needs a gym environment and a
running genesis-node, so here every trial is a simulator run of a
20k-neuron, 64-slot brain on the batched Day-Phase engine, and the score is
computed from the run's activity (``score_of``).

* ``Simulator`` runs many (parameter vector, input condition) trials as one
  batch of brains (axc_batch_*). Each trial starts from the same generated
  topology state and returns the eight outcomes of OUTCOME_NAMES.
* ``tpe_search``: the reference's own search method (Optuna's default TPE
  sampler, restated in tpe.py), every proposal simulated. This is the plain
  search the surrogate is measured against.
* ``plain_search``: uniform random proposals, every one simulated (for scale).
* ``surrogate_search``: the MLP (surrogate.Mlp, hand-written MFMA kernels)
  scores a large pool of random candidates. Only the top k go to the
  simulator, which judges every proposal. New judgements are added to the
  training set and the network is fine-tuned between rounds.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import brain as B
from ._ffi import day_lib
from .batch import BatchEngine
from .engine import DeviceBuffer, VramState, device_copy
from .layout import compute_state_offsets

INPUT_FEATURES = ["input_rate", "input_modulation", "input_period_kticks", "input_driven_fraction"]
FEATURE_NAMES = B.PARAM_NAMES + INPUT_FEATURES
OUTCOME_NAMES = ["rate_exc", "rate_inh", "rate_first_half", "rate_second_half", "weight_log_ratio",
                 "pruned_fraction", "log_threshold_offset", "spread_log_ratio"]
# rate_*: spikes per neuron per tick in percent (excitatory type 0,
#   inhibitory type 1; first and second half of the run over all neurons).
# weight_log_ratio: log((mean |w| at the end + 1) / (mean |w| at the start + 1))
#   over the filled slots; spread_log_ratio: the same for the std of |w|.
# pruned_fraction: filled slots whose |w| ended below the prune threshold
#   (what the next Night's sort&prune would remove).
# log_threshold_offset: log(1 + mean homeostatic threshold offset / 1000).


@dataclass
class Workload:
    n_neurons: int = 20_000
    slots: int = 64
    n_inputs: int = 1_000
    ticks: int = 1_000
    batch_brains: int = 256
    n_topologies: int = 4
    n_conditions: int = 64
    seed: int = 20261203
    graph_ticks: int = 50


@dataclass
class Condition:
    """An input stream: rate p(t) = p (1 + mod sin(2 pi t / period)) on the
    first driven*n_inputs inputs, zero on the rest."""
    rate: float
    modulation: float
    period: float
    driven: float

    def features(self) -> np.ndarray:
        return np.array([self.rate, self.modulation, self.period / 1000.0, self.driven], np.float64)

    def bits(self, n_inputs: int, ticks: int, seed: int) -> np.ndarray:
        rng = np.random.default_rng(seed)
        t = np.arange(ticks)
        pt = self.rate * (1.0 + self.modulation * np.sin(2 * np.pi * t / self.period))
        n_drv = int(round(self.driven * n_inputs))
        words = (n_inputs + 31) // 32
        fire = np.zeros((ticks, words * 32), bool)
        fire[:, :n_drv] = rng.random((ticks, n_drv)) < pt[:, None]
        return np.packbits(fire, axis=1, bitorder="little").view(np.uint32).reshape(ticks, words).copy()


def random_condition(rng: np.random.Generator) -> Condition:
    return Condition(float(rng.uniform(0.01, 0.05)), float(rng.uniform(0.0, 0.9)), float(rng.uniform(20, 400)),
                     float(rng.uniform(0.3, 1.0)))


SEARCH_CONDITION = Condition(0.03, 0.5, 200.0, 0.8)


def brain_spec(wl: Workload, k: int) -> B.BrainSpec:
    """Topology k of the search brain: about one neuron per voxel, somas
    numbered in 8x8x8 bricks (so the grid is a multiple of 8)."""
    side = max(8, int(round(wl.n_neurons ** (1 / 3) / 8)) * 8)
    gz = max(8, -(-wl.n_neurons // (side * side * 8)) * 8)
    return B.BrainSpec(wl.n_neurons, wl.n_inputs, (side, side, gz), wl.slots, (1, 255),
                       seed=(wl.seed * 1_000_003 + k) & 0xFFFFFFFFFFFF, input_ppm=30_000, lut=B.lut_c1(), brick=8)


TARGET_RATE = 3.0  # percent per tick, about four times the heartbeat-only rate


def score_of(out: np.ndarray) -> np.ndarray:
    """Search objective from the eight outcomes (higher is better, at most 0):
    excitatory and inhibitory rates near TARGET_RATE, a stationary rate
    (second half like the first), stable synaptic weights (mean |w| neither
    runs away nor dies) and few weights below the prune threshold. It stands
    in for the reference's episode score, which needs a live environment;
    like that score it punishes runaway and dead weights (its fatality test,
    examples/bayesian_search.py:119-126)."""
    out = np.atleast_2d(out)
    e, i, r1, r2 = (np.maximum(out[:, k], 0.0) for k in range(4))   # predicted rates may dip below 0
    lw, pruned = out[:, 4], out[:, 5]
    le = np.log((e + 0.05) / (TARGET_RATE + 0.05))
    li = np.log((i + 0.05) / (TARGET_RATE + 0.05))
    ls = np.log((r2 + 0.05) / (r1 + 0.05))
    return -(le * le + li * li + ls * ls + lw * lw + 2.0 * pruned)


class _Ptr:
    def __init__(self, p: int):
        self.ptr = p


class Simulator:
    """Runs trials on the batched engine: topologies generated once and kept
    on the device, a pool of brain slots with one batch object per batch
    size (reused across trials: each trial copies its topology in on the
    device and gives the batch its table, dopamine and input stream), input
    streams generated once per condition."""

    def __init__(self, wl: Workload, conditions: list[Condition] | None = None, stream=None):
        self.wl = wl
        self.lib = day_lib()
        t0 = time.perf_counter()
        self.topos = [B.generate(brain_spec(wl, k)) for k in range(wl.n_topologies)]
        b0 = self.topos[0]
        self.N, self.A = b0.padded_n, b0.total_axons
        off = compute_state_offsets(self.N)
        self.off = off
        self.types = [(b.state[off.soma_flags:off.soma_flags + self.N] >> 4).astype(np.int64) for b in self.topos]
        self.n_types = [(max(1, int(np.count_nonzero(t == 0))), max(1, int(np.count_nonzero(t == 1))))
                        for t in self.types]
        self.syn0 = []   # (mean |w|, std |w|) over the filled slots of each topology's initial state
        for b in self.topos:
            w = b.state[off.dendrite_weights:off.dendrite_weights + 2 * 128 * self.N].view(np.int16)
            t = b.state[off.dendrite_targets:off.dendrite_targets + 4 * 128 * self.N].view(np.uint32)
            aw = np.abs(w.astype(np.int64))[t != 0]
            self.syn0.append((float(aw.mean()), float(aw.std())))
        self.topo_dev = [(DeviceBuffer.from_array(b.state), DeviceBuffer.from_array(b.heads)) for b in self.topos]
        rng = np.random.default_rng(wl.seed)
        self.conditions = conditions if conditions is not None else [random_condition(rng)
                                                                      for _ in range(wl.n_conditions)]
        self.cond_bufs = [DeviceBuffer.from_array(c.bits(wl.n_inputs, wl.ticks, wl.seed + 7 * j))
                          for j, c in enumerate(self.conditions)]
        self.pool = [VramState(self.N, self.A) for _ in range(wl.batch_brains)]
        self.slot_topo = [-1] * wl.batch_brains
        self.engines = {}
        self.stream = stream
        self.tuning = None
        self.gen_seconds = time.perf_counter() - t0
        self.sim_seconds = 0.0
        self.phase_seconds = {"upload": 0.0, "step": 0.0, "outcomes": 0.0}
        self.trials = 0

    def add_condition(self, c: Condition) -> int:
        self.conditions.append(c)
        j = len(self.conditions) - 1
        self.cond_bufs.append(DeviceBuffer.from_array(c.bits(self.wl.n_inputs, self.wl.ticks, self.wl.seed + 7 * j)))
        return j

    def features(self, params: np.ndarray, cond_ids) -> np.ndarray:
        """[n][16]: the parameter vector, then the input condition's features."""
        params = np.atleast_2d(params).astype(np.float64)
        table = np.stack([c.features() for c in self.conditions])
        cf = table[np.broadcast_to(np.asarray(cond_ids, np.int64), (params.shape[0],))]
        return np.concatenate([params, cf], axis=1)

    def _members(self, params, cond_ids, topo_ids):
        """Copies each trial's topology into its pool slot (on the device) and
        returns the batch members."""
        mems = []
        for k, (p, cj, tj) in enumerate(zip(params, cond_ids, topo_ids)):
            b = self.topos[int(tj)]
            v = self.pool[k]
            st, hd = self.topo_dev[int(tj)]
            device_copy(v.ptrs.soma_voltage, st.ptr, st.nbytes)
            device_copy(v.ptrs.axon_heads, hd.ptr, hd.nbytes)
            if self.slot_topo[k] != int(tj):  # another topology's targets: the slot's plan is rebuilt
                self.lib.axc_shard_targets_changed(C.byref(v.ptrs))
                self.slot_topo[k] = int(tj)
            mems.append({"vram": v, "lut": B.lut_from_params(p), "d_inputs": _Ptr(self.cond_bufs[int(cj)].ptr),
                         "input_words": (self.wl.n_inputs + 31) // 32, "input_ticks": self.wl.ticks, "v_seg": 1,
                         "virtual_offset": b.virtual_offset, "n_inputs": self.wl.n_inputs,
                         "dopamine": int(p[6])})
        return mems

    def run(self, params: np.ndarray, cond_ids, topo_ids=None) -> np.ndarray:
        """Outcomes [n][8] of n trials, simulated batch_brains at a time."""
        params = np.atleast_2d(np.asarray(params, np.int64))
        n = params.shape[0]
        cond_ids = np.broadcast_to(np.asarray(cond_ids), (n,))
        topo_ids = np.zeros(n, np.int64) if topo_ids is None else np.broadcast_to(np.asarray(topo_ids), (n,))
        out = np.zeros((n, len(OUTCOME_NAMES)), np.float64)
        t0 = time.perf_counter()
        for s in range(0, n, self.wl.batch_brains):
            e = min(n, s + self.wl.batch_brains)
            out[s:e] = self._run_batch(params[s:e], cond_ids[s:e], topo_ids[s:e])
        self.sim_seconds += time.perf_counter() - t0
        self.trials += n
        return out

    def _engine(self, mems):
        n = len(mems)
        be = self.engines.get(n)
        if be is None:
            be = BatchEngine(mems)
            if self.tuning is None:
                self.tuning = be.autotune(8, self.stream)
            else:
                be.set_tuning(**self.tuning)
            self.engines[n] = be
            be.update(mems)  # counters from zero
        else:
            be.update(mems)
        return be

    def _run_batch(self, params, cond_ids, topo_ids) -> np.ndarray:
        wl, lib = self.wl, self.lib
        t0 = time.perf_counter()
        mems = self._members(params, cond_ids, topo_ids)
        be = self._engine(mems)
        t1 = time.perf_counter()
        half = wl.ticks // 2
        mid = None
        for t in range(0, wl.ticks, wl.graph_ticks):
            if t == half:
                lib.gpu_stream_synchronize(self.stream)
                mid = be.stats_all()
            be.step(t, min(wl.graph_ticks, wl.ticks - t), self.stream)
        lib.gpu_stream_synchronize(self.stream)
        end = be.stats_all()
        if mid is None:
            mid = end
        t2 = time.perf_counter()
        syn = be.synapse_stats(params[:, 5], self.stream)
        res = np.stack([self._outcome(params[k], topo_ids[k], mid[1][k], end[1][k], syn[k])
                        for k in range(len(mems))])
        ps = self.phase_seconds
        ps["upload"] += t1 - t0
        ps["step"] += t2 - t1
        ps["outcomes"] += time.perf_counter() - t2
        return res

    def _outcome(self, p, tj, ts_mid, ts_end, syn) -> np.ndarray:
        """Outcomes of one trial from its spikes per type at half time and at
        the end, and its synapse statistics."""
        wl, N = self.wl, self.N
        s0 = self.syn0[int(tj)]
        half = wl.ticks // 2
        ts_mid, ts_end = ts_mid.astype(np.float64), ts_end.astype(np.float64)
        n_exc, n_inh = self.n_types[int(tj)]
        rate_exc = 100.0 * ts_end[0] / (n_exc * wl.ticks)
        rate_inh = 100.0 * ts_end[1] / (n_inh * wl.ticks)
        r1 = 100.0 * ts_mid.sum() / (N * half)
        r2 = 100.0 * (ts_end.sum() - ts_mid.sum()) / (N * (wl.ticks - half))
        filled, sab, ssq, below, sthr = (float(x) for x in syn)
        filled = max(1.0, filled)
        mean1 = sab / filled
        std1 = np.sqrt(max(0.0, ssq / filled - mean1 * mean1))
        lw = np.log((mean1 + 1.0) / (s0[0] + 1.0))
        ls = np.log((std1 + 1.0) / (s0[1] + 1.0))
        pruned = below / filled
        thr_off = np.log1p(max(0.0, sthr / wl.n_neurons) / 1000.0)
        return np.array([rate_exc, rate_inh, r1, r2, lw, pruned, thr_off, ls], np.float64)

    def free(self) -> None:
        for be in self.engines.values():
            be.destroy()
        self.engines = {}
        for v in self.pool:
            v.free()
        for b in self.cond_bufs:
            b.free()
        for st, hd in self.topo_dev:
            st.free()
            hd.free()


# ---------------------------------------------------------------------------
# Searches
# ---------------------------------------------------------------------------
@dataclass
class History:
    """Every judged trial in order, with the wall-clock second (since the
    search started) at which its batch finished."""
    params: list = field(default_factory=list)
    scores: list = field(default_factory=list)
    times: list = field(default_factory=list)

    def add(self, params, scores, t):
        self.params.extend(np.atleast_2d(params).tolist())
        self.scores.extend(np.atleast_1d(scores).tolist())
        self.times.extend([t] * len(np.atleast_1d(scores)))

    def best(self) -> float:
        return max(self.scores) if self.scores else -np.inf

    def time_to(self, target: float) -> float | None:
        for s, t in zip(self.scores, self.times):
            if s >= target:
                return t
        return None


def plain_search(sim: Simulator, n_trials: int, cond_id: int, rng: np.random.Generator) -> History:
    """Uniform random proposals, each judged by the simulator."""
    h = History()
    t0 = time.perf_counter()
    bb = sim.wl.batch_brains
    for s in range(0, n_trials, bb):
        p = B.random_params(rng, min(bb, n_trials - s))
        out = sim.run(p, cond_id)
        h.add(p, score_of(out), time.perf_counter() - t0)
    return h


def tpe_search(sim: Simulator, n_trials: int, cond_id: int, seed: int, batch: int | None = None) -> History:
    """The reference's search method (examples/bayesian_search.py:182-187:
    an Optuna study with its default TPE sampler), restated in tpe.py: each
    batch of proposals is drawn from the Parzen estimators of all trials
    judged so far; the sampler's own time counts in the wall-clock."""
    from .tpe import TPE
    h = History()
    t0 = time.perf_counter()
    bb = batch or sim.wl.batch_brains
    opt = TPE(B.PARAM_RANGES, seed=seed)
    h.sampler_seconds = 0.0
    for s in range(0, n_trials, bb):
        ts = time.perf_counter()
        p = opt.suggest(min(bb, n_trials - s))
        h.sampler_seconds += time.perf_counter() - ts
        sc = score_of(sim.run(p, cond_id))
        opt.tell(p, sc)
        h.add(p, sc, time.perf_counter() - t0)
    return h


def train(net, X: np.ndarray, Y: np.ndarray, steps: int, batch: int, lr: float, rng: np.random.Generator,
          lr_final: float | None = None) -> list:
    """Adam steps on random minibatches (inputs and targets already scaled);
    the learning rate decays geometrically to lr_final if given."""
    from .engine import DeviceBuffer as DB
    n = X.shape[0]
    dX, dY = DB.from_array(np.ascontiguousarray(X, np.float32)), DB.from_array(np.ascontiguousarray(Y, np.float32))
    bx, by = DB(batch * X.shape[1] * 4), DB(batch * Y.shape[1] * 4)
    losses = []
    try:
        for s in range(steps):
            idx = rng.integers(0, n, batch)
            bx.upload(np.ascontiguousarray(X[idx], np.float32))
            by.upload(np.ascontiguousarray(Y[idx], np.float32))
            a = s / max(1, steps - 1)
            l_r = lr if lr_final is None else lr * (lr_final / lr) ** a
            want = (s % 50 == 0) or s == steps - 1
            loss = net.train_step_device(bx.ptr, by.ptr, batch, l_r, want_loss=want)
            if want:
                losses.append((s, loss))
    finally:
        for b in (dX, dY, bx, by):
            b.free()
    return losses


def predict_scores(net, xs, ys, X: np.ndarray) -> np.ndarray:
    return score_of(ys.inv(net.predict(xs.fwd(X)).astype(np.float64)))


def surrogate_search(sim: Simulator, net, xs, ys, cond_id: int, rng: np.random.Generator, rounds: int,
                     candidates: int, top_k: int, X_train: np.ndarray, Y_train: np.ndarray,
                     finetune_steps: int = 200, batch: int = 4096, lr: float = 3e-4) -> tuple[History, dict]:
    """Each round: score `candidates` random parameter vectors with the
    surrogate, simulate the top_k, add them to the training set, fine-tune.
    info["predicted"] / info["simulated"] hold, per round, the network's
    outcome predictions for the judged proposals (made before the judgement)
    and the simulator's outcomes."""
    h = History()
    t0 = time.perf_counter()
    t_score = t_train = 0.0
    X, Y = X_train, Y_train
    predicted, simulated = [], []
    for r in range(rounds):
        ts = time.perf_counter()
        cand = B.random_params(rng, candidates)
        pred_out = ys.inv(net.predict(xs.fwd(sim.features(cand, cond_id))).astype(np.float64))
        top = np.argsort(-score_of(pred_out))[:top_k]
        t_score += time.perf_counter() - ts
        out = sim.run(cand[top], cond_id)
        h.add(cand[top], score_of(out), time.perf_counter() - t0)
        predicted.append(pred_out[top])
        simulated.append(out)
        if r + 1 < rounds and finetune_steps:
            tt = time.perf_counter()
            X = np.concatenate([X, sim.features(cand[top], cond_id)])
            Y = np.concatenate([Y, out])
            train(net, xs.fwd(X), ys.fwd(Y), finetune_steps, min(batch, X.shape[0]), lr, rng)
            t_train += time.perf_counter() - tt
    return h, {"score_seconds": t_score, "finetune_seconds": t_train, "predicted": predicted,
               "simulated": simulated}


# ---------------------------------------------------------------------------
# The experiment: surrogate-guided search against the plain search
# ---------------------------------------------------------------------------
def run_experiment(wl: Workload, n_data: int, n_plain: int, rounds: int, candidates: int, top_k: int,
                   train_steps: int, seed: int = 7, log=print) -> dict:
    """Generates surrogate data with the batched engine, trains the MLP on the
    GPU, then runs both searches on the same simulator and search condition.
    Returns a report with MAE and wall-clock figures."""
    from .surrogate import Mlp, Scaler
    lib = day_lib()
    rng = np.random.default_rng(seed)
    sim = Simulator(wl)
    cond = sim.add_condition(SEARCH_CONDITION)
    log(f"simulator: {wl.n_topologies} topologies of {wl.n_neurons} neurons x {wl.slots} slots, "
        f"{len(sim.conditions)} input conditions, set up in {sim.gen_seconds:.1f}s")

    # 1. Training data: random parameters, random conditions, random topologies.
    t0 = time.perf_counter()
    P = B.random_params(rng, n_data)
    cj = rng.integers(0, wl.n_conditions, n_data)
    tj = rng.integers(0, wl.n_topologies, n_data)
    Y = np.zeros((n_data, len(OUTCOME_NAMES)))
    for s in range(0, n_data, wl.batch_brains):
        e = min(n_data, s + wl.batch_brains)
        Y[s:e] = sim.run(P[s:e], cj[s:e], tj[s:e])
        log(f"data: {e}/{n_data} runs, {time.perf_counter() - t0:.1f}s")
    X = sim.features(P, cj)
    t_data = time.perf_counter() - t0
    n_hold = max(1, n_data // 10)
    Xtr, Ytr, Xho, Yho = X[n_hold:], Y[n_hold:], X[:n_hold], Y[:n_hold]

    # 2. Training on the GPU.
    t0 = time.perf_counter()
    net = Mlp(len(FEATURE_NAMES), 512, 3, len(OUTCOME_NAMES), max_batch=65536, seed=seed)
    xs, ys = Scaler(Xtr), Scaler(Ytr)
    losses = train(net, xs.fwd(Xtr), ys.fwd(Ytr), train_steps, min(4096, Xtr.shape[0]), 1e-3, rng, lr_final=2e-5)
    lib.gpu_device_synchronize()
    t_train = time.perf_counter() - t0
    pred = ys.inv(net.predict(xs.fwd(Xho)).astype(np.float64))
    mae = np.abs(pred - Yho).mean(axis=0)
    mae_const = np.abs(Ytr.mean(axis=0) - Yho).mean(axis=0)
    log(f"training: {train_steps} steps in {t_train:.1f}s, loss {losses[0][1]:.4f} -> {losses[-1][1]:.4f}")
    for k, n in enumerate(OUTCOME_NAMES):
        log(f"  holdout MAE {n:22s} {mae[k]:.4f} (mean predictor {mae_const[k]:.4f})")
    ho_score = score_of(Yho)
    ho_pred_score = score_of(pred)
    rank_corr = float(np.corrcoef(np.argsort(np.argsort(ho_score)), np.argsort(np.argsort(ho_pred_score)))[0, 1])

    # Scoring throughput of the network (1M candidates).
    cand = B.random_params(rng, candidates)
    Xc = xs.fwd(sim.features(cand, cond))
    dx = DeviceBuffer.from_array(Xc)
    dy = DeviceBuffer(candidates * len(OUTCOME_NAMES) * 4)
    net.forward_device(dx.ptr, candidates, dy.ptr)
    lib.gpu_device_synchronize()
    ts = time.perf_counter()
    for _ in range(5):
        net.forward_device(dx.ptr, candidates, dy.ptr)
    lib.gpu_device_synchronize()
    t_fwd = (time.perf_counter() - ts) / 5
    dx.free()
    dy.free()
    flops = 2.0 * (16 * 512 + 512 * 512 * 2 + 512 * 8) * candidates
    log(f"scoring {candidates} candidates: {t_fwd * 1e3:.2f} ms, {flops / t_fwd / 1e12:.1f} TFLOP/s")

    # 3. The plain search: TPE, the reference's method; uniform random
    # proposals beside it for scale.
    hp = tpe_search(sim, n_plain, cond, seed + 1)
    best_plain = hp.best()
    t_plain = hp.time_to(best_plain)
    log(f"plain search (TPE): {n_plain} trials, best {best_plain:.4f} reached at {t_plain:.1f}s "
        f"(of {hp.times[-1]:.1f}s)")
    hr = plain_search(sim, n_plain, cond, rng)
    log(f"random search: {n_plain} trials, best {hr.best():.4f} reached at {hr.time_to(hr.best()):.1f}s")

    # 4. Surrogate-guided search.
    hs, info = surrogate_search(sim, net, xs, ys, cond, rng, rounds, candidates, top_k, Xtr, Ytr)
    t_surr = hs.time_to(best_plain)
    t_surr_random = hs.time_to(hr.best())
    log(f"surrogate search: {len(hs.scores)} trials, best {hs.best():.4f}; TPE best reached at "
        f"{'never' if t_surr is None else f'{t_surr:.1f}s'}")
    predicted, simulated = info.pop("predicted"), info.pop("simulated")
    mae_judged = [np.abs(p - o).mean(axis=0) for p, o in zip(predicted, simulated)]
    for r, m in enumerate(mae_judged):
        log(f"  round {r}: judged MAE " + " ".join(f"{n}={v:.3f}" for n, v in zip(OUTCOME_NAMES, m)))
    first = hs.params[:top_k]
    net.destroy()
    sim.free()
    rep = {
        "workload": {"neurons": wl.n_neurons, "slots": wl.slots, "inputs": wl.n_inputs, "ticks": wl.ticks,
                     "brains_per_batch": wl.batch_brains, "topologies": wl.n_topologies,
                     "input_conditions": wl.n_conditions},
        "data": {"runs": n_data, "seconds": t_data, "holdout": n_hold},
        "simulator_phase_seconds": sim.phase_seconds,
        "training": {"steps": train_steps, "batch": min(4096, Xtr.shape[0]), "seconds": t_train,
                     "loss_first": losses[0][1], "loss_last": losses[-1][1]},
        "holdout_mae": dict(zip(OUTCOME_NAMES, mae.tolist())),
        "holdout_mae_mean_predictor": dict(zip(OUTCOME_NAMES, mae_const.tolist())),
        "holdout_score_rank_corr": rank_corr,
        "judged_mae_per_round": [dict(zip(OUTCOME_NAMES, m.tolist())) for m in mae_judged],
        "scoring": {"candidates": candidates, "ms": t_fwd * 1e3, "tflops": flops / t_fwd / 1e12},
        "plain": {"method": "TPE (tpe.py; the reference's Optuna default sampler)", "trials": n_plain,
                  "best": best_plain, "seconds_to_best": t_plain, "seconds": hp.times[-1],
                  "sampler_seconds": hp.sampler_seconds,
                  "best_after": {str(n): max(hp.scores[:n]) for n in (256, 512, 1024, 2048, 4096, 8192)
                                 if n <= n_plain}},
        "random": {"trials": n_plain, "best": hr.best(), "seconds_to_best": hr.time_to(hr.best()),
                   "seconds": hr.times[-1], "mean_score": float(np.mean(hr.scores))},
        "surrogate": {"rounds": rounds, "top_k": top_k, "trials": len(hs.scores), "best": hs.best(),
                      "seconds_to_plain_best": t_surr, "seconds": hs.times[-1],
                      "round_best": [max(hs.scores[r * top_k:(r + 1) * top_k]) for r in range(rounds)],
                      "round_mean": [float(np.mean(hs.scores[r * top_k:(r + 1) * top_k])) for r in range(rounds)],
                      **info},
        "tpe_mean_score": float(np.mean(hp.scores)),
        "speedup_search_only": (t_plain / t_surr) if t_surr else None,
        "speedup_including_data_and_training": (t_plain / (t_surr + t_data + t_train)) if t_surr else None,
        "speedup_vs_random_search_only": (hr.time_to(hr.best()) / t_surr_random) if t_surr_random else None,
        "first_round_params": np.array(first[:4]).tolist(),
    }
    return rep


def main() -> int:
    import argparse
    import json
    ap = argparse.ArgumentParser(description="surrogate-guided vs plain parameter search")
    ap.add_argument("--data", type=int, default=8192)
    ap.add_argument("--plain", type=int, default=4096)
    ap.add_argument("--rounds", type=int, default=4)
    ap.add_argument("--candidates", type=int, default=1 << 20)
    ap.add_argument("--top-k", type=int, default=256)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--ticks", type=int, default=1000)
    ap.add_argument("--neurons", type=int, default=20_000)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    day_lib().gpu_set_device(0)
    wl = Workload(n_neurons=a.neurons, ticks=a.ticks)
    rep = run_experiment(wl, a.data, a.plain, a.rounds, a.candidates, a.top_k, a.train_steps, a.seed,
                         log=lambda m: print(m, flush=True))
    s = json.dumps(rep, indent=1, default=float)
    if a.out:
        with open(a.out, "w") as f:
            f.write(s + "\n")
    print(s)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
