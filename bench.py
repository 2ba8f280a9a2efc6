#!/usr/bin/env python3
"""Day-Phase throughput benchmark on one MI355X.

Workload (default ``c2``): one brain of 1,500,000 integer GLIF neurons of 4
types (excitatory/inhibitory x fast/slow leak), 100 dendrite slots each
(150M synapses), 128x128x64 voxel grid, axons of 1-255 segments, |w| in
[64,512], 75,000 input axons driven by Bernoulli p=0.02 per tick, GSOP on.
A step is one BSP batch of ``--ticks`` ticks (default 50) through the engine's
batch entry point (one captured HIP graph per batch); inputs for every tick
are resident in HBM before the timed region starts.

Prints one JSON line (metric, value = synaptic events/s, roofline of the
dominant kernel, cpu_baseline of the CPU oracle on the same host, parity of
the first ticks against that oracle). Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "synaptic events/s at 1.5M neurons (% HBM peak); batched vs serial brains/s"
HBM_PEAK_GBS = 8000.0


def _numbering(brick: int) -> str:
    if brick == -1:
        return "the baker's (z-planes, shuffled within each plane)"
    return "bricks of %d^3 voxels" % brick if brick else "x-fastest"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(st: dict, N: int, A: int, words: int, n_out: int = 0) -> dict:
    """Compulsory traffic of the reference algorithm over the counted ticks:
    2N + 18|R| + 2(N-|R|) + 4 sum_R c_i + 2H + 64A + 4 words + 4|F|
    + sum_F (4 c_i + 4 S_i) (+ 5 n_out) per tick, with |R|, c_i, H, |F|, S_i
    counted exactly by the engine. Split per kernel: neuron_tick owns the soma
    and dendrite-scan terms and the soma_to_axon reads of spikers, axon_tick
    the heads and the bitmask, gsop the GSOP terms."""
    T = st["ticks"]
    R = st["integrated"]
    neuron = 2 * N * T + 18 * R + 2 * (N * T - R) + 4 * st["targets_read"] + 2 * st["hits"] + 4 * st["spikes"]
    axon = 64 * A * T + 4 * words * T
    gsop = 4 * st["gsop_targets"] + 4 * st["gsop_slots"]
    readout = 5 * n_out * T
    return {"neuron_tick": neuron, "axon_tick": axon, "gsop": gsop, "readout": readout,
            "total": neuron + axon + gsop + readout}


def cpu_baseline(brain, inputs_host: np.ndarray, seconds: float, max_ticks: int):
    """The CPU oracle on a copy of the initial brain, one tick at a time,
    until ``seconds`` of work or ``max_ticks``. Returns (state, heads, ticks,
    elapsed, stats)."""
    from oracle import oracle as O
    state = brain.state.copy()
    heads = brain.heads.copy()
    ticks = 0
    tot = None
    t0 = time.perf_counter()
    while ticks < max_ticks:
        _, st = O.step_batch(state, brain.padded_n, heads, brain.total_axons, brain.lut, brain.spec.v_seg, ticks, 1,
                             inputs_host[ticks:ticks + 1], brain.virtual_offset, brain.n_inputs)
        tot = st if tot is None else {k: tot[k] + st[k] for k in st}
        ticks += 1
        if time.perf_counter() - t0 >= seconds:
            break
    return state, heads, ticks, time.perf_counter() - t0, tot


def cpu_single_thread(brain, inputs_host: np.ndarray, ticks: int) -> dict:
    """The same oracle on one thread (OpenMP's thread count set to 1 for the
    run, then restored) over the first `ticks` ticks of a fresh copy."""
    import ctypes as C
    from oracle import oracle as O
    O.lib()  # loads the oracle and its OpenMP runtime
    gomp = C.CDLL("libgomp.so.1")
    before = gomp.omp_get_max_threads()
    gomp.omp_set_num_threads(1)
    try:
        state, heads = brain.state.copy(), brain.heads.copy()
        t0 = time.perf_counter()
        _, st = O.step_batch(state, brain.padded_n, heads, brain.total_axons, brain.lut, brain.spec.v_seg, 0, ticks,
                             inputs_host[:ticks], brain.virtual_offset, brain.n_inputs)
        el = time.perf_counter() - t0
    finally:
        gomp.omp_set_num_threads(before)
    return {"value": st["slots_scanned"] / el, "ticks": ticks, "seconds": el}


def bench_nights(lib, brain, vram, eng, batch, nights: int, prune: int) -> dict:
    """Nights of shard_thread.rs:269-378 on the workload brain, each after one
    more Day batch: flags download, GPU sort & prune, weights and targets
    download, host sprouting (csrc/night_host.cpp), upload of the slots the
    sprouting filled (night.execute_night_phase's sequence). Wall-clock per
    phase, averaged."""
    import ctypes as C
    from axicor_amd import night as NT
    from axicor_amd._ffi import ptr
    N = brain.padded_n
    types = NT.sprout_types_from_lut(brain.lut)
    geo = NT.geometry_of(brain)
    ph = {"prune_s": 0.0, "download_s": 0.0, "sprout_s": 0.0, "upload_s": 0.0}
    new_syn = 0
    # Page-locked host copies, allocated once per shard before its first Night.
    st = NT.NightStaging.of(vram)
    flags, w, t, s2a = st.flags, st.weights, st.targets, st.soma_to_axon
    for k in range(nights):
        batch(50)
        eng.synchronize()
        t0 = time.perf_counter()
        lib.gpu_memcpy_device_to_host(ptr(flags), vram.ptrs.soma_flags, N)
        lib.launch_sort_and_prune(C.byref(vram.ptrs), N, prune)
        lib.gpu_device_synchronize()
        t1 = time.perf_counter()
        lib.gpu_memcpy_device_to_host(ptr(w), vram.ptrs.dendrite_weights, w.nbytes)
        lib.gpu_memcpy_device_to_host(ptr(t), vram.ptrs.dendrite_targets, t.nbytes)
        lib.gpu_memcpy_device_to_host(ptr(s2a), vram.ptrs.soma_to_axon, s2a.nbytes)
        t2 = time.perf_counter()
        res = NT.sprout_pass(t, w, flags, s2a, N, geo, types, k, brain.spec.seed, changed=st.changed_buffer(4))
        t3 = time.perf_counter()
        st.upload_slots(vram, res.changed_cols)  # only the new synapses' slots differ from the device
        lib.gpu_device_synchronize()
        t4 = time.perf_counter()
        for key, dt in (("prune_s", t1 - t0), ("download_s", t2 - t1), ("sprout_s", t3 - t2), ("upload_s", t4 - t3)):
            ph[key] += dt
        new_syn += res.new_synapses
        log(f"night {k}: prune {t1 - t0:.3f}s, download {t2 - t1:.2f}s, sprout {t3 - t2:.2f}s "
            f"({res.new_synapses} new synapses), upload {t4 - t3:.2f}s")
    out = {key: v / nights for key, v in ph.items()}
    out.update({"nights": nights, "new_synapses_per_night": new_syn / nights,
                "sprout_threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))})
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_batched(n_brains: int, ticks: int, input_ticks: int, batch_ticks: int, seed: int,
                  per_tick_ticks: int = 200, serial: bool = True) -> dict:
    """Batched brains (128 brains of 8k-250k neurons, slots in {32,64,128},
    their own parameter vectors and input rates) stepped together, against the
    same brains stepped one after another with the plain engine. Every brain's
    final state is compared bit for bit between the two runs."""
    from axicor_amd import brain as B
    from axicor_amd._ffi import day_lib
    from axicor_amd.batch import BatchEngine
    from axicor_amd.engine import DeviceBuffer, ShardEngine, VramState, upload_constant_memory
    lib = day_lib()
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    members = []
    for k in range(n_brains):
        spec, params = B.spec_batch_member(rng)
        b = B.generate(spec)
        members.append((b, int(params[6]), b.inputs(0, input_ticks)))
    neurons = sum(b.spec.n_neurons for b, _, _ in members)
    synapses = sum(b.spec.n_neurons * b.spec.slots_filled for b, _, _ in members)
    log(f"batched: {n_brains} brains, {neurons} neurons, {synapses} synapses, generated in "
        f"{time.perf_counter() - t0:.1f}s")
    mems = []
    for b, dop, inp in members:
        v = VramState(b.padded_n, b.total_axons)
        v.upload_state(b.state)
        v.upload_axon_heads(b.heads)
        mems.append({"vram": v, "lut": b.lut, "d_inputs": DeviceBuffer.from_array(inp), "input_words": b.input_words,
                     "input_ticks": input_ticks, "v_seg": 1, "virtual_offset": b.virtual_offset,
                     "n_inputs": b.n_inputs, "dopamine": dop})
    stream = C.c_void_p()
    lib.gpu_stream_create(C.byref(stream))
    st = stream.value
    be = BatchEngine(mems)
    tune = be.autotune(8, st)
    log(f"batched: autotuned launch shape {tune}")
    lib.gpu_stream_synchronize(st)
    w0 = time.perf_counter()
    for t in range(0, ticks, batch_ticks):
        be.step(t, min(batch_ticks, ticks - t), st)
    lib.gpu_stream_synchronize(st)
    t_batched = time.perf_counter() - w0
    stats = [be.stats(k) for k in range(n_brains)]
    be.destroy()
    res_b = [(m["vram"].download_state(), m["vram"].download_axons()) for m in mems]
    log(f"batched: {ticks} ticks in {t_batched:.2f}s")
    if not serial:
        lib.gpu_stream_destroy(st)
        return {"brains": n_brains, "ticks_per_brain": ticks, "neurons": neurons, "synapses": synapses,
                "batched_brains_per_s": n_brains / t_batched, "batched_seconds": t_batched, "tuning": tune}
    # Serial: the same brains from their initial state, one after another,
    # with the engine's per-tick staged pass and with windows of 4 ticks; the
    # faster of the two is the baseline (small brains favour the per-tick
    # pass).
    saved_window = lib.axc_window_ticks()
    serial_s = {}
    for win in (1, 4):
        lib.axc_set_window_ticks(win)
        for (b, _, _), m in zip(members, mems):
            m["vram"].upload_state(b.state)
            m["vram"].upload_axon_heads(b.heads)
        w0 = time.perf_counter()
        for (b, dop, _), m in zip(members, mems):
            upload_constant_memory(b.lut)
            v = m["vram"]
            for t in range(0, ticks, batch_ticks):
                T = min(batch_ticks, ticks - t)
                row = t % input_ticks
                err = lib.axc_step_day_phase_batch(C.byref(v.ptrs), v.padded_n, v.total_axons, 1, t, T,
                                                   m["d_inputs"].ptr + row * b.input_words * 4, b.input_words,
                                                   b.virtual_offset, b.n_inputs, None, 0, None, None, None, 0, dop,
                                                   st)
                if err:
                    raise RuntimeError(f"serial step failed {err}")
            lib.gpu_stream_synchronize(st)
        serial_s[win] = time.perf_counter() - w0
        log(f"serial (windows of {win}): {ticks} ticks x {n_brains} brains in {serial_s[win]:.2f}s")
    lib.axc_set_window_ticks(saved_window)
    best_window = min(serial_s, key=serial_s.get)
    t_serial = serial_s[best_window]
    identical = all(np.array_equal(rb[0], m["vram"].download_state()) and
                    np.array_equal(rb[1], m["vram"].download_axons()) for rb, m in zip(res_b, mems))
    log(f"batched vs serial: {'bit-identical' if identical else 'MISMATCH'} for all {n_brains} brains")
    # Serial per-tick launches, the reference's structure (one cu_step_day_phase
    # call per tick, compute/shard.rs:75-120), timed on the first pt_ticks
    # ticks of every brain and scaled to the full tick count.
    pt_ticks = min(ticks, per_tick_ticks)
    for (b, _, _), m in zip(members, mems):
        m["vram"].upload_state(b.state)
        m["vram"].upload_axon_heads(b.heads)
    w0 = time.perf_counter()
    for (b, dop, _), m in zip(members, mems):
        upload_constant_memory(b.lut)
        v = m["vram"]
        for t in range(pt_ticks):
            row = t % input_ticks
            err = lib.cu_step_day_phase(C.byref(v.ptrs), v.padded_n, v.total_axons, 1, t,
                                        m["d_inputs"].ptr + row * b.input_words * 4, b.virtual_offset, b.n_inputs,
                                        None, 0, None, None, 0, dop, st)
            if err:
                raise RuntimeError(f"per-tick step failed {err}")
        lib.gpu_stream_synchronize(st)
    t_per_tick = (time.perf_counter() - w0) * ticks / pt_ticks
    log(f"serial per-tick launches: {pt_ticks} ticks x {n_brains} brains in {time.perf_counter() - w0:.2f}s "
        f"({t_per_tick:.1f}s scaled to {ticks} ticks)")
    lib.gpu_stream_destroy(st)
    rates = [s[0]["spikes"] / max(1, s[0]["ticks"] * b.padded_n) for s, (b, _, _) in zip(stats, members)]
    return {
        "brains": n_brains, "ticks_per_brain": ticks, "neurons": neurons, "synapses": synapses,
        "batched_brains_per_s": n_brains / t_batched, "serial_brains_per_s": n_brains / t_serial,
        "speedup": t_serial / t_batched, "bit_identical_all_brains": identical, "tuning": tune,
        "batched_seconds": t_batched, "serial_seconds": t_serial,
        "serial": (f"each brain alone through axc_step_day_phase_batch (HIP-graph batches), one after another; "
                   f"the faster of windows of 1 and 4 ticks ({best_window})"),
        "serial_seconds_by_window": {str(k): v for k, v in serial_s.items()},
        "serial_per_tick_brains_per_s": n_brains / t_per_tick, "speedup_vs_per_tick": t_per_tick / t_batched,
        "serial_per_tick": f"each brain alone, one cu_step_day_phase call per tick (the reference's launch "
                           f"structure), timed over its first {pt_ticks} ticks and scaled to {ticks}",
        "firing_rate_median": float(np.median(rates)),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--ticks", type=int, default=50, help="ticks per step (one BSP batch)")
    ap.add_argument("--workload", default="c2", choices=["c2", "c1", "c2-linear", "c2-baker", "c5", "c5-sub"],
                    help="c2-linear: the c2 brain with somas numbered x-fastest instead of in bricks; "
                         "c2-baker: numbered as genesis-baker numbers a shard (z-planes, shuffled in each); "
                         "c5: the 12M-neuron x 128-slot scale point; c5-sub: its 200k-neuron sub-shard")
    ap.add_argument("--engine-order", action="store_true",
                    help="give the engine the soma positions (axc_shard_set_soma_positions) so that it steps "
                         "in its own brick order whatever the workload's numbering")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-max-ticks", type=int, default=20)
    ap.add_argument("--cpu-single-ticks", type=int, default=3, help="ticks of the one-thread CPU oracle run")
    ap.add_argument("--nights", type=int, default=2,
                    help="after the timed steps, this many Nights (GPU sort & prune, host sprouting), "
                         "each after one more batch of --ticks ticks, timed phase by phase")
    ap.add_argument("--prune", type=int, default=15, help="prune threshold of the Nights")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--window", type=int, default=4,
                    help="ticks per window of the batch graphs' staged pass (axc_set_window_ticks; 1: per tick)")
    ap.add_argument("--timing-ticks", type=int, default=12,
                    help="ticks per batch of the two batches whose kernels are timed one by one with HIP events")
    ap.add_argument("--batched-brains", type=int, default=128, help="0 skips the batched-brains leg")
    ap.add_argument("--batched-ticks", type=int, default=5000)
    ap.add_argument("--batched-only", action="store_true",
                    help="run only the batched-brains leg (for profiling); prints its JSON object")
    ap.add_argument("--batched-no-serial", action="store_true", help="skip the serial runs of the batched leg")
    ap.add_argument("--copy-calibration-mb", default="",
                    help="comma-separated sizes: one device-to-device copy of each (reference points for the HBM "
                         "byte counters)")
    ap.add_argument("--gather-calibration-mb", default="",
                    help="comma-separated sizes: one gather of every 64-byte row of a buffer of each size in random "
                         "order (torch index_select; scattered full-line reads for the HBM byte counters)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: every GPU box here has one GPU and the engine runs one brain per GPU. "
              "Multi-GPU sharding is described in DESIGN.md and not run.", file=sys.stderr)
        return 2

    from axicor_amd import brain as B
    from axicor_amd._ffi import TIMED_KERNELS, day_lib
    from axicor_amd.engine import DeviceBuffer, ShardEngine, VramState, upload_constant_memory

    lib = day_lib()
    lib.gpu_set_device(0)

    if args.batched_only:
        print(json.dumps({"batched": bench_batched(args.batched_brains, args.batched_ticks, args.batched_ticks, 50,
                                                   20261203, serial=not args.batched_no_serial)}), flush=True)
        return 0

    spec = {"c2": B.spec_c2, "c1": B.spec_c1, "c2-linear": lambda: B.spec_c2(brick=0),
            "c2-baker": lambda: B.spec_c2(brick=B.BAKER_ORDER), "c5": B.spec_c5,
            "c5-sub": lambda: B.spec_c5(200_000)}[args.workload]()
    t0 = time.perf_counter()
    brain = B.generate(spec, with_paths=args.nights > 0)
    N, A = brain.padded_n, brain.total_axons
    log(f"generated {args.workload}: padded_n={N} axons={A} state={brain.state.nbytes / 1e9:.2f} GB "
        f"in {time.perf_counter() - t0:.1f}s")

    T = args.ticks
    total_ticks = (args.warmup + args.steps) * T
    cpu_ticks_cap = 0 if args.no_cpu else args.cpu_max_ticks
    n_in_ticks = max(total_ticks, cpu_ticks_cap)
    t0 = time.perf_counter()
    inputs = brain.inputs(0, n_in_ticks)
    words = brain.input_words
    log(f"inputs {inputs.shape} in {time.perf_counter() - t0:.1f}s")

    # ---- CPU baseline (the oracle on the host cores) -----------------------
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        log(f"cpu oracle: up to {args.cpu_seconds}s / {cpu_ticks_cap} ticks on {threads} threads")
        c_state, c_heads, c_ticks, c_el, c_st = cpu_baseline(brain, inputs, args.cpu_seconds, cpu_ticks_cap)
        cpu = {"state": c_state, "heads": c_heads, "ticks": c_ticks, "elapsed": c_el, "stats": c_st,
               "threads": threads}
        log(f"cpu oracle: {c_ticks} ticks in {c_el:.2f}s, {c_st['slots_scanned'] / c_el:.3e} events/s")
        cpu["single"] = cpu_single_thread(brain, inputs, max(1, min(args.cpu_single_ticks, n_in_ticks)))
        log(f"cpu oracle, one thread: {cpu['single']['ticks']} ticks in {cpu['single']['seconds']:.2f}s, "
            f"{cpu['single']['value']:.3e} events/s")

    # ---- GPU ----------------------------------------------------------------
    upload_constant_memory(brain.lut)
    if lib.axc_set_window_ticks(args.window) != 0:
        raise SystemExit(f"--window {args.window}: at most 4 ticks per window")
    vram = VramState(N, A)
    vram.upload_state(brain.state)
    vram.upload_axon_heads(brain.heads)
    if args.engine_order:
        vram.set_soma_positions(brain.soma_pos)
        log("engine order: steps run in the engine's brick order of the soma positions")
    eng = ShardEngine(vram)
    d_in = DeviceBuffer.from_array(inputs)
    ptrs = C.byref(vram.ptrs)
    tick = 0

    def batch(ticks: int):
        nonlocal tick
        err = lib.axc_step_day_phase_batch(ptrs, N, A, brain.spec.v_seg, tick, ticks, d_in.ptr + tick * words * 4,
                                           words, brain.virtual_offset, brain.n_inputs, None, 0, None, None, None,
                                           0, 0, eng.stream)
        if err != 0:
            raise RuntimeError(f"axc_step_day_phase_batch failed: {err}")
        tick += ticks

    parity = None
    if cpu is not None:
        batch(cpu["ticks"])
        eng.synchronize()
        g_state, g_heads = vram.download_state(), vram.download_axons()
        ok = bool(np.array_equal(g_state, cpu["state"]) and np.array_equal(g_heads, cpu["heads"]))
        parity = {"ticks": cpu["ticks"], "bit_exact": ok}
        log(f"parity vs CPU oracle after {cpu['ticks']} ticks: {'bit-exact' if ok else 'MISMATCH'}")
        del g_state, g_heads
        if not ok:
            print(json.dumps({"error": "GPU state differs from the CPU oracle", "parity": parity}))
            return 1

    # Warmup up to a batch boundary so the timed batches replay one graph.
    rem = (args.warmup * T - tick) if tick < args.warmup * T else 0
    while rem > 0:
        k = min(T, rem)
        batch(k)
        rem -= k
    eng.synchronize()
    log(f"warmup done at tick {tick}")

    hip = C.CDLL("libamdhip64.so")
    for mb in [int(x) for x in str(args.copy_calibration_mb).split(",") if x.strip()]:
        nb = mb << 20
        src, dst = DeviceBuffer(nb), DeviceBuffer(nb)
        hip.hipMemcpy(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), C.c_size_t(nb), 3)  # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        src.free()
        dst.free()
        log(f"copy calibration: {mb} MiB device-to-device")
    for mb in [int(x) for x in str(args.gather_calibration_mb).split(",") if x.strip()]:
        # Reads every 64-byte row once, in a random order (plus 8 B of index
        # per row), and writes the rows out in that order.
        import torch
        rows = (mb << 20) // 64
        g = torch.Generator(device="cuda").manual_seed(7)
        src = torch.empty((rows, 64), dtype=torch.uint8, device="cuda").random_(0, 256, generator=g)
        idx = torch.randperm(rows, device="cuda", generator=g)
        torch.cuda.synchronize()
        out = src.index_select(0, idx)
        torch.cuda.synchronize()
        log(f"gather calibration: {mb} MiB of 64-byte rows in random order ({rows} rows; reads {mb} MiB + "
            f"{rows * 8} B of indices, writes {mb} MiB; checksum {int(out[:, 0].sum())})")
        del src, idx, out

    eng.stats_reset()
    start = C.c_void_p()
    stop = C.c_void_p()
    hip.hipEventCreate(C.byref(start))
    hip.hipEventCreate(C.byref(stop))
    eng.synchronize()
    wall0 = time.perf_counter()
    hip.hipEventRecord(start, C.c_void_p(eng.stream))
    for s in range(args.steps):
        batch(T)
    hip.hipEventRecord(stop, C.c_void_p(eng.stream))
    eng.synchronize()
    wall = time.perf_counter() - wall0
    ms = C.c_float()
    hip.hipEventElapsedTime(C.byref(ms), start, stop)
    elapsed = ms.value / 1e3
    st = eng.stats()
    log(f"timed {args.steps} steps x {T} ticks: {elapsed * 1e3:.1f} ms (wall {wall * 1e3:.1f} ms)")

    ab = algorithmic_bytes(st, N, A, words)
    ticks_timed = st["ticks"]
    events = st["slots_scanned"] / elapsed
    tick_frac = ab["total"] / elapsed / (HBM_PEAK_GBS * 1e9)

    # ---- per-kernel durations (HIP events on the engine's stream) ----------
    # The same batches, launched kernel by kernel so each launch is timed.
    lib.axc_kernel_timing_read(None, None, 1)
    lib.axc_kernel_timing_enable(1)
    eng.stats_reset()
    tb = max(2, args.timing_ticks)
    for k in range(2):
        if tick + tb > n_in_ticks:
            tick = 0  # reuse resident inputs from the start; the state just keeps running
        batch(tb)
    eng.synchronize()
    lib.axc_kernel_timing_enable(0)
    kms = (C.c_double * len(TIMED_KERNELS))()
    kn = (C.c_uint64 * len(TIMED_KERNELS))()
    lib.axc_kernel_timing_read(kms, kn, 1)
    st_t = eng.stats()
    ab_t = algorithmic_bytes(st_t, N, A, words)
    fused = 4 * st_t["gsop_targets_fused"] + 4 * st_t["gsop_slots_fused"]
    per_kernel = {TIMED_KERNELS[k]: {"avg_us": kms[k] * 1e3 / kn[k], "launches": int(kn[k])}
                  for k in range(len(TIMED_KERNELS)) if kn[k]}
    wst = eng.window_stats()
    if "window_pass" in per_kernel:
        # Windows of ticks: the dominant kernel is the window pass. Its own
        # compulsory traffic per launch: each scanned soma's filled slots in
        # the reference's format (4 B target + 2 B weight), every soma's
        # flags, timer and fill byte, one 8 B sum word per soma and tick of
        # the window, and the window's byte snapshot (8 B per axon).
        wp = per_kernel["window_pass"]
        dom_bytes = (6 * wst["pass_slots"] + 3 * N * wst["windows"] + 8 * N * st_t["ticks"] +
                     8 * (A + 1) * wst["windows"]) / wp["launches"]
        dom_name, dom_us = "window_pass", wp["avg_us"]
        dom_desc = (f"window_pass (the staged dendrite scan once per window of {args.window} ticks: "
                    "predicted hits and weight sums per soma and tick)")
    else:
        dom_bytes = (ab_t["neuron_tick"] + fused) / per_kernel["neuron_tick"]["launches"]
        dom_name, dom_us = "neuron_tick", per_kernel["neuron_tick"]["avg_us"]
        dom_desc = "neuron_tick (K1 + K5: staged dendrite scan, GLIF soma update)"
    achieved = dom_bytes / (dom_us * 1e-6) / 1e9
    per_kernel[dom_name]["algorithmic_bytes"] = dom_bytes
    per_kernel[dom_name]["achieved_GBs"] = achieved
    per_kernel["axon_tick"]["algorithmic_bytes"] = ab_t["axon_tick"] / per_kernel["axon_tick"]["launches"]
    per_kernel["axon_tick"]["achieved_GBs"] = (per_kernel["axon_tick"]["algorithmic_bytes"] /
                                               (per_kernel["axon_tick"]["avg_us"] * 1e-6) / 1e9)
    if "gsop" in per_kernel:
        g = (ab_t["gsop"] - fused) / per_kernel["gsop"]["launches"]
        per_kernel["gsop"]["algorithmic_bytes"] = g
        per_kernel["gsop"]["achieved_GBs"] = g / (per_kernel["gsop"]["avg_us"] * 1e-6) / 1e9
    log("per-kernel: " + json.dumps(per_kernel))

    traffic = None
    tf = os.path.join(REPO, "profiles", f"{dom_name}_{args.workload}_traffic.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    out = {
        "metric": METRIC,
        "value": events,
        "unit": "synaptic events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.workload}: {spec.n_neurons} GLIF neurons x {spec.slots_filled} slots, "
                         f"{spec.n_inputs} inputs at {spec.input_ppm / 1e6:g}/tick, axons "
                         f"{spec.axon_len[0]}-{spec.axon_len[1]} segments, GSOP on, "
                         f"soma numbering {_numbering(spec.brick)}"),
            "padded_n": N, "total_axons": A, "ticks_per_step": T, "parallelism": "single GPU",
            "engine_order": bool(args.engine_order), "window_ticks": args.window,
        },
        "roofline": {
            "bound": "hbm", "kernel": dom_desc,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
        },
        "ticks_per_s": ticks_timed / elapsed,
        "neuron_ticks_per_s": N * ticks_timed / elapsed,
        "delivered_events_per_s": st["hits"] / elapsed,
        "gsop_updates_per_s": st["gsop_slots"] / elapsed,
        "firing_rate": st["spikes"] / (N * ticks_timed),
        "tick_hbm_frac": tick_frac,
        "algorithmic_bytes_per_tick": ab["total"] / ticks_timed,
        "us_per_tick": elapsed * 1e6 / ticks_timed,
        "per_kernel": per_kernel,
        "parity": parity,
        "window": ({"ticks": args.window,
                    "correction_pieces_per_tick": wst["correction_pieces"] / max(1, st_t["ticks"]),
                    "corrected_entries_per_tick": wst["corrected_entries"] / max(1, st_t["ticks"]),
                    "pass_slots_per_window": wst["pass_slots"] / max(1, wst["windows"])}
                   if wst["windows"] else None),
    }
    if args.nights:
        def night_batch(ticks: int):
            nonlocal tick
            if tick + ticks > n_in_ticks:
                tick = 0  # reuse resident inputs from the start; the state just keeps running
            batch(ticks)
        out["nights"] = bench_nights(lib, brain, vram, eng, night_batch, args.nights, args.prune)
    if args.batched_brains:
        del vram, eng, d_in
        out["batched"] = bench_batched(args.batched_brains, args.batched_ticks, args.batched_ticks, 50, 20261203,
                                       serial=not args.batched_no_serial)
    if cpu is not None:
        out["cpu_baseline"] = {
            "value": cpu["stats"]["slots_scanned"] / cpu["elapsed"], "unit": "synaptic events/s",
            "cores": cpu["threads"], "kind": "port",
            "sample": f"first {cpu['ticks']} ticks of the same workload, CPU oracle (C, OpenMP over somas)",
            "ticks_per_s": cpu["ticks"] / cpu["elapsed"],
            "single_thread": {"value": cpu["single"]["value"], "unit": "synaptic events/s", "cores": 1,
                              "sample": f"first {cpu['single']['ticks']} ticks of the same workload, one thread"},
            "cpu_model": cpu_model(),
            "nproc": os.cpu_count(),
        }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
