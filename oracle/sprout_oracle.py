"""Pure-Python oracle of the Night-Phase sprouting pass, for small brains.

TEST INFRASTRUCTURE ONLY (the checker of axicor_amd/csrc/night_host.cpp).
Restates genesis-baker/src/bake/sprouting.rs:119-378 (run_sprouting_pass),
nudge_axon (51-117), compute_power_index (32-40), AxonSegmentGrid
(spatial_grid.rs:99-185, a dict of cells filled in reverse axon order) and
random_f32 (genesis-core/src/seed.rs:61-64), line by line, with f32
arithmetic done in numpy.float32."""
from __future__ import annotations

import numpy as np

F32 = np.float32
MAXD = 128


def random_f32(seed: int) -> np.float32:
    bits = ((seed & 0xFFFFFFFFFFFFFFFF) >> 41) | 0x3F800000
    return np.array([bits], np.uint32).view(np.float32)[0] - F32(1.0)


def unpack_axon_id(t: int) -> int:
    a = t & 0x00FFFFFF
    return a - 1 if a > 0 else 0


def pos_xyzt(p: int):
    return p & 0x3FF, (p >> 10) & 0x3FF, (p >> 20) & 0xFF, p >> 28


def hash_cell(cx, cy, cz):  # spatial_grid.rs:41-44
    return (cx & 0xFFF) | ((cy & 0xFFF) << 12) | ((cz & 0xFF) << 24)


def sprout_pass(targets, weights, flags, soma_to_axon, padded_n, lengths, paths, tips, dirs, soma_positions,
                types, epoch, master_seed, max_x, max_y, total_ghosts=0, ghost_origins=None, incoming=(),
                zone_hash=0, max_sprouts=4):
    """In place on numpy arrays (targets/weights [128, N]). types: list of dicts
    with w_distance, w_power, w_explore, w_type, initial_synapse_weight,
    is_inhibitory (or None). Returns (new_synapses, handovers, acks, prunes)."""
    N = padded_n
    A = len(tips)
    gs, ge = N, N + total_ghosts
    active = [False] * A
    for t in targets.reshape(-1):
        t = int(t)
        if t != 0:
            a = unpack_axon_id(t)
            if a < A:
                active[a] = True
    a2s = [None] * A
    for s in range(N):
        a = int(soma_to_axon[s])
        if a != 0xFFFFFFFF and a < A:
            a2s[a] = s
    acks = []
    nf = gs
    for ev in incoming:
        while nf < ge and tips[nf] != 0:
            nf += 1
        if nf >= ge:
            break
        acks.append((ev["origin_zone_hash"], ev["local_axon_id"], nf))
        tip = (ev["type_mask"] << 28) | (ev["entry_z"] << 22) | (ev["entry_y"] << 11) | ev["entry_x"]
        d = ((ev["vector_z"] & 0xFF) << 16) | ((ev["vector_y"] & 0xFF) << 8) | (ev["vector_x"] & 0xFF)
        tips[nf] = tip
        dirs[nf] = d
        lengths[nf] = 0
        paths[nf, 0] = tip
        nf += 1
    hand = []

    def nudge(a):
        tip = int(tips[a])
        if tip == 0:
            return
        tx, ty, tz, tm = pos_xyzt(tip)
        d = int(dirs[a])
        dx, dy, dz = [((d >> sh) & 0xFF) - (256 if (d >> sh) & 0x80 else 0) for sh in (0, 8, 16)]
        sh = lambda v: 1 if v > 64 else (-1 if v < -64 else 0)  # noqa: E731
        nx, ny, nz = tx + sh(dx), ty + sh(dy), tz + sh(dz)
        if nx < 0 or nx >= max_x or ny < 0 or ny >= max_y or nz < 0 or nz > 63:
            if len(hand) < 10_000:
                ln = int(lengths[a])
                hand.append(dict(origin_zone_hash=zone_hash, local_axon_id=a, entry_x=min(max(nx, 0), 2047),
                                 entry_y=min(max(ny, 0), 2047), vector_x=dx, vector_y=dy, vector_z=dz,
                                 type_mask=tm, remaining_length=max(256 - ln, 0), entry_z=min(max(nz, 0), 63)))
            tips[a] = 0
            return
        nt = (nx & 0x3FF) | ((ny & 0x3FF) << 10) | ((nz & 0xFF) << 20) | ((tm & 0xF) << 28)
        tips[a] = nt
        ln = int(lengths[a])
        if ln < 256:
            paths[a, ln] = nt
            lengths[a] = (ln + 1) & 0xFF

    for s in range(N):
        if flags[s] & 0x02:
            a = int(soma_to_axon[s])
            if a != 0xFFFFFFFF and a < A:
                nudge(a)
    for a in range(gs, ge):
        nudge(a)

    cells = {}
    for a in range(A - 1, -1, -1):
        for si in range(int(lengths[a])):
            p = int(paths[a, si])
            if p == 0:
                continue
            x, y, z, t = pos_xyzt(p)
            cells.setdefault(hash_cell(x // 2, y // 2, z // 2), []).append((a, si, t))

    # compute_power_index (sprouting.rs:32-40) of a soma, memoised: only the
    # soma being filled changes its weights, and its entry is dropped then.
    power_memo = {}

    def power_index(soma):
        if soma not in power_memo:
            pw = 0
            for sl in range(MAXD):
                pw += abs(int(weights[sl, soma]))
            power_memo[soma] = F32(pw) / (F32(MAXD) * F32(32767.0))
        return power_memo[soma]

    new = 0
    for i in range(N):
        mp = int(soma_positions[i])
        if mp == 0:
            continue
        if flags[i] & 0x0E == 0:
            continue
        mx, my, mz, mt = pos_xyzt(mp)
        cfg = types[mt] if (types is not None and mt < len(types)) else None
        sprouts = 0
        # The axons already on the soma (the reference's scan of all 128
        # slots per candidate), kept as a set that grows with each sprout.
        present = {unpack_axon_id(int(targets[e, i])) for e in range(MAXD) if int(targets[e, i]) != 0}
        for slot in range(MAXD - 1, -1, -1):
            if targets[slot, i] != 0:
                break
            best, best_score = None, F32(-1.0)
            cx, cy, cz = mx // 2, my // 2, mz // 2
            for z in range(cz - 2, cz + 3):
                if z < 0:
                    continue
                for y in range(cy - 2, cy + 3):
                    if y < 0:
                        continue
                    for x in range(cx - 2, cx + 3):
                        if x < 0:
                            continue
                        for (a, si, t) in cells.get(hash_cell(x, y, z), []):
                            if int(soma_to_axon[i]) == a:
                                continue
                            if a in present:
                                continue
                            same = F32(1.0) if mt == t else F32(0.0)
                            noise = random_f32(master_seed + a + epoch)
                            owner = a2s[a]
                            power = F32(1.0) if owner is None else power_index(owner)
                            score = F32(1.0)
                            if cfg is not None:
                                score = F32(cfg["w_distance"]) * F32(1.0)
                                score = score + F32(cfg["w_explore"]) * noise
                                score = score + F32(cfg["w_type"]) * same
                                score = score + F32(cfg["w_power"]) * power
                            if score > best_score:
                                best_score, best = score, (a, si, t)
            if best is None:
                break
            a, si, t = best
            targets[slot, i] = (si << 24) | ((a + 1) & 0x00FFFFFF)
            if types is not None and t < len(types):
                inh, w0 = bool(types[t]["is_inhibitory"]), int(np.int16(np.uint16(types[t]["initial_synapse_weight"])))
            else:
                inh, w0 = False, 74
            weights[slot, i] = -w0 if inh else w0
            present.add(a)
            power_memo.pop(i, None)
            active[a] = True
            new += 1
            sprouts += 1
            if sprouts >= max_sprouts:
                break
    prunes = []
    for g in range(gs, ge):
        if not active[g] and tips[g] != 0:
            origin = int(ghost_origins[g - gs]) if ghost_origins is not None else 0
            if origin != 0:
                prunes.append((origin, g))
                tips[g] = 0
    return new, hand, acks, prunes[:1000]
