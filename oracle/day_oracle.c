/*
 * day_oracle.c — CPU oracle of the Axicor Day Phase and Night sort&prune.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The
 * product (axicor_amd/) never links or calls it.
 *
 * It restates, in plain C, the integer rules of the reference kernels
 * (precedence: kernel source > Rust contract > docs):
 *   K1 cu_reset_burst_counters_kernel   genesis-compute/src/cuda/physics.cu:25-30
 *   push_burst_head                     physics.cu:32-41
 *   K2 cu_inject_inputs_kernel          physics.cu:88-108
 *   K3 cu_apply_spike_batch_kernel      physics.cu:114-132
 *   K4 cu_propagate_axons_kernel        physics.cu:137-154
 *   K5 cu_update_neurons_kernel         physics.cu:159-276
 *   K6 cu_apply_gsop_kernel             physics.cu:281-373 (HIP twin physics.hip:272-347)
 *   K7 cu_record_readout_kernel         physics.cu:378-395
 *   orchestrator order                  physics.cu:441-495
 *   K9 sort_and_prune_kernel            genesis-compute/src/cuda/bindings.cu:450-515
 *   active-tail test                    genesis-core/src/signal.rs:22-25
 *
 * Integer semantics are fixed explicitly: u32 and i32 arithmetic wraps
 * (two's complement; products are formed in unsigned and reinterpreted),
 * `>>` on a negative i32 is arithmetic (gcc defines it so).
 *
 * Decisions where the reference kernel is racy or undefined:
 *  - K5 reads axon heads while spiking threads of the same launch push them
 *    (physics.cu:194 vs 257-262). The oracle uses the snapshot rule: every
 *    dendrite read sees the heads after K4, pushes of this tick apply after
 *    every soma has integrated, K6 sees the pushed heads. This is one legal
 *    interleaving of the reference launch and is deterministic.
 *  - K3 with repeated ids pushes once per occurrence.
 *  - A dendrite target whose axon index is >= total_axons (or whose packed
 *    axon field is 0) reads as an axon with all heads at the sentinel; a
 *    soma_to_axon or mapped soma id out of range is ignored (reads 0).
 *    The reference would access memory out of bounds there.
 */
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#include <stdlib.h>

#define MAXD 128u
#define SENT 0x80000000u

typedef struct {
    int32_t  threshold, rest_potential, leak_rate, homeostasis_penalty;
    uint32_t spontaneous_firing_period_ticks;
    uint16_t initial_synapse_weight, gsop_potentiation, gsop_depression, homeostasis_decay;
    uint8_t  refractory_period, synapse_refractory_period, signal_propagation_length, is_inhibitory;
    uint8_t  inertia_curve[16];
    int32_t  adaptive_leak_max;
    uint16_t adaptive_leak_gain;
    uint8_t  adaptive_mode, _leak_pad[3];
    uint8_t  d1_affinity, d2_affinity, _pad[4];
} __attribute__((aligned(64))) OrcVariant;

_Static_assert(sizeof(OrcVariant) == 64, "VariantParameters is 64 B");
_Static_assert(offsetof(OrcVariant, inertia_curve) == 32, "inertia_curve @32");
_Static_assert(offsetof(OrcVariant, adaptive_leak_max) == 48, "adaptive_leak_max @48");
_Static_assert(offsetof(OrcVariant, d1_affinity) == 58, "d1_affinity @58");

/* Views into a `.state` blob (genesis-compute/src/memory.rs:71-86). */
typedef struct {
    int32_t *v; uint8_t *flags; int32_t *thr; uint8_t *timers; uint32_t *s2a;
    uint32_t *targets; int16_t *weights; uint8_t *dtimers;
} OrcState;

/* Counters of one call, the same quantities the engine reports. */
typedef struct {
    uint64_t ticks, integrated, slots_scanned, targets_read, hits, spikes, gsop_slots, gsop_targets;
} OrcStats;

static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }

static void views(uint8_t *blob, uint32_t n, OrcState *s) {
    size_t N = n;
    s->v = (int32_t *)(blob);
    s->flags = blob + 4 * N;
    s->thr = (int32_t *)(blob + 5 * N);
    s->timers = blob + 9 * N;
    s->s2a = (uint32_t *)(blob + 10 * N);
    s->targets = (uint32_t *)(blob + 14 * N);
    s->weights = (int16_t *)(blob + 14 * N + 512 * N);
    s->dtimers = blob + 14 * N + 768 * N;
}

/* physics.cu:32-41, applied k times. */
static inline void push_k(uint32_t *h, uint32_t k) {
    for (uint32_t r = 0; r < k && r < 8; r++) {
        for (int j = 7; j > 0; j--) h[j] = h[j - 1];
        h[0] = 0;
    }
}

/* signal.rs:22-25: dist = head - seg (u32 wrap) < prop. */
int orc_active_tail(uint32_t head, uint32_t seg, uint32_t prop) {
    return (uint32_t)(head - seg) < prop;
}

static const uint32_t SENT8[8] = {SENT, SENT, SENT, SENT, SENT, SENT, SENT, SENT};

static inline const uint32_t *heads_of(const uint32_t *heads, uint32_t a_total, uint32_t target) {
    uint32_t a = (target & 0x00FFFFFFu) - 1u;
    return a < a_total ? heads + (size_t)a * 8 : SENT8;
}

/* One GSOP weight update (physics.cu:310-371) given the heads of the
 * presynaptic axon, the segment, the postsynaptic variant, the burst multiplier
 * and dopamine. Exposed for known-answer tests. */
int16_t orc_gsop_update(int16_t w, const uint32_t *h, uint32_t seg, const OrcVariant *p,
                        int32_t burst_mult, int16_t dopamine) {
    uint32_t len = p->signal_propagation_length;
    uint32_t min_dist = 0xFFFFFFFFu;
    for (int k = 0; k < 8; k++) {
        uint32_t d = h[k] - seg;
        uint32_t c = d < len ? d : 0xFFFFFFFFu;
        if (c < min_dist) min_dist = c;
    }
    int is_active = min_dist != 0xFFFFFFFFu;
    int32_t sign = (w >= 0) ? 1 : -1;
    int32_t abs_w = (int32_t)w;
    if (abs_w < 0) abs_w = -abs_w;
    uint32_t rank = (uint32_t)abs_w >> 11;
    if (rank > 15) rank = 15;
    int32_t inertia = p->inertia_curve[rank];
    int32_t pot_mod = wmul((int32_t)dopamine, (int32_t)p->d1_affinity) >> 7;
    int32_t dep_mod = wmul((int32_t)dopamine, (int32_t)p->d2_affinity) >> 7;
    int32_t raw_pot = wadd((int32_t)p->gsop_potentiation, pot_mod);
    int32_t raw_dep = wsub((int32_t)p->gsop_depression, dep_mod);
    int32_t final_dep = raw_dep & ~(raw_dep >> 31);
    int32_t delta_pot = wmul(wmul(raw_pot, inertia), burst_mult) >> 7;
    int32_t delta_dep = wmul(wmul(final_dep, inertia), burst_mult) >> 7;
    uint32_t cs = is_active ? (min_dist >> 4) : 0;
    int32_t delta = is_active ? (delta_pot >> cs) : wsub(0, delta_dep);
    delta = wmul(delta, 128) >> (7 + cs);
    int32_t new_abs = wadd(abs_w, delta);
    new_abs = new_abs & ~(new_abs >> 31);
    if (new_abs > 32767) new_abs = 32767;
    return (int16_t)wmul(new_abs, sign);
}

/*
 * One tick of cu_step_day_phase (physics.cu:441-495) with the snapshot rule.
 * blob: `.state` bytes (910*n), heads: u32[8*a_total].
 * bitmask may be NULL (no inputs); spikes may be NULL; mapped/out may be NULL.
 * out (if given) receives n_out bytes for this tick.
 * Returns the number of somas that spiked.
 */
uint32_t orc_step_tick(uint8_t *blob, uint32_t n, uint32_t *heads, uint32_t a_total,
                       const OrcVariant *lut, uint32_t v_seg, uint32_t tick,
                       const uint32_t *bitmask, uint32_t virt_off, uint32_t n_virt,
                       const uint32_t *spikes, uint32_t n_spk,
                       const uint32_t *mapped, uint8_t *out, uint32_t n_out,
                       int16_t dopamine, OrcStats *st, uint8_t *scratch_push) {
    OrcState s;
    views(blob, n, &s);
    size_t N = n;

    /* K1 (physics.cu:25-30). */
    for (size_t i = 0; i < N; i++) s.flags[i] &= 0xF1;

    /* K2 (physics.cu:88-108): bit i of the mask pushes axon virt_off+i. */
    if (n_virt > 0 && bitmask) {
        for (uint32_t i = 0; i < n_virt; i++) {
            if ((bitmask[i >> 5] >> (i & 31)) & 1u) {
                uint32_t a = virt_off + i;
                if (a < a_total) push_k(heads + (size_t)a * 8, 1);
            }
        }
    }
    /* K3 (physics.cu:114-132): ids < total_axons push, once per occurrence. */
    if (n_spk > 0 && spikes) {
        for (uint32_t i = 0; i < n_spk; i++) {
            uint32_t a = spikes[i];
            if (a < a_total) push_k(heads + (size_t)a * 8, 1);
        }
    }
    /* K4 (physics.cu:137-154). */
    for (size_t a = 0; a < a_total; a++) {
        uint32_t *h = heads + a * 8;
        for (int k = 0; k < 8; k++)
            if (h[k] != SENT) h[k] += v_seg;
    }

    /* K5 (physics.cu:159-276) against the post-K4 heads; pushes deferred. */
    uint64_t integrated = 0, scanned = 0, tread = 0, hits = 0;
    #pragma omp parallel for schedule(static) reduction(+:integrated,scanned,tread,hits)
    for (size_t i = 0; i < N; i++) {
        uint8_t flags = s.flags[i];
        uint8_t timer = s.timers[i];
        flags &= (uint8_t)~0x01;
        scratch_push[i] = 0;
        if (timer > 0) {
            s.timers[i] = (uint8_t)(timer - 1);
            s.flags[i] = flags;
            continue;
        }
        integrated++;
        const OrcVariant *p = &lut[(flags >> 4) & 0x0F];
        int32_t v = s.v[i];
        int32_t i_in = 0;
        uint32_t prop = p->signal_propagation_length;
        uint32_t sl;
        for (sl = 0; sl < MAXD; sl++) {
            size_t col = (size_t)sl * N + i;
            uint32_t t = s.targets[col];
            if (t == 0) break;
            uint32_t seg = t >> 24;
            const uint32_t *h = heads_of(heads, a_total, t);
            int hit = 0;
            for (int k = 0; k < 8; k++) hit |= (uint32_t)(h[k] - seg) < prop;
            if (hit) { i_in = wadd(i_in, (int32_t)s.weights[col]); hits++; }
        }
        scanned += sl;
        tread += (sl < MAXD) ? sl + 1 : MAXD;

        int32_t thr = s.thr[i];
        int32_t decayed = wsub(thr, (int32_t)p->homeostasis_decay);
        thr = decayed & ~(decayed >> 31);

        int32_t leak = p->leak_rate;
        if (p->adaptive_mode == 1) {
            int32_t add = wmul(thr, (int32_t)p->adaptive_leak_gain) >> 8;
            leak = wadd(leak, add);
            int32_t over = wsub(leak, p->adaptive_leak_max);
            leak = wsub(leak, over & ~(over >> 31));
        }
        v = wadd(v, i_in);
        int32_t diff = wsub(v, p->rest_potential);
        int32_t sign = (diff > 0) - (diff < 0);
        int32_t abs_diff = wmul(diff, sign);
        int32_t leaked = wsub(abs_diff, leak);
        leaked = leaked & ~(leaked >> 31);
        v = wadd(p->rest_potential, wmul(sign, leaked));

        int32_t eff = wadd(p->threshold, thr);
        int32_t glif = v >= eff ? 1 : 0;
        uint32_t period = p->spontaneous_firing_period_ticks;
        int32_t hb = (period > 0 && ((uint32_t)(tick + (uint32_t)i * 104729u) % period) == 0) ? 1 : 0;
        int32_t fin = glif | hb;

        if (glif) v = p->rest_potential;
        if (glif) thr = wadd(thr, p->homeostasis_penalty);
        uint8_t new_timer = glif ? p->refractory_period : s.timers[i];

        if (fin) {
            flags = (uint8_t)((flags & 0xFC) | 0x03);
            scratch_push[i] = 1;
        } else {
            flags = (uint8_t)(flags & ~0x01);
        }
        s.v[i] = v;
        uint8_t bc = (flags >> 1) & 0x07;
        bc = (uint8_t)(bc + (fin && bc < 7));
        s.flags[i] = (uint8_t)((flags & 0xF0) | (bc << 1) | (uint8_t)fin);
        s.thr[i] = thr;
        s.timers[i] = new_timer;
    }

    /* Deferred pushes of this tick's spikes (physics.cu:253-262). */
    uint32_t nspk = 0;
    for (size_t i = 0; i < N; i++) {
        if (!scratch_push[i]) continue;
        nspk++;
        uint32_t a = s.s2a[i];
        if (a != 0xFFFFFFFFu && a < a_total) push_k(heads + (size_t)a * 8, 1);
    }

    /* K6 (physics.cu:281-373) with the pushed heads. */
    uint64_t gslots = 0, gtargets = 0;
    #pragma omp parallel for schedule(dynamic, 256) reduction(+:gslots,gtargets)
    for (size_t i = 0; i < N; i++) {
        uint8_t flags = s.flags[i];
        if ((flags & 0x01) == 0) continue;
        uint8_t bc = (flags >> 1) & 0x07;
        int32_t bm = bc > 0 ? bc : 1;
        const OrcVariant *p = &lut[(flags >> 4) & 0x0F];
        uint32_t sl;
        for (sl = 0; sl < MAXD; sl++) {
            size_t col = (size_t)sl * N + i;
            uint32_t t = s.targets[col];
            if (t == 0) break;
            const uint32_t *h = heads_of(heads, a_total, t);
            s.weights[col] = orc_gsop_update(s.weights[col], h, t >> 24, p, bm, dopamine);
        }
        gslots += sl;
        gtargets += (sl < MAXD) ? sl + 1 : MAXD;
    }

    /* K7 (physics.cu:378-395). */
    if (n_out > 0 && mapped && out) {
        for (uint32_t j = 0; j < n_out; j++) {
            uint32_t id = mapped[j];
            out[j] = (id != 0xFFFFFFFFu && id < n) ? (s.flags[id] & 0x01) : 0;
        }
    }

    if (st) {
        st->ticks += 1;
        st->integrated += integrated;
        st->slots_scanned += scanned;
        st->targets_read += tread;
        st->hits += hits;
        st->spikes += nspk;
        st->gsop_slots += gslots;
        st->gsop_targets += gtargets;
    }
    return nspk;
}

/*
 * T ticks with per-tick input strides, as ShardEngine::step_day_phase_batch
 * does (genesis-compute/src/compute/shard.rs:40-125): one K1 at the batch
 * start (shard.rs:71; idempotent with the per-tick K1), then T ticks.
 */
void orc_step_batch(uint8_t *blob, uint32_t n, uint32_t *heads, uint32_t a_total,
                    const OrcVariant *lut, uint32_t v_seg, uint32_t tick_base, uint32_t T,
                    const uint32_t *bitmask, uint32_t words_per_tick,
                    uint32_t virt_off, uint32_t n_virt,
                    const uint32_t *spikes, uint32_t max_spikes_per_tick, const uint32_t *spike_counts,
                    const uint32_t *mapped, uint8_t *out, uint32_t n_out,
                    int16_t dopamine, OrcStats *st) {
    uint8_t *scratch = (uint8_t *)malloc(n ? n : 1);
    OrcState s;
    views(blob, n, &s);
    for (size_t i = 0; i < n; i++) s.flags[i] &= 0xF1;
    for (uint32_t t = 0; t < T; t++) {
        orc_step_tick(blob, n, heads, a_total, lut, v_seg, tick_base + t,
                      bitmask ? bitmask + (size_t)t * words_per_tick : NULL, virt_off, n_virt,
                      spikes ? spikes + (size_t)t * max_spikes_per_tick : NULL,
                      spike_counts ? spike_counts[t] : 0,
                      mapped, out ? out + (size_t)t * n_out : NULL, n_out, dopamine, st, scratch);
    }
    free(scratch);
}

/*
 * K9 sort_and_prune_kernel (bindings.cu:450-515), applied to columnar arrays.
 * Insertion sort of the 128 slots: a slot with target==0 is never picked up as
 * a key; a filled key moves left past empty slots and past filled slots with a
 * smaller |w|. Then filled slots with |w| < prune become (0,0,0).
 */
void orc_sort_and_prune(uint8_t *flags, uint32_t *targets, int16_t *weights, uint8_t *timers,
                        uint32_t n, int16_t prune) {
    size_t N = n;
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < N; i++) {
        uint32_t t[MAXD]; int16_t w[MAXD]; uint8_t tm[MAXD];
        flags[i] &= 0xF1;
        for (uint32_t sl = 0; sl < MAXD; sl++) {
            size_t col = (size_t)sl * N + i;
            t[sl] = targets[col]; w[sl] = weights[col]; tm[sl] = timers[col];
        }
        for (int k = 1; k < (int)MAXD; k++) {
            uint32_t kt = t[k]; int16_t kw = w[k]; uint8_t ktm = tm[k];
            if (kt == 0) continue;
            int32_t kabs = kw >= 0 ? kw : -(int32_t)kw;
            int j = k - 1;
            while (j >= 0) {
                int32_t pabs = w[j] >= 0 ? w[j] : -(int32_t)w[j];
                if (t[j] != 0 && pabs >= kabs) break;
                t[j + 1] = t[j]; w[j + 1] = w[j]; tm[j + 1] = tm[j];
                j--;
            }
            t[j + 1] = kt; w[j + 1] = kw; tm[j + 1] = ktm;
        }
        for (uint32_t sl = 0; sl < MAXD; sl++) {
            int32_t a = w[sl] >= 0 ? w[sl] : -(int32_t)w[sl];
            if (t[sl] != 0 && a < prune) { t[sl] = 0; w[sl] = 0; tm[sl] = 0; }
        }
        for (uint32_t sl = 0; sl < MAXD; sl++) {
            size_t col = (size_t)sl * N + i;
            targets[col] = t[sl]; weights[col] = w[sl]; timers[col] = tm[sl];
        }
    }
}

/* State-blob variant of K9 for convenience. */
void orc_sort_and_prune_blob(uint8_t *blob, uint32_t n, int16_t prune) {
    OrcState s;
    views(blob, n, &s);
    orc_sort_and_prune(s.flags, s.targets, s.weights, s.dtimers, n, prune);
}

/* genesis-core/src/seed.rs:33-43 (FNV-1a 64). */
uint64_t orc_seed_from_str(const char *str, size_t len) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < len; i++) { h ^= (uint8_t)str[i]; h *= 0x00000100000001B3ull; }
    return h;
}

/* genesis-core/src/seed.rs:50-57 (128-bit multiply-fold avalanche). */
uint64_t orc_entity_seed(uint64_t master, uint32_t id) {
    uint64_t seed = master + (uint64_t)id + 0x60bee2bee120fc15ull;
    unsigned __int128 tmp = (unsigned __int128)seed * 0xa3b195354a39b70dull;
    unsigned __int128 m1 = (tmp >> 64) ^ tmp;
    tmp = m1 * (unsigned __int128)0x1b03738712fad5c9ull;
    return (uint64_t)((tmp >> 64) ^ tmp);
}

/* genesis-core/src/hash.rs:8-17 (FNV-1a 32). */
uint32_t orc_fnv1a_32(const char *s, size_t len) {
    uint32_t h = 0x811c9dc5u;
    for (size_t i = 0; i < len; i++) { h ^= (uint8_t)s[i]; h *= 0x01000193u; }
    return h;
}

/* genesis-core/src/layout.rs:254-265 (pack) and 268-277 (unpack). */
uint32_t orc_pack_target(uint32_t axon, uint32_t seg) { return (seg << 24) | ((axon + 1) & 0x00FFFFFFu); }
uint32_t orc_unpack_axon(uint32_t t) { uint32_t a = t & 0x00FFFFFFu; return a ? a - 1 : 0; }
