// day_kernels.hip — gfx950 kernels of the Day Phase and the Night sort&prune.
//
// Integer rules restated from genesis-compute/src/cuda/physics.cu (K1-K7) and
// bindings.cu (K9-K12); the per-kernel comments cite the lines. Data stay in
// the reference's SoA layout (columnar dendrites, BurstHeads8 heads), so the
// device pointers handed out by cu_allocate_shard keep their meaning.
//
// Work split of one tick (all on the caller's stream):
//   axon_tick      K2 + K3 counts + K4, one lane per axon, 32 B RMW
//   neuron_tick    K1 + K5 integrate; spikers compacted into a list with a
//                  wave ballot and one atomic per workgroup; the push of the
//                  own axon is counted per axon, not applied
//   push_apply     applies the counted pushes (snapshot rule: every dendrite
//                  read of neuron_tick saw the post-K4 heads)
//   gsop           K6 over the spiker list: one wave per spiker, one lane
//                  per dendrite slot (two rounds of 64)
//   readout        K7
// Everything is integer VALU work; nothing here is a dense contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "day_common.h"

namespace axc {

__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
__device__ __forceinline__ int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }

struct Heads8 { uint4 lo, hi; };

__device__ __forceinline__ Heads8 load_heads(const AxcBurstHeads8* __restrict__ heads, uint32_t a) {
  const uint4* p = reinterpret_cast<const uint4*>(heads + a);
  Heads8 h;
  h.lo = p[0];
  h.hi = p[1];
  return h;
}

__device__ __forceinline__ void store_heads(AxcBurstHeads8* __restrict__ heads, uint32_t a, const Heads8& h) {
  uint4* p = reinterpret_cast<uint4*>(heads + a);
  p[0] = h.lo;
  p[1] = h.hi;
}

// k pushes of push_burst_head (physics.cu:32-41): h[j] <- h[j-k], h[0..k) <- 0.
__device__ __forceinline__ void push_heads(Heads8& h, uint32_t k) {
  if (k == 0) return;
  uint32_t v[8] = {h.lo.x, h.lo.y, h.lo.z, h.lo.w, h.hi.x, h.hi.y, h.hi.z, h.hi.w};
  uint32_t o[8];
#pragma unroll
  for (int j = 0; j < 8; j++) o[j] = (j >= (int)k) ? v[(j - (int)k) & 7] : 0u;
  h.lo = make_uint4(o[0], o[1], o[2], o[3]);
  h.hi = make_uint4(o[4], o[5], o[6], o[7]);
}

// Any head in the active tail of `seg` (physics.cu:198-201; signal.rs:22-25).
__device__ __forceinline__ bool any_hit(const Heads8& h, uint32_t seg, uint32_t prop) {
  return ((h.lo.x - seg) < prop) | ((h.lo.y - seg) < prop) | ((h.lo.z - seg) < prop) |
         ((h.lo.w - seg) < prop) | ((h.hi.x - seg) < prop) | ((h.hi.y - seg) < prop) |
         ((h.hi.z - seg) < prop) | ((h.hi.w - seg) < prop);
}

__device__ __forceinline__ uint32_t dist_or_max(uint32_t head, uint32_t seg, uint32_t len) {
  uint32_t d = head - seg;
  return d < len ? d : 0xFFFFFFFFu;
}

// Heads of a packed target; an axon index outside [0, A) reads as all
// sentinels (the reference would read out of bounds).
__device__ __forceinline__ Heads8 target_heads(const AxcBurstHeads8* __restrict__ heads, uint32_t A, uint32_t t) {
  uint32_t a = (t & 0x00FFFFFFu) - 1u;
  if (a < A) return load_heads(heads, a);
  Heads8 s;
  s.lo = make_uint4(kSentinel, kSentinel, kSentinel, kSentinel);
  s.hi = s.lo;
  return s;
}

__device__ __forceinline__ uint32_t tick_of(const TickParams* prm, uint32_t tick_arg) {
  return prm ? prm->tick_base + tick_arg : tick_arg;
}
__device__ __forceinline__ int32_t dopamine_of(const TickParams* prm, int32_t dop_arg) {
  return prm ? (int32_t)prm->dopamine : dop_arg;
}

// Compact form of eight heads (see kInertHead in day_common.h).
__device__ __forceinline__ uint32_t c16(uint32_t h) { return h < 0xFFFFu ? h : 0xFFFFu; }
__device__ __forceinline__ uint4 compact_heads(const Heads8& h) {
  return make_uint4(c16(h.lo.x) | (c16(h.lo.y) << 16), c16(h.lo.z) | (c16(h.lo.w) << 16),
                    c16(h.hi.x) | (c16(h.hi.y) << 16), c16(h.hi.z) | (c16(h.hi.w) << 16));
}

// Byte form of eight heads, min(h, 255) each, for the staged scan: for a
// segment s and length L with s + L <= 255 (the u8 ranges of both make every
// other case rare), "(h - s) < L" and the freshest distance below L come out
// the same from min(h, 255) as from h, since any h >= 255 fails both tests.
__device__ __forceinline__ uint32_t b8(uint32_t h) { return h < 0xFFu ? h : 0xFFu; }
__device__ __forceinline__ uint2 byte_heads(const Heads8& h) {
  return make_uint2(b8(h.lo.x) | (b8(h.lo.y) << 8) | (b8(h.lo.z) << 16) | (b8(h.lo.w) << 24),
                    b8(h.hi.x) | (b8(h.hi.y) << 8) | (b8(h.hi.z) << 16) | (b8(h.hi.w) << 24));
}

// Smallest (c_k - seg) mod 2^16 over the eight compact heads: < prop means
// "some head in the active tail" (physics.cu:198-201), and for GSOP it is the
// freshest distance min_dist of physics.cu:311-318 whenever it is < len.
// Packed 16-bit subtract and min: nine VALU instructions for eight heads.
typedef unsigned short axc_us2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t min_dist16(const uint4& c, uint32_t seg) {
  const axc_us2 s = {(unsigned short)seg, (unsigned short)seg};
  const axc_us2 d0 = __builtin_bit_cast(axc_us2, c.x) - s;
  const axc_us2 d1 = __builtin_bit_cast(axc_us2, c.y) - s;
  const axc_us2 d2 = __builtin_bit_cast(axc_us2, c.z) - s;
  const axc_us2 d3 = __builtin_bit_cast(axc_us2, c.w) - s;
  const axc_us2 m = __builtin_elementwise_min(__builtin_elementwise_min(d0, d1), __builtin_elementwise_min(d2, d3));
  return min((uint32_t)m.x, (uint32_t)m.y);
}

// min_dist16 over the byte form: the eight bytes are spread to u16 lanes with
// v_perm_b32, then the same packed subtract and min. Equal to min_dist16 on
// the compact form whenever seg + L <= 255 and the result is compared with L.
__device__ __forceinline__ uint32_t min_dist8(const uint2& c, uint32_t seg) {
  const uint32_t p0 = __builtin_amdgcn_perm(0u, c.x, 0x0c010c00u), p1 = __builtin_amdgcn_perm(0u, c.x, 0x0c030c02u);
  const uint32_t p2 = __builtin_amdgcn_perm(0u, c.y, 0x0c010c00u), p3 = __builtin_amdgcn_perm(0u, c.y, 0x0c030c02u);
  const axc_us2 s = {(unsigned short)seg, (unsigned short)seg};
  const axc_us2 d0 = __builtin_bit_cast(axc_us2, p0) - s;
  const axc_us2 d1 = __builtin_bit_cast(axc_us2, p1) - s;
  const axc_us2 d2 = __builtin_bit_cast(axc_us2, p2) - s;
  const axc_us2 d3 = __builtin_bit_cast(axc_us2, p3) - s;
  const axc_us2 m = __builtin_elementwise_min(__builtin_elementwise_min(d0, d1), __builtin_elementwise_min(d2, d3));
  return min((uint32_t)m.x, (uint32_t)m.y);
}

// Axon index of a packed target for the snapshot: out-of-range indices (and
// the empty-slot pattern) go to the inert entry A.
__device__ __forceinline__ uint32_t snap_index(uint32_t t, uint32_t A) {
  uint32_t a = (t & 0x00FFFFFFu) - 1u;
  return a < A ? a : A;
}

// One GSOP weight update (physics.cu:320-371) from the freshest distance md
// (>= len: no active head), with the per-soma terms precomputed.
__device__ __forceinline__ int16_t gsop_weight(int16_t w, uint32_t md, uint32_t len, int32_t raw_pot,
                                               int32_t final_dep, int32_t bm, const uint8_t* curve) {
  const bool act = md < len;
  const int32_t sign = (w >= 0) ? 1 : -1;
  int32_t abs_w = (int32_t)w;
  abs_w = abs_w < 0 ? -abs_w : abs_w;
  uint32_t rank = (uint32_t)abs_w >> 11;
  rank = rank > 15 ? 15 : rank;
  const int32_t inertia = curve[rank];
  const int32_t dpot = wmul(wmul(raw_pot, inertia), bm) >> 7;
  const int32_t ddep = wmul(wmul(final_dep, inertia), bm) >> 7;
  const uint32_t cs = act ? (md >> 4) : 0u;
  int32_t delta = act ? (dpot >> cs) : wsub(0, ddep);
  delta = wmul(delta, 128) >> (7 + cs);
  int32_t na = wadd(abs_w, delta);
  na = na & ~(na >> 31);
  na = na > 32767 ? 32767 : na;
  return (int16_t)wmul(na, sign);
}

struct GsopTerms {
  int32_t raw_pot, final_dep, bm;
  uint32_t len;
};

// Per-soma GSOP terms (physics.cu:289-345) for a soma with flags `flags`.
__device__ __forceinline__ GsopTerms gsop_terms(uint8_t flags, const AxcVariantParameters& p, int32_t dop) {
  GsopTerms g;
  const uint8_t bc = (flags >> 1) & 0x07;
  g.bm = bc > 0 ? bc : 1;
  g.len = p.signal_propagation_length;
  const int32_t pot_mod = wmul(dop, (int32_t)p.d1_affinity) >> 7;
  const int32_t dep_mod = wmul(dop, (int32_t)p.d2_affinity) >> 7;
  g.raw_pot = wadd((int32_t)p.gsop_potentiation, pot_mod);
  const int32_t raw_dep = wsub((int32_t)p.gsop_depression, dep_mod);
  g.final_dep = raw_dep & ~(raw_dep >> 31);
  return g;
}

// ---------------------------------------------------------------------------
// K2 + K3 + K4: pushes from the input bitmask and the spike-schedule counts,
// then propagation (physics.cu:88-154). Pushes commute with each other, so the
// k pushes an axon receives this tick are applied at once, before propagation.
// Also writes the compact snapshot of the propagated heads.
// ---------------------------------------------------------------------------
// axon_inv (may be null): the caller's id of each of the engine's axons,
// which sets the axon's bit in the input bitmask. Returns the pushes applied;
// out and pre (may be null) receive the new heads and the heads before.
__device__ __forceinline__ uint32_t axon_step(AxcBurstHeads8* __restrict__ heads, uint32_t a, uint32_t v_seg,
                                          const uint32_t* __restrict__ bitmask, uint32_t virt_off, uint32_t n_virt,
                                          uint32_t* __restrict__ push_cnt, bool use_cnt, uint4* __restrict__ comp,
                                          uint2* __restrict__ comp8 = nullptr,
                                          const uint32_t* __restrict__ axon_inv = nullptr,
                                          Heads8* out = nullptr, Heads8* pre = nullptr) {
  Heads8 h = load_heads(heads, a);
  if (pre) *pre = h;
  uint32_t k = 0;
  if (bitmask) {
    uint32_t i = (axon_inv ? axon_inv[a] : a) - virt_off;  // wraps for ids below virt_off
    if (i < n_virt) k += (bitmask[i >> 5] >> (i & 31)) & 1u;
  }
  if (use_cnt) {
    uint32_t c = push_cnt[a];
    if (c) {
      k += c;
      push_cnt[a] = 0;
    }
  }
  push_heads(h, k > 8 ? 8 : k);
  h.lo.x += (h.lo.x != kSentinel) ? v_seg : 0u;
  h.lo.y += (h.lo.y != kSentinel) ? v_seg : 0u;
  h.lo.z += (h.lo.z != kSentinel) ? v_seg : 0u;
  h.lo.w += (h.lo.w != kSentinel) ? v_seg : 0u;
  h.hi.x += (h.hi.x != kSentinel) ? v_seg : 0u;
  h.hi.y += (h.hi.y != kSentinel) ? v_seg : 0u;
  h.hi.z += (h.hi.z != kSentinel) ? v_seg : 0u;
  h.hi.w += (h.hi.w != kSentinel) ? v_seg : 0u;
  store_heads(heads, a, h);
  if (comp) comp[a] = compact_heads(h);
  if (comp8) comp8[a] = byte_heads(h);
  if (out) *out = h;
  return k;
}

__global__ __launch_bounds__(256) void axon_tick_kernel(AxcBurstHeads8* __restrict__ heads, uint32_t A,
                                                        uint32_t v_seg, const uint32_t* __restrict__ bitmask_arg,
                                                        uint32_t virt_off, uint32_t n_virt,
                                                        uint32_t* __restrict__ push_cnt, int use_cnt,
                                                        uint32_t* __restrict__ counters,
                                                        const TickParams* __restrict__ prm, uint32_t t_off,
                                                        uint32_t words, uint4* __restrict__ comp,
                                                        uint2* __restrict__ comp8,
                                                        const uint32_t* __restrict__ axon_inv = nullptr) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t* __restrict__ bitmask =
      prm ? (prm->bitmask ? prm->bitmask + (size_t)t_off * words : nullptr) : bitmask_arg;
  // The list counters of the previous tick were consumed by its gsop and
  // push_apply launches; this tick's neuron_tick fills them again.
  if (a < kNumCounters) counters[a] = 0;
  if (a >= A) return;
  axon_step(heads, a, v_seg, bitmask, virt_off, n_virt, push_cnt, use_cnt != 0, comp, comp8, axon_inv);
}

// K3 (physics.cu:114-132): one count per occurrence of an id < A. With
// TickParams the tick's count is read on the device (prm->counts[t_off],
// at most the slice length `stride`), so a captured graph serves any
// schedule; the grid strides over the count.
// axon_map (may be null): the engine's axon id of each of the caller's.
__global__ __launch_bounds__(256) void spike_count_kernel(const uint32_t* __restrict__ ids_arg, uint32_t n_arg,
                                                          uint32_t A, uint32_t* __restrict__ push_cnt,
                                                          const TickParams* __restrict__ prm, uint32_t t_off,
                                                          uint32_t stride, const uint32_t* __restrict__ axon_map = nullptr) {
  const uint32_t* __restrict__ ids = prm ? prm->spikes + (size_t)t_off * stride : ids_arg;
  uint32_t n = n_arg;
  if (prm) {
    n = prm->counts ? prm->counts[t_off] : 0u;
    n = n < stride ? n : stride;
  }
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    const uint32_t a = ids[j];
    if (a < A) atomicAdd(&push_cnt[axon_map ? axon_map[a] : a], 1u);
  }
}

// The soma part of K5 (physics.cu:165-174, 209-275) for soma i after its
// dendritic input i_in: refractory countdown, or homeostasis decay, adaptive
// leak, leak toward rest, threshold test, heartbeat, resets, burst count and
// flags. `flags` already has K1 and the spike bit cleared. Sets fin (spike)
// and, for a spike, the soma's own axon to push. thr_in, v_in and axon_in
// are the soma's threshold offset, voltage and axon as read before the tick
// (used only when it integrates).
// hb_i: the soma's index in the caller's numbering, which sets its heartbeat
// phase (equal to i unless the step runs in the engine's own soma order).
__device__ __forceinline__ void soma_tick_with(const ShardVramPtrs& vram, uint32_t i, uint32_t A,
                                               const AxcVariantParameters& p, uint8_t flags, uint8_t timer, bool refr,
                                               uint32_t tick, int32_t i_in, int32_t thr_in, int32_t v_in,
                                               uint32_t axon_in, bool& fin, bool& push_own, uint32_t& my_axon,
                                               uint32_t hb_i) {
  if (refr) {
    vram.timers[i] = (uint8_t)(timer - 1);
    vram.soma_flags[i] = flags;
    return;
  }
  int32_t thr = thr_in;
  int32_t decayed = wsub(thr, (int32_t)p.homeostasis_decay);
  thr = decayed & ~(decayed >> 31);
  int32_t leak = p.leak_rate;
  if (p.adaptive_mode == 1) {
    int32_t add = wmul(thr, (int32_t)p.adaptive_leak_gain) >> 8;
    leak = wadd(leak, add);
    int32_t over = wsub(leak, p.adaptive_leak_max);
    leak = wsub(leak, over & ~(over >> 31));
  }
  int32_t v = wadd(v_in, i_in);
  int32_t diff = wsub(v, p.rest_potential);
  int32_t sign = (diff > 0) - (diff < 0);
  int32_t abs_diff = wmul(diff, sign);
  int32_t leaked = wsub(abs_diff, leak);
  leaked = leaked & ~(leaked >> 31);
  v = wadd(p.rest_potential, wmul(sign, leaked));
  int32_t eff = wadd(p.threshold, thr);
  bool glif = v >= eff;
  uint32_t period = p.spontaneous_firing_period_ticks;
  bool hb = period > 0 && ((tick + hb_i * 104729u) % period) == 0;
  fin = glif || hb;
  if (glif) {
    v = p.rest_potential;
    thr = wadd(thr, p.homeostasis_penalty);
  }
  uint8_t new_timer = glif ? p.refractory_period : (uint8_t)0;
  if (fin) {
    flags = (uint8_t)((flags & 0xFC) | 0x03);
    my_axon = axon_in;
    push_own = (my_axon != kNoAxon) && (my_axon < A);
  } else {
    flags = (uint8_t)(flags & ~0x01);
  }
  uint8_t bc = (flags >> 1) & 0x07;
  bc = (uint8_t)(bc + ((fin && bc < 7) ? 1 : 0));
  vram.soma_voltage[i] = v;
  vram.soma_flags[i] = (uint8_t)((flags & 0xF0) | (bc << 1) | (fin ? 1 : 0));
  vram.threshold_offset[i] = thr;
  vram.timers[i] = new_timer;
}

__device__ __forceinline__ void soma_tick(const ShardVramPtrs& vram, uint32_t i, uint32_t A,
                                          const AxcVariantParameters& p, uint8_t flags, uint8_t timer, bool refr,
                                          uint32_t tick, int32_t i_in, bool& fin, bool& push_own,
                                          uint32_t& my_axon) {
  const int32_t thr = refr ? 0 : vram.threshold_offset[i];
  const int32_t v = refr ? 0 : vram.soma_voltage[i];
  const uint32_t ax = refr ? 0u : vram.soma_to_axon[i];
  soma_tick_with(vram, i, A, p, flags, timer, refr, tick, i_in, thr, v, ax, fin, push_own, my_axon, i);
}

// ---------------------------------------------------------------------------
// K1 + K5 (physics.cu:25-30, 159-276), one lane per soma, dendrite columns
// read eight slots at a time (eight independent target loads, then eight
// independent 16 B snapshot gathers per lane).
//
// FUSED: the GSOP of the previous tick (K6, physics.cu:281-373) is applied in
// the same pass, slot by slot, to the somas that spiked in the previous tick
// (their spike bit is still set when this tick starts): the weight is updated
// from the post-push snapshot of the previous tick (prev) before this tick's
// integration reads it, which is the order of the reference's separate K6
// launch followed by the next tick's K5. Used inside a batch, where no caller
// can observe the state between ticks; the batch ends with a plain gsop launch.
// ---------------------------------------------------------------------------
// One workgroup's 256 somas [i0, i0+256) of one shard: the per-soma tick and
// the workgroup-level compaction and counters. `lut` is the shard's table,
// already in LDS. `part` (may be null) receives the workgroup's counters.
// Called by the single-shard kernel and by the batched work-list kernel.
template <bool FUSED>
__device__ __forceinline__ void neuron_block(const ShardVramPtrs& vram, uint32_t N, uint32_t A,
                                             const AxcVariantParameters* lut, uint32_t tick, int32_t dop,
                                             const ShardScratch& sc, const uint4* __restrict__ cur,
                                             const uint4* __restrict__ prev, uint32_t i0, uint64_t* part,
                                             bool count_tick) {
  __shared__ uint32_t wave_cnt[2][4];
  __shared__ uint32_t blk_base[2];
  __shared__ uint32_t red[4][8];
  const uint32_t tid = threadIdx.x;
  const uint32_t i = i0 + tid;
  bool fin = false;
  bool push_own = false;
  uint32_t my_axon = 0;
  uint32_t n_int = 0, n_scan = 0, n_tread = 0, n_hits = 0, n_gslots = 0, n_gtgt = 0;

  if (i < N) {
    const uint8_t flags_in = vram.soma_flags[i];
    uint8_t flags = flags_in & 0xF1;  // K1
    const uint8_t timer = vram.timers[i];
    flags &= (uint8_t)~0x01;
    const bool refr = timer > 0;
    const bool pend = FUSED && (flags_in & 1);  // GSOP of the previous tick pending
    const AxcVariantParameters& p = lut[(flags >> 4) & 0x0F];
    const uint32_t prop = p.signal_propagation_length;
    int32_t i_in = 0;

    if (!refr || pend) {
      GsopTerms g;
      if (FUSED && pend) g = gsop_terms(flags_in, p, dop);
      // Byte offsets below are 32-bit lane offsets from wave-uniform bases,
      // so loads use a scalar base register plus one VGPR each.
      const uint32_t toff = i * 4u, woff = i * 2u;
      const char* __restrict__ cbase = reinterpret_cast<const char*>(cur);
      const char* __restrict__ pbase = reinterpret_cast<const char*>(prev);
      uint32_t filled = 0;
      bool alive = true;
      // kScanUnroll slots per step. With kScanPipeline the targets and
      // weights of the next step are requested before this step's snapshot
      // gathers, so each lane has two steps of loads in flight.
      constexpr int U = kScanUnroll;
      uint32_t t[U];
      int16_t w[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const char* row = reinterpret_cast<const char*>(vram.dendrite_targets + (size_t)u * N);
        t[u] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(row + toff));
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const char* row = reinterpret_cast<const char*>(vram.dendrite_weights + (size_t)u * N);
        w[u] = *reinterpret_cast<const int16_t*>(row + woff);
      }
      for (uint32_t s = 0; s < kMaxDendrites && alive; s += U) {
        uint32_t nv = U;
#pragma unroll
        for (int u = U - 1; u >= 0; u--) nv = (t[u] == 0) ? (uint32_t)u : nv;
        alive = nv == U;
        filled += nv;
        uint32_t ao[U];  // byte offset of the slot's snapshot entry; slots past the first empty one read the inert entry
#pragma unroll
        for (int u = 0; u < U; u++) ao[u] = (((uint32_t)u < nv) ? snap_index(t[u], A) : A) << 4;
        uint32_t tn[U];
        int16_t wn[U];
        const bool more = alive && (s + U < kMaxDendrites);
        if (kScanPipeline && more) {
#pragma unroll
          for (int u = 0; u < U; u++) {
            const char* row = reinterpret_cast<const char*>(vram.dendrite_targets + (size_t)(s + U + u) * N);
            tn[u] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(row + toff));
          }
#pragma unroll
          for (int u = 0; u < U; u++) {
            const char* row = reinterpret_cast<const char*>(vram.dendrite_weights + (size_t)(s + U + u) * N);
            wn[u] = *reinterpret_cast<const int16_t*>(row + woff);
          }
        }
        if (FUSED && pend) {
          // Pending GSOP of the previous tick first, so that its snapshot
          // values and this tick's are never live together.
          uint4 q[U];
#pragma unroll
          for (int u = 0; u < U; u++) q[u] = *reinterpret_cast<const uint4*>(pbase + ao[u]);
#pragma unroll
          for (int u = 0; u < U; u++) {
            const int16_t nw = gsop_weight(w[u], min_dist16(q[u], t[u] >> 24), g.len, g.raw_pot, g.final_dep, g.bm,
                                           p.inertia_curve);
            if ((uint32_t)u < nv) {
              w[u] = nw;
              char* row = reinterpret_cast<char*>(vram.dendrite_weights + (size_t)(s + u) * N);
              *reinterpret_cast<int16_t*>(row + woff) = nw;
            }
          }
        }
        __builtin_amdgcn_sched_barrier(0);
        if (!refr) {
          uint4 c[U];
#pragma unroll
          for (int u = 0; u < U; u++) c[u] = *reinterpret_cast<const uint4*>(cbase + ao[u]);
#pragma unroll
          for (int u = 0; u < U; u++) {
            // The inert entry never hits, so slots past the end need no mask.
            const bool hit = min_dist16(c[u], t[u] >> 24) < prop;
            n_hits += hit;
            i_in = wadd(i_in, hit ? (int32_t)w[u] : 0);
          }
        }
        if (kScanPipeline) {
#pragma unroll
          for (int u = 0; u < U; u++) { t[u] = tn[u]; w[u] = wn[u]; }
        } else if (alive && s + U < kMaxDendrites) {
#pragma unroll
          for (int u = 0; u < U; u++) {
            const char* row = reinterpret_cast<const char*>(vram.dendrite_targets + (size_t)(s + U + u) * N);
            t[u] = __builtin_nontemporal_load(reinterpret_cast<const uint32_t*>(row + toff));
          }
#pragma unroll
          for (int u = 0; u < U; u++) {
            const char* row = reinterpret_cast<const char*>(vram.dendrite_weights + (size_t)(s + U + u) * N);
            w[u] = *reinterpret_cast<const int16_t*>(row + woff);
          }
        }
      }
      const uint32_t c_i = filled < kMaxDendrites ? filled + 1 : kMaxDendrites;
      if (!refr) {
        n_scan = filled;
        n_tread = c_i;
      }
      if (FUSED && pend) {
        n_gslots = filled;
        n_gtgt = c_i;
      }
    }

    soma_tick(vram, i, A, p, flags, timer, refr, tick, i_in, fin, push_own, my_axon);
    n_int = refr ? 0u : 1u;
  }

  // Push counts: one atomic per spiking soma with an axon, on distinct
  // addresses unless somas share an axon. The first count of an axon lists it.
  bool own_first = false;
  if (push_own) own_first = atomicAdd(&sc.push_cnt[my_axon], 1u) == 0u;

  // Compaction of spikers and first-owners: ballot per wave, one atomic per
  // workgroup and list.
  const uint32_t lane = tid & 63u, wv = tid >> 6;
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t bs = __ballot(fin);
  uint64_t bo = __ballot(own_first);
  if (lane == 0) {
    wave_cnt[0][wv] = (uint32_t)__popcll(bs);
    wave_cnt[1][wv] = (uint32_t)__popcll(bo);
  }
  uint32_t r[6] = {n_int, n_scan, n_tread, n_hits, n_gslots, n_gtgt};
#pragma unroll
  for (int k = 0; k < 6; k++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r[k] += __shfl_xor(r[k], off, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 6; k++) red[wv][k] = r[k];
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t cs = wave_cnt[0][0] + wave_cnt[0][1] + wave_cnt[0][2] + wave_cnt[0][3];
    uint32_t co = wave_cnt[1][0] + wave_cnt[1][1] + wave_cnt[1][2] + wave_cnt[1][3];
    blk_base[0] = cs ? atomicAdd(&sc.counters[kCntSpikes], cs) : 0u;
    blk_base[1] = co ? atomicAdd(&sc.counters[kCntOwners], co) : 0u;
    if (part) {
      part[kStIntegrated] += red[0][0] + red[1][0] + red[2][0] + red[3][0];
      part[kStScanned] += red[0][1] + red[1][1] + red[2][1] + red[3][1];
      part[kStTargets] += red[0][2] + red[1][2] + red[2][2] + red[3][2];
      part[kStHits] += red[0][3] + red[1][3] + red[2][3] + red[3][3];
      part[kStGsopSlots] += red[0][4] + red[1][4] + red[2][4] + red[3][4];
      part[kStGsopTargets] += red[0][5] + red[1][5] + red[2][5] + red[3][5];
      part[kStSpikes] += cs;
      if (count_tick) part[kStTicks] += 1;
    }
  }
  __syncthreads();
  uint32_t pre_s = blk_base[0], pre_o = blk_base[1];
  for (uint32_t w = 0; w < wv; w++) {
    pre_s += wave_cnt[0][w];
    pre_o += wave_cnt[1][w];
  }
  if (fin) sc.spk_list[pre_s + (uint32_t)__popcll(bs & below)] = i;
  if (own_first) sc.own_list[pre_o + (uint32_t)__popcll(bo & below)] = my_axon;
}

template <bool FUSED>
__global__ __launch_bounds__(256) AXC_NEURON_ATTR void neuron_tick_kernel(ShardVramPtrs vram, uint32_t N, uint32_t A,
                                                          const AxcVariantParameters* __restrict__ lut_g,
                                                          const TickParams* __restrict__ prm, uint32_t tick_arg,
                                                          ShardScratch sc, const uint4* __restrict__ cur,
                                                          const uint4* __restrict__ prev, int32_t dop_arg) {
  __shared__ AxcVariantParameters lut[16];
  const uint32_t tid = threadIdx.x;
  if (tid < 64) reinterpret_cast<uint4*>(lut)[tid] = reinterpret_cast<const uint4*>(lut_g)[tid];
  __syncthreads();
  // XCD-aware order: workgroups are dealt round-robin to the 8 XCDs, so
  // workgroup b runs on XCD b % 8. Each XCD gets one contiguous eighth of the
  // somas and walks it in order, which keeps the snapshot entries its somas
  // gather (axons of nearby somas) in that XCD's own L2. The grid is padded
  // to a multiple of 8; workgroups past the last soma block have no somas.
  const uint32_t per_xcd = gridDim.x / kNumXcd;
  const uint32_t lblock = (blockIdx.x % kNumXcd) * per_xcd + blockIdx.x / kNumXcd;
  neuron_block<FUSED>(vram, N, A, lut, tick_of(prm, tick_arg), dopamine_of(prm, dop_arg), sc, cur, prev,
                      lblock * blockDim.x, sc.stats_part + (size_t)blockIdx.x * kStatsPerBlock, blockIdx.x == 0);
}

// Applies the counted pushes of this tick's spikes (physics.cu:253-262 with
// the snapshot rule). One lane per listed axon; clears the count and refreshes
// the axon's compact snapshot.
// comp8 (may be null): the byte snapshot, refreshed too once the soma pass
// that read this tick's byte entries is done (for gsop_kernel).
__device__ __forceinline__ void push_apply_one(AxcBurstHeads8* __restrict__ heads, const ShardScratch& sc,
                                               uint4* __restrict__ comp, uint32_t j, uint2* __restrict__ comp8 = nullptr) {
  uint32_t a = sc.own_list[j];
  uint32_t c = sc.push_cnt[a];
  sc.push_cnt[a] = 0;
  Heads8 h = load_heads(heads, a);
  push_heads(h, c > 8 ? 8 : c);
  store_heads(heads, a, h);
  if (comp) comp[a] = compact_heads(h);
  if (comp8) comp8[a] = byte_heads(h);
}

// The pushes of spiker j's own axon (the staged pass counted them in
// push_cnt without an owner list): the count is taken with an exchange, so
// somas sharing an axon push it once with the full count. Returns the axon
// pushed, or kNoAxon.
__device__ __forceinline__ uint32_t push_spiker_one(AxcBurstHeads8* __restrict__ heads,
                                                const uint32_t* __restrict__ soma_to_axon, uint32_t A,
                                                const ShardScratch& sc, uint4* __restrict__ comp, uint32_t j,
                                                uint2* __restrict__ comp8) {
  const uint32_t a = soma_to_axon[sc.spk_list[j]];
  if (a == kNoAxon || a >= A) return kNoAxon;
  const uint32_t c = atomicExch(&sc.push_cnt[a], 0u);
  if (c == 0) return kNoAxon;
  Heads8 h = load_heads(heads, a);
  push_heads(h, c > 8 ? 8 : c);
  store_heads(heads, a, h);
  if (comp) comp[a] = compact_heads(h);
  if (comp8) comp8[a] = byte_heads(h);
  return a;
}

__global__ __launch_bounds__(256) void push_spikers_kernel(AxcBurstHeads8* __restrict__ heads,
                                                           const uint32_t* __restrict__ soma_to_axon, uint32_t A,
                                                           ShardScratch sc, uint4* __restrict__ comp,
                                                           uint2* __restrict__ comp8) {
  const uint32_t n = sc.counters[kCntSpikes];
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    push_spiker_one(heads, soma_to_axon, A, sc, comp, j, comp8);
}

__global__ __launch_bounds__(256) void push_apply_kernel(AxcBurstHeads8* __restrict__ heads, ShardScratch sc,
                                                         uint4* __restrict__ comp, uint2* __restrict__ comp8 = nullptr) {
  const uint32_t n = sc.counters[kCntOwners];
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    push_apply_one(heads, sc, comp, j, comp8);
}

// ---------------------------------------------------------------------------
// K6 (physics.cu:281-373) over the spiker list: one wave per spiking soma,
// lane = dendrite slot, heads from the post-push snapshot.
// ---------------------------------------------------------------------------
// GSOP of one spiking soma i by one wave; returns its filled-slot count.
__device__ __forceinline__ uint32_t gsop_soma(const ShardVramPtrs& vram, uint32_t N, uint32_t A,
                                              const AxcVariantParameters* lut, int32_t dop, uint32_t i,
                                              const uint4* __restrict__ comp, uint32_t lane) {
  const uint8_t flags = vram.soma_flags[i];
  const AxcVariantParameters& p = lut[(flags >> 4) & 0x0F];
  const GsopTerms g = gsop_terms(flags, p, dop);
  uint32_t filled = 0;
  bool done = false;
  for (uint32_t base = 0; base < kMaxDendrites && !done; base += 64) {
    const uint32_t s = base + lane;
    const size_t col = (size_t)s * N + i;
    const uint32_t t = vram.dendrite_targets[col];
    const uint64_t empty = __ballot(t == 0);
    const uint32_t first_empty = empty ? (uint32_t)__ffsll((long long)empty) - 1u : 64u;
    if (lane < first_empty) {
      const uint4 c = comp[snap_index(t, A)];
      const int16_t w = vram.dendrite_weights[col];
      vram.dendrite_weights[col] =
          gsop_weight(w, min_dist16(c, t >> 24), g.len, g.raw_pot, g.final_dep, g.bm, p.inertia_curve);
    }
    filled += first_empty;
    done = first_empty < 64;
  }
  return filled;
}

// GSOP of spiking soma i with its filled prefix F known (the staged plan's
// fill): lane l takes slots l and l + 64 at once. comp8: the byte snapshot
// after the pushes (the push kernel refreshed it), exact for segment +
// length <= 255; other slots read the heads.
// Smallest h - seg over the eight u32 heads of axon a (wrapping: heads below
// seg count as far, as in the reference's test); the inert index A is far.
__device__ __forceinline__ uint32_t min_dist_heads(const AxcBurstHeads8* __restrict__ heads, uint32_t A,
                                                   uint32_t a, uint32_t seg) {
  if (a >= A) return 0xFFFFFFFFu;
  const Heads8 h = load_heads(heads, a);
  uint32_t m = min(h.lo.x - seg, h.lo.y - seg);
  m = min(m, min(h.lo.z - seg, h.lo.w - seg));
  m = min(m, min(h.hi.x - seg, h.hi.y - seg));
  return min(m, min(h.hi.z - seg, h.hi.w - seg));
}
// GSOP's distance for a slot: from the byte snapshot where it is exact
// (segment + length <= 255), else from the heads themselves.
__device__ __forceinline__ uint32_t gsop_dist(const AxcBurstHeads8* __restrict__ heads, uint32_t A,
                                              const uint2* __restrict__ comp8, uint32_t a, uint32_t seg,
                                              uint32_t len) {
  return seg + len > 255u ? min_dist_heads(heads, A, a, seg) : min_dist8(comp8[a], seg);
}
__device__ __forceinline__ void gsop_soma_filled(const ShardVramPtrs& vram, uint32_t N, uint32_t A,
                                                 const AxcVariantParameters* lut, int32_t dop, uint32_t i, uint32_t F,
                                                 const uint2* __restrict__ comp8, uint32_t lane) {
  const uint8_t flags = vram.soma_flags[i];
  const AxcVariantParameters& p = lut[(flags >> 4) & 0x0F];
  const GsopTerms g = gsop_terms(flags, p, dop);
  const bool l0 = lane < F, l1 = lane + 64 < F;
  const size_t c0 = (size_t)lane * N + i, c1 = (size_t)(lane + 64) * N + i;
  const uint32_t t0 = l0 ? vram.dendrite_targets[c0] : 0u, t1 = l1 ? vram.dendrite_targets[c1] : 0u;
  const int16_t w0 = l0 ? vram.dendrite_weights[c0] : (int16_t)0, w1 = l1 ? vram.dendrite_weights[c1] : (int16_t)0;
  const uint32_t d0 = gsop_dist(vram.axon_heads, A, comp8, l0 ? snap_index(t0, A) : A, t0 >> 24, g.len);
  const uint32_t d1 = gsop_dist(vram.axon_heads, A, comp8, l1 ? snap_index(t1, A) : A, t1 >> 24, g.len);
  if (l0) vram.dendrite_weights[c0] = gsop_weight(w0, d0, g.len, g.raw_pot, g.final_dep, g.bm, p.inertia_curve);
  if (l1) vram.dendrite_weights[c1] = gsop_weight(w1, d1, g.len, g.raw_pot, g.final_dep, g.bm, p.inertia_curve);
}

// fill: the filled prefix of every soma (the staged plan's), or null to find
// the first empty target of each spiker by ballot. With fill, comp8 is the
// post-push byte snapshot.
__global__ __launch_bounds__(256) void gsop_kernel(ShardVramPtrs vram, uint32_t N, uint32_t A,
                                                   const AxcVariantParameters* __restrict__ lut_g,
                                                   const TickParams* __restrict__ prm, int32_t dop_arg,
                                                   ShardScratch sc, const uint4* __restrict__ comp,
                                                   const uint8_t* __restrict__ fill = nullptr,
                                                   const uint2* __restrict__ comp8 = nullptr) {
  __shared__ AxcVariantParameters lut[16];
  __shared__ uint32_t red[4][2];
  const uint32_t tid = threadIdx.x;
  if (tid < 64) reinterpret_cast<uint4*>(lut)[tid] = reinterpret_cast<const uint4*>(lut_g)[tid];
  __syncthreads();

  const uint32_t lane = tid & 63u, wv = tid >> 6;
  const uint32_t n_spk = sc.counters[kCntSpikes];
  const int32_t dop = dopamine_of(prm, dop_arg);
  const uint32_t waves = gridDim.x * 4;
  uint32_t cnt_slots = 0, cnt_tgt = 0;

  for (uint32_t k = blockIdx.x * 4 + wv; k < n_spk; k += waves) {
    const uint32_t i = sc.spk_list[k];
    uint32_t filled;
    if (fill) {
      filled = fill[i];
      gsop_soma_filled(vram, N, A, lut, dop, i, filled, comp8, lane);
    } else {
      filled = gsop_soma(vram, N, A, lut, dop, i, comp, lane);
    }
    cnt_slots += filled;
    cnt_tgt += filled < kMaxDendrites ? filled + 1 : kMaxDendrites;
  }
  if (lane == 0) {
    red[wv][0] = cnt_slots;
    red[wv][1] = cnt_tgt;
  }
  __syncthreads();
  if (tid == 0) {
    uint64_t* part = sc.stats_part2 + (size_t)blockIdx.x * kStatsPerBlock;
    part[kStGsopSlots] += red[0][0] + red[1][0] + red[2][0] + red[3][0];
    part[kStGsopTargets] += red[0][1] + red[1][1] + red[2][1] + red[3][1];
  }
}

// Fills a compact snapshot (all entries plus the inert entry A) from the heads.
__global__ __launch_bounds__(256) void compact_all_kernel(const AxcBurstHeads8* __restrict__ heads, uint32_t A,
                                                          uint4* __restrict__ comp) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < A) comp[a] = compact_heads(load_heads(heads, a));
  else if (a == A) comp[a] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
}

// K7 (physics.cu:378-395); ids outside [0, N) read 0. perm_inv (may be
// null): the engine's soma of each of the caller's (flags in engine order).
__global__ __launch_bounds__(256) void readout_kernel(const uint8_t* __restrict__ flags, uint32_t N,
                                                      const uint32_t* __restrict__ mapped,
                                                      uint8_t* __restrict__ out_arg, uint32_t n_out,
                                                      const TickParams* __restrict__ prm, uint32_t t_off,
                                                      const uint32_t* __restrict__ perm_inv = nullptr) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_out) return;
  uint8_t* __restrict__ out = prm ? prm->out + (size_t)t_off * n_out : out_arg;
  uint32_t id = mapped[j];
  out[j] = (id != kNoAxon && id < N) ? (flags[perm_inv ? perm_inv[id] : id] & 1u) : 0u;
}

// K1 alone (physics.cu:25-30).
__global__ __launch_bounds__(256) void reset_burst_kernel(uint8_t* __restrict__ flags, uint32_t N) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) flags[i] &= 0xF1;
}

// Zeroes the per-tick list counters (a fill kernel, so the batch graph holds
// no memset node).
__global__ void clear_counters_kernel(uint32_t* __restrict__ counters) {
  if (threadIdx.x < kNumCounters) counters[threadIdx.x] = 0;
}

__global__ void set_params_kernel(TickParams* __restrict__ prm, TickParams v) { *prm = v; }

// K10 init_sentinels_kernel (bindings.cu:560-568).
__global__ __launch_bounds__(256) void init_sentinels_kernel(AxcBurstHeads8* __restrict__ heads, uint32_t A) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= A) return;
  Heads8 h;
  h.lo = make_uint4(kSentinel, kSentinel, kSentinel, kSentinel);
  h.hi = h.lo;
  store_heads(heads, a, h);
}

// ---------------------------------------------------------------------------
// K9 sort_and_prune (bindings.cu:450-515). The reference's per-thread
// insertion sort is a stable sort by the key (filled, |w|) in descending
// order, empties keeping their relative order after the filled slots. Here a
// workgroup loads the 128x64 tile of 64 consecutive somas with coalesced row
// reads into LDS, each wave ranks the 128 slots of one soma at a time
// (rank = #greater keys + #equal keys before it), scatters, prunes, and the
// tile is stored back with coalesced rows.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void sort_prune_kernel(uint8_t* __restrict__ flags,
                                                         uint32_t* __restrict__ targets,
                                                         int16_t* __restrict__ weights,
                                                         uint8_t* __restrict__ timers, uint32_t N,
                                                         int32_t prune) {
  __shared__ uint32_t sT[kMaxDendrites][64];
  __shared__ int16_t sW[kMaxDendrites][64];
  __shared__ uint8_t sM[kMaxDendrites][64];
  __shared__ uint32_t keys[4][kMaxDendrites];
  const uint32_t tid = threadIdx.x;
  const uint32_t n0 = blockIdx.x * 64;
  const uint32_t lane = tid & 63u, wv = tid >> 6;

  for (uint32_t r = wv; r < kMaxDendrites; r += 4) {
    uint32_t i = n0 + lane;
    size_t col = (size_t)r * N + i;
    bool ok = i < N;
    sT[r][lane] = ok ? targets[col] : 0u;
    sW[r][lane] = ok ? weights[col] : (int16_t)0;
    sM[r][lane] = ok ? timers[col] : (uint8_t)0;
  }
  if (tid < 64 && n0 + tid < N) flags[n0 + tid] &= 0xF1;
  __syncthreads();

  // 16 rounds; in each, wave wv handles column c = 4*round + wv. Every wave
  // runs every round so the workgroup barriers below are uniform.
  for (uint32_t rnd = 0; rnd < 16; rnd++) {
    const uint32_t c = rnd * 4 + wv;
    uint32_t t0 = sT[lane][c], t1 = sT[lane + 64][c];
    int16_t w0 = sW[lane][c], w1 = sW[lane + 64][c];
    uint8_t m0 = sM[lane][c], m1 = sM[lane + 64][c];
    int32_t a0 = w0 < 0 ? -(int32_t)w0 : w0, a1 = w1 < 0 ? -(int32_t)w1 : w1;
    uint32_t k0 = t0 ? (0x20000u | (uint32_t)a0) : 0u;
    uint32_t k1 = t1 ? (0x20000u | (uint32_t)a1) : 0u;
    keys[wv][lane] = k0;
    keys[wv][lane + 64] = k1;
    __syncthreads();
    uint32_t r0 = 0, r1 = 0;
    for (uint32_t j = 0; j < kMaxDendrites; j++) {
      uint32_t kj = keys[wv][j];
      r0 += (kj > k0) || (kj == k0 && j < lane);
      r1 += (kj > k1) || (kj == k1 && j < lane + 64);
    }
    // Prune (bindings.cu:497-505) and scatter into the soma's column. The
    // column was read by this wave only, before the barrier above.
    if (t0 != 0 && a0 < prune) { t0 = 0; w0 = 0; m0 = 0; }
    if (t1 != 0 && a1 < prune) { t1 = 0; w1 = 0; m1 = 0; }
    sT[r0][c] = t0; sW[r0][c] = w0; sM[r0][c] = m0;
    sT[r1][c] = t1; sW[r1][c] = w1; sM[r1][c] = m1;
    __syncthreads();
  }

  for (uint32_t r = wv; r < kMaxDendrites; r += 4) {
    uint32_t i = n0 + lane;
    if (i >= N) continue;
    size_t col = (size_t)r * N + i;
    targets[col] = sT[r][lane];
    weights[col] = sW[r][lane];
    timers[col] = sM[r][lane];
  }
}

// K11 ghost_sync_kernel (bindings.cu:341-359).
__global__ __launch_bounds__(256) void ghost_sync_kernel(const AxcBurstHeads8* __restrict__ src,
                                                         AxcBurstHeads8* __restrict__ dst,
                                                         const uint32_t* __restrict__ si,
                                                         const uint32_t* __restrict__ di, uint32_t count) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  uint32_t s = si[j];
  if (s < 0x80000000u) store_heads(dst, di[j], load_heads(src, s));
}

// K12 extract_outgoing_spikes_kernel (bindings.cu:378-418).
__global__ __launch_bounds__(256) void extract_spikes_kernel(const AxcBurstHeads8* __restrict__ heads,
                                                             const uint32_t* __restrict__ si,
                                                             const uint32_t* __restrict__ gid, uint32_t count,
                                                             uint32_t T, uint32_t v_seg,
                                                             uint32_t* __restrict__ out_events,
                                                             uint32_t* __restrict__ out_count) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  uint32_t a = si[j];
  if (a >= 0x80000000u) return;
  Heads8 h = load_heads(heads, a);
  uint32_t g = gid[j];
  uint32_t v[8] = {h.lo.x, h.lo.y, h.lo.z, h.lo.w, h.hi.x, h.hi.y, h.hi.z, h.hi.w};
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (v[k] >= 0x70000000u) continue;
    uint32_t ts = v[k] / v_seg;
    if (ts < T) {
      uint32_t o = atomicAdd(out_count, 1u);
      out_events[2 * o] = g;
      out_events[2 * o + 1] = T - 1 - ts;
    }
  }
}

// Telemetry (physics.cu:400-434) with ids in ascending order: per-block
// counts, one-block exclusive scan, then the write.
__global__ __launch_bounds__(256) void telem_count_kernel(const uint8_t* __restrict__ flags, uint32_t N,
                                                          uint32_t* __restrict__ blk_cnt) {
  __shared__ uint32_t wc[4];
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  uint64_t b = __ballot(i < N && (flags[i] & 1u));
  if ((threadIdx.x & 63) == 0) wc[threadIdx.x >> 6] = (uint32_t)__popcll(b);
  __syncthreads();
  if (threadIdx.x == 0) blk_cnt[blockIdx.x] = wc[0] + wc[1] + wc[2] + wc[3];
}

__global__ __launch_bounds__(1024) void telem_scan_kernel(uint32_t* __restrict__ blk_cnt, uint32_t nb,
                                                          uint32_t* __restrict__ out_count) {
  __shared__ uint32_t carry;
  __shared__ uint32_t tmp[1024];
  // Appends after what *out_count already holds, like the reference's
  // atomicAdd on the caller's counter.
  if (threadIdx.x == 0) carry = *out_count;
  __syncthreads();
  for (uint32_t base = 0; base < nb; base += 1024) {
    uint32_t j = base + threadIdx.x;
    uint32_t v = j < nb ? blk_cnt[j] : 0;
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
      uint32_t x = threadIdx.x >= off ? tmp[threadIdx.x - off] : 0;
      __syncthreads();
      tmp[threadIdx.x] += x;
      __syncthreads();
    }
    if (j < nb) blk_cnt[j] = carry + tmp[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += tmp[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out_count = carry;
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ p) { *p = 0; }

__global__ __launch_bounds__(256) void telem_write_kernel(const uint8_t* __restrict__ flags, uint32_t N,
                                                          const uint32_t* __restrict__ blk_off,
                                                          uint32_t* __restrict__ out_ids) {
  __shared__ uint32_t wc[4];
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  bool s = i < N && (flags[i] & 1u);
  uint64_t b = __ballot(s);
  uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) wc[wv] = (uint32_t)__popcll(b);
  __syncthreads();
  uint32_t off = blk_off[blockIdx.x];
  for (uint32_t w = 0; w < wv; w++) off += wc[w];
  uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (s) out_ids[off + (uint32_t)__popcll(b & below)] = i;
}

}  // namespace axc
