// day_window.hip — the batch graphs' soma pass over windows of ticks.
//
// The staged pass (day_staged.hip) streams every filled slot once per tick.
// But an axon's heads only advance by v_seg per tick until something pushes
// it (physics.cu:137-154 advances every live head; pushes come from K2
// inputs, the K3 schedule and the axon's own somas), so for the axons that
// nothing pushes, the hits of their slots over the next K ticks follow from
// the heads at the start of those ticks. The batch graphs therefore stream
// the slots once per window of K <= kWin ticks:
//   win_pass_kernel     per soma and tick j of the window, the hits that the
//                       start heads predict for tick j and the sum of their
//                       weights (the staged scan, K tests per slot);
// and at each tick j:
//   win_axon_kernel     K2/K3 + K4 (axon_step), keeping per axon its pushes
//                       since the window started (aux); for each such axon
//                       (or one starting near the wrap), the range of the
//                       slots it feeds (the reverse index, built with the
//                       plan) whose prediction can be wrong, as pieces of 16;
//   win_correct_kernel  for those slots, where the actual hit differs from
//                       the predicted one, +-(1 hit, the slot's current
//                       weight) into the soma's sum of tick j;
//   win_soma_kernel     K1 + the soma update of K5 from the corrected sums;
//   win_push_kernel     the spikers' own-axon pushes (kept in aux too);
//   gsop_win_kernel     K6 on the working copy, plus (new - old weight) for
//                       every predicted hit at the window's later ticks.
// The window's start heads are its byte snapshot (two buffers taken in turn:
// a window's ticks write the other one) and, for the axons it pushes, their
// u32 values saved at the first push.
// Inside a window only the axons it has pushed (or that start near the u32
// wrap) keep exact u32 heads tick by tick; the heads of every other axon stay
// at the window's start, and their current byte form is the start snapshot
// advanced (j+1) v_seg. win_end_kernel advances those K ticks at the
// window's end, so between windows (and between API calls) all heads are the
// reference's.
// So soma i's input at tick j is the sum over its slots of w_j * hit_j
// exactly: predicted hits times the start weights, plus the GSOP changes on
// predicted hits, plus the corrections (current weights) on the slots of
// pushed axons, for which alone a prediction can be wrong. Integer sums do
// not depend on the order of the atomics.
//
// A prediction is one rule everywhere (pass, corrections, GSOP changes):
// for a slot of segment s and propagation length L the test of tick j is the
// reference's test (physics.cu:198-201) at segment s - (j+1) v_seg: on the
// window's byte snapshot (min_dist8) when s + L <= 255, on its u32 heads
// otherwise. For an axon that is neither pushed nor near the wrap both are
// the actual test of tick j.
//
// A soma's hit count and weight sum share one 64-bit word, count << 32 plus
// the signed sum (|sum| <= 128 * 32768 < 2^31), so one atomic adds both.
#include "day_common.h"

namespace axc {

__device__ __forceinline__ int64_t win_word(int32_t w, int32_t n) { return ((int64_t)n << 32) + (int64_t)w; }
__device__ __forceinline__ void win_unpack(int64_t V, int32_t& sum, uint32_t& cnt) {
  const int64_t c = (V + ((int64_t)1 << 31)) >> 32;
  cnt = (uint32_t)c;
  sum = (int32_t)(V - (c << 32));
}

// Heads advanced n <= kWin ticks by K4 (physics.cu:137-154): a live head
// moves v_seg per tick and stays where it meets the sentinel.
__device__ __forceinline__ uint32_t adv_head(uint32_t x, uint32_t v, uint32_t n) {
#pragma unroll
  for (uint32_t k = 0; k < kWin; k++)
    if (k < n && x != kSentinel) x += v;
  return x;
}
__device__ __forceinline__ Heads8 advance_heads(const Heads8& h, uint32_t v, uint32_t n) {
  Heads8 r;
  r.lo = make_uint4(adv_head(h.lo.x, v, n), adv_head(h.lo.y, v, n), adv_head(h.lo.z, v, n), adv_head(h.lo.w, v, n));
  r.hi = make_uint4(adv_head(h.hi.x, v, n), adv_head(h.hi.y, v, n), adv_head(h.hi.z, v, n), adv_head(h.hi.w, v, n));
  return r;
}
// The byte form of heads advanced d segments, from their byte form: a byte
// below 255 is the head itself, far below the sentinel and the wrap, so it
// becomes min(b + d, 255); 255 stays.
__device__ __forceinline__ uint2 advance_bytes(const uint2 e, uint32_t d) {
  if (d >= 255u) return make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
  const axc_us2 D = {(unsigned short)d, (unsigned short)d}, M = {255, 255};
  const uint32_t p[4] = {__builtin_amdgcn_perm(0u, e.x, 0x0c010c00u), __builtin_amdgcn_perm(0u, e.x, 0x0c030c02u),
                         __builtin_amdgcn_perm(0u, e.y, 0x0c010c00u), __builtin_amdgcn_perm(0u, e.y, 0x0c030c02u)};
  uint32_t q[4];
#pragma unroll
  for (int k = 0; k < 4; k++)
    q[k] = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(axc_us2, p[k]) + D, M));
  return make_uint2(__builtin_amdgcn_perm(q[1], q[0], 0x06040200u), __builtin_amdgcn_perm(q[3], q[2], 0x06040200u));
}
// Some live head at or above lim: the axon wraps past 2^32 within the window.
__device__ __forceinline__ bool near_wrap(const Heads8& h, uint32_t lim) {
  const uint32_t hv[8] = {h.lo.x, h.lo.y, h.lo.z, h.lo.w, h.hi.x, h.hi.y, h.hi.z, h.hi.w};
  bool r = false;
#pragma unroll
  for (int x = 0; x < 8; x++) r |= hv[x] != kSentinel && hv[x] >= lim;
  return r;
}

// The predicted hits of a slot (axon a in the snapshot's ids, segment seg,
// length L) at ticks 0..K-1 of the window, as bits. The u32 rule reads the
// start heads from hb, which holds them advanced `back` segments already
// (the heads themselves at the window's start, or later for an axon the
// window has not pushed).
__device__ __forceinline__ uint32_t win_pred_mask(const WinState& w, const AxcBurstHeads8* __restrict__ hb,
                                                  uint32_t back, uint32_t A, uint32_t a, uint32_t seg, uint32_t L,
                                                  uint32_t v_seg, uint32_t K) {
  uint32_t m = 0;
  if (seg + L > 255u) {
    if (a >= A) return 0;
#pragma unroll
    for (uint32_t j = 0; j < kWin; j++)
      if (j < K && min_dist_heads(hb, A, a, seg + back - (j + 1) * v_seg) < L) m |= 1u << j;
  } else {
    const uint2 e = w.comp8_w0[a];
#pragma unroll
    for (uint32_t j = 0; j < kWin; j++)
      if (j < K && min_dist8(e, (seg - (j + 1) * v_seg) & 0xFFFFu) < L) m |= 1u << j;
  }
  return m;
}

__device__ __forceinline__ uint64_t aux_pack(uint32_t stamp, uint32_t m, bool hz, uint32_t dmin, uint32_t dmax) {
  return (uint64_t)stamp | ((uint64_t)m << 32) | ((uint64_t)(hz ? 1u : 0u) << 36) | ((uint64_t)dmin << 37) |
         ((uint64_t)dmax << 50);
}
// An axon whose start heads come near the u32 wrap in the next window (the
// window of stamp ctl[1] + 1, whose window pass takes that stamp): its start
// heads saved and its record current for that window, with no pushes and the
// wrap bit, so that its heads stay exact and every slot of it is corrected.
// Lists axon a in the active list of the window with stamp st (once per
// axon and window: where its record turns current). A lane at a time.
__device__ __forceinline__ void act_append(const DirtyMarks& dm, uint32_t A, uint32_t st, uint32_t a) {
  const uint32_t p = st & 1u;
  dm.act[(size_t)p * A + atomicAdd(&dm.ctl[2 + p], 1u)] = a;
}
// The same for the lanes of a whole wave that want it: one atomic per wave.
__device__ __forceinline__ void act_append_wave(const DirtyMarks& dm, uint32_t A, uint32_t st, bool want,
                                                uint32_t a) {
  const uint64_t bal = __ballot(want);
  if (bal == 0) return;
  const uint32_t lane = threadIdx.x & 63u, lead = (uint32_t)__builtin_ctzll(bal), p = st & 1u;
  uint32_t base = 0;
  if (lane == lead) base = atomicAdd(&dm.ctl[2 + p], (uint32_t)__popcll(bal));
  base = __shfl(base, (int)lead, 64);
  if (want) dm.act[(size_t)p * A + base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull))] = a;
}
__device__ __forceinline__ void mark_near_wrap(const Heads8& h, uint32_t a, uint32_t A, uint32_t lim,
                                               const WinState& w, const DirtyMarks& dm) {
  if (lim != 0 && near_wrap(h, lim)) {
    store_heads(w.heads_w0, a, h);
    dm.aux[a] = aux_pack(dm.ctl[1] + 1, 0, true, kNoDrop, 0);
    act_append(dm, A, dm.ctl[1] + 1, a);
  }
}

// The byte snapshot of the heads at a batch's start, for its first window
// (lim: that window's wrap limit, 2^32 - (K+1) v_seg; 0: none).
__global__ __launch_bounds__(256) void win_snapshot_kernel(const AxcBurstHeads8* __restrict__ heads, uint32_t A,
                                                           uint2* __restrict__ comp8, uint32_t lim, WinState w,
                                                           DirtyMarks dm) {
  const uint32_t a = blockIdx.x * 256 + threadIdx.x;
  if (a < A) {
    const Heads8 h = load_heads(heads, a);
    comp8[a] = byte_heads(h);
    mark_near_wrap(h, a, A, lim, w, dm);
  }
  if (a == A) comp8[A] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
}

// The end of a window of K ticks: the heads of the axons it did not push
// advanced the window's K ticks (their u32 heads stayed at its start), so all
// heads are the reference's again; and the next window's near-wrap axons
// marked (lim_next: its limit; 0: no next window in this batch).
__global__ __launch_bounds__(256) void win_end_kernel(AxcBurstHeads8* __restrict__ heads, uint32_t A, uint32_t v_seg,
                                                      uint32_t K, uint32_t lim_next, WinState w, DirtyMarks dm) {
  const uint32_t a = blockIdx.x * 256 + threadIdx.x;
  if (a >= A) return;
  const bool mine = (uint32_t)dm.aux[a] == dm.ctl[1];
  if (mine && lim_next == 0) return;
  Heads8 h = load_heads(heads, a);
  if (!mine) {
    h = advance_heads(h, v_seg, K);
    store_heads(heads, a, h);
  }
  mark_near_wrap(h, a, A, lim_next, w, dm);
}
// Heads shifted out by a push of k: the last min(k, 8) of h; their values at
// the window's start (h advanced `back` past it) go into [dmin, dmax].
__device__ __forceinline__ void note_dropped(const Heads8& h, uint32_t k, uint32_t back, uint32_t& dmin,
                                             uint32_t& dmax) {
  const uint32_t hv[8] = {h.lo.x, h.lo.y, h.lo.z, h.lo.w, h.hi.x, h.hi.y, h.hi.z, h.hi.w};
  const uint32_t kk = k > 8 ? 8u : k;
#pragma unroll
  for (uint32_t x = 0; x < 8; x++) {
    const uint32_t v0 = hv[x] - back;
    if (x >= 8 - kk && hv[x] != kSentinel && v0 < kNoDrop) {
      dmin = min(dmin, v0);
      dmax = max(dmax, v0);
    }
  }
}
// The segment of a reverse entry, and the first index in [r, b) whose
// entry's segment is >= key (an axon's list is sorted by segment).
__device__ __forceinline__ uint32_t rev_seg(const uint64_t e) { return (uint32_t)(e >> 40) & 0xFFu; }
__device__ __forceinline__ uint32_t rev_lower(const uint64_t* __restrict__ rev, uint32_t r, uint32_t b, uint32_t key) {
  while (r < b) {
    const uint32_t mid = (r + b) >> 1;
    if (rev_seg(rev[mid]) < key) r = mid + 1;
    else b = mid;
  }
  return r;
}

// K2/K3 + K4 of tick j of the window (axon_tick_kernel), in two kernels.
// An axon that nothing has pushed since the window started (and that does
// not start near the wrap) keeps its start heads in `heads`; win_axon_kernel
// writes its current byte form from the start snapshot, and lists it as
// active where this tick pushes it. win_active_kernel then steps the listed
// axons, whose heads are exact (at its first push an axon's heads are its
// start heads, saved for the window, advanced j ticks), and makes the tick's
// correction pieces.
// Between pushes an axon's heads only advance, and a push of k shifts them
// by k and puts k new heads in front, so after m pushes (at most 8) since
// the window started the current heads are m new heads followed by the first
// 8 - m start heads advanced j+1 times (the prediction P). The hit test is an
// OR over the heads, so a slot's prediction and its actual test can differ
// only if one of the m new heads or of the start heads shifted out
// (advanced) lies in [s, s + L): only slots with a segment in (x - Lmax, x]
// for such a value x are tested. New heads are at most (j+1) v_seg: a prefix
// of the axon's segment-sorted reverse list (its ends for segments up to 7
// are stored with it). A value >= 255 + Lmax hits no slot. Axons marked near
// the wrap (mark_near_wrap) have every slot tested. comp8: this window's
// current snapshot.

// This tick's pushes of axon a: its K2 input bit and its K3 count. PERM:
// the caller's id of the axon comes from axon_inv (a separate instantiation,
// so that without it the bitmask load does not wait for the loads before it).
template <bool PERM>
__device__ __forceinline__ uint32_t tick_pushes(uint32_t a, const uint32_t* __restrict__ bitmask, uint32_t virt_off,
                                                uint32_t n_virt, const uint32_t* __restrict__ push_cnt, int use_cnt,
                                                const uint32_t* __restrict__ axon_inv) {
  uint32_t k = 0;
  if (bitmask) {
    const uint32_t i = (PERM ? axon_inv[a] : a) - virt_off;  // wraps for ids below virt_off
    if (i < n_virt) k += (*(const __attribute__((address_space(1))) uint32_t*)(bitmask + (i >> 5)) >> (i & 31)) & 1u;
  }
  if (use_cnt) k += push_cnt[a];
  return k;
}

constexpr uint32_t kAxonsPerLane = 4;  // win_axon_kernel: a workgroup takes 1,024 axons
template <bool PERM>
__global__ __launch_bounds__(256) void win_axon_kernel(uint32_t A, uint32_t v_seg, uint32_t virt_off, uint32_t n_virt,
                                                       const uint32_t* __restrict__ push_cnt, int use_cnt,
                                                       uint32_t* __restrict__ counters,
                                                       const TickParams* __restrict__ prm, uint32_t t_off,
                                                       uint32_t words, uint2* __restrict__ comp8,
                                                       const uint32_t* __restrict__ axon_inv, WinState w,
                                                       DirtyMarks dm, uint32_t j) {
  const uint32_t tid = threadIdx.x;
  const uint32_t* __restrict__ bitmask = (prm->bitmask && n_virt) ? prm->bitmask + (size_t)t_off * words : nullptr;
  const uint32_t a0 = blockIdx.x * (256 * kAxonsPerLane) + tid;
  // All of the lane's loads are issued before any of them is used, so that
  // they make one round trip (a load under a branch followed by its use
  // would wait for every load before it). An input bit outside the bitmask
  // reads word 0 and is dropped.
  uint32_t rec[kAxonsPerLane], bw[kAxonsPerLane], pc[kAxonsPerLane], bi[kAxonsPerLane];
  uint2 e[kAxonsPerLane];
#pragma unroll
  for (uint32_t r = 0; r < kAxonsPerLane; r++) {
    const uint32_t a = min(a0 + 256 * r, A ? A - 1 : 0u);
    rec[r] = reinterpret_cast<const uint32_t*>(dm.aux)[2 * (size_t)a];
    e[r] = w.comp8_w0[a];
    bi[r] = (PERM ? axon_inv[a] : a) - virt_off;  // wraps for ids below virt_off
    // (a global load: as a generic one it would also count as an LDS
    // access, and the kernel's next wait on those would take it along)
    bw[r] = bitmask ? *(const __attribute__((address_space(1))) uint32_t*)(bitmask +
                                                                           (bi[r] < n_virt ? bi[r] >> 5 : 0u))
                    : 0u;
    pc[r] = use_cnt ? push_cnt[a] : 0u;
  }
  const uint32_t stamp = dm.ctl[1];
#pragma unroll
  for (uint32_t r = 0; r < kAxonsPerLane; r++) {
    const uint32_t a = a0 + 256 * r;
    const uint32_t k = ((bi[r] < n_virt) ? (bw[r] >> (bi[r] & 31)) & 1u : 0u) + pc[r];
    const bool idle = a < A && rec[r] != stamp;  // not pushed since the window started
    if (idle && k == 0) comp8[a] = advance_bytes(e[r], (j + 1) * v_seg);
    // Pushed now for the first time: win_active_kernel steps it from now on.
    act_append_wave(dm, A, stamp, idle && k != 0, a);
  }
  if (blockIdx.x == 0) {
    if (tid < kNumCounters) counters[tid] = 0;
    if (tid == 0) dm.piece_cnt[(j + 1) & 1] = 0;  // the next tick's (consumed last tick)
  }
}

template <bool PERM>
__global__ __launch_bounds__(256) void win_active_kernel(AxcBurstHeads8* __restrict__ heads, uint32_t A,
                                                         uint32_t v_seg, uint32_t virt_off, uint32_t n_virt,
                                                         uint32_t* __restrict__ push_cnt, int use_cnt,
                                                         const TickParams* __restrict__ prm, uint32_t t_off,
                                                         uint32_t words, uint2* __restrict__ comp8,
                                                         const uint32_t* __restrict__ axon_inv,
                                                         const AxcVariantParameters* __restrict__ lut_g, WinState w,
                                                         DirtyMarks dm, uint32_t j) {
  __shared__ uint32_t wsum[4], wbase;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
  const uint32_t* __restrict__ bitmask = (prm->bitmask && n_virt) ? prm->bitmask + (size_t)t_off * words : nullptr;
  const uint32_t stamp = dm.ctl[1];
  const uint32_t n_act = dm.ctl[2 + (stamp & 1u)];
  const uint32_t* __restrict__ act = dm.act + (size_t)(stamp & 1u) * A;
  uint32_t Lmax = 0;
#pragma unroll
  for (int k = 0; k < 16; k++) Lmax = max(Lmax, (uint32_t)lut_g[k].signal_propagation_length);
  // Whole workgroups run every round: the round's pieces are reserved in the
  // pool with one atomic per workgroup.
  for (uint32_t q0 = blockIdx.x * 256; q0 < n_act; q0 += gridDim.x * 256) {
    const uint32_t q = q0 + tid;
    uint32_t a = 0, rs = 0, cnt = 0;
    if (q < n_act) {
      a = act[q];
      // Everything indexed by the axon is loaded up front (one round trip).
      const uint64_t ax = dm.aux[a];
      Heads8 pre = load_heads(heads, a);
      const uint32_t r0 = w.rev_off[a], r1 = w.rev_off[a + 1];
      const uint4 cv = w.rev_cut[a];
      const uint32_t k = tick_pushes<PERM>(a, bitmask, virt_off, n_virt, push_cnt, use_cnt, axon_inv);
      const bool mine = (uint32_t)ax == stamp;
      uint32_t m = 0, dmin = kNoDrop, dmax = 0;
      bool hz = false;
      if (mine) {
        m = (uint32_t)(ax >> 32) & 0xFu;
        hz = (ax >> 36) & 1u;
        dmin = (uint32_t)(ax >> 37) & 0x1FFFu;
        dmax = (uint32_t)(ax >> 50) & 0x1FFFu;
      }
      if (use_cnt) push_cnt[a] = 0;
      if (!mine) {  // listed by this tick's win_axon_kernel: pre holds the start heads
        store_heads(w.heads_w0, a, pre);
        pre = advance_heads(pre, v_seg, j);
      }
      Heads8 h = pre;
      push_heads(h, k > 8 ? 8 : k);
      h = advance_heads(h, v_seg, 1);
      store_heads(heads, a, h);
      comp8[a] = byte_heads(h);
      if (k || !mine) {
        if (k) {
          note_dropped(pre, k, j * v_seg, dmin, dmax);
          m = min(8u, m + k);
        }
        dm.aux[a] = aux_pack(stamp, m, hz, dmin, dmax);
      }
      uint32_t lo = 256, hi = 0;  // the segments to test, [lo, hi]
      auto take = [&](uint32_t xl, uint32_t xh) {  // values in [xl, xh]
        if (xl < 255u + Lmax) {
          lo = min(lo, xl >= Lmax ? xl - Lmax + 1 : 0u);
          hi = max(hi, min(xh, 255u));
        }
      };
      if (hz) {
        lo = 0;
        hi = 255;
      } else {
        const uint32_t c[8] = {h.lo.x, h.lo.y, h.lo.z, h.lo.w, h.hi.x, h.hi.y, h.hi.z, h.hi.w};
#pragma unroll
        for (uint32_t x = 0; x < 8; x++)
          if (x < m) take(c[x], c[x]);
        if (dmin <= dmax) {
          const uint32_t adv = (j + 1) * v_seg;
          take(dmin + adv, dmax + adv);
        }
      }
      if (lo <= hi) {
        rs = lo > 0 ? rev_lower(w.rev, r0, r1, lo) : r0;
        uint32_t re = 0xFFFFFFFFu;
        if (hi < kRevCuts) {
          const uint32_t wd = hi < 2 ? cv.x : hi < 4 ? cv.y : hi < 6 ? cv.z : cv.w;
          const uint32_t off = (wd >> (16 * (hi & 1))) & 0xFFFFu;
          if (off != 0xFFFFu) re = r0 + off;
        }
        if (re == 0xFFFFFFFFu) re = rev_lower(w.rev, rs, r1, hi + 1);
        cnt = re > rs ? re - rs : 0u;
      }
    }
    // The axon's pieces of up to 16 entries: an inclusive prefix over the
    // wave, the waves' totals added in LDS, one atomic for the workgroup.
    const uint32_t np = (cnt + 15) >> 4;
    uint32_t x = np;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    if (tid == 0) {
      const uint32_t t0 = wsum[0], t1 = wsum[1], t2 = wsum[2], t3 = wsum[3];
      const uint32_t tot = t0 + t1 + t2 + t3;
      wbase = tot ? atomicAdd(&dm.piece_cnt[j & 1], tot) : 0u;
      wsum[0] = 0;
      wsum[1] = t0;
      wsum[2] = t0 + t1;
      wsum[3] = t0 + t1 + t2;
    }
    __syncthreads();
    const uint32_t base = wbase + wsum[wv] + x - np;
    for (uint32_t p = 0; p < np; p++) {
      const uint32_t c = min(16u, cnt - 16 * p);
      dm.pieces[base + p] = (uint64_t)(rs + 16 * p) | ((uint64_t)a << 32) | ((uint64_t)(c - 1) << 56);
    }
    __syncthreads();  // wsum and wbase are reused by the next round
  }
}

// The own-axon pushes of the tick's spikers (push_spikers_kernel), kept in
// aux for the window (comp8: the window's current snapshot).
__global__ __launch_bounds__(256) void win_push_kernel(AxcBurstHeads8* __restrict__ heads,
                                                       const uint32_t* __restrict__ soma_to_axon, uint32_t A,
                                                       ShardScratch sc, uint2* __restrict__ comp8, WinState w,
                                                       DirtyMarks dm, uint32_t v_seg, uint32_t j) {
  const uint32_t n = sc.counters[kCntSpikes];
  const uint32_t stamp = dm.ctl[1];
  // Whole waves run every round (the appends to the active list take one
  // atomic per wave).
  for (uint32_t q0 = blockIdx.x * blockDim.x; q0 < n; q0 += gridDim.x * blockDim.x) {
    const uint32_t q = q0 + threadIdx.x;
    uint32_t a = kNoAxon, c = 0;
    uint64_t ax = 0;
    Heads8 h{};
    if (q < n) {
      a = soma_to_axon[sc.spk_list[q]];
      if (a != kNoAxon && a < A) {
        // Only the lane that takes the count writes this axon's record and
        // heads, so they can be loaded beside the exchange.
        ax = dm.aux[a];
        h = load_heads(heads, a);
        c = atomicExch(&sc.push_cnt[a], 0u);
      }
    }
    const bool mine = (uint32_t)ax == stamp;
    act_append_wave(dm, A, stamp, c != 0 && !mine, a);
    if (c == 0) continue;
    uint32_t m = 0, dmin = kNoDrop, dmax = 0;
    bool hz = false;
    if (mine) {
      m = (uint32_t)(ax >> 32) & 0xFu;
      hz = (ax >> 36) & 1u;
      dmin = (uint32_t)(ax >> 37) & 0x1FFFu;
      dmax = (uint32_t)(ax >> 50) & 0x1FFFu;
    }
    // Unpushed until now: h holds the start heads; the current ones are
    // those advanced j+1 ticks.
    if (!mine) {
      store_heads(w.heads_w0, a, h);
      h = advance_heads(h, v_seg, j + 1);
    }
    note_dropped(h, c, (j + 1) * v_seg, dmin, dmax);
    push_heads(h, c > 8 ? 8 : c);
    store_heads(heads, a, h);
    comp8[a] = byte_heads(h);
    dm.aux[a] = aux_pack(stamp, min(8u, m + c), hz, dmin, dmax);
  }
}

// The predicted hits and weight sums of one loaded chunk c of soma i at
// ticks 0..K-1 (hit_chunk with K shifted segments). Entries: the block's LDS
// copy of the window's byte snapshot (STAGED), else the snapshot itself.
// Slots whose segment + length exceeds the byte form (and spilled slots,
// segment 255 in the plan) are decided from their targets first (rare): in
// the shifted-segment tests their weight and count lanes are zero, and
// their exact hits are added after them.
// one2: 0x00010001, passed in so that min(gap, 1) stays one packed min.
// hits2[j / 2] counts the hits of ticks j (bits 0..7) and j + 1 (bits 8..15):
// a lane's slots in one soma are at most 32.
// The dot products use the clamped form (no accumulator tied to the
// destination); the sums stay far below 2^31, so clamping changes nothing.
template <bool STAGED, uint32_t K, bool ONCE>
__device__ __forceinline__ void win_chunk(const Chunk<STAGED, true>& now, uint32_t c, uint32_t N, uint32_t A,
                                          const StagePlan& sp, const uint2* ent, const WinState& w,
                                          const AxcBurstHeads8* __restrict__ heads, uint32_t i, uint32_t prop,
                                          uint32_t v_seg, uint32_t one2, const uint32_t (&plen)[K],
                                          int32_t (&i_in)[K], int32_t (&hits2)[(K + 1) / 2]) {
  static_assert(K >= 1 && K <= 4, "four predicted ticks per slot in one word");
  constexpr int U = 8;
  const axc_us2 P = {(unsigned short)prop, (unsigned short)prop};
  const axc_us2 kOne = __builtin_bit_cast(axc_us2, one2);
  uint32_t mx;
  if (STAGED) {
    mx = max_byte8(now.k[4], now.k[5]);
  } else {
    mx = 0;
#pragma unroll
    for (int u = 0; u < U; u++) mx = max(mx, now.k[u] >> 24);
  }
  uint32_t slow_m = 0, slow_h = 0;  // bit u: slot u exact; bits 4u..4u+3: its hits
  if (mx + prop > 255u) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (now.seg(u) + prop > 255u) {
        uint32_t t;
        if constexpr (STAGED)
          t = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(sp.tgt + (size_t)c * N * 8) +
                                                 i * 32u + u * 4u);
        else
          t = now.k[u];
        // The slot's own segment picks the rule (a spilled slot's may be
        // small); the inert index never hits.
        slow_m |= 1u << u;
        slow_h |= win_pred_mask(w, heads, 0, A, snap_index(t, A), t >> 24, prop, v_seg, K) << (4 * u);
      }
    }
  }
  constexpr int H = 2;  // slots in pairs (halves of four spill at eight waves per SIMD)
  // The pairs' keep masks: zero lanes for exact-path slots (their weight
  // and count are added after the shifted-segment tests).
  uint32_t keepw[U / H];
#pragma unroll
  for (int p = 0; p < U / H; p++) keepw[p] = 0xFFFFFFFFu;
  if (slow_m) {
#pragma unroll
    for (int p = 0; p < U / H; p++) {
      const uint32_t sm2 = (slow_m >> (H * p)) & 3u;
      keepw[p] = ((sm2 & 1u) ? 0u : 0x0000FFFFu) | ((sm2 & 2u) ? 0u : 0xFFFF0000u);
    }
  }
#pragma unroll
  for (int h0 = 0; h0 < U; h0 += H) {
    const uint32_t keep = keepw[h0 / H];
    const uint32_t wz = now.w2[h0 >> 1] & keep, cm = 0x00010001u & keep, cm8 = 0x01000100u & keep;
    if constexpr (ONCE) {
      // One pass over the heads for all K ticks. D = head - segment + v_seg
      // is the tick-0 distance; tick j adds j v_seg. As u16, D's from heads
      // behind the segment (negative) lie above 2^16 - 255 and the others
      // below 2^16 - 255 - K v_seg, so the least u16 D is the nearest head
      // at or past the segment (hit at tick j if D + j v_seg < L) and the
      // greatest one the nearest head behind it (hit at tick j if
      // D + j v_seg wraps, that is D > 0xFFFF - j v_seg, as long as
      // j v_seg <= L: the caller takes this path only when (K-1) v_seg <= L).
      // The pair interleaved: R[k] holds head k of both slots (low lane: slot
      // h0), S2 their segments, so min and max need no cross-lane step.
      const uint2 e0 = STAGED ? ent[now.entry(h0, A)] : w.comp8_w0[now.entry(h0, A)];
      const uint2 e1 = STAGED ? ent[now.entry(h0 + 1, A)] : w.comp8_w0[now.entry(h0 + 1, A)];
      uint32_t S2;
      if (STAGED) S2 = __builtin_amdgcn_perm(0u, now.k[4 + (h0 >> 2)], 0x0c000c00u | (h0 & 3) | ((h0 & 3) + 1) << 16);
      else S2 = __builtin_amdgcn_perm(now.k[h0 + 1], now.k[h0], 0x0c070c03u);
      const uint32_t v1 = v_seg & 0xFFFFu;
      const axc_us2 V = {(unsigned short)v1, (unsigned short)v1};
      const axc_us2 sv = __builtin_bit_cast(axc_us2, S2) - V;
      axc_us2 MN, MX;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        const uint32_t sel = 0x0c000c00u | (uint32_t)(k & 3) | ((uint32_t)(4 + (k & 3)) << 16);
        const uint32_t r = k < 4 ? __builtin_amdgcn_perm(e1.x, e0.x, sel) : __builtin_amdgcn_perm(e1.y, e0.y, sel);
        const axc_us2 d = __builtin_bit_cast(axc_us2, r) - sv;
        MN = k ? __builtin_elementwise_min(MN, d) : d;
        MX = k ? __builtin_elementwise_max(MX, d) : d;
      }
#pragma unroll
      for (uint32_t j = 0; j < K; j++) {
        // plen[j]: max(L - j v_seg, 0) in both lanes.
        axc_us2 t = __builtin_elementwise_sub_sat(__builtin_bit_cast(axc_us2, plen[j]), MN);
        if (j > 0) {
          const uint32_t qj = (0xFFFFu - j * v1) & 0xFFFFu;
          const axc_us2 Q = {(unsigned short)qj, (unsigned short)qj};
          t = __builtin_elementwise_max(t, __builtin_elementwise_sub_sat(MX, Q));
        }
        const axc_us2 hv = __builtin_elementwise_min(t, kOne);
        i_in[j] = __builtin_amdgcn_sdot2(__builtin_bit_cast(axc_s2, hv), __builtin_bit_cast(axc_s2, wz), i_in[j], true);
        hits2[j / 2] = __builtin_amdgcn_sdot2(__builtin_bit_cast(axc_s2, hv), __builtin_bit_cast(axc_s2, (j & 1) ? cm8 : cm),
                                              hits2[j / 2], true);
      }
      continue;
    }
    // The per-tick tests (short propagation lengths). The pair's entries
    // spread to u16 lanes once: q[u][r] holds heads 2r, 2r+1 of slot h0+u;
    // S[u] the slot's segment in both halves.
    uint32_t q[H][4], S[H];
#pragma unroll
    for (int u = 0; u < H; u++) {
      const uint2 e = STAGED ? ent[now.entry(h0 + u, A)] : w.comp8_w0[now.entry(h0 + u, A)];
      q[u][0] = __builtin_amdgcn_perm(0u, e.x, 0x0c010c00u);
      q[u][1] = __builtin_amdgcn_perm(0u, e.x, 0x0c030c02u);
      q[u][2] = __builtin_amdgcn_perm(0u, e.y, 0x0c010c00u);
      q[u][3] = __builtin_amdgcn_perm(0u, e.y, 0x0c030c02u);
      const int v = h0 + u;
      if (STAGED) S[u] = __builtin_amdgcn_perm(0u, now.k[4 + (v >> 2)], 0x0c000c00u | ((v & 3) * 0x00010001u));
      else S[u] = __builtin_amdgcn_perm(0u, now.k[v], 0x0c030c03u);
    }
#pragma unroll
    for (uint32_t j = 0; j < K; j++) {
      const uint32_t sh = ((j + 1) * v_seg) & 0xFFFFu;
      const axc_us2 D = {(unsigned short)sh, (unsigned short)sh};
      uint32_t d2[2];
#pragma unroll
      for (int x = 0; x < 2; x++) {
        const axc_us2 s = __builtin_bit_cast(axc_us2, S[x]) - D;
        const axc_us2 m0 = __builtin_elementwise_min(__builtin_bit_cast(axc_us2, q[x][0]) - s,
                                                     __builtin_bit_cast(axc_us2, q[x][1]) - s);
        const axc_us2 m1 = __builtin_elementwise_min(__builtin_bit_cast(axc_us2, q[x][2]) - s,
                                                     __builtin_bit_cast(axc_us2, q[x][3]) - s);
        d2[x] = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(m0, m1));
      }
      const uint32_t M = pair_min(d2[0], d2[1]);
      // 1 in the u16 lanes whose distance is below the length.
      const axc_us2 hv = __builtin_elementwise_min(__builtin_elementwise_sub_sat(P, __builtin_bit_cast(axc_us2, M)),
                                                   kOne);
      i_in[j] = __builtin_amdgcn_sdot2(__builtin_bit_cast(axc_s2, hv), __builtin_bit_cast(axc_s2, wz), i_in[j], true);
      hits2[j / 2] = __builtin_amdgcn_sdot2(__builtin_bit_cast(axc_s2, hv), __builtin_bit_cast(axc_s2, (j & 1) ? cm8 : cm),
                                            hits2[j / 2], true);
    }
  }
  if (slow_m) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      if ((slow_m >> u) & 1u) {
        const int32_t wt = now.weight(u);
#pragma unroll
        for (uint32_t j = 0; j < K; j++) {
          if ((slow_h >> (4 * u + j)) & 1u) {
            i_in[j] += wt;
            hits2[j / 2] += (j & 1) ? 256 : 1;
          }
        }
      }
    }
  }
}

// The window pass of soma block lb: per soma and tick j < K of the window
// the predicted hits and their weight sum, added over the four lane groups
// in LDS and stored to acc[j]. Somas refractory for the whole window
// (timer >= K) are not scanned: no tick of the window reads their sums.
// Workgroup 0 also starts the window's bookkeeping: a new stamp (every aux
// record turns stale), the piece counters cleared (nothing reads them during
// the pass).
// BUF: the staged chunks through buffer loads (Chunk::load_buf; shards of at
// most kBufMaxSomas somas).
template <uint32_t K, bool BUF>
__global__ __launch_bounds__(kStageThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
void win_pass_kernel(ShardVramPtrs vram, uint32_t N, uint32_t A, const AxcVariantParameters* __restrict__ lut_g,
                     StagePlan sp, WinState w, DirtyMarks dm, uint32_t v_seg, uint32_t one2) {
  if (blockIdx.x == 0) {
    if (threadIdx.x == 0) {
      const uint32_t st = dm.ctl[1] + 1;
      dm.ctl[1] = st;
      dm.ctl[2 + ((st + 1) & 1u)] = 0;  // the next window's list (its marks come at this window's end)
      w.cnt[0] += 1;
    }
    if (threadIdx.x < 2) dm.piece_cnt[threadIdx.x] = 0;
  }
  extern __shared__ uint2 ent[];  // kStageCap + 1 entries
  __shared__ AxcVariantParameters lut[16];
  __shared__ unsigned long long vsh[K][kBlockSomas];
  __shared__ uint32_t bcnt[2];  // the block's scanned somas and their filled slots
  const uint32_t per_xcd = gridDim.x / kNumXcd;
  const uint32_t lb = (blockIdx.x % kNumXcd) * per_xcd + blockIdx.x / kNumXcd;
  const uint32_t tid = threadIdx.x;
  if (tid < 64) reinterpret_cast<uint4*>(lut)[tid] = reinterpret_cast<const uint4*>(lut_g)[tid];
  const uint32_t cnt = sp.cnt[lb];
  const bool staged = cnt != kStageOverflow;
  const uint32_t g = tid / kBlockSomas, li = tid % kBlockSomas;
  const uint32_t i = lb * kBlockSomas + li;
  uint32_t F = 0;
  uint8_t flags = 0, timer = 0;
  if (i < N) {
    flags = vram.soma_flags[i];
    timer = vram.timers[i];
    F = sp.fill[i];
  }
  if (staged) stage_entries(ent, sp.set + (size_t)lb * kStageCap, cnt, w.comp8_w0, A, tid);
  if (g == 0) {
#pragma unroll
    for (uint32_t j = 0; j < K; j++) vsh[j][li] = 0ull;
  }
  if (tid == 0) bcnt[0] = bcnt[1] = 0;
  __syncthreads();
  const bool scan = i < N && timer < K;
  if (scan) {
    const uint32_t prop = lut[(flags >> 4) & 0x0F].signal_propagation_length;
    int32_t i_in[K], hits2[(K + 1) / 2];
#pragma unroll
    for (uint32_t j = 0; j < K; j++) i_in[j] = 0;
#pragma unroll
    for (uint32_t j = 0; j < (K + 1) / 2; j++) hits2[j] = 0;
    const uint32_t gu = __builtin_amdgcn_readfirstlane(g);
    const uint32_t cg = (gu + kStageGroups - 1) % kStageGroups;
    uint32_t plen[K];
#pragma unroll
    for (uint32_t j = 0; j < K; j++) {
      const uint32_t pj = prop > j * v_seg ? min(prop - j * v_seg, 0xFFFFu) : 0u;
      plen[j] = pj * 0x00010001u;
    }
    // The one-pass test of all K ticks holds while (K-1) v_seg <= L.
    if (prop >= (K - 1) * v_seg) {
#pragma unroll 1
      for (uint32_t c = cg; c * 8 < F; c += kStageGroups) {
        if (staged) {
          Chunk<true, true> now;
          if constexpr (BUF) now.load_buf(sp, N, c, i);
          else now.load(vram, sp, N, c, i);
          win_chunk<true, K, true>(now, c, N, A, sp, ent, w, vram.axon_heads, i, prop, v_seg, one2, plen, i_in, hits2);
        } else {
          Chunk<false, true> now;
          now.load(vram, sp, N, c, i);
          win_chunk<false, K, true>(now, c, N, A, sp, ent, w, vram.axon_heads, i, prop, v_seg, one2, plen, i_in, hits2);
        }
      }
    } else {
#pragma unroll 1
      for (uint32_t c = cg; c * 8 < F; c += kStageGroups) {
        if (staged) {
          Chunk<true, true> now;
          if constexpr (BUF) now.load_buf(sp, N, c, i);
          else now.load(vram, sp, N, c, i);
          win_chunk<true, K, false>(now, c, N, A, sp, ent, w, vram.axon_heads, i, prop, v_seg, one2, plen, i_in, hits2);
        } else {
          Chunk<false, true> now;
          now.load(vram, sp, N, c, i);
          win_chunk<false, K, false>(now, c, N, A, sp, ent, w, vram.axon_heads, i, prop, v_seg, one2, plen, i_in, hits2);
        }
      }
    }
#pragma unroll
    for (uint32_t j = 0; j < K; j++) {
      const int64_t V = win_word(i_in[j], (hits2[j / 2] >> (8 * (j & 1))) & 0xFF);
      if (V != 0) atomicAdd(&vsh[j][li], (unsigned long long)V);
    }
  }
  if (g == 0) {  // whole waves
    uint32_t ns = scan ? 1u : 0u, nf = scan ? F : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      ns += __shfl_xor(ns, off, 64);
      nf += __shfl_xor(nf, off, 64);
    }
    if ((tid & 63u) == 0 && ns) {
      atomicAdd(&bcnt[0], ns);
      atomicAdd(&bcnt[1], nf);
    }
  }
  __syncthreads();
  if (g == 0 && i < N) {
#pragma unroll
    for (uint32_t j = 0; j < K; j++) w.acc[(size_t)j * N + i] = (int64_t)vsh[j][li];
  }
  if (tid == 0 && bcnt[0]) {
    atomicAdd((unsigned long long*)&w.cnt[1], (unsigned long long)bcnt[0]);
    atomicAdd((unsigned long long*)&w.cnt[2], (unsigned long long)bcnt[1]);
  }
}

// The corrections of tick j (the pieces its axon kernel made): a lane per
// reverse entry, the exact prediction rule against the reference's test;
// where they differ, +-(1 hit, the slot's current weight) into the soma's
// sum of tick j. Counters per workgroup in w.cpart (pieces, entries).
__global__ __launch_bounds__(256) void win_correct_kernel(const AxcBurstHeads8* __restrict__ heads,
                                                          const uint8_t* __restrict__ soma_flags, uint32_t N,
                                                          uint32_t A, const AxcVariantParameters* __restrict__ lut_g,
                                                          StagePlan sp, WinState w, DirtyMarks dm, uint32_t v_seg,
                                                          uint32_t j) {
  __shared__ uint32_t prop[16];
  __shared__ uint32_t bsum[4][2];
  const uint32_t tid = threadIdx.x;
  if (tid < 16) prop[tid] = lut_g[tid].signal_propagation_length;
  const uint32_t np = dm.piece_cnt[j & 1];
  const uint64_t* __restrict__ pieces = dm.pieces;
  __syncthreads();
  uint32_t n_p = 0, n_e = 0;
  for (uint32_t x = blockIdx.x * 256 + tid; (x >> 4) < np; x += gridDim.x * 256) {
    const uint64_t pc = pieces[x >> 4];
    const uint32_t l = x & 15u, c = (uint32_t)(pc >> 56) + 1;
    n_p += l == 0 ? 1u : 0u;
    if (l >= c) continue;
    const uint32_t a = (uint32_t)(pc >> 32) & 0x00FFFFFFu;
    const uint64_t en = w.rev[(uint32_t)pc + l];
    const uint32_t i = (uint32_t)en, s = (uint32_t)(en >> 32) & 0xFFu, seg = rev_seg(en);
    const uint32_t L = prop[(uint32_t)(en >> 48) & 0x0Fu];  // the soma's type, kept in the entry
    const bool pred = (win_pred_mask(w, w.heads_w0, 0, A, a, seg, L, v_seg, j + 1) >> j) & 1u;
    const bool hit = any_hit(load_heads(heads, a), seg, L);
    n_e += 1;
    if (pred != hit) {
      const int32_t wt = sp.w[chunk_at(N, s >> 3, i) + (s & 7)];
      const int64_t d = hit ? win_word(wt, 1) : -win_word(wt, 1);
      atomicAdd((unsigned long long*)&w.acc[(size_t)j * N + i], (unsigned long long)d);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    n_p += __shfl_xor(n_p, off, 64);
    n_e += __shfl_xor(n_e, off, 64);
  }
  if ((tid & 63u) == 0) {
    bsum[tid >> 6][0] = n_p;
    bsum[tid >> 6][1] = n_e;
  }
  __syncthreads();
  if (tid == 0) {
    const uint32_t p = bsum[0][0] + bsum[1][0] + bsum[2][0] + bsum[3][0];
    const uint32_t e = bsum[0][1] + bsum[1][1] + bsum[2][1] + bsum[3][1];
    if (p | e) {
      w.cpart[2 * blockIdx.x] += p;
      w.cpart[2 * blockIdx.x + 1] += e;
    }
  }
}

// K1 + the soma part of K5 at tick j of the window, one lane per soma, from
// acc[j]: the staged pass's finish (finish_block) without the scan; the
// spikes as mask words (compact_spikes_kernel makes the list).
__global__ __launch_bounds__(256) void win_soma_kernel(ShardVramPtrs vram, uint32_t N, uint32_t A,
                                                       const AxcVariantParameters* __restrict__ lut_g,
                                                       const TickParams* __restrict__ prm, uint32_t tick_arg,
                                                       ShardScratch sc, StagePlan sp, WinState w, uint32_t j) {
  __shared__ AxcVariantParameters lut[16];
  __shared__ uint32_t red[4][8];
  const uint32_t tid = threadIdx.x, lane = tid & 63u;
  if (tid < 64) reinterpret_cast<uint4*>(lut)[tid] = reinterpret_cast<const uint4*>(lut_g)[tid];
  __syncthreads();
  const uint32_t tick = tick_of(prm, tick_arg);
  uint32_t r[5] = {0, 0, 0, 0, 0};  // integrated, scanned, targets, hits, spikes
  // Grid-stride over blocks of 256 somas (whole waves each round).
  for (uint32_t i0 = blockIdx.x * 256; i0 < N; i0 += gridDim.x * 256) {
    const uint32_t i = i0 + tid;
    bool fin = false, push_own = false;
    uint32_t my_axon = 0;
    if (i < N) {
      // Every load is issued before the refractory test, so that a soma
      // costs one round trip (the loads under the test would wait for the
      // timer's); a refractory soma's values go unused.
      const uint8_t flags_in = vram.soma_flags[i];
      const uint8_t timer = vram.timers[i];
      const uint64_t acc = w.acc[(size_t)j * N + i];
      const int32_t v_ld = vram.soma_voltage[i], thr_ld = vram.threshold_offset[i];
      const uint32_t ax_ld = vram.soma_to_axon[i];
      const uint32_t F = sp.fill[i];
      const uint32_t hb_i = sp.perm ? sp.perm[i] : i;
      const bool refr = timer > 0;
      int32_t sum = 0, v_in = 0, thr_in = 0;
      uint32_t hc = 0, ax_in = 0;
      if (!refr) {
        win_unpack(acc, sum, hc);
        v_in = v_ld;
        thr_in = thr_ld;
        ax_in = ax_ld;
        r[0] += 1;
        r[1] += F;
        r[2] += F < kMaxDendrites ? F + 1 : kMaxDendrites;
        r[3] += hc;
      }
      uint8_t flags = flags_in & 0xF1;  // K1
      flags &= (uint8_t)~0x01;
      soma_tick_with(vram, i, A, lut[(flags_in >> 4) & 0x0F], flags, timer, refr, tick, sum, thr_in, v_in, ax_in,
                     fin, push_own, my_axon, hb_i);
    }
    if (push_own) atomicAdd(&sc.push_cnt[my_axon], 1u);
    const uint64_t bs = __ballot(fin);
    if (lane == 0) sc.spk_mask[i >> 6] = bs;
    r[4] += (uint32_t)__popcll(bs);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r[k] += __shfl_xor(r[k], off, 64);
  }
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < 5; k++) red[tid >> 6][k] = r[k];
  }
  __syncthreads();
  if (tid == 0) {  // no-return atomics: nothing waits for them
    uint64_t* part = sc.stats_part + (size_t)blockIdx.x * kStatsPerBlock;
    const uint32_t slot[5] = {kStIntegrated, kStScanned, kStTargets, kStHits, kStSpikes};
#pragma unroll
    for (int k = 0; k < 5; k++) {
      const uint32_t v = red[0][k] + red[1][k] + red[2][k] + red[3][k];
      if (v) atomicAdd((unsigned long long*)&part[slot[k]], (unsigned long long)v);
    }
    if (blockIdx.x == 0) atomicAdd((unsigned long long*)&part[kStTicks], 1ull);
  }
}

// K6 on the working copy at tick j of a window of K ticks (gsop_chunks_kernel)
// plus, for each slot whose weight changed, the change times its predicted
// hits at the window's later ticks, into acc (summed over the spiker's 16
// lanes first, one atomic per later tick). A spiker whose timer (set by this
// tick's update) is T is refractory at ticks j+1 .. j+T, which ignore its
// sums: only ticks after j + T take the changes.
__global__ __launch_bounds__(256) void gsop_win_kernel(ShardVramPtrs vram, uint32_t N, uint32_t A,
                                                       const AxcVariantParameters* __restrict__ lut_g,
                                                       const TickParams* __restrict__ prm, int32_t dop_arg,
                                                       ShardScratch sc, StagePlan sp, WinState w, DirtyMarks dm,
                                                       uint32_t v_seg, uint32_t j, uint32_t K) {
  __shared__ AxcVariantParameters lut[16];
  __shared__ uint32_t red[4][2];
  const uint32_t tid = threadIdx.x;
  if (tid < 64) reinterpret_cast<uint4*>(lut)[tid] = reinterpret_cast<const uint4*>(lut_g)[tid];
  __syncthreads();
  const uint32_t n = sc.counters[kCntSpikes] * kGsopLanes;
  const int32_t dop = dopamine_of(prm, dop_arg);
  uint32_t cnt_slots = 0, cnt_tgt = 0;
  // Whole waves take part in the shuffles below: every lane of a block runs
  // the same iterations (the bound is the block's, not the lane's).
  for (uint32_t q0 = blockIdx.x * 256; q0 < n; q0 += gridDim.x * 256) {
    const uint32_t q = q0 + tid;
    uint32_t i = 0, c = 0, F = 0, first = K;
    if (q < n) {
      i = sc.spk_list[q / kGsopLanes];
      c = q % kGsopLanes;
      F = sp.fill[i];
      if (c * 8 < F) first = j + 1 + vram.timers[i];  // the first later tick that reads the sum
    }
    // Spikers refractory for the rest of the window (the usual case) take
    // the plain update; the others also keep their weight changes.
    const bool later = first < K;
    int32_t dv[kWin];
#pragma unroll
    for (uint32_t jj = 0; jj < kWin; jj++) dv[jj] = 0;
    if (q < n && c * 8 < F) {
      v4u* wq = reinterpret_cast<v4u*>(sp.w + chunk_at(N, c, i));
      v4u wb{};
      if (later) wb = *wq;
      gsop_chunk(N, A, lut, dop, vram.soma_flags, sp, vram.axon_heads, i, c, F,
                 LazyHeads{dm.aux, dm.ctl[1], j + 1, v_seg});
      if (later) {
        const v4u wa = *wq;
        const uint32_t before[4] = {wb.x, wb.y, wb.z, wb.w}, after[4] = {wa.x, wa.y, wa.z, wa.w};
        const uint32_t L = lut[(vram.soma_flags[i] >> 4) & 0x0F].signal_propagation_length;
        const v4u* tq = reinterpret_cast<const v4u*>(sp.tgt + chunk_at(N, c, i));
        const v4u t0 = tq[0], t1 = tq[1];
        const uint32_t t[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
        const uint32_t nv = F - c * 8 < 8u ? F - c * 8 : 8u;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int32_t w0 = (int32_t)(int16_t)(uint16_t)(before[u >> 1] >> (16 * (u & 1)));
          const int32_t w1 = (int32_t)(int16_t)(uint16_t)(after[u >> 1] >> (16 * (u & 1)));
          if ((uint32_t)u >= nv || w0 == w1) continue;
          // The u32 rule (segment + length > 255) needs the start heads: saved
          // for an axon the window pushed, else still in the heads.
          const uint32_t a = snap_index(t[u], A);
          const AxcBurstHeads8* hb = vram.axon_heads;
          if ((t[u] >> 24) + L > 255u && a < A && (uint32_t)dm.aux[a] == dm.ctl[1]) hb = w.heads_w0;
          const uint32_t m = win_pred_mask(w, hb, 0, A, a, t[u] >> 24, L, v_seg, K);
#pragma unroll
          for (uint32_t jj = 0; jj < kWin; jj++)
            if (jj >= first && ((m >> jj) & 1u)) dv[jj] += w1 - w0;
        }
      }
    }
    if (q < n && c == 0) {
      cnt_slots += F;
      cnt_tgt += F < kMaxDendrites ? F + 1 : kMaxDendrites;
    }
    // The spiker's 16 lanes are an aligned group of a wave (whole waves run
    // every iteration, so the test is alike over the wave).
    if (__any(later)) {
#pragma unroll
      for (uint32_t jj = 0; jj < kWin; jj++) {
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) dv[jj] += __shfl_xor(dv[jj], off, 16);
        if (q < n && (q % kGsopLanes) == 0 && dv[jj] != 0)
          atomicAdd((unsigned long long*)&w.acc[(size_t)jj * N + i], (unsigned long long)win_word(dv[jj], 0));
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    cnt_slots += __shfl_xor(cnt_slots, off, 64);
    cnt_tgt += __shfl_xor(cnt_tgt, off, 64);
  }
  if ((tid & 63u) == 0) {
    red[tid >> 6][0] = cnt_slots;
    red[tid >> 6][1] = cnt_tgt;
  }
  __syncthreads();
  if (tid == 0) {
    uint64_t* part = sc.stats_part2 + (size_t)blockIdx.x * kStatsPerBlock;
    part[kStGsopSlots] += red[0][0] + red[1][0] + red[2][0] + red[3][0];
    part[kStGsopTargets] += red[0][1] + red[1][1] + red[2][1] + red[3][1];
  }
}

// ---- the reverse index (axon -> the slots it feeds), built with the plan ----
// Filled slots per axon of the plan's ids.
__global__ __launch_bounds__(256) void rev_count_kernel(uint32_t N, uint32_t A, StagePlan sp,
                                                        uint32_t* __restrict__ cnt) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const uint32_t F = sp.fill[i];
  for (uint32_t s = 0; s < F; s++) {
    const uint32_t a = snap_index(sp.tgt[chunk_at(N, s >> 3, i) + (s & 7)], A);
    if (a < A) atomicAdd(&cnt[a], 1u);
  }
}

// Exclusive prefix of cnt[0, n) into off[0, n] by one workgroup, 1024 at a
// time; cnt and off may be the same array.
__global__ __launch_bounds__(1024) void rev_scan_kernel(const uint32_t* cnt, uint32_t n, uint32_t* off) {
  __shared__ uint32_t part[1024];
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < n; b0 += 1024) {
    const uint32_t k = b0 + threadIdx.x;
    const uint32_t v = k < n ? cnt[k] : 0u;
    part[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {
      const uint32_t x = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
      __syncthreads();
      part[threadIdx.x] += x;
      __syncthreads();
    }
    if (k < n) off[k] = carry + part[threadIdx.x] - v;
    carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) off[n] = carry;
}

// The entries, at cursor[a]++ (cursor: a copy of rev_off[0, A)), in no
// particular order within an axon's list (rev_sort_kernel sorts them).
// An entry also holds the soma's variant type (bits 48-51, from the caller's
// flags through the plan's order), which picks the propagation length of
// the slot's test: the type changes only with an upload of the state or of
// the flags, which rebuilds the plan and this index.
__global__ __launch_bounds__(256) void rev_fill_kernel(uint32_t N, uint32_t A, StagePlan sp,
                                                       const uint8_t* __restrict__ flags,
                                                       uint32_t* __restrict__ cursor, uint64_t* __restrict__ rev) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const uint32_t F = sp.fill[i];
  const uint64_t type = (uint64_t)((flags[sp.perm ? sp.perm[i] : i] >> 4) & 0x0Fu) << 48;
  for (uint32_t s = 0; s < F; s++) {
    const uint32_t t = sp.tgt[chunk_at(N, s >> 3, i) + (s & 7)];
    const uint32_t a = snap_index(t, A);
    if (a >= A) continue;
    const uint32_t p = atomicAdd(&cursor[a], 1u);
    rev[p] = (uint64_t)i | ((uint64_t)s << 32) | ((uint64_t)(t >> 24) << 40) | type;
  }
}

// Each axon's list sorted by segment (a wave per axon, a counting sort over
// the 256 segments in LDS; the order within one segment is arbitrary), and
// cut[a]: the ends (relative to the list's start) of its entries with a
// segment <= s for s < kRevCuts, as u16.
__global__ __launch_bounds__(256) void rev_sort_kernel(uint32_t A, const uint32_t* __restrict__ rev_off,
                                                       const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                                                       uint4* __restrict__ cut) {
  __shared__ uint32_t hist[4][256];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  uint32_t* h = hist[wv];
  for (uint32_t a = blockIdx.x * 4 + wv; a < A; a += gridDim.x * 4) {
    const uint32_t r0 = rev_off[a], r1 = rev_off[a + 1];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) h[lane * 4 + k] = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t p = r0 + lane; p < r1; p += 64) atomicAdd(&h[(uint32_t)(in[p] >> 40) & 0xFFu], 1u);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    // Exclusive prefix over the bins: lane l owns bins 4l .. 4l+3.
    uint32_t c[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) c[k] = h[lane * 4 + k];
    const uint32_t tot = c[0] + c[1] + c[2] + c[3];
    uint32_t x = tot;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    uint32_t base = x - tot;
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
      h[lane * 4 + k] = base;
      base += c[k];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {  // the end of the entries with segment <= s, s < kRevCuts (0xFFFF: past u16)
      uint32_t e[kRevCuts];
#pragma unroll
      for (uint32_t k = 0; k < kRevCuts; k++) e[k] = min(h[k + 1], 0xFFFFu);
      cut[a] = make_uint4(e[0] | (e[1] << 16), e[2] | (e[3] << 16), e[4] | (e[5] << 16), e[6] | (e[7] << 16));
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    for (uint32_t p = r0 + lane; p < r1; p += 64) {
      const uint64_t e = in[p];
      out[r0 + atomicAdd(&h[(uint32_t)(e >> 40) & 0xFFu], 1u)] = e;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
  }
}

}  // namespace axc
