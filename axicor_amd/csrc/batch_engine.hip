// batch_engine.hip — many independent brains through one set of work-list
// kernels per tick (no counterpart in the reference, whose host runs one
// ShardEngine per shard, genesis-compute/src/compute/shard.rs:18-125).
//
// Each brain keeps the reference layout (its own cu_allocate_shard arrays)
// and its own 16-entry parameter table. A batch of ticks is captured in one
// HIP graph. With staging (the default) the graph loads every brain's working
// copy of the weights, then per tick launches, for all brains together, the
// axon phase over a flat list of (brain, 256-axon) items, the staged soma
// phase over (brain, kBlockSomas-soma) items in the autotuner's dispatch
// order, the spike-list compaction, a prefix of the brains' spike counts, the
// own-axon pushes and the chunked GSOP over the flattened spike lists; it
// ends by storing the working copies back. Without staging, per tick: axon
// phase, direct soma phase (the previous tick's GSOP fused in), owner-list
// pushes, and one GSOP launch at the end. Each brain runs exactly the device
// code of the single-shard path (axon_step, staged_block, push_spiker_one,
// gsop_chunk; neuron_block, push_apply_one, gsop_soma), so every brain is
// bit-identical to running it alone.
//
// Included by day_engine.hip (one translation unit).

namespace axc {

// Device view of one brain.
struct BrainDev {
  ShardVramPtrs vram;
  uint32_t N, A, v_seg, virt_off, n_virt, words, input_ticks;
  int32_t dop;
  const AxcVariantParameters* lut;  // device, 16 entries
  const uint32_t* bitmask;          // device [input_ticks][words] or null
  ShardScratch sc;
  uint64_t* stats;                  // [kStatsPerBlock] accumulated with atomics
  uint64_t* type_spikes;            // [16] spikes per variant (flags >> 4)
  StagePlan sp;                     // staged soma pass (comp8 null when the batch does not stage)
};

// Work item: brain index and first element (axon or soma) of a 256 chunk.
struct WorkItem {
  uint32_t brain;
  uint32_t start;
};

struct BatchTick {
  uint32_t tick_base;
};

// Axon phase over all brains (K2 + K4 with the counters reset, see
// axon_tick_kernel). Persistent: block g takes items g, g + G, ...
__global__ __launch_bounds__(256) void batch_axon_kernel(const BrainDev* __restrict__ brains,
                                                         const WorkItem* __restrict__ items, uint32_t n_items,
                                                         const BatchTick* __restrict__ bt, uint32_t t_off,
                                                         uint32_t snap) {
  const uint32_t tick = bt->tick_base + t_off;
  for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
    const WorkItem w = items[it];
    const BrainDev& b = brains[w.brain];
    const uint32_t a = w.start + threadIdx.x;
    if (w.start == 0 && threadIdx.x < kNumCounters) b.sc.counters[threadIdx.x] = 0;
    if (a >= b.A) continue;
    const uint32_t* bm = b.bitmask ? b.bitmask + (size_t)(tick % b.input_ticks) * b.words : nullptr;
    // The staged pass reads the byte snapshot (and the heads for exact
    // fallbacks); the compact one only serves the direct pass.
    axon_step(b.vram.axon_heads, a, b.v_seg, bm, b.virt_off, bm ? b.n_virt : 0u, b.sc.push_cnt, false,
              b.sp.comp8 ? nullptr : b.sc.comp[snap & 1], b.sp.comp8);
  }
}

// Soma phase over all brains. Block g on XCD g % 8 walks a contiguous eighth
// of the item list (the XCD-aware order of neuron_tick_kernel).
template <bool FUSED>
__global__ __launch_bounds__(256) AXC_NEURON_ATTR void batch_neuron_kernel(const BrainDev* __restrict__ brains,
                                                                        const WorkItem* __restrict__ items,
                                                                        uint32_t n_items,
                                                                        const BatchTick* __restrict__ bt,
                                                                        uint32_t t_off, uint32_t snap) {
  __shared__ AxcVariantParameters lut[16];
  __shared__ uint64_t cnt[kStatsPerBlock];
  const uint32_t tick = bt->tick_base + t_off;
  const uint32_t G = gridDim.x;
  const uint32_t per_xcd_blocks = G / kNumXcd;
  const uint32_t xcd = blockIdx.x % kNumXcd, k = blockIdx.x / kNumXcd;
  const uint32_t span = (n_items + kNumXcd - 1) / kNumXcd;
  const uint32_t lo = xcd * span, hi = min(n_items, lo + span);
  for (uint32_t it = lo + k; it < hi; it += per_xcd_blocks) {
    const WorkItem w = items[it];
    const BrainDev& b = brains[w.brain];
    __syncthreads();  // the previous item's LDS table and counters are done with
    if (threadIdx.x < 64) reinterpret_cast<uint4*>(lut)[threadIdx.x] = reinterpret_cast<const uint4*>(b.lut)[threadIdx.x];
    if (threadIdx.x < kStatsPerBlock) cnt[threadIdx.x] = 0;
    __syncthreads();
    neuron_block<FUSED>(b.vram, b.N, b.A, lut, tick, b.dop, b.sc, b.sc.comp[snap & 1], b.sc.comp[(snap & 1) ^ 1],
                        w.start, cnt, w.start == 0);
    __syncthreads();
    if (threadIdx.x < kStatsPerBlock && cnt[threadIdx.x]) atomicAdd((unsigned long long*)&b.stats[threadIdx.x],
                                                                   (unsigned long long)cnt[threadIdx.x]);
    // Spikes per variant type, for per-population firing rates.
    const uint32_t i = w.start + threadIdx.x;
    if (i < b.N) {
      const uint8_t f = b.vram.soma_flags[i];
      if (f & 1) atomicAdd((unsigned long long*)&b.type_spikes[f >> 4], 1ull);
    }
  }
}

// Push phase: every thread of the grid strides over each brain's owner list.
__global__ __launch_bounds__(256) void batch_push_kernel(const BrainDev* __restrict__ brains, uint32_t n_brains,
                                                         uint32_t snap) {
  const uint32_t gt = blockIdx.x * blockDim.x + threadIdx.x, gs = gridDim.x * blockDim.x;
  for (uint32_t bi = 0; bi < n_brains; bi++) {
    const BrainDev& b = brains[bi];
    const uint32_t n = b.sc.counters[kCntOwners];
    for (uint32_t j = gt; j < n; j += gs) push_apply_one(b.vram.axon_heads, b.sc, b.sc.comp[snap & 1], j, b.sp.comp8);
  }
}

// GSOP of the last tick of a batch for every brain: every wave of the grid
// strides over each brain's spike list.
__global__ __launch_bounds__(256) void batch_gsop_kernel(const BrainDev* __restrict__ brains, uint32_t n_brains,
                                                         uint32_t snap) {
  __shared__ AxcVariantParameters lut[16];
  __shared__ uint32_t red[4][2];
  const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint32_t gw = blockIdx.x * 4 + wv, W = gridDim.x * 4;
  for (uint32_t bi = 0; bi < n_brains; bi++) {
    const BrainDev& b = brains[bi];
    __syncthreads();
    if (threadIdx.x < 64) reinterpret_cast<uint4*>(lut)[threadIdx.x] = reinterpret_cast<const uint4*>(b.lut)[threadIdx.x];
    __syncthreads();
    const uint32_t n = b.sc.counters[kCntSpikes];
    uint32_t cs = 0, ct = 0;
    for (uint32_t k = gw; k < n; k += W) {
      const uint32_t filled = gsop_soma(b.vram, b.N, b.A, lut, b.dop, b.sc.spk_list[k], b.sc.comp[snap & 1], lane);
      cs += filled;
      ct += filled < kMaxDendrites ? filled + 1 : kMaxDendrites;
    }
    if (lane == 0) {
      red[wv][0] = cs;
      red[wv][1] = ct;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const uint64_t s = red[0][0] + red[1][0] + red[2][0] + red[3][0];
      const uint64_t t = red[0][1] + red[1][1] + red[2][1] + red[3][1];
      if (s) atomicAdd((unsigned long long*)&b.stats[kStGsopSlots], (unsigned long long)s);
      if (t) atomicAdd((unsigned long long*)&b.stats[kStGsopTargets], (unsigned long long)t);
    }
  }
}

__global__ void batch_set_tick_kernel(BatchTick* __restrict__ bt, uint32_t tick_base) { bt->tick_base = tick_base; }

// Staged soma phase over all brains: one workgroup per work item of
// kBlockSomas somas (staged_block, weights from each brain's working copy), the items of
// one XCD a contiguous eighth of the list (grid: the item count rounded up to
// a multiple of 8). GSOP counts are taken here, for batch_gsop_chunks_kernel.
// (A persistent grid walking the items measured slower: the values the
// compiler keeps across items spill.)
// `items` is the dispatch order (the autotuner's choice): XCD x takes
// entries [x * per_xcd, (x + 1) * per_xcd) in order; padding entries carry
// brain kNoBrain.
constexpr uint32_t kNoBrain = 0xFFFFFFFFu;
__global__ __launch_bounds__(kStageThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
void batch_staged_kernel(const BrainDev* __restrict__ brains, const WorkItem* __restrict__ items, uint32_t n_items,
                         const BatchTick* __restrict__ bt, uint32_t t_off, uint32_t snap) {
  __shared__ AxcVariantParameters lut[16];
  const uint32_t per_xcd = gridDim.x / kNumXcd;
  const uint32_t it = (blockIdx.x % kNumXcd) * per_xcd + blockIdx.x / kNumXcd;
  if (it >= n_items) return;  // the whole workgroup
  const WorkItem w = items[it];
  if (w.brain == kNoBrain) return;
  const BrainDev& b = brains[w.brain];
  if (threadIdx.x < 64)
    reinterpret_cast<uint4*>(lut)[threadIdx.x] = reinterpret_cast<const uint4*>(b.lut)[threadIdx.x];
  staged_block<true>(b.vram, b.N, b.A, lut, bt->tick_base + t_off, b.sc, b.sp,
                     w.start / kBlockSomas, b.stats, w.start == 0, b.type_spikes, true);
}

// The staged pass's work per item, for the autotuner's balanced orders: the
// chunks of 8 slots its somas scan plus its staged entries / 32 (the
// gathers of one entry cost about a 32nd of a chunk's loads and tests).
__global__ __launch_bounds__(kBlockSomas) void batch_stage_cost_kernel(const BrainDev* __restrict__ brains,
                                                                       const WorkItem* __restrict__ items,
                                                                       uint32_t* __restrict__ cost) {
  __shared__ uint32_t part[kBlockSomas / 64];
  const WorkItem w = items[blockIdx.x];
  const BrainDev& b = brains[w.brain];
  const uint32_t i = w.start + threadIdx.x;
  uint32_t c = i < b.N ? (b.sp.fill[i] + 7u) / 8u : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63u) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (uint32_t k = 0; k < kBlockSomas / 64; k++) t += part[k];
    const uint32_t n = b.sp.cnt[w.start / kBlockSomas];
    cost[blockIdx.x] = t + (n == kStageOverflow ? kStageCap : n) / 32u;
  }
}

// Exclusive prefix of the brains' list counts of this tick (one workgroup):
// counter kCntSpikes or kCntOwners of each brain's workspace.
__global__ __launch_bounds__(1024) void batch_spike_prefix_kernel(const BrainDev* __restrict__ brains,
                                                                  uint32_t n_brains, uint32_t* __restrict__ prefix,
                                                                  uint32_t counter = kCntSpikes) {
  __shared__ uint32_t part[1024];
  uint32_t carry = 0;
  for (uint32_t b0 = 0; b0 < n_brains; b0 += 1024) {
    const uint32_t bi = b0 + threadIdx.x;
    const uint32_t v = bi < n_brains ? brains[bi].sc.counters[counter] : 0u;
    part[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
      const uint32_t x = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
      __syncthreads();
      part[threadIdx.x] += x;
      __syncthreads();
    }
    if (bi < n_brains) prefix[bi] = carry + part[threadIdx.x] - v;
    carry += part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) prefix[n_brains] = carry;
}

// The brain whose range [prefix[b], prefix[b+1]) holds k (k < prefix[n]).
__device__ __forceinline__ uint32_t brain_of(const uint32_t* __restrict__ prefix, uint32_t n_brains, uint32_t k) {
  uint32_t lo = 0, hi = n_brains;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (prefix[mid] <= k) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Push phase over the flattened spike lists of all brains (prefix of the
// spike counts from batch_spike_prefix_kernel): each spiker's own axon, as
// push_spikers_kernel.
__global__ __launch_bounds__(256) void batch_push_flat_kernel(const BrainDev* __restrict__ brains, uint32_t n_brains,
                                                              const uint32_t* __restrict__ prefix, uint32_t snap) {
  const uint32_t total = prefix[n_brains];
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < total; k += gridDim.x * blockDim.x) {
    const uint32_t bi = brain_of(prefix, n_brains, k);
    const BrainDev& b = brains[bi];
    push_spiker_one(b.vram.axon_heads, b.vram.soma_to_axon, b.A, b.sc, nullptr, k - prefix[bi], b.sp.comp8);
  }
}

// compact_spikes_kernel for every brain of the batch: kItemWords lanes per
// staged work item (one per mask word of its kBlockSomas somas), one
// returning atomic per wave (or per item where a wave's items span brains) on
// the brain's spike counter.
constexpr uint32_t kItemWords = kBlockSomas / 64;
static_assert(kItemWords == 2 || kItemWords == 4, "an item's mask words are a power-of-two lane group");
__global__ __launch_bounds__(256) void batch_compact_spikes_kernel(const BrainDev* __restrict__ brains,
                                                                   const WorkItem* __restrict__ items,
                                                                   uint32_t n_items) {
  const uint32_t q = blockIdx.x * 256 + threadIdx.x, l = q % kItemWords;
  const uint32_t it = q / kItemWords;
  uint64_t m = 0;
  uint32_t word = 0, bi = 0;
  if (it < n_items) {
    const WorkItem w = items[it];
    bi = w.brain;
    word = w.start / 64u + l;
    if (w.start + l * 64u < brains[bi].N) m = brains[bi].sc.spk_mask[word];
  }
  const uint32_t c = (uint32_t)__popcll(m);
  // The items of a brain are consecutive, so a wave's items are mostly of
  // one brain: then one atomic reserves the whole wave's range (per item,
  // the atomics queue on the brain's one counter).
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t b0 = __shfl(bi, 0, 64);
  const bool one_brain = __all(it >= n_items || bi == b0);
  const uint32_t width = one_brain ? 64u : kItemWords, seg = lane & (width - 1);
  uint32_t x = c;  // inclusive prefix over the wave, or over the item's lanes
  for (uint32_t off = 1; off < width; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, (int)width);
    if (seg >= off) x += y;
  }
  const uint32_t total = __shfl(x, (int)(width - 1), (int)width);
  uint32_t base = 0;
  // (lane 63 may lie past the last item: the wave's brain is lane 0's)
  if (seg == width - 1 && total) base = atomicAdd(&brains[one_brain ? b0 : bi].sc.counters[kCntSpikes], total);
  if (total == 0) return;  // alike over each prefix segment
  uint32_t pos = __shfl(base, (int)(width - 1), (int)width) + x - c;
  uint32_t* list = brains[bi].sc.spk_list;
  for (uint64_t r = m; r; r &= r - 1) list[pos++] = word * 64u + (uint32_t)__builtin_ctzll(r);
}

// GSOP of this tick for every brain's spikers over the flattened spike lists
// (prefix from batch_spike_prefix_kernel), kGsopLanes lanes per spiker, one
// chunk of 8 slots each, on the brains' working copies of the weights.
__global__ __launch_bounds__(256) void batch_gsop_chunks_kernel(const BrainDev* __restrict__ brains,
                                                                uint32_t n_brains, const uint32_t* __restrict__ prefix,
                                                                uint32_t snap) {
  const uint32_t total = prefix[n_brains] * kGsopLanes;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
    const uint32_t k = q / kGsopLanes, c = q % kGsopLanes;
    const uint32_t bi = brain_of(prefix, n_brains, k);
    const BrainDev& b = brains[bi];
    const uint32_t i = b.sc.spk_list[k - prefix[bi]];
    const uint32_t F = b.sp.fill[i];
    if (c * 8 < F) gsop_chunk(b.N, b.A, b.lut, b.dop, b.vram.soma_flags, b.sp, b.vram.axon_heads, i, c, F);
  }
}

// Every brain's working copy of the weights, loaded from (IN) or stored to
// its reference rows: one workgroup per soma work item.
template <bool IN>
__global__ __launch_bounds__(256) void batch_copy_weights_kernel(const BrainDev* __restrict__ brains,
                                                                 const WorkItem* __restrict__ items, uint32_t n_items) {
  for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
    const WorkItem w = items[it];
    const BrainDev& b = brains[w.brain];
    const uint32_t i = w.start + threadIdx.x;
    if (i < b.N) copy_weights_soma<IN>(b.vram.dendrite_weights, b.N, b.sp, i, i);
  }
}

// Synapse statistics per brain, one workgroup per soma work item: over the
// slots with a target, the count, sum |w|, sum w^2 and the count with
// |w| < prune[brain]; plus the sum of the threshold offsets. Integer sums, so
// the result does not depend on the order of the atomics.
constexpr int kSynStats = 5;
__global__ __launch_bounds__(256) void batch_synapse_stats_kernel(const BrainDev* __restrict__ brains,
                                                                  const WorkItem* __restrict__ items,
                                                                  const int16_t* __restrict__ prune,
                                                                  unsigned long long* __restrict__ out) {
  __shared__ unsigned long long red[kSynStats][256];
  const WorkItem it = items[blockIdx.x];
  const BrainDev& b = brains[it.brain];
  const uint32_t i = it.start + threadIdx.x;
  unsigned long long v[kSynStats] = {0, 0, 0, 0, 0};
  if (i < b.N) {
    const uint32_t* __restrict__ tg = b.vram.dendrite_targets;
    const int16_t* __restrict__ wt = b.vram.dendrite_weights;
    const int32_t pr = prune[it.brain];
    for (uint32_t s = 0; s < 128; s++) {
      const size_t o = (size_t)s * b.N + i;
      if (tg[o] == 0) continue;
      const int32_t w = wt[o];
      const uint32_t aw = (uint32_t)(w < 0 ? -w : w);
      v[0] += 1;
      v[1] += aw;
      v[2] += (unsigned long long)aw * aw;
      v[3] += (int32_t)aw < pr ? 1 : 0;
    }
    v[4] = (unsigned long long)(long long)b.vram.threshold_offset[i];
  }
  for (int k = 0; k < kSynStats; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int k = 0; k < kSynStats; k++) red[k][threadIdx.x] += red[k][threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x < kSynStats) atomicAdd(&out[(size_t)it.brain * kSynStats + threadIdx.x], red[threadIdx.x][0]);
}

}  // namespace axc
