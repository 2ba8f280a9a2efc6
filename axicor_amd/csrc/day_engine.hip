// day_engine.hip — host side of libaxicor_day.so: the C-ABI of
// include/axicor_day.h over HIP.
//
// Mirrors genesis-compute/src/cuda/bindings.cu (memory, streams, I/O, Night,
// multi-zone helpers) and physics.cu:441-506 (tick orchestration, parameter
// table), with the engine's own additions: a per-shard workspace (spike
// lists, push counters, work counters), the batch entry point that captures a
// whole BSP batch in a HIP graph, and the host sort&prune.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <mutex>
#include <unordered_map>
#include <vector>
#include <map>
#include <tuple>
#include <thread>
#include <algorithm>
#include "day_common.h"
#include "day_kernels.hip"
#include "day_staged.hip"
#include "day_window.hip"
#include "batch_engine.hip"
#include <atomic>

using namespace axc;

static_assert(sizeof(AxcVariantParameters) == 64, "VariantParameters is 64 B (layout.rs:47)");
static_assert(alignof(AxcVariantParameters) == 64, "VariantParameters align 64 (layout.rs:48)");
static_assert(offsetof(AxcVariantParameters, spontaneous_firing_period_ticks) == 16, "@16");
static_assert(offsetof(AxcVariantParameters, initial_synapse_weight) == 20, "@20");
static_assert(offsetof(AxcVariantParameters, refractory_period) == 28, "@28");
static_assert(offsetof(AxcVariantParameters, inertia_curve) == 32, "@32");
static_assert(offsetof(AxcVariantParameters, adaptive_leak_max) == 48, "@48");
static_assert(offsetof(AxcVariantParameters, adaptive_leak_gain) == 52, "@52");
static_assert(offsetof(AxcVariantParameters, adaptive_mode) == 54, "@54");
static_assert(offsetof(AxcVariantParameters, d1_affinity) == 58, "@58");
static_assert(offsetof(AxcVariantParameters, d2_affinity) == 59, "@59");
static_assert(sizeof(AxcBurstHeads8) == 32 && alignof(AxcBurstHeads8) == 32, "BurstHeads8 (layout.rs:73-74)");
static_assert(sizeof(ShardVramPtrs) == 9 * sizeof(void*), "ShardVramPtrs is 9 pointers (ffi.rs:25-42)");

namespace {

constexpr uint32_t kThreads = 256;

inline uint32_t blocks_for(uint64_t n, uint32_t t = kThreads) { return (uint32_t)((n + t - 1) / t); }
// Workgroups of the grid-stride per-tick kernels: eight per CU, so that
// short-lived workgroups do not wait on the launch rate.
constexpr uint32_t kPersistBlocks = 2048;
// Soma blocks of neuron_tick, padded to a multiple of the XCD count.
inline uint32_t neuron_blocks(uint32_t N) { return (blocks_for(N) + kNumXcd - 1) / kNumXcd * kNumXcd; }
// Blocks of the staged soma pass (kBlockSomas somas each), and that count
// padded to a multiple of the XCD count (its grid).
inline uint32_t stage_blocks(uint32_t N) { return blocks_for(N, kBlockSomas); }
inline uint32_t staged_grid(uint32_t N) { return (stage_blocks(N) + kNumXcd - 1) / kNumXcd * kNumXcd; }

// ---- per-device variant table (the reference's __constant__ VARIANT_LUT[16],
// physics.cu:82; process-global per device like constant memory) -----------
struct DeviceLut {
  AxcVariantParameters* dev = nullptr;
};
std::mutex g_lut_mu;
std::map<int, DeviceLut> g_luts;

// The table of the current device if cu_upload_constant_memory (or another
// entry point) has created it, else null; allocates nothing.
const AxcVariantParameters* device_lut_if_present() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_lut_mu);
  auto it = g_luts.find(dev);
  return it != g_luts.end() ? it->second.dev : nullptr;
}

AxcVariantParameters* device_lut() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_lut_mu);
  DeviceLut& d = g_luts[dev];
  if (!d.dev) {
    if (hipMalloc((void**)&d.dev, 16 * sizeof(AxcVariantParameters)) != hipSuccess) return nullptr;
    (void)hipMemset(d.dev, 0, 16 * sizeof(AxcVariantParameters));
    (void)hipDeviceSynchronize();
  }
  return d.dev;
}

// ---- per-shard workspace, keyed by the state base pointer ------------------
struct GraphKey {
  uint32_t T, words, virt_off, n_virt, max_spk, n_out;
  bool has_bitmask, has_spikes, has_out;
  const void* mapped;
  const void* heads;  // the captured kernels hold the arrays' addresses
  uint32_t v_seg;
  bool staged;
  bool ordered;  // in the engine's soma order (the graph holds the private arrays)
  uint32_t win;      // ticks per window (1: the per-tick staged pass)
  const void* rev;   // the window's reverse index (the graph holds its address)
  // The per-tick spike counts are not part of the key: the graph's K3
  // kernels read them from the shard's device copy (TickParams::counts).
  bool operator==(const GraphKey& o) const {
    return T == o.T && words == o.words && virt_off == o.virt_off && n_virt == o.n_virt && staged == o.staged &&
           max_spk == o.max_spk && n_out == o.n_out && has_bitmask == o.has_bitmask &&
           has_spikes == o.has_spikes && has_out == o.has_out && mapped == o.mapped && heads == o.heads &&
           v_seg == o.v_seg && ordered == o.ordered && win == o.win && rev == o.rev;
  }
};

struct Shard {
  uint32_t N = 0, A = 0;
  int device = 0;
  void* state_base = nullptr;
  ShardScratch sc{};
  uint32_t stats_blocks = 0;
  uint32_t flip = 0;          // snapshot buffer of the next per-tick call
  TickParams* prm = nullptr;  // device
  bool owned_by_library = false;
  // One cached graph per shard (the common case: the same batch shape every
  // BSP step); replaced when the shape changes.
  bool has_graph = false;
  GraphKey gkey{};
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  uint64_t graph_captures = 0;  // graphs recorded for this shard so far
  // Device copy of a graph batch's per-tick spike counts ([counts_cap]).
  uint32_t* d_counts = nullptr;
  uint32_t counts_cap = 0;
  // Event recorded after the last asynchronous host upload into the
  // targets; the plan rebuild waits for it on its own stream.
  hipEvent_t upload_ev = nullptr;
  bool upload_pending = false;
  // Staged dendrite scan (day_staged.hip): the plan, allocated at the first
  // step, and whether the targets changed since it was built.
  StagePlan sp{};
  std::atomic<bool> stage_dirty{true};
  // Bumped whenever the targets may have changed (for plans kept elsewhere:
  // the batches that include this shard's arrays).
  std::atomic<uint64_t> targets_version{1};
  // The engine's own soma order (axc_shard_set_soma_positions): engine soma
  // e is the caller's soma order[e]; empty for the caller's order. The plan
  // and the private soma arrays and heads (shadow) of a step follow it; the
  // caller's arrays are copied in at the start of a step and out at its end.
  std::vector<uint32_t> order;
  uint32_t* d_perm = nullptr;      // [N] order
  uint32_t* d_perm_inv = nullptr;  // [N] its inverse
  uint32_t* d_axon_map = nullptr;  // [A] the engine's id of each of the caller's axons
  uint32_t* d_axon_inv = nullptr;  // [A] its inverse
  uint32_t order_virt_off = 0, order_n_virt = 0;  // the virtual input range the axon map orders
  std::vector<uint32_t> owned;     // the axons owned by somas, ascending (moved by the order)
  ShardVramPtrs shadow{};          // private voltage, flags, threshold, timers, soma_to_axon, heads
  bool ordered() const { return !order.empty(); }
  // Windows of ticks (day_window.hip): allocated at the first windowed
  // batch; the reverse index follows the plan (rebuilt with it).
  WinState win{};
  DirtyMarks dirty{};
  size_t rev_cap = 0;
  bool rev_ready = false;
  uint64_t rev_version = 0;  // targets_version the reverse index was built at
};

// Staged scan on/off for the whole process (on by default).
std::atomic<bool> g_staging{true};
// Ticks per window of the batch graphs' staged pass (day_window.hip); 1 is
// the per-tick pass.
std::atomic<uint32_t> g_window{1};

std::mutex g_mu;
std::unordered_map<const void*, Shard*> g_shards;

// Optional per-kernel timing of the per-tick path: HIP events recorded on the
// caller's stream around each launch, summed by axc_kernel_timing_read.
enum TimedKernel {
  kKAxon = 0, kKNeuron, kKPush, kKGsop, kKReadout, kKSpikes, kKWeights,
  kKWinStart, kKWinPass, kKWinCorrect, kKWinSoma, kKWinEnd, kNumTimed
};
struct TimingState {
  bool enabled = false;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;
  double ms[kNumTimed] = {0};
  uint64_t launches[kNumTimed] = {0};
};
std::mutex g_tmu;
TimingState g_timing;

struct TimedScope {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t st;
  int which;
  bool on;
  TimedScope(int w, hipStream_t s, bool active) : st(s), which(w), on(active) {
    if (!on) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { on = false; return; }
    (void)hipEventRecord(a, st);
  }
  ~TimedScope() {
    if (!on) return;
    (void)hipEventRecord(b, st);
    std::lock_guard<std::mutex> lk(g_tmu);
    g_timing.pending.push_back({which, {a, b}});
  }
};

// The device arrays of a shard's soma order (the host order stays).
void free_order_arrays(Shard* s) {
  for (void* q : {(void*)s->d_perm, (void*)s->d_perm_inv, (void*)s->d_axon_map, (void*)s->d_axon_inv,
                  (void*)s->shadow.soma_voltage,
                  (void*)s->shadow.soma_flags, (void*)s->shadow.threshold_offset, (void*)s->shadow.timers,
                  (void*)s->shadow.soma_to_axon, (void*)s->shadow.axon_heads})
    if (q) (void)hipFree(q);
  s->d_perm = s->d_perm_inv = s->d_axon_map = s->d_axon_inv = nullptr;
  s->shadow = ShardVramPtrs{};
}

void free_window(Shard* s) {
  for (void* q : {(void*)s->win.acc, (void*)s->win.comp8_alt, (void*)s->win.heads_w0, (void*)s->win.rev_off,
                  (void*)s->win.rev, (void*)s->win.rev_cut, (void*)s->win.cnt, (void*)s->win.cpart,
                  (void*)s->dirty.aux, (void*)s->dirty.ctl, (void*)s->dirty.pieces, (void*)s->dirty.piece_cnt,
                  (void*)s->dirty.act})
    if (q) (void)hipFree(q);
  s->win = WinState{};
  s->dirty = DirtyMarks{};
  s->rev_cap = 0;
  s->rev_ready = false;
}

void free_scratch(Shard* s) {
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
  s->has_graph = false;
  (void)hipFree(s->sc.spk_list);
  (void)hipFree(s->sc.spk_mask);
  (void)hipFree(s->sc.own_list);
  (void)hipFree(s->sc.push_cnt);
  (void)hipFree(s->sc.counters);
  (void)hipFree(s->sc.stats_part);
  (void)hipFree(s->sc.stats_part2);
  (void)hipFree(s->sc.comp[0]);
  (void)hipFree(s->sc.comp[1]);
  (void)hipFree(s->prm);
  (void)hipFree(s->d_counts);
  s->sc = ShardScratch{};
  s->prm = nullptr;
  s->d_counts = nullptr;
  s->counts_cap = 0;
  for (void* q : {(void*)s->sp.code, (void*)s->sp.seg, (void*)s->sp.fill, (void*)s->sp.set, (void*)s->sp.cnt,
                  (void*)s->sp.comp8, (void*)s->sp.tgt, (void*)s->sp.w})
    if (q) (void)hipFree(q);
  s->sp = StagePlan{};
  s->stage_dirty = true;
  free_order_arrays(s);
  free_window(s);
}

// Makes `st` wait for the last asynchronous host upload into the shard's
// targets (gpu_memcpy_host_to_device_async on another stream), so that a plan
// is never built from targets still in flight.
void wait_target_upload(Shard* s, hipStream_t st) {
  hipEvent_t ev = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (s->upload_pending) ev = s->upload_ev;
    s->upload_pending = false;
  }
  if (ev) (void)hipStreamWaitEvent(st, ev, 0);
}

// Whether some axon that the shard's order moves is one of the step's
// virtual inputs (whose bitmask index is their axon index): such a step
// cannot run in the engine's order.
bool order_moves_inputs(const Shard* s, uint32_t virt_off, uint32_t n_virt) {
  if (!s->ordered() || n_virt == 0) return false;
  auto it = std::lower_bound(s->owned.begin(), s->owned.end(), virt_off);
  return it != s->owned.end() && (uint64_t)*it < (uint64_t)virt_off + n_virt;
}

// The device side of a shard's soma order, at a plan rebuild: the order and
// its inverse, the axon map from the current soma_to_axon (the k-th axon
// owned by a soma in the engine's order takes the k-th smallest owned id),
// the virtual inputs [virt_off, virt_off + n_virt) ordered among themselves
// by the first soma they reach in the engine's order (other axons keep their
// ids), the private arrays; the plan is then built in that order. Without an
// order the plan is in the caller's.
hipError_t setup_order(Shard* s, const ShardVramPtrs& v, hipStream_t st, uint32_t virt_off, uint32_t n_virt) {
  s->sp.perm = nullptr;
  s->sp.perm_inv = nullptr;
  s->sp.axon_map = nullptr;
  if (!s->ordered()) {
    free_order_arrays(s);
    s->owned.clear();
    return hipSuccess;
  }
  const size_t N = s->N, A = s->A;
  hipError_t e = hipStreamSynchronize(st);
  std::vector<uint32_t> s2a(N);
  if (e == hipSuccess) e = hipMemcpy(s2a.data(), v.soma_to_axon, N * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  std::vector<uint8_t> seen(A, 0);
  std::vector<uint32_t> by_order;
  for (size_t k = 0; k < N; k++) {
    const uint32_t a = s2a[s->order[k]];
    if (a < A && !seen[a]) {
      seen[a] = 1;
      by_order.push_back(a);
    }
  }
  s->owned = by_order;
  std::sort(s->owned.begin(), s->owned.end());
  std::vector<uint32_t> amap(A), inv(N);
  for (size_t a = 0; a < A; a++) amap[a] = (uint32_t)a;
  for (size_t k = 0; k < by_order.size(); k++) amap[by_order[k]] = s->owned[k];
  for (size_t k = 0; k < N; k++) inv[s->order[k]] = (uint32_t)k;
  if (!s->d_perm) {
    ShardVramPtrs& h = s->shadow;
    e = hipMalloc((void**)&s->d_perm, N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_perm_inv, N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_axon_map, (A ? A : 1) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&s->d_axon_inv, (A ? A : 1) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&h.soma_voltage, N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&h.soma_flags, N);
    if (e == hipSuccess) e = hipMalloc((void**)&h.threshold_offset, N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&h.timers, N);
    if (e == hipSuccess) e = hipMalloc((void**)&h.soma_to_axon, N * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&h.axon_heads, (A ? A : 1) * sizeof(AxcBurstHeads8));
    if (e != hipSuccess) {
      free_order_arrays(s);
      return e;
    }
  }
  e = hipMemcpy(s->d_perm, s->order.data(), N * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->d_perm_inv, inv.data(), N * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  // Virtual inputs (when none of them is owned by a soma): ordered by the
  // first soma they reach, those reaching none after the others, ties by id.
  s->order_virt_off = virt_off;
  s->order_n_virt = n_virt;
  const bool inputs_free = n_virt > 0 && (uint64_t)virt_off + n_virt <= A && !order_moves_inputs(s, virt_off, n_virt);
  if (inputs_free) {
    std::vector<uint32_t> key(n_virt, 0xFFFFFFFFu);
    uint32_t* d_key = s->sc.spk_list;  // scratch: N entries, rewritten by the next tick
    if (n_virt > N) {
      d_key = nullptr;
      if ((e = hipMalloc((void**)&d_key, (size_t)n_virt * 4)) != hipSuccess) return e;
    }
    e = hipMemcpy(d_key, key.data(), (size_t)n_virt * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      input_key_kernel<<<blocks_for(N), kThreads>>>(v.dendrite_targets, (uint32_t)N, s->d_perm_inv, virt_off, n_virt,
                                                     d_key);
      e = hipMemcpy(key.data(), d_key, (size_t)n_virt * 4, hipMemcpyDeviceToHost);
    }
    if (d_key != s->sc.spk_list) (void)hipFree(d_key);
    if (e != hipSuccess) return e;
    std::vector<uint32_t> idx(n_virt);
    for (uint32_t k = 0; k < n_virt; k++) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return key[x] < key[y]; });
    for (uint32_t k = 0; k < n_virt; k++) amap[virt_off + idx[k]] = virt_off + k;
  }
  std::vector<uint32_t> ainv(A);
  for (size_t a = 0; a < A; a++) ainv[amap[a]] = (uint32_t)a;
  if (A) e = hipMemcpy(s->d_axon_map, amap.data(), A * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && A) e = hipMemcpy(s->d_axon_inv, ainv.data(), A * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) return e;
  // The dendrite arrays stay the caller's (the plan reads them through the order).
  s->shadow.dendrite_targets = v.dendrite_targets;
  s->shadow.dendrite_weights = v.dendrite_weights;
  s->shadow.dendrite_timers = v.dendrite_timers;
  s->sp.perm = s->d_perm;
  s->sp.perm_inv = s->d_perm_inv;
  s->sp.axon_map = s->d_axon_map;
  return hipSuccess;
}

// The caller's state into the engine's order (soma arrays, heads) or back.
void order_copy(Shard* s, const ShardVramPtrs& ref, bool in, hipStream_t st) {
  const uint32_t N = s->N, A = s->A;
  if (in) {
    order_somas_kernel<true><<<blocks_for(N), kThreads, 0, st>>>(ref, s->shadow, N, A, s->d_perm, s->d_axon_map);
    if (A) order_heads_kernel<true><<<blocks_for(A), kThreads, 0, st>>>(ref.axon_heads, s->shadow.axon_heads, A,
                                                                        s->d_axon_map);
  } else {
    if (A) order_heads_kernel<false><<<blocks_for(A), kThreads, 0, st>>>(ref.axon_heads, s->shadow.axon_heads, A,
                                                                         s->d_axon_map);
    order_somas_kernel<false><<<blocks_for(N), kThreads, 0, st>>>(ref, s->shadow, N, A, s->d_perm, s->d_axon_map);
  }
}

// Allocates the staging plan of a shard (first step) and rebuilds it on
// `st` when the targets changed. Returns whether the staged scan is used.
bool ensure_stage(Shard* s, const ShardVramPtrs& v, hipStream_t st, hipError_t* err, uint32_t virt_off = 0,
                  uint32_t n_virt = 0) {
  *err = hipSuccess;
  if (!g_staging.load() || s->N == 0) return false;
  const size_t N = s->N, A = s->A;
  // An ordered plan orders the step's virtual inputs: another input range
  // rebuilds it.
  if (s->ordered() && n_virt > 0 && (s->order_virt_off != virt_off || s->order_n_virt != n_virt))
    s->stage_dirty = true;
  const size_t blocks = staged_grid((uint32_t)N);
  if (!s->sp.comp8) {
    static std::once_flag attr_once;
    std::call_once(attr_once, [] {
      const int bytes = (int)((kStageCap + 1) * sizeof(uint2));
      for (const void* f : {(const void*)neuron_staged_kernel<false>, (const void*)neuron_staged_kernel<true>,
                            (const void*)batch_staged_kernel, (const void*)win_pass_kernel<1, false>,
                            (const void*)win_pass_kernel<2, false>, (const void*)win_pass_kernel<3, false>,
                            (const void*)win_pass_kernel<4, false>, (const void*)win_pass_kernel<1, true>,
                            (const void*)win_pass_kernel<2, true>, (const void*)win_pass_kernel<3, true>,
                            (const void*)win_pass_kernel<4, true>})
        (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    });
    StagePlan p{};
    hipError_t e = hipMalloc((void**)&p.code, kMaxDendrites * N * sizeof(uint16_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p.seg, kMaxDendrites * N);
    if (e == hipSuccess) e = hipMalloc((void**)&p.fill, N);
    if (e == hipSuccess) e = hipMalloc((void**)&p.set, blocks * kStageCap * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p.cnt, blocks * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p.comp8, (A + 1) * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc((void**)&p.tgt, kMaxDendrites * N * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&p.w, kMaxDendrites * N * sizeof(int16_t));
    // Padding blocks keep an empty set; entry A is the inert axon.
    if (e == hipSuccess) e = hipMemset(p.cnt, 0, blocks * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemset(p.comp8, 0xFF, (A + 1) * sizeof(uint2));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      for (void* q : {(void*)p.code, (void*)p.seg, (void*)p.fill, (void*)p.set, (void*)p.cnt, (void*)p.comp8,
                      (void*)p.tgt, (void*)p.w})
        if (q) (void)hipFree(q);
      *err = e;
      return false;
    }
    s->sp = p;
    s->stage_dirty = true;
  }
  if (s->stage_dirty.exchange(false)) {
    wait_target_upload(s, st);
    if ((*err = setup_order(s, v, st, virt_off, n_virt)) != hipSuccess) {
      s->stage_dirty = true;
      return false;
    }
    stage_build_kernel<<<stage_blocks((uint32_t)N), kBlockSomas, 0, st>>>(v.dendrite_targets, (uint32_t)N,
                                                                          (uint32_t)A, s->sp);
    s->rev_ready = false;
    *err = hipGetLastError();
    if (*err != hipSuccess) {
      s->stage_dirty = true;
      return false;
    }
  }
  return true;
}

// A host write of [dst, dst+size) may have changed some shard's targets.
// For an asynchronous write (`async`, already enqueued on `st`) an event is
// recorded after it, which the next plan rebuild waits for.
// Slots given by column (slot * padded_n + soma) written in the reference
// rows (axc_shard_write_slots); columns past the rows are skipped.
__global__ __launch_bounds__(256) void write_slots_kernel(uint32_t* __restrict__ targets, int16_t* __restrict__ weights,
                                                          uint64_t n_cols, const uint32_t* __restrict__ cols,
                                                          const uint32_t* __restrict__ t, const int16_t* __restrict__ w,
                                                          uint32_t n) {
  for (uint32_t k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) {
    const uint32_t c = cols[k];
    if (c < n_cols) {
      targets[c] = t[k];
      weights[c] = w[k];
    }
  }
}

void note_device_write(const void* dst, size_t size, bool async = false, hipStream_t st = nullptr) {
  if (!dst || size == 0) return;
  const char* a = (const char*)dst;
  std::lock_guard<std::mutex> lk(g_mu);
  for (auto& kv : g_shards) {
    Shard* s = kv.second;
    const char* t0 = (const char*)s->state_base + (size_t)s->N * 14;
    const char* t1 = t0 + (size_t)s->N * 4 * kMaxDendrites;
    // The flags column too: the reverse index keeps each soma's type.
    const char* f0 = (const char*)s->state_base + (size_t)s->N * 4;
    const char* f1 = f0 + (size_t)s->N;
    if ((a < t1 && a + size > t0) || (a < f1 && a + size > f0)) {
      if (async) {
        if (!s->upload_ev) (void)hipEventCreateWithFlags(&s->upload_ev, hipEventDisableTiming);
        if (s->upload_ev && hipEventRecord(s->upload_ev, st) == hipSuccess) s->upload_pending = true;
      }
      s->stage_dirty = true;
      s->targets_version++;
    }
  }
}

// Frees a shard's workspace and its bookkeeping.
void destroy_shard(Shard* s) {
  free_scratch(s);
  if (s->upload_ev) (void)hipEventDestroy(s->upload_ev);
  delete s;
}

void mark_stage_dirty(const void* state_base) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_shards.find(state_base);
  if (it != g_shards.end()) {
    it->second->stage_dirty = true;
    it->second->targets_version++;
  }
}

// Targets version of the registered shard at state_base (0: not registered).
uint64_t targets_version_of(const void* state_base) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_shards.find(state_base);
  return it != g_shards.end() ? it->second->targets_version.load() : 0;
}

hipError_t alloc_scratch(Shard* s) {
  const size_t N = s->N ? s->N : 1, A = s->A ? s->A : 1;
  s->stats_blocks = std::max(neuron_blocks((uint32_t)N), staged_grid((uint32_t)N));
  hipError_t e;
  if ((e = hipMalloc((void**)&s->sc.spk_list, N * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&s->sc.spk_mask, (size_t)staged_grid((uint32_t)N) * (kBlockSomas / 64) * 8)) != hipSuccess)
    return e;
  if ((e = hipMalloc((void**)&s->sc.own_list, A * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&s->sc.push_cnt, A * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&s->sc.counters, kNumCounters * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&s->sc.stats_part, (size_t)s->stats_blocks * kStatsPerBlock * 8)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&s->sc.stats_part2, (size_t)kListBlocks * kStatsPerBlock * 8)) != hipSuccess) return e;
  if ((e = hipMalloc((void**)&s->prm, sizeof(TickParams))) != hipSuccess) return e;
  for (int k = 0; k < 2; k++) {
    if ((e = hipMalloc((void**)&s->sc.comp[k], (A + 1) * sizeof(uint4))) != hipSuccess) return e;
    // Entry A is the inert axon; entries [0, A) are rewritten every tick.
    if ((e = hipMemset(s->sc.comp[k], 0xFF, (A + 1) * sizeof(uint4))) != hipSuccess) return e;
  }
  if ((e = hipMemset(s->sc.push_cnt, 0, A * 4)) != hipSuccess) return e;
  if ((e = hipMemset(s->sc.counters, 0, kNumCounters * 4)) != hipSuccess) return e;
  if ((e = hipMemset(s->sc.stats_part, 0, (size_t)s->stats_blocks * kStatsPerBlock * 8)) != hipSuccess) return e;
  if ((e = hipMemset(s->sc.stats_part2, 0, (size_t)kListBlocks * kStatsPerBlock * 8)) != hipSuccess) return e;
  if ((e = hipMemset(s->prm, 0, sizeof(TickParams))) != hipSuccess) return e;
  return hipDeviceSynchronize();
}

// Finds the workspace of a shard; for pointer sets the library did not
// allocate (a caller's own hipMalloc), creates one sized for (N, A).
Shard* shard_for(const ShardVramPtrs* v, uint32_t N, uint32_t A, hipError_t* err) {
  *err = hipSuccess;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_shards.find(v->soma_voltage);
  if (it != g_shards.end()) {
    Shard* s = it->second;
    if (s->N == N && s->A == A) return s;
    if (s->owned_by_library) {  // the arrays were sized for (s->N, s->A)
      *err = hipErrorInvalidValue;
      return nullptr;
    }
    free_scratch(s);
    s->N = N;
    s->A = A;
    if ((*err = alloc_scratch(s)) != hipSuccess) {  // unregister rather than keep a half-built workspace
      g_shards.erase(it);
      destroy_shard(s);
      return nullptr;
    }
    return s;
  }
  Shard* s = new Shard();
  s->N = N;
  s->A = A;
  (void)hipGetDevice(&s->device);
  s->state_base = v->soma_voltage;
  if ((*err = alloc_scratch(s)) != hipSuccess) {
    destroy_shard(s);
    return nullptr;
  }
  g_shards[v->soma_voltage] = s;
  return s;
}

// Enqueues one tick. prm == nullptr: tick, dopamine and the per-tick
// pointers are taken from the arguments; otherwise kernels read them from prm
// (tick = tick_base + t_off, slices at base + t_off*stride) and the pointer
// arguments only say whether inputs/spikes/outputs are present.
// snap: compact snapshot buffer of this tick (the other one holds the
// previous tick's). fuse_prev: apply the previous tick's GSOP inside this
// tick's neuron pass. gsop_now: launch this tick's GSOP (otherwise the next
// tick of the batch applies it). staged: the staged soma pass; pw: with the
// plan's working copy of the weights (loaded by the caller).
hipError_t enqueue_tick(const ShardVramPtrs& v, Shard* s, uint32_t v_seg, uint32_t tick,
                        const uint32_t* bitmask, uint32_t words, uint32_t virt_off, uint32_t n_virt,
                        const uint32_t* spikes, uint32_t max_spk, uint32_t n_spk, const uint32_t* mapped,
                        uint8_t* out, uint32_t n_out, int32_t dopamine, const TickParams* prm, hipStream_t st,
                        const AxcVariantParameters* lut, uint32_t snap, bool fuse_prev, bool gsop_now,
                        bool staged, bool pw = false, bool allow_timing = true) {
  const uint32_t N = s->N, A = s->A;
  uint4* cur = s->sc.comp[snap & 1];
  uint4* prev = s->sc.comp[(snap & 1) ^ 1];
  bool timed;
  {
    std::lock_guard<std::mutex> lk(g_tmu);
    timed = allow_timing && g_timing.enabled && prm == nullptr;  // never inside a captured graph
  }
  // K3: per call, launched for the given count; in a graph (prm), for
  // every tick whenever the batch has a schedule, with the count read on the
  // device and a grid that strides over it.
  const bool k3 = prm ? (spikes != nullptr && max_spk > 0) : (n_spk > 0 && spikes != nullptr);
  if (k3) {
    TimedScope ts(kKSpikes, st, timed);
    const uint32_t grid = prm ? std::min<uint32_t>(blocks_for(max_spk), 64u) : blocks_for(n_spk);
    spike_count_kernel<<<grid, kThreads, 0, st>>>(spikes, n_spk, A, s->sc.push_cnt, prm, tick, max_spk,
                                                  staged ? s->sp.axon_map : nullptr);
  }
  const bool k2 = n_virt > 0 && bitmask != nullptr;
  {
    TimedScope ts(kKAxon, st, timed);
    axon_tick_kernel<<<blocks_for(A > kNumCounters ? A : kNumCounters), kThreads, 0, st>>>(
        v.axon_heads, A, v_seg, k2 ? bitmask : nullptr, virt_off, k2 ? n_virt : 0u, s->sc.push_cnt, k3 ? 1 : 0,
        s->sc.counters, prm, tick, words, staged ? nullptr : cur, staged ? s->sp.comp8 : nullptr,
        (staged && s->sp.perm) ? s->d_axon_inv : nullptr);
  }
  if (N > 0 && staged) {
    // The staged pass never fuses the previous tick's GSOP (gsop_now is set
    // for every tick of a staged batch).
    TimedScope ts(kKNeuron, st, timed);
    const size_t lds = (kStageCap + 1) * sizeof(uint2);
    if (pw)
      neuron_staged_kernel<true><<<staged_grid(N), kStageThreads, lds, st>>>(v, N, A, lut, prm, tick, s->sc, s->sp);
    else
      neuron_staged_kernel<false><<<staged_grid(N), kStageThreads, lds, st>>>(v, N, A, lut, prm, tick, s->sc,
                                                                                s->sp);
  } else if (N > 0) {
    TimedScope ts(kKNeuron, st, timed);
    if (fuse_prev)
      neuron_tick_kernel<true><<<neuron_blocks(N), kThreads, 0, st>>>(v, N, A, lut, prm, tick, s->sc, cur, prev,
                                                                   dopamine);
    else
      neuron_tick_kernel<false><<<neuron_blocks(N), kThreads, 0, st>>>(v, N, A, lut, prm, tick, s->sc, cur, prev,
                                                                    dopamine);
  }
  {
    TimedScope ts(kKPush, st, timed);
    if (staged) {
      // The staged pass leaves its spikes as mask words: the spike list first.
      const uint32_t n_words = (N + 63) / 64;
      if (N > 0) compact_spikes_kernel<<<blocks_for(n_words), kThreads, 0, st>>>(n_words, s->sc);
      push_spikers_kernel<<<kListBlocks / 4, kThreads, 0, st>>>(v.axon_heads, v.soma_to_axon, A, s->sc, nullptr,
                                                                s->sp.comp8);
    } else {
      push_apply_kernel<<<kListBlocks / 4, kThreads, 0, st>>>(v.axon_heads, s->sc, cur, nullptr);
    }
  }
  if (gsop_now) {
    TimedScope ts(kKGsop, st, timed);
    if (pw)
      gsop_chunks_kernel<<<kListBlocks, kThreads, 0, st>>>(v, N, A, lut, prm, dopamine, s->sc, s->sp);
    else
      gsop_kernel<<<kListBlocks, kThreads, 0, st>>>(v, N, A, lut, prm, dopamine, s->sc, cur,
                                                    staged ? s->sp.fill : nullptr, staged ? s->sp.comp8 : nullptr);
  }
  if (n_out > 0 && mapped != nullptr && out != nullptr) {
    TimedScope ts(kKReadout, st, timed);
    readout_kernel<<<blocks_for(n_out), kThreads, 0, st>>>(v.soma_flags, N, mapped, out, n_out, prm, tick,
                                                           (staged && s->sp.perm) ? s->d_perm_inv : nullptr);
  }
  return hipGetLastError();
}

// ---- windows of ticks (day_window.hip) -------------------------------------
// The window length a batch uses: g_window, or 1 where the byte form of the
// shifted segments could wrap (K v_seg must stay far below 2^16).
uint32_t window_for(uint32_t v_seg) {
  const uint32_t k = g_window.load();
  if (k < 2) return 1;
  return (uint64_t)k * v_seg < 0x8000u ? k : 1;
}

// Allocates the window buffers (first windowed batch) and builds the
// reverse index of the current plan when it is missing or stale. Outside
// graph capture: it synchronizes `st` once to size the index (a larger
// index moves, and the batch graph holding its address is recorded again).
hipError_t ensure_window(Shard* s, const uint8_t* flags, hipStream_t st) {
  const size_t N = s->N, A = s->A;
  hipError_t e = hipSuccess;
  if (!s->win.acc) {
    WinState w{};
    DirtyMarks d{};
    e = hipMalloc((void**)&w.acc, (size_t)kWin * (N ? N : 1) * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc((void**)&w.comp8_alt, (A + 1) * sizeof(uint2));
    if (e == hipSuccess) e = hipMalloc((void**)&w.heads_w0, (A ? A : 1) * sizeof(AxcBurstHeads8));
    if (e == hipSuccess) e = hipMalloc((void**)&w.rev_off, (A + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&w.rev_cut, (A ? A : 1) * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc((void**)&d.aux, (A ? A : 1) * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void**)&d.ctl, 4 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&d.act, 2 * (A ? A : 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc((void**)&d.piece_cnt, 2 * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(d.piece_cnt, 0, 2 * sizeof(uint32_t), st);
    if (e == hipSuccess) e = hipMalloc((void**)&w.cnt, kWinCounters * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void**)&w.cpart, (size_t)kListBlocks * 2 * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemsetAsync(w.cnt, 0, kWinCounters * sizeof(uint64_t), st);
    if (e == hipSuccess) e = hipMemsetAsync(w.cpart, 0, (size_t)kListBlocks * 2 * sizeof(uint64_t), st);
    // Stamps start at 0 and the first window's is 1: no record is current.
    if (e == hipSuccess) e = hipMemsetAsync(d.aux, 0, (A ? A : 1) * sizeof(uint64_t), st);
    // The second snapshot's inert entry (the first's is the plan's).
    if (e == hipSuccess) e = hipMemsetAsync(w.comp8_alt + A, 0xFF, sizeof(uint2), st);
    if (e == hipSuccess) e = hipMemsetAsync(d.ctl, 0, 4 * sizeof(uint32_t), st);
    s->win = w;
    s->dirty = d;
    if (e != hipSuccess) {
      free_window(s);
      return e;
    }
  }
  const uint64_t ver = s->targets_version.load();
  if (s->rev_ready && s->rev_version == ver) return hipSuccess;
  // Counts per axon into rev_off, their exclusive prefix in place, the
  // total read back to size rev, then the entries (at a cursor per axon).
  if ((e = hipMemsetAsync(s->win.rev_off, 0, (A + 1) * sizeof(uint32_t), st)) != hipSuccess) return e;
  if (N) rev_count_kernel<<<blocks_for(N), kThreads, 0, st>>>((uint32_t)N, (uint32_t)A, s->sp, s->win.rev_off);
  rev_scan_kernel<<<1, 1024, 0, st>>>(s->win.rev_off, (uint32_t)A, s->win.rev_off);
  uint32_t total = 0;
  if ((e = hipMemcpyAsync(&total, s->win.rev_off + A, 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
  if (total > s->rev_cap || !s->win.rev) {
    // The index, and the correction pieces of a tick: an axon's range of
    // c entries makes ceil(c / 16) pieces, so at most total / 16 + A.
    (void)hipFree(s->win.rev);
    (void)hipFree(s->dirty.pieces);
    s->win.rev = nullptr;
    s->dirty.pieces = nullptr;
    s->rev_cap = 0;
    const size_t cap = (size_t)total + total / 8 + 1024;
    if ((e = hipMalloc((void**)&s->win.rev, cap * sizeof(uint64_t))) != hipSuccess) return e;
    if ((e = hipMalloc((void**)&s->dirty.pieces, (cap / 16 + A + 16) * sizeof(uint64_t))) != hipSuccess) return e;
    s->rev_cap = cap;
  }
  // The entries into a temporary array, then sorted by segment into rev.
  uint64_t* tmp = nullptr;
  uint32_t* cursor = nullptr;
  e = hipMalloc((void**)&tmp, ((size_t)total + 1) * sizeof(uint64_t));
  if (e == hipSuccess) e = hipMalloc((void**)&cursor, (A + 1) * sizeof(uint32_t));
  if (e == hipSuccess)
    e = hipMemcpyAsync(cursor, s->win.rev_off, (A + 1) * sizeof(uint32_t), hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) {
    (void)hipFree(tmp);
    (void)hipFree(cursor);
    return e;
  }
  if (N) rev_fill_kernel<<<blocks_for(N), kThreads, 0, st>>>((uint32_t)N, (uint32_t)A, s->sp, flags, cursor, tmp);
  if (A) rev_sort_kernel<<<std::min<uint32_t>(blocks_for(A, 4), 8192u), kThreads, 0, st>>>((uint32_t)A,
                                                                                           s->win.rev_off, tmp,
                                                                                           s->win.rev,
                                                                                           s->win.rev_cut);
  if (e == hipSuccess) e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(st);
  (void)hipFree(tmp);
  (void)hipFree(cursor);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return e;
  s->rev_ready = true;
  s->rev_version = ver;
  return hipSuccess;
}

// The T ticks of a batch in windows of `win` ticks, on the device
// parameters prm (tick_base, schedule counts, slices), as the batch graphs
// run them; `timed` times each launch (never while capturing).
hipError_t enqueue_windows(const ShardVramPtrs& v, Shard* s, uint32_t v_seg, uint32_t T, uint32_t win, uint32_t words,
                           uint32_t virt_off, uint32_t n_virt, bool has_bitmask, bool has_spikes, uint32_t max_spk,
                           const uint32_t* mapped, bool has_out, uint32_t n_out, const TickParams* prm,
                           hipStream_t st, const AxcVariantParameters* lut, bool timed) {
  const uint32_t N = s->N, A = s->A;
  const bool k2 = has_bitmask && n_virt > 0;
  const bool k3 = has_spikes && max_spk > 0;
  const size_t lds = (kStageCap + 1) * sizeof(uint2);
  const uint32_t* axon_inv = s->sp.perm ? s->d_axon_inv : nullptr;
  // The byte snapshot of the heads at the batch's start into the plan's
  // buffer; the windows then take the two buffers in turn as their start
  // snapshot, writing the current one into the other.
  // The wrap limit of the window starting at t0 (2^32 - (K+1) v_seg; 0: no
  // window there).
  auto wrap_lim = [&](uint32_t t0) -> uint32_t {
    if (t0 >= T || v_seg == 0) return 0u;
    const uint32_t K = std::min(win, T - t0);
    return (uint32_t)((1ull << 32) - (uint64_t)(K + 1) * v_seg);
  };
  {
    TimedScope ts(kKWinStart, st, timed);
    win_snapshot_kernel<<<blocks_for((uint64_t)A + 1), kThreads, 0, st>>>(v.axon_heads, A, s->sp.comp8, wrap_lim(0),
                                                                         s->win, s->dirty);
  }
  uint32_t wi = 0;
  for (uint32_t t0 = 0; t0 < T; t0 += win, wi++) {
    const uint32_t K = std::min(win, T - t0);
    WinState w = s->win;
    w.comp8_w0 = (wi & 1) ? s->win.comp8_alt : s->sp.comp8;
    uint2* cur = (wi & 1) ? s->sp.comp8 : s->win.comp8_alt;
    StagePlan spc = s->sp;  // GSOP's current snapshot
    spc.comp8 = cur;
    {
      TimedScope ts(kKWinPass, st, timed);
      const uint32_t g = staged_grid(N);
      const uint32_t one2 = 0x00010001u;
      auto pass = [&](auto kern) { kern<<<g, kStageThreads, lds, st>>>(v, N, A, lut, s->sp, w, s->dirty, v_seg, one2); };
      if (N <= kBufMaxSomas) {
        if (K == 1) pass(win_pass_kernel<1, true>);
        else if (K == 2) pass(win_pass_kernel<2, true>);
        else if (K == 3) pass(win_pass_kernel<3, true>);
        else pass(win_pass_kernel<4, true>);
      } else {
        if (K == 1) pass(win_pass_kernel<1, false>);
        else if (K == 2) pass(win_pass_kernel<2, false>);
        else if (K == 3) pass(win_pass_kernel<3, false>);
        else pass(win_pass_kernel<4, false>);
      }
    }
    for (uint32_t j = 0; j < K; j++) {
      const uint32_t t = t0 + j;
      if (k3) {
        TimedScope ts(kKSpikes, st, timed);
        spike_count_kernel<<<std::min<uint32_t>(blocks_for(max_spk), 64u), kThreads, 0, st>>>(
            nullptr, 0, A, s->sc.push_cnt, prm, t, max_spk, s->sp.axon_map);
      }
      {
        TimedScope ts(kKAxon, st, timed);
        const uint32_t nv = k2 ? n_virt : 0u;
        const int uc = k3 ? 1 : 0;
        if (axon_inv) {
          win_axon_kernel<true><<<blocks_for(A ? A : 1, 256 * kAxonsPerLane), kThreads, 0, st>>>(
              A, v_seg, virt_off, nv, s->sc.push_cnt, uc, s->sc.counters, prm, t, words, cur, axon_inv, w, s->dirty,
              j);
          win_active_kernel<true><<<kListBlocks / 4, kThreads, 0, st>>>(v.axon_heads, A, v_seg, virt_off, nv,
                                                                       s->sc.push_cnt, uc, prm, t, words, cur,
                                                                       axon_inv, lut, w, s->dirty, j);
        } else {
          win_axon_kernel<false><<<blocks_for(A ? A : 1, 256 * kAxonsPerLane), kThreads, 0, st>>>(
              A, v_seg, virt_off, nv, s->sc.push_cnt, uc, s->sc.counters, prm, t, words, cur, nullptr, w, s->dirty,
              j);
          win_active_kernel<false><<<kListBlocks / 4, kThreads, 0, st>>>(v.axon_heads, A, v_seg, virt_off, nv,
                                                                        s->sc.push_cnt, uc, prm, t, words, cur,
                                                                        nullptr, lut, w, s->dirty, j);
        }
      }
      {
        TimedScope ts(kKWinCorrect, st, timed);
        win_correct_kernel<<<kListBlocks, kThreads, 0, st>>>(v.axon_heads, v.soma_flags, N, A, lut, s->sp, w, s->dirty,
                                                             v_seg, j);
      }
      {
        TimedScope ts(kKWinSoma, st, timed);
        win_soma_kernel<<<std::min<uint32_t>(blocks_for(N), kPersistBlocks), kThreads, 0, st>>>(v, N, A, lut, prm, t,
                                                                                             s->sc, s->sp, w, j);
      }
      {
        TimedScope ts(kKPush, st, timed);
        const uint32_t n_words = (N + 63) / 64;
        compact_spikes_kernel<<<blocks_for(n_words), kThreads, 0, st>>>(n_words, s->sc);
        win_push_kernel<<<kListBlocks / 4, kThreads, 0, st>>>(v.axon_heads, v.soma_to_axon, A, s->sc, cur, w, s->dirty,
                                                             v_seg, j);
      }
      {
        TimedScope ts(kKGsop, st, timed);
        gsop_win_kernel<<<kListBlocks, kThreads, 0, st>>>(v, N, A, lut, prm, 0, s->sc, spc, w, s->dirty, v_seg, j, K);
      }
      if (has_out && n_out > 0) {
        TimedScope ts(kKReadout, st, timed);
        readout_kernel<<<blocks_for(n_out), kThreads, 0, st>>>(v.soma_flags, N, mapped, nullptr, n_out, prm, t,
                                                               s->sp.perm ? s->d_perm_inv : nullptr);
      }
    }
    {
      TimedScope ts(kKWinEnd, st, timed);
      win_end_kernel<<<blocks_for(A), kThreads, 0, st>>>(v.axon_heads, A, v_seg, K, wrap_lim(t0 + K), w, s->dirty);
    }
  }
  return hipGetLastError();
}

// ---- the allocation-free per-tick entry point ------------------------------
// Device scratch of axc_step_day_phase_noalloc: the direct scan's workspace
// (list counters, the compact head snapshot, push counts, owner and spike
// lists, the kernels' counter partials) for one shard of at most
// kNoAllocScratchBytes, carved out per call. Every GPU the library runs on
// holds its own copy until the process ends. Host code never passes the
// array's name to a kernel (on the host it is a stand-in, not a device
// address): it uses the per-device address from hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_na_mu;
std::map<int, unsigned char*> g_na_ptr;  // device -> address of that device's scratch

// Bytes of scratch a shard of (N, A) needs, each piece 256-byte aligned.
struct NoAllocLayout {
  size_t counters, comp, push_cnt, own_list, spk_list, stats_part, stats_part2, total;
};
NoAllocLayout noalloc_layout(uint32_t N, uint32_t A) {
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  NoAllocLayout l{};
  size_t o = 0;
  l.counters = o; o += up(kNumCounters * 4);
  l.comp = o; o += up(((size_t)A + 1) * sizeof(uint4));
  l.push_cnt = o; o += up((size_t)(A ? A : 1) * 4);
  l.own_list = o; o += up((size_t)(A ? A : 1) * 4);
  l.spk_list = o; o += up((size_t)(N ? N : 1) * 4);
  l.stats_part = o; o += up((size_t)neuron_blocks(N ? N : 1) * kStatsPerBlock * 8);
  l.stats_part2 = o; o += up((size_t)kListBlocks * kStatsPerBlock * 8);
  l.total = o;
  return l;
}

// The address of the current device's scratch, looked up once per device
// (with that device current); null on failure (*err set).
unsigned char* noalloc_scratch_ptr(hipError_t* err) {
  int dev = 0;
  *err = hipGetDevice(&dev);
  if (*err != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_na_mu);
  auto it = g_na_ptr.find(dev);
  if (it != g_na_ptr.end()) return it->second;
  void* p = nullptr;
  *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (*err != hipSuccess || !p) {
    if (*err == hipSuccess) *err = hipErrorInvalidSymbol;
    return nullptr;
  }
  g_na_ptr[dev] = (unsigned char*)p;
  return (unsigned char*)p;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) p[j] = v;
}

}  // namespace

extern "C" {

// ===========================================================================
// Shard memory (bindings.cu:571-691)
// ===========================================================================
int32_t cu_allocate_shard(uint32_t padded_n, uint32_t total_axons, ShardVramPtrs* out) {
  if (!out) return (int32_t)hipErrorInvalidValue;
  const size_t N = padded_n;
  const size_t total = N * AXC_STATE_BYTES_PER_NEURON;
  void* base = nullptr;
  hipError_t e = hipMalloc(&base, total ? total : 64);
  if (e != hipSuccess) {
    fprintf(stderr, "[cu_allocate_shard] hipMalloc state failed: %s\n", hipGetErrorString(e));
    return (int32_t)e;
  }
  if ((e = hipMemset(base, 0, total ? total : 64)) != hipSuccess) {
    (void)hipFree(base);
    return (int32_t)e;
  }
  char* b = (char*)base;
  size_t off = 0;
  out->soma_voltage = (int32_t*)(b + off); off += 4 * N;
  out->soma_flags = (uint8_t*)(b + off); off += 1 * N;
  out->threshold_offset = (int32_t*)(b + off); off += 4 * N;
  out->timers = (uint8_t*)(b + off); off += 1 * N;
  out->soma_to_axon = (uint32_t*)(b + off); off += 4 * N;
  out->dendrite_targets = (uint32_t*)(b + off); off += 512 * N;
  out->dendrite_weights = (int16_t*)(b + off); off += 256 * N;
  out->dendrite_timers = (uint8_t*)(b + off);
  e = hipMalloc((void**)&out->axon_heads, (size_t)(total_axons ? total_axons : 1) * sizeof(AxcBurstHeads8));
  if (e != hipSuccess) {
    fprintf(stderr, "[cu_allocate_shard] hipMalloc axon_heads failed: %s\n", hipGetErrorString(e));
    (void)hipFree(base);
    return (int32_t)e;
  }
  if (total_axons) init_sentinels_kernel<<<blocks_for(total_axons), kThreads>>>(out->axon_heads, total_axons);
  e = hipDeviceSynchronize();
  Shard* s = e == hipSuccess ? shard_for(out, padded_n, total_axons, &e) : nullptr;
  if (!s) {  // nothing is left allocated on failure
    (void)hipFree(out->axon_heads);
    (void)hipFree(base);
    *out = ShardVramPtrs{};
    return (int32_t)e;
  }
  s->owned_by_library = true;
  return 0;
}

int32_t cu_upload_state_blob(const ShardVramPtrs* vram, const void* blob, size_t size) {
  if (!vram || (!blob && size)) return (int32_t)hipErrorInvalidValue;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_shards.find(vram->soma_voltage);
    if (it != g_shards.end() && size > (size_t)it->second->N * AXC_STATE_BYTES_PER_NEURON)
      return (int32_t)hipErrorInvalidValue;
  }
  hipError_t e = hipMemcpy(vram->soma_voltage, blob, size, hipMemcpyHostToDevice);
  mark_stage_dirty(vram->soma_voltage);
  if (e != hipSuccess) {
    fprintf(stderr, "[cu_upload_state_blob] hipMemcpy failed: %s\n", hipGetErrorString(e));
    return (int32_t)e;
  }
  return 0;
}

int32_t cu_upload_axons_blob(const ShardVramPtrs* vram, const void* blob, size_t size) {
  if (!vram || (!blob && size)) return (int32_t)hipErrorInvalidValue;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_shards.find(vram->soma_voltage);
    if (it != g_shards.end() && size > (size_t)it->second->A * sizeof(AxcBurstHeads8))
      return (int32_t)hipErrorInvalidValue;
  }
  hipError_t e = hipMemcpy(vram->axon_heads, blob, size, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    fprintf(stderr, "[cu_upload_axons_blob] hipMemcpy failed: %s\n", hipGetErrorString(e));
    return (int32_t)e;
  }
  return 0;
}

void cu_free_shard(ShardVramPtrs* vram) {
  if (!vram) return;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_shards.find(vram->soma_voltage);
    if (it != g_shards.end()) {
      (void)hipDeviceSynchronize();
      destroy_shard(it->second);
      g_shards.erase(it);
    }
  }
  if (vram->soma_voltage) {
    (void)hipFree(vram->soma_voltage);
    vram->soma_voltage = nullptr;
    vram->soma_flags = nullptr;
    vram->threshold_offset = nullptr;
    vram->timers = nullptr;
    vram->soma_to_axon = nullptr;
    vram->dendrite_targets = nullptr;
    vram->dendrite_weights = nullptr;
    vram->dendrite_timers = nullptr;
  }
  if (vram->axon_heads) {
    (void)hipFree(vram->axon_heads);
    vram->axon_heads = nullptr;
  }
}

// ===========================================================================
// Day phase (physics.cu:441-506)
// ===========================================================================
void cu_reset_burst_counters(const ShardVramPtrs* ptrs, uint32_t padded_n, void* stream) {
  if (!ptrs || padded_n == 0) return;
  reset_burst_kernel<<<blocks_for(padded_n), kThreads, 0, (hipStream_t)stream>>>(ptrs->soma_flags, padded_n);
}

int32_t cu_step_day_phase(const ShardVramPtrs* vram, uint32_t padded_n, uint32_t total_axons, uint32_t v_seg,
                          uint32_t current_tick, const uint32_t* input_bitmask, uint32_t virtual_offset,
                          uint32_t num_virtual_axons, const uint32_t* incoming_spikes,
                          uint32_t num_incoming_spikes, const uint32_t* mapped_soma_ids,
                          uint8_t* output_history, uint32_t num_outputs, int16_t dopamine, void* stream) {
  if (!vram) return (int32_t)hipErrorInvalidValue;
  if (num_virtual_axons > 0 && input_bitmask &&
      (uint64_t)virtual_offset + num_virtual_axons > (uint64_t)total_axons)
    return (int32_t)hipErrorInvalidValue;
  hipError_t e;
  Shard* s = shard_for(vram, padded_n, total_axons, &e);
  if (!s) return (int32_t)e;
  const AxcVariantParameters* lut = device_lut();
  if (!lut) return (int32_t)hipErrorOutOfMemory;
  const bool staged = ensure_stage(s, *vram, (hipStream_t)stream, &e, virtual_offset,
                                   input_bitmask ? num_virtual_axons : 0u);
  if (e != hipSuccess) return (int32_t)e;
  hipStream_t st = (hipStream_t)stream;
  if (staged && s->ordered()) {
    // In the engine's order: the caller's state in, one tick on the private
    // arrays and the working copy of the weights, the state back out.
    if (order_moves_inputs(s, virtual_offset, input_bitmask ? num_virtual_axons : 0u))
      return (int32_t)hipErrorInvalidValue;
    order_copy(s, *vram, true, st);
    copy_weights_kernel<true><<<blocks_for(padded_n), kThreads, 0, st>>>(vram->dendrite_weights, padded_n, s->sp);
    e = enqueue_tick(s->shadow, s, v_seg, current_tick, input_bitmask, 0, virtual_offset, num_virtual_axons,
                     incoming_spikes, 0, num_incoming_spikes, mapped_soma_ids, output_history, num_outputs,
                     (int32_t)dopamine, nullptr, st, lut, s->flip, false, true, true, true);
    copy_weights_kernel<false><<<blocks_for(padded_n), kThreads, 0, st>>>(vram->dendrite_weights, padded_n, s->sp);
    order_copy(s, *vram, false, st);
    if (e == hipSuccess) e = hipGetLastError();
    return (int32_t)e;
  }
  e = enqueue_tick(*vram, s, v_seg, current_tick, input_bitmask, 0, virtual_offset, num_virtual_axons,
                   incoming_spikes, 0, num_incoming_spikes, mapped_soma_ids, output_history, num_outputs,
                   (int32_t)dopamine, nullptr, st, lut, s->flip, false, true, staged);
  s->flip ^= 1u;
  return (int32_t)e;
}

int32_t axc_step_day_phase_noalloc(const ShardVramPtrs* vram, uint32_t padded_n, uint32_t total_axons,
                                   uint32_t v_seg, uint32_t current_tick, const uint32_t* input_bitmask,
                                   uint32_t virtual_offset, uint32_t num_virtual_axons,
                                   const uint32_t* incoming_spikes, uint32_t num_incoming_spikes,
                                   const uint32_t* mapped_soma_ids, uint8_t* output_history, uint32_t num_outputs,
                                   int16_t dopamine, void* stream) {
  if (!vram) return (int32_t)hipErrorInvalidValue;
  if (num_virtual_axons > 0 && input_bitmask &&
      (uint64_t)virtual_offset + num_virtual_axons > (uint64_t)total_axons)
    return (int32_t)hipErrorInvalidValue;
  const NoAllocLayout l = noalloc_layout(padded_n, total_axons);
  if (l.total > kNoAllocScratchBytes) {
    fprintf(stderr,
            "[axc_step_day_phase_noalloc] padded_n=%u total_axons=%u needs %zu bytes of scratch; the limit is %zu "
            "(use cu_step_day_phase)\n",
            padded_n, total_axons, l.total, kNoAllocScratchBytes);
    return (int32_t)hipErrorInvalidValue;
  }
  const AxcVariantParameters* lut = device_lut_if_present();
  if (!lut) return (int32_t)hipErrorNotInitialized;  // cu_upload_constant_memory first
  hipError_t e;
  unsigned char* base = noalloc_scratch_ptr(&e);
  if (!base) return (int32_t)e;
  // A workspace view over the scratch; no registry entry, no staging plan.
  Shard w;
  w.N = padded_n;
  w.A = total_axons;
  w.sc.counters = (uint32_t*)(base + l.counters);
  w.sc.comp[0] = w.sc.comp[1] = (uint4*)(base + l.comp);
  w.sc.push_cnt = (uint32_t*)(base + l.push_cnt);
  w.sc.own_list = (uint32_t*)(base + l.own_list);
  w.sc.spk_list = (uint32_t*)(base + l.spk_list);
  w.sc.stats_part = (uint64_t*)(base + l.stats_part);
  w.sc.stats_part2 = (uint64_t*)(base + l.stats_part2);
  w.sc.spk_mask = nullptr;  // staged pass only
  hipStream_t st = (hipStream_t)stream;
  // The scratch keeps the last call's contents (any shard's): the push counts
  // must start at zero and the snapshot's entry A must be the inert axon.
  fill_u32_kernel<<<std::min<uint32_t>(blocks_for(total_axons ? total_axons : 1), 1024u), kThreads, 0, st>>>(
      w.sc.push_cnt, total_axons, 0u);
  fill_u32_kernel<<<1, 64, 0, st>>>((uint32_t*)(w.sc.comp[0] + total_axons), 4u, 0xFFFFFFFFu);
  e = enqueue_tick(*vram, &w, v_seg, current_tick, input_bitmask, 0, virtual_offset, num_virtual_axons,
                   incoming_spikes, 0, num_incoming_spikes, mapped_soma_ids, output_history, num_outputs,
                   (int32_t)dopamine, nullptr, st, lut, 0, false, true, false, false, false);
  return (int32_t)e;
}

int32_t cu_upload_constant_memory(const AxcVariantParameters* lut) {
  if (!lut) return (int32_t)hipErrorInvalidValue;
  AxcVariantParameters* d = device_lut();
  if (!d) return (int32_t)hipErrorOutOfMemory;
  // The table is read by in-flight ticks on any stream of this device: wait
  // for them, then overwrite (the reference's symbol copy has the same effect
  // on a device without concurrent streams).
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return (int32_t)e;
  return (int32_t)hipMemcpy(d, lut, 16 * sizeof(AxcVariantParameters), hipMemcpyHostToDevice);
}

// ===========================================================================
// Runtime helpers (bindings.cu:210-291)
// ===========================================================================
void* gpu_malloc(size_t size) {
  void* p = nullptr;
  if (hipMalloc(&p, size) != hipSuccess) return nullptr;
  return p;
}
void gpu_free(void* p) {
  if (p) (void)hipFree(p);
}
void* gpu_host_alloc(size_t size) {
  void* p = nullptr;
  if (hipHostMalloc(&p, size, hipHostMallocDefault) != hipSuccess) return nullptr;
  return p;
}
void gpu_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}
void gpu_memcpy_host_to_device_async(void* dst, const void* src, size_t size, void* stream) {
  (void)hipMemcpyAsync(dst, src, size, hipMemcpyHostToDevice, (hipStream_t)stream);
  note_device_write(dst, size, true, (hipStream_t)stream);
}
void gpu_memcpy_device_to_host_async(void* dst, const void* src, size_t size, void* stream) {
  (void)hipMemcpyAsync(dst, src, size, hipMemcpyDeviceToHost, (hipStream_t)stream);
}
bool gpu_memcpy_host_to_device(void* dst, const void* src, size_t size) {
  hipError_t e = hipMemcpy(dst, src, size, hipMemcpyHostToDevice);
  note_device_write(dst, size);
  if (e != hipSuccess) fprintf(stderr, "[gpu_memcpy_host_to_device] %s (size %zu)\n", hipGetErrorString(e), size);
  return e == hipSuccess;
}
bool gpu_memcpy_device_to_host(void* dst, const void* src, size_t size) {
  hipError_t e = hipMemcpy(dst, src, size, hipMemcpyDeviceToHost);
  if (e != hipSuccess) fprintf(stderr, "[gpu_memcpy_device_to_host] %s (src=%p, size=%zu)\n", hipGetErrorString(e), src, size);
  return e == hipSuccess;
}
int32_t gpu_stream_create(void** out_stream) {
  if (!out_stream) return (int32_t)hipErrorInvalidValue;
  return (int32_t)hipStreamCreateWithFlags((hipStream_t*)out_stream, hipStreamNonBlocking);
}
int32_t gpu_stream_destroy(void* stream) { return (int32_t)hipStreamDestroy((hipStream_t)stream); }
void gpu_stream_synchronize(void* stream) { (void)hipStreamSynchronize((hipStream_t)stream); }
void gpu_set_device(int32_t id) {
  hipError_t e = hipSetDevice(id);
  if (e != hipSuccess) fprintf(stderr, "FATAL HIP ERROR in gpu_set_device: %s\n", hipGetErrorString(e));
}
void gpu_device_synchronize(void) { (void)hipDeviceSynchronize(); }
void gpu_synchronize(void) { (void)hipDeviceSynchronize(); }
void gpu_load_constants(const void* host_ptr) {
  (void)cu_upload_constant_memory((const AxcVariantParameters*)host_ptr);
}
void update_constant_memory_hot_reload(const AxcVariantParameters* nv, void* stream) {
  AxcVariantParameters* d = device_lut();
  if (!d || !nv) return;
  (void)hipMemcpyAsync(d, nv, 16 * sizeof(AxcVariantParameters), hipMemcpyHostToDevice, (hipStream_t)stream);
}

// ===========================================================================
// Night (bindings.cu:450-515, 699-710)
// ===========================================================================
void launch_sort_and_prune(const ShardVramPtrs* ptrs, uint32_t padded_n, int16_t prune_threshold) {
  if (!ptrs || padded_n == 0) return;
  mark_stage_dirty(ptrs->soma_voltage);
  sort_prune_kernel<<<blocks_for(padded_n, 64), kThreads>>>(ptrs->soma_flags, ptrs->dendrite_targets,
                                                             ptrs->dendrite_weights, ptrs->dendrite_timers,
                                                             padded_n, (int32_t)prune_threshold);
}

// ===========================================================================
// Multi-zone helpers (bindings.cu:341-436)
// ===========================================================================
void launch_extract_outgoing_spikes(const AxcBurstHeads8* heads, const uint32_t* src, const uint32_t* dst_ghost,
                                    uint32_t count, uint32_t T, uint32_t v_seg, void* out_events,
                                    uint32_t* out_count, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (!out_count) return;
  zero_u32_kernel<<<1, 1, 0, st>>>(out_count);
  if (count == 0 || v_seg == 0) return;
  extract_spikes_kernel<<<blocks_for(count), kThreads, 0, st>>>(heads, src, dst_ghost, count, T, v_seg,
                                                                 (uint32_t*)out_events, out_count);
}

void launch_ghost_sync(const AxcBurstHeads8* src, AxcBurstHeads8* dst, const uint32_t* si, const uint32_t* di,
                       uint32_t count, void* stream) {
  if (count == 0) return;
  ghost_sync_kernel<<<blocks_for(count), kThreads, 0, (hipStream_t)stream>>>(src, dst, si, di, count);
}

// ===========================================================================
// Telemetry (physics.cu:400-434, bindings.cu:780-796)
// ===========================================================================
void gpu_reset_telemetry_count(const ShardVramPtrs* ptrs, void* stream) {
  (void)ptrs;
  (void)stream;
}

void launch_extract_telemetry(const ShardVramPtrs* ptrs, uint32_t padded_n, uint32_t* out_ids,
                              uint32_t* out_count, void* stream) {
  if (!ptrs || !out_ids || !out_count || padded_n == 0) return;
  hipError_t e;
  Shard* s = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_shards.find(ptrs->soma_voltage);
    if (it != g_shards.end()) s = it->second;
  }
  if (!s) {
    s = shard_for(ptrs, padded_n, 1, &e);
    if (!s) return;
  }
  hipStream_t st = (hipStream_t)stream;
  const uint32_t nb = blocks_for(padded_n);
  // The spike list buffer (N entries) holds the per-block counts here; it is
  // rewritten by the next tick before use.
  telem_count_kernel<<<nb, kThreads, 0, st>>>(ptrs->soma_flags, padded_n, s->sc.spk_list);
  telem_scan_kernel<<<1, 1024, 0, st>>>(s->sc.spk_list, nb, out_count);
  telem_write_kernel<<<nb, kThreads, 0, st>>>(ptrs->soma_flags, padded_n, s->sc.spk_list, out_ids);
}

// ===========================================================================
// I/O buffers (bindings.cu:717-778)
// ===========================================================================
int32_t cu_allocate_io_buffers(uint32_t input_words, uint32_t schedule_capacity, uint32_t output_capacity,
                               uint32_t** d_input_bitmask, uint32_t** d_incoming_spikes,
                               uint8_t** d_output_history) {
  if (!d_input_bitmask || !d_incoming_spikes || !d_output_history) return -1;
  *d_input_bitmask = nullptr;
  *d_incoming_spikes = nullptr;
  *d_output_history = nullptr;
  bool ok = input_words == 0 || hipMalloc((void**)d_input_bitmask, (size_t)input_words * 4) == hipSuccess;
  ok = ok && (schedule_capacity == 0 ||
              hipMalloc((void**)d_incoming_spikes, (size_t)schedule_capacity * 4) == hipSuccess);
  ok = ok && (output_capacity == 0 || hipMalloc((void**)d_output_history, (size_t)output_capacity) == hipSuccess);
  if (ok) return 0;
  // On failure nothing stays allocated and every output is null.
  for (void** q : {(void**)d_input_bitmask, (void**)d_incoming_spikes, (void**)d_output_history}) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
  return -1;
}

int32_t cu_dma_h2d_io(uint32_t* d_in, const uint32_t* h_in, uint32_t words, uint32_t* d_spk, const uint32_t* h_spk,
                      uint32_t cap, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (words > 0 && d_in && h_in) (void)hipMemcpyAsync(d_in, h_in, (size_t)words * 4, hipMemcpyHostToDevice, st);
  if (cap > 0 && d_spk && h_spk) (void)hipMemcpyAsync(d_spk, h_spk, (size_t)cap * 4, hipMemcpyHostToDevice, st);
  return 0;
}

int32_t axc_dma_h2d_schedule(uint32_t* d_spikes, const uint32_t* h_spikes, uint32_t max_spk, const uint32_t* h_counts,
                             uint32_t T, void* stream) {
  if (T == 0) return 0;
  if (!h_counts) return (int32_t)hipErrorInvalidValue;
  uint32_t widest = 0;
  for (uint32_t t = 0; t < T; t++) {
    if (h_counts[t] > max_spk) return (int32_t)hipErrorInvalidValue;
    widest = std::max(widest, h_counts[t]);
  }
  if (widest == 0) return 0;
  if (!d_spikes || !h_spikes) return (int32_t)hipErrorInvalidValue;
  const size_t pitch = (size_t)max_spk * 4;
  return (int32_t)hipMemcpy2DAsync(d_spikes, pitch, h_spikes, pitch, (size_t)widest * 4, T, hipMemcpyHostToDevice,
                                   (hipStream_t)stream);
}

int32_t cu_dma_d2h_io(uint8_t* h_out, const uint8_t* d_out, uint32_t cap, void* stream) {
  if (cap > 0 && d_out && h_out)
    (void)hipMemcpyAsync(h_out, d_out, cap, hipMemcpyDeviceToHost, (hipStream_t)stream);
  return 0;
}

// ===========================================================================
// Legacy launchers (ffi.rs:193-233): routed to the current kernels on the
// null stream. The reference's launch_apply_spike_batch passes a host pointer
// as the count (bindings.cu:316-317); here the count argument is used.
// ===========================================================================
static ShardVramPtrs legacy_ptrs(const AxcLegacyVramState& v) {
  ShardVramPtrs p;
  p.soma_voltage = v.voltage;
  p.soma_flags = v.flags;
  p.threshold_offset = v.threshold_offset;
  p.timers = v.refractory_timer;
  p.soma_to_axon = v.soma_to_axon;
  p.dendrite_targets = v.dendrite_targets;
  p.dendrite_weights = v.dendrite_weights;
  p.dendrite_timers = v.dendrite_timers;
  p.axon_heads = v.axon_heads;
  return p;
}

void launch_inject_inputs(AxcLegacyVramState v, const uint32_t* bitmask, uint32_t current_tick,
                          uint32_t total_virtual_axons) {
  (void)current_tick;
  if (!bitmask || total_virtual_axons == 0 || v.total_axons == 0) return;
  // Inputs at the start of the axon array, stride 1 (bindings.cu:295-302),
  // without propagation: a push-only pass.
  hipError_t e;
  ShardVramPtrs p = legacy_ptrs(v);
  Shard* s = shard_for(&p, v.padded_n, v.total_axons, &e);
  if (!s) return;
  uint32_t n = total_virtual_axons < v.total_axons ? total_virtual_axons : v.total_axons;
  axon_tick_kernel<<<blocks_for(v.total_axons), kThreads>>>(v.axon_heads, n, 0u, bitmask, 0u, n, s->sc.push_cnt, 0,
                                                            s->sc.counters, nullptr, 0u, 0u, nullptr, nullptr);
}
void launch_apply_spike_batch(AxcLegacyVramState v, const void* schedule, uint32_t count) {
  if (!schedule || count == 0 || v.total_axons == 0) return;
  hipError_t e;
  ShardVramPtrs p = legacy_ptrs(v);
  Shard* s = shard_for(&p, v.padded_n, v.total_axons, &e);
  if (!s) return;
  spike_count_kernel<<<blocks_for(count), kThreads>>>((const uint32_t*)schedule, count, v.total_axons, s->sc.push_cnt,
                                                      nullptr, 0u, 0u);
  axon_tick_kernel<<<blocks_for(v.total_axons), kThreads>>>(v.axon_heads, v.total_axons, 0u, nullptr, 0u, 0u,
                                                            s->sc.push_cnt, 1, s->sc.counters, nullptr, 0u, 0u, nullptr, nullptr);
}
void launch_propagate_axons(AxcLegacyVramState v, uint32_t v_seg) {
  if (v.total_axons == 0) return;
  hipError_t e;
  ShardVramPtrs p = legacy_ptrs(v);
  Shard* s = shard_for(&p, v.padded_n, v.total_axons, &e);
  if (!s) return;
  axon_tick_kernel<<<blocks_for(v.total_axons), kThreads>>>(v.axon_heads, v.total_axons, v_seg, nullptr, 0u, 0u,
                                                            s->sc.push_cnt, 0, s->sc.counters, nullptr, 0u, 0u, nullptr, nullptr);
}
void launch_update_neurons(AxcLegacyVramState v, const void* constants_ptr, uint32_t current_tick) {
  (void)constants_ptr;
  if (v.padded_n == 0) return;
  hipError_t e;
  ShardVramPtrs p = legacy_ptrs(v);
  Shard* s = shard_for(&p, v.padded_n, v.total_axons, &e);
  const AxcVariantParameters* lut = device_lut();
  if (!s || !lut) return;
  // Needs the snapshot of the current heads: rebuilt from them first.
  compact_all_kernel<<<blocks_for(v.total_axons + 1), kThreads>>>(p.axon_heads, v.total_axons, s->sc.comp[0]);
  neuron_tick_kernel<false><<<neuron_blocks(v.padded_n), kThreads>>>(p, v.padded_n, v.total_axons, lut, nullptr,
                                                                  current_tick, s->sc, s->sc.comp[0], s->sc.comp[1], 0);
  push_apply_kernel<<<kListBlocks / 4, kThreads>>>(p.axon_heads, s->sc, s->sc.comp[0]);
}
void launch_apply_gsop(AxcLegacyVramState v) {
  if (v.padded_n == 0) return;
  hipError_t e;
  ShardVramPtrs p = legacy_ptrs(v);
  Shard* s = shard_for(&p, v.padded_n, v.total_axons, &e);
  const AxcVariantParameters* lut = device_lut();
  if (!s || !lut) return;
  compact_all_kernel<<<blocks_for(v.total_axons + 1), kThreads>>>(p.axon_heads, v.total_axons, s->sc.comp[0]);
  gsop_kernel<<<kListBlocks, kThreads>>>(p, v.padded_n, v.total_axons, lut, nullptr, 0, s->sc, s->sc.comp[0]);
}
void launch_record_readout(AxcLegacyVramState v, const uint32_t* mapped, uint8_t* out, uint32_t current_tick,
                           uint32_t total_pixels) {
  if (!mapped || !out || total_pixels == 0) return;
  readout_kernel<<<blocks_for(total_pixels), kThreads>>>(v.flags, v.padded_n, mapped,
                                                         out + (size_t)current_tick * total_pixels, total_pixels,
                                                         nullptr, 0u);
}

// ===========================================================================
// Engine extensions
// ===========================================================================
int32_t axc_shard_stats_reset(const ShardVramPtrs* vram, void* stream) {
  if (!vram) return (int32_t)hipErrorInvalidValue;
  Shard* s = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_shards.find(vram->soma_voltage);
    if (it != g_shards.end()) s = it->second;
  }
  if (!s) return (int32_t)hipErrorInvalidValue;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = hipMemsetAsync(s->sc.stats_part, 0, (size_t)s->stats_blocks * kStatsPerBlock * 8, st);
  if (e == hipSuccess) e = hipMemsetAsync(s->sc.stats_part2, 0, (size_t)kListBlocks * kStatsPerBlock * 8, st);
  if (e == hipSuccess && s->win.cnt) e = hipMemsetAsync(s->win.cnt, 0, kWinCounters * sizeof(uint64_t), st);
  if (e == hipSuccess && s->win.cpart)
    e = hipMemsetAsync(s->win.cpart, 0, (size_t)kListBlocks * 2 * sizeof(uint64_t), st);
  return (int32_t)e;
}

int32_t axc_shard_window_stats(const ShardVramPtrs* vram, uint64_t* out) {
  if (!vram || !out) return (int32_t)hipErrorInvalidValue;
  Shard* s = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_shards.find(vram->soma_voltage);
    if (it != g_shards.end()) s = it->second;
  }
  if (!s) return (int32_t)hipErrorInvalidValue;
  for (uint32_t k = 0; k < kWinCounters; k++) out[k] = 0;
  if (!s->win.cnt) return 0;
  hipError_t e = hipMemcpy(out, s->win.cnt, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost);
  std::vector<uint64_t> part((size_t)kListBlocks * 2);
  if (e == hipSuccess) e = hipMemcpy(part.data(), s->win.cpart, part.size() * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return (int32_t)e;
  for (size_t k = 0; k < kListBlocks; k++) {
    out[3] += part[2 * k];
    out[4] += part[2 * k + 1];
  }
  return 0;
}

int32_t axc_shard_stats_read(const ShardVramPtrs* vram, AxcShardStats* out) {
  if (!vram || !out) return (int32_t)hipErrorInvalidValue;
  Shard* s = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_shards.find(vram->soma_voltage);
    if (it != g_shards.end()) s = it->second;
  }
  if (!s) return (int32_t)hipErrorInvalidValue;
  std::vector<uint64_t> a((size_t)s->stats_blocks * kStatsPerBlock), b((size_t)kListBlocks * kStatsPerBlock);
  hipError_t e = hipMemcpy(a.data(), s->sc.stats_part, a.size() * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(b.data(), s->sc.stats_part2, b.size() * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return (int32_t)e;
  AxcShardStats r{};
  auto add = [&](const std::vector<uint64_t>& v, size_t nb) {
    for (size_t k = 0; k < nb; k++) {
      const uint64_t* q = &v[k * kStatsPerBlock];
      r.integrated += q[kStIntegrated];
      r.slots_scanned += q[kStScanned];
      r.targets_read += q[kStTargets];
      r.hits += q[kStHits];
      r.spikes += q[kStSpikes];
      r.gsop_slots += q[kStGsopSlots];
      r.gsop_targets += q[kStGsopTargets];
      r.ticks += q[kStTicks];
    }
  };
  add(a, s->stats_blocks);
  for (size_t k = 0; k < s->stats_blocks; k++) {
    r.gsop_slots_fused += a[k * kStatsPerBlock + kStGsopSlots];
    r.gsop_targets_fused += a[k * kStatsPerBlock + kStGsopTargets];
  }
  add(b, kListBlocks);
  *out = r;
  return 0;
}

int32_t axc_step_day_phase_batch(const ShardVramPtrs* vram, uint32_t padded_n, uint32_t total_axons, uint32_t v_seg,
                                 uint32_t tick_base, uint32_t T, const uint32_t* d_bitmask, uint32_t words,
                                 uint32_t virt_off, uint32_t n_virt, const uint32_t* d_spikes, uint32_t max_spk,
                                 const uint32_t* h_counts, const uint32_t* d_mapped, uint8_t* d_out, uint32_t n_out,
                                 int16_t dopamine, void* stream) {
  if (!vram) return (int32_t)hipErrorInvalidValue;
  if (n_virt > 0 && d_bitmask && (uint64_t)virt_off + n_virt > (uint64_t)total_axons)
    return (int32_t)hipErrorInvalidValue;
  if (T == 0) return 0;
  // A tick's count beyond its slice (or counts without a schedule) would read
  // past the schedule buffer: refused.
  if (h_counts)
    for (uint32_t t = 0; t < T; t++)
      if (h_counts[t] > (d_spikes ? max_spk : 0u)) return (int32_t)hipErrorInvalidValue;
  hipError_t e;
  Shard* s = shard_for(vram, padded_n, total_axons, &e);
  if (!s) return (int32_t)e;
  const AxcVariantParameters* lut = device_lut();
  if (!lut) return (int32_t)hipErrorOutOfMemory;
  hipStream_t st = (hipStream_t)stream;

  GraphKey key;
  key.T = T;
  key.words = words;
  key.virt_off = virt_off;
  key.n_virt = n_virt;
  key.max_spk = max_spk;
  key.n_out = n_out;
  key.has_bitmask = d_bitmask != nullptr;
  key.has_spikes = d_spikes != nullptr;
  key.has_out = d_out != nullptr && d_mapped != nullptr && n_out > 0;
  key.mapped = d_mapped;
  key.heads = vram->axon_heads;
  key.v_seg = v_seg;
  key.staged = ensure_stage(s, *vram, st, &e, virt_off, d_bitmask ? n_virt : 0u);
  if (e != hipSuccess) return (int32_t)e;
  // In the engine's soma order the ticks run on the private arrays, copied
  // in from the caller's at the start and out at the end.
  key.ordered = key.staged && s->ordered();
  if (key.ordered && order_moves_inputs(s, virt_off, d_bitmask ? n_virt : 0u)) return (int32_t)hipErrorInvalidValue;
  const ShardVramPtrs& V = key.ordered ? s->shadow : *vram;
  // The direct pass applies tick t's GSOP inside tick t+1's soma pass; with
  // the staged pass every tick ends with the GSOP kernel, on the plan's
  // working copy of the weights, loaded before the first tick and stored
  // back after the last.
  const bool fuse = !key.staged;
  const bool pw = key.staged;
  // Staged batches stream the slots once per window of key.win ticks.
  key.win = (key.staged && padded_n > 0) ? window_for(v_seg) : 1u;
  key.rev = nullptr;
  if (key.win > 1) {
    if ((e = ensure_window(s, vram->soma_flags, st)) != hipSuccess) return (int32_t)e;
    key.rev = s->win.rev;
  }
  std::vector<uint32_t> counts(T, 0);
  if (d_spikes && h_counts) counts.assign(h_counts, h_counts + T);

  // The per-batch burst reset of shard.rs:71 (idempotent with the per-tick one).
  reset_burst_kernel<<<blocks_for(padded_n ? padded_n : 1), kThreads, 0, st>>>(vram->soma_flags, padded_n);
  bool timed;
  {
    std::lock_guard<std::mutex> lk(g_tmu);
    timed = g_timing.enabled;
  }
  // The batch's device parameters (the graphs, and windowed batches also
  // when timed): tick base, dopamine, slices, per-tick schedule counts.
  auto upload_params = [&]() -> hipError_t {
    TickParams pv;
    pv.tick_base = tick_base;
    pv.dopamine = (int32_t)dopamine;
    pv.bitmask = d_bitmask;
    pv.spikes = d_spikes;
    pv.out = d_out;
    pv.counts = nullptr;
    hipError_t r;
    if (key.has_spikes) {
      // The counts go to the shard's device copy, read by the graph's K3
      // kernels (stream order keeps the previous replay's reads before this).
      if (s->counts_cap < T) {
        (void)hipStreamSynchronize(st);
        (void)hipFree(s->d_counts);
        s->d_counts = nullptr;
        s->counts_cap = 0;
        if ((r = hipMalloc((void**)&s->d_counts, (size_t)T * 4)) != hipSuccess) return r;
        s->counts_cap = T;
      }
      if ((r = hipMemcpyAsync(s->d_counts, counts.data(), (size_t)T * 4, hipMemcpyHostToDevice, st)) != hipSuccess)
        return r;
      pv.counts = s->d_counts;
    }
    set_params_kernel<<<1, 1, 0, st>>>(s->prm, pv);
    return hipGetLastError();
  };
  if (timed) {
    // Kernel by kernel, so that each launch can be timed: the same kernels
    // and tick structure as the graph below.
    if (key.win > 1 && (e = upload_params()) != hipSuccess) return (int32_t)e;
    if (key.ordered) order_copy(s, *vram, true, st);
    if (pw) {
      TimedScope ts(kKWeights, st, true);
      copy_weights_kernel<true><<<blocks_for(padded_n), kThreads, 0, st>>>(vram->dendrite_weights, padded_n, s->sp);
    }
    if (key.win > 1) {
      e = enqueue_windows(V, s, v_seg, T, key.win, words, virt_off, n_virt, key.has_bitmask, key.has_spikes, max_spk,
                          d_mapped, key.has_out, n_out, s->prm, st, lut, true);
    } else {
      for (uint32_t t = 0; t < T && e == hipSuccess; t++) {
        e = enqueue_tick(V, s, v_seg, tick_base + t, d_bitmask ? d_bitmask + (size_t)t * words : nullptr, words,
                         virt_off, n_virt, d_spikes ? d_spikes + (size_t)t * max_spk : nullptr, max_spk,
                         counts[t], d_mapped, d_out ? d_out + (size_t)t * n_out : nullptr, n_out,
                         (int32_t)dopamine, nullptr, st, lut, t, t > 0 && fuse, t + 1 == T || !fuse, key.staged, pw);
      }
    }
    if (pw && e == hipSuccess) {
      TimedScope ts(kKWeights, st, true);
      copy_weights_kernel<false><<<blocks_for(padded_n), kThreads, 0, st>>>(vram->dendrite_weights, padded_n, s->sp);
      e = hipGetLastError();
    }
    if (key.ordered) order_copy(s, *vram, false, st);
    if (e == hipSuccess) e = hipGetLastError();
    return (int32_t)e;
  }
  if ((e = upload_params()) != hipSuccess) return (int32_t)e;

  if (!(s->has_graph && s->gkey == key)) {
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->gexec = nullptr;
    s->graph = nullptr;
    s->has_graph = false;
    hipStream_t cap;
    if ((e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking)) != hipSuccess) return (int32_t)e;
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      if (key.ordered) order_copy(s, *vram, true, cap);
      if (pw)
        copy_weights_kernel<true><<<blocks_for(padded_n), kThreads, 0, cap>>>(vram->dendrite_weights, padded_n,
                                                                               s->sp);
      if (key.win > 1) {
        e = enqueue_windows(V, s, v_seg, T, key.win, words, virt_off, n_virt, key.has_bitmask, key.has_spikes,
                            max_spk, d_mapped, key.has_out, n_out, s->prm, cap, lut, false);
      } else {
        for (uint32_t t = 0; t < T && e == hipSuccess; t++) {
          // GSOP of tick t runs inside tick t+1's neuron pass; the last tick's
          // GSOP is launched on its own so the batch leaves a complete state.
          e = enqueue_tick(V, s, v_seg, t, d_bitmask, words, virt_off, n_virt, d_spikes, max_spk,
                           0, d_mapped, d_out, n_out, 0, s->prm, cap, lut, t, t > 0 && fuse,
                           t + 1 == T || !fuse, key.staged, pw);
        }
      }
      if (pw && e == hipSuccess)
        copy_weights_kernel<false><<<blocks_for(padded_n), kThreads, 0, cap>>>(vram->dendrite_weights, padded_n,
                                                                                s->sp);
      if (key.ordered && e == hipSuccess) order_copy(s, *vram, false, cap);
      hipGraph_t g = nullptr;
      hipError_t e2 = hipStreamEndCapture(cap, &g);
      if (e == hipSuccess) e = e2;
      if (e == hipSuccess) {
        e = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
        s->graph = g;
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    (void)hipStreamDestroy(cap);
    if (e != hipSuccess) return (int32_t)e;
    s->gkey = key;
    s->has_graph = true;
    s->graph_captures++;
  }
  e = hipGraphLaunch(s->gexec, st);
  return (int32_t)e;
}

void axc_host_sort_and_prune(uint8_t* flags, uint32_t* targets, int16_t* weights, uint8_t* timers, uint32_t padded_n,
                             int16_t prune) {
  const size_t N = padded_n;
  auto work = [&](size_t lo, size_t hi) {
    uint32_t key[kMaxDendrites], t[kMaxDendrites], idx[kMaxDendrites];
    int16_t w[kMaxDendrites];
    uint8_t m[kMaxDendrites];
    for (size_t i = lo; i < hi; i++) {
      flags[i] &= 0xF1;
      uint32_t nf = 0, ne = 0;
      uint32_t empties[kMaxDendrites];
      for (uint32_t s = 0; s < kMaxDendrites; s++) {
        size_t col = (size_t)s * N + i;
        t[s] = targets[col];
        w[s] = weights[col];
        m[s] = timers[col];
        int32_t a = w[s] < 0 ? -(int32_t)w[s] : w[s];
        key[s] = (uint32_t)a;
        if (t[s]) idx[nf++] = s; else empties[ne++] = s;
      }
      // Filled slots by descending |w|, ties in slot order; empty slots after
      // them in slot order: the outcome of the reference insertion sort.
      std::stable_sort(idx, idx + nf, [&](uint32_t x, uint32_t y) { return key[x] > key[y]; });
      for (uint32_t e = 0; e < ne; e++) idx[nf + e] = empties[e];
      for (uint32_t r = 0; r < kMaxDendrites; r++) {
        uint32_t s = idx[r];
        bool pr = t[s] != 0 && (int32_t)key[s] < (int32_t)prune;
        size_t col = (size_t)r * N + i;
        targets[col] = pr ? 0u : t[s];
        weights[col] = pr ? (int16_t)0 : w[s];
        timers[col] = pr ? (uint8_t)0 : m[s];
      }
    }
  };
  unsigned nt = std::thread::hardware_concurrency();
  nt = nt == 0 ? 1 : (nt > 16 ? 16 : nt);
  if (N < 4096) nt = 1;
  std::vector<std::thread> th;
  size_t chunk = (N + nt - 1) / nt;
  for (unsigned k = 0; k < nt; k++) {
    size_t lo = k * chunk, hi = lo + chunk < N ? lo + chunk : N;
    if (lo >= hi) break;
    th.emplace_back(work, lo, hi);
  }
  for (auto& x : th) x.join();
}

void axc_staging_enable(int enable) { g_staging = enable != 0; }

int32_t axc_set_window_ticks(uint32_t ticks) {
  if (ticks > kWin) return (int32_t)hipErrorInvalidValue;
  g_window = ticks < 1 ? 1u : ticks;
  return 0;
}
uint32_t axc_window_ticks(void) { return g_window.load(); }

int32_t axc_shard_set_soma_positions(const ShardVramPtrs* vram, const uint32_t* positions, uint32_t n) {
  if (!vram) return (int32_t)hipErrorInvalidValue;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_shards.find(vram->soma_voltage);
  if (it == g_shards.end()) return (int32_t)hipErrorInvalidValue;
  Shard* s = it->second;
  if (positions && n != s->N) return (int32_t)hipErrorInvalidValue;
  (void)hipDeviceSynchronize();  // no step of the shard may be in flight with the old order
  s->order.clear();
  if (positions) {
    // 8x8x8-voxel bricks, z-major over bricks and x-fastest inside one (the
    // caller's index breaks ties); somas at position 0 (padding) last.
    std::vector<std::pair<uint64_t, uint32_t>> key(n);
    for (uint32_t i = 0; i < n; i++) {
      const uint32_t p = positions[i];
      const uint64_t x = p & 0x3FF, y = (p >> 10) & 0x3FF, z = (p >> 20) & 0xFF;
      const uint64_t brick = ((z >> 3) << 14) | ((y >> 3) << 7) | (x >> 3);
      const uint64_t inner = ((z & 7) << 6) | ((y & 7) << 3) | (x & 7);
      key[i] = {p ? (brick << 9) | inner : ~0ull, i};
    }
    std::stable_sort(key.begin(), key.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    s->order.resize(n);
    bool identity = true;
    for (uint32_t k = 0; k < n; k++) {
      s->order[k] = key[k].second;
      identity = identity && key[k].second == k;
    }
    if (identity) s->order.clear();
  }
  // The plan (and a cached graph, whose kernels hold the arrays) follow the new order.
  s->stage_dirty = true;
  s->targets_version++;
  if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  s->gexec = nullptr;
  s->graph = nullptr;
  s->has_graph = false;
  return 0;
}

int32_t axc_shard_graph_captures(const ShardVramPtrs* vram, uint64_t* out) {
  if (!vram || !out) return (int32_t)hipErrorInvalidValue;
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_shards.find(vram->soma_voltage);
  if (it == g_shards.end()) return (int32_t)hipErrorInvalidValue;
  *out = it->second->graph_captures;
  return 0;
}

void axc_shard_targets_changed(const ShardVramPtrs* vram) {
  if (vram) mark_stage_dirty(vram->soma_voltage);
}

int32_t axc_shard_write_slots(const ShardVramPtrs* vram, uint32_t padded_n, const uint32_t* d_cols,
                              const uint32_t* d_targets, const int16_t* d_weights, uint32_t n, void* stream) {
  if (!vram || !vram->dendrite_targets || !vram->dendrite_weights || (n && (!d_cols || !d_targets || !d_weights)))
    return (int32_t)hipErrorInvalidValue;
  hipStream_t st = (hipStream_t)stream;
  if (n) {
    write_slots_kernel<<<std::min<uint32_t>(blocks_for(n), 1024u), kThreads, 0, st>>>(
        vram->dendrite_targets, vram->dendrite_weights, (uint64_t)kMaxDendrites * padded_n, d_cols, d_targets,
        d_weights, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int32_t)e;
  }
  // As an asynchronous copy into the targets: the plan is rebuilt at the next
  // step, after this stream's write.
  note_device_write(vram->dendrite_targets, (size_t)kMaxDendrites * padded_n * 4, true, st);
  return 0;
}

void axc_kernel_timing_enable(int enable) {
  std::lock_guard<std::mutex> lk(g_tmu);
  g_timing.enabled = enable != 0;
}

int32_t axc_kernel_timing_read(double* ms_out, uint64_t* launches_out, int reset) {
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pend;
  {
    std::lock_guard<std::mutex> lk(g_tmu);
    pend.swap(g_timing.pending);
  }
  hipError_t err = hipSuccess;
  for (auto& p : pend) {
    hipError_t e = hipEventSynchronize(p.second.second);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.second.first, p.second.second);
    if (e != hipSuccess) err = e;
    std::lock_guard<std::mutex> lk(g_tmu);
    g_timing.ms[p.first] += ms;
    g_timing.launches[p.first] += 1;
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  std::lock_guard<std::mutex> lk(g_tmu);
  for (int k = 0; k < kNumTimed; k++) {
    if (ms_out) ms_out[k] = g_timing.ms[k];
    if (launches_out) launches_out[k] = g_timing.launches[k];
    if (reset) {
      g_timing.ms[k] = 0;
      g_timing.launches[k] = 0;
    }
  }
  return (int32_t)err;
}

const char* axc_version(void) { return "axicor_amd day-phase engine 0.1 (gfx950)"; }

}  // extern "C"

// ===========================================================================
// Batched brains (no counterpart in the reference)
// ===========================================================================
namespace {

struct Batch {
  int device = 0;
  std::vector<AxcBrainDesc> desc;
  std::vector<BrainDev> host_dev;
  std::vector<Shard*> shards;          // per-brain workspaces (not in the shard registry)
  BrainDev* d_brains = nullptr;
  AxcVariantParameters* d_luts = nullptr;
  uint64_t* d_stats = nullptr;         // [n][kStatsPerBlock]
  uint64_t* d_type_spikes = nullptr;   // [n][16]
  unsigned long long* d_syn = nullptr; // [n][kSynStats] synapse statistics
  int16_t* d_prune = nullptr;          // [n] prune thresholds of the last axc_batch_synapse_stats
  WorkItem* d_axon_items = nullptr;
  WorkItem* d_soma_items = nullptr;
  WorkItem* d_stage_items = nullptr;  // (brain, first soma) of each kBlockSomas block, for the staged pass
  uint32_t n_axon_items = 0, n_soma_items = 0, n_stage_items = 0;
  // The staged pass's dispatch order (AxcBatchTuning::stage_order): the
  // stage items themselves for 0, else a balanced list of 8 equal shares
  // padded with kNoBrain entries (capacity 8 * n_stage_items).
  WorkItem* d_stage_order = nullptr;
  uint32_t n_stage_order = 0, order_of = 0xFFFFFFFFu;
  BatchTick* d_tick = nullptr;
  AxcBatchTuning tune{0, 0, 0, 0};
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  uint32_t graph_T = 0;
  AxcBatchTuning graph_tune{0, 0, 0, 0};
  // Staged soma pass: each brain's plan lives in its workspace (shards[k]);
  // built_version[k] is the targets version it was built from.
  bool staged = false;
  uint32_t* d_prefix = nullptr;  // [2][n+1] spike- and owner-count prefixes of the current tick
  std::vector<uint64_t> built_version;
};

// The staged pass's dispatch order for stage_order k (see AxcBatchTuning):
// k = 0 is the stage items in brain order; otherwise every brain's blocks
// are cut into runs of k, and the runs, heaviest first, go to the XCD share
// with the least work so far (longest-processing-time), each share keeping
// its runs in that order; shares are padded to equal length.
hipError_t build_stage_order(Batch* b, uint32_t k, hipStream_t st) {
  if (b->order_of == k) return hipSuccess;
  if (k == 0) {
    b->n_stage_order = b->n_stage_items;
    hipError_t e = hipMemcpyAsync(b->d_stage_order, b->d_stage_items, (size_t)b->n_stage_items * sizeof(WorkItem),
                                  hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) b->order_of = 0;
    return e;
  }
  const uint32_t n = b->n_stage_items;
  uint32_t* d_cost = nullptr;
  hipError_t e = hipMalloc((void**)&d_cost, (size_t)n * 4);
  if (e != hipSuccess) return e;
  batch_stage_cost_kernel<<<n, kBlockSomas, 0, st>>>(b->d_brains, b->d_stage_items, d_cost);
  std::vector<uint32_t> cost(n);
  std::vector<WorkItem> items(n);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(cost.data(), d_cost, (size_t)n * 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(items.data(), b->d_stage_items, (size_t)n * sizeof(WorkItem), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d_cost);
  if (e != hipSuccess) return e;
  // Runs: [first item, end) within one brain, at most k items.
  struct Run { uint32_t first, end; uint64_t w; };
  std::vector<Run> runs;
  for (uint32_t q = 0; q < n;) {
    uint32_t r = q;
    uint64_t w = 0;
    while (r < n && r - q < k && items[r].brain == items[q].brain) w += cost[r++];
    runs.push_back(Run{q, r, w});
    q = r;
  }
  std::stable_sort(runs.begin(), runs.end(), [](const Run& x, const Run& y) { return x.w > y.w; });
  std::vector<std::vector<WorkItem>> share(kNumXcd);
  std::vector<uint64_t> load(kNumXcd, 0);
  for (const Run& r : runs) {
    const uint32_t x = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
    load[x] += r.w;
    for (uint32_t q = r.first; q < r.end; q++) share[x].push_back(items[q]);
  }
  size_t per = 0;
  for (const auto& v : share) per = std::max(per, v.size());
  std::vector<WorkItem> order(per * kNumXcd, WorkItem{kNoBrain, 0});
  for (uint32_t x = 0; x < kNumXcd; x++) std::copy(share[x].begin(), share[x].end(), order.begin() + x * per);
  e = hipMemcpy(b->d_stage_order, order.data(), order.size() * sizeof(WorkItem), hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    b->n_stage_order = (uint32_t)order.size();
    b->order_of = k;
  }
  return e;
}

hipError_t batch_capture(Batch* b, uint32_t T) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
  hipStream_t cap;
  hipError_t e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking);
  if (e != hipSuccess) return e;
  const uint32_t n = (uint32_t)b->desc.size();
  e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    if (b->staged) {
      // Per tick: axons, staged somas, pushes, GSOP after the pushes, on the
      // brains' working copies of the weights (loaded first, stored last).
      const size_t lds = (kStageCap + 1) * sizeof(uint2);
      batch_copy_weights_kernel<true><<<b->tune.soma_blocks, kThreads, 0, cap>>>(b->d_brains, b->d_soma_items,
                                                                                 b->n_soma_items);
      for (uint32_t t = 0; t < T; t++) {
        batch_axon_kernel<<<b->tune.axon_blocks, kThreads, 0, cap>>>(b->d_brains, b->d_axon_items, b->n_axon_items,
                                                                     b->d_tick, t, t);
        batch_staged_kernel<<<(b->n_stage_order + kNumXcd - 1) / kNumXcd * kNumXcd, kStageThreads, lds, cap>>>(
            b->d_brains, b->d_stage_order, b->n_stage_order, b->d_tick, t, t);
        batch_compact_spikes_kernel<<<blocks_for((uint64_t)b->n_stage_items * kItemWords), kThreads, 0, cap>>>(
            b->d_brains, b->d_stage_items, b->n_stage_items);
        batch_spike_prefix_kernel<<<1, 1024, 0, cap>>>(b->d_brains, n, b->d_prefix, kCntSpikes);
        batch_push_flat_kernel<<<b->tune.list_blocks, kThreads, 0, cap>>>(b->d_brains, n, b->d_prefix, t);
        batch_gsop_chunks_kernel<<<b->tune.list_blocks * 8, kThreads, 0, cap>>>(b->d_brains, n, b->d_prefix, t);
      }
      batch_copy_weights_kernel<false><<<b->tune.soma_blocks, kThreads, 0, cap>>>(b->d_brains, b->d_soma_items,
                                                                                  b->n_soma_items);
    } else {
    for (uint32_t t = 0; t < T; t++) {
      batch_axon_kernel<<<b->tune.axon_blocks, kThreads, 0, cap>>>(b->d_brains, b->d_axon_items, b->n_axon_items,
                                                                   b->d_tick, t, t);
      if (t == 0)
        batch_neuron_kernel<false><<<b->tune.soma_blocks, kThreads, 0, cap>>>(b->d_brains, b->d_soma_items,
                                                                              b->n_soma_items, b->d_tick, t, t);
      else
        batch_neuron_kernel<true><<<b->tune.soma_blocks, kThreads, 0, cap>>>(b->d_brains, b->d_soma_items,
                                                                             b->n_soma_items, b->d_tick, t, t);
      batch_push_kernel<<<b->tune.list_blocks, kThreads, 0, cap>>>(b->d_brains, n, t);
    }
    batch_gsop_kernel<<<b->tune.list_blocks, kThreads, 0, cap>>>(b->d_brains, n, T - 1);
    }
    e = hipGetLastError();
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(cap, &g);
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) {
      e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
      b->graph = g;
    } else if (g) {
      (void)hipGraphDestroy(g);
    }
  }
  (void)hipStreamDestroy(cap);
  if (e == hipSuccess) {
    b->graph_T = T;
    b->graph_tune = b->tune;
  }
  return e;
}

hipError_t batch_run(Batch* b, uint32_t tick_base, uint32_t T, hipStream_t st) {
  if (!(b->gexec && b->graph_T == T && b->graph_tune.axon_blocks == b->tune.axon_blocks &&
        b->graph_tune.soma_blocks == b->tune.soma_blocks && b->graph_tune.list_blocks == b->tune.list_blocks &&
        b->graph_tune.stage_order == b->tune.stage_order)) {
    hipError_t e = b->staged ? build_stage_order(b, b->tune.stage_order, st) : hipSuccess;
    if (e == hipSuccess) e = batch_capture(b, T);
    if (e != hipSuccess) return e;
  }
  batch_set_tick_kernel<<<1, 1, 0, st>>>(b->d_tick, tick_base);
  return hipGraphLaunch(b->gexec, st);
}

void batch_free(Batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  for (Shard* s : b->shards) destroy_shard(s);
  (void)hipFree(b->d_brains);
  (void)hipFree(b->d_luts);
  (void)hipFree(b->d_stats);
  (void)hipFree(b->d_type_spikes);
  (void)hipFree(b->d_syn);
  (void)hipFree(b->d_prune);
  (void)hipFree(b->d_axon_items);
  (void)hipFree(b->d_soma_items);
  (void)hipFree(b->d_stage_items);
  (void)hipFree(b->d_stage_order);
  (void)hipFree(b->d_prefix);
  (void)hipFree(b->d_tick);
  delete b;
}

}  // namespace

extern "C" {

int32_t axc_batch_create(const AxcBrainDesc* brains, uint32_t n, void** out) {
  if (!brains || n == 0 || !out) return (int32_t)hipErrorInvalidValue;
  Batch* b = new Batch();
  (void)hipGetDevice(&b->device);
  b->desc.assign(brains, brains + n);
  hipError_t e = hipSuccess;
  std::vector<WorkItem> ax, so, sg;
  for (uint32_t k = 0; k < n && e == hipSuccess; k++) {
    const AxcBrainDesc& d = brains[k];
    if (d.num_virtual_axons && (uint64_t)d.virtual_offset + d.num_virtual_axons > d.total_axons) {
      e = hipErrorInvalidValue;
      break;
    }
    if (d.d_input_bitmask && d.input_ticks == 0) {
      e = hipErrorInvalidValue;
      break;
    }
    Shard* s = new Shard();
    s->N = d.padded_n;
    s->A = d.total_axons;
    s->device = b->device;
    b->shards.push_back(s);
    if ((e = alloc_scratch(s)) != hipSuccess) break;
    // Staging plan of this brain, built now from its current targets.
    if (g_staging.load() && d.padded_n > 0) {
      ensure_stage(s, d.vram, nullptr, &e);
      if (e != hipSuccess) break;
      b->staged = true;
    }
    b->built_version.push_back(targets_version_of(d.vram.soma_voltage));
    for (uint32_t a = 0; a < d.total_axons; a += kThreads) ax.push_back(WorkItem{k, a});
    for (uint32_t i = 0; i < d.padded_n; i += kThreads) so.push_back(WorkItem{k, i});
    for (uint32_t i = 0; i < d.padded_n; i += kBlockSomas) sg.push_back(WorkItem{k, i});
  }
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_luts, (size_t)n * 16 * sizeof(AxcVariantParameters));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_stats, (size_t)n * kStatsPerBlock * 8);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_type_spikes, (size_t)n * 16 * 8);
  if (e == hipSuccess) e = hipMemset(b->d_stats, 0, (size_t)n * kStatsPerBlock * 8);
  if (e == hipSuccess) e = hipMemset(b->d_type_spikes, 0, (size_t)n * 16 * 8);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_tick, sizeof(BatchTick));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_syn, (size_t)n * kSynStats * 8);
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_prune, (size_t)n * sizeof(int16_t));
  if (e == hipSuccess) {
    b->n_axon_items = (uint32_t)ax.size();
    b->n_soma_items = (uint32_t)so.size();
    e = hipMalloc((void**)&b->d_axon_items, ax.size() * sizeof(WorkItem));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_soma_items, so.size() * sizeof(WorkItem));
    if (e == hipSuccess) e = hipMemcpy(b->d_axon_items, ax.data(), ax.size() * sizeof(WorkItem), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->d_soma_items, so.data(), so.size() * sizeof(WorkItem), hipMemcpyHostToDevice);
    b->n_stage_items = (uint32_t)sg.size();
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_stage_items, sg.size() * sizeof(WorkItem));
    if (e == hipSuccess)
      e = hipMemcpy(b->d_stage_items, sg.data(), sg.size() * sizeof(WorkItem), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_stage_order, kNumXcd * (sg.size() + 1) * sizeof(WorkItem));
  }
  if (e == hipSuccess) {
    b->host_dev.resize(n);
    for (uint32_t k = 0; k < n && e == hipSuccess; k++) {
      const AxcBrainDesc& d = brains[k];
      e = hipMemcpy(b->d_luts + (size_t)k * 16, d.lut, 16 * sizeof(AxcVariantParameters), hipMemcpyHostToDevice);
      BrainDev& h = b->host_dev[k];
      h.vram = d.vram;
      h.N = d.padded_n;
      h.A = d.total_axons;
      h.v_seg = d.v_seg;
      h.virt_off = d.virtual_offset;
      h.n_virt = d.num_virtual_axons;
      h.words = d.input_words;
      h.input_ticks = d.input_ticks ? d.input_ticks : 1;
      h.dop = d.dopamine;
      h.lut = b->d_luts + (size_t)k * 16;
      h.bitmask = d.d_input_bitmask;
      h.sc = b->shards[k]->sc;
      h.stats = b->d_stats + (size_t)k * kStatsPerBlock;
      h.type_spikes = b->d_type_spikes + (size_t)k * 16;
      h.sp = b->shards[k]->sp;
    }
  }
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_prefix, (size_t)2 * (n + 1) * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc((void**)&b->d_brains, n * sizeof(BrainDev));
  if (e == hipSuccess) e = hipMemcpy(b->d_brains, b->host_dev.data(), n * sizeof(BrainDev), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    batch_free(b);
    return (int32_t)e;
  }
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  b->tune.axon_blocks = (uint32_t)cus * 8;
  b->tune.soma_blocks = (uint32_t)cus * 4;
  b->tune.list_blocks = (uint32_t)cus * 2;
  *out = b;
  return 0;
}

int32_t axc_batch_step(void* h, uint32_t tick_base, uint32_t ticks, void* stream) {
  if (!h || ticks == 0) return (int32_t)hipErrorInvalidValue;
  Batch* b = (Batch*)h;
  hipStream_t st = (hipStream_t)stream;
  if (b->staged) {
    // Rebuild the plans of brains whose targets changed since (Night uploads,
    // sort&prune, state uploads).
    for (size_t k = 0; k < b->desc.size(); k++) {
      const AxcBrainDesc& d = b->desc[k];
      const uint64_t v = targets_version_of(d.vram.soma_voltage);
      if (v == b->built_version[k] || d.padded_n == 0) continue;
      stage_build_kernel<<<stage_blocks(d.padded_n), kBlockSomas, 0, st>>>(d.vram.dendrite_targets, d.padded_n,
                                                                      d.total_axons, b->shards[k]->sp);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return (int32_t)e;
      b->built_version[k] = v;
    }
  }
  return (int32_t)batch_run(b, tick_base, ticks, st);
}

int32_t axc_batch_set_tuning(void* h, const AxcBatchTuning* t) {
  if (!h || !t || t->axon_blocks == 0 || t->list_blocks == 0 || t->soma_blocks == 0 ||
      t->soma_blocks % kNumXcd != 0 || t->stage_order > (1u << 20))
    return (int32_t)hipErrorInvalidValue;
  ((Batch*)h)->tune = *t;
  return 0;
}

int32_t axc_batch_get_tuning(void* h, AxcBatchTuning* t) {
  if (!h || !t) return (int32_t)hipErrorInvalidValue;
  *t = ((Batch*)h)->tune;
  return 0;
}

int32_t axc_batch_autotune(void* h, uint32_t ticks, void* stream, AxcBatchTuning* best_out) {
  if (!h || ticks == 0) return (int32_t)hipErrorInvalidValue;
  Batch* b = (Batch*)h;
  hipStream_t st = (hipStream_t)stream;
  const uint32_t n = (uint32_t)b->desc.size();
  // Snapshot every brain (state, heads) so that tuning leaves no trace.
  std::vector<void*> bak(2 * n, nullptr);
  hipError_t e = hipSuccess;
  for (uint32_t k = 0; k < n && e == hipSuccess; k++) {
    const AxcBrainDesc& d = b->desc[k];
    const size_t sb = (size_t)d.padded_n * AXC_STATE_BYTES_PER_NEURON, hb = (size_t)d.total_axons * 32;
    e = hipMalloc(&bak[2 * k], sb);
    if (e == hipSuccess) e = hipMalloc(&bak[2 * k + 1], hb);
    if (e == hipSuccess) e = hipMemcpyAsync(bak[2 * k], d.vram.soma_voltage, sb, hipMemcpyDeviceToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(bak[2 * k + 1], d.vram.axon_heads, hb, hipMemcpyDeviceToDevice, st);
  }
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  const AxcBatchTuning base = b->tune;
  AxcBatchTuning best = base;
  float best_ms = 1e30f;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  // Candidates: on the staged path the soma pass's dispatch order (brain
  // order, or balanced runs of 4, 32 or 256 blocks), the list grid (pushes,
  // GSOP) and the axon grid; on the direct path the soma, axon and list
  // grids.
  std::vector<AxcBatchTuning> cand;
  if (b->staged) {
    for (uint32_t so : {0u, 4u, 32u, 256u})
      for (uint32_t lm : {1u, 2u, 4u})
        for (uint32_t am : {4u, 16u}) cand.push_back(AxcBatchTuning{(uint32_t)cus * am, base.soma_blocks, (uint32_t)cus * lm, so});
  } else {
    for (uint32_t sm : {2u, 4u, 8u, 16u})
      for (uint32_t lm : {1u, 2u, 4u})
        for (uint32_t am : {4u, 16u})
          cand.push_back(AxcBatchTuning{(uint32_t)cus * am, (uint32_t)cus * sm, (uint32_t)cus * lm, 0u});
  }
  for (const AxcBatchTuning& c : cand) {
    {
      if (e != hipSuccess) break;
      b->tune = c;
      e = batch_run(b, 0, ticks, st);  // capture + warm
      if (e == hipSuccess) (void)hipEventRecord(e0, st);
      if (e == hipSuccess) e = batch_run(b, ticks, ticks, st);
      if (e == hipSuccess) (void)hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0;
      if (e == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
      if (e == hipSuccess && ms < best_ms) {
        best_ms = ms;
        best = b->tune;
      }
      // Restore before the next candidate.
      for (uint32_t k = 0; k < n && e == hipSuccess; k++) {
        const AxcBrainDesc& d = b->desc[k];
        const size_t sb = (size_t)d.padded_n * AXC_STATE_BYTES_PER_NEURON, hb = (size_t)d.total_axons * 32;
        e = hipMemcpyAsync(d.vram.soma_voltage, bak[2 * k], sb, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d.vram.axon_heads, bak[2 * k + 1], hb, hipMemcpyDeviceToDevice, st);
      }
    }
  }
  // The snapshot buffers follow the restored state; counters restart at zero.
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  for (void* p : bak) (void)hipFree(p);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e == hipSuccess) {
    e = hipMemset(b->d_stats, 0, (size_t)n * kStatsPerBlock * 8);
    if (e == hipSuccess) e = hipMemset(b->d_type_spikes, 0, (size_t)n * 16 * 8);
  }
  b->tune = best;
  if (best_out) *best_out = best;
  return (int32_t)e;
}

int32_t axc_batch_synapse_stats(void* h, const int16_t* prune_thresholds, AxcSynapseStats* out, void* stream) {
  if (!h || !out) return (int32_t)hipErrorInvalidValue;
  Batch* b = (Batch*)h;
  const size_t n = b->desc.size();
  hipStream_t st = (hipStream_t)stream;
  std::vector<int16_t> pr(n, 0);
  if (prune_thresholds) pr.assign(prune_thresholds, prune_thresholds + n);
  std::vector<unsigned long long> v(n * kSynStats);
  hipError_t e = hipMemcpyAsync(b->d_prune, pr.data(), n * sizeof(int16_t), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(b->d_syn, 0, n * kSynStats * 8, st);
  if (e == hipSuccess) {
    batch_synapse_stats_kernel<<<b->n_soma_items, kThreads, 0, st>>>(b->d_brains, b->d_soma_items, b->d_prune,
                                                                     b->d_syn);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(v.data(), b->d_syn, n * kSynStats * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return (int32_t)e;
  for (size_t k = 0; k < n; k++) {
    out[k].filled = v[k * kSynStats + 0];
    out[k].sum_abs_weight = v[k * kSynStats + 1];
    out[k].sum_sq_weight = v[k * kSynStats + 2];
    out[k].below_prune = v[k * kSynStats + 3];
    out[k].sum_threshold_offset = (int64_t)v[k * kSynStats + 4];
  }
  return 0;
}

int32_t axc_batch_stats(void* h, uint32_t brain, AxcShardStats* out, uint64_t* type_spikes16) {
  if (!h || !out) return (int32_t)hipErrorInvalidValue;
  Batch* b = (Batch*)h;
  if (brain >= b->desc.size()) return (int32_t)hipErrorInvalidValue;
  uint64_t v[kStatsPerBlock];
  hipError_t e = hipMemcpy(v, b->d_stats + (size_t)brain * kStatsPerBlock, sizeof(v), hipMemcpyDeviceToHost);
  if (e == hipSuccess && type_spikes16)
    e = hipMemcpy(type_spikes16, b->d_type_spikes + (size_t)brain * 16, 16 * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return (int32_t)e;
  out->integrated = v[kStIntegrated];
  out->slots_scanned = v[kStScanned];
  out->targets_read = v[kStTargets];
  out->hits = v[kStHits];
  out->spikes = v[kStSpikes];
  out->gsop_slots = v[kStGsopSlots];
  out->gsop_targets = v[kStGsopTargets];
  out->ticks = v[kStTicks];
  out->gsop_slots_fused = 0;
  out->gsop_targets_fused = 0;
  return 0;
}

int32_t axc_batch_update(void* h, const AxcBrainDesc* brains, uint32_t n) {
  if (!h || !brains) return (int32_t)hipErrorInvalidValue;
  Batch* b = (Batch*)h;
  if (n != b->desc.size()) return (int32_t)hipErrorInvalidValue;
  for (uint32_t k = 0; k < n; k++) {
    const AxcBrainDesc& d = brains[k];
    const AxcBrainDesc& o = b->desc[k];
    if (d.vram.soma_voltage != o.vram.soma_voltage || d.vram.axon_heads != o.vram.axon_heads ||
        d.padded_n != o.padded_n || d.total_axons != o.total_axons || !d.lut)
      return (int32_t)hipErrorInvalidValue;
    if (d.num_virtual_axons && (uint64_t)d.virtual_offset + d.num_virtual_axons > d.total_axons)
      return (int32_t)hipErrorInvalidValue;
    if (d.d_input_bitmask && d.input_ticks == 0) return (int32_t)hipErrorInvalidValue;
  }
  // The previous steps may still read the tables and descriptors.
  hipError_t e = hipDeviceSynchronize();
  std::vector<AxcVariantParameters> luts((size_t)n * 16);
  for (uint32_t k = 0; k < n && e == hipSuccess; k++) {
    const AxcBrainDesc& d = brains[k];
    b->desc[k] = d;
    std::copy(d.lut, d.lut + 16, luts.begin() + (size_t)k * 16);
    BrainDev& hd = b->host_dev[k];
    hd.v_seg = d.v_seg;
    hd.virt_off = d.virtual_offset;
    hd.n_virt = d.num_virtual_axons;
    hd.words = d.input_words;
    hd.input_ticks = d.input_ticks ? d.input_ticks : 1;
    hd.dop = d.dopamine;
    hd.bitmask = d.d_input_bitmask;
  }
  if (e == hipSuccess)
    e = hipMemcpy(b->d_luts, luts.data(), luts.size() * sizeof(AxcVariantParameters), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(b->d_brains, b->host_dev.data(), n * sizeof(BrainDev), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(b->d_stats, 0, (size_t)n * kStatsPerBlock * 8);
  if (e == hipSuccess) e = hipMemset(b->d_type_spikes, 0, (size_t)n * 16 * 8);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return (int32_t)e;
}

int32_t axc_batch_stats_all(void* h, AxcShardStats* out, uint64_t* type_spikes16) {
  if (!h || !out) return (int32_t)hipErrorInvalidValue;
  Batch* b = (Batch*)h;
  const size_t n = b->desc.size();
  std::vector<uint64_t> v(n * kStatsPerBlock);
  hipError_t e = hipMemcpy(v.data(), b->d_stats, v.size() * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess && type_spikes16) e = hipMemcpy(type_spikes16, b->d_type_spikes, n * 16 * 8, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return (int32_t)e;
  for (size_t k = 0; k < n; k++) {
    const uint64_t* q = &v[k * kStatsPerBlock];
    AxcShardStats& r = out[k];
    r.integrated = q[kStIntegrated];
    r.slots_scanned = q[kStScanned];
    r.targets_read = q[kStTargets];
    r.hits = q[kStHits];
    r.spikes = q[kStSpikes];
    r.gsop_slots = q[kStGsopSlots];
    r.gsop_targets = q[kStGsopTargets];
    r.ticks = q[kStTicks];
    r.gsop_slots_fused = 0;
    r.gsop_targets_fused = 0;
  }
  return 0;
}

void axc_batch_destroy(void* h) {
  if (!h) return;
  (void)hipDeviceSynchronize();
  batch_free((Batch*)h);
}

}  // extern "C"
