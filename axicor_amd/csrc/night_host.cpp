// night_host.cpp — the Night-Phase sprouting pass on host cores.
//
// Restates genesis-baker/src/bake/sprouting.rs:119-378 (run_sprouting_pass)
// with its helpers nudge_axon (51-117) and compute_power_index (32-40), the
// segment grid of spatial_grid.rs:99-185 and random_f32 of
// genesis-core/src/seed.rs:61-64. The integer rules and the f32 scoring are
// the reference's, evaluated in the same order (scores compared with a strict
// `>` in the same candidate order, f32 arithmetic without contraction), so
// the pass is bit-identical to the reference's on the same inputs:
//   1. axon liveness from the targets; axon -> soma map (last soma wins);
//   2. incoming handovers become ghost axons in free ghost slots (acks);
//   3. local axons of somas with flags & 0x02 and every ghost axon are
//      nudged one voxel along their direction (handover when leaving);
//   4. a segment grid of 2-voxel cells over all axon paths, axons inserted in
//      reverse order;
//   5. every soma with a position and flags & 0x0E fills empty slots from the
//      tail, at most max_sprouts, each with the best-scoring segment within 2
//      cells (self and duplicate axons excluded), weight +-initial weight of
//      the presynaptic type;
//   6. ghost axons left without synapses are reported for pruning.
// Somas are processed in index order because a sprout changes the weights
// that later power indices read; the candidate scan of each empty slot is
// split over threads with a merge that keeps the sequential choice (step 5).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>
#include <omp.h>
#include <sys/mman.h>
#include <new>
#include <algorithm>

namespace {

// Wall time of each phase on stderr when AXC_NIGHT_PROFILE is set.
struct PhaseClock {
  bool on = std::getenv("AXC_NIGHT_PROFILE") != nullptr;
  double t = omp_get_wtime();
  void mark(const char* what) {
    if (!on) return;
    const double now = omp_get_wtime();
    std::fprintf(stderr, "[sprout] %-28s %8.1f ms\n", what, (now - t) * 1e3);
    t = now;
  }
};

constexpr uint32_t kMaxD = 128;

// genesis-core/src/types.rs:16-46 (PackedPosition: x 10 | y 10 | z 8 | type 4).
inline uint32_t px(uint32_t p) { return p & 0x3FF; }
inline uint32_t py(uint32_t p) { return (p >> 10) & 0x3FF; }
inline uint32_t pz(uint32_t p) { return (p >> 20) & 0xFF; }
inline uint32_t ptype(uint32_t p) { return p >> 28; }
inline uint32_t pack_pos(uint32_t x, uint32_t y, uint32_t z, uint32_t t) {
  return (x & 0x3FF) | ((y & 0x3FF) << 10) | ((z & 0xFF) << 20) | ((t & 0xF) << 28);
}

// genesis-core/src/seed.rs:61-64.
inline float random_f32(uint64_t seed) {
  uint32_t bits = (uint32_t)(seed >> 41) | 0x3F800000u;
  float f;
  std::memcpy(&f, &bits, 4);
  return f - 1.0f;
}

inline uint32_t unpack_axon(uint32_t t) {  // layout.rs:268-271 (saturating -1)
  uint32_t a = t & 0x00FFFFFFu;
  return a ? a - 1 : 0;
}

// One candidate segment of the grid (8 B): its axon (24 bits) and segment
// index (8 bits); its type, a flag for an owner soma that sprouts in this
// pass (whose power index may still change), a flag for an axon without an
// owner (power index 1.0), and the owner's sum of |weights| at the start of
// the pass (at most 128 * 32768 = 2^22), from which its power index follows.
struct SegRef {
  static constexpr uint32_t kOwnerSprouts = 0x10, kNoOwner = 0x20;
  uint32_t axon_seg;
  uint32_t meta;  // type | flags | sum_abs << 8
  uint32_t axon() const { return axon_seg & 0x00FFFFFFu; }
  uint32_t seg() const { return axon_seg >> 24; }
  uint32_t type() const { return meta & 15u; }
  bool owner_sprouts() const { return (meta & kOwnerSprouts) != 0; }
  bool no_owner() const { return (meta & kNoOwner) != 0; }
  uint32_t sum_abs() const { return meta >> 8; }
};
static_assert(sizeof(SegRef) == 8, "grid references are 8 B");

// Large host arrays on transparent huge pages where the kernel grants them
// (fewer first-touch faults for the grid's hundreds of MB).
struct FreeDeleter {
  void operator()(void* p) const { std::free(p); }
};
template <class T>
std::unique_ptr<T[], FreeDeleter> huge_array(size_t n) {
  constexpr size_t kHuge = size_t(2) << 20;
  const size_t bytes = ((n ? n : 1) * sizeof(T) + kHuge - 1) / kHuge * kHuge;
  void* p = std::aligned_alloc(kHuge, bytes);
  if (!p) throw std::bad_alloc();
  madvise(p, bytes, MADV_HUGEPAGE);
  return std::unique_ptr<T[], FreeDeleter>(static_cast<T*>(p));
}

}  // namespace

extern "C" {

// Sprouting preferences of one neuron type (genesis-core config/blueprints.rs:112-124).
struct AxcSproutType {
  float w_distance, w_power, w_explore, w_type;
  uint16_t initial_synapse_weight;
  uint8_t is_inhibitory;
  uint8_t _pad;
};

// genesis-core/src/ipc.rs:30-67 (20 B, 12 B, 8 B).
struct AxcHandoverEvent {
  uint32_t origin_zone_hash, local_axon_id;
  uint16_t entry_x, entry_y;
  int8_t vector_x, vector_y, vector_z;
  uint8_t type_mask;
  uint16_t remaining_length;
  uint8_t entry_z, _padding;
};
static_assert(sizeof(AxcHandoverEvent) == 20, "AxonHandoverEvent is 20 B (ipc.rs:45-48)");
struct AxcHandoverAck {
  uint32_t target_zone_hash, src_axon_id, dst_ghost_id;
};
struct AxcHandoverPrune {
  uint32_t target_zone_hash, dst_ghost_id;
};

struct AxcSproutResult {
  uint32_t new_synapses;
  uint32_t handovers_out;  // events written to handovers[0..]
  uint32_t acks;
  uint32_t prunes;
};

// One sprouting pass. Arrays as in run_sprouting_pass: targets/weights
// columnar [128][padded_n]; flags [padded_n] (taken before sort & prune);
// ghost_origins [total_ghosts]; handovers: in [0, n_in) incoming events, out
// the outgoing ones (capacity max_handovers); tips/dirs/lengths [total_axons],
// paths [total_axons][256]; soma_positions [padded_n]. types (n_types
// entries, may be null: all scores 1.0 and weight +74 as in the reference).
// acks (capacity n_in) and prunes (capacity max_prunes) receive the outputs.
// changed_cols (may be null; capacity max_changed) receives the column
// (slot * padded_n + soma) of every new synapse in the order made, so that a
// caller can upload only those slots (axc_shard_write_slots).
int axc_sprout_pass(uint32_t* targets, int16_t* weights, const uint8_t* flags, const uint32_t* ghost_origins,
                    AxcHandoverEvent* handovers, uint32_t n_in, uint32_t max_handovers, uint32_t* tips,
                    uint32_t* dirs, const uint32_t* soma_to_axon, uint32_t padded_n, uint32_t total_ghosts,
                    uint32_t total_axons, uint32_t max_x, uint32_t max_y, const AxcSproutType* types,
                    uint32_t n_types, uint64_t epoch, uint8_t* lengths, uint32_t* paths,
                    const uint32_t* soma_positions, uint64_t master_seed, uint32_t zone_hash, uint16_t max_sprouts,
                    AxcHandoverAck* acks, AxcHandoverPrune* prunes, uint32_t max_prunes, AxcSproutResult* out,
                    uint32_t* changed_cols, uint32_t max_changed) {
  const size_t N = padded_n, A = total_axons;
  const size_t ghost_start = N, ghost_end = N + total_ghosts;
  if (ghost_end > A) return -1;
  if (A >= (1u << 24)) return -1;  // a target holds axon + 1 in 24 bits (layout.rs:268-271)
  PhaseClock clk;

  // 1. Liveness and the axon -> soma map (every writer stores the same 1).
  // Liveness is read only for ghost axons (step 6): a shard without ghosts
  // skips the scan of its targets.
  std::vector<uint8_t> active(A, 0);
  if (ghost_end > ghost_start) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < (int64_t)((size_t)kMaxD * N); k++) {
      if (targets[k] != 0) {
        uint32_t a = unpack_axon(targets[k]);
        if (a < A) __atomic_store_n(&active[a], (uint8_t)1, __ATOMIC_RELAXED);
      }
    }
  }
  std::vector<uint32_t> axon_to_soma(A, 0xFFFFFFFFu);
  for (size_t s = 0; s < N; s++) {
    uint32_t a = soma_to_axon[s];
    if (a != 0xFFFFFFFFu && a < A) axon_to_soma[a] = (uint32_t)s;
  }

  clk.mark("liveness, axon owners");
  // 2. Absorb incoming handovers into free ghost slots.
  uint32_t n_acks = 0;
  size_t next_free = ghost_start;
  for (uint32_t k = 0; k < n_in; k++) {
    const AxcHandoverEvent ev = handovers[k];
    while (next_free < ghost_end && tips[next_free] != 0) next_free++;
    if (next_free >= ghost_end) break;  // ghost capacity exceeded
    acks[n_acks++] = AxcHandoverAck{ev.origin_zone_hash, ev.local_axon_id, (uint32_t)next_free};
    uint32_t tip = ((uint32_t)ev.type_mask << 28) | ((uint32_t)ev.entry_z << 22) | ((uint32_t)ev.entry_y << 11) |
                   (uint32_t)ev.entry_x;
    uint32_t dir = ((uint32_t)(uint8_t)ev.vector_z << 16) | ((uint32_t)(uint8_t)ev.vector_y << 8) |
                   (uint32_t)(uint8_t)ev.vector_x;
    tips[next_free] = tip;
    dirs[next_free] = dir;
    lengths[next_free] = 0;
    paths[next_free * 256] = tip;
    next_free++;
  }

  clk.mark("handovers");
  // 3. Nudge living local axons and every ghost axon.
  uint32_t n_hand = 0;
  auto nudge = [&](size_t a) {
    const uint32_t tip = tips[a];
    if (tip == 0) return;
    const int32_t tx = (int32_t)px(tip), ty = (int32_t)py(tip), tz = (int32_t)pz(tip);
    const uint8_t type_mask = (uint8_t)ptype(tip);
    const uint32_t d = dirs[a];
    const int8_t dx = (int8_t)(d & 0xFF), dy = (int8_t)((d >> 8) & 0xFF), dz = (int8_t)((d >> 16) & 0xFF);
    const int32_t sx = dx > 64 ? 1 : (dx < -64 ? -1 : 0);
    const int32_t sy = dy > 64 ? 1 : (dy < -64 ? -1 : 0);
    const int32_t sz = dz > 64 ? 1 : (dz < -64 ? -1 : 0);
    const int32_t nx = tx + sx, ny = ty + sy, nz = tz + sz;
    if (nx < 0 || nx >= (int32_t)max_x || ny < 0 || ny >= (int32_t)max_y || nz < 0 || nz > 63) {
      if (n_hand < max_handovers) {
        AxcHandoverEvent ev{};
        ev.origin_zone_hash = zone_hash;
        ev.local_axon_id = (uint32_t)a;
        ev.entry_x = (uint16_t)std::min(std::max(nx, 0), 2047);
        ev.entry_y = (uint16_t)std::min(std::max(ny, 0), 2047);
        ev.vector_x = dx;
        ev.vector_y = dy;
        ev.vector_z = dz;
        ev.type_mask = type_mask;
        const uint16_t len = lengths[a];
        ev.remaining_length = (uint16_t)(len < 256 ? 256 - len : 0);
        ev.entry_z = (uint8_t)std::min(std::max(nz, 0), 63);
        ev._padding = 0;
        handovers[n_hand++] = ev;
      }
      tips[a] = 0;
      return;
    }
    const uint32_t next = pack_pos((uint32_t)nx, (uint32_t)ny, (uint32_t)nz, type_mask);
    tips[a] = next;
    const uint32_t len = lengths[a];
    paths[a * 256 + len] = next;          // len <= 255 always (u8)
    lengths[a] = (uint8_t)(len + 1);      // 255 + 1 wraps to 0 as the u8 cast does
  };
  for (size_t s = 0; s < N; s++) {
    if (flags[s] & 0x02) {
      uint32_t a = soma_to_axon[s];
      if (a != 0xFFFFFFFFu && a < A) nudge(a);
    }
  }
  for (size_t a = ghost_start; a < ghost_end; a++) nudge(a);

  clk.mark("nudges");
  // Power index of every soma (sprouting.rs:32-40), explore noise of every
  // axon (random_f32 of seed + axon + epoch, sprouting.rs:280-282), and the
  // somas that sprout candidates (placed, bursting, an empty last slot) in
  // index order.
  // The power index is (float)sum|w| / (128 * 32767.0f); the integer sums
  // are kept so that a sprouting soma's index follows from its new weights.
  auto power_of = [](uint32_t sum_abs) -> float { return (float)sum_abs / ((float)kMaxD * 32767.0f); };
  std::vector<uint32_t> sum_abs(N);
  std::vector<float> power(N);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < (int64_t)N; s++) {
    uint32_t pw = 0;
    for (uint32_t k = 0; k < kMaxD; k++) {
      const int32_t w = weights[(size_t)k * N + s];
      pw += (uint32_t)(w < 0 ? -w : w);
    }
    sum_abs[s] = pw;
    power[s] = power_of(pw);
  }
  std::vector<uint32_t> sprouters;
  for (size_t i = 0; i < N; i++)
    if (soma_positions[i] != 0 && (flags[i] & 0x0E) != 0 && targets[(size_t)(kMaxD - 1) * N + i] == 0)
      sprouters.push_back((uint32_t)i);
  std::vector<uint8_t> may_change(N, 0);
  for (uint32_t i : sprouters) may_change[i] = 1;

  clk.mark("power index, sprouters");
  // 4. Segment grid, cells of 2 voxels, axons in reverse order, dense CSR
  // with the insertion order of the reference's per-cell vectors. Built in
  // parallel: thread t takes the t-th contiguous range of axons in reverse
  // order and places its segments after those of threads t' < t in every
  // cell, which is the sequential insertion order. A segment is left out
  // when its axon's previous segment or the one before (same type) lies in
  // the same cell: such a segment comes after that one in the cell and
  // scores the same (the score depends on the axon and its type only), so
  // the scan's strict `>` never picks it. (Later segments of an axon that
  // re-enters a cell stay in; they are never picked either.)
  // Each thread walks its axons' paths once: it counts per cell and appends
  // (cell, reference) to a list of its own in walk order; the lists are then
  // scattered into the grid.
  const uint32_t cell = 2;
  // What a score reads of an axon: its explore noise (random_f32 of seed +
  // axon + epoch, sprouting.rs:280-282) and its owner's power index at the
  // start of the pass (1.0 without an owner, sprouting.rs:285-290).
  auto noise_of = [&](uint32_t a) { return types ? random_f32(master_seed + (uint64_t)a + epoch) : 0.0f; };
  auto pw_of = [&](SegRef r) { return r.no_owner() ? 1.0f : power_of(r.sum_abs()); };
  const int nth = omp_get_max_threads();
  // Axon range of thread t in reverse order: axons [A - hi(t), A - lo(t)).
  auto lo_of = [&](int t) { return (size_t)((A * (size_t)t) / (size_t)nth); };
  std::vector<std::unique_ptr<uint32_t[], FreeDeleter>> t_cell(nth);
  std::vector<std::unique_ptr<SegRef[], FreeDeleter>> t_ref(nth);
  std::vector<size_t> t_len(nth, 0);
  // Every grid entry is written by the scatter below: left uninitialized,
  // so that the scattering threads touch its pages first (hundreds of MB at
  // 1.5M neurons).
  std::unique_ptr<SegRef[], FreeDeleter> refs;
  // The walk stores each kept segment's cell as z/2 << 18 | y/2 << 9 | x/2
  // (the grid's extent is known only after it) and the greatest cell
  // coordinates; the count pass turns the keys into dense cell ids.
  std::vector<uint32_t> t_max(3 * (size_t)nth, 0);
  // Range t is taken by thread t % (threads granted): every range is built
  // whatever number of threads the runtime grants.
#pragma omp parallel num_threads(nth)
  for (int t = omp_get_thread_num(); t < nth; t += omp_get_num_threads()) {
    size_t cap = 1;  // the branch-free walk writes one entry past the last kept one
    for (size_t k = lo_of(t); k < lo_of(t + 1); k++) cap += lengths[A - 1 - k];
    t_cell[t] = huge_array<uint32_t>(cap);
    t_ref[t] = huge_array<SegRef>(cap);
    uint32_t* oc = t_cell[t].get();
    SegRef* orf = t_ref[t].get();
    size_t n = 0;
    uint32_t mx = 0, my = 0, mz = 0;
    for (size_t k = lo_of(t); k < lo_of(t + 1); k++) {
      const uint32_t ai = (uint32_t)(A - 1 - k);
      const uint32_t owner = axon_to_soma[ai];
      const uint32_t meta = owner == 0xFFFFFFFFu
                                ? SegRef::kNoOwner
                                : (sum_abs[owner] << 8) | (may_change[owner] ? SegRef::kOwnerSprouts : 0u);
      const uint32_t* path = paths + (size_t)ai * 256;
      uint32_t k1 = ~0u, k2 = ~0u;
      for (uint32_t s = 0, len = lengths[ai]; s < len; s++) {
        const uint32_t p = path[s];
        const uint32_t x = px(p) / cell, y = py(p) / cell, z = pz(p) / cell;
        const uint32_t fk = (z << 18) | (y << 9) | x;
        const uint32_t key = (fk << 4) | ptype(p);
        const bool live = p != 0;
        oc[n] = fk;
        orf[n] = SegRef{ai | (s << 24), meta | ptype(p)};
        n += (size_t)(live & (key != k1) & (key != k2));
        mx = live ? std::max(mx, x) : mx;
        my = live ? std::max(my, y) : my;
        mz = live ? std::max(mz, z) : mz;
        k2 = live ? k1 : k2;
        k1 = live ? key : k1;
      }
    }
    t_len[t] = n;
    t_max[3 * (size_t)t] = mx;
    t_max[3 * (size_t)t + 1] = my;
    t_max[3 * (size_t)t + 2] = mz;
  }
  clk.mark("grid: walk");
  uint32_t gx = 1, gy = 1, gz = 1;
  for (int t = 0; t < nth; t++)
    if (t_len[t]) {
      gx = std::max(gx, t_max[3 * (size_t)t] + 1);
      gy = std::max(gy, t_max[3 * (size_t)t + 1] + 1);
      gz = std::max(gz, t_max[3 * (size_t)t + 2] + 1);
    }
  const size_t ncell = (size_t)gx * gy * gz;
  std::vector<uint32_t> cnt((size_t)nth * ncell, 0);
  std::vector<uint32_t> start(ncell + 1, 0);
#pragma omp parallel num_threads(nth)
  for (int t = omp_get_thread_num(); t < nth; t += omp_get_num_threads()) {
    uint32_t* mine = cnt.data() + (size_t)t * ncell;
    uint32_t* oc = t_cell[t].get();
    for (size_t k = 0; k < t_len[t]; k++) {
      const uint32_t fk = oc[k];
      const uint32_t c = (uint32_t)(((size_t)(fk >> 18) * gy + ((fk >> 9) & 511u)) * gx + (fk & 511u));
      oc[k] = c;
      mine[c]++;
    }
  }
  clk.mark("grid: counts");
  // Exclusive offsets: cell-major, thread order within a cell.
  size_t run = 0;
  for (size_t c = 0; c < ncell; c++) {
    start[c] = (uint32_t)run;
    for (int t = 0; t < nth; t++) {
      const uint32_t k = cnt[(size_t)t * ncell + c];
      cnt[(size_t)t * ncell + c] = (uint32_t)run;
      run += k;
    }
  }
  start[ncell] = (uint32_t)run;
  refs = huge_array<SegRef>(run);
  clk.mark("grid: offsets");
#pragma omp parallel num_threads(nth)
  for (int t = omp_get_thread_num(); t < nth; t += omp_get_num_threads()) {
    uint32_t* posn = cnt.data() + (size_t)t * ncell;
    const uint32_t* oc = t_cell[t].get();
    const SegRef* orf = t_ref[t].get();
    for (size_t k = 0; k < t_len[t]; k++) refs[posn[oc[k]]++] = orf[k];
    t_cell[t].reset();
    t_ref[t].reset();
  }
  clk.mark("grid: scatter");
  // Per cell, bounds on what a score reads of its references: explore
  // noise (random_f32 is nondecreasing in seed + axon + epoch, so the
  // cell's least and greatest axon bound it unless that sum wraps between
  // them), power index (separately for owners that sprout in this pass,
  // whose index may still rise by their sprouts' weights; power_of is
  // nondecreasing in the sum) and the presynaptic types. Phase A uses them
  // to pass over cells that cannot change its result (below).
  struct CellBound {
    float nmin, nmax, pmin, pmax, qmin, qmax;
    uint16_t types;
    uint8_t any_p, any_q;
  };
  std::vector<CellBound> cb(ncell);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t c = 0; c < (int64_t)ncell; c++) {
    uint32_t a_lo = ~0u, a_hi = 0, p_lo = ~0u, p_hi = 0, q_lo = ~0u, q_hi = 0, tmask = 0, meta_or = 0;
    for (uint32_t r = start[c]; r < start[c + 1]; r++) {
      const SegRef sr = refs[r];
      a_lo = std::min(a_lo, sr.axon());
      a_hi = std::max(a_hi, sr.axon());
      tmask |= 1u << sr.type();
      meta_or |= sr.meta;
      if (sr.owner_sprouts()) {
        q_lo = std::min(q_lo, sr.sum_abs());
        q_hi = std::max(q_hi, sr.sum_abs());
      } else if (!sr.no_owner()) {
        p_lo = std::min(p_lo, sr.sum_abs());
        p_hi = std::max(p_hi, sr.sum_abs());
      }
    }
    CellBound b{0.0f, 1.0f, 1e30f, -1e30f, 1e30f, -1e30f, (uint16_t)tmask, 0, 0};
    if (start[c + 1] > start[c]) {
      const uint64_t x_lo = master_seed + (uint64_t)a_lo + epoch, x_hi = master_seed + (uint64_t)a_hi + epoch;
      if (x_lo <= x_hi) {
        b.nmin = noise_of(a_lo);
        b.nmax = noise_of(a_hi);
      }
    }
    if (p_lo <= p_hi) {
      b.pmin = power_of(p_lo);
      b.pmax = power_of(p_hi);
      b.any_p = 1;
    }
    if (meta_or & SegRef::kNoOwner) {
      b.pmin = std::min(b.pmin, 1.0f);
      b.pmax = std::max(b.pmax, 1.0f);
      b.any_p = 1;
    }
    if (q_lo <= q_hi) {
      b.qmin = power_of(q_lo);
      b.qmax = power_of(q_hi);
      b.any_q = 1;
    }
    cb[c] = b;
  }
  // The most a sprouting owner's power index can rise in phase B: its
  // max_sprouts new synapses of at most the largest initial weight, with
  // room for the f32 rounding of the quotient.
  uint32_t w0max = 74;
  if (types) {
    w0max = 0;
    for (uint32_t k = 0; k < n_types; k++) w0max = std::max<uint32_t>(w0max, types[k].initial_synapse_weight);
  }
  const float power_rise =
      ((float)max_sprouts * (float)std::min<uint32_t>(w0max, 32768u) + 4.0f) / ((float)kMaxD * 32767.0f) + 1e-6f;

  clk.mark("segment grid");
  // 5. Synaptogenesis. The reference visits the somas in index order, and
  // for each empty slot scans the candidate segments of the 5x5x5 cells
  // around the soma (z, y, x ascending, each cell's segments in grid order)
  // keeping the first best score (strict `>`). A soma's choice depends on the
  // other somas only through the power index of candidate owners
  // (sprouting.rs:285-290), which changes when an owner sprouts earlier in
  // the order (its weights gain the new synapses; scores only rise with
  // power). Everything else it reads is its own. So the pass runs in two
  // phases with the reference's results bit for bit:
  //  A. every sprouting soma (in parallel, one scan each) keeps the best
  //     K = max(1, max_sprouts) distinct axons among its candidates whose
  //     owner cannot change before it (no owner, a later owner, or an owner
  //     that does not sprout), ranked by (score, scan position), and lists
  //     the other candidates (an earlier owner that sprouts) with the part of
  //     their score before the power term;
  //  B. in index order, each empty slot takes the best of the unpicked kept
  //     candidates and the listed ones rescored with the owners' current
  //     power index; after its sprouts the soma's power index is updated.
  // The soma's present axons are found in a small hash set.
  const uint32_t K = max_sprouts > 0 ? max_sprouts : 1u;
  struct Cand {
    float score;     // full score (kept candidates) or the part before the power term (listed ones)
    uint32_t pos;    // scan position: earlier wins a tie
    SegRef ref;
    uint32_t owner;  // the owner soma (listed candidates)
  };
  auto better = [](float s1, uint32_t p1, float s2, uint32_t p2) { return s1 > s2 || (s1 == s2 && p1 < p2); };
  const size_t n_spr = sprouters.size();
  std::vector<Cand> kept((size_t)n_spr * K);
  std::vector<uint32_t> n_kept(n_spr, 0);
  std::vector<std::vector<Cand>> listed(n_spr);

#pragma omp parallel
  {
    constexpr uint32_t kHash = 512;  // the soma's present axons (<= 128)
    uint32_t hset[kHash];
    std::vector<size_t> cells;
    cells.reserve(125);
#pragma omp for schedule(dynamic, 16)
    for (int64_t q = 0; q < (int64_t)n_spr; q++) {
      const uint32_t i = sprouters[(size_t)q];
      const uint32_t my_pos = soma_positions[i];
      const uint32_t my_type = ptype(my_pos);
      const AxcSproutType* cfg = (types && my_type < n_types) ? &types[my_type] : nullptr;
      for (uint32_t h = 0; h < kHash; h++) hset[h] = 0xFFFFFFFFu;
      for (uint32_t s = 0; s < kMaxD; s++) {
        const uint32_t t = targets[(size_t)s * N + i];
        if (!t) continue;
        const uint32_t a = unpack_axon(t);
        uint32_t h = (a * 2654435761u) >> 23;
        while (hset[h] != 0xFFFFFFFFu && hset[h] != a) h = (h + 1) & (kHash - 1);
        hset[h] = a;
      }
      auto present = [&](uint32_t a) {
        uint32_t h = (a * 2654435761u) >> 23;
        while (hset[h] != 0xFFFFFFFFu) {
          if (hset[h] == a) return true;
          h = (h + 1) & (kHash - 1);
        }
        return false;
      };
      const uint32_t my_axon = soma_to_axon[i];
      const int32_t cx = (int32_t)(px(my_pos) / cell), cy = (int32_t)(py(my_pos) / cell),
                    cz = (int32_t)(pz(my_pos) / cell);
      cells.clear();
      for (int32_t z = cz - 2; z <= cz + 2; z++) {
        if (z < 0 || z >= (int32_t)gz) continue;
        for (int32_t y = cy - 2; y <= cy + 2; y++) {
          if (y < 0 || y >= (int32_t)gy) continue;
          for (int32_t x = cx - 2; x <= cx + 2; x++) {
            if (x < 0 || x >= (int32_t)gx) continue;
            cells.push_back(((size_t)z * gy + (size_t)y) * gx + (size_t)x);
          }
        }
      }
      Cand* top = &kept[(size_t)q * K];
      uint32_t nt = 0;
      std::vector<Cand>& lst = listed[(size_t)q];
      uint32_t pos = 0;
      for (size_t c : cells) {
        // Once the kept list is full, a cell whose every score is at most the
        // last kept one cannot change the result: its candidates come later
        // in the scan (equal scores lose), its kept candidates cannot enter,
        // and a listed one (rescored in phase B) could only be picked by
        // beating every unpicked kept entry. The bound follows the score's
        // own f32 steps with each term at its extreme over the cell (each
        // step is monotone in its operands).
        if (nt == K) {
          const CellBound& b = cb[c];
          float bound = 1.0f;
          if (cfg) {
            const float n_ext = cfg->w_explore >= 0.0f ? b.nmax : b.nmin;
            const bool has_mine = (b.types >> my_type) & 1u, has_other = (b.types & ~(1u << my_type)) != 0;
            const float same_ext = cfg->w_type >= 0.0f ? (has_mine ? 1.0f : 0.0f) : (has_other ? 0.0f : 1.0f);
            float p_ext;
            if (cfg->w_power >= 0.0f)
              p_ext = std::max(b.any_p ? b.pmax : -1e30f, b.any_q ? b.qmax + power_rise : -1e30f);
            else
              p_ext = std::min(b.any_p ? b.pmin : 1e30f, b.any_q ? b.qmin : 1e30f);
            float acc = cfg->w_distance * 1.0f;
            acc = acc + cfg->w_explore * n_ext;
            acc = acc + cfg->w_type * same_ext;
            acc = acc + cfg->w_power * p_ext;
            bound = acc;
          }
          if (bound <= top[K - 1].score) {
            pos += start[c + 1] - start[c];
            continue;
          }
        }
        for (uint32_t r = start[c]; r < start[c + 1]; r++, pos++) {
          const SegRef sr = refs[r];
          const uint32_t a = sr.axon();
          float score = 1.0f;
          if (cfg) {
            // Left to right, products first, no fused multiply-add (the
            // library is built with -ffp-contract=off).
            const float same = (my_type == sr.type()) ? 1.0f : 0.0f;
            float acc = cfg->w_distance * 1.0f;
            acc = acc + cfg->w_explore * noise_of(a);
            acc = acc + cfg->w_type * same;
            if (sr.owner_sprouts()) {
              const uint32_t owner = axon_to_soma[a];
              if (owner < i) {  // rescored in phase B
                if (a != my_axon && !present(a)) lst.push_back(Cand{acc, pos, sr, owner});
                continue;
              }
            }
            acc = acc + cfg->w_power * pw_of(sr);
            score = acc;
          }
          // The scan starts from a best score of -1: lower (or NaN) never wins;
          // nor does a reference that cannot enter a full list (an axon
          // already listed has a score at least the last entry's).
          if (!(score > -1.0f)) continue;
          if (nt == K && !better(score, pos, top[K - 1].score, top[K - 1].pos)) continue;
          if (a == my_axon || present(a)) continue;
          // The axon's best reference so far replaces its entry; a new axon
          // enters if the list has room or it beats the last entry.
          uint32_t k = 0;
          while (k < nt && top[k].ref.axon() != a) k++;
          if (k < nt) {
            if (!better(score, pos, top[k].score, top[k].pos)) continue;
          } else if (nt < K) {
            k = nt++;
          } else {
            if (!better(score, pos, top[K - 1].score, top[K - 1].pos)) continue;
            k = K - 1;
          }
          top[k] = Cand{score, pos, sr, 0};
          while (k > 0 && better(top[k].score, top[k].pos, top[k - 1].score, top[k - 1].pos)) {
            std::swap(top[k], top[k - 1]);
            k--;
          }
        }
      }
      n_kept[(size_t)q] = nt;
    }
  }

  clk.mark("phase A (parallel scan)");
  uint32_t new_syn = 0;
  std::vector<uint32_t> picked;
  for (size_t q = 0; q < n_spr; q++) {
    const uint32_t i = sprouters[q];
    const uint32_t my_type = ptype(soma_positions[i]);
    const AxcSproutType* cfg = (types && my_type < n_types) ? &types[my_type] : nullptr;
    const Cand* top = &kept[q * K];
    const std::vector<Cand>& lst = listed[q];
    picked.clear();
    auto is_picked = [&](uint32_t a) { return std::find(picked.begin(), picked.end(), a) != picked.end(); };
    uint32_t sprouts = 0;
    for (int32_t slot = (int32_t)kMaxD - 1; slot >= 0; slot--) {
      const size_t col = (size_t)slot * N + i;
      if (targets[col] != 0) break;
      const Cand* best = nullptr;
      float best_score = -1.0f;
      for (uint32_t k = 0; k < n_kept[q]; k++) {
        if (is_picked(top[k].ref.axon())) continue;
        best = &top[k];  // the kept list is ranked: its first unpicked entry is its best
        best_score = top[k].score;
        break;
      }
      for (const Cand& c : lst) {
        if (is_picked(c.ref.axon())) continue;
        const float sc = c.score + cfg->w_power * power[c.owner];
        if (!(sc > -1.0f)) continue;
        if (!best || better(sc, c.pos, best_score, best->pos)) {
          best = &c;
          best_score = sc;
        }
      }
      if (!best) break;
      const SegRef br = best->ref;
      const uint32_t b_axon = br.axon(), b_type = br.type();
      targets[col] = (br.seg() << 24) | ((b_axon + 1u) & 0x00FFFFFFu);
      bool inhib = false;
      int16_t w0 = 74;
      if (types) {
        if (b_type < n_types) {
          inhib = types[b_type].is_inhibitory != 0;
          w0 = (int16_t)types[b_type].initial_synapse_weight;
        }
      }
      const int32_t w_old = weights[col];
      weights[col] = inhib ? (int16_t)-w0 : w0;
      const int32_t w_new = weights[col];
      sum_abs[i] += (uint32_t)(w_new < 0 ? -w_new : w_new) - (uint32_t)(w_old < 0 ? -w_old : w_old);
      if (changed_cols && new_syn < max_changed) changed_cols[new_syn] = (uint32_t)col;
      picked.push_back(b_axon);
      active[b_axon] = 1;
      new_syn++;
      sprouts++;
      if (sprouts >= max_sprouts) break;
    }
    if (sprouts) power[i] = power_of(sum_abs[i]);  // this soma's weights changed
  }

  clk.mark("phase B (index order)");
  // 6. Orphan ghost axons.
  uint32_t n_prunes = 0;
  for (size_t g = ghost_start; g < ghost_end; g++) {
    if (!active[g] && tips[g] != 0) {
      const uint32_t origin = ghost_origins ? ghost_origins[g - ghost_start] : 0;
      if (origin != 0) {
        if (n_prunes < max_prunes && n_prunes < 1000) prunes[n_prunes++] = AxcHandoverPrune{origin, (uint32_t)g};
        tips[g] = 0;
      }
    }
  }
  clk.mark("orphan ghosts");
  if (out) *out = AxcSproutResult{new_syn, n_hand, n_acks, n_prunes};
  return 0;
}

}  // extern "C"
