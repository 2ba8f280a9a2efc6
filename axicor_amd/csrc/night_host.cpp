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
#include <cstring>
#include <vector>
#include <omp.h>
#include <algorithm>

namespace {

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

struct SegRef {
  uint32_t axon_id;
  uint16_t seg_idx;
  uint8_t type_idx;
};

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
int axc_sprout_pass(uint32_t* targets, int16_t* weights, const uint8_t* flags, const uint32_t* ghost_origins,
                    AxcHandoverEvent* handovers, uint32_t n_in, uint32_t max_handovers, uint32_t* tips,
                    uint32_t* dirs, const uint32_t* soma_to_axon, uint32_t padded_n, uint32_t total_ghosts,
                    uint32_t total_axons, uint32_t max_x, uint32_t max_y, const AxcSproutType* types,
                    uint32_t n_types, uint64_t epoch, uint8_t* lengths, uint32_t* paths,
                    const uint32_t* soma_positions, uint64_t master_seed, uint32_t zone_hash, uint16_t max_sprouts,
                    AxcHandoverAck* acks, AxcHandoverPrune* prunes, uint32_t max_prunes, AxcSproutResult* out) {
  const size_t N = padded_n, A = total_axons;
  const size_t ghost_start = N, ghost_end = N + total_ghosts;
  if (ghost_end > A) return -1;

  // 1. Liveness and the axon -> soma map.
  std::vector<uint8_t> active(A, 0);
  for (size_t k = 0; k < (size_t)kMaxD * N; k++) {
    if (targets[k] != 0) {
      uint32_t a = unpack_axon(targets[k]);
      if (a < A) active[a] = 1;
    }
  }
  std::vector<uint32_t> axon_to_soma(A, 0xFFFFFFFFu);
  for (size_t s = 0; s < N; s++) {
    uint32_t a = soma_to_axon[s];
    if (a != 0xFFFFFFFFu && a < A) axon_to_soma[a] = (uint32_t)s;
  }

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

  // 4. Segment grid, cells of 2 voxels, axons in reverse order, dense CSR
  // with the insertion order of the reference's per-cell vectors. Built in
  // parallel: thread t takes the t-th contiguous range of axons in reverse
  // order and places its segments after those of threads t' < t in every
  // cell, which is the sequential insertion order. Only the first segment of
  // an axon in a cell is kept: the others follow it directly in that cell
  // and score the same (the score depends on the axon and its type only), so
  // the scan's strict `>` never picks them.
  const uint32_t cell = 2;
  uint32_t gx = 1, gy = 1, gz = 1;
#pragma omp parallel for schedule(static) reduction(max : gx, gy, gz)
  for (int64_t a = 0; a < (int64_t)A; a++) {
    for (uint32_t s = 0; s < lengths[a]; s++) {
      uint32_t p = paths[a * 256 + s];
      if (!p) continue;
      gx = std::max(gx, px(p) / cell + 1);
      gy = std::max(gy, py(p) / cell + 1);
      gz = std::max(gz, pz(p) / cell + 1);
    }
  }
  const size_t ncell = (size_t)gx * gy * gz;
  auto cid = [&](uint32_t p) { return ((size_t)(pz(p) / cell) * gy + py(p) / cell) * gx + px(p) / cell; };
  const int nth = omp_get_max_threads();
  // Axon range of thread t in reverse order: axons [A - hi(t), A - lo(t)).
  auto lo_of = [&](int t) { return (size_t)((A * (size_t)t) / (size_t)nth); };
  std::vector<uint32_t> cnt((size_t)nth * ncell, 0);
  std::vector<uint32_t> start(ncell + 1, 0);
  std::vector<SegRef> refs;
  {
    // Per (thread, cell) counts of all segments.
#pragma omp parallel num_threads(nth)
    {
      const int t = omp_get_thread_num();
      uint32_t* my = cnt.data() + (size_t)t * ncell;
      for (size_t k = lo_of(t); k < lo_of(t + 1); k++) {
        const size_t ai = A - 1 - k;
        for (uint32_t s = 0; s < lengths[ai]; s++) {
          uint32_t p = paths[ai * 256 + s];
          if (p) my[cid(p)]++;
        }
      }
    }
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
    std::vector<SegRef> all(run);
#pragma omp parallel num_threads(nth)
    {
      const int t = omp_get_thread_num();
      uint32_t* pos = cnt.data() + (size_t)t * ncell;
      for (size_t k = lo_of(t); k < lo_of(t + 1); k++) {
        const size_t ai = A - 1 - k;
        for (uint32_t s = 0; s < lengths[ai]; s++) {
          uint32_t p = paths[ai * 256 + s];
          if (p) all[pos[cid(p)]++] = SegRef{(uint32_t)ai, (uint16_t)s, (uint8_t)ptype(p)};
        }
      }
    }
    // Keep the first of each run of one axon (and type) within a cell.
    std::vector<uint32_t> kept(ncell + 1, 0);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)ncell; c++) {
      uint32_t k = 0;
      for (uint32_t r = start[c]; r < start[c + 1]; r++)
        k += (r == start[c] || all[r].axon_id != all[r - 1].axon_id || all[r].type_idx != all[r - 1].type_idx);
      kept[c + 1] = k;
    }
    for (size_t c = 0; c < ncell; c++) kept[c + 1] += kept[c];
    refs.resize(kept[ncell]);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)ncell; c++) {
      uint32_t o = kept[c];
      for (uint32_t r = start[c]; r < start[c + 1]; r++)
        if (r == start[c] || all[r].axon_id != all[r - 1].axon_id || all[r].type_idx != all[r - 1].type_idx)
          refs[o++] = all[r];
    }
    start.swap(kept);
  }

  // 5. Synaptogenesis. Three exact rewrites of the reference loop keep its
  // results bit for bit:
  //  - the power index of every soma (sprouting.rs:32-40) is kept in an
  //    array and recomputed when that soma sprouts (its weights change only
  //    then; the reference computes it at first use, from the same weights);
  //  - random_f32 of each axon for this epoch (the explore term) is computed
  //    once per axon instead of once per candidate segment;
  //  - a soma's presynaptic axons are marked in an array (mark[a] == i + 1)
  //    instead of a hash set.
  // The candidate scan of one empty slot (the 5x5x5 cells in z, y, x order,
  // each cell's segments in grid order) is split over threads by cells; each
  // keeps its first best score, and the merge takes the highest score, ties
  // to the earliest position: the strict `>` of the sequential scan.
  auto power_index = [&](uint32_t soma) -> float {
    uint32_t pw = 0;
    for (uint32_t s = 0; s < kMaxD; s++) {
      int32_t w = weights[(size_t)s * N + soma];
      pw += (uint32_t)(w < 0 ? -w : w);
    }
    return (float)pw / ((float)kMaxD * 32767.0f);
  };
  std::vector<float> power(N);
#pragma omp parallel for schedule(static)
  for (int64_t s = 0; s < (int64_t)N; s++) power[s] = power_index((uint32_t)s);
  std::vector<float> noise(types ? A : 0);
  if (types) {
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < (int64_t)A; a++) noise[a] = random_f32(master_seed + (uint64_t)a + epoch);
  }
  std::vector<uint32_t> mark(A, 0);
  struct Best {  // a thread's first best candidate
    float score;
    SegRef ref;
    bool found;
  };
  uint32_t new_syn = 0;
  std::vector<size_t> cells;
  cells.reserve(125);
  for (size_t i = 0; i < N; i++) {
    const uint32_t my_pos = soma_positions[i];
    if (my_pos == 0) continue;
    if ((flags[i] & 0x0E) == 0) continue;
    const uint32_t my_type = ptype(my_pos);
    const AxcSproutType* cfg = (types && my_type < n_types) ? &types[my_type] : nullptr;
    const uint32_t stamp = (uint32_t)i + 1;
    for (uint32_t s = 0; s < kMaxD; s++) {
      uint32_t t = targets[(size_t)s * N + i];
      if (t) {
        const uint32_t a = unpack_axon(t);
        if (a < A) mark[a] = stamp;  // axons >= A match no segment of the grid
      }
    }
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
    const int ncells = (int)cells.size();
    uint32_t sprouts = 0;
    for (int32_t slot = (int32_t)kMaxD - 1; slot >= 0; slot--) {
      const size_t col = (size_t)slot * N + i;
      if (targets[col] != 0) break;
      std::vector<Best> part;
#pragma omp parallel
      {
#pragma omp single
        part.assign((size_t)omp_get_num_threads(), Best{-1.0f, SegRef{}, false});
        Best b{-1.0f, SegRef{}, false};
        // Contiguous ranges of cells per thread, in thread order.
#pragma omp for schedule(static)
        for (int k = 0; k < ncells; k++) {
          const size_t c = cells[(size_t)k];
          for (uint32_t r = start[c]; r < start[c + 1]; r++) {
            const SegRef& sr = refs[r];
            const uint32_t a = sr.axon_id;
            if (a == my_axon || mark[a] == stamp) continue;
            float score = 1.0f;
            if (cfg) {
              const uint32_t owner = axon_to_soma[a];
              const float pw = owner == 0xFFFFFFFFu ? 1.0f : power[owner];
              const float same = (my_type == sr.type_idx) ? 1.0f : 0.0f;
              // Left to right, products first, no fused multiply-add
              // (the library is built with -ffp-contract=off).
              float acc = cfg->w_distance * 1.0f;
              acc = acc + cfg->w_explore * noise[a];
              acc = acc + cfg->w_type * same;
              acc = acc + cfg->w_power * pw;
              score = acc;
            }
            if (score > b.score) b = Best{score, sr, true};
          }
        }
        part[(size_t)omp_get_thread_num()] = b;
      }
      // Threads hold increasing ranges of cells: the first strictly better
      // one in thread order is the sequential scan's choice.
      Best best{-1.0f, SegRef{}, false};
      for (const Best& b : part)
        if (b.found && b.score > best.score) best = b;
      if (!best.found) break;
      targets[col] = ((uint32_t)best.ref.seg_idx << 24) | ((best.ref.axon_id + 1u) & 0x00FFFFFFu);
      bool inhib = false;
      int16_t w0 = 74;
      if (types) {
        if (best.ref.type_idx < n_types) {
          inhib = types[best.ref.type_idx].is_inhibitory != 0;
          w0 = (int16_t)types[best.ref.type_idx].initial_synapse_weight;
        }
      }
      weights[col] = inhib ? (int16_t)-w0 : w0;
      mark[best.ref.axon_id] = stamp;
      active[best.ref.axon_id] = 1;
      power[i] = power_index((uint32_t)i);  // this soma's weights changed
      new_syn++;
      sprouts++;
      if (sprouts >= max_sprouts) break;
    }
  }

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
  if (out) *out = AxcSproutResult{new_syn, n_hand, n_acks, n_prunes};
  return 0;
}

}  // extern "C"
