// brain_gen.cpp — seeded synthetic brains in the genesis-baker binary layouts.
//
// Produces a `.state` blob (headerless SoA, genesis-compute/src/memory.rs:71-86,
// written the way genesis-baker/src/bake/layout.rs:127-159 writes it), an
// `.axons` blob (raw BurstHeads8[], layout.rs:162-169), the 16-entry variant
// table and per-tick input bitmasks. All randomness comes from
// entity_seed(master, id) (genesis-core/src/seed.rs:50-57); Bernoulli draws
// compare a 64-bit hash against floor(p * 2^64) with p in parts per million.
//
// Topology model (a generator of our own; the reference baker's cone-traced
// growth is not reproduced): neurons fill a gx*gy*gz voxel grid in index
// order, neuron i sits in the voxel of rank floor(i*V/n), voxels ranked
// x-fastest or brick by brick (GenSpec::brick). Axon i belongs to neuron i
// (soma_to_axon[i] = i); input axons sit at [padded_n, padded_n+n_in)
// (genesis-compute/src/memory.rs:128-130 with zero ghosts) spread evenly over
// the grid. Each filled dendrite slot picks a presynaptic axon whose voxel is
// at Chebyshev distance r from the neuron's voxel, r geometric with the given
// mean, on a uniformly chosen face of the cube of half-width r, clamped to
// the grid; the segment is uniform in [0, len(axon)). Weights have |w|
// uniform in [w_min, w_max] and the sign of the presynaptic type
// (inputs are excitatory). Initial heads: h0 = 0x80000000 - len*v_seg, the
// other seven at the sentinel (genesis-baker/src/bake/axon_growth.rs:511-514).
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <cmath>
#include <algorithm>

namespace {

inline uint64_t entity_seed(uint64_t master, uint32_t id) {
  uint64_t seed = master + (uint64_t)id + 0x60bee2bee120fc15ull;
  unsigned __int128 tmp = (unsigned __int128)seed * 0xa3b195354a39b70dull;
  unsigned __int128 m1 = (tmp >> 64) ^ tmp;
  tmp = m1 * (unsigned __int128)0x1b03738712fad5c9ull;
  return (uint64_t)((tmp >> 64) ^ tmp);
}

// floor(x * n / 2^64) for a 64-bit hash x: uniform integer in [0, n).
inline uint32_t uniform(uint64_t x, uint32_t n) {
  return (uint32_t)(((unsigned __int128)x * n) >> 64);
}

constexpr uint64_t kPpmUnit = 18446744073709ull;  // floor(2^64 / 10^6)

// Domain tags so that each kind of draw uses its own stream.
constexpr uint64_t kTagType = 0x7479706500000001ull;
constexpr uint64_t kTagVolt = 0x766f6c7400000002ull;
constexpr uint64_t kTagLen = 0x6c656e6700000003ull;
constexpr uint64_t kTagDend = 0x64656e6400000004ull;
constexpr uint64_t kTagInput = 0x696e707400000005ull;

}  // namespace

extern "C" {

// Mirrors genesis-core layout.rs VariantParameters (64 B); the caller passes
// 16 of them as raw bytes.
struct GenVariant {
  int32_t threshold, rest_potential, leak_rate, homeostasis_penalty;
  uint32_t spontaneous_firing_period_ticks;
  uint16_t initial_synapse_weight, gsop_potentiation, gsop_depression, homeostasis_decay;
  uint8_t refractory_period, synapse_refractory_period, signal_propagation_length, is_inhibitory;
  uint8_t inertia_curve[16];
  int32_t adaptive_leak_max;
  uint16_t adaptive_leak_gain;
  uint8_t adaptive_mode, _leak_pad[3];
  uint8_t d1_affinity, d2_affinity, _pad[4];
} __attribute__((aligned(64)));
static_assert(sizeof(GenVariant) == 64, "VariantParameters is 64 B");

struct GenSpec {
  uint32_t n_neurons;       // real neurons (padded_n = align64(n_neurons))
  uint32_t n_inputs;        // virtual input axons
  uint32_t gx, gy, gz;      // voxel grid
  uint32_t slots_filled;    // filled dendrite slots per real neuron (<= 128)
  uint32_t axon_len_min, axon_len_max;  // segments, in [1, 255]
  int32_t w_min, w_max;     // |w| range
  uint32_t v_seg;
  uint32_t n_types;         // variants used: type of neuron from type_ppm_cdf
  uint32_t type_ppm_cdf[16];  // cumulative ppm, last = 1'000'000
  uint32_t mean_dist_milli;  // mean Chebyshev distance of targets, in 1/1000 voxel
  uint64_t master_seed;
  uint32_t brick;            // 0: voxels numbered x-fastest; B > 0: BxBxB bricks, x-fastest
                             // inside a brick and over bricks (dims multiples of B);
                             // kBakerOrder: z-planes in order, each plane's voxels in a
                             // seeded shuffle (the baker's numbering)
};

// padded_n and total_axons of a spec (memory.rs:32 align_to_warp, boot.rs:65-71).
void axc_gen_sizes(const GenSpec* spec, uint32_t* padded_n, uint32_t* total_axons) {
  uint32_t pn = (spec->n_neurons + 63u) & ~63u;
  uint32_t a = pn + spec->n_inputs;
  a = (a + 31u) & ~31u;
  *padded_n = pn;
  *total_axons = a;
}

// Voxel numbering: rank <-> (x, y, z). With bricks, the voxels of one brick
// get consecutive ranks, so consecutive somas are spatial neighbours in all
// three axes. In the baker's order (genesis-baker/src/bake/neuron_placement.rs
// :60-87 sorts the placed somas by z only, each plane's positions taken from a
// shuffled pool) consecutive ranks share a z-plane and nothing else: rank r
// is plane r / (gx gy), at position perm[r % (gx gy) + plane offset] of that
// plane, perm a seeded Fisher-Yates shuffle per plane (our own shuffle; the
// baker's ChaCha8 stream is not reproduced).
constexpr uint32_t kBakerOrder = 0xFFFFFFFFu;
struct VoxelOrder {
  uint32_t gx, gy, gz, b;
  const uint32_t* perm = nullptr;  // baker order: [gz][gx*gy] in-plane index of each rank
  const uint32_t* inv = nullptr;   // its inverse
  void unrank(uint64_t r, uint32_t* x, uint32_t* y, uint32_t* z) const {
    if (b == kBakerOrder) {
      const uint64_t P = (uint64_t)gx * gy;
      const uint32_t q = perm[r];
      *x = q % gx;
      *y = q / gx;
      *z = (uint32_t)(r / P);
      return;
    }
    if (b == 0) {
      *x = (uint32_t)(r % gx);
      *y = (uint32_t)((r / gx) % gy);
      *z = (uint32_t)(r / ((uint64_t)gx * gy));
      return;
    }
    const uint64_t b3 = (uint64_t)b * b * b;
    const uint64_t br = r / b3, off = r % b3;
    const uint32_t nbx = gx / b, nby = gy / b;
    const uint32_t bx = (uint32_t)(br % nbx), by = (uint32_t)((br / nbx) % nby), bz = (uint32_t)(br / ((uint64_t)nbx * nby));
    *x = bx * b + (uint32_t)(off % b);
    *y = by * b + (uint32_t)((off / b) % b);
    *z = bz * b + (uint32_t)(off / ((uint64_t)b * b));
  }
  uint64_t rank(uint32_t x, uint32_t y, uint32_t z) const {
    if (b == kBakerOrder) {
      const uint64_t P = (uint64_t)gx * gy;
      return (uint64_t)z * P + inv[(uint64_t)z * P + (uint64_t)x + (uint64_t)gx * y];
    }
    if (b == 0) return (uint64_t)x + (uint64_t)gx * ((uint64_t)y + (uint64_t)gy * (uint64_t)z);
    const uint32_t nbx = gx / b, nby = gy / b;
    const uint64_t br = (uint64_t)(x / b) + (uint64_t)nbx * ((uint64_t)(y / b) + (uint64_t)nby * (uint64_t)(z / b));
    const uint64_t off = (uint64_t)(x % b) + (uint64_t)b * ((uint64_t)(y % b) + (uint64_t)b * (uint64_t)(z % b));
    return br * b * b * b + off;
  }
};

// The voxel order of the input axons' positions: the somas' order, except in
// the baker's order, whose ranks are not spatial within a plane: then 8x8x8
// bricks (x-fastest if the grid is not made of whole bricks), so that the
// inputs a soma draws from its target voxel are near it.
VoxelOrder input_order(const GenSpec* spec, const VoxelOrder& vo) {
  if (spec->brick != kBakerOrder) return vo;
  const bool bricks = spec->gx % 8 == 0 && spec->gy % 8 == 0 && spec->gz % 8 == 0;
  return VoxelOrder{spec->gx, spec->gy, spec->gz, bricks ? 8u : 0u};
}

// The per-plane shuffles of the baker's order (empty for the other orders).
void baker_tables(const GenSpec* spec, std::vector<uint32_t>& perm, std::vector<uint32_t>& inv) {
  if (spec->brick != kBakerOrder) return;
  const uint64_t P = (uint64_t)spec->gx * spec->gy;
  perm.resize(P * spec->gz);
  inv.resize(P * spec->gz);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < (int64_t)spec->gz; z++) {
    uint32_t* p = perm.data() + (size_t)z * P;
    for (uint64_t k = 0; k < P; k++) p[k] = (uint32_t)k;
    uint64_t h = entity_seed(spec->master_seed ^ 0x706c616e65000007ull, (uint32_t)z);
    for (uint64_t k = P - 1; k > 0; k--) {
      h = entity_seed(h, (uint32_t)k);
      std::swap(p[k], p[h % (k + 1)]);
    }
    uint32_t* v = inv.data() + (size_t)z * P;
    for (uint64_t k = 0; k < P; k++) v[p[k]] = (uint32_t)k;
  }
}

// Fills `state` (910*padded_n bytes), `heads` (8*total_axons u32) and
// optionally `axon_len` (total_axons bytes, 0 = no axon) and `soma_pos`
// (padded_n u32, PackedPosition genesis-core/src/types.rs:16-46, 0 for padding).
// `lut` holds 16 variants (threshold/rest give the initial voltage range,
// is_inhibitory the weight sign). Returns 0 or -1 on a bad spec.
int axc_gen_brain(const GenSpec* spec, const GenVariant* lut, uint8_t* state, uint32_t* heads,
                  uint8_t* axon_len, uint32_t* soma_pos) {
  if (spec->n_neurons == 0 || spec->slots_filled > 128 || spec->axon_len_min < 1 ||
      spec->axon_len_max > 255 || spec->axon_len_min > spec->axon_len_max || spec->w_min < 0 ||
      spec->w_max > 32767 || spec->w_min > spec->w_max || spec->n_types == 0 || spec->n_types > 16 ||
      spec->gx == 0 || spec->gy == 0 || spec->gz == 0)
    return -1;
  if (spec->brick && spec->brick != kBakerOrder &&
      (spec->gx % spec->brick || spec->gy % spec->brick || spec->gz % spec->brick))
    return -1;
  std::vector<uint32_t> bperm, binv;
  baker_tables(spec, bperm, binv);
  const VoxelOrder vo{spec->gx, spec->gy, spec->gz, spec->brick, bperm.data(), binv.data()};
  const VoxelOrder vin = input_order(spec, vo);
  uint32_t pn, A;
  axc_gen_sizes(spec, &pn, &A);
  const uint32_t n = spec->n_neurons, nin = spec->n_inputs;
  const size_t N = pn;
  const uint64_t V = (uint64_t)spec->gx * spec->gy * spec->gz;
  const uint64_t ms = spec->master_seed;

  int32_t* volt = (int32_t*)state;
  uint8_t* flags = state + 4 * N;
  int32_t* thr = (int32_t*)(state + 5 * N);
  uint8_t* timers = state + 9 * N;
  uint32_t* s2a = (uint32_t*)(state + 10 * N);
  uint32_t* targets = (uint32_t*)(state + 14 * N);
  int16_t* weights = (int16_t*)(state + 526 * N);
  uint8_t* dtimers = state + 782 * N;
  std::memset(state, 0, 910 * N);

  // Types and lengths first: dendrites need the presynaptic type and length.
  std::vector<uint8_t> type(N, 0);
  std::vector<uint8_t> len(A, 0);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; i++) {
    uint32_t u = uniform(entity_seed(ms ^ kTagType, (uint32_t)i), 1000000u);
    uint32_t t = 0;
    while (t + 1 < spec->n_types && u >= spec->type_ppm_cdf[t]) t++;
    type[i] = (uint8_t)t;
  }
  const uint32_t span = spec->axon_len_max - spec->axon_len_min + 1;
#pragma omp parallel for schedule(static)
  for (int64_t a = 0; a < (int64_t)A; a++) {
    bool real = (a < (int64_t)n) || (a >= (int64_t)pn && a < (int64_t)pn + nin);
    len[a] = real ? (uint8_t)(spec->axon_len_min + uniform(entity_seed(ms ^ kTagLen, (uint32_t)a), span)) : 0;
  }

  // Geometric distance law: P(r >= k+1 | r >= k) = q, mean = q/(1-q), r >= 0.
  double mean = spec->mean_dist_milli / 1000.0;
  double q = mean / (1.0 + mean);
  uint64_t q_thr = (uint64_t)(q * 18446744073709551615.0);
  const uint32_t gx = spec->gx, gy = spec->gy, gz = spec->gz;
  const uint64_t total_src = (uint64_t)n + nin;

#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t ii = 0; ii < (int64_t)N; ii++) {
    uint32_t i = (uint32_t)ii;
    if (i >= n) {  // padding neuron: type 0, no axon, no dendrites
      s2a[i] = 0xFFFFFFFFu;
      if (soma_pos) soma_pos[i] = 0;
      continue;
    }
    const GenVariant& p = lut[type[i]];
    uint64_t vs = entity_seed(ms ^ kTagVolt, i);
    int64_t lo = p.rest_potential, hi = p.threshold;
    int32_t v0 = (hi > lo) ? (int32_t)(lo + (int64_t)uniform(vs, (uint32_t)(hi - lo))) : (int32_t)lo;
    volt[i] = v0;
    flags[i] = (uint8_t)(type[i] << 4);
    thr[i] = 0;
    timers[i] = 0;
    s2a[i] = i;
    uint64_t vox = (uint64_t)i * V / n;
    uint32_t x, y, z;
    vo.unrank(vox, &x, &y, &z);
    if (soma_pos) soma_pos[i] = (x & 0x3FF) | ((y & 0x3FF) << 10) | ((z & 0xFF) << 20) | ((uint32_t)(type[i] & 0xF) << 28);

    uint64_t ds = entity_seed(ms ^ kTagDend, i);
    for (uint32_t s = 0; s < spec->slots_filled; s++) {
      uint64_t h = entity_seed(ds, s * 8u);
      uint32_t r = 0;
      while (r < 64 && h < q_thr) { r++; h = entity_seed(ds, s * 8u + 1u + (r << 12)); }
      uint64_t h2 = entity_seed(ds, s * 8u + 2u);
      uint32_t face = uniform(h2, 6);
      int32_t c[3];
      uint64_t h3 = entity_seed(ds, s * 8u + 3u);
      c[0] = (int32_t)uniform(h3, 2 * r + 1) - (int32_t)r;
      c[1] = (int32_t)uniform(entity_seed(ds, s * 8u + 4u), 2 * r + 1) - (int32_t)r;
      c[2] = (int32_t)uniform(entity_seed(ds, s * 8u + 5u), 2 * r + 1) - (int32_t)r;
      c[face >> 1] = (face & 1) ? (int32_t)r : -(int32_t)r;
      int64_t tx = (int64_t)x + c[0], ty = (int64_t)y + c[1], tz = (int64_t)z + c[2];
      tx = tx < 0 ? 0 : (tx >= gx ? gx - 1 : tx);
      ty = ty < 0 ? 0 : (ty >= gy ? gy - 1 : ty);
      tz = tz < 0 ? 0 : (tz >= gz ? gz - 1 : tz);
      uint64_t tv = vo.rank((uint32_t)tx, (uint32_t)ty, (uint32_t)tz);
      uint64_t h6 = entity_seed(ds, s * 8u + 6u);
      // Source population: inputs with probability nin/(n+nin).
      uint32_t axon;
      bool from_input = nin > 0 && uniform(h6, (uint32_t)total_src) >= n;
      uint64_t h7 = entity_seed(ds, s * 8u + 7u);
      if (from_input) {
        const uint64_t tvi = vin.b == vo.b ? tv : vin.rank((uint32_t)tx, (uint32_t)ty, (uint32_t)tz);
        uint64_t k = ((unsigned __int128)tvi * nin + uniform(h7, nin)) / V;
        if (k >= nin) k = nin - 1;
        axon = pn + (uint32_t)k;
      } else {
        uint64_t j = ((unsigned __int128)tv * n + uniform(h7, n)) / V;
        if (vo.b == kBakerOrder) {
          // Ranks are not spatial within a plane here, so take a soma of
          // voxel tv itself (the formula above can land on rank tv - 1),
          // other than soma i: another in the voxel, else one in the voxel
          // next to it along x.
          auto somas_of = [&](uint64_t r, uint64_t* a, uint64_t* b) {
            *a = ((unsigned __int128)r * n + V - 1) / V;
            *b = ((unsigned __int128)(r + 1) * n + V - 1) / V;
          };
          uint64_t lo, hi;
          somas_of(tv, &lo, &hi);
          if (hi > lo) j = lo + uniform(h7, (uint32_t)(hi - lo));
          if (j == i && hi - lo > 1) j = lo + (j - lo + 1) % (hi - lo);
          if (j == i) {
            const uint32_t nx = tx + 1 < (int64_t)gx ? (uint32_t)tx + 1 : (tx > 0 ? (uint32_t)tx - 1 : (uint32_t)tx);
            somas_of(vo.rank(nx, (uint32_t)ty, (uint32_t)tz), &lo, &hi);
            if (hi > lo && lo != i) j = lo;
          }
        }
        if (j >= n) j = n - 1;
        if (j == i) j = (j + 1) % n;
        axon = (uint32_t)j;
      }
      uint32_t L = len[axon];
      uint32_t seg = uniform(entity_seed(h7, 1), L ? L : 1);
      bool inhib = (!from_input) && lut[type[axon]].is_inhibitory;
      int32_t mag = spec->w_min + (int32_t)uniform(entity_seed(h7, 2), (uint32_t)(spec->w_max - spec->w_min + 1));
      size_t col = (size_t)s * N + i;
      targets[col] = (seg << 24) | ((axon + 1u) & 0x00FFFFFFu);
      weights[col] = (int16_t)(inhib ? -mag : mag);
      dtimers[col] = 0;
    }
  }

#pragma omp parallel for schedule(static)
  for (int64_t a = 0; a < (int64_t)A; a++) {
    uint32_t* h = heads + (size_t)a * 8;
    for (int k = 0; k < 8; k++) h[k] = 0x80000000u;
    if (len[a]) h[0] = 0x80000000u - (uint32_t)len[a] * spec->v_seg;
  }
  if (axon_len) std::memcpy(axon_len, len.data(), A);
  return 0;
}

// Axon geometry for the Night Phase: paths [A][256] (PackedPosition of each
// segment, types.rs:16-46), tips [A] (last segment) and dirs [A] (packed i8
// growth direction, sprouting.rs:72-75). Axon a < n starts at its soma's
// voxel, input axon k at the voxel of rank floor(k*V/n_in); each segment moves
// one voxel along every axis whose direction component exceeds 64 in
// magnitude, plus a one-voxel seeded jitter on one axis, clamped to the grid.
// Requires the soma_pos and axon_len of axc_gen_brain.
int axc_gen_paths(const GenSpec* spec, const uint32_t* soma_pos, const uint8_t* axon_len, uint32_t* paths,
                  uint32_t* tips, uint32_t* dirs) {
  uint32_t pn, A;
  axc_gen_sizes(spec, &pn, &A);
  const uint32_t n = spec->n_neurons, nin = spec->n_inputs;
  const uint64_t V = (uint64_t)spec->gx * spec->gy * spec->gz;
  std::vector<uint32_t> bperm, binv;
  baker_tables(spec, bperm, binv);
  const VoxelOrder vo{spec->gx, spec->gy, spec->gz, spec->brick, bperm.data(), binv.data()};
  const VoxelOrder vin = input_order(spec, vo);
  const uint64_t ms = spec->master_seed ^ 0x7061746800000006ull;
#pragma omp parallel for schedule(static)
  for (int64_t ai = 0; ai < (int64_t)A; ai++) {
    const uint32_t a = (uint32_t)ai;
    uint32_t* row = paths + (size_t)a * 256;
    std::memset(row, 0, 256 * sizeof(uint32_t));
    tips[a] = 0;
    dirs[a] = 0;
    const uint32_t L = axon_len[a];
    if (L == 0) continue;
    int32_t x, y, z;
    uint32_t type;
    if (a < n) {
      const uint32_t p = soma_pos[a];
      x = (int32_t)(p & 0x3FF);
      y = (int32_t)((p >> 10) & 0x3FF);
      z = (int32_t)((p >> 20) & 0xFF);
      type = p >> 28;
    } else {
      const uint32_t k = a - pn;
      uint32_t ux, uy, uz;
      vin.unrank((uint64_t)k * V / (nin ? nin : 1), &ux, &uy, &uz);
      x = (int32_t)ux;
      y = (int32_t)uy;
      z = (int32_t)uz;
      type = 0;
    }
    const uint64_t h = entity_seed(ms, a);
    const int8_t d[3] = {(int8_t)(uniform(h, 255) - 127), (int8_t)(uniform(entity_seed(h, 1), 255) - 127),
                         (int8_t)(uniform(entity_seed(h, 2), 255) - 127)};
    int32_t c[3] = {x, y, z};
    const int32_t lim[3] = {(int32_t)spec->gx - 1, (int32_t)spec->gy - 1, std::min<int32_t>((int32_t)spec->gz - 1, 255)};
    for (uint32_t s = 0; s < L; s++) {
      row[s] = (uint32_t)(c[0] & 0x3FF) | ((uint32_t)(c[1] & 0x3FF) << 10) | ((uint32_t)(c[2] & 0xFF) << 20) |
               ((type & 0xF) << 28);
      const uint64_t j = entity_seed(h, 16 + s);
      const uint32_t axis = uniform(j, 3);
      const int32_t jit = (int32_t)uniform(entity_seed(j, 1), 3) - 1;
      for (int k = 0; k < 3; k++) {
        int32_t step = d[k] > 64 ? 1 : (d[k] < -64 ? -1 : 0);
        if ((uint32_t)k == axis) step += jit;
        c[k] = std::min(std::max(c[k] + step, 0), lim[k]);
      }
    }
    tips[a] = row[L - 1];
    dirs[a] = ((uint32_t)(uint8_t)d[2] << 16) | ((uint32_t)(uint8_t)d[1] << 8) | (uint32_t)(uint8_t)d[0];
  }
  return 0;
}

// Bernoulli input bitmasks for ticks tick0 .. tick0+T-1: out[t][words] with
// words = ceil(n_inputs/32); bit i of tick t is set iff
// entity_seed(entity_seed(master ^ TAG, tick), i) < ppm * floor(2^64/10^6).
void axc_gen_inputs(uint64_t master_seed, uint32_t n_inputs, uint32_t ppm, uint32_t tick0, uint32_t T,
                    uint32_t* out) {
  const uint32_t words = (n_inputs + 31) / 32;
  const uint64_t thr = (uint64_t)ppm * kPpmUnit;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < (int64_t)T; t++) {
    uint64_t ts = entity_seed(master_seed ^ kTagInput, tick0 + (uint32_t)t);
    uint32_t* row = out + (size_t)t * words;
    for (uint32_t w = 0; w < words; w++) {
      uint32_t bits = 0;
      for (uint32_t b = 0; b < 32; b++) {
        uint32_t i = w * 32 + b;
        if (i < n_inputs && entity_seed(ts, i) < thr) bits |= 1u << b;
      }
      row[w] = bits;
    }
  }
}

uint64_t axc_gen_entity_seed(uint64_t master, uint32_t id) { return entity_seed(master, id); }

}  // extern "C"
