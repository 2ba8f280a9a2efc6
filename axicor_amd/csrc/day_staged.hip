// day_staged.hip — the dendrite scan of K5 (physics.cu:182-206) with each
// soma block's presynaptic snapshot entries staged in LDS.
//
// A block of 256 consecutive somas reads its dendrite columns once per tick
// and gathers one snapshot entry per filled slot. Its distinct presynaptic
// axons are few (about 6,500 for the 25,600 slots of a 256-soma block of the
// 1.5M-neuron brain with somas numbered by 8x8x8 bricks). So the gathers that
// limit the direct scan become ~6,500 global loads per block plus LDS reads:
//   stage_build_kernel  (whenever the targets change) per block: the distinct
//                       axons of the filled slots (LDS hash set), each slot's
//                       block-local entry (u16) and segment (u8), and each
//                       soma's filled prefix;
//   neuron_staged_kernel (every tick) loads the block's byte snapshot entries
//                       into LDS, then four groups of 256 lanes scan slot
//                       chunks 8g..8g+7, 8(g+4).. of the block's somas (the
//                       slots before the first empty target, as the reference
//                       loop's `break`), and group 0 finishes each soma with
//                       the GLIF update and a spike mask word per wave
//                       (compact_spikes_kernel then builds the spike list).
// Each tick's GSOP (K6) runs after the pushes: gsop_chunks_kernel on the
// batch graphs' working copy of the weights (16 lanes per spiking soma, one
// chunk of 8 slots each), gsop_kernel on the reference rows for single
// ticks. Applying it inside the next tick's staged pass (as the direct pass
// does) measured slower even from the staged byte entries (this tick's heads
// at segment + v_seg, no dependent gathers): at 1% firing most waves hold one
// or two somas that spiked, whose slot pass runs with the other lanes idle (a
// pass before the scan: 575 us against 306 + 136 us for scan + GSOP at 1.5M
// neurons; inside the scan loop the registers spill).
// The snapshot entries are the byte form of the heads (min(h, 255) x 8, see
// byte_heads); a slot whose segment + propagation length exceeds 255 reads the
// u32 heads from global memory instead, so every hit is exact.
// Blocks whose distinct axons exceed kStageCap stage the first kStageCap of
// them; the slots of the others take the exact path from the u32 heads. Only
// blocks whose hash set overflows read the plan's targets and gather the
// byte snapshot from global memory.
#include "day_common.h"

namespace axc {

// Lane groups per 256-soma block; each scans every kStageGroups-th chunk of
// 8 slots. (Two groups at half the occupancy, with the next chunk's loads
// issued early, measured slower.)
constexpr uint32_t kStageGroups = 4;
// Somas per staged block (the block's lanes are kStageGroups groups of this
// many). 128-soma blocks (3 or 4 workgroups per CU, kStageCap 5248 or 4544)
// measured 25-30% slower at 1.5M neurons: their staging costs more per soma
// (32 entries per soma against 25) than the extra workgroups recover.
#ifndef AXC_BLOCK_SOMAS
#define AXC_BLOCK_SOMAS 256
#endif
constexpr uint32_t kBlockSomas = AXC_BLOCK_SOMAS;
static_assert(kBlockSomas == 128 || kBlockSomas == 256, "staged blocks of 128 or 256 somas");
constexpr uint32_t kStageThreads = kBlockSomas * kStageGroups;
constexpr uint32_t kGroup0Waves = kBlockSomas / 64;  // the waves of group 0
constexpr uint32_t kHashSize = 16384;     // LDS hash set of the build kernel
constexpr uint32_t kHashEmpty = 0xFFFFFFFFu;
constexpr uint32_t kHashMaxProbe = 256;

__device__ __forceinline__ uint32_t hash_slot(uint32_t key) { return (key * 2654435761u) >> (32 - 14); }

// A caller's target in the plan's axon ids (axons outside [0, A) and the
// empty target unchanged).
__device__ __forceinline__ uint32_t plan_target(uint32_t t, uint32_t A, const uint32_t* __restrict__ axon_map) {
  if (!axon_map) return t;
  const uint32_t a = (t & 0x00FFFFFFu) - 1u;
  return a < A ? (t & 0xFF000000u) | ((axon_map[a] + 1u) & 0x00FFFFFFu) : t;
}

// One workgroup of 256 lanes per soma block (grid = ceil(N/256)). The
// block's distinct axons are gathered in an LDS hash set, then sorted, so
// that the set (and the LDS entries) are in axon order: consecutive lanes of
// the per-tick staging loop then gather neighbouring snapshot entries, which
// share cache lines. A slot's block-local entry is the rank of its axon.
__global__ __launch_bounds__(kBlockSomas) void stage_build_kernel(const uint32_t* __restrict__ targets, uint32_t N,
                                                          uint32_t A, StagePlan sp) {
  __shared__ uint32_t keys[kHashSize];
  __shared__ uint32_t sorted[kHashSize];
  __shared__ uint32_t scan[kBlockSomas];
  __shared__ uint32_t overflow;
  const uint32_t tid = threadIdx.x, b = blockIdx.x;
  const uint32_t i = b * kBlockSomas + tid;
  const uint32_t r = (sp.perm && i < N) ? sp.perm[i] : i;  // the caller's soma
  for (uint32_t j = tid; j < kHashSize; j += kBlockSomas) keys[j] = kHashEmpty;
  if (tid == 0) overflow = 0;
  __syncthreads();
  uint32_t F = 0;
  if (i < N) {
    for (; F < kMaxDendrites; F++) {
      const uint32_t t = plan_target(targets[(size_t)F * N + r], A, sp.axon_map);
      if (t == 0) break;
      sp.tgt[((size_t)(F >> 3) * N + i) * 8 + (F & 7)] = t;
      const uint32_t key = snap_index(t, A);
      uint32_t h = hash_slot(key), probe = 0;
      for (; probe < kHashMaxProbe; probe++) {
        const uint32_t prev = atomicCAS(&keys[h], kHashEmpty, key);
        if (prev == kHashEmpty || prev == key) break;
        h = (h + 1) & (kHashSize - 1);
      }
      if (probe == kHashMaxProbe) overflow = 1;
    }
    sp.fill[i] = (uint8_t)F;
    for (uint32_t s = F; s & 7; s++) sp.tgt[((size_t)(s >> 3) * N + i) * 8 + (s & 7)] = 0;
  }
  __syncthreads();
  // Compact the table (lane tid owns positions [kPer tid, kPer tid + kPer)).
  constexpr uint32_t kPer = kHashSize / kBlockSomas;
  uint32_t mine = 0;
  for (uint32_t j = 0; j < kPer; j++) mine += keys[tid * kPer + j] != kHashEmpty;
  scan[tid] = mine;
  __syncthreads();
  for (uint32_t off = 1; off < kBlockSomas; off <<= 1) {
    const uint32_t v = tid >= off ? scan[tid - off] : 0u;
    __syncthreads();
    scan[tid] += v;
    __syncthreads();
  }
  const uint32_t total = scan[kBlockSomas - 1];
  // Only a failed hash insert sends the block to the unstaged scan. Past
  // kStageCap distinct axons the first kStageCap (in axon order) are staged
  // and the slots of the rest are spilled (see below).
  const bool fits = overflow == 0;
  const uint32_t staged_n = total < kStageCap ? total : kStageCap;
  if (tid == 0) sp.cnt[b] = fits ? staged_n : kStageOverflow;
  if (!fits) return;  // block-uniform
  uint32_t next = scan[tid] - mine;
  for (uint32_t j = 0; j < kPer; j++) {
    const uint32_t key = keys[tid * kPer + j];
    if (key != kHashEmpty) sorted[next++] = key;
  }
  // Bitonic sort of sorted[0, S), S = the next power of two >= total,
  // padded with kHashEmpty (which sorts last).
  uint32_t S = 1;
  while (S < total) S <<= 1;
  for (uint32_t j = total + tid; j < S; j += kBlockSomas) sorted[j] = kHashEmpty;
  __syncthreads();
  for (uint32_t k = 2; k <= S; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t x = tid; x < S; x += kBlockSomas) {
        const uint32_t y = x ^ j;
        if (y > x) {
          const uint32_t a = sorted[x], c = sorted[y];
          const bool up = (x & k) == 0;
          if (up ? a > c : a < c) {
            sorted[x] = c;
            sorted[y] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  __syncthreads();  // also when S == 1 and the loops above did not run
  for (uint32_t j = tid; j < staged_n; j += kBlockSomas) sp.set[(size_t)b * kStageCap + j] = sorted[j];
  if (i >= N) return;
  for (uint32_t s = 0; s < F; s++) {
    const uint32_t t = plan_target(targets[(size_t)s * N + r], A, sp.axon_map);
    const uint32_t key = snap_index(t, A);
    uint32_t lo = 0, hi = total;  // the key is present: sorted[lo] == key at the end
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (sorted[mid] <= key) lo = mid;
      else hi = mid;
    }
    const size_t c = ((size_t)(s >> 3) * N + i) * 8 + (s & 7);  // chunk-interleaved
    // A spilled slot (its axon past the staged ones) gets the inert entry
    // and segment 255, which sends it down the scan's exact path (segment
    // + length > 255), where the target and the u32 heads decide it.
    const bool spilled = lo >= kStageCap;
    sp.code[c] = (uint16_t)(spilled ? kStageCap : lo);
    sp.seg[c] = (uint8_t)(spilled ? 255u : (t >> 24));
  }
  // The rest of the last chunk: the inert entry, segment 0 (never a hit).
  for (uint32_t s = F; s & 7; s++) {
    const size_t c = ((size_t)(s >> 3) * N + i) * 8 + (s & 7);
    sp.code[c] = (uint16_t)kStageCap;
    sp.seg[c] = 0;
  }
}

__device__ __forceinline__ char* weight_row(const ShardVramPtrs& vram, uint32_t N, uint32_t s) {
  return reinterpret_cast<char*>(vram.dendrite_weights + (size_t)s * N);
}
// Snapshot entry of axon a (< 2^28): a scalar base plus a 32-bit lane offset.
__device__ __forceinline__ uint4 snap_at(const uint4* __restrict__ snap, uint32_t a) {
  return *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(snap) + (a << 4));
}

// One chunk of 8 slots of soma i as loaded: staged, the plan's 8 entries
// (u16, one 16-byte load) and 8 segments (u8, one 8-byte load); otherwise
// the plan's 8 targets (segment << 24 | axon + 1, two 16-byte loads). Plus
// the 8 weights, two per register: PW, from the plan's working copy (one
// 16-byte load); otherwise from the 8 reference rows.
typedef uint32_t v4u __attribute__((ext_vector_type(4)));
// The most somas whose [16][N][8] plan arrays (16 B per chunk) stay within
// the 32-bit offsets of Chunk::load_buf.
constexpr uint32_t kBufMaxSomas = 1u << 24;
typedef uint32_t v2u __attribute__((ext_vector_type(2)));

__device__ __forceinline__ size_t chunk_at(uint32_t N, uint32_t c, uint32_t i) { return ((size_t)c * N + i) * 8; }

template <bool STAGED, bool PW>
struct Chunk {
  uint32_t k[STAGED ? 6 : 8];
  uint32_t w2[4];
  __device__ __forceinline__ void load(const ShardVramPtrs& vram, const StagePlan& sp, uint32_t N, uint32_t c,
                                       uint32_t i) {
    // Streamed once per tick: nontemporal, so that the byte snapshot and
    // the block sets the staging gathers read keep their places in L2.
    if (STAGED) {
      const v4u e = __builtin_nontemporal_load(
          reinterpret_cast<const v4u*>(reinterpret_cast<const char*>(sp.code + (size_t)c * N * 8) + i * 16u));
      const v2u g = __builtin_nontemporal_load(
          reinterpret_cast<const v2u*>(reinterpret_cast<const char*>(sp.seg + (size_t)c * N * 8) + i * 8u));
      k[0] = e.x; k[1] = e.y; k[2] = e.z; k[3] = e.w; k[4] = g.x; k[5] = g.y;
    } else {
      const v4u* q =
          reinterpret_cast<const v4u*>(reinterpret_cast<const char*>(sp.tgt + (size_t)c * N * 8) + i * 32u);
      const v4u t0 = __builtin_nontemporal_load(q), t1 = __builtin_nontemporal_load(q + 1);
      k[0] = t0.x; k[1] = t0.y; k[2] = t0.z; k[3] = t0.w; k[4] = t1.x; k[5] = t1.y; k[6] = t1.z; k[7] = t1.w;
    }
    if (PW) {
      const v4u wv = __builtin_nontemporal_load(
          reinterpret_cast<const v4u*>(reinterpret_cast<const char*>(sp.w + (size_t)c * N * 8) + i * 16u));
      w2[0] = wv.x; w2[1] = wv.y; w2[2] = wv.z; w2[3] = wv.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const char* r0 = reinterpret_cast<const char*>(vram.dendrite_weights + (size_t)(c * 8 + 2 * j) * N);
        const char* r1 = reinterpret_cast<const char*>(vram.dendrite_weights + (size_t)(c * 8 + 2 * j + 1) * N);
        w2[j] = (uint32_t)__builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(r0 + i * 2u)) |
                ((uint32_t)__builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(r1 + i * 2u)) << 16);
      }
    }
  }
  // load() of the staged form with the working copy (STAGED && PW) through
  // buffer loads: the chunk's row offset in a scalar register (c is the same
  // in every lane of the wave) and one 32-bit offset per lane, so that the
  // loop keeps no 64-bit address per lane and array. Every [16][N][8] array
  // must lie within 2^32 bytes (kBufMaxSomas).
  __device__ __forceinline__ void load_buf(const StagePlan& sp, uint32_t N, uint32_t c, uint32_t i) {
    static_assert(STAGED && PW, "the staged form with the working copy");
    constexpr int kFlags = 0x00020000, kNt = 2;  // raw buffer (gfx9 word 3); nontemporal
    const __amdgpu_buffer_rsrc_t rc = __builtin_amdgcn_make_buffer_rsrc(sp.code, (short)0, -1, kFlags);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(sp.seg, (short)0, -1, kFlags);
    const __amdgpu_buffer_rsrc_t rw = __builtin_amdgcn_make_buffer_rsrc(sp.w, (short)0, -1, kFlags);
    const uint32_t row8 = __builtin_amdgcn_readfirstlane(c * N * 8u);
    const v4u e = __builtin_bit_cast(v4u, __builtin_amdgcn_raw_buffer_load_b128(rc, i * 16u, row8 * 2u, kNt));
    const v2u g = __builtin_bit_cast(v2u, __builtin_amdgcn_raw_buffer_load_b64(rs, i * 8u, row8, kNt));
    const v4u wv = __builtin_bit_cast(v4u, __builtin_amdgcn_raw_buffer_load_b128(rw, i * 16u, row8 * 2u, kNt));
    k[0] = e.x; k[1] = e.y; k[2] = e.z; k[3] = e.w; k[4] = g.x; k[5] = g.y;
    w2[0] = wv.x; w2[1] = wv.y; w2[2] = wv.z; w2[3] = wv.w;
  }
  __device__ __forceinline__ uint32_t seg(int u) const {
    return STAGED ? ((k[4 + (u >> 2)] >> (8 * (u & 3))) & 0xFFu) : (k[u] >> 24);
  }
  // Staged: the block-local entry; otherwise the axon's snapshot index.
  __device__ __forceinline__ uint32_t entry(int u, uint32_t A) const {
    return STAGED ? ((k[u >> 1] >> (16 * (u & 1))) & 0xFFFFu) : snap_index(k[u], A);
  }
  __device__ __forceinline__ int32_t weight(int u) const {
    return (int32_t)(int16_t)(uint16_t)(w2[u >> 1] >> (16 * (u & 1)));
  }
};

typedef short axc_s2 __attribute__((ext_vector_type(2)));

// Distances of the eight heads of byte entry e from segment s, as u16 lanes
// (h - s; heads below s wrap to >= 65281), reduced to two lanes. S = s in
// both halves.
__device__ __forceinline__ uint32_t dist_pair_lanes(const uint2& e, uint32_t S) {
  const axc_us2 s = __builtin_bit_cast(axc_us2, S);
  const axc_us2 d0 = __builtin_bit_cast(axc_us2, __builtin_amdgcn_perm(0u, e.x, 0x0c010c00u)) - s;
  const axc_us2 d1 = __builtin_bit_cast(axc_us2, __builtin_amdgcn_perm(0u, e.x, 0x0c030c02u)) - s;
  const axc_us2 d2 = __builtin_bit_cast(axc_us2, __builtin_amdgcn_perm(0u, e.y, 0x0c010c00u)) - s;
  const axc_us2 d3 = __builtin_bit_cast(axc_us2, __builtin_amdgcn_perm(0u, e.y, 0x0c030c02u)) - s;
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_elementwise_min(d0, d1),
                                                                __builtin_elementwise_min(d2, d3)));
}
// [min of ru, min of rv] from two dist_pair_lanes results.
__device__ __forceinline__ uint32_t pair_min(uint32_t ru, uint32_t rv) {
  const axc_us2 lo = __builtin_bit_cast(axc_us2, __builtin_amdgcn_perm(rv, ru, 0x05040100u));
  const axc_us2 hi = __builtin_bit_cast(axc_us2, __builtin_amdgcn_perm(rv, ru, 0x07060302u));
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(lo, hi));
}
// [max of ru, max of rv]: pair_min with the larger lane.
__device__ __forceinline__ uint32_t pair_max(uint32_t ru, uint32_t rv) {
  const axc_us2 lo = __builtin_bit_cast(axc_us2, __builtin_amdgcn_perm(rv, ru, 0x05040100u));
  const axc_us2 hi = __builtin_bit_cast(axc_us2, __builtin_amdgcn_perm(rv, ru, 0x07060302u));
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(lo, hi));
}
// Largest of the eight bytes of (x, y).
__device__ __forceinline__ uint32_t max_byte8(uint32_t x, uint32_t y) {
  const axc_us2 a = __builtin_bit_cast(axc_us2, x & 0x00FF00FFu), b = __builtin_bit_cast(axc_us2, (x >> 8) & 0x00FF00FFu);
  const axc_us2 c = __builtin_bit_cast(axc_us2, y & 0x00FF00FFu), d = __builtin_bit_cast(axc_us2, (y >> 8) & 0x00FF00FFu);
  const axc_us2 m = __builtin_elementwise_max(__builtin_elementwise_max(a, b), __builtin_elementwise_max(c, d));
  return max((uint32_t)m.x, (uint32_t)m.y);
}

// Hit test and weighted sum of one loaded chunk c (slots 8c .. 8c + 7) of
// soma i. The plan gives the slots of the last chunk at or past the fill the
// inert entry and segment 0, so every slot of a chunk is tested alike.
// Slots are taken in pairs: the two minimum distances in one register, the
// hit flags (distance < length) as two u16 lanes, and their weights added
// with one 2-lane dot product. hit_cnt counts the hits in its two lanes.
template <bool STAGED, bool PW>
__device__ __forceinline__ void hit_chunk(const Chunk<STAGED, PW>& now, uint32_t c, uint32_t N, uint32_t A,
                                          const StagePlan& sp, const uint2* ent,
                                          const AxcBurstHeads8* __restrict__ heads,
                                          uint32_t i, uint32_t prop, int32_t& i_in, axc_us2& hit_cnt) {
  constexpr int U = 8;
  const axc_us2 P = {(unsigned short)prop, (unsigned short)prop};
  const axc_us2 kHalf = {0x7FFF, 0x7FFF}, k15 = {15, 15};
  // Segment u of the chunk in both u16 halves.
  uint32_t S[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    if (STAGED) S[u] = __builtin_amdgcn_perm(0u, now.k[4 + (u >> 2)], 0x0c000c00u | ((u & 3) * 0x00010001u));
    else S[u] = __builtin_amdgcn_perm(0u, now.k[u], 0x0c030c03u);
  }
  uint32_t M[U / 2];
  constexpr int H = 4;  // entries read in two halves of four
#pragma unroll
  for (int h0 = 0; h0 < U; h0 += H) {
    uint2 e[H];
#pragma unroll
    for (int u = 0; u < H; u++) {
      if (STAGED) e[u] = ent[now.entry(h0 + u, A)];
      else e[u] = sp.comp8[now.entry(h0 + u, A)];
    }
#pragma unroll
    for (int u = 0; u < H; u += 2)
      M[(h0 + u) / 2] = pair_min(dist_pair_lanes(e[u], S[h0 + u]), dist_pair_lanes(e[u + 1], S[h0 + u + 1]));
  }
  // Segment + length beyond the byte form (rare), or a spilled slot: the
  // slot's exact distance from its target (segment included) and the heads.
  uint32_t mx;
  if (STAGED) {
    mx = max_byte8(now.k[4], now.k[5]);
  } else {
    mx = 0;
#pragma unroll
    for (int u = 0; u < U; u++) mx = max(mx, now.k[u] >> 24);
  }
  if (mx + prop > 255u) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t sg = S[u] & 0xFFu;
      if (sg + prop > 255u) {
        uint32_t t;
        if constexpr (STAGED)
          t = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(sp.tgt + (size_t)c * N * 8) +
                                                 i * 32u + u * 4u);
        else
          t = now.k[u];
        // A slot at or past the fill has the empty target: no hit.
        const uint32_t d = min(t ? min_dist_heads(heads, A, snap_index(t, A), t >> 24) : 0xFFFFu, 0xFFFFu);
        const uint32_t sh = 16 * (u & 1);
        M[u / 2] = (M[u / 2] & ~(0xFFFFu << sh)) | (d << sh);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < U / 2; j++) {
    // 1 in the u16 lanes whose distance is below the length.
    const axc_us2 gap = __builtin_elementwise_sub_sat(P, __builtin_bit_cast(axc_us2, M[j]));
    const axc_us2 h = (gap + kHalf) >> k15;
    i_in = __builtin_amdgcn_sdot2(__builtin_bit_cast(axc_s2, h), __builtin_bit_cast(axc_s2, now.w2[j]), i_in, false);
    hit_cnt += h;
  }
}

// The hit sum of soma i over the slots of group gu: chunks of 8 slots
// c = (gu + 3) % 4 + 4k below the soma's filled prefix F, one at a time.
// (Loading each lane's first chunk before the staging gathers, or all of a
// lane's chunks at once in a one-workgroup-per-CU pipelined pass with two
// LDS entry buffers, measured slower.)
template <bool STAGED, bool PW>
__device__ __forceinline__ void scan_slots(const ShardVramPtrs& vram, uint32_t N, uint32_t A, const StagePlan& sp,
                                           const uint2* ent, uint32_t gu, uint32_t i,
                                           uint32_t F, uint32_t prop, int32_t& i_in, uint32_t& hits) {
  // Group 0, which finishes the somas afterwards, takes chunks 3, 7, 11, ...
  // so that a partly filled last chunk lands on another group.
  const uint32_t cg = (gu + kStageGroups - 1) % kStageGroups;
  axc_us2 hit_cnt = {0, 0};
#pragma unroll 1
  for (uint32_t c = cg; c * 8 < F; c += kStageGroups) {
    Chunk<STAGED, PW> now;
    now.load(vram, sp, N, c, i);
    hit_chunk<STAGED, PW>(now, c, N, A, sp, ent, vram.axon_heads, i, prop, i_in, hit_cnt);
  }
  hits += (uint32_t)hit_cnt.x + (uint32_t)hit_cnt.y;
}

// Loads the block's byte snapshot entries into ent[0, cnt); ent[kStageCap]
// is the inert entry. Eight independent loads in flight per lane.
__device__ __forceinline__ void stage_entries(uint2* ent, const uint32_t* __restrict__ bset, uint32_t cnt,
                                              const uint2* __restrict__ comp8, uint32_t A, uint32_t tid) {
  constexpr uint32_t SU = 8;
  for (uint32_t j0 = 0; j0 < cnt; j0 += kStageThreads * SU) {
    uint32_t ix[SU];
    uint2 v[SU];
#pragma unroll
    for (uint32_t k = 0; k < SU; k++) {
      // Unconditional loads (entry 0 stands in past the end), so that all
      // eight are in flight at once.
      const uint32_t j = j0 + k * kStageThreads + tid;
      const uint32_t a = bset[j < cnt ? j : 0u];
      ix[k] = j < cnt ? a : A;
    }
#pragma unroll
    for (uint32_t k = 0; k < SU; k++) v[k] = comp8[ix[k]];
#pragma unroll
    for (uint32_t k = 0; k < SU; k++) {
      const uint32_t j = j0 + k * kStageThreads + tid;
      if (j < cnt) ent[j] = v[k];
    }
  }
  if (tid == 0) ent[kStageCap] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
}

// Workgroup scratch of one soma block's finish.
struct BlockLds {
  int32_t pre_v[kBlockSomas], pre_thr[kBlockSomas];  // group 0's somas: voltage, threshold offset
  uint32_t pre_ax[kBlockSomas];                      // and axon, read before the scan
  int32_t red_in[kStageGroups][kBlockSomas];
  uint16_t red_hits[kStageGroups][kBlockSomas];
  uint32_t red[kGroup0Waves][8];  // group 0's waves' counters
  uint32_t done;                  // group 0's waves that have written theirs
};

// The end of a block's tick, after the scan: the groups' hit sums are
// added and group 0 finishes its somas (GLIF update, flags), each of its
// waves then storing its spikes as one mask word (sc.spk_mask, turned into
// the spike list by compact_spikes_kernel); the block's counters go to part[] with atomics
// whose results nobody waits for. The other groups leave after the barrier.
// type_spikes (may be null) counts spikes per variant type; count_gsop adds
// the slot and target counts of the tick's GSOP (for callers whose GSOP
// kernel does not count).
__device__ __forceinline__ void finish_block(BlockLds& L, const ShardVramPtrs& vram, uint32_t N, uint32_t A,
                                             const AxcVariantParameters& p, uint32_t tick, const ShardScratch& sc,
                                             uint32_t tid, uint32_t i, uint32_t F, uint8_t flags_in, uint8_t timer,
                                             bool refr, int32_t i_in, uint32_t hits, uint64_t* part,
                                             bool count_tick, uint64_t* type_spikes, bool count_gsop,
                                             uint32_t hb_i) {
  const uint32_t g = tid / kBlockSomas, li = tid % kBlockSomas;
  L.red_in[g][li] = i_in;
  L.red_hits[g][li] = (uint16_t)hits;
  __syncthreads();
  if (g != 0) return;  // whole waves

  bool fin = false, push_own = false;
  uint32_t my_axon = 0;
  uint32_t n_int = 0, n_scan = 0, n_tread = 0, n_hits = 0, n_gslots = 0, n_gtgt = 0;
  if (i < N) {
    int32_t sum = 0;
    uint32_t hsum = 0;
#pragma unroll
    for (uint32_t k = 0; k < kStageGroups; k++) {
      sum = wadd(sum, L.red_in[k][li]);
      hsum += L.red_hits[k][li];
    }
    if (!refr) {
      n_scan = F;
      n_tread = F < kMaxDendrites ? F + 1 : kMaxDendrites;
      n_hits = hsum;
    }
    uint8_t flags = flags_in & 0xF1;  // K1
    flags &= (uint8_t)~0x01;
    soma_tick_with(vram, i, A, p, flags, timer, refr, tick, sum, L.pre_thr[li], L.pre_v[li], L.pre_ax[li], fin,
                   push_own, my_axon, hb_i);
    n_int = refr ? 0u : 1u;
    if (fin && type_spikes) atomicAdd((unsigned long long*)&type_spikes[flags_in >> 4], 1ull);
    if (fin && count_gsop) {
      n_gslots = F;
      n_gtgt = F < kMaxDendrites ? F + 1 : kMaxDendrites;
    }
  }
  // The push is counted here and applied by push_spikers_kernel, which walks
  // the spike list (no owner list: nothing waits for this atomic).
  if (push_own) atomicAdd(&sc.push_cnt[my_axon], 1u);

  // This wave's spikers, as one mask word (zero included, so that no bit of
  // an earlier tick stays); compact_spikes_kernel turns the words into the
  // spike list. (Appending here with a returning atomic kept the block's LDS
  // held for the atomic's round trip: 13.7 us of a 250 us pass.)
  const uint32_t lane = tid & 63u;
  const uint64_t bs = __ballot(fin);
  if (lane == 0) sc.spk_mask[i >> 6] = bs;
  const uint32_t cs = (uint32_t)__popcll(bs);
  // This wave's counters; the last of group 0's waves to get here adds the
  // block's to part[] (one set of atomics per block, no barrier).
  uint32_t r[7] = {n_int, n_scan, n_tread, n_hits, n_gslots, n_gtgt, cs};
#pragma unroll
  for (int k = 0; k < 6; k++) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) r[k] += __shfl_xor(r[k], off, 64);
  }
  if (lane == 0) {
    const uint32_t wv = tid >> 6;
#pragma unroll
    for (int k = 0; k < 7; k++) L.red[wv][k] = r[k];
    const uint32_t before =
        __hip_atomic_fetch_add(&L.done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (before == kGroup0Waves - 1) {
      const uint32_t slot[7] = {kStIntegrated, kStScanned, kStTargets, kStHits, kStGsopSlots, kStGsopTargets,
                                kStSpikes};
#pragma unroll
      for (int k = 0; k < 7; k++) {
        uint32_t v = 0;
#pragma unroll
        for (uint32_t w = 0; w < kGroup0Waves; w++) v += L.red[w][k];
        if (v) atomicAdd((unsigned long long*)&part[slot[k]], (unsigned long long)v);
      }
      if (count_tick) atomicAdd((unsigned long long*)&part[kStTicks], 1ull);
    }
  }
}

// K1 + K5 of one tick for soma block lb (256 somas) of one shard, by a
// workgroup of kStageThreads lanes (see the file comment). `lut` is the
// shard's table, already in LDS (visible after the first barrier here).
// The tick's GSOP (K6) is left to gsop_kernel, after the pushes. PW: the
// weights are read from the plan's working copy. See finish_block for the
// counters.
template <bool PW>
__device__ __forceinline__ void staged_block(const ShardVramPtrs& vram, uint32_t N, uint32_t A,
                                             const AxcVariantParameters* lut, uint32_t tick, const ShardScratch& sc,
                                             const StagePlan& sp, uint32_t lb,
                                             uint64_t* part, bool count_tick, uint64_t* type_spikes,
                                             bool count_gsop = false) {
  extern __shared__ uint2 ent[];  // kStageCap + 1 entries
  __shared__ BlockLds L;
  const uint32_t tid = threadIdx.x;
  const uint32_t cnt = sp.cnt[lb];
  const bool staged = cnt != kStageOverflow;
  const uint32_t* __restrict__ bset = sp.set + (size_t)lb * kStageCap;
  const uint32_t g = tid / kBlockSomas, li = tid % kBlockSomas;
  const uint32_t i = lb * kBlockSomas + li;
  // The somas' bytes are requested before the staging gathers, which hide
  // their latency.
  uint32_t F = 0;
  uint8_t flags_in = 0, timer = 0;
  int32_t v_in = 0, thr_in = 0;
  uint32_t ax_in = 0, hb_i = i;
  if (i < N) {
    flags_in = vram.soma_flags[i];
    timer = vram.timers[i];
    F = sp.fill[i];
    if (g == 0) {  // group 0 finishes the somas: their state, parked in LDS over the scan
      v_in = vram.soma_voltage[i];
      thr_in = vram.threshold_offset[i];
      ax_in = vram.soma_to_axon[i];
      if (sp.perm) hb_i = sp.perm[i];  // the caller's index (heartbeat phase)
    }
  }
  if (staged) stage_entries(ent, bset, cnt, sp.comp8, A, tid);
  if (g == 0) {
    L.pre_v[li] = v_in;
    L.pre_thr[li] = thr_in;
    L.pre_ax[li] = ax_in;
  }
  if (tid == 0) L.done = 0;
  __syncthreads();

  int32_t i_in = 0;
  uint32_t hits = 0;
  const bool refr = i >= N || timer > 0;
  const AxcVariantParameters& p = lut[(flags_in >> 4) & 0x0F];
  // The group index is the same for the 64 lanes of a wave: scalar.
  const uint32_t gu = __builtin_amdgcn_readfirstlane(g);
  if (!refr) {
    if (staged) scan_slots<true, PW>(vram, N, A, sp, ent, gu, i, F, p.signal_propagation_length, i_in, hits);
    else scan_slots<false, PW>(vram, N, A, sp, ent, gu, i, F, p.signal_propagation_length, i_in, hits);
  }
  finish_block(L, vram, N, A, p, tick, sc, tid, i, F, flags_in, timer, refr, i_in, hits, part, count_tick,
               type_spikes, count_gsop, hb_i);
}

// The single-shard soma pass: one workgroup per soma block, blocks dealt to
// XCDs in contiguous eighths (see neuron_tick_kernel). Eight waves per SIMD:
// 2048 / kStageThreads workgroups per CU.
template <bool PW>
__global__ __launch_bounds__(kStageThreads) __attribute__((amdgpu_waves_per_eu(8, 8)))
void neuron_staged_kernel(ShardVramPtrs vram, uint32_t N, uint32_t A, const AxcVariantParameters* __restrict__ lut_g,
                          const TickParams* __restrict__ prm, uint32_t tick_arg, ShardScratch sc,
                          StagePlan sp) {
  __shared__ AxcVariantParameters lut[16];
  const uint32_t per_xcd = gridDim.x / kNumXcd;
  const uint32_t lb = (blockIdx.x % kNumXcd) * per_xcd + blockIdx.x / kNumXcd;
  if (threadIdx.x < 64) reinterpret_cast<uint4*>(lut)[threadIdx.x] = reinterpret_cast<const uint4*>(lut_g)[threadIdx.x];
  staged_block<PW>(vram, N, A, lut, tick_of(prm, tick_arg), sc, sp, lb,
                   sc.stats_part + (size_t)blockIdx.x * kStatsPerBlock, blockIdx.x == 0, nullptr);
}

// The working copy of the weights: soma i's chunks c with 8c < fill[i], to
// (IN) or from the reference rows. Slots of such a chunk at or past the fill
// are copied as they are (no pass writes them). One lane per soma, chunks in
// turn: each reference row is read or written by consecutive lanes.
// i: the plan's soma, r: the same soma in the caller's numbering (its rows).
template <bool IN>
__device__ __forceinline__ void copy_weights_soma(int16_t* __restrict__ rows, uint32_t N, const StagePlan& sp,
                                                  uint32_t i, uint32_t r) {
  const uint32_t F = sp.fill[i];
  for (uint32_t c = 0; c * 8 < F; c++) {
    v4u* q = reinterpret_cast<v4u*>(sp.w + chunk_at(N, c, i));
    if (IN) {
      uint32_t w2[4];
#pragma unroll
      for (int j = 0; j < 4; j++)
        w2[j] = (uint32_t)(uint16_t)rows[(size_t)(c * 8 + 2 * j) * N + r] |
                ((uint32_t)(uint16_t)rows[(size_t)(c * 8 + 2 * j + 1) * N + r] << 16);
      *q = v4u{w2[0], w2[1], w2[2], w2[3]};
    } else {
      const v4u v = *q;
      const uint32_t w2[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int j = 0; j < 4; j++) {
        rows[(size_t)(c * 8 + 2 * j) * N + r] = (int16_t)(uint16_t)(w2[j] & 0xFFFFu);
        rows[(size_t)(c * 8 + 2 * j + 1) * N + r] = (int16_t)(uint16_t)(w2[j] >> 16);
      }
    }
  }
}

// One lane per soma of the caller's numbering, so that the rows are read
// (IN) or written coalesced; in the engine's order each lane's chunks are
// those of the plan's soma perm_inv[r].
template <bool IN>
__global__ __launch_bounds__(256) void copy_weights_kernel(int16_t* __restrict__ rows, uint32_t N, StagePlan sp) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r < N) copy_weights_soma<IN>(rows, N, sp, sp.perm_inv ? sp.perm_inv[r] : r, r);
}

// Per virtual input axon (ids [virt_off, virt_off + n_virt)) the first of
// its postsynaptic somas in the engine's order, for ordering the inputs like
// the somas they reach (key[k] starts at ~0).
__global__ __launch_bounds__(256) void input_key_kernel(const uint32_t* __restrict__ targets, uint32_t N,
                                                        const uint32_t* __restrict__ perm_inv, uint32_t virt_off,
                                                        uint32_t n_virt, uint32_t* __restrict__ key) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= N) return;
  const uint32_t e = perm_inv[r];
  for (uint32_t s = 0; s < kMaxDendrites; s++) {
    const uint32_t t = targets[(size_t)s * N + r];
    if (t == 0) break;
    const uint32_t k = (t & 0x00FFFFFFu) - 1u - virt_off;
    if (k < n_virt) atomicMin(&key[k], e);
  }
}

// GSOP (K6, physics.cu:281-373) of one chunk c of spiking soma i on the
// working copy: 8 slots from one 16-byte weight load and two 16-byte target
// loads; the heads from the post-push byte snapshot (comp8), or the compact
// one (comp) where segment + length > 255. F: the soma's fill (> 8c).
// Heads inside a window of ticks (day_window.hip): an axon whose record in
// aux is not current (stamp) was not pushed since the window started, and
// its u32 heads are still the window's start heads; its current ones are
// those advanced `steps` ticks. aux null: the heads are current.
struct LazyHeads {
  const uint64_t* aux;
  uint32_t stamp, steps, v_seg;
};
__device__ __forceinline__ uint32_t lazy_adv(uint32_t x, uint32_t v, uint32_t n) {
#pragma unroll
  for (uint32_t k = 0; k < 8; k++)
    if (k < n && x != kSentinel) x += v;
  return x;
}
// min_dist_heads on the current heads of axon a.
__device__ __forceinline__ uint32_t min_dist_lazy(const AxcBurstHeads8* __restrict__ heads, uint32_t A, uint32_t a,
                                                  uint32_t seg, const LazyHeads& lz) {
  if (a >= A) return 0xFFFFFFFFu;
  const Heads8 h = load_heads(heads, a);
  const uint32_t n = (lz.aux && (uint32_t)lz.aux[a] != lz.stamp) ? lz.steps : 0u;
  const uint32_t hv[8] = {h.lo.x, h.lo.y, h.lo.z, h.lo.w, h.hi.x, h.hi.y, h.hi.z, h.hi.w};
  uint32_t m = 0xFFFFFFFFu;
#pragma unroll
  for (int x = 0; x < 8; x++) m = min(m, lazy_adv(hv[x], lz.v_seg, n) - seg);
  return m;
}
__device__ __forceinline__ void gsop_chunk(uint32_t N, uint32_t A, const AxcVariantParameters* lut, int32_t dop,
                                           const uint8_t* __restrict__ flags_all, const StagePlan& sp,
                                           const AxcBurstHeads8* __restrict__ heads, uint32_t i, uint32_t c,
                                           uint32_t F, const LazyHeads lz = LazyHeads{nullptr, 0, 0, 0}) {
  const uint8_t flags = flags_all[i];
  const AxcVariantParameters& p = lut[(flags >> 4) & 0x0F];
  const GsopTerms g = gsop_terms(flags, p, dop);
  const uint32_t nv = F - c * 8 < 8u ? F - c * 8 : 8u;
  const v4u* tq = reinterpret_cast<const v4u*>(sp.tgt + chunk_at(N, c, i));
  const v4u t0 = tq[0], t1 = tq[1];
  const uint32_t t[8] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w};
  v4u* wq = reinterpret_cast<v4u*>(sp.w + chunk_at(N, c, i));
  const v4u wv = *wq;
  uint32_t w2[4] = {wv.x, wv.y, wv.z, wv.w};
  uint32_t d[8];
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const uint32_t a = (uint32_t)u < nv ? snap_index(t[u], A) : A, seg = t[u] >> 24;
    d[u] = seg + g.len > 255u ? min_dist_lazy(heads, A, a, seg, lz) : min_dist8(sp.comp8[a], seg);
  }
#pragma unroll
  for (int u = 0; u < 8; u++) {
    if ((uint32_t)u >= nv) continue;
    const uint32_t sh = 16 * (u & 1);
    const int16_t w = (int16_t)(uint16_t)(w2[u >> 1] >> sh);
    const int16_t nw = gsop_weight(w, d[u], g.len, g.raw_pot, g.final_dep, g.bm, p.inertia_curve);
    w2[u >> 1] = (w2[u >> 1] & ~(0xFFFFu << sh)) | ((uint32_t)(uint16_t)nw << sh);
  }
  *wq = v4u{w2[0], w2[1], w2[2], w2[3]};
}

// GSOP of a tick on the working copy: lane group of kGsopLanes per spiking
// soma, lane j of a group takes chunk j (the chunks below the soma's fill).
constexpr uint32_t kGsopLanes = kMaxDendrites / 8;
__global__ __launch_bounds__(256) void gsop_chunks_kernel(ShardVramPtrs vram, uint32_t N, uint32_t A,
                                                          const AxcVariantParameters* __restrict__ lut_g,
                                                          const TickParams* __restrict__ prm, int32_t dop_arg,
                                                          ShardScratch sc, StagePlan sp) {
  __shared__ AxcVariantParameters lut[16];
  __shared__ uint32_t red[4][2];
  const uint32_t tid = threadIdx.x;
  if (tid < 64) reinterpret_cast<uint4*>(lut)[tid] = reinterpret_cast<const uint4*>(lut_g)[tid];
  __syncthreads();
  const uint32_t n = sc.counters[kCntSpikes] * kGsopLanes;
  const int32_t dop = dopamine_of(prm, dop_arg);
  uint32_t cnt_slots = 0, cnt_tgt = 0;
  for (uint32_t q = blockIdx.x * 256 + tid; q < n; q += gridDim.x * 256) {
    const uint32_t i = sc.spk_list[q / kGsopLanes], c = q % kGsopLanes;
    const uint32_t F = sp.fill[i];
    if (c * 8 < F) gsop_chunk(N, A, lut, dop, vram.soma_flags, sp, vram.axon_heads, i, c, F);
    if (c == 0) {
      cnt_slots += F;
      cnt_tgt += F < kMaxDendrites ? F + 1 : kMaxDendrites;
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

// The tick's spike list from the staged pass's mask words (bit l of word w:
// soma 64w + l): one returning atomic per workgroup of 256 words reserves
// its range, and each lane writes its word's somas in ascending order.
// (One atomic per wave, all on one counter, measured 6.5 us at 1.5M somas.)
__global__ __launch_bounds__(256) void compact_spikes_kernel(uint32_t n_words, ShardScratch sc) {
  __shared__ uint32_t wsum[4], base;
  const uint32_t w = blockIdx.x * 256 + threadIdx.x, lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint64_t m = w < n_words ? sc.spk_mask[w] : 0ull;
  const uint32_t c = (uint32_t)__popcll(m);
  uint32_t x = c;  // inclusive prefix over the wave
#pragma unroll
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) wsum[wv] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    base = total ? atomicAdd(&sc.counters[kCntSpikes], total) : 0u;
  }
  __syncthreads();
  uint32_t pos = base + x - c;
  for (uint32_t k = 0; k < wv; k++) pos += wsum[k];
  for (uint64_t r = m; r; r &= r - 1) sc.spk_list[pos++] = w * 64u + (uint32_t)__builtin_ctzll(r);
}

// The engine's soma order in and out of its private soma arrays (sh) from
// and to the caller's (ref): soma i of the engine is the caller's perm[i].
// IN also derives the private soma_to_axon in the engine's axon ids.
template <bool IN>
__global__ __launch_bounds__(256) void order_somas_kernel(ShardVramPtrs ref, ShardVramPtrs sh, uint32_t N, uint32_t A,
                                                          const uint32_t* __restrict__ perm,
                                                          const uint32_t* __restrict__ axon_map) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  const uint32_t r = perm[i];
  if (IN) {
    sh.soma_voltage[i] = ref.soma_voltage[r];
    sh.soma_flags[i] = ref.soma_flags[r];
    sh.threshold_offset[i] = ref.threshold_offset[r];
    sh.timers[i] = ref.timers[r];
    const uint32_t a = ref.soma_to_axon[r];
    sh.soma_to_axon[i] = a < A ? axon_map[a] : a;
  } else {
    ref.soma_voltage[r] = sh.soma_voltage[i];
    ref.soma_flags[r] = sh.soma_flags[i];
    ref.threshold_offset[r] = sh.threshold_offset[i];
    ref.timers[r] = sh.timers[i];
  }
}

// The heads in and out of the engine's axon order: the caller's axon a is
// the engine's axon_map[a].
template <bool IN>
__global__ __launch_bounds__(256) void order_heads_kernel(AxcBurstHeads8* __restrict__ ref, AxcBurstHeads8* __restrict__ sh,
                                                          uint32_t A, const uint32_t* __restrict__ axon_map) {
  const uint32_t a = blockIdx.x * 256 + threadIdx.x;
  if (a >= A) return;
  const uint32_t e = axon_map[a];
  const uint4* src = reinterpret_cast<const uint4*>(IN ? ref + a : sh + e);
  uint4* dst = reinterpret_cast<uint4*>(IN ? sh + e : ref + a);
  const uint4 x0 = src[0], x1 = src[1];
  dst[0] = x0;
  dst[1] = x1;
}

}  // namespace axc
