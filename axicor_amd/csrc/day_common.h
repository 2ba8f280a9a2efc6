// day_common.h — shared definitions of the Day-Phase kernels and host code.
#pragma once
#include <stdint.h>
#include <stddef.h>
#include "../../include/axicor_day.h"

namespace axc {

constexpr uint32_t kMaxDendrites = AXC_MAX_DENDRITES;
constexpr uint32_t kSentinel = AXC_AXON_SENTINEL;
constexpr uint32_t kNoAxon = AXC_NO_AXON;

// Per-tick list counters (device, reset at the start of every tick).
constexpr uint32_t kCntSpikes = 0;  // entries of spk_list
constexpr uint32_t kCntOwners = 1;  // entries of own_list
constexpr uint32_t kNumCounters = 4;

// Per-workgroup partial counters (device, accumulated across ticks).
constexpr uint32_t kStatsPerBlock = 8;
constexpr uint32_t kStIntegrated = 0;
constexpr uint32_t kStScanned = 1;
constexpr uint32_t kStTargets = 2;
constexpr uint32_t kStHits = 3;
constexpr uint32_t kStSpikes = 4;
constexpr uint32_t kStGsopSlots = 5;
constexpr uint32_t kStGsopTargets = 6;
constexpr uint32_t kStTicks = 7;

// Compact head snapshot: per axon the eight heads as u16, each min(h, 0xFFFF)
// (one 16 B load). For every segment s <= 255 and length p <= 255 the tests
// of the reference give the same answer on the snapshot as on the u32 heads:
// (h - s) mod 2^32 < p  <=>  (c - s) mod 2^16 < p, and the smallest such
// distance is the same, because any h >= 510 fails the test either way and a
// c < s wraps to >= 2^16 - 255 in u16 as h < s does in u32. Entry A (one past
// the last axon) is all 0xFFFF and stands for an axon with no live head.
constexpr uint32_t kInertHead = 0xFFFFu;

// Dendrite scan shape of neuron_tick (slots per step, next-step prefetch).
#ifndef AXC_SCAN_UNROLL
#define AXC_SCAN_UNROLL 8
#endif
#ifndef AXC_SCAN_PIPELINE
#define AXC_SCAN_PIPELINE 0
#endif
#ifndef AXC_NEURON_ATTR
#define AXC_NEURON_ATTR
#endif
constexpr int kScanUnroll = AXC_SCAN_UNROLL;
constexpr bool kScanPipeline = AXC_SCAN_PIPELINE != 0;
static_assert(AXC_MAX_DENDRITES % AXC_SCAN_UNROLL == 0, "the scan steps must tile the 128 slots");

// XCDs of an MI355X; workgroups are dealt to them round-robin.
constexpr uint32_t kNumXcd = 8;

// Workgroups of the list-driven kernels (gsop, push_apply): fixed so that a
// HIP graph of a batch does not depend on the number of spikes.
constexpr uint32_t kListBlocks = 2048;

// Per-batch values of a graph-replayed batch, written by a one-lane kernel
// before each replay: kernels add their tick offset t to tick_base and index
// the per-tick input/output slices at base + t*stride. A null TickParams
// pointer means "take the kernel arguments as they are" (the per-tick entry
// point).
struct TickParams {
  uint32_t tick_base;
  int32_t dopamine;
  const uint32_t* bitmask;  // [T][words] input bitmasks, or null
  const uint32_t* spikes;   // [T][max_spikes] spike schedule, or null
  uint8_t* out;             // [T][n_out] output history, or null
  const uint32_t* counts;   // [T] entries of each tick's schedule slice (device), or null
};

// Engine-private per-shard device buffers (no counterpart in the reference).
struct ShardScratch {
  uint32_t* spk_list;    // [N] somas that spiked this tick
  uint64_t* spk_mask;    // [ceil(N/256)*4] the staged pass's spikes, bit l of word w = soma 64w + l
  uint32_t* own_list;    // [A] axons receiving pushes this tick (first owner)
  uint32_t* push_cnt;    // [A] pending pushes per axon, zero between uses
  uint32_t* counters;    // [kNumCounters]
  uint64_t* stats_part;  // [ceil(N/256)][kStatsPerBlock] neuron_tick partials
  uint64_t* stats_part2; // [kListBlocks][kStatsPerBlock] gsop partials (slots, targets)
  uint4* comp[2];        // [A+1] compact head snapshots, used alternately tick by tick
};

// Per-block staging plan of the dendrite scan (engine-private, rebuilt from
// the reference's dendrite_targets whenever they change). Soma block b
// (somas [256b, 256b+256)) reads the byte snapshot entries of the distinct
// presynaptic axons of its filled slots from LDS: set[b*kStageCap + j] is the
// axon of block-local entry j (A for targets outside [0, A)).
#ifndef AXC_STAGE_CAP
#define AXC_STAGE_CAP 8704
#endif
constexpr uint32_t kStageCap = AXC_STAGE_CAP;    // LDS entries per block (8 B each)
constexpr uint32_t kStageOverflow = 0xFFFFFFFFu; // cnt of a block whose set does not fit
struct StagePlan {
  uint16_t* code;  // [16][N][8] block-local entry of each filled slot's axon: slot s of soma i at
                   // ((s / 8) * N + i) * 8 + s % 8, so one soma's chunk of 8 slots is 16 contiguous bytes
  uint8_t* seg;    // [16][N][8] segment of each filled slot (target >> 24), same order
  uint8_t* fill;   // [N] filled prefix: slots before the first empty target
  uint32_t* set;   // [blocks][kStageCap]
  uint32_t* cnt;   // [blocks] entries, or kStageOverflow
  uint2* comp8;    // [A+1] byte snapshot: per axon 8 x min(h, 255); entry A inert
  uint32_t* tgt;   // [16][N][8] targets of the filled slots (reference format), same order
  int16_t* w;      // [16][N][8] weights of the filled slots' chunks, same order: the working copy of
                   // the batch graphs, loaded from the reference rows at a batch's start and stored
                   // back at its end
  // The engine's own soma order (null: the caller's). Soma i of the plan is
  // the caller's soma perm[i]; axon a of the caller is the plan's axon
  // axon_map[a] (its targets and sets are in those ids).
  const uint32_t* perm;
  const uint32_t* perm_inv;
  const uint32_t* axon_map;
};

// Windows of ticks of the batch graphs (day_window.hip): the slots are
// streamed once per window of up to kWin ticks.
#ifndef AXC_WIN_TICKS
#define AXC_WIN_TICKS 4
#endif
constexpr uint32_t kWin = AXC_WIN_TICKS;

// Per axon and window (aux[a], valid while its low 32 bits equal the
// window's stamp ctl[1]): bits 32-35 the pushes since the window started
// (at most 8: the start heads shifted out), bit 36 a start head near the
// u32 wrap, bits 37-49 and 50-62 the least and greatest window-start value
// of the live heads shifted out so far (8191 and 0 for none; values >= 8191
// are left out: they hit no slot). Each tick's corrections are pieces of up
// to 16 reverse entries, start | axon << 32 | (count - 1) << 56, at
// pieces[k], k < piece_cnt[tick & 1]; the pool holds the most pieces all
// axons can make in a tick (an axon's range makes at most ceil(its list
// length / 16)).
constexpr uint32_t kNoDrop = 8191;
// The axons whose record is current in a window (pushed since it started,
// or marked near the wrap) are listed once each in act, by the parity of
// the window's stamp: act[(stamp & 1) * A + k], k < ctl[2 + (stamp & 1)].
struct DirtyMarks {
  uint64_t* aux;        // [A]
  uint32_t* ctl;        // [4]: (unused), stamp, active counts of even and odd stamps
  uint64_t* pieces;     // the pool
  uint32_t* piece_cnt;  // [2] pieces of even and odd ticks
  uint32_t* act;        // [2][A]
};

struct WinState {
  int64_t* acc;               // [kWin][N] per soma and tick of the window: hits << 32 plus the hit weights' sum
  uint2* comp8_w0;            // [A+1] byte snapshot of the window's start heads (sp.comp8 or comp8_alt, by
                              // window); entry A inert
  uint2* comp8_alt;           // [A+1] the other byte snapshot buffer
  AxcBurstHeads8* heads_w0;   // [A] the window's start heads, written for the axons the window pushes (or
                              // that start near the wrap) at their first push
  uint32_t* rev_off;          // [A+1] the reverse index: axon a's slots at rev[rev_off[a], rev_off[a+1]),
                              // sorted by segment
  uint64_t* rev;              // soma | slot << 32 | segment << 40
  uint4* rev_cut;             // [A] per segment s < kRevCuts the end (u16, relative to rev_off[a]) of axon
                              // a's entries with a segment <= s; 0xFFFF past the u16 range
  uint64_t* cnt;              // [kWinCounters] work counters, accumulated across batches
  uint64_t* cpart;            // [kListBlocks][2] the correction kernel's per-workgroup pieces and entries
};
// Window work counters: windows, somas scanned by the window passes, their
// filled slots (then, from cpart: correction pieces and reverse entries
// tested, summed over ticks).
constexpr uint32_t kWinCounters = 5;
constexpr uint32_t kRevCuts = 8;

}  // namespace axc
