/*
 * axicor_day.h — C-ABI of the MI355X Day-Phase engine (libaxicor_day.so).
 *
 * Drop-in for the `extern "C"` block that genesis-compute's Rust crate binds
 * (genesis-compute/src/ffi.rs:48-273).  Every symbol of that block is exported
 * with the same name, argument order, argument meaning and error convention:
 *   - functions returning int32_t return 0 on success, otherwise a nonzero
 *     hipError_t value (cu_allocate_io_buffers: -1), exactly like the
 *     cudaError_t values of genesis-compute/src/cuda/bindings.cu;
 *   - `void` launchers report nothing (the caller checks with a later sync).
 * Streams are opaque `void*` (ffi.rs:5 `CudaStream = *mut c_void`); on this
 * library they are hipStream_t handles created by gpu_stream_create.
 *
 * Symbols below the "engine extensions" line have no counterpart in the
 * reference; each says which caller it is for.
 *
 * Plain C, no HIP or torch types: a Rust, C, C++ or ctypes caller includes or
 * binds this file as is.
 */
#ifndef AXICOR_DAY_H
#define AXICOR_DAY_H

#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------- */
/* Binary contracts                                                          */
/* ------------------------------------------------------------------------- */

#define AXC_MAX_DENDRITES 128u          /* genesis-core/src/constants.rs:8 */
#define AXC_AXON_SENTINEL 0x80000000u   /* genesis-core/src/constants.rs:13 */
#define AXC_TARGET_AXON_MASK 0x00FFFFFFu /* constants.rs:21 */
#define AXC_TARGET_SEG_SHIFT 24u         /* constants.rs:22 */
#define AXC_NO_AXON 0xFFFFFFFFu          /* soma_to_axon "no axon" marker */
#define AXC_STATE_BYTES_PER_NEURON 910u  /* memory.rs:32-54: 4+1+4+1+4+128*(4+2+1) */

/* Eight burst heads of one axon, 32 B, align 32
 * (genesis-core/src/layout.rs:51-74; physics.cu:6-9). */
typedef struct AxcBurstHeads8 {
    uint32_t h[8];
} __attribute__((aligned(32))) AxcBurstHeads8;

/* Per-neuron-type parameter block, 64 B, align 64
 * (genesis-core/src/layout.rs:10-48 == physics.cu:46-79). Field offsets are
 * part of the ABI and are checked by static asserts in the library. */
typedef struct AxcVariantParameters {
    int32_t  threshold;                        /* @0  */
    int32_t  rest_potential;                   /* @4  */
    int32_t  leak_rate;                        /* @8  */
    int32_t  homeostasis_penalty;              /* @12 */
    uint32_t spontaneous_firing_period_ticks;  /* @16 */
    uint16_t initial_synapse_weight;           /* @20 */
    uint16_t gsop_potentiation;                /* @22 */
    uint16_t gsop_depression;                  /* @24 */
    uint16_t homeostasis_decay;                /* @26 */
    uint8_t  refractory_period;                /* @28 */
    uint8_t  synapse_refractory_period;        /* @29 */
    uint8_t  signal_propagation_length;        /* @30 */
    uint8_t  is_inhibitory;                    /* @31 */
    uint8_t  inertia_curve[16];                /* @32 */
    int32_t  adaptive_leak_max;                /* @48 */
    uint16_t adaptive_leak_gain;               /* @52 */
    uint8_t  adaptive_mode;                    /* @54 */
    uint8_t  _leak_pad[3];                     /* @55 */
    uint8_t  d1_affinity;                      /* @58 */
    uint8_t  d2_affinity;                      /* @59 */
    uint8_t  _pad[4];                          /* @60 */
} __attribute__((aligned(64))) AxcVariantParameters;

/* Device pointers of one shard, in `.state` blob order
 * (genesis-compute/src/ffi.rs:25-42). The first seven arrays live in one
 * allocation whose base is soma_voltage; the columnar dendrite arrays put slot
 * s of neuron i at index s*padded_n + i. */
typedef struct ShardVramPtrs {
    int32_t  *soma_voltage;      /* [N]      i32 */
    uint8_t  *soma_flags;        /* [N]      u8: [7:4] type, [3:1] burst, [0] spike */
    int32_t  *threshold_offset;  /* [N]      i32 */
    uint8_t  *timers;            /* [N]      u8 refractory counter */
    uint32_t *soma_to_axon;      /* [N]      u32, 0xFFFFFFFF = none */
    uint32_t *dendrite_targets;  /* [128*N]  u32 packed (seg<<24)|(axon+1), 0 = empty */
    int16_t  *dendrite_weights;  /* [128*N]  i16 */
    uint8_t  *dendrite_timers;   /* [128*N]  u8 */
    AxcBurstHeads8 *axon_heads;  /* [A]      32 B per axon */
} ShardVramPtrs;

/* ------------------------------------------------------------------------- */
/* Shard memory — ffi.rs:57-87, bindings.cu:571-691                          */
/* ------------------------------------------------------------------------- */

/* ffi.rs:57. One allocation for the 8 SoA arrays (zeroed) plus one for the
 * heads (all sentinel). Also sets up the engine's private per-shard workspace. */
int32_t cu_allocate_shard(uint32_t padded_n, uint32_t total_axons, ShardVramPtrs *out_vram);

/* ffi.rs:72. Synchronous copy of a `.state` blob to the base pointer. The
 * blob must not exceed 910*padded_n bytes (hipErrorInvalidValue otherwise). */
int32_t cu_upload_state_blob(const ShardVramPtrs *vram, const void *state_blob, size_t state_size);

/* ffi.rs:80. Synchronous copy of a `.axons` blob (raw AxcBurstHeads8[]) to
 * axon_heads; shorter blobs leave the remaining heads as they are. */
int32_t cu_upload_axons_blob(const ShardVramPtrs *vram, const void *axons_blob, size_t axons_size);

/* ffi.rs:87. Frees both allocations and the workspace, nulls the pointers. */
void cu_free_shard(ShardVramPtrs *vram);

/* ------------------------------------------------------------------------- */
/* Day phase — ffi.rs:63-109, physics.cu:441-506                             */
/* ------------------------------------------------------------------------- */

/* ffi.rs:63. flags &= 0xF1 on every neuron, asynchronously on `stream`. */
void cu_reset_burst_counters(const ShardVramPtrs *ptrs, uint32_t padded_n, void *stream);

/* ffi.rs:90-106. One tick, enqueued asynchronously on `stream`:
 * inject inputs, apply the spike schedule, propagate heads, integrate every
 * soma (GLIF), push the heads of spiking somas, GSOP on spiking somas, record
 * outputs. Dendrite reads see the heads after propagation; pushes of this
 * tick's spikes apply after every soma has integrated.
 * Returns 0, or hipErrorInvalidValue for inputs the reference would turn into
 * out-of-bounds writes (virtual_offset + num_virtual_axons > total_axons),
 * or the launch error. */
int32_t cu_step_day_phase(const ShardVramPtrs *vram, uint32_t padded_n, uint32_t total_axons,
                          uint32_t v_seg, uint32_t current_tick,
                          const uint32_t *input_bitmask, uint32_t virtual_offset,
                          uint32_t num_virtual_axons,
                          const uint32_t *incoming_spikes, uint32_t num_incoming_spikes,
                          const uint32_t *mapped_soma_ids, uint8_t *output_history,
                          uint32_t num_outputs, int16_t dopamine, void *stream);

/* ffi.rs:109. Process-wide (per device) 16-entry parameter table. */
int32_t cu_upload_constant_memory(const AxcVariantParameters *lut);

/* ------------------------------------------------------------------------- */
/* Runtime helpers — ffi.rs:114-158, bindings.cu:210-291                     */
/* ------------------------------------------------------------------------- */
void *gpu_malloc(size_t size);                                     /* ffi.rs:114 */
void  gpu_free(void *dev_ptr);                                     /* ffi.rs:115 */
void *gpu_host_alloc(size_t size);                                 /* ffi.rs:117 */
void  gpu_host_free(void *ptr);                                    /* ffi.rs:118 */
void  gpu_memcpy_host_to_device_async(void *dst, const void *src, size_t size, void *stream); /* :120 */
void  gpu_memcpy_device_to_host_async(void *dst, const void *src, size_t size, void *stream); /* :127 */
bool  gpu_memcpy_host_to_device(void *dst_dev, const void *src_host, size_t size);           /* :134 */
bool  gpu_memcpy_device_to_host(void *dst_host, const void *src_dev, size_t size);           /* :140 */
int32_t gpu_stream_create(void **out_stream);                      /* ffi.rs:146 */
int32_t gpu_stream_destroy(void *stream);                          /* ffi.rs:147 */
void  gpu_stream_synchronize(void *stream);                        /* ffi.rs:149 */
void  gpu_set_device(int32_t device_id);                           /* ffi.rs:150 */
void  gpu_device_synchronize(void);                                /* ffi.rs:151 */
void  gpu_synchronize(void);                                       /* ffi.rs:154 */
void  gpu_load_constants(const void *host_ptr);                    /* ffi.rs:157 */
void  update_constant_memory_hot_reload(const AxcVariantParameters *new_variants, void *stream); /* :158 */

/* ------------------------------------------------------------------------- */
/* Night phase — ffi.rs:160-164, bindings.cu:450-515,699-710                 */
/* ------------------------------------------------------------------------- */

/* ffi.rs:160. Per neuron: flags &= 0xF1; stable sort of the 128 slots by
 * descending |weight| with empty (target==0) slots after the filled ones;
 * then every filled slot with |weight| < prune_threshold becomes (0,0,0).
 * Runs on the null stream; the caller synchronises (shard_thread.rs:295-300). */
void launch_sort_and_prune(const ShardVramPtrs *ptrs, uint32_t padded_n, int16_t prune_threshold);

/* ------------------------------------------------------------------------- */
/* Multi-zone helpers — ffi.rs:166-185, bindings.cu:341-436                  */
/* ------------------------------------------------------------------------- */

/* ffi.rs:166. For each link i with src_indices[i] < 0x80000000 and each head
 * h < 0x70000000 of that axon with t = h / v_seg < sync_batch_ticks, appends
 * the 8 B event {dst_ghost_ids[i], sync_batch_ticks-1-t} to out_events and
 * bumps *out_count (zeroed first). Event order is unspecified. */
void launch_extract_outgoing_spikes(const AxcBurstHeads8 *axon_heads, const uint32_t *src_indices,
                                    const uint32_t *dst_ghost_ids, uint32_t count,
                                    uint32_t sync_batch_ticks, uint32_t v_seg,
                                    void *out_events, uint32_t *out_count, void *stream);

/* ffi.rs:178. dst_heads[dst_indices[i]] = src_heads[src_indices[i]] when
 * src_indices[i] < 0x80000000. */
void launch_ghost_sync(const AxcBurstHeads8 *src_heads, AxcBurstHeads8 *dst_heads,
                       const uint32_t *src_indices, const uint32_t *dst_indices,
                       uint32_t count, void *stream);

/* ------------------------------------------------------------------------- */
/* Telemetry — ffi.rs:235-246, physics.cu:400-434                            */
/* ------------------------------------------------------------------------- */

/* ffi.rs:235. No-op, as in bindings.cu:780-782. */
void gpu_reset_telemetry_count(const ShardVramPtrs *ptrs, void *stream);

/* ffi.rs:240. Appends the ids of somas with flags&1 to out_ids and adds their
 * number to *out_count_pinned (the caller zeroes it). Unlike the reference,
 * ids come out in ascending order. */
void launch_extract_telemetry(const ShardVramPtrs *ptrs, uint32_t padded_n, uint32_t *out_ids,
                              uint32_t *out_count_pinned, void *stream);

/* ------------------------------------------------------------------------- */
/* I/O buffers — ffi.rs:248-272, bindings.cu:717-778                         */
/* ------------------------------------------------------------------------- */
int32_t cu_allocate_io_buffers(uint32_t input_words, uint32_t schedule_capacity,
                               uint32_t output_capacity, uint32_t **d_input_bitmask,
                               uint32_t **d_incoming_spikes, uint8_t **d_output_history);
int32_t cu_dma_h2d_io(uint32_t *d_input_bitmask, const uint32_t *h_input_bitmask, uint32_t input_words,
                      uint32_t *d_incoming_spikes, const uint32_t *h_incoming_spikes,
                      uint32_t schedule_capacity, void *stream);
int32_t cu_dma_d2h_io(uint8_t *h_output_history, const uint8_t *d_output_history,
                      uint32_t output_capacity, void *stream);

/* ------------------------------------------------------------------------- */
/* Legacy launchers — ffi.rs:193-233. Declared by the reference's Rust crate */
/* and never called by genesis-node; two of them (update_neurons, apply_gsop)*/
/* have no definition in the reference at all. They take the old by-value    */
/* SoA_State struct (bindings.cu:16-36 == layout.rs VramState).              */
/* ------------------------------------------------------------------------- */
typedef struct AxcLegacyVramState {
    uint32_t padded_n;
    uint32_t total_axons;
    int32_t  *voltage;
    uint8_t  *flags;
    int32_t  *threshold_offset;
    uint8_t  *refractory_timer;
    uint32_t *soma_to_axon;
    uint32_t *dendrite_targets;
    int16_t  *dendrite_weights;
    uint8_t  *dendrite_timers;
    AxcBurstHeads8 *axon_heads;
    uint32_t *input_bitmask;
    uint8_t  *output_history;
    uint32_t *telemetry_count;
    uint32_t *telemetry_spikes;
} AxcLegacyVramState;

void launch_inject_inputs(AxcLegacyVramState vram, const uint32_t *bitmask, uint32_t current_tick,
                          uint32_t total_virtual_axons);                       /* ffi.rs:193 */
void launch_apply_spike_batch(AxcLegacyVramState vram, const void *tick_schedule,
                              uint32_t tick_spikes_count);                     /* ffi.rs:202 */
void launch_propagate_axons(AxcLegacyVramState vram, uint32_t v_seg);           /* ffi.rs:209 */
void launch_update_neurons(AxcLegacyVramState vram, const void *constants_ptr,
                           uint32_t current_tick);                             /* ffi.rs:215 */
void launch_apply_gsop(AxcLegacyVramState vram);                               /* ffi.rs:222 */
void launch_record_readout(AxcLegacyVramState vram, const uint32_t *mapped_soma_ids,
                           uint8_t *output_history, uint32_t current_tick,
                           uint32_t total_pixels);                             /* ffi.rs:226 */

/* ========================================================================= */
/* Engine extensions (no counterpart in the reference)                       */
/* ========================================================================= */

/* Per-tick work counters of one shard, accumulated on the device since the
 * last axc_shard_stats_reset (all exact integers):
 *   ticks            ticks stepped
 *   integrated       sum over ticks of |R| (somas not refractory at tick start)
 *   slots_scanned    sum over ticks of sum_{i in R} S_i (filled slots evaluated)
 *   targets_read     sum over ticks of sum_{i in R} min(S_i+1,128)
 *   hits             sum over ticks of H (slot hit tests that were true)
 *   spikes           sum over ticks of |F| (somas with the spike bit set)
 *   gsop_slots       sum over ticks of sum_{i in F} S_i
 *   gsop_targets     sum over ticks of sum_{i in F} min(S_i+1,128)
 * For benchmarks and metrics; read after the stream is synchronised. */
typedef struct AxcShardStats {
    uint64_t ticks;
    uint64_t integrated;
    uint64_t slots_scanned;
    uint64_t targets_read;
    uint64_t hits;
    uint64_t spikes;
    uint64_t gsop_slots;
    uint64_t gsop_targets;
    uint64_t gsop_slots_fused;    /* the part of gsop_slots applied inside neuron_tick */
    uint64_t gsop_targets_fused;  /* the part of gsop_targets applied inside neuron_tick */
} AxcShardStats;

int32_t axc_shard_stats_reset(const ShardVramPtrs *vram, void *stream);
int32_t axc_shard_stats_read(const ShardVramPtrs *vram, AxcShardStats *out);

/* A whole BSP batch (the loop of ShardEngine::step_day_phase_batch,
 * genesis-compute/src/compute/shard.rs:75-120) in one call: ticks
 * tick_base .. tick_base+sync_batch_ticks-1, with per-tick pointer strides
 * input_words_per_tick (bitmask), max_spikes_per_tick (schedule) and
 * num_outputs (output history), and h_spike_counts[t] spikes at tick t
 * (host array, may be NULL when the schedule is NULL).
 * Results are bit-identical to sync_batch_ticks calls of cu_step_day_phase.
 * For callers that own the tick loop (the Rust ShardEngine, the benchmarks):
 * the engine captures the batch in a HIP graph and replays it. The graph
 * depends on the batch's shape (T, slice sizes, pointers, v_seg), not on the
 * spike counts: those are copied to the device each call and read there, so
 * changing schedules replay the same graph. Returns hipErrorInvalidValue
 * when some h_spike_counts[t] exceeds max_spikes_per_tick, or is nonzero
 * without a schedule. */
int32_t axc_step_day_phase_batch(const ShardVramPtrs *vram, uint32_t padded_n, uint32_t total_axons,
                                 uint32_t v_seg, uint32_t tick_base, uint32_t sync_batch_ticks,
                                 const uint32_t *d_input_bitmask, uint32_t input_words_per_tick,
                                 uint32_t virtual_offset, uint32_t num_virtual_axons,
                                 const uint32_t *d_incoming_spikes, uint32_t max_spikes_per_tick,
                                 const uint32_t *h_spike_counts,
                                 const uint32_t *d_mapped_soma_ids, uint8_t *d_output_history,
                                 uint32_t num_outputs, int16_t dopamine, void *stream);

/* The tick of cu_step_day_phase without allocating: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and error codes as cu_step_day_phase, computed by the
 * direct dendrite scan, with its workspace in one static device buffer of
 * 16 MiB per GPU instead of per-shard allocations. It needs about
 * 16 (A + 1) + 8 A + 4.25 N + 128 KiB bytes of it (N = padded_n,
 * A = total_axons; e.g. N = A = 500,000) and returns hipErrorInvalidValue
 * beyond that, hipErrorNotInitialized before cu_upload_constant_memory.
 * Every call on one GPU shares that buffer: calls that overlap on one GPU
 * (other streams, other threads) race and must be serialised by the caller. */
int32_t axc_step_day_phase_noalloc(const ShardVramPtrs *vram, uint32_t padded_n, uint32_t total_axons,
                                   uint32_t v_seg, uint32_t current_tick, const uint32_t *input_bitmask,
                                   uint32_t virtual_offset, uint32_t num_virtual_axons,
                                   const uint32_t *incoming_spikes, uint32_t num_incoming_spikes,
                                   const uint32_t *mapped_soma_ids, uint8_t *output_history, uint32_t num_outputs,
                                   int16_t dopamine, void *stream);

/* Soma positions of the shard (PackedPosition u32 per soma, as the baker's
 * shard.pos holds them, genesis-core/src/types.rs:16-46; n = padded_n; NULL
 * returns to the caller's order). The staged steps then run in the engine's
 * own soma order: somas by 8x8x8-voxel bricks, axons owned by somas in the
 * same order, so that a block's presynaptic axons are few and close whatever
 * the caller's numbering (the baker numbers somas by z-plane only,
 * genesis-baker/src/bake/neuron_placement.rs:60-87). The caller's arrays keep
 * the caller's order between calls: each step copies the soma arrays and
 * heads into private ones and back. Results are identical in any order. A
 * step whose virtual input axons are owned by somas is refused
 * (hipErrorInvalidValue). Engine extension. */
int32_t axc_shard_set_soma_positions(const ShardVramPtrs *vram, const uint32_t *positions, uint32_t n);

/* Graphs recorded so far for the shard at vram (one per new batch shape). */
int32_t axc_shard_graph_captures(const ShardVramPtrs *vram, uint64_t *out);

/* The spike schedule's upload without the reference's full-capacity copy
 * (compute/shard.rs:55-67 hands cu_dma_h2d_io max_spikes_per_tick * T words
 * every batch, bindings.cu:760-764, even when the schedule is empty): one 2D
 * copy of the first max_t(h_spike_counts[t]) ids of each tick's slice of the
 * host schedule h_spikes ([T][max_spikes_per_tick]) into the same slices of
 * d_spikes, nothing when every count is 0. The kernels read only the first
 * count ids of a slice, so the result equals the full copy. Asynchronous on
 * `stream` (pinned h_spikes), hipErrorInvalidValue for a count beyond the
 * slice. For callers of axc_step_day_phase_batch and cu_step_day_phase. */
int32_t axc_dma_h2d_schedule(uint32_t *d_spikes, const uint32_t *h_spikes, uint32_t max_spikes_per_tick,
                             const uint32_t *h_spike_counts, uint32_t ticks, void *stream);

/* Host-side Night Phase sort & prune on columnar host arrays, the same rule as
 * launch_sort_and_prune; for hosts that run the Night Phase on CPU cores. */
void axc_host_sort_and_prune(uint8_t *flags, uint32_t *targets, int16_t *weights, uint8_t *timers,
                             uint32_t padded_n, int16_t prune_threshold);

/* Staged dendrite scan. The engine keeps a private plan derived from each
 * shard's dendrite_targets: per block of 256 somas, the distinct presynaptic
 * axons (whose head snapshots the soma pass stages in LDS) and each slot's
 * block-local index and segment. The plan is rebuilt at the next step after
 * the targets change through cu_upload_state_blob, launch_sort_and_prune or
 * gpu_memcpy_host_to_device(_async) into the targets array or the flags
 * array (the windows keep each soma's variant type, soma_flags >> 4). A
 * caller that writes dendrite_targets or the variant types any other way
 * (its own copies or kernels) calls axc_shard_targets_changed before the
 * next step. axc_staging_enable(0)
 * switches the whole process to the direct scan (same results). */
void axc_staging_enable(int enable);
void axc_shard_targets_changed(const ShardVramPtrs *vram);
/* Writes n slots of the shard's reference rows on `stream`:
 * dendrite_targets[cols[k]] = targets[k] and dendrite_weights[cols[k]] =
 * weights[k], cols[k] = slot * padded_n + soma (device arrays; columns past
 * 128 * padded_n are skipped), and rebuilds the staging plan at the next step
 * after that write. No counterpart in the reference, whose Night uploads the
 * whole targets and weights (shard_thread.rs:339-350): for a host Night that
 * changed a few slots (the new synapses of axc_sprout_pass). */
int32_t axc_shard_write_slots(const ShardVramPtrs *vram, uint32_t padded_n, const uint32_t *cols,
                              const uint32_t *targets, const int16_t *weights, uint32_t n, void *stream);

/* Windows of ticks (no counterpart in the reference; same results).
 * axc_step_day_phase_batch with the staged scan streams the dendrite slots
 * once per window of `ticks` ticks instead of once per tick: it predicts
 * every slot's hits over the window from the heads at the window's start and
 * corrects, tick by tick, the slots of the axons pushed since then (inputs,
 * schedule, own spikes), through a reverse index (axon -> slots) kept with
 * the plan. 1 (the default) is the per-tick pass; at most 4. A batch whose
 * ticks x v_seg reach 2^15 runs per tick. Process-wide, like
 * axc_staging_enable. Returns 0, or hipErrorInvalidValue for ticks > 4. */
int32_t axc_set_window_ticks(uint32_t ticks);
uint32_t axc_window_ticks(void);
/* The shard's window work counters since axc_shard_stats_reset, in out[5]:
 * windows, somas scanned by the window passes (those not refractory for the
 * whole window), their filled slots, correction pieces (up to 16 reverse-index
 * entries of one pushed axon each) and the entries they test, both summed
 * over ticks. */
int32_t axc_shard_window_stats(const ShardVramPtrs *vram, uint64_t *out);

/* Per-kernel timing, for benchmarks: when enabled, HIP events are recorded
 * on the caller's stream around each launch of cu_step_day_phase, and
 * axc_step_day_phase_batch launches its kernels one by one instead of
 * replaying its graph (same kernels, same results) so that they can be timed.
 * axc_kernel_timing_read waits for the events and returns the summed
 * milliseconds and launch counts of 12 kernels in this order:
 * axon_tick, neuron_tick, push_apply, gsop, readout, spike_count,
 * weights_copy (the batch's working copy of the weights, in and out),
 * window_start, window_pass, window_correct, window_soma, window_end (the
 * windowed batches' own kernels, see axc_set_window_ticks). */
void axc_kernel_timing_enable(int enable);
int32_t axc_kernel_timing_read(double *ms_out, uint64_t *launches_out, int reset);

/* ---- Batched brains -------------------------------------------------- */
/* Many independent brains, each in its own cu_allocate_shard arrays with its
 * own 16-entry parameter table, stepped together: per tick three work-list
 * kernels cover every brain, a batch of ticks is one captured HIP graph.
 * Every brain's result is bit-identical to stepping it alone with
 * axc_step_day_phase_batch. For callers that run many brains at once
 * (parameter searches, ensembles). */
typedef struct AxcBrainDesc {
    ShardVramPtrs vram;                 /* from cu_allocate_shard */
    uint32_t padded_n, total_axons, v_seg;
    uint32_t virtual_offset, num_virtual_axons;
    const AxcVariantParameters *lut;    /* host pointer, 16 entries, copied at create */
    const uint32_t *d_input_bitmask;    /* device [input_ticks][input_words], or NULL */
    uint32_t input_words;               /* u32 words per tick */
    uint32_t input_ticks;               /* tick t reads row t % input_ticks */
    int16_t dopamine;
} AxcBrainDesc;

/* Launch shape: workgroups of the axon, soma and list kernels (the soma count
 * must be a multiple of 8, one share per XCD), and the staged soma pass's
 * dispatch order: 0 deals the brains' 256-soma blocks to the 8 XCDs in
 * contiguous eighths of the brain order; k > 0 cuts every brain into runs
 * of k blocks and deals the runs, heaviest first, to the XCD with the least
 * work so far (work = the blocks' dendrite chunks and staged entries). */
typedef struct AxcBatchTuning {
    uint32_t axon_blocks, soma_blocks, list_blocks, stage_order;
} AxcBatchTuning;

int32_t axc_batch_create(const AxcBrainDesc *brains, uint32_t n_brains, void **out_batch);
/* Ticks tick_base .. tick_base+ticks-1 for every brain, asynchronously. */
int32_t axc_batch_step(void *batch, uint32_t tick_base, uint32_t ticks, void *stream);
/* Times launch shapes on `ticks` ticks from a snapshot of every brain (the
 * staged pass's dispatch order, the list and the axon grids), restores the
 * snapshot, keeps the fastest shape (also returned). */
int32_t axc_batch_autotune(void *batch, uint32_t ticks, void *stream, AxcBatchTuning *best);
int32_t axc_batch_set_tuning(void *batch, const AxcBatchTuning *tuning);
int32_t axc_batch_get_tuning(void *batch, AxcBatchTuning *tuning);
/* Counters of one brain since creation (or the last autotune) and its spikes
 * per variant type (16 entries, may be NULL). */
int32_t axc_batch_stats(void *batch, uint32_t brain, AxcShardStats *out, uint64_t *type_spikes16);

/* Synapse statistics of every brain in its current state: over the dendrite
 * slots with a target (target != 0), their number, sum |w|, sum w^2 and the
 * number with |w| below the brain's prune threshold (what the next Night's
 * sort&prune, bindings.cu:497-505, would remove), plus the sum of the soma
 * threshold offsets. prune_thresholds: host array of n_brains entries, or
 * NULL for 0. out: host array of n_brains entries. Synchronises `stream`. */
typedef struct AxcSynapseStats {
    uint64_t filled, sum_abs_weight, sum_sq_weight, below_prune;
    int64_t sum_threshold_offset;
} AxcSynapseStats;
int32_t axc_batch_synapse_stats(void *batch, const int16_t *prune_thresholds, AxcSynapseStats *out, void *stream);
/* New per-brain tables, dopamine and inputs for the same brains (each
 * brains[k].vram, padded_n and total_axons must be those given at create),
 * and counters reset to zero: a batch object reused across trials without
 * reallocating its workspaces. A brain whose targets changed (by the
 * library's uploads or axc_shard_targets_changed) gets its plan rebuilt at
 * the next step. */
int32_t axc_batch_update(void *batch, const AxcBrainDesc *brains, uint32_t n_brains);
/* axc_batch_stats of every brain in one call: out[n_brains],
 * type_spikes16 [n_brains][16] (may be NULL). */
int32_t axc_batch_stats_all(void *batch, AxcShardStats *out, uint64_t *type_spikes16);
void axc_batch_destroy(void *batch);

/* Library/version string. */
const char *axc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* AXICOR_DAY_H */
