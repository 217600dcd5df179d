/* C-ABI of the MI355X (gfx950) compute_dmsc implementation.
 *
 * Plain pointers and sizes only; no torch types. Device buffers are HIP device pointers on the
 * context's device; `stream` is the HIP stream the work is ordered on (NULL = default stream).
 *
 * Output slots follow the reference DMSComplex order (reference dmsc/csrc/dmsc_gpu.cpp:202-247):
 *   0 shape, 1 max_pts, 2 min_pts, 3 sad_pts, 4 e_max, 5 e_min, 6 p_max, 7 p_min, 8 ppairs_max,
 *   9 ppairs_min, 10 grad_indices, 11 peaks, 12 ridges, 13 ridge_offsets, 14 ridge_arc_offsets,
 *   15 basins, 16 valleys, 17 valley_offsets, 18 valley_arc_offsets.
 * All int32 except p_max/p_min (float32). Outputs are allocated through the caller's allocator
 * callback, so the caller owns them (the reference returns freshly allocated tensors,
 * dmsc_gpu.cpp:141-158); slot 0 (shape) is returned by value in `shape`.
 */
#ifndef DMSC_HIP_H
#define DMSC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* dmsc_stream_t; /* == hipStream_t */

typedef enum {
  DMSC_OK = 0,
  DMSC_ERR_INVALID_ARGUMENT = 1, /* reference: TORCH_CHECK in input_validation.hxx:9-26 */
  DMSC_ERR_TOO_LARGE = 2,        /* 32-bit cell ids, input_validation.hxx:24-26; or int32 arc offsets */
  DMSC_ERR_HIP = 3,              /* a HIP runtime call or kernel launch failed */
  DMSC_ERR_ALLOC = 4,            /* device scratch or the caller's allocator failed */
  DMSC_ERR_INTERNAL = 5          /* an internal consistency check failed */
} dmsc_status;

/* Flags, with the meaning of the reference extract_dmsc arguments (dmsc_gpu.cpp:128-130). The
 * trace flags are given in primal terms; the dual swap (dmsc_gpu.cpp:71-74) happens inside. */
enum {
  DMSC_RETURN_GRADIENT = 1u << 0,
  DMSC_DUAL = 1u << 1,
  DMSC_TRACE_MAX_ARCS = 1u << 2,   /* ridges   */
  DMSC_TRACE_MIN_ARCS = 1u << 3,   /* valleys  */
  DMSC_TRACE_MAX_GROUPS = 1u << 4, /* peaks    */
  DMSC_TRACE_MIN_GROUPS = 1u << 5  /* basins   */
};

enum { DMSC_NUM_OUTPUTS = 19 };
enum { DMSC_DTYPE_INT32 = 0, DMSC_DTYPE_FLOAT32 = 1 };

/* Allocator callback: return device memory for output `slot` with the given shape and dtype (ndim
 * 1 or 2, row-major), possibly with zero elements (the pointer is then never dereferenced). Return
 * NULL on failure. Never called concurrently. dmsc_extract calls it from the calling thread; for a
 * large batch, dmsc_extract_batch runs all parts but the first on worker threads of the library (see
 * dmsc_extract_batch) and the calls for those parts come from those threads while the calling thread
 * works too, one at a time with each other and with the calling thread's own calls.
 *   field >= 0: the outputs of one field (0 for dmsc_extract, the batch index when a batch group
 *     holds a single field); called once per slot 1..18 with that field's shape.
 *   field == -1 (dmsc_extract_batch): one buffer of dims[0] elements holding `slot` of several
 *     consecutive fields back to back in field order; out[b].data[slot] then points into it and
 *     out[b].dims / ndim give field b's own shape. */
typedef void* (*dmsc_alloc_fn)(void* user, int32_t field, int32_t slot, int32_t ndim, const int64_t* dims,
                               int32_t dtype);

typedef struct {
  void* data[DMSC_NUM_OUTPUTS]; /* device pointers from the allocator (data[0] unused) */
  int64_t dims[DMSC_NUM_OUTPUTS][2];
  int32_t ndim[DMSC_NUM_OUTPUTS];
  int32_t dtype[DMSC_NUM_OUTPUTS];
  int32_t shape[2]; /* slot 0: {H, W} */
} dmsc_outputs;

/* Threading: every entry point locks its context for its whole host-side run, so calls on one
 * context from several threads take turns and never share scratch. Error messages are kept per
 * calling thread (dmsc_last_error returns the calling thread's). */
typedef struct dmsc_context dmsc_context;

/* Creates a context bound to HIP device `device`; it owns reusable device scratch (grow-only).
 * Replaces the per-call gpu::Workspace of the reference (dmsc_gpu.cpp:148, gpu/workspace.hxx). */
dmsc_status dmsc_context_create(int device, dmsc_context** out);
void dmsc_context_destroy(dmsc_context* ctx);

/* Full pipeline on one contiguous H x W float32 field resident on the device.
 * Replaces extract_single_dmsc_gpu_t (reference dmsc_gpu.cpp:53-82) with the whole pipeline on
 * the GPU (gradient, critical cells, V-path forests, persistence, arc geometry, labels, assembly).
 * Synchronous: the outputs are complete when it returns (as in the reference). */
dmsc_status dmsc_extract(dmsc_context* ctx, const float* d_field, int32_t H, int32_t W, float persistence_threshold,
                         uint32_t flags, dmsc_stream_t stream, dmsc_alloc_fn alloc, void* alloc_user,
                         dmsc_outputs* out);

/* Batch of fields of independent sizes and thresholds (fields[b] is a contiguous H[b] x W[b]
 * device array). Generalises extract_batched_dmsc_gpu (reference dmsc_gpu.cpp:84-126), which
 * requires one shape and one threshold. out must hold B entries. Fields are processed in groups
 * (up to 2^28 cells per group): all fields of a group share every kernel launch and every
 * persistence / geometry round, so a batch of small fields costs about as many launches as one
 * field. A large batch runs as parts of about equal size side by side, as many as pipe_groups (4 by
 * default) with at least pipe_min_slots (H+1)(W+1) slots (2^23 by default) each: the first on the
 * calling thread and `stream`, each other one on a library worker thread with a context and stream
 * of its own, ordered after the work queued on `stream` before the call, with `stream` ordered after
 * them when the call returns (dmsc_set_param "pipe_groups" 1 keeps the whole batch on the calling
 * thread). Results equal dmsc_extract on each field alone. */
dmsc_status dmsc_extract_batch(dmsc_context* ctx, int32_t B, const float* const* d_fields, const int32_t* H,
                               const int32_t* W, const float* persistence_thresholds, uint32_t flags,
                               dmsc_stream_t stream, dmsc_alloc_fn alloc, void* alloc_user, dmsc_outputs* out);

/* The persistence pairs of one field's complex, as produced in the output slots (device pointers). */
typedef struct {
  int64_t n_sad;              /* number of output saddles (slot 3 length) */
  int64_t n_min, n_max;       /* slot 2 / slot 1 lengths */
  const int32_t* sad_pts;     /* slot 3 */
  const float* p_min;         /* slot 7 */
  const int32_t* ppairs_min;  /* slot 9 */
  const int32_t* min_pts;     /* slot 2 */
  const float* p_max;         /* slot 6 */
  const int32_t* ppairs_max;  /* slot 8 */
  const int32_t* max_pts;     /* slot 1 */
} dmsc_pairs;

/* Gradient of loss_b = sum over the pairs of field b with 0 <= persistence < max_persistence of
 * persistence^2, scaled by d_grad_scale[b] (device float), written to d_grad[b] (H[b] x W[b]
 * float32, fully overwritten). The pairs must come from an unsimplified complex (threshold < 0),
 * in which every extremum dies in at most one pair; flags: DMSC_DUAL as in the extraction.
 * Deterministic (no float atomics) and asynchronous on `stream`: two launches per 16 fields
 * (scatter, gather) over a context-owned scratch of 9 H W floats per field, which is all zeros
 * between calls; a call on another stream waits for the previous call's work. No reference
 * counterpart (the reference lists differentiation as future work, readme.md:237-242). */
dmsc_status dmsc_persistence_backward(dmsc_context* ctx, int32_t B, const float* const* d_fields, const int32_t* H,
                                      const int32_t* W, uint32_t flags, const dmsc_pairs* pairs,
                                      float max_persistence, const float* d_grad_scale, float* const* d_grad,
                                      dmsc_stream_t stream);

/* dmsc_persistence_backward without any device allocation: no counterpart in the reference. For
 * callers that cannot allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc. Same arguments, results and ordering as
 * dmsc_persistence_backward (the same kernels, bit for bit), but the scratch is the library's
 * fixed per-device array of DMSC_NOALLOC_SCRATCH_FLOATS floats (64 MiB, held by every GPU the
 * library runs on until the process ends), set up by dmsc_context_create. A field needs 9 H W of
 * them (at most 1864135 pixels); a larger field returns DMSC_ERR_TOO_LARGE before anything is
 * queued; fields are processed in as many chunks as the array needs. The array is all zeros again
 * after each call. Every call on a GPU shares that GPU's array: calls on one context are ordered,
 * but calls from different contexts that overlap on one GPU race. Devices 0-63 only. */
enum { DMSC_NOALLOC_SCRATCH_FLOATS = 1 << 24 };
dmsc_status dmsc_persistence_backward_noalloc(dmsc_context* ctx, int32_t B, const float* const* d_fields,
                                              const int32_t* H, const int32_t* W, uint32_t flags,
                                              const dmsc_pairs* pairs, float max_persistence,
                                              const float* d_grad_scale, float* const* d_grad, dmsc_stream_t stream);

/* Loss of each field: d_loss[b] = sum over the pairs of field b with 0 <= persistence <
 * max_persistence of persistence^2 (both sides; summed in double in a fixed order, so the result
 * is reproducible), as float32 on the device. One kernel launch per 16 fields, asynchronous on
 * `stream`. No reference counterpart (the reference lists differentiation as future work,
 * readme.md:237-242). */
dmsc_status dmsc_persistence_loss(dmsc_context* ctx, int32_t B, const dmsc_pairs* pairs, float max_persistence,
                                  float* d_loss, dmsc_stream_t stream);

/* Kernel timing with HIP events recorded on the work stream around launches (no reference
 * counterpart; the benchmark uses it to measure the dominant kernels' durations live).
 * dmsc_profile_enable(ctx, level) clears and starts accumulating: level 1 times only the key
 * stretches ("gradient", "persist_sweep" = each batch of persistence rounds launched back to back,
 * "persistence_backward"), level 2 times every stage (each timing event is a stream marker costing
 * a few microseconds of GPU time, so level 2 slows the round loops down), 0 stops. Entries are
 * (name, total ms, launches) over all calls since; a batch run as several parts adds all parts'
 * times (each part on its own stream, so they overlap in time). */
dmsc_status dmsc_profile_enable(dmsc_context* ctx, int32_t on);
int32_t dmsc_profile_count(dmsc_context* ctx);
dmsc_status dmsc_profile_entry(dmsc_context* ctx, int32_t index, const char** name, double* total_ms,
                               int64_t* launches);

/* Scheduling parameters of a context; outputs never depend on them (no reference counterpart: the
 * autotuner, dmsc_amd/autotune.py, picks them per GPU). Names and ranges:
 *   "round_grid"    64..65536  most blocks of a persistence / geometry round kernel (default 2048)
 *   "small_list"    0..2^30    worklists up to this size run one item per thread (default 65536)
 *   "rounds_first"  1..4096    rounds before the first host check of a round loop (default 4)
 *   "rounds_max"    1..4096    most rounds between host checks (default 16)
 *   "expand_levels" 1..4096    arc expansion levels between host checks (default 8)
 *   "graph_rounds"  0..1       1: each batch of rounds is captured once as a hipGraph and replayed (default 0)
 *   "tail_items"    0..65536   at most this many pending items: one workgroup runs every remaining round
 *                              of a round loop in one launch (default 2048; 0 = never)
 *   "grad_rows"     8, 16, 32  slot rows of a gradient tile (tiles are 64 slots wide; default 16)
 *   "side_rounds"   0..4096    first persistence rounds run side by side, vertex side first (default 1)
 *   "pipe_groups"   1..8       parts of a large batch run side by side, each on its own thread and
 *                              stream (default 4; 1 = one pipeline on the calling thread)
 *   "pipe_min_slots" 0..2^31-1 fewest slots per part of a split batch (default 2^23; 0 = no limit)
 * Unknown names or values out of range return DMSC_ERR_INVALID_ARGUMENT. */
dmsc_status dmsc_set_param(dmsc_context* ctx, const char* name, int64_t value);
dmsc_status dmsc_get_param(dmsc_context* ctx, const char* name, int64_t* value);

/* Message of the calling thread's last failing call on this context (valid until that thread's
 * next call on it). */
const char* dmsc_last_error(dmsc_context* ctx);

/* Copies an internal intermediate of the last dmsc_extract call into host memory, for stage-by-
 * stage parity tests. No reference counterpart. Returns the element count (int32 unless stated),
 * or -1 for an unknown name; copies min(count, capacity) elements, so host_dst must hold capacity
 * elements of the name's type. Names: "counts" (n_min, n_sad, n_max), "min_ids", "sad_ids",
 * "max_ids", "asc_order", "sad_cmax", "sad_cmin", "pers_min" (float), "pers_max" (float),
 * "pair_min", "pair_max", "face_pj", "vert_pj" (uint64 per face / vertex: V-path terminal in the
 * low 32 bits, depth in the high 32), "rounds" (persistence rounds, then the geometry rounds of
 * each side), "persist_pending" (per host check of the persistence rounds: round reached, pending
 * saddles of each side), "geo_pending" (the same for the geometry rounds, first entry round 0 with
 * the cancellations of each side), "arc_stats" (int64: raw pieces copied and output arc cells of the vertex
 * side, then of the face side). After a batch run as several parts, "counts" and "arc_stats" cover
 * the whole batch, "rounds" is the largest of the parts', and the arrays and logs are the first part's. */
int64_t dmsc_debug_fetch(dmsc_context* ctx, const char* name, void* host_dst, int64_t capacity);

#ifdef __cplusplus
}
#endif

#endif /* DMSC_HIP_H */
