/* gaffa_hip.h - C ABI of the MI355X (gfx950) search hot path.
 *
 * This is the drop-in boundary: plain C, plain pointers and sizes, no C++ or
 * torch types.  gaffa's own C++ host layer (plans, result types, programs)
 * calls these entry points where gaffa's CUDA translation units launch their
 * kernels; the functions each one replaces are cited beside it.  All device
 * pointers are HIP device allocations on the current device; tables (delays,
 * FFA task descriptors) are computed on the host by the caller and uploaded.
 *
 * Every entry point returns GAFFA_HIP_OK (0) or a status naming the C++
 * exception class gaffa throws for the same condition; gaffa_hip_last_error()
 * returns the message of the last failure on the calling thread.  Kernels are
 * enqueued on the given stream (NULL = the default stream) and never
 * synchronise unless the entry point says so.
 */
#ifndef GAFFA_HIP_H
#define GAFFA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GAFFA_HIP_ABI_VERSION 3

typedef enum gaffa_hip_status {
  GAFFA_HIP_OK = 0,
  GAFFA_HIP_INVALID_ARGUMENT = 1, /* std::invalid_argument */
  GAFFA_HIP_OVERFLOW = 2,         /* std::overflow_error */
  GAFFA_HIP_OUT_OF_RANGE = 3,     /* std::out_of_range */
  GAFFA_HIP_RUNTIME = 4,          /* std::runtime_error (HIP / device failure) */
  GAFFA_HIP_LOGIC = 5,            /* std::logic_error */
  GAFFA_HIP_NOT_SUPPORTED = 6,
} gaffa_hip_status;

typedef struct gaffa_hip_stream_opaque* gaffa_hip_stream; /* == hipStream_t */
typedef struct gaffa_hip_event_opaque* gaffa_hip_event;   /* == hipEvent_t */
typedef struct gaffa_hip_graph_opaque* gaffa_hip_graph;   /* == hipGraphExec_t */

const char* gaffa_hip_last_error(void);
int gaffa_hip_abi_version(void);

/* ------------------------------------------------------------- runtime ----
 * Replace src/gaffa/cuda_memory.cu:21-116 (CudaDeviceMemory allocation and
 * release), the cudaMemcpy calls of dedispersion_cuda.cu:619-626/718-728 and
 * ffa_bindings.cpp:146-152, and vector_add.cu:64-89 (device queries). */
typedef struct gaffa_hip_device_info {
  char name[256];
  char arch[64];           /* gcnArchName, e.g. "gfx950:sramecc+:xnack-" */
  int32_t compute_units;
  int32_t clock_khz;
  int32_t memory_clock_khz;
  int32_t memory_bus_width;
  int64_t total_memory;
  int32_t l2_bytes;
  int32_t lds_bytes_per_block;
  int32_t runtime_version;
} gaffa_hip_device_info;

int gaffa_hip_device_count(int32_t* count);
int gaffa_hip_device_info_get(int32_t device, gaffa_hip_device_info* out);
int gaffa_hip_set_device(int32_t device);
int gaffa_hip_get_device(int32_t* device);
int gaffa_hip_malloc(void** ptr, uint64_t bytes);
int gaffa_hip_free(void* ptr);
/* Copies in any direction (hipMemcpyDefault).  With a NULL stream the copy is
 * synchronous; otherwise it is enqueued on the stream. */
int gaffa_hip_memcpy(void* dst, const void* src, uint64_t bytes,
                     gaffa_hip_stream stream);
int gaffa_hip_memset(void* dst, int32_t value, uint64_t bytes,
                     gaffa_hip_stream stream);
int gaffa_hip_stream_create(gaffa_hip_stream* stream);
int gaffa_hip_stream_destroy(gaffa_hip_stream stream);
int gaffa_hip_stream_synchronize(gaffa_hip_stream stream);
int gaffa_hip_device_synchronize(void);
int gaffa_hip_event_create(gaffa_hip_event* event);
int gaffa_hip_event_destroy(gaffa_hip_event event);
int gaffa_hip_event_record(gaffa_hip_event event, gaffa_hip_stream stream);
int gaffa_hip_event_elapsed_ms(gaffa_hip_event start, gaffa_hip_event stop,
                               float* ms);
int gaffa_hip_event_synchronize(gaffa_hip_event event);
/* Work enqueued on `stream` after this call waits for `event`. */
int gaffa_hip_stream_wait_event(gaffa_hip_stream stream, gaffa_hip_event event);
/* Page-locked host memory (hipHostMalloc) for asynchronous copies. */
int gaffa_hip_host_malloc(void** ptr, uint64_t bytes);
int gaffa_hip_host_free(void* ptr);
/* dst[0 .. count) = value with a kernel (graph-safe reset of counters; no
 * gaffa counterpart). */
int gaffa_hip_fill_u32(uint32_t* dst, uint32_t value, uint64_t count,
                       gaffa_hip_stream stream);
/* Stream capture into a replayable graph (no counterpart in gaffa, whose
 * CUDA path launches kernels one by one). */
int gaffa_hip_graph_begin(gaffa_hip_stream stream);
int gaffa_hip_graph_end(gaffa_hip_stream stream, gaffa_hip_graph* graph);
int gaffa_hip_graph_launch(gaffa_hip_graph graph, gaffa_hip_stream stream);
int gaffa_hip_graph_destroy(gaffa_hip_graph graph);

/* -------------------------------------------------------- dedispersion ----
 * Input: uint8 filterbank samples, [nsamples][nchans] (nifs == 1, channels
 * ascending as gaffa's reader stores them).  Output: uint32 sums,
 * out[dm * out_stride + t] for t < out_nsamples.  Samples at t >= nsamples
 * read as zero (dedispersion.cpp:206-212, 367-372, 431-437). */

/* Brute force, out[dm][t] = sum_c x[t + delays[dm][c]][chan_begin + c].
 * Replaces multi_dm_kernel (dedispersion_cuda.cu:320-348) as launched by
 * dedisperse_multi_dm_cuda_device_impl; also serves the single-DM case
 * (single_dm_kernel, :269-293) with ndm == 1.  max_block_spread bounds
 * max - min of the delays inside any 16-DM x 32-channel block (the host
 * computes it from its table); it sizes the LDS slab. */
typedef struct gaffa_hip_dedisperse_multi_args {
  const uint8_t* samples;
  uint64_t nsamples;
  uint64_t nchans;
  uint64_t chan_begin;
  uint64_t chan_end;
  const int32_t* delays; /* [ndm][chan_end - chan_begin], all >= 0 */
  uint64_t ndm;
  uint32_t* out;
  uint64_t out_nsamples;
  uint64_t out_stride;
  int32_t max_block_spread;
} gaffa_hip_dedisperse_multi_args;

int gaffa_hip_dedisperse_multi_u8(const gaffa_hip_dedisperse_multi_args* args,
                                  gaffa_hip_stream stream);

/* Channel-major copy of a filterbank for subband dedispersion:
 * out[c * out_stride + t] = samples[t * nchans + chan_begin + c] for
 * t < nsamples and 0 for nsamples <= t < out_stride (the zero padding stands
 * for gaffa's out-of-range reads).  out_stride: multiple of 16, at least
 * gaffa_hip_subband_channel_stride(); out 16-byte aligned.  No gaffa
 * counterpart (its kernels read the sample-major filterbank). */
typedef struct gaffa_hip_transpose_args {
  const uint8_t* samples; /* [nsamples][nchans] */
  uint64_t nsamples;
  uint64_t nchans;
  uint64_t chan_begin;
  uint64_t chan_end;
  uint8_t* out;           /* [chan_end - chan_begin][out_stride] */
  uint64_t out_stride;
  uint64_t t_begin;       /* time range written: [t_begin, t_begin + t_count), */
  uint64_t t_count;       /* multiples of 256 except at out_stride; 0 = all */
} gaffa_hip_transpose_args;

/* dst[r][c] = src[r][reverse ? nchans - 1 - c : c] for rows r < rows (uint8
 * filterbank rows; reverse turns a descending-frequency file into gaffa's
 * ascending channel order).  No gaffa counterpart (its reader reverses on
 * the host, io/filterbank.cpp:79-118). */
int gaffa_hip_copy_rows_u8(const uint8_t* src, uint8_t* dst, uint64_t rows,
                           uint64_t nchans, int32_t reverse, gaffa_hip_stream stream);

int gaffa_hip_filterbank_transpose_u8(const gaffa_hip_transpose_args* args,
                                      gaffa_hip_stream stream);

/* Time chunk gaffa_hip_dedisperse_subband_u8 uses for this geometry: the
 * u16 subband windows (chunk + the group's residual halo per subband) must
 * fit LDS.  residual_window as in gaffa_hip_dedisperse_subband_args (0 =
 * nsub * max_residual).  0 when nothing fits: the call would fail with
 * GAFFA_HIP_OVERFLOW. */
uint32_t gaffa_hip_subband_time_tile(uint64_t nsel, uint64_t subband_channels,
                                     uint64_t max_group_ndm, int32_t max_residual,
                                     int64_t residual_window);

/* Row length (bytes) a channel-major copy needs so that every stage-1 read
 * of gaffa_hip_dedisperse_subband_u8 stays in the row. */
uint64_t gaffa_hip_subband_channel_stride(uint64_t nsamples, int32_t max_coarse,
                                          int32_t max_residual);

/* One nominal DM group of a multi-segment subband launch: coarse delays row
 * `coarse_row` of coarse_delays, residual delays and output rows
 * [first_dm, first_dm + ndm). */
typedef struct gaffa_hip_subband_group {
  uint32_t coarse_row;
  uint32_t first_dm;
  uint32_t ndm;
  uint32_t reserved; /* 0 */
} gaffa_hip_subband_group;

/* Two-stage subband dedispersion fused in LDS: stage 1 sums each subband at
 * the coarse delays of its nominal DM, reading the channel-major rows made by
 * gaffa_hip_filterbank_transpose_u8; stage 2 adds the subbands at the residual
 * delays; the intermediate never reaches HBM.  Replaces subband_stage1_kernel
 * + subband_stage2_kernel and their tile loop (dedispersion_cuda.cu:350-417,
 * 659-707).  nsub = ceil(nsel / subband_channels).  Groups: either `groups`
 * (device, ngroups entries; ndm = output rows, n_coarse_rows = rows of
 * coarse_delays, several DM segments in one launch) or, with groups NULL,
 * gaffa's nominal groups of one plan: group k = DMs [k * ndm_per_nominal,
 * ...) with coarse row k.  max_residual bounds every residual delay,
 * max_coarse every coarse delay; residual_window bounds, over the groups,
 * the sum over subbands of the group's largest residual in that subband,
 * each rounded up to a multiple of 8 (0 = nsub * max_residual rounded up:
 * always safe, may need a smaller chunk).  Any group size whose windows fit
 * LDS works (no ndm_per_nominal limit besides max_group_ndm <= 1024). */
typedef struct gaffa_hip_dedisperse_subband_args {
  const uint8_t* channels;  /* channel-major [nsel][channel_stride] */
  uint64_t channel_stride;  /* bytes between channel rows, multiple of 4 */
  uint64_t channel_length;  /* readable bytes from `channels` along a row
                               (0 = channel_stride; less when `channels`
                               starts inside the rows) */
  uint64_t nsamples;        /* valid samples per row */
  uint64_t nsel;            /* channels (the selected range) */
  uint64_t subband_channels;
  uint64_t ndm_per_nominal;
  uint64_t ndm;
  const int32_t* coarse_delays;   /* [n_nominal][nsel] */
  const int32_t* residual_delays; /* [ndm][nsub] */
  uint32_t* out;
  uint64_t out_nsamples;
  uint64_t out_stride;
  int32_t max_residual;
  int32_t max_coarse;
  const gaffa_hip_subband_group* groups; /* NULL = gaffa's groups of one plan */
  uint64_t ngroups;
  uint64_t n_coarse_rows;
  uint64_t max_group_ndm; /* largest group (groups mode; else ndm_per_nominal) */
  int64_t residual_window;
} gaffa_hip_dedisperse_subband_args;

int gaffa_hip_dedisperse_subband_u8(
    const gaffa_hip_dedisperse_subband_args* args, gaffa_hip_stream stream);

/* Aligned dynamic spectrum for one DM: out[t][c] = samples[t + delays[c]]
 * [chan_begin + c] for t < out_nsamples (out [out_nsamples][chan_end -
 * chan_begin], element_size 1, 2 or 4 bytes, copied bit for bit).
 * max_tile_spread bounds max - min of the delays inside any channel_tile
 * consecutive channels (sizes the LDS slab).  Replaces
 * dedisperse_spectrum_kernel (dedispersion_cuda.cu:295-318). */
typedef struct gaffa_hip_spectrum_args {
  const void* samples;   /* [nsamples][nchans] */
  uint64_t nsamples;
  uint64_t nchans;
  uint64_t chan_begin;
  uint64_t chan_end;
  uint32_t element_size;
  const int32_t* delays; /* [chan_end - chan_begin], >= 0 */
  void* out;
  uint64_t out_nsamples;
  uint32_t channel_tile; /* 8..256, power of two */
  int32_t max_tile_spread;
} gaffa_hip_spectrum_args;

int gaffa_hip_dedisperse_spectrum(const gaffa_hip_spectrum_args* args,
                                  gaffa_hip_stream stream);

/* Synchronises `stream` and reports (then clears) a refused dedispersion
 * block: a delay table whose spread exceeded the bound the caller passed.
 * No counterpart in gaffa (its kernels read global memory directly). */
int gaffa_hip_dedisperse_check(gaffa_hip_stream stream);

/* ------------------------------------------------- detrend + normalise ----
 * In-place batch preprocessing of a [nseries][stride] plane: optional
 * uint32 -> float conversion (input_is_u32), riptide fast running-median
 * detrend (mean scrunch by `scrunch_factor`, exact edge-padded median of
 * `median_window` low-resolution points, linear interpolation back,
 * subtract), and optional normalise to zero mean / unit population sigma.
 * Replaces convert_uint32_to_float_kernel (time_series_cuda.cu:50-58) and the
 * seven kernels of CudaPreprocessProgram (preprocessing_cuda.cu:274-516).
 * scrunch_factor == 1 selects the exact running median at full resolution.
 * `workspace` must hold gaffa_hip_preprocess_workspace_bytes(...) bytes.
 * `status` (device, int32[nseries]) receives per-series deferred errors:
 * 0 ok, 1 non-finite input, 2 constant input rejected, 3 non-finite
 * variance.  The caller reads it after synchronising (preprocessing_cuda.cu:
 * 781-817 reports the same conditions at synchronize()). */
typedef struct gaffa_hip_preprocess_args {
  void* data; /* uint32 (input_is_u32) or float, becomes float */
  uint64_t nseries;
  uint64_t nsamples;
  uint64_t stride; /* elements between series starts */
  int32_t input_is_u32;
  int32_t detrend;
  uint64_t scrunch_factor;
  uint64_t median_window; /* odd */
  int32_t normalise;
  int32_t reject_constant;
  void* workspace;
  uint64_t workspace_bytes;
  int32_t* status;
} gaffa_hip_preprocess_args;

uint64_t gaffa_hip_preprocess_workspace_bytes(uint64_t nseries,
                                              uint64_t nsamples,
                                              uint64_t scrunch_factor);
int gaffa_hip_detrend_normalise_batch(const gaffa_hip_preprocess_args* args,
                                      gaffa_hip_stream stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same conversion, detrend and normalise as
 * gaffa_hip_detrend_normalise_batch, with its scratch in a library-owned
 * device array of GAFFA_HIP_NOALLOC_SCRATCH_BYTES (held on every device the
 * library runs on until the process ends) instead of a workspace argument.
 * `workspace` must be NULL and `workspace_bytes` 0; batches whose scratch
 * (gaffa_hip_preprocess_workspace_bytes) exceeds the maximum are rejected
 * with GAFFA_HIP_OVERFLOW.  The call allocates nothing.  All calls on one
 * device share the array: calls that overlap on one device (other streams,
 * other threads) race and must be serialised by the caller. */
#define GAFFA_HIP_NOALLOC_SCRATCH_BYTES (64ull << 20)
int gaffa_hip_detrend_normalise_batch_noalloc(const gaffa_hip_preprocess_args* args,
                                              gaffa_hip_stream stream);
uint64_t gaffa_hip_noalloc_scratch_bytes(void);

/* ------------------------------------------------------------------ FFA ----
 * Device-side layout of one FFA execution plan (host-built, see
 * gaffa_amd/csrc/host/ffa_gpu.cpp).  A plan is shared by every series of a
 * batch; a batch of series with different lengths and plans is run as one
 * "series" row holding all of them, each group naming its input's offset and
 * length and each task the series its peaks belong to.  Per series, `prepared` holds each prepare group's downsampled copy
 * at prep_offset[group] (factor-1 groups read the input series directly) and
 * `work` holds two ping-pong regions of rows*bins floats per task. */
typedef struct gaffa_hip_ffa_group {
  double factor;            /* downsample factor, 1.0 = no copy */
  uint64_t prepared_nsamples;
  uint64_t prep_offset;     /* float offset inside a series' prepared region */
  uint64_t input_offset;    /* float offset of this group's input series in a
                               series row (0 when every series has one plan) */
  uint64_t input_nsamples;  /* length of that input series */
} gaffa_hip_ffa_group;

#define GAFFA_HIP_FFA_MAX_WIDTHS 16

typedef struct gaffa_hip_ffa_task {
  uint32_t group;
  uint32_t bins;
  uint32_t rows;
  uint32_t rows_eval;
  uint32_t depth;           /* depth of the LDS subtrees, >= 1 */
  uint32_t nwidths;
  uint64_t work_offset;     /* float offset of this task's rows*bins region */
  float inv_stdnoise;       /* 1.0f / ffa_task_stdnoise(task) */
  uint32_t series;          /* added to the peak's series index (batches of
                               series with their own plans in one row) */
  uint32_t periodogram_row; /* first row of this task in the series'
                               periodogram (plan order, rows_eval each) */
  uint32_t regions;         /* bit k: work region (0 or 1) holding the
                               depth-k rows of this task's tree */
  uint32_t terminal_levels; /* tree levels the detection kernel merges from
                               global memory: 1 (it reads the depth-1 rows)
                               or 2 (depth-2 rows, where that saves a global
                               pass); 2 needs depth >= 2 */
  uint32_t width[GAFFA_HIP_FFA_MAX_WIDTHS];
  float height[GAFFA_HIP_FFA_MAX_WIDTHS];
  float baseline[GAFFA_HIP_FFA_MAX_WIDTHS];
} gaffa_hip_ffa_task;

typedef struct gaffa_hip_ffa_item {
  uint32_t task;
  uint32_t a; /* subtree: node start row; level: output depth | levels << 16
                 (1, 2 or 3 tree levels merged in one pass); terminal: first row */
  uint32_t b; /* subtree: node rows;      level: first row; terminal: rows */
  uint32_t c; /* level: rows in this item */
} gaffa_hip_ffa_item;

/* One significant (row, width) detection, 20 bytes (cf. FfaCudaPeak,
 * ffa_cuda_kernels.cuh:70-79).  The host expands it with make_ffa_peak. */
typedef struct gaffa_hip_ffa_peak {
  uint32_t series;
  uint32_t task;
  uint32_t shift;
  uint32_t phase_width; /* phase | (width_index << 16) */
  float snr;
} gaffa_hip_ffa_peak;

/* One row of one subtree pass (32 bytes), in float offsets from the pass's
 * source row buffer.  A pass runs one or two tree levels:
 *   one level   out[i] = A[i] + B[(i + rb) mod bins]
 *   two levels  out[i] = (A[i] + B[(i + rb) mod bins])
 *                        + (C[(i + rc) mod bins] + D[(i + rd) mod bins]),
 * the two one-level merges' FP32 sums in their order (the depth-(d + 1) row
 * in between is never stored): A + B is the head child row and C + D the
 * tail child row read at the output merge's roll rc.  A child that is a leaf
 * (or a row that is a leaf above the pass) adds the -0.0 row
 * (GAFFA_HIP_FFA_COPY_ROW), which leaves every value as it is.  A subtree
 * item of L levels has ceil(L / 2) passes, the node's own (top) pass first;
 * every pass has two levels except the bottom one when L is odd.  Each pass
 * holds 16 wave lists of 2 * ceil(ceil(rows / 16) / 2) entries: wave w's
 * rows w, w + 16, ..., then padding entries (out = GAFFA_HIP_FFA_SKIP_ROW) up
 * to the list length.  The item's `c` field is the index of its first
 * entry; items with the same rows and bins share a table.  No gaffa
 * counterpart: gaffa recurses per row (ffa.cpp:110-251). */
#define GAFFA_HIP_FFA_COPY_ROW 0xFFFFFFFFu
#define GAFFA_HIP_FFA_SKIP_ROW 0xFFFFFFFFu
typedef struct gaffa_hip_ffa_merge {
  uint32_t head;  /* A: row * bins */
  uint32_t tail;  /* B: row * bins + rb, or GAFFA_HIP_FFA_COPY_ROW */
  uint32_t wrap;  /* bins - rb: phases whose B index does not wrap */
  uint32_t out;   /* output row * bins, or GAFFA_HIP_FFA_SKIP_ROW: padding */
  uint32_t head2; /* C: row * bins + rc (two-level passes; else unused) */
  uint32_t wrap2; /* bins - rc */
  uint32_t tail2; /* D: row * bins + rd, or GAFFA_HIP_FFA_COPY_ROW */
  uint32_t wrap3; /* bins - rd */
} gaffa_hip_ffa_merge;

/* Leaves of one output row of a level pass of L = 1..3 tree levels: the
 * 2^L depth-(d + L) rows it sums and the combined roll of each (mod bins,
 * roll[0] = 0), in the order of its one-level merges; entries past 2^L are
 * 0.  No gaffa counterpart (gaffa merges one level per pass). */
typedef struct gaffa_hip_ffa_level_row {
  uint32_t row[8];
  uint32_t roll[8];
} gaffa_hip_ffa_level_row;

/* A level item's leaf runs: leaf j of its rows spans rows first[j] ..
 * first[j] + count[j] - 1 (smallest to largest over the item's rows); the
 * runs are staged back to back, `staged` rows in all; the item's rows'
 * leaves are level_leaf_rows[row_offset .. row_offset + item rows). */
typedef struct gaffa_hip_ffa_level_span {
  uint32_t row_offset;
  uint32_t staged;
  uint32_t first[8];
  uint32_t count[8];
  uint32_t reserved[2];
} gaffa_hip_ffa_level_span;

typedef struct gaffa_hip_ffa_plan_desc {
  const gaffa_hip_ffa_group* groups; /* device */
  uint32_t ngroups;
  const gaffa_hip_ffa_task* tasks;   /* device */
  uint32_t ntasks;
  const gaffa_hip_ffa_item* subtree_items; /* device; c = first merge */
  uint32_t nsubtree_items;
  const gaffa_hip_ffa_merge* subtree_merges; /* device */
  const gaffa_hip_ffa_item* level_items;   /* device, sorted by depth desc */
  const gaffa_hip_ffa_level_span* level_spans;    /* device, one per level item */
  const gaffa_hip_ffa_level_row* level_leaf_rows; /* device, spans' row tables */
  const uint32_t* level_item_offsets;      /* host, [max_depth + 1] */
  uint32_t max_depth;
  const gaffa_hip_ffa_item* terminal_items; /* device */
  uint32_t nterminal_items;
  uint32_t max_bins;
  uint32_t subtree_lds_bytes;  /* 8 * subtree_half_floats + 1024 + 4 * Z, Z =
                                  max(256, max_bins + 64) rounded up to 64 */
  uint32_t subtree_half_floats; /* max over subtree items of rows * bins */
  uint32_t max_subtree_desc;    /* max over subtree items of merges per item */
  uint32_t terminal_lds_bytes; /* max over tasks of ..._terminal_lds_bytes() */
  uint32_t level_lds_bytes;    /* dynamic LDS of a level launch: items that need
                                  at most this much (..._level_lds_bytes) merge
                                  from LDS, the others from global memory (0:
                                  every item from global memory); at most
                                  155648 */
  uint32_t prep_blocks;        /* max over downsampled groups of
                                  ceil(prepared_nsamples / outputs per block) */
  uint32_t prep_fused_halo;    /* > 0: every downsampled group reads the one
                                  series (input_offset 0, prep_input_nsamples
                                  samples) and a window reaches at most this
                                  many samples (ceil(max factor) + 2, <= 4096)
                                  past its start: one pass per series */
  uint64_t prep_input_nsamples;
  uint64_t max_prepared_nsamples;
  uint64_t prepared_stride; /* floats per series in `prepared` */
  uint64_t work_stride;     /* floats per series per work region */
} gaffa_hip_ffa_plan_desc;

/* Raises the launch limits (dynamic LDS) the plan's kernels need, so that a
 * later stream capture of its launches records kernels only.  No gaffa
 * counterpart. */
int gaffa_hip_ffa_reserve_launch(const gaffa_hip_ffa_plan_desc* plan);

/* Geometry the host needs to build a plan descriptor: rows per terminal and
 * level work item, prepared outputs per prepare workgroup for a factor, and
 * the LDS bytes of a terminal tile / subtree node (0xFFFFFFFF = too many
 * rows for a subtree).  No gaffa counterpart (its launch geometry is
 * compiled into ffa_cuda_executor.cuh). */
uint32_t gaffa_hip_ffa_terminal_rows(void);
uint32_t gaffa_hip_ffa_level_rows(void);
uint32_t gaffa_hip_ffa_prepare_outputs_per_block(double factor);
uint32_t gaffa_hip_ffa_terminal_lds_bytes(uint32_t bins, uint32_t max_width);
/* The same for terminal items of `rows` rows (1 .. gaffa_hip_ffa_terminal_rows_max();
 * 0xFFFFFFFF otherwise), and the largest row counts level and terminal items
 * may hold (the host picks rows per item within them; no gaffa counterpart). */
uint32_t gaffa_hip_ffa_terminal_lds_bytes_rows(uint32_t bins, uint32_t max_width, uint32_t rows);
uint32_t gaffa_hip_ffa_terminal_rows_max(void);
uint32_t gaffa_hip_ffa_level_rows_max(void);
/* LDS bytes a level work item of `count` rows from `first` (depth `depth` of
 * a task of `task_rows` rows and `bins` bins, merged over `levels` = 1..3
 * levels) needs to merge from LDS: its leaf rows' runs plus a guard
 * (UINT64_MAX for invalid arguments).  No gaffa counterpart (gaffa merges
 * every level from global memory, ffa_cuda_program.cuh:230-336). */
uint64_t gaffa_hip_ffa_level_lds_bytes(uint32_t task_rows, uint32_t bins, uint32_t depth,
                                       uint32_t levels, uint32_t first, uint32_t count);
/* The same item's span (all but row_offset, which the caller sets) and its
 * rows' leaves (`rows`: count entries), with the kernels' own row-selection
 * arithmetic (gaffa's ffa.cpp:133-143 in FP32); returns the LDS need as
 * gaffa_hip_ffa_level_lds_bytes.  Null outputs are not written. */
uint64_t gaffa_hip_ffa_level_item(uint32_t task_rows, uint32_t bins, uint32_t depth,
                                  uint32_t levels, uint32_t first, uint32_t count,
                                  gaffa_hip_ffa_level_span* span, gaffa_hip_ffa_level_row* rows);
uint32_t gaffa_hip_ffa_subtree_lds_bytes(uint32_t rows, uint32_t bins);
/* Fills the merge table of a subtree node of `rows` rows (host memory; a
 * null `out` only counts) with the kernels' own row-selection arithmetic
 * (gaffa's ffa.cpp:133-143 in FP32).  Returns the entry count, 0 for rows
 * outside 1..512 or bins outside 2..65535. */
uint32_t gaffa_hip_ffa_subtree_merges(uint32_t rows, uint32_t bins, gaffa_hip_ffa_merge* out);

typedef struct gaffa_hip_ffa_search_args {
  const gaffa_hip_ffa_plan_desc* plan;
  const float* series;  /* [nseries][series_stride], preprocessed */
  uint64_t nseries;
  uint64_t nsamples;
  uint64_t series_stride;
  uint32_t series_index_base; /* added to gaffa_hip_ffa_peak::series */
  float snr_threshold;
  float* prepared;      /* [nseries][plan->prepared_stride] */
  float* work;          /* [2][nseries][plan->work_stride] */
  gaffa_hip_ffa_peak* peaks;
  uint64_t peak_capacity;
  uint32_t* peak_count; /* device; incremented, never reset here */
  /* Optional dense periodogram (null = none): the S/N of every evaluated
   * (row, width) of every task, periodogram[s * periodogram_stride +
   * (task.periodogram_row + row) * periodogram_widths + width_index] for
   * series s of this call.  No gaffa counterpart (its search keeps only the
   * peaks above threshold); read by the candidate scorer. */
  float* periodogram;
  uint64_t periodogram_stride;
  uint32_t periodogram_widths;
} gaffa_hip_ffa_search_args;

/* Downsample every prepare group of every series (weighted-sum, serial FP32
 * per output exactly as downsample_weighted_sum_cpu).  Replaces
 * downsample_weighted_sum_float_kernel (time_series_cuda.cu:15-48) and
 * prepare_copy_float_kernel (ffa_cuda_kernels.cuh:130-143). */
int gaffa_hip_ffa_prepare_batch(const gaffa_hip_ffa_search_args* args,
                                gaffa_hip_stream stream);

/* Transform + boxcar detection of every task of every series; appends
 * significant peaks to args->peaks (peak_count may exceed peak_capacity:
 * the caller grows the buffer and repeats, as ffa_cuda_executor.cuh:491-531).
 * Replaces the per-level copy/merge/subtree kernels and the detection
 * kernels (ffa_cuda_kernels.cuh:145-737) and the tile runner that launches
 * them (ffa_cuda_executor.cuh:385-531). */
int gaffa_hip_ffa_search_batch(const gaffa_hip_ffa_search_args* args,
                               gaffa_hip_stream stream);

/* The four launch stages of gaffa_hip_ffa_search_batch, enqueued in this
 * order on one stream: prepare_batch, subtree_batch (every level of the
 * task trees that fits in LDS), levels_batch (the remaining levels below the
 * root, through global memory), detect_batch (root merge + boxcar S/N +
 * peak append).  Separate entry points let a caller time each stage with
 * events; the sequence is equivalent to search_batch.  No gaffa counterpart
 * (its executor launches per-level kernels itself). */
int gaffa_hip_ffa_subtree_batch(const gaffa_hip_ffa_search_args* args,
                                gaffa_hip_stream stream);
int gaffa_hip_ffa_levels_batch(const gaffa_hip_ffa_search_args* args,
                               gaffa_hip_stream stream);
int gaffa_hip_ffa_detect_batch(const gaffa_hip_ffa_search_args* args,
                               gaffa_hip_stream stream);

/* Transform only: writes the first rows_eval final rows of task `task` for
 * every series, [nseries][rows_eval][bins] (testing and ffa_transform_block
 * parity; replaces ffa_transform_block_cuda, ffa_cuda.h:240-263). */
int gaffa_hip_ffa_transform_rows(const gaffa_hip_ffa_search_args* args,
                                 uint32_t task, float* out,
                                 gaffa_hip_stream stream);

/* ------------------------------------------------- candidate selection ----
 * gaffa's candidate selection (select_candidates_cpu, candidate.cpp:175-212)
 * over the compact peaks of a DM search, where `series` is the DM index:
 * single-linkage clusters along frequency (a gap wider than the radius
 * between consecutive present trial frequencies splits), then along DM index
 * inside each frequency cluster (and width, unless clustered across widths).
 * gaffa runs it on the host only; these entry points replace it for peaks
 * already on the device.  Trial periods enter as dense ranks (equal periods
 * share a rank, ascending period = ascending rank), and frequencies as
 * descending ranks: rdesc = nranks - 1 - rank.
 *
 * 1. gaffa_hip_candidate_ranks: peak_rank[i] = period_rank[task_first_row[
 *    task] + shift] of peak i, present[rank] = 1 for every rank a peak has
 *    (present is cleared first), status = 0 or the GAFFA_HIP_CANDIDATE_*
 *    bits of the peaks that fail (a task, row, DM or width out of range, or
 *    a non-finite S/N).
 * 2. The caller numbers the frequency clusters over the present ranks in
 *    rdesc order: cluster_of_rdesc[rd] (the CPU's frequency gaps).
 * 3. gaffa_hip_candidate_clusters: one record per DM cluster, in cluster
 *    order (frequency cluster, then DM / width): its descending-rank and DM
 *    index ranges, top S/N and peak count, plus a tie record (cluster, peak
 *    index) for every peak whose S/N equals its cluster's top; the caller
 *    picks each representative with gaffa's peak order.  segments and ties
 *    hold npeaks entries; the workspace is
 *    gaffa_hip_candidate_workspace_bytes(npeaks) bytes.  Steps 1 and 3 are
 *    asynchronous on `stream`; step 2 needs `present` on the host. */
#define GAFFA_HIP_CANDIDATE_BAD_PEAK 1u
#define GAFFA_HIP_CANDIDATE_NONFINITE 2u

typedef struct gaffa_hip_candidate_segment {
  uint32_t rdesc_min, rdesc_max; /* frequency range as descending ranks */
  uint32_t dm_min, dm_max;       /* DM index range */
  float snr;                     /* top S/N */
  uint32_t count;                /* peaks in the cluster */
} gaffa_hip_candidate_segment;

typedef struct gaffa_hip_candidate_tie {
  uint32_t segment; /* DM cluster */
  uint32_t peak;    /* index into the compact peaks */
} gaffa_hip_candidate_tie;

typedef struct gaffa_hip_candidate_args {
  const gaffa_hip_ffa_peak* peaks; /* device [npeaks] */
  uint64_t npeaks;
  const uint32_t* period_rank;     /* device [task_first_row[ntasks]] */
  const uint64_t* task_first_row;  /* device [ntasks + 1] */
  uint32_t ntasks;
  uint32_t nranks;
  uint32_t ndm;
  uint32_t nwidths;
  uint32_t* peak_rank;             /* device [npeaks] (step 1 out, step 3 in) */
  uint8_t* present;                /* device [nranks] */
  uint32_t* status;                /* device [1] */
  const uint32_t* cluster_of_rdesc; /* device [nranks] */
  uint32_t nclusters;
  uint32_t dm_cluster_radius;      /* DM index gap that still joins */
  uint32_t cluster_across_widths;  /* 0: widths kept apart */
  void* workspace;                 /* device */
  uint64_t workspace_bytes;
  gaffa_hip_candidate_segment* segments; /* device [npeaks] */
  uint32_t* nsegments;             /* device [1] */
  gaffa_hip_candidate_tie* ties;   /* device [npeaks] */
  uint32_t* nties;                 /* device [1] */
} gaffa_hip_candidate_args;

int gaffa_hip_candidate_workspace_bytes(uint64_t npeaks, uint64_t* bytes);
int gaffa_hip_candidate_ranks(const gaffa_hip_candidate_args* args, gaffa_hip_stream stream);
int gaffa_hip_candidate_clusters(const gaffa_hip_candidate_args* args, gaffa_hip_stream stream);

/* ------------------------------------------------------------- folding ----
 * Fractional-phase-overlap folds (folding.cpp:300-420): sample t of a job's
 * input (data + input_offset + t * row_stride, channels contiguous) covers
 * phase [i s, i s + s) bins, i = start_sample + t, s = tsamp / period * nbin,
 * and adds value * overlap (FP64) to each overlapped bin; out = sum /
 * exposure where exposure > 0, else 0.  Adjacent input channels are averaged
 * into output_channels (nchans % output_channels == 0).  Layouts:
 * out [njobs][nsubint][output_channels][nbin] float, exposure
 * [njobs][nsubint][nbin] double (may be NULL).  nsubint = ceil(nsamples /
 * subint_samples).  Every FP64 sum is formed in the CPU's order, so results
 * equal gaffa's fold_spectrum_cpu bit for bit when period > tsamp.
 * Replaces fold_spectrum_cpu / fold_time_series_cpu for device data (gaffa
 * folds on the host only). */
typedef enum gaffa_hip_fold_dtype {
  GAFFA_HIP_FOLD_U8 = 0,
  GAFFA_HIP_FOLD_U16 = 1,
  GAFFA_HIP_FOLD_U32 = 2,
  GAFFA_HIP_FOLD_F32 = 3,
} gaffa_hip_fold_dtype;

typedef struct gaffa_hip_fold_job {
  uint64_t input_offset; /* elements from data to this job's sample 0 */
  double period;         /* seconds */
  uint64_t start_sample; /* absolute index of sample 0 (phase reference) */
} gaffa_hip_fold_job;

typedef struct gaffa_hip_fold_args {
  const void* data;
  int32_t dtype;         /* gaffa_hip_fold_dtype */
  uint64_t nsamples;     /* samples per job */
  uint32_t nchans;
  uint64_t row_stride;   /* elements between consecutive samples */
  double tsamp;
  uint32_t nbin;
  uint64_t subint_samples;
  uint64_t nsubint;
  uint32_t output_channels;
  const gaffa_hip_fold_job* jobs; /* device [njobs] */
  uint32_t njobs;
  float* out;
  double* exposure;
  uint64_t channel_stride; /* elements between channels, 0 = 1 */
  double divisor;          /* the input holds sums of this many channels, 0 = 1 */
} gaffa_hip_fold_args;

int gaffa_hip_fold_batch(const gaffa_hip_fold_args* args, gaffa_hip_stream stream);

/* Subband sums of the aligned dynamic spectrum at one DM, the input of a
 * full fold onto nsel / subband_channels output channels:
 *   out[k * out_stride + t] = sum_{c in subband k} channels[c][t + delays[c]]
 * for t < out_nsamples, from the channel-major copy
 * (gaffa_hip_filterbank_transpose_u8).  Replaces dedisperse_spectrum_cuda
 * (dedispersion_cuda.cu:295-318) plus the per-output-channel averaging of
 * fold_spectrum_cpu (folding.cpp:354-360): folding these sums with
 * divisor = subband_channels (gaffa_hip_fold_args) gives gaffa's cube bit for
 * bit, since the uint8 channel sums are exact.  delays: device [nsel], each in
 * [0, max_delay]; rows readable for ceil(out_nsamples / 512) * 512 +
 * max_delay + 528 bytes; out 16-byte aligned, out_stride a multiple of 8. */
typedef struct gaffa_hip_subband_spectrum_args {
  const uint8_t* channels; /* [nsel][channel_stride] */
  uint64_t channel_stride;
  uint64_t channel_length; /* readable bytes per row, 0 = channel_stride */
  uint64_t nsel;
  uint64_t subband_channels;
  const int32_t* delays;
  int32_t max_delay;
  uint16_t* out;           /* [nsel / subband_channels][out_stride] */
  uint64_t out_stride;
  uint64_t out_nsamples;
} gaffa_hip_subband_spectrum_args;

int gaffa_hip_subband_spectrum_u8(const gaffa_hip_subband_spectrum_args* args,
                                  gaffa_hip_stream stream);

/* --------------------------------------------------- synthetic inputs ----
 * Seeded, integer-exact synthetic filterbank (no gaffa counterpart): sample
 * (t, c) = clip(pop1024 - 384 + floor((dither12 + signal_q + rfi_q)/4096)).
 * The host builds the pulsar tables (delays, 4096-bin 1/4096-unit profiles,
 * 2^-20-sample periods and phases). */
typedef struct gaffa_hip_synth_args {
  uint8_t* out; /* [t_count][nchans] */
  uint64_t nchans;
  uint64_t t_begin;
  uint64_t t_count;
  uint64_t key;            /* splitmix64(seed) */
  uint32_t npulsars;
  const int32_t* delays;   /* device [npulsars][nchans] */
  const int32_t* profiles; /* device [npulsars][4096] */
  const int64_t* period_q; /* device [npulsars] */
  const int64_t* phase0_q; /* device [npulsars] */
  const int32_t* rfi_q;    /* device [nchans] */
} gaffa_hip_synth_args;

/* Seeded batch of float time series packed in one row (no gaffa
 * counterpart; feeds the batched FFA): series s occupies
 * out[offsets[s] .. offsets[s] + lengths[s]) with N(0,1) noise (hashed
 * Box-Muller) plus, where amplitude[s] != 0, a Gaussian pulse train of the
 * given period (s), start phase and phase sigma; the rest of the row up to
 * row_length (padding) is 0.  offsets ascending, offsets[0] == 0. */
typedef struct gaffa_hip_synth_series_args {
  float* out;
  uint64_t row_length;
  uint32_t nseries;
  uint64_t key;
  double tsamp;
  const uint64_t* offsets;  /* device [nseries] */
  const uint64_t* lengths;  /* device [nseries] */
  const double* period;     /* device [nseries] */
  const double* phase0;     /* device [nseries] */
  const float* sigma;       /* device [nseries], phase units */
  const float* amplitude;   /* device [nseries], 0 = noise only */
} gaffa_hip_synth_series_args;

int gaffa_hip_synth_series(const gaffa_hip_synth_series_args* args,
                           gaffa_hip_stream stream);

int gaffa_hip_synth_filterbank(const gaffa_hip_synth_args* args,
                               gaffa_hip_stream stream);

/* ----------------------------------------------------- candidate scorer ----
 * No gaffa counterpart: a learned score that decides which search
 * candidates go on to full folding.  Inputs per candidate: a 32-subint x
 * 64-phase folded cube (zero mean, unit variance) and a 128-trial
 * periodogram patch (mean removed, S/N units).  Network (fixed shape):
 *   conv1..conv4: circular 1D convolutions over phase, kernel 3,
 *                 32 -> 64 -> 64 -> 64 -> 64 channels (the subints are
 *                 conv1's input channels), each + bias, ReLU;
 *   pool:         mean and max over phase (128) ++ the patch (128) = 256;
 *   fc1:          256 -> 512, + bias, ReLU;  fc2: 512 -> 1 logit.
 * Weights and activations are bfloat16 (round to nearest even), every
 * product sum accumulates in FP32 on MFMA (v_mfma_f32_16x16x32_bf16).
 * Parameters live in one FP32 master vector (layout below); a bfloat16
 * copy feeds the kernels and is refreshed by gaffa_hip_scorer_adam /
 * gaffa_hip_scorer_sync_params. */

#define GAFFA_HIP_SCORER_SUBINTS 32
#define GAFFA_HIP_SCORER_PHASE_BINS 64
#define GAFFA_HIP_SCORER_PATCH 128
#define GAFFA_HIP_SCORER_CHANNELS 64
#define GAFFA_HIP_SCORER_HIDDEN 512
#define GAFFA_HIP_SCORER_LAYERS 6

/* Float offsets of each layer's weight and bias in the parameter vector.
 * Conv weights are [out][tap][in] (tap 0 reads phase p - 1), fc weights
 * [out][in]. */
typedef struct gaffa_hip_scorer_layout {
  uint64_t weight[GAFFA_HIP_SCORER_LAYERS];
  uint64_t bias[GAFFA_HIP_SCORER_LAYERS];
  uint64_t fan_in[GAFFA_HIP_SCORER_LAYERS];
  uint64_t nparams;        /* padded to a multiple of 8 */
} gaffa_hip_scorer_layout;

int gaffa_hip_scorer_get_layout(gaffa_hip_scorer_layout* layout);

/* Device workspace for up to `batch` candidates per call. */
uint64_t gaffa_hip_scorer_workspace_bytes(uint32_t batch);

typedef struct gaffa_hip_scorer_net {
  float* params;      /* [nparams] FP32 master */
  uint16_t* params16; /* [nparams] bfloat16 copy used by the kernels */
  float* grads;       /* [nparams], written by backward */
  float* adam_m;      /* [nparams] */
  float* adam_v;      /* [nparams] */
  void* workspace;    /* gaffa_hip_scorer_workspace_bytes(capacity) */
  uint32_t capacity;  /* batch capacity of the workspace */
} gaffa_hip_scorer_net;

/* Forward pass of `batch` candidates.  Candidate b reads cube
 * rows index[b] (or b when index is null) of cubes [N][32][64] and
 * patches [N][128] (FP32, device).  Writes logits[b] (device, FP32) and
 * keeps the activations in the workspace for gaffa_hip_scorer_backward. */
int gaffa_hip_scorer_forward(const gaffa_hip_scorer_net* net, const float* cubes,
                             const float* patches, const uint32_t* index, uint32_t batch,
                             float* logits, gaffa_hip_stream stream);

/* Backward pass of the binary cross-entropy with logits, mean over the
 * batch, for the last forward: labels[b] (device, 0 or 1, read at the same
 * index as the inputs).  Writes every parameter gradient to net->grads and
 * the loss to *loss (device). */
int gaffa_hip_scorer_backward(const gaffa_hip_scorer_net* net, const float* labels,
                              const uint32_t* index, uint32_t batch, float* loss,
                              gaffa_hip_stream stream);

/* One Adam step (bias-corrected, step >= 1) on the FP32 master, then the
 * bfloat16 copy is refreshed. */
int gaffa_hip_scorer_adam(const gaffa_hip_scorer_net* net, float lr, float beta1, float beta2,
                          float eps, uint32_t step, gaffa_hip_stream stream);

/* params16 = bfloat16(params), after the master was written directly. */
int gaffa_hip_scorer_sync_params(const gaffa_hip_scorer_net* net, gaffa_hip_stream stream);

/* Scorer input preparation.  Patches: out[i][q] = periodogram[series[i] *
 * series_stride + clamp(row[i] - 64 + q, 0, rows - 1) * widths + width[i]]
 * for q < 128 (the trials around the candidate's, edges repeated), minus
 * the patch mean (the values stay in S/N units, so the network sees how
 * strong the candidate is).  normalise_rows: each row of x [rows][length]
 * to zero mean and unit variance (FP32 mean and variance; a constant row
 * becomes 0). */
typedef struct gaffa_hip_scorer_patch_args {
  const float* periodogram;
  uint64_t series_stride;
  uint64_t rows;
  uint32_t widths;
  const uint32_t* series; /* device [n] */
  const uint32_t* row;    /* device [n] */
  const uint32_t* width;  /* device [n] */
  uint32_t n;
  float* out;             /* device [n][128] */
} gaffa_hip_scorer_patch_args;

int gaffa_hip_scorer_gather_patches(const gaffa_hip_scorer_patch_args* args,
                                    gaffa_hip_stream stream);
int gaffa_hip_scorer_normalise_rows(float* x, uint32_t rows, uint32_t length,
                                    gaffa_hip_stream stream);

/* The MFMA GEMM the scorer is built from, exposed for testing:
 * C(i, j) = sum_k A(i, k) B(k, j) over bfloat16 A(i, k) = a[i*a_row + k*a_col]
 * and B(k, j) = b[k*b_row + j*b_col], FP32 accumulation; then + bias[j]
 * (if bias), max(., 0) (if relu), stored to c32 and/or c16 (bfloat16) at
 * [i*ldc + j].  Split over k_splits workgroups along k (partial sums in
 * `partial`, [k_splits][m][n] floats, added in split order). */
typedef struct gaffa_hip_gemm_args {
  const uint16_t* a;
  int64_t a_row, a_col;
  const uint16_t* b;
  int64_t b_row, b_col;
  uint32_t m, n, k;
  const float* bias;
  uint32_t relu;
  float* c32;
  uint16_t* c16;
  uint64_t ldc;
  uint32_t k_splits;
  float* partial;
} gaffa_hip_gemm_args;

int gaffa_hip_gemm_bf16(const gaffa_hip_gemm_args* args, gaffa_hip_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* GAFFA_HIP_H */
