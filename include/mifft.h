/*
 * mifft.h -- C ABI of the MI355X (gfx950) FFT engine.
 *
 * This is the drop-in boundary for wgpuFFT's public plan surface
 * (FftConfig / FftPlan / FftRecorder, src/config.rs, src/plan.rs).  wgpu
 * objects become HIP objects: a wgpu::Buffer becomes a device pointer plus its
 * size in bytes, a CommandEncoder / compute pass becomes a hipStream_t, and
 * "recording" becomes enqueueing on that stream (optionally captured into a
 * hipGraph).  Every call returns an mifft_status; nothing aborts or panics.
 *
 * Layouts are the reference's (README.md:111-128): axis 0 is contiguous,
 * complex data is interleaved [re, im], df64 is [re_hi, re_lo, im_hi, im_lo],
 * and R2C output is the packed spectrum [floor(N0/2)+1, N1, ...].
 */
#ifndef MIFFT_H
#define MIFFT_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MIFFT_MAX_RANK 8
#define MIFFT_MAX_FORCED_AXES 8

/* Status codes: one per wgpuFFT FftError variant on this path (src/error.rs:12-180). */
typedef enum mifft_status {
    MIFFT_SUCCESS = 0,
    MIFFT_ERR_ZERO_LENGTH = 1,                 /* FftError::ZeroLength */
    MIFFT_ERR_ZERO_BATCH = 2,                  /* FftError::ZeroBatch */
    MIFFT_ERR_LENGTH_TOO_LARGE = 3,            /* FftError::LengthTooLarge */
    MIFFT_ERR_UNSUPPORTED_LENGTH = 4,          /* FftError::UnsupportedLength */
    MIFFT_ERR_PRECISION_UNSUPPORTED = 5,       /* FftError::PrecisionUnsupported */
    MIFFT_ERR_INVALID_TUNING = 6,              /* FftError::InvalidTuning */
    MIFFT_ERR_EMPTY_AXES = 7,                  /* FftError::EmptyAxes */
    MIFFT_ERR_INVALID_AXIS = 8,                /* FftError::InvalidAxis */
    MIFFT_ERR_DUPLICATE_AXIS = 9,              /* FftError::DuplicateAxis */
    MIFFT_ERR_INVALID_REAL_DIRECTION = 10,     /* FftError::InvalidRealTransformDirection */
    MIFFT_ERR_UNSUPPORTED_REAL_AXES = 11,      /* FftError::UnsupportedRealAxes */
    MIFFT_ERR_BUFFER_TOO_SMALL = 12,           /* FftError::BufferViewTooSmall */
    MIFFT_ERR_WORKSPACE_TOO_SMALL = 13,        /* FftError::WorkspaceTooSmall */
    MIFFT_ERR_IN_PLACE_UNSUPPORTED = 14,       /* FftError::InPlaceUnsupported */
    MIFFT_ERR_ALLOCATION_FAILED = 15,          /* FftError::GpuPlanResourceAllocationFailed */
    MIFFT_ERR_HOST_BUFFER_LENGTH = 16,         /* FftError::HostBufferLength */
    MIFFT_ERR_NULL_ARGUMENT = 17,              /* no counterpart: a NULL handle or pointer */
    MIFFT_ERR_INVALID_ENUM = 18,               /* no counterpart: enum value out of range */
    MIFFT_ERR_RANK_TOO_LARGE = 19,             /* no counterpart: rank > MIFFT_MAX_RANK */
    MIFFT_ERR_HIP = 20,                        /* a HIP runtime call failed */
    MIFFT_ERR_RECORDER_STATE = 21,             /* recorder used after end */
    MIFFT_ERR_OVER_BUDGET = 22,                /* server admission: request exceeds the memory budget */
    MIFFT_ERR_BUSY = 23,                       /* server: no free slot, try again */
    MIFFT_ERR_UNALIGNED = 24,                  /* FftError::BufferViewOffsetUnaligned: device
                                                  buffers must be 16-byte aligned */
    MIFFT_ERR_LAYOUT_ZERO_STRIDE = 25,         /* FftError::BufferLayoutZeroStride */
    MIFFT_ERR_LAYOUT_BATCH_STRIDE = 26,        /* FftError::BufferLayoutBatchStrideTooSmall */
    MIFFT_ERR_LAYOUT_OUT_OF_BOUNDS = 27,       /* FftError::BufferLayoutOutOfBounds */
    MIFFT_ERR_LAYOUT_TOO_LARGE = 28            /* FftError::BufferLayoutTooLarge */
} mifft_status;

/* FftTransformKind (plan.rs:24-29). */
typedef enum mifft_kind { MIFFT_C2C = 0, MIFFT_R2C = 1, MIFFT_C2R = 2 } mifft_kind;
/* FftDirection (config.rs:42-47). */
typedef enum mifft_direction { MIFFT_FORWARD = 0, MIFFT_INVERSE = 1 } mifft_direction;
/* Normalization (config.rs:59-71); default INVERSE. */
typedef enum mifft_normalization {
    MIFFT_NORM_NONE = 0,
    MIFFT_NORM_FORWARD = 1,
    MIFFT_NORM_INVERSE = 2,
    MIFFT_NORM_ORTHOGONAL = 3
} mifft_normalization;
/* FftPrecision (config.rs:5-17). */
typedef enum mifft_precision { MIFFT_F32 = 0, MIFFT_F64 = 1, MIFFT_DF64 = 2 } mifft_precision;
/* AxisKind (runtime/axis_policy.rs). */
typedef enum mifft_axis_kind {
    MIFFT_AXIS_MIXED = 0,
    MIFFT_AXIS_RADER = 1,
    MIFFT_AXIS_BLUESTEIN = 2
} mifft_axis_kind;
/* C2cRoute (runtime/c2c.rs:7631-7650). */
typedef enum mifft_route {
    MIFFT_ROUTE_DIRECT_DFT = 0,
    MIFFT_ROUTE_MIXED_RADIX = 1,
    MIFFT_ROUTE_RADER = 2,
    MIFFT_ROUTE_BLUESTEIN = 3,
    MIFFT_ROUTE_AXIS_SEQUENCE = 4
} mifft_route;

/* FftTuning (tuning.rs:35-77); mifft_tuning_init() gives the reference defaults. */
typedef struct mifft_tuning {
    uint32_t workgroup_size;               /* 64; power of two */
    uint32_t fused_workgroup_size;         /* 256; power of two */
    uint64_t rader_max_prime;              /* 8192; >= 2 */
    uint64_t direct_max_prime;             /* 127 */
    int32_t force_rader_axes[MIFFT_MAX_FORCED_AXES];
    int32_t n_force_rader_axes;
    int32_t force_bluestein_axes[MIFFT_MAX_FORCED_AXES];
    int32_t n_force_bluestein_axes;
    int32_t fuse_long_axes;                /* 1 */
    int32_t fuse_small_volumes;            /* 1 */
    uint64_t fused_min_convolution_length; /* 128 */
} mifft_tuning;

/* FftConfig (config.rs:75-84); mifft_config_init() mirrors FftConfig::new_nd. */
typedef struct mifft_config {
    uint64_t shape[MIFFT_MAX_RANK];
    int32_t rank;
    int32_t axes[MIFFT_MAX_RANK];
    int32_t naxes;
    uint64_t batch;
    int32_t direction;     /* mifft_direction */
    int32_t normalization; /* mifft_normalization */
    int32_t precision;     /* mifft_precision */
    mifft_tuning tuning;
} mifft_config;

typedef struct mifft_plan mifft_plan;
typedef struct mifft_recorder mifft_recorder;

/* ------------------------------------------------------------- config */
/* FftTuning::default (tuning.rs:55-77). */
void mifft_tuning_init(mifft_tuning* t);
/* FftConfig::new_nd (config.rs:90-103): all axes, batch 1, forward, Inverse, f32. */
mifft_status mifft_config_init(mifft_config* c, int32_t rank, const uint64_t* shape);
/* FftConfig::validate (config.rs:178-211) with FftTuning::validate_for_config (tuning.rs:280-336). */
mifft_status mifft_config_validate(const mifft_config* c);
/* FftConfig::scale_f64 (config.rs:274-286); scale uses prod(shape). */
mifft_status mifft_config_scale(const mifft_config* c, double* scale);
/* FftConfig::total_complex_len (config.rs:218-229). */
mifft_status mifft_config_total_len(const mifft_config* c, uint64_t* len);

/* --------------------------------------------------------------- plans */
/* FftPlan::c2c / r2c / c2r (plan.rs:221,327,355) on the current HIP device.
 * Plan creation uploads device tables (twiddles, chirps, filter spectra) and
 * synchronises; the plan is immutable afterwards. */
mifft_status mifft_plan_create(const mifft_config* c, mifft_kind kind, mifft_plan** out);
void mifft_plan_destroy(mifft_plan* p);
/* FftPlan::c2c_with_diagnostics / c2c_checked_with_diagnostics (plan.rs:227-261):
 * mifft_plan_create, and on failure the FftPlanCreationError diagnostics as
 * JSON (error.rs:210-228): {"error": <FftError variant>, "route": {...},
 * "stages": [], "blockers": [{kind, route, stage, layout, helper_buffer,
 * reason, required_bytes, actual_bytes, limit_bytes}], "active_tuning": {...}}
 * with the blocker FftError::diagnostics gives the error (error.rs:498-850);
 * on success the plan's diagnostics.  Writes at most cap bytes
 * (NUL-terminated) and the full size to *needed. */
mifft_status mifft_plan_create_with_diagnostics(const mifft_config* c, mifft_kind kind, mifft_plan** out, char* buf,
                                                size_t cap, size_t* needed);
/* The diagnostics of the last failing plan creation or execution on this
 * thread (FftPlanCreationError / FftExecutionError::diagnostics,
 * error.rs:180-228): for an execution, the plan's diagnostics with the
 * blockers of the buffers it was given; "null" when the last failure was
 * neither. */
mifft_status mifft_last_error_diagnostics(char* buf, size_t cap, size_t* needed);

/* Queries (plan.rs:390-492, 815-820). */
mifft_kind mifft_plan_kind(const mifft_plan* p);
mifft_status mifft_plan_config(const mifft_plan* p, mifft_config* out);
/* FftPlan::factors: radices of the first transformed axis; returns the count. */
int32_t mifft_plan_factors(const mifft_plan* p, uint64_t* out, int32_t cap);
/* FftPlan::axis_factors: radices of transformed axis `axis_index` (in config order). */
int32_t mifft_plan_axis_factors(const mifft_plan* p, int32_t axis_index, uint64_t* out, int32_t cap);
/* FftPlan::axis_kinds: one mifft_axis_kind per transformed axis; returns the count. */
int32_t mifft_plan_axis_kinds(const mifft_plan* p, int32_t* out, int32_t cap);
mifft_route mifft_plan_route(const mifft_plan* p);
uint64_t mifft_plan_workspace_bytes(const mifft_plan* p);
uint64_t mifft_plan_required_input_bytes(const mifft_plan* p);
uint64_t mifft_plan_required_output_bytes(const mifft_plan* p);
uint64_t mifft_plan_required_bytes(const mifft_plan* p);
/* FftPlan::packed_shape: rank entries for R2C/C2R, 0 for C2C. */
int32_t mifft_plan_packed_shape(const mifft_plan* p, uint64_t* out, int32_t cap);
int32_t mifft_plan_supports_in_place(const mifft_plan* p);
/* Number of kernel launches one execution enqueues, and a text description
 * of them ("reg N=4096 16x16x16 lines=2048; ..."), for diagnostics. */
int32_t mifft_plan_launch_count(const mifft_plan* p);
mifft_status mifft_plan_describe(const mifft_plan* p, char* buf, size_t cap);
/* FftPlan::diagnostics (plan.rs:938-1004, diagnostics.rs:8-140) as a JSON
 * object: route {transform, route, execution_kind, reason_codes,
 * attempted_routes}, stages [{label, kind, route}] (one per launch),
 * blockers [], buffer_requirements [{role, required_bytes, format}],
 * active_tuning {FftTuning fields}, device_limits.  Writes at most cap bytes
 * (NUL-terminated) and the full size including the NUL to *needed. */
mifft_status mifft_plan_diagnostics(const mifft_plan* p, char* buf, size_t cap, size_t* needed);
/* FftPlan::diagnostics_for_workspace (plan.rs:1023-1055): the plan's
 * diagnostics with the blockers executing it over these buffers would meet
 * (an input or output smaller than required, a pointer not 16-byte aligned, a
 * workspace smaller than mifft_plan_workspace_bytes); pointers may be NULL
 * (alignment not checked).  An empty "blockers" list means the execution
 * would be accepted. */
mifft_status mifft_plan_diagnostics_for_workspace(const mifft_plan* p, const void* in, uint64_t in_bytes,
                                                  const void* out, uint64_t out_bytes, const void* workspace,
                                                  uint64_t workspace_bytes, char* buf, size_t cap, size_t* needed);
/* Text of launch `index` (0 <= index < mifft_plan_launch_count). */
mifft_status mifft_plan_launch_describe(const mifft_plan* p, int32_t index, char* buf, size_t cap);
/* The tuning-table key of launch `index` ("" when the launch takes no table
 * entry) and the name of the registered line-kernel variant it runs ("" for
 * runtime-schedule and non-line kernels).  No single counterpart in the
 * reference, whose pipeline-cache keys are internal (pipeline_cache.rs:2048-2840);
 * this names the autotuner's choice per launch so that a test can force and
 * check every variant.  Either buffer may be NULL. */
mifft_status mifft_plan_launch_tuning(const mifft_plan* p, int32_t index, char* key, size_t key_cap, char* variant,
                                      size_t variant_cap);
/* Plan-owned device tables, copied back for inspection (the reference
 * builds its tables on the host and checks them in its tests,
 * twiddle.rs:198-237, bluestein_axis.rs:818-867, rader_axis.rs:1291-1336).
 * Each table has a text description of what it holds, space-separated:
 *   "stage-twiddles prec=<f32|f64> radices=<R0,R1,..> count=<c>": entry
 *     off_s + (m-1) L_s + k = W_{L_s R_s}^{k m} for stages s >= 1 (L_s the
 *     product of the radices before s, off_s the entries of stages 1..s-1);
 *   "chirp prec=.. n=<n>": entry i = W_{2n}^{i^2 mod 2n};
 *   "bluestein-filter n=<n> prec=.. m=<m>": FFT_m(b)/m, b[i] = conj(chirp[i])
 *     for |i| < n, indices mod m, zero elsewhere;
 *   "rader-filter p=<p> ginv=<h> prec=.. m=<m>": FFT_m(b)/m, b[t] =
 *     W_p^{h^t mod p} for t < p-1 and b[m-t] = b[p-1-t] when m > p-1;
 *   "rader-perm-in p=<p> g=<g>", "rader-perm-out p=<p> ginv=<h>": u32
 *     entries g^q mod p, h^q mod p;
 *   "roots prec=.. n=<n> count=<c>": entry k = W_n^k;
 *   "roots-strided prec=f64 n=<n> count=<c> mult=<s>": entry k = W_n^{k s mod n}.
 * W_n^t = exp(-2 pi i t / n); complex entries are interleaved [re, im] of
 * the stated precision, generated in FP64 on the device with the phase
 * reduced in integers and rounded once.  mifft_plan_table writes the
 * description (desc may be NULL) and the size to *bytes, and with `host`
 * non-NULL copies the table (synchronously) into host_cap >= *bytes bytes. */
int32_t mifft_plan_table_count(const mifft_plan* p);
mifft_status mifft_plan_table(const mifft_plan* p, int32_t index, char* desc, size_t desc_cap, void* host,
                              uint64_t host_cap, uint64_t* bytes);
/* twiddle_lut_f32 / twiddle_lut_f64 (runtime/twiddle.rs:13-37) on the
 * device: dst[k] = W_n^k for k < count (k taken mod n), FP64 with the phase
 * reduced in integers, rounded once to `precision` (MIFFT_F32 or MIFFT_F64),
 * enqueued on `stream`. */
mifft_status mifft_twiddle_table(void* dst, uint64_t n, uint64_t count, int32_t precision, hipStream_t stream);
/* The compiled register-kernel registry (the autotuner's search space: radix
 * split, threads per line, lines per workgroup, pipelined and fold forms):
 * its size and the name of entry `index`, the form a tuning-table choice uses. */
int32_t mifft_registry_count(void);
mifft_status mifft_registry_name(int32_t index, char* buf, size_t cap);
/* Enqueues only launch `index` of one execution (no counterpart in the
 * reference: per-kernel timing for roofline measurement).  The buffers are
 * those the full execution would use, validated as mifft_execute validates
 * them (sizes, 16-byte alignment); in == out only for plans that support in
 * place (MIFFT_ERR_IN_PLACE_UNSUPPORTED otherwise: a real plan's single launch
 * cannot take the input copy a whole in-place execution takes).  `workspace`
 * may be NULL to use the plan's own. */
mifft_status mifft_execute_launch(const mifft_plan* p, hipStream_t stream, int32_t index, const void* in,
                                  uint64_t in_bytes, void* out, uint64_t out_bytes, void* workspace,
                                  uint64_t workspace_bytes);

/* ------------------------------------------------------------ execution */
/* FftPlan::execute_checked (plan.rs:510-518): input -> output on `stream`.
 * The input is left unchanged.  Buffer sizes are validated like
 * BufferView::validate_min_size (BufferViewTooSmall). */
mifft_status mifft_execute(const mifft_plan* p, hipStream_t stream, const void* in, uint64_t in_bytes,
                           void* out, uint64_t out_bytes);
/* FftPlan::execute_with_workspace (plan.rs:535-550): caller-owned workspace,
 * which lets one plan run concurrently on several streams. */
mifft_status mifft_execute_with_workspace(const mifft_plan* p, hipStream_t stream, const void* in,
                                          uint64_t in_bytes, void* out, uint64_t out_bytes,
                                          void* workspace, uint64_t workspace_bytes);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): mifft_execute_with_workspace with the workspace taken
 * from the library's static device scratch, MIFFT_STATIC_SCRATCH_BYTES per
 * GPU (held by every GPU the library runs on until the process ends).  It
 * calls no hipMalloc / hipFree.  In place on a real plan the input copy also
 * goes into the scratch.  An execution needing more scratch than that fails
 * with MIFFT_ERR_WORKSPACE_TOO_SMALL before any GPU work.  Every call on one
 * GPU shares that GPU's scratch: executions that overlap on one GPU (other
 * streams, other threads) race and must not be issued; calls in order on one
 * stream are safe.  The plan's kernels write every scratch element they read,
 * so nothing needs clearing between calls. */
#define MIFFT_STATIC_SCRATCH_BYTES (80ull << 20)
mifft_status mifft_execute_static_scratch(const mifft_plan* p, hipStream_t stream, const void* in,
                                          uint64_t in_bytes, void* out, uint64_t out_bytes);
/* FftPlan::execute_in_place (plan.rs:827-835). */
mifft_status mifft_execute_in_place(const mifft_plan* p, hipStream_t stream, void* buf, uint64_t bytes);

/* ------------------------------------------------------------ recorder */
/* FftRecorder (plan.rs:38-95): executions recorded into one recorder run in
 * order on its stream; with capture_graph != 0 they are captured into one
 * hipGraph, instantiated by mifft_recorder_end and replayed by
 * mifft_recorder_launch.  Caller work may be enqueued on
 * mifft_recorder_stream() between records (FftRecorder::compute_pass). */
mifft_status mifft_recorder_begin(hipStream_t stream, int32_t capture_graph, mifft_recorder** out);
hipStream_t mifft_recorder_stream(const mifft_recorder* r);
mifft_status mifft_record(mifft_recorder* r, const mifft_plan* p, const void* in, uint64_t in_bytes,
                          void* out, uint64_t out_bytes);
mifft_status mifft_record_in_place(mifft_recorder* r, const mifft_plan* p, void* buf, uint64_t bytes);
mifft_status mifft_recorder_end(mifft_recorder* r);
mifft_status mifft_recorder_launch(mifft_recorder* r, hipStream_t stream);
void mifft_recorder_destroy(mifft_recorder* r);

/* ------------------------------------------------------- batched launch */
/* Many 1D f32 C2C transforms of uneven lengths and line counts in one
 * persistent launch (two when a length needs a four-step split).  No
 * counterpart entry point in the reference: it replaces recording one plan per
 * transform into an FftRecorder (plan.rs:38-95).  A transform's result does
 * not depend on the other items of the batch (bit for bit).  Buffers are
 * bound at creation. */
typedef struct mifft_batch_item {
    uint64_t n;            /* transform length */
    uint64_t batch;        /* lines (contiguous, n apart) */
    int32_t direction;     /* mifft_direction */
    int32_t normalization; /* mifft_normalization */
    int32_t algorithm;     /* 0 route choice, 1 force Rader, 2 force Bluestein */
    int32_t reserved;
    const void* in;        /* device, interleaved complex f32, 16-byte aligned */
    void* out;             /* may equal in */
} mifft_batch_item;
typedef struct mifft_batch mifft_batch;

mifft_status mifft_batch_create(const mifft_batch_item* items, int32_t count, const mifft_tuning* tuning,
                                mifft_batch** out);
/* Enqueues the persistent launch(es) on `stream`. */
mifft_status mifft_batch_execute(mifft_batch* b, hipStream_t stream);
/* Captures one execution into a hipGraph kept by the batch (first call only). */
mifft_status mifft_batch_capture(mifft_batch* b, hipStream_t stream);
/* Replays the captured graph on `stream`. */
mifft_status mifft_batch_launch(mifft_batch* b, hipStream_t stream);
mifft_status mifft_batch_info(const mifft_batch* b, uint64_t* tiles, uint64_t* phases, uint64_t* lds_bytes,
                              uint64_t* workspace_bytes, uint32_t* grid);
void mifft_batch_destroy(mifft_batch* b);

/* -------------------------------------------------------------- serving */
/* Continuous-batching request loop on one GPU (no counterpart in the
 * reference, which leaves submission to the caller).  A tick packs the
 * admitted requests the batched kernel can run (1D / 2D C2C f32 and f64, R2C,
 * below the batch limit) into one persistent batched launch per pass and runs
 * the others through resident plans; up to `streams` ticks overlap on the device;
 * requests are admitted while input + output + workspace bytes of the work on
 * the GPU stay within the budget.  Request results do not depend on what else
 * is in flight. */
typedef struct mifft_server_request {
    int32_t kind;          /* mifft_kind */
    int32_t precision;     /* mifft_precision */
    int32_t rank;          /* 1 or 2 */
    int32_t direction;     /* mifft_direction */
    int32_t normalization; /* mifft_normalization */
    int32_t reserved;
    uint64_t shape[2];
    uint64_t batch;
    const void* in;
    void* out;
    uint64_t in_bytes;
    uint64_t out_bytes;
} mifft_server_request;

typedef struct mifft_server_result {
    double arrival_s; /* trace time of arrival */
    double admit_s;   /* when its tick was enqueued */
    double done_s;    /* when its completion was observed */
    int32_t status;   /* MIFFT_SUCCESS or MIFFT_ERR_OVER_BUDGET */
    int32_t path;     /* 0 batched launch, 1 resident plan, 2 rejected */
} mifft_server_result;

typedef struct mifft_server mifft_server;

mifft_status mifft_server_create(uint64_t memory_budget_bytes, int32_t streams, int32_t max_tick_requests,
                                 mifft_server** out);
/* Requests of at least `max_points` points (shape product x batch) run through
 * their resident plans beside the batched launch instead of inside it
 * (default 2^19: such a transform fills the GPU by itself).  0 batches nothing;
 * UINT64_MAX batches everything the batched kernel can run. */
mifft_status mifft_server_set_batch_limit(mifft_server* s, uint64_t max_points);
/* Builds and keeps resident the plans of every distinct request shape, plus
 * the per-stream workspaces, descriptor arenas and pinned staging. */
mifft_status mifft_server_prepare(mifft_server* s, const mifft_server_request* reqs, int32_t n);
/* Replays an arrival trace (seconds from start).  mode 0: continuous
 * batching; mode 1: sequential per-plan dispatch on one stream (the baseline). */
mifft_status mifft_server_replay(mifft_server* s, const mifft_server_request* reqs, const double* arrival_s,
                                 int32_t n, int32_t mode, mifft_server_result* results);
/* Online serving (no counterpart in the reference): a request of any shape is
 * submitted as it arrives; its plans are built on first sight of its shape
 * and kept resident (a bounded cache, least recently used shapes evicted when
 * nothing queued or in flight uses them); it waits in a FIFO queue and is
 * admitted, by the rules of mifft_server_replay, into the next tick on a free
 * stream slot while the bytes on the GPU stay within the budget.  Submission
 * and polling advance the loop; nothing runs in the background.  The caller
 * keeps the request's buffers until its completion is polled.  Results do not
 * depend on what else is in flight (bit for bit).  A request that can never
 * fit fails at submission with MIFFT_ERR_OVER_BUDGET. */
typedef struct mifft_server_completion {
    uint64_t ticket;  /* from mifft_server_submit */
    double submit_s;  /* seconds from the server's first submission */
    double admit_s;   /* its tick was enqueued */
    double done_s;    /* its completion was observed */
    int32_t status;   /* MIFFT_SUCCESS */
    int32_t path;     /* 0 batched launch, 1 resident plan */
} mifft_server_completion;
mifft_status mifft_server_submit(mifft_server* s, const mifft_server_request* req, uint64_t* ticket);
/* Advances the loop and returns up to `cap` completions in completion order;
 * with wait != 0 blocks (polling the device) until at least one completion is
 * available or nothing is queued or in flight. */
mifft_status mifft_server_poll(mifft_server* s, int32_t wait, mifft_server_completion* out, int32_t cap,
                               int32_t* count);
/* Requests queued, in flight, and completed but not yet polled. */
mifft_status mifft_server_pending(const mifft_server* s, uint64_t* queued, uint64_t* in_flight, uint64_t* completed);
/* Resident request shapes of the online path (default 256, at least 1) and how many are cached now. */
mifft_status mifft_server_set_plan_cache(mifft_server* s, uint64_t max_shapes);
uint64_t mifft_server_cached_shapes(const mifft_server* s);
/* Waits for the server's work on the GPU, then frees it. */
void mifft_server_destroy(mifft_server* s);

/* ------------------------------------------------------ logical views */
/* FftLogicalLayout (src/runtime/logical_io.rs:33-38): logical element i of
 * batch b sits at element element_offset + b * batch_stride + i * element_stride
 * of its buffer, in elements of the endpoint's format (complex for C2C and the
 * packed side of R2C / C2R, f32 for the real side).  batch_stride 0 means the
 * per-batch span, (elements per batch - 1) * element_stride + 1. */
typedef struct mifft_logical_layout {
    uint64_t element_offset;
    uint64_t element_stride;
    uint64_t batch_stride;
} mifft_logical_layout;

/* FftLogicalLayout::contiguous (logical_io.rs:90-96). */
void mifft_logical_layout_init(mifft_logical_layout* l);
/* FftLogicalLayout::required_element_span (logical_io.rs:123-145): the
 * elements the layout spans for `batch` batches of `per_batch` elements;
 * MIFFT_ERR_LAYOUT_ZERO_STRIDE / _BATCH_STRIDE / _TOO_LARGE on a bad layout. */
mifft_status mifft_logical_layout_span(const mifft_logical_layout* l, uint64_t per_batch, uint64_t batch,
                                       uint64_t* elements);
/* Workspace bytes of mifft_execute_logical_with_workspace for these layouts:
 * the plan's workspace plus contiguous staging for each strided side. */
mifft_status mifft_plan_logical_workspace_bytes(const mifft_plan* p, const mifft_logical_layout* in_layout,
                                                const mifft_logical_layout* out_layout, uint64_t* bytes);
/* FftPlan::diagnostics_for_views / _for_logical_views(_with_workspace)
 * (plan.rs:1057-1300): the same for logical layouts (NULL: contiguous):
 * layout blockers (zero element stride, batch stride too small, span too
 * large), views smaller than their layouts span, and with workspace_bytes
 * non-NULL a caller workspace smaller than mifft_plan_logical_workspace_bytes. */
mifft_status mifft_plan_diagnostics_for_views(const mifft_plan* p, const void* in, uint64_t in_bytes,
                                              const mifft_logical_layout* in_layout, const void* out,
                                              uint64_t out_bytes, const mifft_logical_layout* out_layout,
                                              const uint64_t* workspace_bytes, char* buf, size_t cap,
                                              size_t* needed);
/* FftPlan::execute_logical_views (plan.rs:734-746, c2c.rs:1564-1584,
 * real.rs:997-1030): the plan over caller layouts.  A side that is contiguous
 * (stride 1, batches back to back, 16-byte aligned start) is used in place; a
 * strided side is gathered into / scattered from contiguous staging, and the
 * elements between strided ones are left untouched.  NULL layouts are
 * contiguous.  Without a workspace the plan grows its own staging (one
 * in-flight execution per plan). */
mifft_status mifft_execute_logical(const mifft_plan* p, hipStream_t stream, const void* in, uint64_t in_bytes,
                                   const mifft_logical_layout* in_layout, void* out, uint64_t out_bytes,
                                   const mifft_logical_layout* out_layout);
/* FftPlan::execute_logical_views_with_workspace (plan.rs:748-763). */
mifft_status mifft_execute_logical_with_workspace(const mifft_plan* p, hipStream_t stream, const void* in,
                                                  uint64_t in_bytes, const mifft_logical_layout* in_layout,
                                                  void* out, uint64_t out_bytes,
                                                  const mifft_logical_layout* out_layout, void* workspace,
                                                  uint64_t workspace_bytes);
/* FftPlan::record_logical_views (plan.rs:917-936). */
mifft_status mifft_record_logical(mifft_recorder* r, const mifft_plan* p, const void* in, uint64_t in_bytes,
                                  const mifft_logical_layout* in_layout, void* out, uint64_t out_bytes,
                                  const mifft_logical_layout* out_layout);

/* ------------------------------------------------------------ autotuner */
/* No single counterpart in the reference: wgpuFFT compiles in per-shape
 * constants measured offline (src/runtime/axis_plan.rs:2092-2110,
 * register_fft.rs:64-118).  Here the planner consults a process-wide tuning
 * table of measured choices (a registered kernel variant: radix split,
 * threads per line, lines per workgroup; or a runtime-schedule LDS shape),
 * keyed by the shape of each line pass; without an entry it uses its
 * defaults.
 *
 * mifft_autotune builds the plan of (config, kind) and, pass by pass, times
 * the whole plan with each candidate of that pass swapped in (`reps`
 * executions per trial, three interleaved trials, HIP events on `stream`,
 * input and output buffers rotated beyond the 256 MB MALL, one workspace, in
 * place when the plan allows it), keeping each pass's fastest candidate for
 * the passes after it, and records the winners.  Plans created afterwards
 * use the table.  *passes_tuned (optional) receives the
 * number of passes that had more than one candidate; report (optional,
 * report_cap bytes) receives one text line per pass with every candidate's
 * time. */
mifft_status mifft_autotune(const mifft_config* config, mifft_kind kind, int32_t reps, hipStream_t stream,
                            int32_t* passes_tuned, char* report, uint64_t report_cap);
/* The same over `count` (1..16) plans executed back to back, as a caller runs
 * them: in place on one array when every plan allows it (a forward+inverse
 * pair), otherwise each plan reading the previous plan's output (R2C then
 * C2R).  Shapes are tuned in launch order; a candidate applies to every pass
 * of its shape in every plan and the whole sequence is timed. */
mifft_status mifft_autotune_sequence(const mifft_config* configs, const mifft_kind* kinds, int32_t count,
                                     int32_t reps, hipStream_t stream, int32_t* shapes_tuned, char* report,
                                     uint64_t report_cap);
/* Text form "key = choice", one entry per line ('#' starts a comment). */
mifft_status mifft_tuning_table_save(const char* path);
/* Adds the entries of a saved table; a malformed line or an unknown variant
 * fails with MIFFT_ERR_INVALID_TUNING and adds nothing. */
mifft_status mifft_tuning_table_load(const char* path);
mifft_status mifft_tuning_table_parse(const char* text, int32_t* entries);
void mifft_tuning_table_clear(void);
int32_t mifft_tuning_table_size(void);

/* ------------------------------------------------------ host (CPU) plans */
/* CpuFftPlan (src/cpu.rs:52-361): the same FftConfig semantics and buffer
 * layouts on host memory, for machines or callers without the GPU.  Buffers
 * are scalar words: f32 for F32 and Df64 plans ([re_hi, re_lo, im_hi, im_lo]
 * per Df64 element), f64 for F64 plans; R2C output / C2R input is the packed
 * spectrum [floor(N0/2)+1, N1, ...].  Work is split over host threads by
 * whole lines and tiles, so results do not depend on the thread count; long
 * contiguous lines in batches of fewer than 16 lines run as a four-step split
 * (cpu.rs:683-796), strided axes through tiles of neighbouring columns
 * (cpu.rs:526-681).  Executions of one plan may run concurrently. */
typedef struct mifft_cpu_plan mifft_cpu_plan;
/* CpuFftPlan::c2c / r2c / c2r (cpu.rs:88-210): r2c needs the forward and c2r
 * the inverse direction (MIFFT_ERR_INVALID_REAL_DIRECTION), every axis in
 * order (MIFFT_ERR_UNSUPPORTED_REAL_AXES) and F32 or F64
 * (MIFFT_ERR_PRECISION_UNSUPPORTED for Df64). */
mifft_status mifft_cpu_plan_create(const mifft_config* c, mifft_kind kind, mifft_cpu_plan** out);
void mifft_cpu_plan_destroy(mifft_cpu_plan* p);
/* CpuFftPlan::kind / config / packed_shape (cpu.rs:212-225). */
mifft_kind mifft_cpu_plan_kind(const mifft_cpu_plan* p);
mifft_status mifft_cpu_plan_config(const mifft_cpu_plan* p, mifft_config* out);
int32_t mifft_cpu_plan_packed_shape(const mifft_cpu_plan* p, uint64_t* out, int32_t cap);
/* CpuFftPlan::required_input_len / required_output_len (cpu.rs:227-243), in scalar words. */
uint64_t mifft_cpu_plan_required_input_len(const mifft_cpu_plan* p);
uint64_t mifft_cpu_plan_required_output_len(const mifft_cpu_plan* p);
/* CpuFftPlan::execute (cpu.rs:245-268): F32 and Df64 plans on f32 words;
 * lengths must equal the required lengths (MIFFT_ERR_HOST_BUFFER_LENGTH);
 * an F64 plan gives MIFFT_ERR_PRECISION_UNSUPPORTED. */
mifft_status mifft_cpu_execute(const mifft_cpu_plan* p, const float* in, uint64_t in_len, float* out,
                               uint64_t out_len);
/* CpuFftPlan::execute_f64 (cpu.rs:270-283): F64 plans on f64 words. */
mifft_status mifft_cpu_execute_f64(const mifft_cpu_plan* p, const double* in, uint64_t in_len, double* out,
                                   uint64_t out_len);
/* CpuFftPlan::execute_in_place / execute_in_place_f64 (cpu.rs:285-334): C2C
 * only (MIFFT_ERR_IN_PLACE_UNSUPPORTED for real plans). */
mifft_status mifft_cpu_execute_in_place(const mifft_cpu_plan* p, float* data, uint64_t len);
mifft_status mifft_cpu_execute_in_place_f64(const mifft_cpu_plan* p, double* data, uint64_t len);
/* Host threads the plans may use (0: every hardware thread, the reference's
 * available_parallelism, cpu.rs:829-832), the count in effect, and the
 * instruction set of the engine in use ("avx2+fma" or "x86-64"). */
void mifft_cpu_set_max_threads(int32_t n);
int32_t mifft_cpu_max_threads(void);
const char* mifft_cpu_isa(void);

/* --------------------------------------------------------------- errors */
const char* mifft_status_string(mifft_status s);
/* Thread-local detail text of the last failing call on this thread. */
const char* mifft_last_error_detail(void);

/* --------------------------------------------------------------- misc */
const char* mifft_version(void);
/* Synthetic inputs on the device: SplitMix64 uniform[-1,1) and N(0,1) f32
 * (element k uses SplitMix64 output k of `seed`). */
mifft_status mifft_fill_uniform_f32(float* dst, uint64_t count, uint64_t seed, hipStream_t stream);
mifft_status mifft_fill_normal_f32(float* dst, uint64_t count, uint64_t seed, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* MIFFT_H */
