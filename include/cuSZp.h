/*
 * cuSZp-compatible C ABI of the MI355X (gfx950) fused error-bounded compressor.
 *
 * Drop-in for the reference's public interface:
 *   - generic entries        include/cuSZp.h:38-45 (reference)   -> cuSZp_compress / cuSZp_decompress
 *   - per-dimension entries  include/cuSZp/cuSZp_entry_{1,2,3}D_f{32,64}.h:6-11 (reference)
 *   - enums                  include/cuSZp.h:20-36 (reference), identical values
 *
 * Every symbol below is exported with C linkage from libcuszp_amd.so. Device
 * pointers are plain HIP device pointers (hipMalloc / PyTorch-ROCm tensors);
 * `stream` is a hipStream_t (0 = null stream).
 *
 * Differences from the reference, all deliberate:
 *   - dims is passed as the POD `cuszp_dims_t` {x, y, z} (x fastest), which has
 *     the layout of CUDA's uint3 and of pycuSZp's ctypes Uint3;
 *   - *cmpSize is the true stream length (the reference under-reports it for
 *     single-threadblock grids and for 2D/3D block counts not divisible by 4);
 *   - decompression writes +0.0 into rate-0 blocks instead of leaving them
 *     untouched, i.e. it equals the reference's output into a zeroed buffer;
 *   - invalid arguments (bad enum, 2D/3D dims.x % 4 != 0, nbEle != x*y*z,
 *     errorBound <= 0 or not finite) print a message and return without
 *     launching; the *_ex entries return a status code instead.
 */
#ifndef CUSZP_AMD_CUSZP_H
#define CUSZP_AMD_CUSZP_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
#define CUSZP_DEFAULT(v) = v
extern "C" {
#else
#define CUSZP_DEFAULT(v)
#endif

/* include/cuSZp.h:20-36 (reference) */
typedef enum {
    CUSZP_DIM_1D = 1,
    CUSZP_DIM_2D = 2,
    CUSZP_DIM_3D = 3
} cuszp_dim_t;

typedef enum {
    CUSZP_MODE_FIXED = 0,
    CUSZP_MODE_PLAIN = 1,
    CUSZP_MODE_OUTLIER = 2
} cuszp_mode_t;

typedef enum {
    CUSZP_TYPE_FLOAT = 0,
    CUSZP_TYPE_DOUBLE = 1
} cuszp_type_t;

/* Layout-compatible with CUDA's uint3 and python/pycuSZp.py:8-9 (reference). */
typedef struct {
    unsigned int x, y, z;
} cuszp_dims_t;

typedef enum {
    CUSZP_SUCCESS = 0,
    CUSZP_ERROR_INVALID_ARGUMENT = 1,   /* bad enum, null pointer, bad error bound */
    CUSZP_ERROR_INVALID_DIMS = 2,       /* 2D/3D: x % 4 != 0, zero extent, nbEle != x*y*z */
    CUSZP_ERROR_TOO_LARGE = 3,          /* stream or block count beyond the format's limits */
    CUSZP_ERROR_OUTPUT_TOO_SMALL = 4,   /* compressed stream exceeds cmpCapacity */
    CUSZP_ERROR_CORRUPT_STREAM = 5,     /* cmpSize shorter than the header describes */
    CUSZP_ERROR_HIP = 6,                /* a HIP runtime call failed */
    CUSZP_ERROR_INTERNAL = 7,           /* device-side scan watchdog fired; output invalid */
    /* Not an error: the stream is written and *cmpSize is valid, but a 2D/3D
     * outlier-mode block holds a residual of magnitude 2^31 (saturated or
     * wrapped quantisation codes). Its rate byte 32 is read as rate 0 by every
     * outlier decoder, the reference's included (src/cuSZp_kernels_2D_f32.cu:2320),
     * so that block decompresses to zeros, beyond the error bound. */
    CUSZP_WARNING_BOUND_EXCEEDED = 8
} cuszp_status_t;

/* ---- generic entries: include/cuSZp.h:38-45 (reference) ---------------- */
void cuSZp_compress(void* d_oriData, unsigned char* d_cmpBytes,
                    size_t nbEle, size_t* cmpSize, float errorBound,
                    cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode,
                    hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress(void* d_decData, unsigned char* d_cmpBytes,
                      size_t nbEle, size_t cmpSize, float errorBound,
                      cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode,
                      hipStream_t stream CUSZP_DEFAULT(0));

/* ---- 1D entries: include/cuSZp/cuSZp_entry_1D_f32.h:6-11, _1D_f64.h:6-11 (reference) */
void cuSZp_compress_1D_fixed_f32(float* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_1D_fixed_f32(float* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_1D_plain_f32(float* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_1D_plain_f32(float* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_1D_outlier_f32(float* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_1D_outlier_f32(float* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_1D_fixed_f64(double* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_1D_fixed_f64(double* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_1D_plain_f64(double* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_1D_plain_f64(double* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_1D_outlier_f64(double* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_1D_outlier_f64(double* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));

/* ---- 2D entries: include/cuSZp/cuSZp_entry_2D_f32.h:6-11, _2D_f64.h:6-11 (reference) */
void cuSZp_compress_2D_fixed_f32(float* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_2D_fixed_f32(float* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_2D_plain_f32(float* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_2D_plain_f32(float* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_2D_outlier_f32(float* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_2D_outlier_f32(float* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_2D_fixed_f64(double* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_2D_fixed_f64(double* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_2D_plain_f64(double* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_2D_plain_f64(double* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_2D_outlier_f64(double* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_2D_outlier_f64(double* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));

/* ---- 3D entries: include/cuSZp/cuSZp_entry_3D_f32.h:6-11, _3D_f64.h:6-11 (reference) */
void cuSZp_compress_3D_fixed_f32(float* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_3D_fixed_f32(float* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_3D_plain_f32(float* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_3D_plain_f32(float* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_3D_outlier_f32(float* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_3D_outlier_f32(float* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, float errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_3D_fixed_f64(double* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_3D_fixed_f64(double* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_3D_plain_f64(double* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_3D_plain_f64(double* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_compress_3D_outlier_f64(double* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));
void cuSZp_decompress_3D_outlier_f64(double* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, cuszp_dims_t dims, double errorBound, hipStream_t stream CUSZP_DEFAULT(0));

/* ---- extensions (no reference counterpart; the reference returns void and never checks) ---- */

/* Upper bound of the compressed size: header + every block at 32 planes + 4
 * outlier bytes. The reference sizes its outputs at nbEle*sizeof(T)
 * (examples/cuSZp_test_f32_1D.cpp:61), which incompressible f32 data overflows. */
size_t cuSZp_max_compressed_size(size_t nbEle, cuszp_dim_t dim, cuszp_dims_t dims,
                                 cuszp_type_t type, cuszp_mode_t mode);

/* Status-returning form of cuSZp_compress: validates arguments, never writes
 * past d_cmpBytes + cmpCapacity, returns once *cmpSize is valid. errorBound is
 * used as double for f64 and rounded to float for f32. */
int cuSZp_compress_ex(const void* d_oriData, unsigned char* d_cmpBytes, size_t cmpCapacity,
                      size_t nbEle, size_t* cmpSize, double errorBound,
                      cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode,
                      hipStream_t stream);

/* Status-returning form of cuSZp_decompress: rejects streams whose header
 * describes more bytes than cmpSize (those blocks decode as zero). */
int cuSZp_decompress_ex(void* d_decData, const unsigned char* d_cmpBytes, size_t nbEle,
                        size_t cmpSize, double errorBound,
                        cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode,
                        hipStream_t stream);

/* Stream-ordered forms: enqueue and return without synchronising. The stream
 * length is written to the device word *d_cmpSize (compress) when the kernel
 * finishes; errors are reported by a later cuSZp_last_async_status(stream). */
int cuSZp_compress_async(const void* d_oriData, unsigned char* d_cmpBytes, size_t cmpCapacity,
                         size_t nbEle, unsigned long long* d_cmpSize, double errorBound,
                         cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode,
                         hipStream_t stream);
int cuSZp_decompress_async(void* d_decData, const unsigned char* d_cmpBytes, size_t nbEle,
                           size_t cmpSize, double errorBound,
                           cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode,
                           hipStream_t stream);
/* Synchronises `stream` and returns the error status accumulated by its async calls. */
int cuSZp_last_async_status(hipStream_t stream);

const char* cuSZp_status_string(int status);

/* ---- scan diagnostics (no reference counterpart) ----
 * The inter-tile scan is a single-pass decoupled lookback. A tile that waits
 * too long for a predecessor claims it and computes that tile's byte count
 * itself (forward progress without assuming that every workgroup is
 * resident); CUSZP_SCAN_INFO_STOLEN records that this happened. It is
 * information, not an error: the output is the same either way. */
#define CUSZP_SCAN_INFO_STOLEN 8u
typedef struct {
    unsigned int info;               /* CUSZP_SCAN_INFO_* bits seen since the last report */
    unsigned long long polls;        /* lookback polls that found their window unresolved (*) */
    unsigned long long waits;        /* lookbacks that polled at least once (*) */
    unsigned long long claims;       /* predecessors whose byte count a waiting tile computed (*) */
    unsigned long long wait_ticks;   /* 100 MHz clock ticks spent in those waits (*) */
} cuszp_scan_report_t;               /* (*) counted only while the stats switch is on */
/* What the single-array entries on `stream` (this device) reported since the
 * last call; clears it. Synchronises `stream`. */
int cuSZp_scan_report(hipStream_t stream, cuszp_scan_report_t* report);
/* Scan settings for launches made after the call (process-wide; a captured
 * batch graph keeps the settings of its capture):
 *   steal_ticks  wait before claiming a pending predecessor, in 100 MHz clock
 *                ticks (negative: the default, 10,000 = 100 us; 0: claim at
 *                the first unresolved poll);
 *   hold_every   test stress: tile t with t % hold_every == hold_every - 1
 *   hold_ticks     publishes only after waiting hold_ticks (0: off; at most
 *                  10,000,000 = 100 ms, else CUSZP_ERROR_INVALID_ARGUMENT), so
 *                  that its successors claim it;
 *   stats        nonzero: count polls, waits, claims and wait time (costs one
 *                device atomic per waiting tile). */
int cuSZp_scan_config(int steal_ticks, unsigned int hold_every, unsigned int hold_ticks, int stats);
/* Test fault injection: the next `count` growths of a scan-state status array
 * (single-array pool, batch plan, serving lane) fail as if the device were out
 * of memory; the call that hits one returns CUSZP_ERROR_HIP and the next call
 * allocates again. 0 turns it off. */
int cuSZp_debug_fail_allocs(unsigned int count);

/* ---- batched form (no reference counterpart) ----
 * Many arrays of any sizes, dtypes, modes, dims and error bounds compress (or
 * decompress) in ONE kernel launch with one workgroup per tile of the
 * concatenated arrays. Every array's stream is byte-identical to its
 * single-array stream. A plan validates the arrays once,
 * uploads their descriptors, owns the scan state and can hold the launch as a
 * captured hipGraph, so execute() is a single graph launch. */
typedef struct {
    const void* d_in;       /* compress: values; decompress: compressed stream */
    void* d_out;            /* compress: stream buffer; decompress: values */
    size_t nbEle;
    size_t capacity;        /* compress: bytes at d_out; decompress: stream length (cmpSize) */
    double errorBound;      /* absolute; f32 arrays use (float)errorBound, like cuSZp_compress_ex */
    cuszp_dim_t dim;
    cuszp_dims_t dims;      /* ignored for 1D */
    cuszp_type_t type;
    cuszp_mode_t mode;
} cuszp_batch_item_t;

typedef struct cuszp_batch_plan_s* cuszp_batch_plan_t;

enum { CUSZP_BATCH_COMPRESS = 0, CUSZP_BATCH_DECOMPRESS = 1 };
enum {
    CUSZP_BATCH_GRAPH = 1,  /* capture the launch in a hipGraph; execute() replays it */
    CUSZP_BATCH_TUNE = 2    /* autotune the launch shape when the plan is created (runs the batch) */
};

/* Validates every item (status of the first bad one is returned, its index in
 * *bad_index if non-null), allocates the plan's device state on the current
 * device and enqueues its set-up on `stream`. */
int cuSZp_batch_plan_create(cuszp_batch_plan_t* plan, const cuszp_batch_item_t* items, size_t count,
                            int direction, int flags, hipStream_t stream, size_t* bad_index);
/* Enqueues the batch on `stream` and returns (no host synchronisation). */
int cuSZp_batch_plan_execute(cuszp_batch_plan_t plan, hipStream_t stream);
/* Synchronises `stream`; copies each array's stream length (compress) and
 * status to the host arrays (either may be null); returns the first non-zero
 * status or CUSZP_SUCCESS. An array with CUSZP_WARNING_BOUND_EXCEEDED keeps
 * its stream length. */
int cuSZp_batch_plan_results(cuszp_batch_plan_t plan, size_t* cmpSizes, int* statuses, hipStream_t stream);
/* Device array of per-array stream lengths written by each execute (compress). */
const unsigned long long* cuSZp_batch_plan_device_sizes(cuszp_batch_plan_t plan);
/* Times the candidate launch shapes on `stream` (runs the batch several times):
 * workgroups of 4 or 8 waves of 64 lanes, and lookback windows of 128 or 256
 * predecessor tiles per round trip; keeps the fastest for this plan and for
 * later plans of the same size class. Returns its workgroup size (waves) in
 * *waves and the workgroups of its (first) launch, one per tile, in *grid. */
int cuSZp_batch_plan_tune(cuszp_batch_plan_t plan, hipStream_t stream, int* waves, unsigned* grid);
void cuSZp_batch_plan_destroy(cuszp_batch_plan_t plan);
/* cuSZp_scan_report for the plan's launches (synchronises `stream`). */
int cuSZp_batch_plan_scan_report(cuszp_batch_plan_t plan, hipStream_t stream, cuszp_scan_report_t* report);

/* One-shot batched calls: create a plan, execute, read results, destroy. */
int cuSZp_compress_batch(const cuszp_batch_item_t* items, size_t count, size_t* cmpSizes, int* statuses,
                         hipStream_t stream);
int cuSZp_decompress_batch(const cuszp_batch_item_t* items, size_t count, int* statuses, hipStream_t stream);

/* ---- serving form (no reference counterpart) ----
 * A request server with continuous batching. Requests are compress or
 * decompress operations on single arrays, submitted as they arrive; each
 * result is identical to the single-array call's. Outputs live in a device
 * pool the caller provides (sub-allocated in 256-byte units); admission is
 * first come, first served, by output bytes: the oldest waiting request starts
 * as soon as its output fits in the pool. Admitted requests run in batches on
 * up to `lanes` streams at once (one batched launch per direction per batch),
 * so requests that arrive while a batch runs start on another lane without
 * waiting for it. A completed request keeps its output in the pool until
 * cuSZp_server_release. All calls on one server are serialised internally. */
typedef struct cuszp_server_s* cuszp_server_t;
typedef struct {
    unsigned long long id;  /* as returned by cuSZp_server_submit */
    int status;             /* CUSZP_SUCCESS or the request's error */
    int direction;          /* CUSZP_BATCH_COMPRESS / CUSZP_BATCH_DECOMPRESS */
    size_t offset;          /* byte offset of the output in the pool */
    void* d_out;            /* pool + offset */
    size_t bytes;           /* compress: stream length; decompress: nbEle * sizeof(value) */
} cuszp_completion_t;
/* lanes: 1..4 concurrent batches; max_batch: requests per batch (1..2^20). */
int cuSZp_server_create(cuszp_server_t* server, void* d_pool, size_t pool_bytes, int lanes, size_t max_batch);
/* item->d_out is ignored (the output is allocated in the pool); for
 * decompression item->capacity is the stream length. Validation errors are
 * returned here (no id is issued); requests whose output exceeds the pool
 * return CUSZP_ERROR_TOO_LARGE. Launches what it can before returning. */
int cuSZp_server_submit(cuszp_server_t server, int direction, const cuszp_batch_item_t* item,
                        unsigned long long* id);
/* Retires finished batches, admits and launches waiting requests, and copies
 * up to max_done completions (oldest first) to done[]. With wait != 0 it
 * blocks until at least one completion is available or nothing is running. */
int cuSZp_server_poll(cuszp_server_t server, cuszp_completion_t* done, size_t max_done, size_t* n_done, int wait);
/* Returns a completed request's output bytes to the pool. */
int cuSZp_server_release(cuszp_server_t server, unsigned long long id);
/* Requests waiting for pool space, admitted or running, and pool bytes held. */
int cuSZp_server_stats(cuszp_server_t server, size_t* queued, size_t* in_flight, size_t* pool_used);
/* Waits for running batches and frees the server's own resources (not the pool). */
void cuSZp_server_destroy(cuszp_server_t server);

/* Load replay for measuring a server: request i is submitted at arrival_s
 * seconds after the call starts (arrival times non-decreasing), every
 * completion is released at once, and latency_s[i] = completion seen -
 * arrival. sequential != 0 instead runs the requests one at a time in arrival
 * order through cuSZp_compress_ex / cuSZp_decompress_ex (output at the start
 * of the pool). The server must hold no other request (CUSZP_ERROR_INVALID_ARGUMENT
 * otherwise). *makespan_s = time until the last completion. */
typedef struct {
    int direction;
    cuszp_batch_item_t item;
    double arrival_s;
} cuszp_request_t;
int cuSZp_server_replay(cuszp_server_t server, const cuszp_request_t* requests, size_t count, int sequential,
                        double* latency_s, int* statuses, double* makespan_s);

/* ---- relative error bounds (the step before the path) ----
 * The reference CLI turns a relative bound into an absolute one on the host
 * (examples/cuSZp.cpp:176-210 of the reference); these compute the same on the
 * device array. cuSZp_value_range: min and max of the elements, NaNs skipped
 * unless element 0 is NaN (then both are NaN), as the CLI's comparison loop
 * gives. cuSZp_rel_error_bound: *errorBound = rel * (max - min) in float
 * arithmetic (f64 arrays: the difference in double, rounded to float). Both
 * synchronise `stream`. */
int cuSZp_value_range(const void* d_data, size_t nbEle, cuszp_type_t type, double* min_val, double* max_val,
                      hipStream_t stream);
int cuSZp_rel_error_bound(const void* d_data, size_t nbEle, cuszp_type_t type, float rel, float* errorBound,
                          hipStream_t stream);

/* Reconstruction quality of d_decData against d_oriData (nbEle values of
 * `type`), the figures of the reference's computePSNR
 * (src/cuSZp_utility.cu:891-950 of the reference) computed on the device:
 * metrics[0] mse, [1] value range of the original, [2] PSNR (dB), [3] L2 norm
 * of the error, [4] that norm over the L2 norm of d_decData, [5] NRMSE,
 * [6] max |error|, [7] max |error| / |original| over nonzero originals,
 * [8] correlation coefficient. f32 data uses the reference's float per-element
 * arithmetic with double sums. Synchronises `stream`. */
int cuSZp_quality(const void* d_oriData, const void* d_decData, size_t nbEle, cuszp_type_t type, double* metrics,
                  hipStream_t stream);
/* Mean SSIM of d_decData against d_oriData over 7x7x7 windows moved by 2 along
 * each axis, as the reference's computeSSIM (src/cuSZp_utility.cu:761-889 of
 * the reference) on a dims.x * dims.y * dims.z volume (x fastest). Every
 * extent must be at least 7 (CUSZP_ERROR_INVALID_DIMS). Synchronises `stream`. */
int cuSZp_ssim(const void* d_oriData, const void* d_decData, cuszp_dims_t dims, cuszp_type_t type, double* ssim,
               hipStream_t stream);

/* ---- self-describing container (no reference counterpart) ----
 * The reference writes header-less raw streams (examples/cuSZp.cpp:306-317 of
 * the reference): reading one back needs the type, mode, dims, bound and
 * length from elsewhere. A container is a 64-byte little-endian header
 * followed by the unchanged raw stream:
 *   [0,8)   magic 89 43 53 5A 50 0D 0A 1A ("\x89CSZP\r\n\x1a")
 *   [8,10)  format version (1)        [10,12) header bytes (64)
 *   [12]    dim (1..3)   [13] type (0 f32, 1 f64)   [14] mode (0..2)
 *   [15]    flags: bit 0 = the bound was given relative to the value range
 *   [16,28) dims x, y, z (u32 each; 0 for 1D)
 *   [28,32) CRC-32 (IEEE 802.3, as zlib's crc32) of the 64 header bytes with
 *           this field zero
 *   [32,40) nbEle (u64)
 *   [40,48) the absolute bound the stream was written with (f64; for f32
 *           streams the float value)
 *   [48,56) the relative bound it came from (f64; 0 when absolute)
 *   [56,64) stream bytes (u64)
 * The raw stream starts at byte 64 and is exactly what cuSZp_compress_ex
 * writes, so any raw-stream reader can decompress it from there. */
#define CUSZP_CONTAINER_HEADER_BYTES 64
#define CUSZP_CONTAINER_VERSION 1
typedef struct {
    cuszp_dim_t dim;
    cuszp_type_t type;
    cuszp_mode_t mode;
    int relative;                      /* 1: errorBound came from relBound */
    cuszp_dims_t dims;                 /* {0,0,0} for 1D */
    unsigned long long nbEle;
    double errorBound;                 /* absolute bound of the stream */
    double relBound;                   /* 0 unless relative */
    unsigned long long streamBytes;    /* raw stream length after the header */
} cuszp_container_info_t;
/* Encodes `info` into the 64 header bytes at `header` (host memory). Rejects
 * bad enums, dims the compressor would reject and non-positive bounds
 * (CUSZP_ERROR_INVALID_ARGUMENT / CUSZP_ERROR_INVALID_DIMS). */
int cuSZp_container_write_header(const cuszp_container_info_t* info, unsigned char* header);
/* Decodes and validates a header from `len` host bytes (len >= 64): magic,
 * version, CRC, enums, geometry and bound. Returns CUSZP_ERROR_CORRUPT_STREAM
 * for anything that is not a valid version-1 header. */
int cuSZp_container_read_header(const unsigned char* header, size_t len, cuszp_container_info_t* info);
/* Compresses into a container at d_container (device memory, capacity bytes:
 * 64 + cuSZp_max_compressed_size always suffices). relBound is recorded in
 * the header when > 0 (errorBound must then be the absolute bound derived
 * from it, e.g. by cuSZp_rel_error_bound). *containerSize = 64 + stream
 * bytes. Blocking, like cuSZp_compress_ex. */
int cuSZp_compress_container(const void* d_oriData, unsigned char* d_container, size_t capacity, size_t nbEle,
                             size_t* containerSize, double errorBound, double relBound, cuszp_dim_t dim,
                             cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode, hipStream_t stream);
/* Decompresses the container of containerSize bytes at d_container (device
 * memory) into d_decData, which holds at most decCapacity values of the
 * container's type (CUSZP_ERROR_OUTPUT_TOO_SMALL if the container describes
 * more). The decoded header goes to *info when info is not null. Blocking. */
int cuSZp_decompress_container(void* d_decData, size_t decCapacity, const unsigned char* d_container,
                               size_t containerSize, cuszp_container_info_t* info, hipStream_t stream);

/* ---- compression that allocates nothing (no counterpart in the reference) ----
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops and harnesses that forbid hipMalloc. Same stream as
 * cuSZp_compress_ex, but the scan state lives in a fixed device array of the
 * library (about 514 KiB per GPU, held until the process ends) and the call
 * only launches kernels: no hipMalloc/hipFree, no memset, no copy, no
 * synchronisation. When the launches have finished, d_result[0] holds the
 * stream length and d_result[1] the call's status (CUSZP_SUCCESS,
 * CUSZP_ERROR_OUTPUT_TOO_SMALL, CUSZP_ERROR_INTERNAL or
 * CUSZP_WARNING_BOUND_EXCEEDED); d_result is two device words of the caller.
 * At most 2^16 tiles: 2^31 f32 or 2^30 f64 values; larger arrays return
 * CUSZP_ERROR_TOO_LARGE before anything is launched. Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other streams or
 * host threads) race, so the caller must serialise them. */
int cuSZp_compress_noalloc(const void* d_oriData, unsigned char* d_cmpBytes, size_t cmpCapacity,
                           size_t nbEle, unsigned long long* d_result, double errorBound,
                           cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode,
                           hipStream_t stream);

/* Library version string ("cuszp_amd <semver> gfx950"). */
const char* cuSZp_version(void);

#ifdef __cplusplus
}  /* extern "C" */

/* C++ convenience: accept HIP's uint3/dim3 where the reference takes uint3. */
#include <hip/hip_vector_types.h>
inline cuszp_dims_t cuszp_dims(const uint3& d) { cuszp_dims_t r = {d.x, d.y, d.z}; return r; }
inline void cuSZp_compress(void* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize,
                           float errorBound, cuszp_dim_t dim, uint3 dims, cuszp_type_t type,
                           cuszp_mode_t mode, hipStream_t stream = 0) {
    cuSZp_compress(d_oriData, d_cmpBytes, nbEle, cmpSize, errorBound, dim, cuszp_dims(dims), type, mode, stream);
}
inline void cuSZp_decompress(void* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize,
                             float errorBound, cuszp_dim_t dim, uint3 dims, cuszp_type_t type,
                             cuszp_mode_t mode, hipStream_t stream = 0) {
    cuSZp_decompress(d_decData, d_cmpBytes, nbEle, cmpSize, errorBound, dim, cuszp_dims(dims), type, mode, stream);
}
#endif

#undef CUSZP_DEFAULT
#endif /* CUSZP_AMD_CUSZP_H */
