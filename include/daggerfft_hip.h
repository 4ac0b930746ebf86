/*
 * daggerfft_hip.h -- C ABI of libdaggerfft_hip, the MI355X (gfx950) 3D FFT
 * library behind the DaggerFFT.jl transform path.
 *
 * Every entry point returns an int status (DFFT_OK == 0); nothing throws or
 * aborts across this boundary.  All device work is enqueued on the caller's
 * HIP stream (passed as void*; NULL = the default stream) and nothing
 * synchronises the device implicitly except where stated.  Plans may be
 * created and destroyed from several threads at once; one plan may execute
 * concurrently on distinct streams with distinct workspaces.
 *
 * Layout: arrays are column-major (Julia order).  An array of size
 * n[0] x n[1] x n[2] holds element (i,j,k) at i + n[0]*(j + n[1]*k); complex
 * elements are interleaved (re, im).  Axes are 0-based here (Julia's dims are
 * 1-based; the Julia shim subtracts 1).
 *
 * What each entry point replaces in the reference (Julia, sanatgp/DaggerFFT.jl):
 *   dfft_plan_create     plan_transform / gpu_plan_transform + the plan caches
 *                        get_or_create_plan / get_or_create_gpu_plan
 *                        (src/fft.jl:126-166, src/fftgpu.jl:83-116), and the
 *                        per-stage structure of fft!/ifft!/gpu_fft!/gpu_ifft!
 *                        for Pencil and Slab (src/fft.jl:650-776,
 *                        src/fftgpu.jl:497-619)
 *   dfft_execute         the three-stage driver bodies: apply_fft! /
 *                        apply_gpu_fft! (`mul!(out, plan, in)`, src/fft.jl:168-174,
 *                        src/fftgpu.jl:118-124) per stage plus the on-device
 *                        LOCAL_COPIES redistribution (src/fftgpu.jl:367-385)
 *   dfft_plan_destroy    cleanup_workspace! / gpu_cleanup_workspace!
 *                        (src/fft.jl:778-781, src/fftgpu.jl:630-634)
 *   dfft_box_*           pack_kernel!, unpack_kernel!, local_copy_kernel!
 *                        (src/fftgpu.jl:452-495) and pack_data!/unpack_data!/
 *                        local_copy! (src/fft.jl:180-231)
 *   dfft_batch_*         no reference counterpart: many independent transforms
 *                        per launch (one launch per axis-pass level)
 *   dfft_server_*        no reference counterpart: continuous-batching request
 *                        loop with resident plans and a device memory arena
 */
#ifndef DAGGERFFT_HIP_H
#define DAGGERFFT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define DFFT_OK                    0
#define DFFT_ERR_ARG               1  /* bad argument (ArgumentError in the reference) */
#define DFFT_ERR_UNSUPPORTED_SIZE  2  /* a line too long for one LDS tile: n > 8192 except 16384; with a prime factor > 7, n > 8192 (C2C) or n > 4096 (other ops) */
#define DFFT_ERR_HIP               3  /* a HIP runtime call failed */
#define DFFT_ERR_OOM               4  /* device allocation failed */
#define DFFT_ERR_BUDGET            5  /* serving admission: request exceeds the HBM budget */
#define DFFT_ERR_STATE             6  /* object used in the wrong state */
#define DFFT_ERR_NOT_READY         7  /* serving: request not finished yet */
#define DFFT_ERR_UNSUPPORTED       8  /* transform kind not implemented on the GPU path */
#define DFFT_ERR_SCRATCH           9  /* dfft_execute_static: the plan's workspace exceeds DFFT_STATIC_SCRATCH_BYTES */
#define DFFT_ERR_WISDOM           10  /* a wisdom file stamped for another HIP build or GPU architecture */

/* ---- transform kinds (FFT / RFFT / IRFFT tags, src/fft.jl:17-24; R2R README.md:59-66) */
#define DFFT_C2C  0
#define DFFT_R2C  1
#define DFFT_C2R  2
#define DFFT_R2R  3

/* ---- r2r kinds: FFTW's fftw_r2r_kind numbering (R2R_SUPPORTED_KINDS, src/fft.jl:1-11) */
#define DFFT_R2HC     0
#define DFFT_HC2R     1
#define DFFT_DHT      2
#define DFFT_REDFT00  3
#define DFFT_REDFT01  4
#define DFFT_REDFT10  5
#define DFFT_REDFT11  6
#define DFFT_RODFT00  7
#define DFFT_RODFT01  8
#define DFFT_RODFT10  9
#define DFFT_RODFT11 10

/* ---- precision --------------------------------------------------------- */
#define DFFT_F32  0   /* Float32 / ComplexF32 data, twiddles generated in FP64+ and rounded once */
#define DFFT_F64  1   /* Float64 / ComplexF64 */

/* ---- decomposition (Pencil / Slab, src/fft.jl:13-15) ------------------- */
#define DFFT_PENCIL 0
#define DFFT_SLAB   1

/* ---- plan flags -------------------------------------------------------- */
#define DFFT_FLAG_UNNORMALIZED   0x1u  /* sign=+1 C2C/C2R without the 1/n scaling (FFTW bfft) */
#define DFFT_FLAG_NORMALIZE_R2R  0x2u  /* scale R2R by 1/prod(logical size) (README ifft with REDFT01) */
#define DFFT_FLAG_DESTROY_INPUT  0x4u  /* C2R may overwrite its input (no workspace needed) */
#define DFFT_FLAG_NO_TUNE        0x8u  /* skip the autotuner cache; use the heuristic schedule */
#define DFFT_FLAG_GRAPH          0x10u /* replay the pass list as a hipGraph */
#define DFFT_FLAG_STAGE_ORDER    0x20u /* C2C: run the axis passes in the reference's stage order
                                          (dims[0..] forward, reversed inverse); by default the
                                          contiguous axis runs first -- same result, the passes commute */
#define DFFT_FLAG_PLANE_PIPELINE 0x40u /* Slab: software-pipeline the 2D stage over z-plane chunks
                                          (grouped launches re-reading each chunk from the Infinity Cache) */
#define DFFT_FLAG_BATCH_GROUPED  0x80u /* batch: grouped launches per axis-pass level (three to six) */
#define DFFT_FLAG_BATCH_LANES    0x100u /* batch: each transform's own plan on four forked streams */
#define DFFT_FLAG_MEASURE        0x200u /* time the schedules of passes the cache lacks at plan
                                         * creation (FFTW.MEASURE, src/fft.jl:127), and a batch's
                                         * launch schedule; synchronises the device and saves the
                                         * cache if dfft_tune_autosave is set */
#define DFFT_FLAG_DCT_FULL_LENGTH 0x400u /* REDFT10/REDFT01 always through the n-point complex FFT */
#define DFFT_FLAG_DCT_HALF_LENGTH 0x800u /* REDFT10/REDFT01 of n = 4k through the n/2-point complex
                                          * FFT of packed pairs wherever a kernel exists (n = 16..1024);
                                          * by default only from n = 256 (REDFT10) / 512 (REDFT01) in
                                          * arrays of 4M elements or more, where it measured faster.
                                          * With both flags: full length */
#define DFFT_FLAG_BATCH_PERSISTENT 0x1000u /* batch: the grouped schedule as one cooperative launch,
                                            * a grid barrier between axis-pass levels */

typedef struct dfft_plan dfft_plan;

/*
 * Create a plan.  Replaces plan_transform (src/fft.jl:126-144) /
 * gpu_plan_transform (src/fftgpu.jl:85-105) and the Pencil/Slab stage lists
 * of fft!/ifft! (src/fft.jl:651-776).
 *   rank      1..3
 *   n         logical (real-space) size per axis, n[0] fastest
 *   kind      DFFT_C2C / R2C / C2R / R2R
 *   sign      -1 forward, +1 inverse (C2C: ifft, scaled 1/n per dim unless
 *             DFFT_FLAG_UNNORMALIZED).  R2C requires -1, C2R requires +1.
 *             For R2R the sign only sets the stage order.
 *   r2r_kinds one FFTW kind per entry of dims (R2R only; else may be NULL)
 *   precision DFFT_F32 / DFFT_F64
 *   decomp    DFFT_PENCIL / DFFT_SLAB
 *   dims      0-based axes to transform, in stage order (ndims 1..3, distinct).
 *             R2C/C2R halve dims[0] to n/2+1 (FFTW rfft convention).
 */
int dfft_plan_create(dfft_plan **plan, int rank, const int64_t *n, int kind, int sign,
                     const int *r2r_kinds, int precision, int decomp,
                     const int *dims, int ndims, uint32_t flags);

/* Bytes of device workspace dfft_execute needs when d_work != NULL (0 if none). */
int dfft_plan_workspace_size(const dfft_plan *plan, size_t *bytes);

/* Bytes of the input and output arrays for this plan. */
int dfft_plan_io_size(const dfft_plan *plan, size_t *in_bytes, size_t *out_bytes);

/*
 * Execute on `stream`.  Replaces `mul!(out, plan, in)` per stage plus the
 * single-device redistribution (src/fft.jl:170, src/fftgpu.jl:120,367-385).
 * d_in == d_out runs in place (C2C and R2R only).  d_work may be NULL: the
 * plan then uses (and owns) its own workspace.  Concurrent executions of one
 * plan need distinct d_work buffers.
 */
int dfft_execute(const dfft_plan *plan, const void *d_in, void *d_out, void *d_work, void *stream);

/*
 * dfft_execute without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * caller passes the input and output arrays; a plan that needs a workspace
 * (C2R without DFFT_FLAG_DESTROY_INPUT) uses the library's static scratch
 * buffer of DFFT_STATIC_SCRATCH_BYTES (256 MiB) on the current device, which
 * each device holds from its first such call until the process ends.
 *   - DFFT_ERR_SCRATCH if the plan's workspace is larger (nothing is launched;
 *     dfft_last_error gives both sizes);
 *   - DFFT_ERR_ARG if the current device is not the plan's device;
 *   - every call on one device shares that device's scratch buffer and it
 *     keeps the last call's contents (the plan's first pass overwrites the
 *     part it uses before any pass reads it), so calls that overlap on one
 *     device -- on different streams, or from different threads -- race;
 *   - plans created with DFFT_FLAG_GRAPH run their passes as plain launches
 *     here (instantiating the graph would allocate).
 */
#define DFFT_STATIC_SCRATCH_BYTES ((size_t)256 << 20)
int dfft_execute_static(const dfft_plan *plan, const void *d_in, void *d_out, void *stream);

int dfft_plan_destroy(dfft_plan *plan);

/* Execute one axis pass of the plan (0 <= pass < num_passes), so a caller can
 * time or profile each axis pass on its own.  Running passes 0..n-1 in order
 * on one stream is exactly dfft_execute.  No reference counterpart (the
 * reference times whole stages with NVTX ranges, src/fft.jl:654-686). */
int dfft_execute_pass(const dfft_plan *plan, int pass, const void *d_in, void *d_out, void *d_work,
                      void *stream);

/* Number of axis passes and their names/geometry, for profiling and tests.
 * dfft_plan_describe writes one JSON object per pass: axis, op, L (line
 * length), radices, TPL/TC (threads per line, lines or columns per tile),
 * variant (0 runtime schedule, >= 1 compile-time), launch shape, LDS bytes,
 * algorithmic bytes, swz (tile order), bluestein / rader (the convolution
 * length of a Bluestein or Rader pass, 0 otherwise), paired_loads (1: the
 * first stage's loads of a ComplexF64 C2C pass along a power-of-two stride of
 * 8 KiB or more go out a pair at a time, 0: all at once) and the kernel's
 * name as profilers print it. */
int dfft_plan_num_passes(const dfft_plan *plan, int *npasses);
int dfft_plan_describe(const dfft_plan *plan, int pass, char *buf, size_t buflen);

/* How dfft_execute runs the plan: one launch per axis pass, or (Slab plans with
 * DFFT_FLAG_PLANE_PIPELINE of rank-3 C2C/R2C/C2R arrays with dims[2] == 2) a software pipeline over chunks
 * of z-planes -- launch i runs the dims[0] pass of chunk i together with the
 * dims[1] pass of chunk i-1, which is re-read while still in the Infinity
 * Cache -- followed by the dims[2] pass.  `launches`: kernel launches per
 * execution. */
#define DFFT_SCHEDULE_PASSES          0
#define DFFT_SCHEDULE_PLANE_PIPELINE  1
int dfft_plan_schedule(const dfft_plan *plan, int *schedule, int *launches);

/* Algorithmic HBM bytes of the plan: sum over passes of bytes read + written. */
int dfft_plan_algorithmic_bytes(const dfft_plan *plan, double *bytes);

const char *dfft_strerror(int status);
/* The calling thread's last DFFT_ERR_HIP cause (the failing HIP call and HIP's
 * message) or DFFT_ERR_SCRATCH cause (the two sizes). */
const char *dfft_last_error(void);
int dfft_version(int *major, int *minor);

/*
 * Box maps of the redistribution, on device.  `esize` is the element size in
 * bytes (4, 8 or 16).  Offsets, lo and cnt are 0-based element counts.
 *   dfft_box_pack:   buf[offset + i + c0*(j + c1*k)] = src[lo + (i,j,k)]
 *                    (pack_kernel!, src/fftgpu.jl:452-465; pack_data!, src/fft.jl:180-194)
 *   dfft_box_unpack: dst[lo + (i,j,k)] = buf[offset + i + c0*(j + c1*k)]
 *                    (unpack_kernel!, src/fftgpu.jl:467-480; unpack_data!, src/fft.jl:196-214)
 *   dfft_box_copy:   dst[dlo + (i,j,k)] = src[slo + (i,j,k)]
 *                    (local_copy_kernel!, src/fftgpu.jl:482-495; local_copy!, src/fft.jl:216-231)
 *   dfft_permute:    out = in with axes permuted: out axis d is in axis perm[d]
 *                    (the pencil transpose as an LDS-tiled axis transpose)
 */
int dfft_box_pack(void *d_buf, const void *d_src, const int64_t *src_shape, const int64_t *lo,
                  const int64_t *cnt, int64_t offset, int esize, void *stream);
int dfft_box_unpack(void *d_dst, const void *d_buf, const int64_t *dst_shape, const int64_t *lo,
                    const int64_t *cnt, int64_t offset, int esize, void *stream);
int dfft_box_copy(void *d_dst, const int64_t *dst_shape, const int64_t *dlo, const void *d_src,
                  const int64_t *src_shape, const int64_t *slo, const int64_t *cnt, int esize,
                  void *stream);
int dfft_permute(void *d_out, const void *d_in, const int64_t *in_shape, const int *perm, int esize,
                 void *stream);

/* Plain device-to-device copy of `bytes` (a multiple of 16; both pointers
 * 16-byte aligned; ranges must not overlap): 16 B per lane with streaming
 * loads and stores.  No reference counterpart: the benchmark measures the HBM
 * copy ceiling the axis passes are compared with. */
int dfft_device_copy(void *d_dst, const void *d_src, size_t bytes, void *stream);

/* Seeded SplitMix64 uniform[0,1) fill on device (the benchmark input generator):
 * real element e <- u(seed, e); complex element e <- (u(seed, 2e), u(seed, 2e+1)). */
int dfft_fill_uniform(void *d_out, int64_t nelem, int is_complex, int precision, uint64_t seed,
                      void *stream);

/* ---- autotuner (the FFTW wisdom analogue: PLAN_CACHE, import_wisdom /
 * export_wisdom, src/fft.jl:113-124,162) ---------------------------------
 * dfft_autotune times candidate schedules (radix order, threads per line,
 * lines/columns per workgroup, LDS padding) of every axis pass of the plan the
 * same arguments describe, on the current device, and records the fastest in
 * a process-wide cache that later dfft_plan_create / dfft_batch_create calls
 * use (unless DFFT_FLAG_NO_TUNE).  level: 0 off, 1 quick, 2 full.  It
 * allocates scratch buffers of the plan's size and synchronises `stream`. */
int dfft_autotune(int rank, const int64_t *n, int kind, int sign, const int *r2r_kinds, int precision,
                  int decomp, const int *dims, int ndims, uint32_t flags, int level, void *stream);
/* Wisdom files start with a stamp naming the HIP version the library was
 * compiled with and the GPU architecture (hip=... arch=...).  dfft_tune_load
 * refuses a stamped file of another build or architecture with DFFT_ERR_WISDOM (dfft_last_error says
 * which); unstamped hand-written files load as schedule overrides.  The
 * library imports <library dir>/wisdom/<arch>.txt (or $DFFT_WISDOM) at its
 * first plan creation, as the reference imports its wisdom at module init
 * (src/fft.jl:116-120). */
int dfft_tune_save(const char *path);
int dfft_tune_load(const char *path);
int dfft_tune_clear(void);
int dfft_tune_entries(int *count);
/* Save the whole cache to `path` whenever tuning adds to it (dfft_autotune,
 * DFFT_FLAG_MEASURE plans): the reference's save_wisdom() after every new plan
 * (src/fft.jl:122-124,162).  NULL or "" turns it off; $DFFT_WISDOM_SAVE sets it
 * at the first plan creation. */
int dfft_tune_autosave(const char *path);
/* Plan serialisation: write the schedules of this plan's passes (the tuned
 * ones or the built-in choice) as a stamped wisdom file.  A process that
 * loads it and creates a plan with the same arguments gets the same kernels,
 * radix lists, tile shapes and launch geometry. */
int dfft_plan_export_wisdom(const dfft_plan *plan, const char *path);

/* ---- batched launch ------------------------------------------------------
 * No reference counterpart: many independent transforms per kernel launch
 * (the reference runs one cuFFT plan per chunk, src/fftgpu.jl:118-124).
 * Each transform is rank 3 over dims (0,1,2); C2C (either sign) or R2C
 * (forward, even n[0]).  Three schedules:
 *   DFFT_FLAG_BATCH_GROUPED  level l (axis pass l of every transform) runs as
 *                            one launch, or as two -- ComplexF64, then
 *                            ComplexF32, each kernel holding one precision's
 *                            schedules -- when all its passes run
 *                            compile-time schedules; three to six launches
 *                            per batch, and no workgroup waits for another.
 *                            C2C passes of the lengths 16, 32, 48, 64, 80,
 *                            96, 112, 128, 144, 160, 192, 224, 240, 256 run
 *                            compile-time schedules in 256-thread tiles;
 *                            other passes the runtime engine.  Results equal
 *                            each transform alone up to rounding;
 *   DFFT_FLAG_BATCH_LANES    each transform's own plan (the compile-time
 *                            kernels) on one of four streams forked from and
 *                            joined back into the caller's stream; results
 *                            equal each transform run alone bit for bit;
 *   DFFT_FLAG_BATCH_PERSISTENT  the grouped schedule as ONE cooperative
 *                            launch (hipLaunchCooperativeKernel): as many
 *                            256-thread workgroups as the device holds at
 *                            once loop over each level's tiles and meet at a
 *                            grid barrier between levels.  Same tiles and
 *                            results as DFFT_FLAG_BATCH_GROUPED, bit for bit.
 *                            Every level must fit 256-thread tiles
 *                            (DFFT_ERR_UNSUPPORTED_SIZE otherwise, e.g. a
 *                            4096-point contiguous axis);
 *                            DFFT_ERR_UNSUPPORTED on a device without
 *                            cooperative launches.  With DFFT_FLAG_BATCH_LANES
 *                            or DFFT_FLAG_GRAPH: DFFT_ERR_ARG.  Opt-in only:
 *                            measured 25% slower than the grouped launches,
 *                            and after a process's first cooperative launch
 *                            work spread over several streams runs slower
 *                            for the rest of the process (measured: the
 *                            lanes schedule 2.17 -> 4.03 ms; it concerns the
 *                            serving loop too);
 * No schedule flag: the schedule cache's pick for this mix of descriptors;
 * with DFFT_FLAG_MEASURE and no pick cached, the grouped and lanes schedules
 * are timed on scratch arrays of the batch's sizes (the lanes' plans with
 * their passes autotuned) and the faster is cached (and saved with the cache
 * when a save path is set).  With no pick: grouped when every axis length of every
 * transform is one of the compile-time grouped lengths and every transform
 * is C2C, else lanes.  DFFT_FLAG_BATCH_GROUPED with DFFT_FLAG_BATCH_LANES:
 * DFFT_ERR_ARG.
 * With DFFT_FLAG_GRAPH at creation the launches are recorded once as a
 * hipGraph (the lanes as parallel branches) and replayed by every execute,
 * re-recorded when the buffer pointers change. */
typedef struct {
    int64_t n[3];       /* size of this 3D transform (n[0] fastest) */
    int kind;           /* DFFT_C2C or DFFT_R2C */
    int sign;           /* -1 / +1 (C2C); -1 (R2C) */
    int precision;      /* DFFT_F32 / DFFT_F64 */
    uint32_t flags;     /* DFFT_FLAG_UNNORMALIZED */
} dfft_batch_desc;

typedef struct dfft_batch dfft_batch;

int dfft_batch_create(dfft_batch **batch, int count, const dfft_batch_desc *descs, uint32_t flags);
/* d_in[t], d_out[t]: device pointers of transform t (in place when equal, C2C only). */
int dfft_batch_execute(dfft_batch *batch, void *const *d_in, void *const *d_out, void *stream);
int dfft_batch_workspace_size(const dfft_batch *batch, size_t *bytes);
/* workgroups (tiles) over all launches of one execution, and the launch count */
int dfft_batch_num_items(const dfft_batch *batch, int64_t *items);
int dfft_batch_num_launches(const dfft_batch *batch, int *launches);
int dfft_batch_algorithmic_bytes(const dfft_batch *batch, double *bytes);
int dfft_batch_destroy(dfft_batch *batch);

/* ---- serving loop ----------------------------------------------------------
 * No reference counterpart as a whole.  It serves a stream of independent 3D
 * transform requests with what the reference rebuilds per call kept
 * resident: the plan cache (GPU_PLAN_CACHE / get_or_create_gpu_plan, src/fftgpu.jl:83,107-116)
 * and the device staging buffers of a workspace (create_gpu_workspace,
 * src/fftgpu.jl:227-249, released by gpu_cleanup_workspace!,
 * src/fftgpu.jl:630-634).
 *
 * Clients submit requests from any thread; submission never blocks on the
 * device.  A scheduler thread admits queued requests in arrival order as
 * in-flight slots and memory under the HBM budget free up (a later request
 * that fits may overtake one waiting for memory; the oldest waiting request
 * keeps its memory reserved, so it is never starved), launches each on the
 * least-loaded of `lanes` HIP streams -- work of requests that arrive while
 * others run overlaps on the device -- and retires it when its completion
 * event fires.  Plans are built once per (shape, kind, sign, precision,
 * decomp, flags), at the first submission of that shape, and kept until
 * dfft_server_destroy.  Device memory is an arena that grows in chunks up to
 * the budget and is carved into request buffers; nothing is allocated or
 * freed on the device while requests run.  A request's result is the same
 * as dfft_execute of the same plan on the same input (the same kernels run). */
typedef struct dfft_server dfft_server;

typedef struct {
    size_t hbm_budget;      /* device bytes the server may hold: arena + plans (0: 90% of free memory) */
    int max_inflight;       /* requests admitted and not yet retired (0: 32) */
    int lanes;              /* HIP streams the requests are spread over (0: 4) */
    uint32_t flags;         /* reserved, must be 0 */
    int reserved;
} dfft_server_config;

typedef struct {
    int64_t n[3];           /* 3D size, n[0] fastest; transforms run over dims (0,1,2) */
    int kind;               /* DFFT_C2C, DFFT_R2C or DFFT_C2R */
    int sign;               /* -1 / +1 for C2C; -1 for R2C; +1 for C2R */
    int precision;          /* DFFT_F32 / DFFT_F64 */
    int decomp;             /* DFFT_PENCIL / DFFT_SLAB */
    uint32_t flags;         /* plan flags (DFFT_FLAG_UNNORMALIZED, DFFT_FLAG_NO_TUNE, ...) */
    uint32_t req_flags;     /* 0 or DFFT_REQ_DISCARD */
    uint64_t seed;          /* d_in == NULL: the input is dfft_fill_uniform(seed) of the input shape */
    const void *d_in;       /* the caller's device input, read and left untouched; or NULL */
    void *d_out;            /* the caller's device output; or NULL: the result stays in a server
                               buffer until dfft_server_release */
    double t_arrival;       /* arrival time on the server clock (dfft_server_clock); 0: time of submission */
} dfft_request;

/* req_flags: the result is not read.  The server returns a server-held
 * output buffer to its arena as soon as the request completes (a caller's
 * d_out is written as usual); the request's record (state, times) stays until
 * dfft_server_release, and its d_result reads NULL.  For callers that keep no
 * results, such as throughput runs, so finished results do not hold memory
 * that queued requests are waiting for. */
#define DFFT_REQ_DISCARD  0x1u

#define DFFT_REQ_QUEUED   0
#define DFFT_REQ_RUNNING  1
#define DFFT_REQ_DONE     2
#define DFFT_REQ_FAILED   3

typedef struct {
    int state;              /* DFFT_REQ_* */
    int status;             /* DFFT_OK once done, the error once failed, DFFT_ERR_NOT_READY before */
    int lane;               /* stream it ran on (-1 before admission) */
    int reserved;
    double t_arrival;       /* server clock, seconds */
    double t_admit;
    double t_done;          /* when the scheduler saw its completion event */
    const void *d_result;   /* once done: the caller's d_out or the server buffer holding the result */
    size_t result_bytes;
    double algorithmic_bytes;
} dfft_request_info;

typedef struct {
    int64_t submitted, completed, failed;
    int64_t queued, running, peak_running;
    int64_t plans;          /* resident plans */
    int64_t overtakes;      /* admissions that overtook a request waiting for memory */
    int64_t buffer_allocs;  /* request buffers carved from the arena */
    int64_t arena_chunks;   /* device allocations of the arena (2 GiB chunks for buffers under
                               256 MiB, 8 GiB chunks for larger ones) */
    int64_t arena_waits;    /* requests whose admission waited for a contiguous range */
    size_t budget;
    size_t bytes_held;      /* arena chunks + plans */
    size_t bytes_busy;      /* buffers of running requests and of unreleased results */
    size_t peak_bytes_held;
} dfft_server_stats;

int dfft_server_create(dfft_server **server, const dfft_server_config *config);
/* Seconds on the server's clock (monotonic, 0 at creation). */
int dfft_server_clock(const dfft_server *server, double *now);
/* Build and keep the plans of `count` request shapes ahead of their first submission. */
int dfft_server_prepare(dfft_server *server, const dfft_request *requests, int count);
/* Queue a request; *ticket identifies it.  DFFT_ERR_BUDGET if it can never fit
 * the budget; plan errors (DFFT_ERR_UNSUPPORTED_SIZE, ...) are returned here.
 * The caller's d_in must hold its data when submitted (the server's streams do
 * not wait for the caller's), and d_in/d_out must stay valid until the
 * request is done. */
int dfft_server_submit(dfft_server *server, const dfft_request *request, int64_t *ticket);
/* Wait up to timeout_s (< 0: forever) for one request: DFFT_OK when it is
 * done or failed (see dfft_server_info), DFFT_ERR_NOT_READY on time-out. */
int dfft_server_wait(dfft_server *server, int64_t ticket, double timeout_s);
/* Next finished request not yet returned by this call, in completion order. */
int dfft_server_wait_any(dfft_server *server, double timeout_s, int64_t *ticket);
int dfft_server_info(dfft_server *server, int64_t ticket, dfft_request_info *info);
/* Copy a finished request's result (result_bytes) to d_dst, ordered on `stream`. */
int dfft_server_copy_result(dfft_server *server, int64_t ticket, void *d_dst, size_t bytes, void *stream);
/* Return a finished request's buffers to the pool and forget its ticket;
 * a queued request is cancelled.  DFFT_ERR_STATE while it runs. */
int dfft_server_release(dfft_server *server, int64_t ticket);
/* Wait until nothing is queued or running. */
int dfft_server_drain(dfft_server *server, double timeout_s);
int dfft_server_stats_get(dfft_server *server, dfft_server_stats *stats);
/* Waits for running requests, then frees everything the server holds. */
int dfft_server_destroy(dfft_server *server);

#ifdef __cplusplus
}
#endif

#endif /* DAGGERFFT_HIP_H */
