/*
 * flagfft_ext.h — MI355X-build extensions beside the drop-in API of flagfft.h.
 *
 * None of these has a counterpart in the reference's include/flagfft.h; they
 * expose what this build adds (host-side plan inspection for tests and tools,
 * per-plan launch statistics for byte accounting, and tuning overrides).
 */
#pragma once

#include <stddef.h>

#include "flagfft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Library version string ("flagfft-gfx950 <major>.<minor>"). */
const char *flagfftxVersion(void);

/* Run only the planner (no device is touched) for the same arguments that
 * flagfftPlanMany would validate with default (contiguous) layouts, and write
 * the plan tree text into buf (truncated to len-1 bytes, NUL-terminated).
 * Returns the same validation codes as flagfftPlanMany (reference:
 * src/exec/c_api_plan.cpp:274-336, src/exec/c_api_internal.cpp:60-120). */
flagfftResult flagfftxPlanDescribe(int rank, const int *n, flagfftType type, int batch, char *buf, size_t len);

/* Number of kernel launches one Exec of `direction` enqueues and the sum of
 * their staged HBM bytes (read + write of every pass, workspaces included). */
flagfftResult flagfftxPlanStats(flagfftHandle plan, int direction, int *launches, long long *pass_bytes);

/* Per-launch HIP-event timing of a plan's kernels, recorded on the plan's
 * stream around every launch while enabled; enabling or disabling resets the
 * accumulated totals. */
flagfftResult flagfftxSetProfiling(flagfftHandle plan, int enable);

/* Waits for the recorded events and reports, for each launch of one Exec of
 * `direction`: accumulated milliseconds, number of timed calls, and the
 * launch's algorithmic HBM bytes (read + write).  *nsteps receives the launch
 * count; at most `capacity` entries are written to each array. */
flagfftResult flagfftxGetProfile(flagfftHandle plan, int direction, int capacity, double *ms_total, long long *calls,
                                 long long *bytes, int *nsteps);

/* Label of launch `step` of `direction` (kernel family, shape); NULL when out of range. */
const char *flagfftxStepLabel(flagfftHandle plan, int direction, int step);

/* Plan-time tunables for plans created afterwards by this thread (0 keeps
 * the default).  Used by the autotuner to sweep decompositions. */
typedef struct flagfftxTuning_t {
  int leaf_max;       /* longest single-pass LDS transform */
  int pass_max;       /* longest super-radix of a multi-pass chain */
  int units;          /* transforms (or butterflies) per workgroup; for an
                         LDS-resident Bluestein core the form instead: 1 / 2
                         Stockham stages (prefetched / per-stage twiddles),
                         3 / 4 in-place stages (table / binary-powered
                         twiddles), 5 one workgroup per sequence, 6 / 7 the
                         convolution split over 2 / 4 workgroups */
  int ept;            /* elements per thread: 16 or 32 (fp32), 8 or 16 (fp64) */
  int chunk_bytes;    /* contiguous bytes per row segment in strided passes */
  int npasses;        /* explicit super-radix split, `npasses` entries of `passes` */
  int passes[4];
  int nradices;       /* explicit single-pass radix order */
  int radices[16];
  int stage_form;     /* kernel of the passes of a multi-pass chain: 0 automatic,
                         1 register-resident two-stage passes where compiled,
                         2 LDS Stockham super-stages only */
  int variant;        /* instance of a kernel family compiled in several launch
                         shapes, 1-based (0: the default): the split-precision
                         1024-point kernel's prefetch depth / occupancy */
} flagfftxTuning;

flagfftResult flagfftxSetTuning(const flagfftxTuning *tuning); /* NULL restores defaults */

/* Tuned-plan database.  Plan creation (1D, vector arithmetic, no explicit
 * flagfftxSetTuning on the thread) looks up the newest valid row of table
 * tuned_measurements for (gfx950, nx, batch bucket, input dtype, direction,
 * layout) and applies its overrides; C2C/Z2Z use the forward row for both
 * directions.  The file is $FLAGFFT_TUNE_DB, else tuned_plans.sqlite beside
 * libflagfft.so; FLAGFFT_TUNE_DISABLE=1 turns lookups off.  Same table, key
 * columns and batch buckets as the reference's cache
 * (src/utils/plan_json.cpp:170-207, src/exec/tuned_plan.cpp:21-58).
 * Store appends a row (the autotuner's winner, with its measured median and
 * the default plan's for comparison). */
flagfftResult flagfftxTunedStore(int nx, flagfftType type, int batch, const flagfftxTuning *tuning,
                                 double median_us, double gbps, double default_us);
flagfftResult flagfftxTunedDbPath(char *buf, size_t len);

/* Graph mode: each Exec replays the plan's kernel launches as one hipGraph
 * (recorded at the first Exec of each direction; later Execs patch only the
 * nodes that touch the caller's buffers).  Results are identical to the
 * default mode.  While profiling is enabled Exec launches directly. */
flagfftResult flagfftxSetGraphMode(flagfftHandle plan, int enable);

/* flagfftExecC2C without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The plan's workspaces
 * are taken from one fixed 64 MiB array per device, held from the library's
 * first use on that device until the process ends, instead of the plan's own;
 * a plan needing more (e.g. a multi-pass C2C past 8M elements per exec)
 * returns FLAGFFT_INVALID_SIZE.  Results are bit-identical to
 * flagfftExecC2C.  All calls on one device share that array: calls that
 * overlap on one device (other streams, other threads) race and must be
 * serialised by the caller. */
flagfftResult flagfftxExecC2CNoAlloc(flagfftHandle plan, flagfftComplex *idata, flagfftComplex *odata, int direction);

/* Arithmetic of a plan.  VECTOR is the FP32/FP64 path of flagfftPlan1d.  The
 * split-precision modes run the two radix-32 stages of a length-1024 C2C
 * transform as dense complex matrix products on the matrix cores: every FP32
 * operand is split into hi + lo halves (BF16 or FP16), three products per real
 * product (hi*hi + hi*lo + lo*hi) accumulate in FP32. */
typedef enum flagfftxCompute_t {
  FLAGFFTX_COMPUTE_VECTOR = 0,
  FLAGFFTX_COMPUTE_BF16X3 = 1,
  FLAGFFTX_COMPUTE_FP16X3 = 2
} flagfftxCompute;

/* flagfftPlan1d with an arithmetic mode.  The split-precision modes accept
 * FLAGFFT_C2C with nx == 1024 only (FLAGFFT_NOT_SUPPORTED otherwise); an
 * unknown mode is FLAGFFT_INVALID_VALUE.  Exec, SetStream and Destroy are the
 * usual calls. */
flagfftResult flagfftxPlan1dCompute(flagfftHandle *plan, int nx, flagfftType type, int batch, int compute);

/* Rank-3 transforms of `batch` contiguous nx x ny x nz arrays (row-major, nz
 * contiguous; real transforms use nx x ny x (nz/2+1) complex on the
 * spectrum side), executed with the usual flagfftExec* calls: rows along nz,
 * then the ny-point and nx-point columns in place of transposes.  No
 * counterpart in the reference, whose flagfftPlan3d and rank-3
 * flagfftPlanMany return FLAGFFT_NOT_SUPPORTED (ctest/test_plan.cpp:151-159);
 * those two keep that answer here.  Errors: null plan or batch <= 0 ->
 * FLAGFFT_INVALID_VALUE, a size <= 0 -> FLAGFFT_INVALID_SIZE, bad type ->
 * FLAGFFT_INVALID_TYPE, no device -> FLAGFFT_INVALID_DEVICE. */
flagfftResult flagfftxPlan3d(flagfftHandle *plan, int nx, int ny, int nz, flagfftType type, int batch);

/* flagfftPlanMany with the advanced layouts as well: for FLAGFFT_C2C /
 * FLAGFFT_Z2Z of rank 1 or 2, any element stride >= 1, distance >= 1 and
 * embed (embed[rank-1] >= n[rank-1]; element y of row x of batch entry b at
 * b * dist + (x * embed[1] + y) * stride); for the real types of rank 2,
 * padded rows (unit strides, embed[1] >= the row length of that side,
 * dist = n[0] * embed[1], an even real-side pitch when n[1] is even; the
 * in-place layout of 2 (n[1]/2+1) reals per row is one case).  The reference
 * answers all of these with FLAGFFT_NOT_SUPPORTED
 * (src/exec/c_api_internal.cpp:60-120; flagfftPlanMany keeps that answer
 * here).  Layouts flagfftPlanMany accepts behave exactly as there; other real
 * layouts and rank 3 stay FLAGFFT_NOT_SUPPORTED; a complex stride, distance
 * or embed out of range is FLAGFFT_INVALID_VALUE.  No counterpart in the
 * reference.  The output layout must not overlap the input unless both are
 * the same layout in the same buffer (or the real in-place layout). */
flagfftResult flagfftxPlanMany(flagfftHandle *plan, int rank, int *n, int *inembed, int istride, int idist,
                               int *onembed, int ostride, int odist, flagfftType type, int batch);

/* ---- serving loop ---------------------------------------------------------
 * A server owns a device arena of `hbm_budget_bytes` for request outputs and
 * workspaces, a cache of plans and twiddle tables resident on the device, and
 * a dispatcher thread.  Submit enqueues a request (device input pointer; the
 * input must stay valid until the request completes) and returns at once.  A
 * request is admitted when its output and workspace fit in the arena (FIFO),
 * so requests that arrive while others run start as memory frees up.
 *   BATCHED: admitted single-pass requests are formed into batches, one per
 *     free batch stream (up to max_inflight batches overlap on the device),
 *     and run together in one persistent mixed-length kernel launch that
 *     drains a tile queue; longer or larger requests run their own plan's
 *     kernels on two further streams: those above 64 MiB in arrival order on
 *     the first, the others on the less loaded one, so they pass the large
 *     ones and overlap with the batches.
 *   PLAIN: each admitted request is launched on its own, in arrival order, on
 *     one stream, with its cached plan (what per-request Exec calls do).
 * The output stays in the arena until Release.  Latency is measured from
 * Submit to the completion of the request's last kernel, as seen by the
 * dispatcher. */
typedef struct flagfftxServer_st *flagfftxServer;

typedef enum flagfftxServeMode_t { FLAGFFTX_SERVE_BATCHED = 0, FLAGFFTX_SERVE_PLAIN = 1 } flagfftxServeMode;

typedef struct flagfftxServerStats_t {
  long long completed;       /* requests completed since the last reset */
  long long transforms;      /* sum of their batch counts */
  double bytes;              /* algorithmic bytes (input + output) of those requests */
  double span_s;             /* first arrival to last completion */
  double p50_ms, p99_ms, mean_ms, max_ms;  /* request latency */
  long long batches;         /* dispatches (BATCHED) or launches (PLAIN) */
  long long queue_launches;  /* persistent-kernel launches */
  long long queued_requests; /* requests served by the persistent kernel */
  double arena_peak_bytes;   /* high-water mark of the arena */
  double table_bytes;        /* resident plan tables */
} flagfftxServerStats;

flagfftResult flagfftxServerCreate(flagfftxServer *server, size_t hbm_budget_bytes, int max_inflight, int mode);
/* Build (and keep) the plans for one transform kind ahead of traffic. */
flagfftResult flagfftxServerPrepare(flagfftxServer server, flagfftType type, int nx, int direction);
/* nx * batch above 2^31 - 1 -> FLAGFFT_INVALID_SIZE (split such a request). */
flagfftResult flagfftxServerSubmit(flagfftxServer server, flagfftType type, int nx, int batch, int direction,
                                   const void *idata, long long *ticket);
/* done = 1 when complete; then *odata is the output in the arena. */
flagfftResult flagfftxServerPoll(flagfftxServer server, long long ticket, int *done, void **odata,
                                 double *latency_ms);
flagfftResult flagfftxServerWait(flagfftxServer server, long long ticket, void **odata, double *latency_ms);
/* Return a completed request's output memory to the arena. */
flagfftResult flagfftxServerRelease(flagfftxServer server, long long ticket);
flagfftResult flagfftxServerGetStats(flagfftxServer server, flagfftxServerStats *stats);
flagfftResult flagfftxServerResetStats(flagfftxServer server);
/* Waits for every submitted request, then frees the server.  Unlike plan
 * handles, a server handle is not looked up in a table: it must not be used
 * after Destroy, nor destroyed while another thread is inside a call on it. */
flagfftResult flagfftxServerDestroy(flagfftxServer server);

#ifdef __cplusplus
}
#endif
