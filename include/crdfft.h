/* crdfft: the C-ABI of the MI355X (gfx950) batched FFT path.
 *
 * Drop-in boundary for the reference's crd-hesap-fft plan/execute interface. Every entry point is plain C: opaque
 * handles, raw device pointers, sizes and an opaque HIP stream (hipStream_t passed as void*; NULL = the null
 * stream). All device buffers belong to the caller; a plan owns its twiddle tables and its workspace.
 *
 * What each entry point replaces in the reference (paths relative to engine/numerics/hesap-fft/):
 *   crdfft_plan_create          FftPlan<T>(IAllocator*, usize n) and RealFftPlan<T>(IAllocator*, usize n):
 *                               include/crd/hesap/fft/fft.hpp:224-304, real_fft.hpp:21-35. Table construction
 *                               (fft.hpp:251-260) runs in FP64 on the device and is narrowed to FP32.
 *   crdfft_plan_destroy         ~FftPlan (fft.hpp:308-341).
 *   crdfft_execute              FftTypeKernels<T>::execute via FftPlan::execute (fft_kernels.hpp:181-185,
 *                               fft.hpp:409-413), batched over contiguous transforms the way detail::run_batch
 *                               splits a batch (detail/batch_arms.hpp:59-77); for R2C/C2R plans RealFftPlan::rfft
 *                               and ::irfft (real_fft.hpp:41-93).
 *   crdfft_execute_batched      FftTypeKernels<T>::execute_batched via FftPlan::execute_batched: element-major
 *                               data[i*batch + t] (fft_kernels.hpp:181-185, fft.hpp:462-470).
 *   crdfft_execute_reference    FftPlan::execute_reference, the radix-2 oracle transform (fft.hpp:415-460).
 *   crdfft_self_check           fft_self_check (src/fft_self_check.cpp:207-211): known-answer inputs through the
 *                               plan's own kernels, answers from an FP64 transform on the host.
 *   crdfft_plan_describe /      FftPlan::state() and fft_algorithm_name (fft.hpp:473, fft_kernels.hpp:41-64): the
 *   crdfft_plan_get_info        planner's choice, as data.
 *
 * Conventions (fft.hpp:16, real_fft.hpp:39-93): forward X[k] = sum_j x[j] exp(-2 pi i jk/n); inverse uses +i;
 * C2C is unnormalized in both directions; C2R (irfft) scales by 1/n. Complex data is interleaved FP32 {re, im}.
 * Errors: the reference asserts on contract violations and aborts on allocation failure (fft.hpp:52, :237); here
 * every entry point returns a status and never aborts the host process. NaN and Inf propagate as they do on the CPU.
 * Threading: a plan is immutable after creation. Its scratch (four-step intermediate, Bluestein and N-D work
 * buffers) is per stream: the buffer allocated with the plan serves the first stream that executes it, and the first
 * execute on any other stream allocates that stream's own (up to 64 streams; beyond, CRDFFT_OUT_OF_MEMORY), so host
 * threads may share one plan on distinct streams (the reference keeps one plan per thread, fft.hpp:17-18). Calls on
 * one stream are ordered by the stream. The first execute on a new stream must not be inside a graph capture.
 */
#ifndef CRDFFT_H
#define CRDFFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRDFFT_API __attribute__((visibility("default")))

typedef enum crdfft_status
{
    CRDFFT_OK = 0,
    CRDFFT_INVALID_SIZE = 1,        /* n outside what the plan type supports (reference: CRD_ASSERT, fft.hpp:237) */
    CRDFFT_INVALID_ARGUMENT = 2,    /* null handle/pointer, bad enum, size mismatch */
    CRDFFT_OUT_OF_MEMORY = 3,       /* device allocation refused (reference: kFftPlanOutOfMemory, fft.hpp:52) */
    CRDFFT_HIP_ERROR = 4,           /* a HIP runtime call failed; crdfft_last_hip_error() holds its code */
    CRDFFT_SELF_CHECK_FAILED = 5,   /* known answers not met (reference: fft_self_check names the routine) */
    CRDFFT_UNSUPPORTED = 6,         /* a layout or precision this build does not run */
    CRDFFT_NO_DEVICE = 7            /* no HIP device visible */
} crdfft_status;

/* FftDirection (fft_kernels.hpp:20-24) */
typedef enum crdfft_direction
{
    CRDFFT_FORWARD = 0,
    CRDFFT_INVERSE = 1
} crdfft_direction;

typedef enum crdfft_type
{
    CRDFFT_C2C = 0, /* FftPlan */
    CRDFFT_R2C = 1, /* RealFftPlan::rfft: n reals -> n/2+1 complex */
    CRDFFT_C2R = 2  /* RealFftPlan::irfft: n/2+1 complex -> n reals, scaled 1/n */
} crdfft_type;

typedef enum crdfft_precision
{
    CRDFFT_FP32 = 0,        /* plain FP32 arithmetic */
    CRDFFT_SPLIT_BF16 = 1,  /* radix stages as DFT-matrix MFMA products, operands split hi+lo in BF16 */
    CRDFFT_SPLIT_FP16 = 2   /* same in FP16 with per-transform power-of-two scaling */
} crdfft_precision;

/* Algorithms a plan can run (the GPU counterpart of FftAlgorithm, fft_kernels.hpp:28-39). */
typedef enum crdfft_algorithm
{
    CRDFFT_ALGO_COPY = 0,       /* n == 1 */
    CRDFFT_ALGO_SINGLE = 1,     /* one LDS-staged kernel per transform (n <= 16384) */
    CRDFFT_ALGO_FOUR_STEP = 2,  /* n = n1*n2: column pass with twiddle, row pass with transposed store */
    CRDFFT_ALGO_REFERENCE = 3,  /* the radix-2 reference only */
    CRDFFT_ALGO_BLUESTEIN = 4,  /* chirp-z over a power-of-two plan of m >= 2n-1 (any n) */
    CRDFFT_ALGO_FUSED = 5,      /* n = n1*n2 <= 2^17: both four-step passes of a transform in one workgroup of a
                                   persistent launch, the intermediate in a per-workgroup slot (chunk = slots) */
    CRDFFT_ALGO_COOP = 6        /* n = n1*n1 (4096, 65536): both passes of a transform split over a group of
                                   workgroups of one XCD, the intermediate shared through that XCD's L2
                                   (chunk = workgroups of the launch) */
} crdfft_algorithm;

/* Plan flags */
#define CRDFFT_FLAG_NONE 0u
#define CRDFFT_FLAG_NO_SELF_CHECK 1u       /* skip the known-answer check at creation */
#define CRDFFT_FLAG_NORMALIZE_INVERSE 2u   /* C2C inverse scaled by 1/n (ifft_normalized, fft.hpp:1015-1029) */
#define CRDFFT_FLAG_SEPARATE_PASSES 4u     /* Bluestein, DCT, R2C and C2R plans: the separate passes even where one
                                              kernel would do (a differential check of the one-kernel forms) */
#define CRDFFT_FLAG_CHIRP_Z 8u             /* Bluestein plans: the chirp-z algorithm even for the 2/3/5/7-smooth n
                                              that otherwise run as one mixed-radix launch */

typedef struct crdfft_plan_s* crdfft_plan;

/* What the planner chose (the GPU counterpart of FftPlanState's size fields, fft_kernels.hpp:69-141). */
typedef struct crdfft_plan_info
{
    uint64_t n;           /* transform length (real length for R2C/C2R) */
    uint64_t batch;
    int32_t type;         /* crdfft_type */
    int32_t precision;    /* crdfft_precision */
    int32_t algorithm;    /* crdfft_algorithm of the complex transform that runs */
    uint64_t complex_n;   /* length of the complex transform (n/2 for R2C/C2R) */
    uint64_t n1, n2;      /* four-step split (0 for single) */
    int32_t e;            /* values per thread of the single/row kernel */
    int32_t e1, tile1;    /* column kernel: values per thread, columns per workgroup */
    int32_t e2, tile2;    /* row kernel: values per thread, rows per workgroup */
    uint64_t chunk;       /* transforms per column+row pass pair (four-step) */
    uint64_t workspace_bytes;
    uint64_t table_bytes;
    uint32_t hints;       /* cache hints: bit 0 non-temporal input loads, bit 1 non-temporal output stores;
                             four-step bit 5 (CRDFFT_HINT_TWO_PIPELINES): alternate chunks run on a second stream */
    int32_t check;        /* crdfft_check_status of the creation-time self-check */
} crdfft_plan_info;

/* What the creation-time known-answer check found (KernelCheckStatus, engine/foundation/core/include/crd/core/
 * kernel_dispatch.hpp:83-99). The plan's selected kernels are checked first; when they fail, the plan drops to its
 * floor, the radix-2 reference kernels of crdfft_execute_reference (bit reversal + butterfly stages over device
 * memory, with the plan's 1/n and the separate real-transform passes around them), is checked again and runs
 * there: DEMOTED. When the floor fails too, creation returns CRDFFT_SELF_CHECK_FAILED (the reference runs its
 * portable table then; a device whose simplest kernels misexecute has nothing safer to offer). Fault seam:
 * CRDFFT_FAULT_SELF_CHECK=1 in the environment fails the selected kernels' check (the plan demotes), =2 fails
 * every check. Element-major batches (crdfft_execute_batched) keep their own kernels. */
typedef enum crdfft_check_status
{
    CRDFFT_CHECK_NOT_RUN = 0, /* CRDFFT_FLAG_NO_SELF_CHECK, or a plan kind the check does not cover */
    CRDFFT_CHECK_PASSED = 1,  /* the selected kernels passed and run */
    CRDFFT_CHECK_DEMOTED = 2  /* the selected kernels failed; the floor passed and runs */
} crdfft_check_status;

/* Tuning overrides for crdfft_plan_create_tuned; a zero field keeps the planner's choice. */
#define CRDFFT_HINTS_NONE 0x100u /* crdfft_tuning.hints: no cache hints (plain loads and stores) */
/* Four-step plans: alternate chunks run on a second stream of the plan's (one per caller stream, forked from and
 * joined back to the caller's stream with events), each chunk pipeline with its own workspace, so one chunk's row
 * pass overlaps the next chunk's column pass. The planner sets it for C2C plans with columns of 2048 points and more
 * and at least two chunks; the first execute on a caller stream creates that stream's pipeline, and an execute
 * captured into a graph before the pipeline exists records both halves on the captured stream instead. */
#define CRDFFT_HINT_TWO_PIPELINES 0x20u
typedef struct crdfft_tuning
{
    uint64_t n1;          /* four-step split */
    int32_t e, e1, tile1, e2, tile2;
    uint64_t chunk;       /* transforms per pass pair */
    uint32_t hints;       /* cache hints (crdfft_plan_info.hints); 0 keeps the planner's, CRDFFT_HINTS_NONE
                             asks for none */
    int32_t algorithm;    /* CRDFFT_ALGO_FOUR_STEP, CRDFFT_ALGO_FUSED or CRDFFT_ALGO_COOP for n > 16384 (COOP also
                             at 4096); 0 keeps the planner's.
                             FUSED: tile1 = columns per phase-A tile (threads = n1/e1 * tile1), tile2 derived,
                             chunk caps the persistent workgroups (slots).
                             COOP: hints bit 16 runs every workgroup alone (a differential check) */
} crdfft_tuning;

/* A plan of `batch` transforms of length n. idist/odist: distance between consecutive transforms in elements of
 * the input/output type (complex for C2C; real for R2C input and C2R output), 0 = packed (n, or n/2+1 for the
 * complex side of R2C/C2R). C2C: any n >= 1 that is a power of two up to 2^24. R2C/C2R: n a power of two >= 4. */
CRDFFT_API crdfft_status crdfft_plan_create(crdfft_plan* plan, uint64_t n, uint64_t batch, uint64_t idist,
                                            uint64_t odist, int32_t type, int32_t precision, uint32_t flags);
CRDFFT_API crdfft_status crdfft_plan_create_tuned(crdfft_plan* plan, uint64_t n, uint64_t batch, uint64_t idist,
                                                  uint64_t odist, int32_t type, int32_t precision, uint32_t flags,
                                                  const crdfft_tuning* tuning);
CRDFFT_API crdfft_status crdfft_plan_destroy(crdfft_plan plan);
/* CRDFFT_ALGO_COOP plans: synchronizes `stream` and reports the plan's last launch on it: waits that gave up (0 when
 * every group barrier completed, so the output is valid) and the mode (1 = groups formed, 2 = every workgroup alone,
 * which is also valid). Other plans report 0 and 0. */
CRDFFT_API crdfft_status crdfft_plan_coop_status(crdfft_plan plan, void* stream, uint32_t* waits_given_up,
                                                 uint32_t* mode);
/* BluesteinPlan<T>(IAllocator*, usize n) (bluestein.hpp:29-70): a C2C plan of any n >= 1 run as a chirp-z
 * transform over a power-of-two plan of m >= 2n-1 (crdfft_plan_info.n1 = m). Forward unnormalized; unlike FftPlan,
 * the inverse is scaled by 1/n (bluestein.hpp:86). Executed with crdfft_execute; in == out is allowed. A
 * 2/3/5/7-smooth n <= 16384 in packed rows runs as one mixed-radix launch of the persistent kernel (crdfft_many_*;
 * crdfft_plan_info.n1 = 0: no convolution) unless CRDFFT_FLAG_CHIRP_Z. Otherwise, for m <= 16384 (n <= 8192) one
 * kernel runs the whole chirp-z transform per sequence (x read once, X written once, no workspace:
 * crdfft_plan_info.e > 0, workspace_bytes 0); longer n run the five passes over an m-point workspace. */
CRDFFT_API crdfft_status crdfft_bluestein_plan_create(crdfft_plan* plan, uint64_t n, uint64_t batch, uint64_t idist,
                                                      uint64_t odist, uint32_t flags);

/* The plan's transforms on device memory, asynchronously on `stream`. in == out runs in place (C2C only). */
CRDFFT_API crdfft_status crdfft_execute(crdfft_plan plan, const void* in, void* out, int32_t direction,
                                        void* stream);
/* Element-major batch in place: element i of transform t at data[i*batch + t]; batch <= the plan's batch may
 * differ from it (the reference passes it per call). C2C plans only. */
CRDFFT_API crdfft_status crdfft_execute_batched(crdfft_plan plan, void* data, uint64_t batch, int32_t direction,
                                                void* stream);
/* The radix-2 reference transform of the plan's batch (fft.hpp:417-460): bit reversal, then log2(n) butterfly
 * stages. in == out runs it in place (the bit reversal by pair swaps, as the reference does on the caller's span);
 * otherwise out of place. C2C plans only. */
CRDFFT_API crdfft_status crdfft_execute_reference(crdfft_plan plan, const void* in, void* out, int32_t direction,
                                                  void* stream);
/* crdfft_execute without any device allocation. No counterpart in the reference: it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Instead of the
 * plan's per-stream workspace (which a first execute on a new stream allocates) the four-step passes use the
 * library's static device scratch, crdfft_noalloc_scratch_bytes() (64 MiB) per device, in chunks that fit it, on
 * the caller's stream alone. Every call on one device shares that scratch: calls that overlap on one device race,
 * so the caller serializes them. Plain FP32 plans of one kernel or the four-step only (C2C, R2C, C2R):
 * CRDFFT_UNSUPPORTED for the fused, cooperative, split-precision and Bluestein forms; CRDFFT_INVALID_SIZE when one
 * transform's four-step intermediate (8 * complex_n bytes) exceeds the scratch. Same results, bit for bit, as
 * crdfft_execute on the same plan. */
CRDFFT_API crdfft_status crdfft_execute_noalloc(crdfft_plan plan, const void* in, void* out, int32_t direction,
                                                void* stream);
CRDFFT_API uint64_t crdfft_noalloc_scratch_bytes(void);
/* Known-answer check of the plan's kernels (synchronous; allocates its own small buffers). */
CRDFFT_API crdfft_status crdfft_self_check(crdfft_plan plan);

/* Per-phase kernel times of one execute (the GPU counterpart of the reference's CRD_HESAP_FFT_PROFILE phase
 * counters, fft_kernels.hpp:199-216, dumped by src/fft.cpp prof::dump): hipEvents recorded on `stream` around every
 * launch, accumulated into *acc (zero it before the first call). Synchronises `stream` before returning.
 * Phases: 0 = single kernel or four-step column pass, 1 = four-step row pass, 2 = real-transform pre/post pass,
 * 3 = copies. points[] counts complex points read and written once by the phase's launches. */
typedef struct crdfft_phase_times
{
    double ms[4];
    uint64_t launches[4];
    uint64_t points[4];
} crdfft_phase_times;
CRDFFT_API crdfft_status crdfft_execute_profiled(crdfft_plan plan, const void* in, void* out, int32_t direction,
                                                 void* stream, crdfft_phase_times* acc);

CRDFFT_API crdfft_status crdfft_plan_get_info(crdfft_plan plan, crdfft_plan_info* info);
/* The planner's choice for a shape without creating a plan or touching a device. */
CRDFFT_API crdfft_status crdfft_plan_describe(uint64_t n, uint64_t batch, int32_t type, int32_t precision,
                                              crdfft_plan_info* info);
/* The same with tuning overrides (what crdfft_plan_create_tuned would build; the fused and cooperative forms count
 * slots for a 256-CU device here and recount them on the plan's device). */
CRDFFT_API crdfft_status crdfft_plan_describe_tuned(uint64_t n, uint64_t batch, int32_t type, int32_t precision,
                                                    const crdfft_tuning* tuning, crdfft_plan_info* info);

/* ---- N-dimensional transforms: NdFftPlan<T>(IAllocator*, ConstSpan<usize> dims) (nd_fft.hpp:28-196) ----
 * Row-major volumes of `ndim` dims (each >= 1), `batch` volumes back to back, in place. Forward is the unnormalized
 * N-D DFT; the inverse is scaled by 1/prod(dims) (nd_fft.hpp:104). Power-of-two axes up to 4096 run as strided
 * column transforms in place; other axes are transposed to contiguous lines, transformed by a 1D plan (power of two,
 * else Bluestein, as the reference picks, nd_fft.hpp:66-78) and transposed back. */
typedef struct crdfft_nd_plan_s* crdfft_nd_plan;
#define CRDFFT_ND_TRANSPOSE_ALL 1u /* flag: run every axis but the last through transposed lines */
CRDFFT_API crdfft_status crdfft_nd_plan_create(crdfft_nd_plan* plan, const uint64_t* dims, uint32_t ndim,
                                               uint64_t batch, uint32_t flags);
CRDFFT_API crdfft_status crdfft_nd_execute(crdfft_nd_plan plan, void* data, int32_t direction, void* stream);
CRDFFT_API crdfft_status crdfft_nd_plan_destroy(crdfft_nd_plan plan);

/* ---- mixed-size batches: one persistent launch over transforms of many 2/3/5/7-smooth lengths ----
 * No single reference call does this: the reference runs such a batch as one plan per length, through Bluestein for
 * every length that is not a power of two (bluestein.hpp), one job per contiguous range (detail/batch_arms.hpp).
 * Here the transforms are grouped by length into tiles, the tiles balanced over one workgroup per CU, and each length
 * runs its own radix plan (radices 2-10, 12, 16) chosen by the autotuner. C2C, FP32, unnormalized. */
typedef struct crdfft_many_plan_s* crdfft_many_plan;

#define CRDFFT_MAX_STAGES 16
/* A radix plan of one length: the stage radices in order (product n) and transforms per tile (0 = as many as fit
 * 16384 points). */
typedef struct crdfft_radix_plan
{
    uint64_t n;
    uint32_t nstages;
    uint32_t radix[CRDFFT_MAX_STAGES];
    uint32_t group;
} crdfft_radix_plan;

typedef struct crdfft_many_info
{
    uint64_t transforms;
    uint64_t lengths;       /* distinct lengths */
    uint64_t tiles;
    uint64_t points;        /* complex points over all transforms */
    uint32_t workgroups;    /* persistent grid */
    double imbalance;       /* max over workgroups of planned cost / mean */
} crdfft_many_info;

/* The radix plans the autotuner chooses among for a 2/3/5/7-smooth n (INVALID_SIZE otherwise); candidate 0 is the
 * planner's default. Writes at most `capacity`, sets *count to the number that exist. */
CRDFFT_API crdfft_status crdfft_radix_candidates(uint64_t n, crdfft_radix_plan* out, uint32_t capacity,
                                                 uint32_t* count);
/* `count` transforms, transform t of lengths[t] at element offset offsets[t] of both in and out (NULL: packed in
 * order). choices: radix plans per length (a length without one gets candidate 0). */
CRDFFT_API crdfft_status crdfft_many_plan_create(crdfft_many_plan* plan, const uint64_t* lengths,
                                                 const uint64_t* offsets, uint64_t count,
                                                 const crdfft_radix_plan* choices, uint64_t nchoices, uint32_t flags);
CRDFFT_API crdfft_status crdfft_many_plan_destroy(crdfft_many_plan plan);
CRDFFT_API crdfft_status crdfft_many_plan_get_info(crdfft_many_plan plan, crdfft_many_info* info);
/* One persistent kernel launch on `stream`. in == out is allowed. */
CRDFFT_API crdfft_status crdfft_many_execute(crdfft_many_plan plan, const void* in, void* out, int32_t direction,
                                             void* stream);
/* Captures one crdfft_many_execute on `stream` into a hipGraph the plan owns (replacing an earlier one), and
 * instantiates it; crdfft_many_graph_launch replays it on any stream. */
CRDFFT_API crdfft_status crdfft_many_graph_capture(crdfft_many_plan plan, const void* in, void* out,
                                                   int32_t direction, void* stream);
CRDFFT_API crdfft_status crdfft_many_graph_launch(crdfft_many_plan plan, void* stream);

/* One autotuner measurement: a length, the candidate index, its radix plan and the measured time per transform. */
typedef struct crdfft_tune_record
{
    crdfft_radix_plan plan;
    uint32_t candidate;
    uint32_t transforms;    /* transforms in the timed launch */
    double us_per_transform;
} crdfft_tune_record;

/* The autotuner of the mixed-size path: for every distinct length of `lengths`, times every candidate radix plan
 * (crdfft_radix_candidates) in the persistent kernel on a synthetic batch of 2^24 points of that length (median of
 * `reps` launches, hipEvents) and writes the fastest plan per length to choices (capacity >= distinct lengths). The
 * batch as given is then timed with those plans and with the planner's defaults, and the defaults are returned if
 * they run it faster. Every per-length measurement is appended to records when records != NULL (at most
 * record_capacity). Synchronous. */
CRDFFT_API crdfft_status crdfft_many_autotune(const uint64_t* lengths, uint64_t count, uint32_t reps,
                                              crdfft_radix_plan* choices, uint64_t capacity, uint64_t* nchoices,
                                              crdfft_tune_record* records, uint64_t record_capacity,
                                              uint64_t* nrecords);

/* ---- the autotuner of single-length plans and the learned cost model ----
 * One timed plan: the shape, the resolved knobs and the median time of one execute. */
typedef struct crdfft_plan_record
{
    uint64_t n;           /* transform length (the volume for N-D) */
    uint64_t batch;
    int32_t type;
    int32_t precision;
    int32_t algorithm;
    uint32_t ndim;        /* 1 for 1D; N-D: dims[0..ndim) */
    uint64_t dims[3];
    uint64_t stride;      /* 1 = contiguous transforms */
    crdfft_tuning tuning; /* N-D: hints bit 2 = CRDFFT_ND_TRANSPOSE_ALL */
    double ms;
} crdfft_plan_record;

#define CRDFFT_FEATURES 32
/* The candidate knob settings of a power-of-two plan (candidate 0: the planner's default). */
CRDFFT_API crdfft_status crdfft_plan_candidates(uint64_t n, uint64_t batch, int32_t type, crdfft_tuning* out,
                                                uint32_t capacity, uint32_t* count);
/* Times one plan (median of `reps` executes after a warm one, hipEvents, device buffers of the plan's shape). */
CRDFFT_API crdfft_status crdfft_time_plan(uint64_t n, uint64_t batch, int32_t type, const crdfft_tuning* tuning,
                                          uint32_t reps, crdfft_plan_record* record);
/* Times an N-D plan (flags: CRDFFT_ND_TRANSPOSE_ALL or 0). */
CRDFFT_API crdfft_status crdfft_time_nd(const uint64_t* dims, uint32_t ndim, uint64_t batch, uint32_t flags,
                                        uint32_t reps, crdfft_plan_record* record);
/* Exhaustive autotuning: times every candidate, returns the fastest; records every measurement. */
CRDFFT_API crdfft_status crdfft_autotune_plan(uint64_t n, uint64_t batch, int32_t type, uint32_t reps,
                                              crdfft_tuning* best, crdfft_plan_record* records,
                                              uint32_t record_capacity, uint32_t* nrecords);
/* CRDFFT_FEATURES floats per record: the factorization of n (exponents of 2, 3, 5, 7), batch, type, precision,
 * stride, algorithm and knobs, N-D dims, all scaled to about [0, 1]. */
CRDFFT_API crdfft_status crdfft_plan_features(const crdfft_plan_record* records, uint64_t count, float* out);

/* A 3-layer perceptron (inputs -> hidden -> hidden -> 1, ReLU, FP32) whose inference and training (forward,
 * backward, Adam) run in this library's own HIP kernels. Parameters, flat: W1[hidden][inputs], b1[hidden],
 * W2[hidden][hidden], b2[hidden], w3[hidden], b3. Host arrays in and out. */
typedef struct crdfft_mlp_s* crdfft_mlp;
CRDFFT_API crdfft_status crdfft_mlp_create(crdfft_mlp* mlp, uint32_t inputs, uint32_t hidden, uint64_t seed);
CRDFFT_API crdfft_status crdfft_mlp_destroy(crdfft_mlp mlp);
CRDFFT_API crdfft_status crdfft_mlp_param_count(crdfft_mlp mlp, uint64_t* count);
CRDFFT_API crdfft_status crdfft_mlp_get_params(crdfft_mlp mlp, float* params);
CRDFFT_API crdfft_status crdfft_mlp_set_params(crdfft_mlp mlp, const float* params);
/* y[i] = f(x[i]) for n rows of `inputs` floats. */
CRDFFT_API crdfft_status crdfft_mlp_predict(crdfft_mlp mlp, const float* x, uint64_t n, float* y);
/* Mean squared error over n rows and its gradient with respect to every parameter (flat layout). */
CRDFFT_API crdfft_status crdfft_mlp_loss_grad(crdfft_mlp mlp, const float* x, const float* t, uint64_t n,
                                              float* loss, float* grad);
/* Adam on the mean squared error: `epochs` passes over n rows in seeded shuffled minibatches of `batch`. */
CRDFFT_API crdfft_status crdfft_mlp_train(crdfft_mlp mlp, const float* x, const float* t, uint64_t n,
                                          uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                                          float* final_loss);

/* ---- the spectral front end over the real FFT (hesap-dsp spectral.hpp) ----
 * A plan of real frames of nperseg samples (a power of two, 4..32768) at hop = nperseg - noverlap: frame s is
 * x[s*hop, s*hop + nperseg), and a signal of len samples has (len - noverlap)/hop frames (0 if len < nperseg), as
 * segment_periodograms_arm / stft_arm count them (spectral.hpp:53-56, :342-345). Every call is one fused launch
 * (plus a small reduction for Welch) on device memory: frame load, optional mean removal, window, rfft, epilogue. */
typedef struct crdfft_spectral_plan_s* crdfft_spectral_plan;
typedef enum crdfft_window
{
    CRDFFT_WINDOW_BOXCAR = 0,
    CRDFFT_WINDOW_HANN = 1 /* periodic hann (windows.hpp:120-124, sym = false), FP64-built */
} crdfft_window;
#define CRDFFT_SPECTRAL_DETREND 1u /* subtract each frame's mean (detrend='constant') */
CRDFFT_API crdfft_status crdfft_spectral_plan_create(crdfft_spectral_plan* plan, uint64_t nperseg, uint64_t noverlap,
                                                    int32_t window, uint32_t flags);
CRDFFT_API crdfft_status crdfft_spectral_plan_destroy(crdfft_spectral_plan plan);
CRDFFT_API uint64_t crdfft_spectral_frames(crdfft_spectral_plan plan, uint64_t len);
/* stft (spectral.hpp:334-386, window hann, no detrend): out = frames x (nperseg/2+1) complex, unscaled rfft of each
 * windowed frame. */
CRDFFT_API crdfft_status crdfft_stft(crdfft_spectral_plan plan, const float* x, uint64_t len, void* out, void* stream);
/* spectrogram / segment periodograms (spectral.hpp:43-126, 175-184): out = frames x (nperseg/2+1) floats, the
 * density |X|^2 / (fs * sum w^2), doubled except DC and Nyquist. periodogram (spectral.hpp:186-222) is one boxcar
 * frame of the whole signal (nperseg = len, noverlap 0). */
CRDFFT_API crdfft_status crdfft_spectrogram(crdfft_spectral_plan plan, const float* x, uint64_t len, double fs,
                                            float* out, void* stream);
/* welch_psd (spectral.hpp:140-172): psd = nperseg/2+1 floats, the frames' density periodograms averaged. The sum runs
 * in a fixed order (lane-group partial rows, then strided row groups added in group order), so the result is bit-identical run to run. Uses
 * the plan's scratch for the partial rows (per stream). All zeros when there is no frame. */
CRDFFT_API crdfft_status crdfft_welch(crdfft_spectral_plan plan, const float* x, uint64_t len, double fs, float* psd,
                                      void* stream);

/* ---- real-to-real transforms (DctPlan<f32>, hesap-fft dct.hpp:34-170) ----
 * A plan of `batch` rows of n reals (n a power of two, 2 <= n <= 2^25), row t at in + t*idist and out + t*odist
 * (0 = packed, n). Conventions of scipy norm=None (dct.hpp:10-17):
 *   DCT-II : y[k] = 2 sum_n x[n] cos(pi (2n+1) k / 2N)
 *   DCT-III: y[k] = x[0] + 2 sum_{n>=1} x[n] cos(pi n (2k+1) / 2N)            (dct3(dct2(x)) = 2N x)
 *   DST-II : y[k] = 2 sum_n x[n] sin(pi (2n+1) (k+1) / 2N)
 *   DST-III: y[k] = (-1)^k x[N-1] + 2 sum_{n<=N-2} x[n] sin(pi (n+1) (2k+1) / 2N)  (dst3(dst2(x)) = 2N x)
 * computed by Makhoul's reduction to one complex FFT of n/2 points. n <= 32768 runs one kernel per call (each row
 * read once and written once, no workspace); longer n, or CRDFFT_FLAG_SEPARATE_PASSES, run a shuffle / rotation
 * pass, the library's R2C or C2R plan and an output pass over a chunked workspace (per stream). in == out is
 * allowed when idist == odist. */
typedef struct crdfft_dct_plan_s* crdfft_dct_plan;
typedef enum crdfft_r2r_kind
{
    CRDFFT_DCT2 = 0, /* DctPlan::dct2 */
    CRDFFT_DCT3 = 1, /* DctPlan::dct3 */
    CRDFFT_DST2 = 2, /* DctPlan::dst2 */
    CRDFFT_DST3 = 3  /* DctPlan::dst3 */
} crdfft_r2r_kind;
/* DctPlan<T>(IAllocator*, usize n) (dct.hpp:38-57), for a batch of rows. flags: CRDFFT_FLAG_NO_SELF_CHECK,
 * CRDFFT_FLAG_SEPARATE_PASSES. */
CRDFFT_API crdfft_status crdfft_dct_plan_create(crdfft_dct_plan* plan, uint64_t n, uint64_t batch, uint64_t idist,
                                                uint64_t odist, uint32_t flags);
CRDFFT_API crdfft_status crdfft_dct_plan_destroy(crdfft_dct_plan plan);
/* n, batch, complex_n = n/2, e (> 0: the one-kernel form), workspace and table bytes. */
CRDFFT_API crdfft_status crdfft_dct_plan_get_info(crdfft_dct_plan plan, crdfft_plan_info* info);
/* DctPlan::dct2 / dct3 / dst2 / dst3 (dct.hpp:62-170) of every row, asynchronously on `stream`. */
CRDFFT_API crdfft_status crdfft_dct_execute(crdfft_dct_plan plan, int32_t kind, const float* in, float* out,
                                            void* stream);

CRDFFT_API const char* crdfft_status_string(crdfft_status status);
CRDFFT_API int32_t crdfft_last_hip_error(void);
CRDFFT_API const char* crdfft_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CRDFFT_H */
