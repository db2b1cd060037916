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
 * Threading: a plan is immutable after creation except for its workspace; calls on one plan from two streams at
 * once race on the workspace of four-step plans (one plan per stream, as the reference keeps one plan per thread,
 * fft.hpp:17-18).
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
    CRDFFT_ALGO_REFERENCE = 3   /* the radix-2 reference only */
} crdfft_algorithm;

/* Plan flags */
#define CRDFFT_FLAG_NONE 0u
#define CRDFFT_FLAG_NO_SELF_CHECK 1u       /* skip the known-answer check at creation */
#define CRDFFT_FLAG_NORMALIZE_INVERSE 2u   /* C2C inverse scaled by 1/n (ifft_normalized, fft.hpp:1015-1029) */

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
} crdfft_plan_info;

/* Tuning overrides for crdfft_plan_create_tuned; a zero field keeps the planner's choice. */
typedef struct crdfft_tuning
{
    uint64_t n1;          /* four-step split */
    int32_t e, e1, tile1, e2, tile2;
    uint64_t chunk;       /* transforms per pass pair */
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

/* The plan's transforms on device memory, asynchronously on `stream`. in == out runs in place (C2C only). */
CRDFFT_API crdfft_status crdfft_execute(crdfft_plan plan, const void* in, void* out, int32_t direction,
                                        void* stream);
/* Element-major batch in place: element i of transform t at data[i*batch + t]; batch <= the plan's batch may
 * differ from it (the reference passes it per call). C2C plans only. */
CRDFFT_API crdfft_status crdfft_execute_batched(crdfft_plan plan, void* data, uint64_t batch, int32_t direction,
                                                void* stream);
/* The radix-2 reference transform of the plan's batch, out of place (in != out). C2C plans only. */
CRDFFT_API crdfft_status crdfft_execute_reference(crdfft_plan plan, const void* in, void* out, int32_t direction,
                                                  void* stream);
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

CRDFFT_API const char* crdfft_status_string(crdfft_status status);
CRDFFT_API int32_t crdfft_last_hip_error(void);
CRDFFT_API const char* crdfft_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CRDFFT_H */
