// The planner: a pure function of (n, batch, type, precision, tuning) to the kernels a plan runs. Host-only, no
// device calls, so crdfft_plan_describe can answer without a GPU.
//
// It plays the part of the reference's planned_forward / four_step_split / level_forward (engine/numerics/
// hesap-fft/include/crd/hesap/fft/fft.hpp:63-210): there nine CPU algorithms are picked by size and CPU level;
// here the choice is between one LDS-resident kernel (n <= 16384) and a two-kernel four-step (n1 x n2, each side
// LDS-resident), with the values per thread, the tile widths and the chunk of transforms per pass pair as knobs.
#pragma once

#include <stdint.h>

#include "../../include/crdfft.h"

namespace crdfft
{

struct PlanShape
{
    uint64_t n = 0;         // user length (real length for R2C/C2R)
    uint64_t batch = 0;
    int32_t type = CRDFFT_C2C;
    int32_t precision = CRDFFT_FP32;
    int32_t algorithm = CRDFFT_ALGO_COPY;
    uint64_t complex_n = 0;
    uint64_t n1 = 0, n2 = 0;
    int32_t e = 0;
    int32_t e1 = 0, tile1 = 0;
    int32_t e2 = 0, tile2 = 0;
    uint64_t chunk = 0;
    uint64_t workspace_bytes = 0;
    uint32_t hints = 0; // kHintNtLoad | kHintNtStore
    int32_t threads = 0; // fused four-step: workgroup size (tile1 columns, tile2 rows per tile)
};

// The largest complex length one LDS-resident kernel runs.
constexpr uint64_t kSingleMax = 16384;
// The largest complex length of the plain path (two LDS-resident passes of at most 4096 each).
constexpr uint64_t kFourStepMax = uint64_t{1} << 24;
// Bytes of four-step intermediate per column/row pass pair: each pass then launches at least 32K workgroups, so
// launch gaps and the partial last wave cost little (measured on MI355X: 64 MiB chunks ran 15% slower than 128 MiB
// and those slower than 256 MiB at n = 2^16; the intermediate does not measurably stay in the Infinity Cache).
constexpr uint64_t kChunkBytes = uint64_t{256} << 20;
// Plan hints of the four-step: bit 0 streams the column pass's input past the caches, bit 1 the row pass's output;
// bit 3 runs the column pass as the persistent prefetching launch; bit 5 runs alternate chunks on a second stream
// (two column/row pipelines in flight, each with its own half-size workspace).
constexpr uint32_t kDefaultFourStepHints = 3;
constexpr uint32_t kHintPersistentColumns = 8;
constexpr uint32_t kHintTwoPipes = 32;
// Columns this long make the column pass latency-bound (one workgroup fills a CU's LDS): such C2C plans default to
// two pipelines, so one chunk's row pass runs beside the next chunk's column pass. Measured on MI355X, back-to-back
// executes: n = 2^20 x 64 5 % less time, 2^22 x 16 4 %, 2^24 x 4 2.5 %; 2^18 x 256 (1024-point columns) 2 % more.
constexpr uint64_t kTwoPipesMinN1 = 2048;
// The largest complex length of the fused four-step (one transform's intermediate per workgroup slot).
constexpr uint64_t kFusedMax = uint64_t{1} << 17;
// Workgroup size of the fused four-step unless tuned: 512 threads (two workgroups per CU hide each other's memory
// waits).
constexpr int kFusedThreads = 512;
// CUs assumed by crdfft_plan_describe for the fused slot count (plan creation asks the device).
constexpr uint64_t kDescribeCUs = 256;
// Up to this length the column pass reads its twiddle W_N^{n2 k1} from one full table (n entries, at most 2 MiB,
// L2-resident): one load and one complex multiply per value instead of two multiplies.
constexpr uint64_t kFullTwiddleMax = uint64_t{1} << 18;

bool is_pow2(uint64_t n);
uint32_t log2_exact(uint64_t n);

crdfft_status choose_shape(uint64_t n, uint64_t batch, int32_t type, int32_t precision, const crdfft_tuning* tuning,
                           PlanShape* out);
void shape_to_info(const PlanShape& s, crdfft_plan_info* info);

// Values per thread and columns per workgroup of the element-major kernel for n (e depends on n only).
bool element_major_shape(uint64_t n, uint64_t batch, int* e, int* c);

// The instantiated kernel shapes (mirrors of the kernels' tables, for the planner and the autotuner).
bool have_row_kernel(uint64_t n, int e);
bool have_col_kernel(uint64_t n1, int e, int c);
bool have_rowtr_kernel(uint64_t n2, int e, int r);
bool have_cols_pf(uint64_t n1, int e, int c);
// Fused four-step of n1 x n2 at threads per workgroup; lds_bytes receives its LDS per workgroup.
bool have_fused_kernel(uint64_t n1, int e1, uint64_t n2, int e2, int threads, unsigned* lds_bytes = nullptr);
// The cooperative kernel of n = n1^2 (at *e values per thread, or the first instantiated when *e is 0): its values
// per thread and workgroup size, false where none is instantiated.
bool coop_shape_of(uint64_t n, uint64_t* n1, int* e, int* threads, unsigned* lds_bytes);
constexpr uint32_t kCoopSpinLimit = 1u << 18;

// The demotion floor of a plan (the radix-2 reference kernels, crdfft_check_status): switch the plan onto it
// (building the reference's W_n table), and whether it runs there.
crdfft_status plan_use_floor(crdfft_plan p);
bool plan_on_floor(crdfft_plan p);

} // namespace crdfft
