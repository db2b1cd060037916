// Host-side view of the HIP kernels of the batched FFT path (fft_kernels.hip): launch descriptors and launchers.
// Every launcher checks its shape against what the kernel assumes and returns false for a shape it has no
// instantiation for; it never launches a grid the kernel's indexing does not cover.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace crdfft
{

struct cfloat
{
    float x, y;
};

// Contiguous transforms of one size in one workgroup each (or several per workgroup below 256 threads):
// in/out at t*dist + i, `count` transforms.
struct RowArgs
{
    const cfloat* in;
    cfloat* out;
    const cfloat* tw; // stage twiddles of n (stage_twiddle_entries(n, e) entries)
    uint64_t count;
    uint64_t idist;
    uint64_t odist;
    float scale; // applied to every output (1 = unscaled, exact)
    uint32_t hints; // kHintNtLoad / kHintNtStore
};

// Cache hints: non-temporal loads of the input / stores of the output (streamed data that should not displace a
// cache-resident intermediate).
constexpr uint32_t kHintNtLoad = 1u;
constexpr uint32_t kHintNtStore = 2u;
// Plan knob beside the cache hints: run the four-step's column pass as a persistent launch that prefetches its next
// tile (k_cols_pf), where that kernel is instantiated.
constexpr uint32_t kHintPersistent = 8u;
// Launch-time flag of a C2C four-step whose column pass has T1 = n1/e1 threads per column equal to the row pass's
// rows per workgroup and tile1 columns equal to its n2/e2 threads per row: the intermediate is laid out in tiles of
// T1 rows x tile1 columns, ordered [k1 / T1][n2 / tile1], so the column pass stores each of its row slots as one run
// of T1*tile1 points and the row pass reads its rows as runs of as many (instead of tile1-point row segments at the
// matrix's row stride). Not a plan hint: the executor sets it on both passes' arguments.
constexpr uint32_t kHintTiled = 64u;

// Column transforms of length n1 over `ncols` adjacent columns of a row-major n1 x ncols matrix, for `count`
// matrices at t*dist: four-step pass 1 (with the factored twiddle W_N^{n2 k1} when thi != nullptr) and the
// element-major batch (ncols = batch, no twiddle).
struct ColArgs
{
    const cfloat* in;
    cfloat* out;
    const cfloat* tw;  // stage twiddles of n1
    const cfloat* thi;   // W_N^{j << hshift}
    const cfloat* tlo;   // W_N^{j}, j < 2^hshift
    const cfloat* tstep; // W_N^{n2 * (n1/E) * m} at [m * ncols + n2], m < E (the per-slot step of the twiddle)
    const cfloat* tfull; // when set: W_N^{n2 k1} at [k1 * ncols + n2] (used instead of thi/tlo/tstep)
    uint32_t hshift;
    uint64_t ncols;
    uint64_t count;
    uint64_t idist;
    uint64_t odist;
    uint32_t hints;
    // k_cols_c2r only: the input rows are the h + 1 bins of C2R transforms (h = n1 * ncols); each column value is
    // formed as pre_scale * Z[j] from X[j], X[h-j] and wreal[j] = W_n^j on load
    const cfloat* wreal = nullptr;
    float pre_scale = 1.0f;
};

// Row transforms of length n2 over the n1 rows of a row-major n1 x n2 matrix, stored transposed:
// element k2 of row k1 goes to out[k1 + n1*k2] (four-step pass 2).
struct RowTrArgs
{
    const cfloat* in;
    cfloat* out;
    const cfloat* tw; // stage twiddles of n2
    uint64_t n1;
    uint64_t count;
    uint64_t idist;
    uint64_t odist;
    float scale; // applied to every output (1 = unscaled, exact)
    uint32_t hints;
    const cfloat* wreal = nullptr; // k_rowtr_r2c only: W_n^k, k <= h (n = 2h), and out holds h + 1 bins per row
};

// Both four-step passes of n = n1 * n2 in one persistent launch (k_fused): workgroup b runs transforms b, b + grid,
// ... through its own n-point slot of `ws` (slots >= grid slots of n1 * n2 points).
struct FusedArgs
{
    const cfloat* in;
    cfloat* out;
    cfloat* ws;
    const cfloat* tw1;   // stage twiddles of n1
    const cfloat* tw2;   // stage twiddles of n2
    const cfloat* tfull; // W_n^{c k1} at [k1 * n2 + c]
    uint64_t slots;
    uint64_t count;
    uint64_t idist;
    uint64_t odist;
    float scale;
    uint32_t hints;
};

// One square transform n = n1 * n1 split over a group of workgroups of one XCD (k_coop): each member runs one
// column tile and one row tile of every transform its group takes, and the group's n-point slot is shared through
// that XCD's L2. Groups form at launch from the XCC_ID each workgroup reads; a launch that does not become resident
// as a whole (or kHintCoopSolo) runs every workgroup alone on its own slot.
struct CoopArgs
{
    const cfloat* in;
    cfloat* out;
    cfloat* slots;     // gridDim.x slots of n points
    uint32_t* ctl;     // coop_ctl_words(grid) control words, cleared by launch_coop
    const cfloat* tw;  // stage twiddles of n1
    const cfloat* thi; // W_n^{n1 j}, j < n1
    const cfloat* tlo; // W_n^{j}, j < n1
    uint64_t count;
    uint64_t idist;
    uint64_t odist;
    float scale;
    uint32_t hints;
    uint32_t spin_limit; // polls (each followed by a short sleep) before a wait gives up and flags an error
};
constexpr uint32_t kHintCoopSolo = 16u;
// control words: grid arrival, mode, error, 8 per-XCD ticket counters, then two barrier counters per unit (each on
// its own 128-byte line)
constexpr uint32_t kCoopCtlGrid = 0, kCoopCtlMode = 32, kCoopCtlError = 64, kCoopCtlXcd = 96, kCoopCtlUnits = 352;
inline uint64_t coop_ctl_words(uint64_t grid) { return kCoopCtlUnits + 64 * grid; }

struct LaunchShape
{
    unsigned threads = 0;     // workgroup size
    unsigned per_group = 0;   // rows (transforms, columns) one workgroup covers
    unsigned lds_bytes = 0;   // dynamic LDS per workgroup
};

// The values per thread a kernel of size n runs with (E), chosen by the planner among the instantiated ones.
bool row_kernel_shape(uint64_t n, int e, LaunchShape* out);
bool col_kernel_shape(uint64_t n1, int e, int c, LaunchShape* out);
bool rowtr_kernel_shape(uint64_t n2, int e, int r, LaunchShape* out);

bool fused_kernel_shape(uint64_t n1, int e1, uint64_t n2, int e2, int threads, LaunchShape* out);
bool coop_kernel_shape(uint64_t n1, int e, int threads, LaunchShape* out);
// Resident workgroups per CU of the cooperative kernel (sets its LDS attribute first).
hipError_t coop_blocks_per_cu(uint64_t n1, int e, int threads, int* blocks);
// Clears the control words and launches `grid` workgroups (grid <= the slots behind a.slots).
hipError_t launch_coop(uint64_t n1, int e, int threads, bool inverse, const CoopArgs& a, unsigned grid,
                       hipStream_t s);
// Resident workgroups per CU of the fused kernel (sets its LDS attribute first).
hipError_t fused_blocks_per_cu(uint64_t n1, int e1, uint64_t n2, int e2, int threads, int* blocks);

hipError_t launch_rows(uint64_t n, int e, bool inverse, const RowArgs& a, hipStream_t s);
hipError_t launch_fused(uint64_t n1, int e1, uint64_t n2, int e2, int threads, bool inverse, const FusedArgs& a,
                        uint64_t grid, hipStream_t s);
hipError_t launch_cols(uint64_t n1, int e, int c, bool inverse, const ColArgs& a, hipStream_t s);

// Element-major batches of n = n1 * n2 > 4096 (data[i*batch + t], fft.hpp:462-470): the four-step over strided
// columns. Each pass transforms `ncols` adjacent columns of `count` matrices, matrix t at in + t*idist with rows
// in_ld apart, and writes them to out + t*odist with rows out_ld apart (64-bit offsets throughout). With thi set,
// the pass multiplies value k1 of matrix t by W_n^{t*k1} = thi[e >> hshift] * tlo[e & (2^hshift - 1)],
// e = (t0 + t)*k1.
struct EmArgs
{
    const cfloat* in;
    cfloat* out;
    const cfloat* tw; // stage twiddles of the column length
    const cfloat* thi;
    const cfloat* tlo;
    uint32_t hshift;
    uint64_t ncols;
    uint64_t count;
    uint64_t idist;
    uint64_t odist;
    uint64_t in_ld;
    uint64_t out_ld;
    uint64_t t0 = 0; // matrix index of the first matrix (the twiddle exponent counts from it)
};
// (values per thread, columns per workgroup) of the element-major column kernel of length len, or false.
bool em_col_shape(uint64_t len, int* e, int* c);
hipError_t launch_cols_em(uint64_t len, bool inverse, const EmArgs& a, hipStream_t s);
// The inverse column pass of a C2R plan's half-length four-step with the pre-rotation fused into its loads.
bool have_cols_c2r(uint64_t n1, int e, int c);
hipError_t launch_cols_c2r(uint64_t n1, int e, int c, const ColArgs& a, hipStream_t s);
// The persistent prefetching column pass (same arithmetic as launch_cols; ncols a multiple of c).
bool have_cols_pf(uint64_t n1, int e, int c);
hipError_t launch_cols_pf(uint64_t n1, int e, int c, bool inverse, const ColArgs& a, hipStream_t s);
hipError_t launch_rowtr(uint64_t n2, int e, int r, bool inverse, const RowTrArgs& a, hipStream_t s);
// The forward row pass of an R2C plan's half-length four-step with the Hermitian recombine fused in (k_rowtr_r2c):
// out (a.odist complex apart) gets the h + 1 bins of each real transform; a.wreal = W_n^k, k <= h.
bool have_rowtr_r2c(uint64_t n2, int e, int r);
hipError_t launch_rowtr_r2c(uint64_t n2, int e, int r, const RowTrArgs& a, hipStream_t s);

// Entries of the stage twiddle table of an n-point kernel at e values per thread, and the FP64 device build of it.
uint64_t stage_twiddle_entries(uint64_t n, int e);
hipError_t build_stage_twiddles(cfloat* table, uint64_t n, int e, hipStream_t s);
// out[j] = W_d^{j * mult} for j < count, computed in FP64 on the device and narrowed.
hipError_t build_roots(cfloat* out, uint64_t count, uint64_t mult, uint64_t d, hipStream_t s);

// The full four-step twiddle table out[k1 * ncols + c] = W_n^{c k1}, k1 < rows.
hipError_t build_full_twiddle(cfloat* out, uint64_t rows, uint64_t ncols, uint64_t n, hipStream_t s);
// The column pass's step table: out[m * ncols + n2] = W_n^{n2 * step * m}, m < e (step = n1 / e).
hipError_t build_step_table(cfloat* out, uint64_t ncols, int e, uint64_t step, uint64_t n, hipStream_t s);

// Bluestein passes (bluestein_kernels.hip): chirp pre-multiply + zero pad, pointwise product with FFT_m(b), chirp
// post-multiply (with the inverse's conjugation and 1/n), and the FP64-built chirp / b tables.
hipError_t launch_blue_pre(const cfloat* x, cfloat* w, const cfloat* chirp, uint64_t n, uint64_t m, uint64_t count,
                           uint64_t idist, bool inverse, hipStream_t s);
hipError_t launch_blue_mul(cfloat* w, const cfloat* bhat, uint64_t m, uint64_t count, hipStream_t s);
hipError_t launch_blue_post(const cfloat* w, cfloat* y, const cfloat* chirp, uint64_t n, uint64_t m, uint64_t count,
                            uint64_t odist, bool inverse, hipStream_t s);
hipError_t build_blue_tables(cfloat* chirp, cfloat* b, uint64_t n, uint64_t m, hipStream_t s);

// The whole Bluestein transform in one kernel for m <= 16384 (fft_kernels.hip): one workgroup slot per transform
// loads x * conj(c) (zero past n), runs FFT_m, multiplies by FFT_m(b), runs the unnormalized IFFT_m and stores
// conj(c) * w * post_scale (the inverse conjugated) for k < n. tw is the stage table of (m, e).
struct BlueArgs
{
    const cfloat* in;
    cfloat* out;
    const cfloat* chirp; // n entries
    const cfloat* bhat;  // m entries
    const cfloat* tw;
    uint64_t n;
    uint64_t count;
    uint64_t idist;
    uint64_t odist;
};
// The values per thread of the one-kernel form for convolution length m, or 0 where it has no instantiation.
int blue_fused_e(uint64_t m);
hipError_t launch_blue_fused(uint64_t m, int e, bool inverse, const BlueArgs& a, hipStream_t s);

// out[t][c][r] = in[t][r][c] (nd_kernels.hip), and data *= scale.
hipError_t launch_transpose(const cfloat* in, cfloat* out, uint64_t rows, uint64_t cols, uint64_t count,
                            uint64_t idist, uint64_t odist, hipStream_t s);
hipError_t launch_scale(cfloat* data, uint64_t count, float scale, hipStream_t s);
// data[t*dist + i] *= scale for i < n, t < count.
hipError_t launch_scale_rows(cfloat* data, uint64_t n, uint64_t count, uint64_t dist, float scale, hipStream_t s);

// The radix-2 reference transform (bit reversal + log2 n butterfly stages over global memory) of `count`
// contiguous transforms; `w` holds W_n^k for k < n/2.
hipError_t launch_reference(const cfloat* in, cfloat* out, const cfloat* w, uint64_t n, uint64_t count,
                            uint64_t idist, uint64_t odist, bool inverse, hipStream_t s);

// out[t*odist + i] = in[t*idist + i] for i < n (n = 1 transforms, and host-side copies of a batch).
// Streaming copy of `bytes` (a multiple of 16) with 16-byte vector accesses: the device-copy ceiling probe.
hipError_t launch_stream_copy(void* dst, const void* src, uint64_t bytes, hipStream_t s);

hipError_t launch_copy(const cfloat* in, cfloat* out, uint64_t n, uint64_t count, uint64_t idist, uint64_t odist,
                       hipStream_t s);

} // namespace crdfft
