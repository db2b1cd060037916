// Shared layout of the mixed-size persistent launch (many_kernels.hip) and its host planner (many.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace crdfft
{

constexpr int kManyThreads = 1024;      // one workgroup per CU
constexpr int kManyTilePoints = 16384;  // complex points per tile (LDS: interleaved complex + pad)
constexpr int kManyMaxStages = 16;
constexpr int kManyTwLo = 128;          // twiddle of a length: W_n^e = hi[e >> 7] * lo[e & 127]
constexpr int kManyTwEntries = 2 * kManyTwLo;

// The radices the kernel has butterflies for (radices 14 and 15 would need more than the 128 VGPRs of a
// 1024-thread workgroup beside the prefetched tile).
inline bool many_radix_supported(uint32_t r)
{
    switch (r)
    {
        case 1: case 2: case 3: case 4: case 5: case 6: case 7: case 8: case 9: case 10: case 12: case 16:
            return true;
        default:
            return false;
    }
}

// x / d for x * d < 2^32 as one multiply-high: magic = ceil(2^32 / d), 0 for d == 1.
inline uint32_t many_magic(uint32_t d) { return d <= 1 ? 0u : static_cast<uint32_t>(((uint64_t{1} << 32) + d - 1) / d); }

// One Stockham stage of a length: radix R, butterflies per transform n / R, sub-DFT size p before the stage, and
// the stage twiddle W_{pR} = W_n^{twstep}.
struct ManyStage
{
    uint32_t radix;
    uint32_t nr, nr_magic;
    uint32_t p, p_magic;
    uint32_t twstep;
};

// One distinct length and its radix plan.
struct ManySize
{
    uint32_t n;
    uint32_t n_magic;
    uint32_t nstages;
    uint32_t reserved;
    uint64_t tw_offset; // lo[128] then hi[128] at tw[tw_offset ...]
    ManyStage stage[kManyMaxStages];
};

// `count` transforms of one length, contiguous at element offset base of both in and out.
struct ManyTile
{
    uint32_t size_id;
    uint32_t count;
    uint64_t base;
};

struct ManyArgs
{
    const cfloat* in;
    cfloat* out;
    const ManySize* sizes;
    const ManyTile* tiles;
    const cfloat* tw;
    const uint32_t* wg_begin; // workgroups + 1 entries
};

hipError_t launch_many(const ManyArgs& a, unsigned workgroups, bool inverse, hipStream_t s);

} // namespace crdfft
