// Shared layout of the mixed-size persistent launch (many_kernels.hip) and its host planner (many.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"

namespace crdfft
{

constexpr int kManyThreads = 1024;      // one workgroup per CU
constexpr int kManyTilePoints = 16384;  // complex points per tile (LDS: split re/im + pad)
constexpr int kManyMaxStages = 16;

// One distinct length and its radix plan.
struct ManySize
{
    uint32_t n;
    uint32_t nstages;
    uint32_t radix[kManyMaxStages];
    uint64_t tw_offset; // W_n^e, e < n, at tw[tw_offset + e]
};

// `count` transforms of one length whose element offsets are offsets[first .. first + count).
struct ManyTile
{
    uint32_t size_id;
    uint32_t count;
    uint64_t first;
};

struct ManyArgs
{
    const cfloat* in;
    cfloat* out;
    const ManySize* sizes;
    const ManyTile* tiles;
    const uint64_t* offsets;
    const cfloat* tw;
    const uint32_t* wg_begin; // workgroups + 1 entries
};

unsigned many_lds_bytes();
// Raises the kernels' dynamic LDS limit (once per process, before any launch or capture).
hipError_t many_prepare();
hipError_t launch_many(const ManyArgs& a, unsigned workgroups, bool inverse, hipStream_t s);

} // namespace crdfft
