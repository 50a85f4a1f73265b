// rx_launch.h — launch-shape helpers shared by the ray kernels.
#pragma once

#include "rx_common.h"
#include "rx_device.h"

namespace rx {

// Blocks are dealt round-robin to the 8 XCDs (block b runs on XCD b % 8).
// Remapping b -> (b % 8) * (G / 8) + b / 8 gives each XCD a contiguous range
// of logical blocks, so neighbouring rays (neighbouring pixels, queue
// slots written by the same wave) share one XCD's L2.  Grids are rounded
// up to a multiple of 8 by grid_for().
__device__ __forceinline__ int xcd_block() {
    int g = gridDim.x, b = blockIdx.x;
    if ((g & 7) != 0) return b;
    return (b & 7) * (g >> 3) + (b >> 3);
}

// Grid for a grid-stride kernel over up to n items with BS threads per
// block: enough blocks to fill the chip `per_cu` times, never more than
// the items need, always a multiple of 8.
inline int grid_for(long long n, int bs, int per_cu) {
    long long need = (n + bs - 1) / bs;
    long long cap = (long long)device_cu_count() * per_cu;
    long long g = need < cap ? need : cap;
    if (g < 8) g = 8;
    return (int)((g + 7) / 8 * 8);
}

// Stack depth classes: a lane's stack never holds more than depth + 1
// entries (one deferred sibling per level plus the two children just
// pushed), so trees up to depth 23 use 24 entries and deeper ones 64.
constexpr int STACK_SHALLOW = 24;
constexpr int STACK_DEEP = 64;
constexpr int BS_SHALLOW = 256;
constexpr int BS_DEEP = 64;

inline bool use_deep_stack(int depth) { return depth + 1 > STACK_SHALLOW; }

}  // namespace rx
