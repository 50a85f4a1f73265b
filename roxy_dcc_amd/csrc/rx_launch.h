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
// pushed).  Trees up to depth 23 get 24 entries, deeper ones 64.  The first
// STACK_LDS entries live in LDS (8 x 8 B per lane: 16 KB per 256-thread
// block, so LDS no longer caps occupancy), the rest in a global spill
// column per thread; traced rays rarely go past 8 entries.
constexpr int STACK_LDS = 8;
constexpr int STACK_SHALLOW = 24;
constexpr int STACK_DEEP = 64;
constexpr int BS_SHALLOW = 256;
constexpr int BS_DEEP = 64;
constexpr int TRAV_BLOCKS_PER_CU_SHALLOW = 8;
constexpr int TRAV_BLOCKS_PER_CU_DEEP = 32;

inline bool use_deep_stack(int depth) { return depth + 1 > STACK_SHALLOW; }

// Register budgets of the persistent traversal kernels, as waves per SIMD:
// the closest-hit kernel fits 8 (64 VGPRs) without spilling, the shadow
// kernel with its light loop 6 (80 VGPRs).  Dependent node gathers want
// as many waves in flight as the registers allow.
#define EXTEND_WAVES_PER_EU 8
#define CONNECT_WAVES_PER_EU 6

// Grid of a traversal kernel over up to n rays, and the spill area (int2
// entries) that grid needs.
// per_cu = 0 takes the default (the most blocks the register budget keeps
// resident); a tuned value is clamped to 1 .. that default, which is what
// the spill area is sized for.
inline int trav_grid(long long n, bool deep, int per_cu = 0) {
    int lim = deep ? TRAV_BLOCKS_PER_CU_DEEP : TRAV_BLOCKS_PER_CU_SHALLOW;
    int pc = per_cu <= 0 || per_cu > lim ? lim : per_cu;
    return grid_for(n, deep ? BS_DEEP : BS_SHALLOW, pc);
}
// The render's traversal kernels on shallow trees (rx_render.hip) keep
// the top of the pass's BVH in LDS: the first TL_NODES interior nodes in
// breadth-first order (a treelet, 56 B each), which take about half of a
// ray's interior-node fetches on a 70k-triangle mesh and a third on a
// 2M-triangle scene, off the vector-memory path.  Links into the treelet
// are TL_REF + slot (node indices stay below TL_REF).  One copy serves a
// whole block, so the blocks are large: 1024 threads for the closest-hit
// kernel (2 per CU at 8 waves per SIMD), 512 for the shadow kernel (3 per
// CU at 6), with STACK_LDS_TL stack entries per lane in LDS:
// 1024 x 6 x 8 B + 511 x 56 B = 76.6 KB and 512 x 6 x 8 B + 28.6 KB =
// 52.6 KB, within a CU's 160 KB at those block counts (TL_NODES and
// TL_REF are in rx_device.h).  Over the link-free tree copy the
// closest-hit treelet keeps boxes only (48 B per slot), and the lanes get
// a seventh LDS stack entry: 1024 x 7 x 8 B + 511 x 48 B = 80.0 KB.
#ifndef RX_STACK_LDS_TL
#define RX_STACK_LDS_TL 6
#endif
constexpr int STACK_LDS_TL = RX_STACK_LDS_TL;
constexpr int BS_EXTEND = 1024;
constexpr int EXTEND_BLOCKS_PER_CU = 2;
constexpr int BS_CONNECT = 512;
constexpr int CONNECT_BLOCKS_PER_CU = 3;

// Grid of a render traversal kernel with `bs`-thread blocks, `lim` of
// which fit a CU (per_cu clamped to 1 .. lim as in trav_grid).
inline int render_trav_grid(long long n, int bs, int lim, int per_cu) {
    int pc = per_cu <= 0 || per_cu > lim ? lim : per_cu;
    return grid_for(n, bs, pc);
}

// Spill area of the render kernels: the largest resident grid of either
// stack class times the entries a lane may spill.
inline size_t render_spill_entries(bool deep) {
    long long cu = device_cu_count();
    if (deep)
        return (size_t)(STACK_DEEP - STACK_LDS) * (size_t)(cu * TRAV_BLOCKS_PER_CU_DEEP * BS_DEEP + 8LL * BS_DEEP);
    long long ext = cu * EXTEND_BLOCKS_PER_CU * BS_EXTEND + 8LL * BS_EXTEND;
    long long con = cu * CONNECT_BLOCKS_PER_CU * BS_CONNECT + 8LL * BS_CONNECT;
    return (size_t)(STACK_SHALLOW - STACK_LDS_TL) * (size_t)(ext > con ? ext : con);
}

inline size_t trav_spill_entries(bool deep) {
    long long threads = deep ? (long long)device_cu_count() * TRAV_BLOCKS_PER_CU_DEEP * BS_DEEP
                             : (long long)device_cu_count() * TRAV_BLOCKS_PER_CU_SHALLOW * BS_SHALLOW;
    // grid_for rounds up to a multiple of 8 blocks
    threads += 8LL * (deep ? BS_DEEP : BS_SHALLOW);
    return (size_t)((deep ? STACK_DEEP : STACK_SHALLOW) - STACK_LDS) * (size_t)threads;
}

// Library-owned spill area for the query entry points (grown on demand,
// one per device, kept until the process ends).
int query_spill(bool deep, int2** out);

}  // namespace rx
