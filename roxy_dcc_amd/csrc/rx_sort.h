// rx_sort.h — stable LSD radix sort of (key, value) u32 pairs for the ray
// queues, with the length read on the device.
//
// The queue length of a bounce is only known on the device, so every
// kernel here reads it from `count` (clamped to the capacity the grid was
// sized for) and the host never waits for it; nothing is cleared with a
// memory fill, so a pass that contains the sort can be captured in a HIP
// graph.  8-bit digits (four passes for the 30-bit ray keys), three
// kernels per pass:
//   k_sort_hist     per-tile digit counts, digit-major hist[d][tile]
//   k_sort_scan     one block per digit: exclusive scan of its row over the
//                   tiles, and the digit's total
//   k_sort_scatter  digit bases (scan of the totals + the tile's row
//                   prefix), stable ranks within the tile, the tile sorted
//                   in LDS, then written out run by run (coalesced)
// A tile is 4096 items, 1024 per wave in rounds of 64 consecutive items,
// so (wave, round, lane) is the items' order: a lane's rank among the
// earlier items with its digit is the wave's running count for the digit
// plus its rank among the round's lanes with that digit (eight ballots).
// Queues shorter than `min_n` are not sorted: each pass's scatter kernel
// then copies them through unchanged, so whatever the number of passes
// the caller's permutation reaches the result buffers as it was.
#pragma once

#include "rx_device.h"

namespace rx {

constexpr int SORT_BS = 256;
constexpr int SORT_WAVES = SORT_BS / 64;
constexpr int SORT_ROUNDS = 16;
constexpr int SORT_TILE = SORT_BS * SORT_ROUNDS;
constexpr int SORT_BITS = 8;
constexpr int SORT_RADIX = 1 << SORT_BITS;
constexpr int SORT_DPT = SORT_RADIX / SORT_BS;  // digits per thread in the per-block digit loops
static_assert(SORT_RADIX % SORT_BS == 0, "digits spread evenly over the block's threads");

inline int sort_tiles(long long cap) { return (int)((cap + SORT_TILE - 1) / SORT_TILE); }
inline int sort_passes(unsigned key_bits) {
    const int p = (int)((key_bits + SORT_BITS - 1) / SORT_BITS);
    return p < 1 ? 1 : p;
}
// Scratch of a sort of up to `cap` items: hist (SORT_RADIX x tiles) and
// the digit totals, in u32.
inline size_t sort_scratch_words(long long cap) { return (size_t)SORT_RADIX * sort_tiles(cap) + SORT_RADIX; }

__device__ __forceinline__ int sort_len(const int32_t* __restrict__ count, long long cap) {
    const int n = *count;
    return n < 0 ? 0 : (n > cap ? (int)cap : n);
}

// Lanes of the wave whose digit equals this lane's (among `valid` lanes).
__device__ __forceinline__ unsigned long long digit_peers(uint32_t d, bool valid) {
    unsigned long long m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < SORT_BITS; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long on = __ballot(bit);
        m &= bit ? on : ~on;
    }
    return m;
}

__device__ __forceinline__ int lane_rank(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Exclusive scan of one u32 per thread over a SORT_BS block; *total gets
// the sum.  Every thread calls it (uses __syncthreads).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t x, uint32_t* wave_sums, uint32_t* total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t incl = x;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wave_sums[wv] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < SORT_WAVES; ++k) {
        const uint32_t s = wave_sums[k];
        before += k < wv ? s : 0u;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + incl - x;
}

__global__ __launch_bounds__(SORT_BS) void k_sort_hist(const uint32_t* __restrict__ keys,
                                                       const int32_t* __restrict__ count, long long cap, int min_n,
                                                       int shift, uint32_t* __restrict__ hist, int tiles_cap) {
    __shared__ uint32_t h[SORT_RADIX];
    const int n = sort_len(count, cap);
    if (n < min_n) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // the grid is sized for the capacity: each block takes tiles
    // blockIdx.x, blockIdx.x + gridDim.x, ... up to the length
    for (int tile = blockIdx.x; (long long)tile * SORT_TILE < n; tile += gridDim.x) {
        const long long t0 = (long long)tile * SORT_TILE;
#pragma unroll
        for (int q = 0; q < SORT_DPT; ++q) h[q * SORT_BS + threadIdx.x] = 0u;
        __syncthreads();
        for (int r = 0; r < SORT_ROUNDS; ++r) {
            const long long i = t0 + (long long)wv * (64 * SORT_ROUNDS) + r * 64 + lane;
            const bool valid = i < n;
            const uint32_t d = valid ? (keys[i] >> shift) & (SORT_RADIX - 1) : 0u;
            const unsigned long long m = digit_peers(d, valid);
            if (valid && lane_rank(m) == 0) atomicAdd(&h[d], (uint32_t)__popcll(m));
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < SORT_DPT; ++q) {
            const int d = q * SORT_BS + threadIdx.x;
            hist[(size_t)d * tiles_cap + tile] = h[d];
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(SORT_BS) void k_sort_scan(const int32_t* __restrict__ count, long long cap, int min_n,
                                                       uint32_t* __restrict__ hist, int tiles_cap,
                                                       uint32_t* __restrict__ total) {
    __shared__ uint32_t wave_sums[SORT_WAVES];
    const int n = sort_len(count, cap);
    if (n < min_n) return;
    const int nt = (n + SORT_TILE - 1) / SORT_TILE;
    uint32_t* row = hist + (size_t)blockIdx.x * tiles_cap;
    const int per = (nt + SORT_BS - 1) / SORT_BS;
    const int b = threadIdx.x * per, e = min(nt, b + per);
    uint32_t s = 0;
    for (int k = b; k < e; ++k) s += row[k];
    uint32_t all;
    uint32_t run = block_exclusive_scan(s, wave_sums, &all);
    for (int k = b; k < e; ++k) {
        const uint32_t c = row[k];
        row[k] = run;
        run += c;
    }
    if (threadIdx.x == 0) total[blockIdx.x] = all;
}

__global__ __launch_bounds__(SORT_BS) void k_sort_scatter(const uint32_t* __restrict__ key_in,
                                                          const uint32_t* __restrict__ val_in,
                                                          uint32_t* __restrict__ key_out,
                                                          uint32_t* __restrict__ val_out,
                                                          const int32_t* __restrict__ count, long long cap, int min_n,
                                                          int shift, const uint32_t* __restrict__ hist, int tiles_cap,
                                                          const uint32_t* __restrict__ total) {
    __shared__ uint32_t wave_sums[SORT_WAVES];
    __shared__ uint32_t wcnt[SORT_WAVES][SORT_RADIX];
    __shared__ uint32_t gbase[SORT_RADIX];     // output position of the tile's first item of each digit
    __shared__ uint32_t lbase[SORT_RADIX];     // its position in the tile's sorted order
    __shared__ uint32_t skey[SORT_TILE], sval[SORT_TILE];
    const int n = sort_len(count, cap);
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    if (n < min_n) {
        // not sorted: carried through unchanged
        for (long long i = (long long)blockIdx.x * SORT_BS + t; i < n; i += (long long)gridDim.x * SORT_BS) {
            key_out[i] = key_in[i];
            val_out[i] = val_in[i];
        }
        return;
    }
    // the grid is sized for the capacity: each block takes tiles
    // blockIdx.x, blockIdx.x + gridDim.x, ... up to the length
    for (int tile = blockIdx.x; (long long)tile * SORT_TILE < n; tile += gridDim.x) {
        const long long t0 = (long long)tile * SORT_TILE;
        // this tile's first output position for each digit: thread t holds
        // digits t*DPT .. t*DPT + DPT - 1 (consecutive, so one scan of the
        // per-thread sums orders them all)
        {
            uint32_t tot[SORT_DPT], part = 0;
#pragma unroll
            for (int q = 0; q < SORT_DPT; ++q) {
                tot[q] = total[t * SORT_DPT + q];
                part += tot[q];
            }
            uint32_t all;
            uint32_t before = block_exclusive_scan(part, wave_sums, &all);
#pragma unroll
            for (int q = 0; q < SORT_DPT; ++q) {
                const int d = t * SORT_DPT + q;
                gbase[d] = before + hist[(size_t)d * tiles_cap + tile];
                before += tot[q];
#pragma unroll
                for (int w = 0; w < SORT_WAVES; ++w) wcnt[w][d] = 0u;
            }
        }
        __syncthreads();
        uint32_t k[SORT_ROUNDS], v[SORT_ROUNDS], off[SORT_ROUNDS];
        const long long w0 = t0 + (long long)wv * (64 * SORT_ROUNDS) + lane;
#pragma unroll
        for (int r = 0; r < SORT_ROUNDS; ++r) {
            const long long i = w0 + r * 64;
            const bool valid = i < n;
            k[r] = valid ? key_in[i] : 0u;
            v[r] = valid ? val_in[i] : 0u;
        }
#pragma unroll
        for (int r = 0; r < SORT_ROUNDS; ++r) {
            const bool valid = w0 + r * 64 < n;
            const uint32_t d = (k[r] >> shift) & (SORT_RADIX - 1);
            const unsigned long long m = digit_peers(d, valid);
            const int rank = lane_rank(m);
            // all lanes of one digit read the wave's running count before its
            // first lane (rank 0) moves it on: one wave, LDS in program order
            const uint32_t run = wcnt[wv][d];
            if (valid && rank == 0) wcnt[wv][d] = run + (uint32_t)__popcll(m);
            off[r] = run + (uint32_t)rank;
        }
        __syncthreads();
        // position of each wave's first item of each digit in the tile's
        // sorted order: the digit's start in the tile plus the earlier waves'
        {
            uint32_t cnt[SORT_DPT], part = 0;
#pragma unroll
            for (int q = 0; q < SORT_DPT; ++q) {
                const int d = t * SORT_DPT + q;
                cnt[q] = 0;
#pragma unroll
                for (int w = 0; w < SORT_WAVES; ++w) cnt[q] += wcnt[w][d];
                part += cnt[q];
            }
            uint32_t all;
            uint32_t before = block_exclusive_scan(part, wave_sums, &all);
#pragma unroll
            for (int q = 0; q < SORT_DPT; ++q) {
                const int d = t * SORT_DPT + q;
                lbase[d] = before;
                uint32_t s = before;
#pragma unroll
                for (int w = 0; w < SORT_WAVES; ++w) {
                    const uint32_t c = wcnt[w][d];
                    wcnt[w][d] = s;
                    s += c;
                }
                before += cnt[q];
            }
        }
        __syncthreads();
        // the tile, sorted, into LDS
#pragma unroll
        for (int r = 0; r < SORT_ROUNDS; ++r) {
            if (w0 + r * 64 < n) {
                const uint32_t pos = wcnt[wv][(k[r] >> shift) & (SORT_RADIX - 1)] + off[r];
                if (pos < (uint32_t)SORT_TILE) {
                    skey[pos] = k[r];
                    sval[pos] = v[r];
                }
            }
        }
        __syncthreads();
        // and out: consecutive threads write consecutive items of a digit's run
        const int m_tile = (int)min((long long)SORT_TILE, n - t0);
        for (int j = t; j < m_tile; j += SORT_BS) {
            const uint32_t key = skey[j];
            const uint32_t d = (key >> shift) & (SORT_RADIX - 1);
            const uint32_t dest = gbase[d] + ((uint32_t)j - lbase[d]);
            // always < n when the digit counts are consistent; the test keeps
            // a slip from ever writing past the buffers
            if (dest < (uint32_t)n) {
                key_out[dest] = key;
                val_out[dest] = sval[j];
            }
        }
        __syncthreads();
    }
}

// Copies the first *count pairs (the end of a sort with an odd number of
// passes, for callers that want the result in the first buffers).
__global__ void k_sort_copy(const uint32_t* __restrict__ ks, const uint32_t* __restrict__ vs,
                            uint32_t* __restrict__ kd, uint32_t* __restrict__ vd, const int32_t* __restrict__ count,
                            long long cap) {
    const int n = sort_len(count, cap);
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}

// Enqueues the sort of the first *count (at most cap) pairs of (keys,
// vals) by their low key_bits; (keys2, vals2) are the second buffers.  The
// sorted pairs end in (keys, vals) after an even number of passes and in
// (keys2, vals2) after an odd one: *res_keys / *res_vals say which.
// scratch holds sort_scratch_words(cap) u32.
inline hipError_t enqueue_ray_sort(uint32_t* keys, uint32_t* vals, uint32_t* keys2, uint32_t* vals2,
                                   const int32_t* count, long long cap, int min_n, unsigned key_bits,
                                   uint32_t* scratch, hipStream_t s, uint32_t** res_keys, uint32_t** res_vals) {
    const int tiles = sort_tiles(cap);
    // one block per tile of the capacity (blocks past the length return at
    // once): measured faster than fewer blocks looping over the tiles
    const int grid = tiles;
    uint32_t* hist = scratch;
    uint32_t* tot = scratch + (size_t)SORT_RADIX * tiles;
    const int passes = sort_passes(key_bits);
    uint32_t *ki = keys, *vi = vals, *ko = keys2, *vo = vals2;
    for (int p = 0; p < passes; ++p) {
        const int shift = SORT_BITS * p;
        k_sort_hist<<<grid, SORT_BS, 0, s>>>(ki, count, cap, min_n, shift, hist, tiles);
        k_sort_scan<<<SORT_RADIX, SORT_BS, 0, s>>>(count, cap, min_n, hist, tiles, tot);
        k_sort_scatter<<<grid, SORT_BS, 0, s>>>(ki, vi, ko, vo, count, cap, min_n, shift, hist, tiles, tot);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        uint32_t* tk = ki;
        ki = ko;
        ko = tk;
        uint32_t* tv = vi;
        vi = vo;
        vo = tv;
    }
    if (res_keys) *res_keys = ki;
    if (res_vals) *res_vals = vi;
    return hipSuccess;
}

}  // namespace rx
