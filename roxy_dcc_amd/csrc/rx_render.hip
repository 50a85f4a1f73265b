// rx_render.hip — wavefront path tracer reproducing render_kernel's
// estimator (rendering/taichi/trace.py:139-283), for one scene or a batch
// of scenes.
//
// A pass traces, for every job of the batch, ns = min(samples_left,
// samples_per_pass) samples of each of its pixels: job j owns the path
// slots [path_offset, path_offset + W*H*samples_per_pass), slot
// path_offset + s*W*H + pixel for its sample s.  Per bounce:
//   extend   closest hit for every queued ray (SoA queue in HBM)
//   shade    miss/emission/material, next-event request, BSDF sample,
//            Russian roulette; survivors and next-event requests are
//            appended to the next queues with one atomic per wave
//            (ballot + popcount)
//   connect  sphere-light sampling and any-hit shadow rays for each
//            next-event request, added to the path's colour
// then accumulate folds each job's samples into its running means in
// sample order, and advance moves every job's sample window.  All
// per-pass state lives on the device, so a pass can be captured in a HIP
// graph and replayed.  Randomness is Philox4x32-10 keyed by the job's seed
// and counted by (pixel, sample, bounce, draw): a pixel's samples do not
// depend on the pass size, the queue order or the other jobs.
#include <stdlib.h>

#include <mutex>
#include <new>

#include "rx_launch.h"
#include "rx_sort.h"

namespace {

using namespace rx;

constexpr int BS = 256;
constexpr int MAX_JOBS = 4096;

// Atomics on one address from every wave of a launch serialise (about
// 11.5 ns each on MI355X, where device-scope atomics are resolved past the
// per-XCD L2s); spread over 64 addresses they cost well under 1 ns.  So
// counts that every wave adds to go to NPART partial slots, picked by
// blockIdx % NPART, and are folded into their destination once per pass.
constexpr int NPART = 64;

// The same holds for handing out queue slots: a persistent kernel's waves
// take rays from FETCH_SEGS counters, one per contiguous segment of the
// queue, instead of one, so refills do not queue behind one address.
#ifndef RX_FETCH_SEGS
#define RX_FETCH_SEGS 32
#endif
constexpr int FETCH_SEGS = RX_FETCH_SEGS;
constexpr int FETCH_STRIDE = 32;  // ints: 128 bytes between counters

// Adaptive-stop error sums: |change of the running mean| per sample index,
// summed over pixels and channels in fixed point (2^22 units per colour
// unit, each term capped at 2^14) so the sum, and the stop decision, do not
// depend on the order of the atomics; EPART partial slots per job.
constexpr int EPART = 16;

// Camera rays test only the analytic primitives their pixel's 8x8 tile can
// see (k_prim_tiles): up to PRIM_TILE_MAX primitives, PRIM_TILE_WORDS
// 64-bit masks per tile.
constexpr int PRIM_TILE_WORDS = 4;
constexpr int PRIM_TILE_MAX = 64 * PRIM_TILE_WORDS;
constexpr double ERR_SCALE = 4194304.0;
constexpr float ERR_CAP = 16384.0f;

// Path state and queues, SoA over the workspace.
struct ws_view {
    rx_job* jobs;
    int32_t* path_off;   // n_jobs + 1
    int32_t* pix_off;    // n_jobs + 1
    int n_jobs;
    long long P;         // path slots
    long long NP;        // pixel slots
    // per path
    float *thr_x, *thr_y, *thr_z, *col_x, *col_y, *col_z, *ior;
    float *fn_x, *fn_y, *fn_z, *fa_x, *fa_y, *fa_z, *fd;
    int32_t *p_job, *p_rays;
    // rays are counted per path (p_rays, summed by k_accumulate over the
    // samples a job keeps) when a pass holds an adaptive job, whose stop
    // drops samples; otherwise the traversal kernels add their counts to
    // the jobs' ray partials directly (lane_rays; set per pass)
    int ray_per_path;
    // 1: the primary queue's slots hold no ray records (camera rays are
    // traced pixel-major, and the camera-ray k_shade recomputes each slot's
    // ray from its pixel and sample; set per pass)
    int cam_rays;
    // 1 for the pixel-major camera-ray launch of k_extend: its position i
    // is a pixel and sample (pm_camera_ray), whose ray the kernel computes
    // instead of reading a record (set per launch)
    int cam_gen;
    // two ray queues: 32 B per slot, {origin, path index bits} {direction, 0}
    gv4f* q_ray[2];
    // hits of the current queue: 16 B per slot, {t, coded id bits, u, v}
    gv4f* h_rec;
    // next-event requests
    // next-event requests, 64 B each: {hit point, path bits}, {normal,
    // scale}, {albedo, 0}, {throughput before the bounce, 0}
    gv4f* e_rec;
    // queue lengths: [0], [1] ray queues, [2] next-event queue
    int32_t* counts;
    // hand-out counters of the persistent kernels: FETCH_SEGS per kernel
    // (closest-hit, then shadow), FETCH_STRIDE ints apart
    int32_t* fetch_seg;
    // per-job queue offsets of the primary rays (n_jobs + 1)
    int32_t* q_off;
    // partial sums of the traversal counters and of each job's rays, one
    // slot per NPART blocks (blockIdx % NPART), folded once per pass
    unsigned long long* ctr_part;   // NPART x 16 (an rx_counters per 128 bytes)
    unsigned long long* ray_part;   // NPART x n_jobs
    // samples of this pass each job accumulates (k_decide), and the
    // adaptive-stop error sums: EPART x n_jobs x err_stride, fixed point
    int32_t* ns_eff;
    unsigned long long* err_part;
    int err_stride;
    // traversal stack spill area (entries past the LDS part)
    int2* spill;
    // reordering of the ray queues for k_extend: the sort keys and slot
    // indices (double buffers) and the radix sort's digit counts
    // (rx_sort.h); hit_slot (set per launch) maps a position of the traced
    // order back to the queue slot its hit belongs to, ray_slot (if set) to
    // the slot its ray is read from (null: position i is slot i)
    uint32_t *sk_in, *sk_out, *si_in, *si_out;
    uint32_t* sort_scratch;
    const uint32_t* hit_slot;
    const uint32_t* ray_slot;
    // order in which k_connect takes the next-event requests (set per
    // launch; null: request i is position i)
    const uint32_t* nee_slot;
    // top of the tree for the traversal kernels' LDS (k_treelet): slot k's
    // child boxes at tl_box[3 k ..], its links at tl_link[k], the slot
    // count at tl_count[0]; built for the BVH whose nodes are tl_nodes
    // (set per pass by the host; null: no treelet)
    gv4f* tl_box;
    gv2i* tl_link;
    int32_t* tl_count;
    const rx_node* tl_nodes;
    // eye-relative copies of the (single) job's scene for its camera rays
    // (k_eye_prep; null when the workspace has no room for them or the pass
    // holds several jobs): node boxes minus the eye, triangles as {e - v0,
    // e1, e2, cross(e - v0, e1)}, the root record's box minus the eye, and a
    // 32-byte record per primitive (eye_prim)
    rx_node* eye_nodes;
    rx_tri* eye_tris;
    rx_bvh_header* eye_header;
    gv4f* eye_prims;
    // with the copies, for 1..PRIM_TILE_MAX analytic primitives: per 8x8
    // pixel tile of the camera (tile_cols per row), the primitives a camera
    // ray through the tile can hit, as PRIM_TILE_WORDS 64-bit masks
    // (k_prim_tiles); null otherwise
    uint64_t* tile_prims;
    int tile_cols;
    // link-free copy of the (single) job's BVH (imp_usable; null when not
    // used): slot h of heap order (root 0, children 2h+1 and 2h+2) holds,
    // for an interior node, its children's boxes as the first three of its
    // 64-byte record's four 16-byte words (the fourth unused), a leaf
    // child's box marked on its x axis (imp_mark); imp_nodes in
    // world space, eye_imp minus the camera position (with the eye-relative
    // copies); for a leaf, its packed (first << 8 | count) in imp_leaf.
    // imp_records = 2^depth - 1 interior slots.
    gv4f* imp_nodes;
    gv4f* eye_imp;
    int32_t* imp_leaf;
    int imp_records;
};

void carve(rx_arena& ar, int n_jobs, long long P, bool deep, int err_stride, ws_view& w, bool sort = true) {
    w.jobs = ar.take<rx_job>(n_jobs);
    w.spill = ar.take<int2>(render_spill_entries(deep));
    w.tl_box = ar.take<gv4f>(3 * TL_NODES);
    w.tl_link = ar.take<gv2i>(TL_NODES);
    w.tl_count = ar.take<int32_t>(4);
    w.tl_nodes = nullptr;
    w.eye_nodes = nullptr;
    w.eye_tris = nullptr;
    w.eye_header = nullptr;
    w.eye_prims = nullptr;
    w.tile_prims = nullptr;
    w.tile_cols = 0;
    w.imp_nodes = w.eye_imp = nullptr;
    w.imp_leaf = nullptr;
    w.imp_records = 0;
    w.path_off = ar.take<int32_t>(n_jobs + 1);
    w.pix_off = ar.take<int32_t>(n_jobs + 1);
    w.counts = ar.take<int32_t>(8);
    w.fetch_seg = ar.take<int32_t>(2 * FETCH_SEGS * FETCH_STRIDE);
    w.q_off = ar.take<int32_t>(n_jobs + 1);
    w.ctr_part = ar.take<unsigned long long>(NPART * 16);
    w.ray_part = ar.take<unsigned long long>((size_t)NPART * n_jobs);
    w.ns_eff = ar.take<int32_t>(n_jobs);
    w.err_stride = err_stride;
    w.err_part = ar.take<unsigned long long>((size_t)EPART * n_jobs * err_stride);
    float** pf[] = {&w.thr_x, &w.thr_y, &w.thr_z, &w.col_x, &w.col_y, &w.col_z, &w.ior,
                    &w.fn_x,  &w.fn_y,  &w.fn_z,  &w.fa_x,  &w.fa_y,  &w.fa_z,  &w.fd};
    for (float** f : pf) *f = ar.take<float>(P);
    w.p_job = ar.take<int32_t>(P);
    w.p_rays = ar.take<int32_t>(P);
    w.ray_per_path = 1;
    w.cam_rays = 0;
    w.cam_gen = 0;
    for (int b = 0; b < 2; ++b) w.q_ray[b] = ar.take<gv4f>(2 * P);
    w.h_rec = ar.take<gv4f>(P);
    w.e_rec = ar.take<gv4f>(4 * P);
    w.hit_slot = nullptr;
    w.ray_slot = nullptr;
    w.nee_slot = nullptr;
    w.sk_in = w.sk_out = w.si_in = w.si_out = nullptr;
    w.sort_scratch = nullptr;
    if (sort) {
        w.sk_in = ar.take<uint32_t>(P);
        w.sk_out = ar.take<uint32_t>(P);
        w.si_in = ar.take<uint32_t>(P);
        w.si_out = ar.take<uint32_t>(P);
        w.sort_scratch = ar.take<uint32_t>(sort_scratch_words(P));
    }
    w.n_jobs = n_jobs;
    w.P = P;
}

// Job owning slot `x` of a prefix-offset table (largest j with off[j] <= x).
__device__ __forceinline__ int owner(const int32_t* __restrict__ off, int n, long long x) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if ((long long)off[mid] <= x)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Set bits of the wave mask m in the lanes below this one (v_mbcnt).
__device__ __forceinline__ int lanes_below(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ int job_of_path(const ws_view& w, int p) { return w.n_jobs == 1 ? 0 : w.p_job[p]; }

// Ray queue records: two 16-byte words per slot, {origin, path index
// bits} and {direction, tag bits}, so a ray is one pair of wide loads or
// stores; the tag of a camera ray is its pixel's 8x8 tile (camera_tile),
// 0 otherwise.
__device__ __forceinline__ void put_ray(const ws_view& w, int b, int slot, int path, f3 o, f3 d, int tag = 0) {
    RX_GLOBAL gv4f* q = (RX_GLOBAL gv4f*)w.q_ray[b] + 2 * (long long)slot;
    q[0] = gv4f{o.x, o.y, o.z, __int_as_float(path)};
    q[1] = gv4f{d.x, d.y, d.z, __int_as_float(tag)};
}

__device__ __forceinline__ int get_ray(const ws_view& w, int b, int slot, f3& o, f3& d) {
    const RX_GLOBAL gv4f* q = (const RX_GLOBAL gv4f*)w.q_ray[b] + 2 * (long long)slot;
    gv4f a = q[0], c = q[1];
    o = mk(a.x, a.y, a.z);
    d = mk(c.x, c.y, c.z);
    return __float_as_int(a.w);
}

// get_ray() that also returns the record's tag
__device__ __forceinline__ int get_ray_tag(const ws_view& w, int b, int slot, f3& o, f3& d, int& tag) {
    const RX_GLOBAL gv4f* q = (const RX_GLOBAL gv4f*)w.q_ray[b] + 2 * (long long)slot;
    gv4f a = q[0], c = q[1];
    o = mk(a.x, a.y, a.z);
    d = mk(c.x, c.y, c.z);
    tag = __float_as_int(c.w);
    return __float_as_int(a.w);
}

__device__ __forceinline__ int get_ray_path(const ws_view& w, int b, int slot) {
    return __float_as_int(((const RX_GLOBAL float*)w.q_ray[b])[8 * (long long)slot + 3]);
}

// Hit records: {t, id, u, v} with the kind folded into the id: a BVH
// triangle id >= 0, primitive k as -(k + 2), a miss as -1.
__device__ __forceinline__ int code_hit(int kind, int id) {
    return kind == RX_KIND_BVH ? id : (kind == RX_KIND_PRIM ? -(id + 2) : -1);
}

__device__ __forceinline__ void put_hit(const ws_view& w, int i, float t, int kind, int id, float u, float v) {
    ((RX_GLOBAL gv4f*)w.h_rec)[i] = gv4f{t, __int_as_float(code_hit(kind, id)), u, v};
}

__device__ __forceinline__ void get_hit(const ws_view& w, int i, float& t, int& kind, int& id, float& u, float& v) {
    gv4f h = ((const RX_GLOBAL gv4f*)w.h_rec)[i];
    int c = __float_as_int(h.y);
    t = h.x;
    u = h.z;
    v = h.w;
    kind = c >= 0 ? RX_KIND_BVH : (c == -1 ? RX_KIND_MISS : RX_KIND_PRIM);
    id = c >= -1 ? c : -(c + 2);
}

__device__ __forceinline__ int pass_samples(const rx_job& j) {
    int spp = j.params.samples_per_pass;
    return j.samples_left < spp ? j.samples_left : spp;
}

// Pixel of slot k (0 <= k < W*H) of a sample's path slots, and back.  When
// both image sides are multiples of 8, slots run through 8x8 pixel tiles
// in row-major tile order, so the 64 rays of a wave leave from one
// compact block of pixels and share node fetches; otherwise slots are
// scanline order.  Either way a bijection on [0, W*H): sampling is keyed
// by the pixel, so images do not depend on the slot order.
__device__ __forceinline__ bool tiled(const rx_camera& c) { return ((c.width | c.height) & 7) == 0; }

// The 8x8 tile of pixel (px, py), row-major over the image's tiles.
__device__ __forceinline__ int camera_tile(const rx_camera& c, int px, int py) {
    return (py >> 3) * ((c.width + 7) >> 3) + (px >> 3);
}

__device__ __forceinline__ int slot_pixel(const rx_camera& c, int k) {
    if (!tiled(c)) return k;
    int tile = k >> 6, in = k & 63, tiles_x = c.width >> 3;
    int tx = tile % tiles_x, ty = tile / tiles_x;
    return (ty * 8 + (in >> 3)) * c.width + tx * 8 + (in & 7);
}

__device__ __forceinline__ u4 rng(const rx_render_params& prm, int pix, int frame, int depth, uint32_t draw) {
    return philox4x32_10(u4{(uint32_t)pix, (uint32_t)frame, (uint32_t)depth, draw}, prm.seed_lo, prm.seed_hi);
}

// Appends to a queue: every lane of the wave calls this with `want`; the
// wave reserves popcount(ballot) slots with one atomic and each wanting
// lane gets base + its rank among the wanting lanes.
__device__ __forceinline__ int wave_append(int32_t* counter, bool want) {
    unsigned long long m = __ballot(want);
    int lane = threadIdx.x & 63;
    int total = __popcll(m);
    int base = 0;
    if (lane == 0 && total > 0) base = atomicAdd(counter, total);
    base = __shfl(base, 0, 64);
    return base + lanes_below(m);
}

// End of the interior phase of a traversal step.  Lanes walk interior
// nodes until they hold a leaf; the wave then intersects its lanes'
// leaves together.  Waiting for the last lane to find its leaf issues
// wide node fetches for a handful of lanes, and the vector-memory return
// path is paid per instruction, so the phase ends once no more than
// INNER_EXIT_NUM / INNER_EXIT_DEN of the lanes that entered it are still
// walking (after at least one step, so every lane progresses); those
// lanes carry on in the next step.  A lane's own sequence of pops and
// tests is unchanged.  Measured on MI355X (70k- and 2M-triangle scenes,
// Mrays/s): running to the last lane 817 / 353, 2/3 827 / 378, 1/2
// 874 / 395, 3/8 885 / 399, 1/3 880 / 393, 1/4 880 / 391, 1/8 855 / 372.
#ifndef INNER_EXIT_NUM
#define INNER_EXIT_NUM 3
#endif
#ifndef INNER_EXIT_DEN
#define INNER_EXIT_DEN 8
#endif
__device__ __forceinline__ bool inner_done(int n_in, int entered, int it) {
    if (n_in == 0) return true;
    return it > 0 && n_in * INNER_EXIT_DEN <= entered * INNER_EXIT_NUM;
}

// Rays a wave of a persistent kernel keeps in flight at once: all 64 lanes
// when the queue has work for every resident wave, otherwise an even share
// (n / waves, at least 1).  When few rays are left (the later bounces),
// handing 64 of them to each of the first waves would leave most waves
// idle and every SIMD waiting on one or two waves' dependent fetches;
// spreading them keeps the same rays in flight with less divergence per
// wave and more waves to hide latency.  It changes which lane runs a ray,
// never the ray's own sequence of fetches and tests.
__device__ __forceinline__ int wave_cap(int n, int block_size) {
    long long waves = (long long)gridDim.x * (block_size / 64);
    long long c = (n + waves - 1) / waves;
    return c >= 64 ? 64 : (c < 1 ? 1 : (int)c);
}

// Hands queue slots to idle lanes: the wave takes up to (cap - busy lanes)
// consecutive slots with one atomic, in lane order.  The queue [0, n) is
// cut into FETCH_SEGS equal segments with a counter each; a wave starts at
// its home segment and moves to the next one when that runs dry, so
// every slot is handed out once and the waves' atomics spread over
// FETCH_SEGS addresses.  A full-width wave still waits until refill_min
// (default REFILL_MIN) lanes are free (or all are) before it fetches.
// Returns this lane's
// slot (>= n when it gets none) and sets `exhausted` once every segment
// has run dry.  seg / tried: the wave's current segment and how many it
// has found empty (wave-uniform).
#ifndef RX_REFILL_MIN
#define RX_REFILL_MIN 48
#endif
constexpr int REFILL_MIN = RX_REFILL_MIN;
#ifndef RX_XCD_SEGMENTS
#define RX_XCD_SEGMENTS 1
#endif
__device__ __forceinline__ int home_segment(int block_size) {
    const int wpb = block_size / 64, wave = threadIdx.x >> 6;
    if (RX_XCD_SEGMENTS && FETCH_SEGS % 8 == 0) {
        // blocks go to the 8 XCDs round-robin (block b on XCD b % 8): the
        // waves of one XCD start in its own eighth of the segments, so each
        // XCD's L2 holds the nodes of one stretch of the (sorted) queue
        constexpr int per = FETCH_SEGS / 8;
        const long long k = (long long)(blockIdx.x >> 3) * wpb + wave;
        return (blockIdx.x & 7) * per + (int)(k % per);
    }
    return (int)(((long long)blockIdx.x * wpb + wave) % FETCH_SEGS);
}

__device__ __forceinline__ int wave_fetch(int32_t* ctr, int n, int cap, bool idle, bool& exhausted, int& seg,
                                          int& tried, int refill_min = REFILL_MIN) {
    unsigned long long m = __ballot(idle);
    int lane = threadIdx.x & 63;
    int n_idle = __popcll(m);
    int want = cap - (64 - n_idle);
    if (want > n_idle) want = n_idle;
    if (want <= 0) return n;
    if (cap >= 32 && want < refill_min && n_idle < 64) return n;
    if (n <= 0) tried = FETCH_SEGS;
    int rank = lanes_below(m);
    while (tried < FETCH_SEGS) {
        const long long s0 = (long long)n * seg / FETCH_SEGS, s1 = (long long)n * (seg + 1) / FETCH_SEGS;
        int base = 0;
        if (lane == 0) base = atomicAdd(&ctr[seg * FETCH_STRIDE], want);
        base = __shfl(base, 0, 64);
        const long long avail = s1 - s0 - base;
        if (avail >= want) {
            if (avail == want) {
                seg = seg + 1 == FETCH_SEGS ? 0 : seg + 1;
                ++tried;
            }
            return (idle && rank < want) ? (int)(s0 + base + rank) : n;
        }
        seg = seg + 1 == FETCH_SEGS ? 0 : seg + 1;
        ++tried;
        if (avail > 0) return (idle && rank < (int)avail) ? (int)(s0 + base + rank) : n;
    }
    exhausted = true;
    return n;
}

// Appends to two queues with one atomic per block and queue: every thread
// of the block calls this at the same point (uniform control flow), with
// whether it wants a slot in queue a and in queue b.  Slots are handed out
// in thread order within the block.  One pair of barriers serves both
// queues, and the two atomics leave in one instruction (lanes 0 and 1 of
// wave 0), so their round trips overlap.
template <int SB = BS>
__device__ __forceinline__ void block_append_pair(int32_t* ca, bool wa, int32_t* cb, bool wb,
                                                  int32_t* lds /* 2 (SB / 64 + 1) */, int& ra, int& rb) {
    constexpr int NW = SB / 64;
    unsigned long long ma = __ballot(wa), mb = __ballot(wb);
    int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        lds[wv] = __popcll(ma);
        lds[NW + 1 + wv] = __popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        int32_t* l = lds + threadIdx.x * (NW + 1);
        int32_t tot = 0;
        for (int k = 0; k < NW; ++k) {
            int32_t c = l[k];
            l[k] = tot;
            tot += c;
        }
        l[NW] = tot > 0 ? atomicAdd(threadIdx.x == 0 ? ca : cb, tot) : 0;
    }
    __syncthreads();
    ra = lds[NW] + lds[wv] + lanes_below(ma);
    rb = lds[2 * NW + 1] + lds[NW + 1 + wv] + lanes_below(mb);
    __syncthreads();
}

// ---------------------------------------------------------------- pass control
__global__ void k_set_job(rx_job* jobs, rx_job j, int32_t* path_off, int32_t* pix_off, int32_t paths, int32_t pix) {
    jobs[0] = j;
    path_off[0] = 0;
    path_off[1] = paths;
    pix_off[0] = 0;
    pix_off[1] = pix;
}

// Start of a pass (one block of PB threads): clears the queue lengths and
// the counter partials, and lays out the primary-ray queue.  Job j's
// active paths this pass are its first pass_samples * npix path slots, so
// they get queue slots q_off[j] + local with q_off an exclusive prefix sum
// over jobs: compaction without atomics, in path order.
constexpr int PB = 256;
__global__ __launch_bounds__(PB) void k_pass_begin(ws_view w) {
    __shared__ int32_t part[PB];
    const int t = threadIdx.x;
    if (t < 8) w.counts[t] = 0;
    for (int k = t; k < 2 * FETCH_SEGS; k += PB) w.fetch_seg[k * FETCH_STRIDE] = 0;
    for (int k = t; k < NPART * 16; k += PB) w.ctr_part[k] = 0ull;
    const int per = (w.n_jobs + PB - 1) / PB;
    const int j0 = t * per, j1 = min(w.n_jobs, j0 + per);
    int32_t sum = 0;
    for (int j = j0; j < j1; ++j) {
        const rx_job& J = w.jobs[j];
        sum += pass_samples(J) * J.camera.width * J.camera.height;
    }
    part[t] = sum;
    __syncthreads();
    if (t == 0) {
        int32_t run = 0;
        for (int k = 0; k < PB; ++k) {
            int32_t c = part[k];
            part[k] = run;
            run += c;
        }
        w.q_off[w.n_jobs] = run;
        w.counts[0] = run;
    }
    __syncthreads();
    int32_t run = part[t];
    for (int j = j0; j < j1; ++j) {
        w.q_off[j] = run;
        const rx_job& J = w.jobs[j];
        run += pass_samples(J) * J.camera.width * J.camera.height;
    }
}

// Clears the per-job ray partials (once, before a job list's first pass;
// k_advance clears them after every fold).
__global__ void k_clear_ray_part(ws_view w) {
    long long n = (long long)NPART * w.n_jobs;
    long long ne = (long long)EPART * w.n_jobs * w.err_stride;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n || k < ne;
         k += (long long)gridDim.x * blockDim.x) {
        if (k < n) w.ray_part[k] = 0ull;
        if (k < ne) w.err_part[k] = 0ull;
    }
}

// End of a pass: moves every job's sample window, folds the job's ray
// partials into its count, and (block 0) folds the traversal counter
// partials into the caller's counters.
__global__ void k_advance(ws_view w, rx_counters* __restrict__ ctr) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < w.n_jobs) {
        rx_job& J = w.jobs[j];
        int ns = w.ns_eff[j];
        J.first_sample += ns;
        J.samples_left = J.adaptive_stopped ? 0 : J.samples_left - ns;
        unsigned long long r = 0;
        for (int k = 0; k < NPART; ++k) {
            r += w.ray_part[(size_t)k * w.n_jobs + j];
            w.ray_part[(size_t)k * w.n_jobs + j] = 0ull;
        }
        J.rays += r;
    }
    if (blockIdx.x == 0 && threadIdx.x < 8 && ctr) {
        unsigned long long sum = 0;
        for (int k = 0; k < NPART; ++k) sum += w.ctr_part[k * 16 + threadIdx.x];
        if (sum) atomicAdd(reinterpret_cast<unsigned long long*>(ctr) + threadIdx.x, sum);
    }
}

// This block's partial slot of the traversal counters.
__device__ __forceinline__ rx_counters* ctr_slot(const ws_view& w) {
    return reinterpret_cast<rx_counters*>(w.ctr_part + (blockIdx.x & (NPART - 1)) * 16);
}

// A lane's rays not yet added to its job's ray partials (passes without
// per-path counts): added when the lane's next ray belongs to another job
// (queues run job by job, so that is rare) and at the kernel's end, with
// one atomic per wave when the wave's lanes share the job.
struct lane_rays {
    int job = 0;
    unsigned long long n = 0;
    __device__ __forceinline__ unsigned long long* slot(const ws_view& w, int j) const {
        return w.ray_part + (size_t)(blockIdx.x & (NPART - 1)) * w.n_jobs + j;
    }
    __device__ __forceinline__ void add(const ws_view& w, int j, unsigned long long k) {
        if (j != job) {
            if (n) atomicAdd(slot(w, job), n);
            job = j;
            n = 0;
        }
        n += k;
    }
    // every lane of the wave
    __device__ __forceinline__ void flush(const ws_view& w) {
        const int j0 = __builtin_amdgcn_readfirstlane(job);
        if (__ballot(n != 0 && job != j0) == 0ull)
            wave_add(slot(w, j0), n);
        else if (n)
            atomicAdd(slot(w, job), n);
        n = 0;
    }
};

// ---------------------------------------------------------------- generate
__global__ __launch_bounds__(BS) void k_generate(ws_view w) {
    const int stride = gridDim.x * BS;
    for (long long base = (long long)xcd_block() * BS; base < w.P; base += stride) {
        long long p = base + threadIdx.x;
        bool active = false;
        int jx = 0, tag = 0;
        f3 org = mk(0.0f, 0.0f, 0.0f), rd = org;
        if (p < w.P) {
            jx = w.n_jobs == 1 ? 0 : owner(w.path_off, w.n_jobs, p);
            const rx_job& J = w.jobs[jx];
            int npix = J.camera.width * J.camera.height;
            long long local = p - w.path_off[jx];
            int s = (int)(local / npix), pix = slot_pixel(J.camera, (int)(local % npix));
            if (s < pass_samples(J)) {
                active = true;
                rd = camera_dir(J.camera, pix % J.camera.width, pix / J.camera.width, J.first_sample + s);
                org = ld3(J.camera.pos);
                tag = camera_tile(J.camera, pix % J.camera.width, pix / J.camera.width);
                // throughput, colour and IOR start as 1, 0 and 1: the
                // camera-ray k_shade takes them as constants and writes them,
                // and the first-hit depth (0 for a miss).  A job of depth 0
                // is never shaded: its samples stay black with no first hit.
                if (J.params.max_depth <= 0) {
                    w.col_x[p] = 0.0f;
                    w.col_y[p] = 0.0f;
                    w.col_z[p] = 0.0f;
                    w.fd[p] = 0.0f;
                }
                if (w.n_jobs > 1) w.p_job[p] = jx;  // job_of_path reads it only then
                if (w.ray_per_path) w.p_rays[p] = 0;
            }
        }
        if (active && !w.cam_rays) {
            int slot = w.q_off[jx] + (int)(p - w.path_off[jx]);
            put_ray(w, 0, slot, (int)p, org, rd, tag);
        }
    }
}

// Camera ray at position i of a pixel-major camera-ray launch: within
// each job, position q_off[j] + k * ns + s is sample s of pixel slot k (ns
// samples of the job this pass), so a wave traces the samples of one or a
// few pixels — nearly identical rays — instead of 64 neighbouring pixels
// of one sample.  The ray is k_generate's for that pixel and sample (same
// camera_dir inputs).  Returns its path; sets the pixel's tile tag and the
// (sample-major) queue slot of its hit, where k_shade reads it.  ONE: the
// pass holds one job, whose record `jobs` (read-only, unaliased) then
// gives through scalar loads.
template <bool ONE>
__device__ __forceinline__ int pm_camera_ray(const ws_view& w, const rx_job* __restrict__ jobs, int i, f3& ro, f3& rd,
                                             int& tag, int& hit) {
    const int jx = ONE ? 0 : (w.n_jobs == 1 ? 0 : owner(w.q_off, w.n_jobs, i));
    const rx_job& J = jobs[jx];
    const int npix = J.camera.width * J.camera.height;
    const int ns = pass_samples(J);
    const int li = i - w.q_off[jx];
    const int k = li / ns, s = li - k * ns;
    const int local = s * npix + k;
    const int pix = slot_pixel(J.camera, k);
    const int px = pix % J.camera.width, py = pix / J.camera.width;
    rd = camera_dir(J.camera, px, py, J.first_sample + s);
    ro = ld3(J.camera.pos);
    tag = camera_tile(J.camera, px, py);
    hit = w.q_off[jx] + local;
    return w.path_off[jx] + local;
}

// Closest analytic primitive along a ray: the linear scan of
// rendering/taichi/intersect.py:259-287 (ascending index, strict
// 0.001 < t < closest, so the lowest index wins a tie).
template <bool COUNT>
__device__ __forceinline__ void closest_prim(const rx_scene& sc, f3 ro, f3 rd, float& prim_t, int& prim_id,
                                             trav_count& cnt) {
    prim_t = 1e9f;
    prim_id = -1;
    const int np = sc.n_prims;
    const rx_prim* prims = sc.prims;
    for (int k = 0; k < np; ++k) {
        float t = prim_hit(gload(prims + k), ro, rd);
        if (COUNT) cnt.prims++;
        if (t > 0.001f && t < prim_t) {
            prim_t = t;
            prim_id = k;
        }
    }
}

// The same scan when every active lane's ray belongs to job j (wave-uniform):
// the job's primitive records are read with scalar loads, once per wave.
template <bool COUNT>
__device__ __forceinline__ void closest_prim_job(const rx_job* jobs, int j, f3 ro, f3 rd, float& prim_t, int& prim_id,
                                                 trav_count& cnt) {
    const RX_CONST rx_job* J = (const RX_CONST rx_job*)(jobs + j);
    const int np = J->scene.n_prims;
    const rx_prim* prims = J->scene.prims;
    prim_t = 1e9f;
    prim_id = -1;
    for (int k = 0; k < np; ++k) {
        float t = prim_hit(cload(prims + k), ro, rd);
        if (COUNT) cnt.prims++;
        if (t > 0.001f && t < prim_t) {
            prim_t = t;
            prim_id = k;
        }
    }
}

// closest_prim for the rays of the active lanes, with scalar loads when they
// all belong to one job.
template <bool COUNT>
__device__ __forceinline__ void closest_prim_lanes(const ws_view& w, int jx, f3 ro, f3 rd, float& prim_t,
                                                   int& prim_id, trav_count& cnt) {
    int j0 = __builtin_amdgcn_readfirstlane(jx);
    if (__ballot(jx != j0) == 0ull)
        closest_prim_job<COUNT>(w.jobs, j0, ro, rd, prim_t, prim_id, cnt);
    else
        closest_prim<COUNT>(w.jobs[jx].scene, ro, rd, prim_t, prim_id, cnt);
}

// Any primitive occluding (0.001 < t < max_t) for the active lanes' shadow
// rays, in index order with the first occluder ending the scan
// (intersect.py:226-241); scalar loads when the lanes share a job.
template <bool COUNT>
__device__ __forceinline__ bool any_prim_lanes(const ws_view& w, int jx, f3 so, f3 ld, float max_t,
                                               trav_count& cnt) {
    bool occ = false;
    int j0 = __builtin_amdgcn_readfirstlane(jx);
    if (__ballot(jx != j0) == 0ull) {
        const RX_CONST rx_job* J = (const RX_CONST rx_job*)(w.jobs + j0);
        const int np = J->scene.n_prims;
        const rx_prim* prims = J->scene.prims;
        for (int k = 0; k < np && !occ; ++k) {
            float t = prim_hit(cload(prims + k), so, ld);
            if (COUNT) cnt.prims++;
            if (0.001f < t && t < max_t) occ = true;
        }
    } else {
        const rx_scene& sc = w.jobs[jx].scene;
        const int np = sc.n_prims;
        for (int k = 0; k < np && !occ; ++k) {
            float t = prim_hit(gload(sc.prims + k), so, ld);
            if (COUNT) cnt.prims++;
            if (0.001f < t && t < max_t) occ = true;
        }
    }
    return occ;
}

// Closest-primitive pre-pass over ray queue `qi`: writes each ray's
// primitive hit into the hit records, where k_extend<.., PRE = true> picks it
// up.  The BVH walk never reads the primitive result (it starts from
// closest = 1e9 and is compared with it only at the end, intersect.py:292),
// so computing it first and separately leaves every hit unchanged.
template <bool COUNT>
__global__ __launch_bounds__(BS) void k_prim_scan(ws_view w, int qi) {
    const int n = w.counts[qi];
    trav_count cnt{0, 0, 0};
    const int stride = gridDim.x * BS;
    for (int base = xcd_block() * BS; base < n; base += stride) {
        int i = base + threadIdx.x;
        if (i < n) {
            f3 ro, rd;
            int p = get_ray(w, qi, i, ro, rd);
            float t;
            int id;
            closest_prim_lanes<COUNT>(w, job_of_path(w, p), ro, rd, t, id, cnt);
            // the primitive result alone: {prim_t, prim_id}
            ((RX_GLOBAL gv4f*)w.h_rec)[i] = gv4f{t, __int_as_float(id), 0.0f, 0.0f};
        }
    }
    if (COUNT) wave_add(&ctr_slot(w)->prim_tests, cnt.prims);
}

// ---------------------------------------------------------------- eye-relative copies
// The camera rays of a single-job render all start at the camera position
// e.  k_eye_prep writes, once per call, what their tests compute first from
// e alone: node boxes minus e, triangles as {e - v0, e1, e2, cross(e - v0,
// e1)}, the root box minus e and the primitives' eye-relative records
// (eye_prim); the camera-ray launch of k_extend reads these instead
// (slab_rel, tri_hit_rel, prim_hit_rel: the remaining operations of the
// same tests, so the same bits).
__global__ __launch_bounds__(BS) void k_eye_prep(ws_view w) {
    const rx_job& J = w.jobs[0];
    const rx_scene& sc = J.scene;
    const f3 e = ld3(J.camera.pos);
    const int nn = sc.n_tris > 0 && w.eye_nodes ? sc.n_nodes : 0;
    long long n = nn > sc.n_tris ? nn : sc.n_tris;
    if (sc.n_prims > n) n = sc.n_prims;
    if (n < 1) n = 1;
    for (long long k = (long long)blockIdx.x * BS + threadIdx.x; k < n; k += (long long)gridDim.x * BS) {
        if (k < nn) {
            const RX_GLOBAL gv4f* src = (const RX_GLOBAL gv4f*)(sc.nodes + k);
            RX_GLOBAL gv4f* dst = (RX_GLOBAL gv4f*)(w.eye_nodes + k);
            dst[0] = src[0] - eye_planes(0, e);
            dst[1] = src[1] - eye_planes(1, e);
            dst[2] = src[2] - eye_planes(2, e);
            dst[3] = src[3];  // the links
        }
        if (k < sc.n_tris) {
            f3 v0, e1, e2;
            load_tri(sc.tris, (int)k, v0, e1, e2);
            const f3 sv = e - v0;         // tri_hit: s = ro - v0
            const f3 q = cross(sv, e1);   // tri_hit: q = cross(s, e1)
            RX_GLOBAL gv4f* dst = (RX_GLOBAL gv4f*)(w.eye_tris + k);
            dst[0] = gv4f{sv.x, sv.y, sv.z, e1.x};
            dst[1] = gv4f{e1.y, e1.z, e2.x, e2.y};
            dst[2] = gv4f{e2.z, q.x, q.y, q.z};
        }
        if (k < sc.n_prims) {
            gv4f r0, r1;
            eye_prim(gload(sc.prims + k), e, r0, r1);
            ((RX_GLOBAL gv4f*)w.eye_prims)[2 * k] = r0;
            ((RX_GLOBAL gv4f*)w.eye_prims)[2 * k + 1] = r1;
        }
        if (k == 0 && sc.n_tris > 0) {
            rx_bvh_header h = gload(sc.header);
            const float ec[3] = {e.x, e.y, e.z};
            for (int a = 0; a < 3; ++a) {
                h.bmin[a] = h.bmin[a] - ec[a];
                h.bmax[a] = h.bmax[a] - ec[a];
            }
            *w.eye_header = h;
        }
    }
}

// The link-free tree copy (imp_nodes, eye_imp, imp_leaf).  A leaf child's
// box (lo, hi) is marked on its x axis: lo.x and hi.x swapped when they
// differ, lo.x a quiet NaN when they are equal.  The slab tests take
// fminf / fmaxf of the distances to the two x planes, the same pair either
// way round; with one of them NaN, fminf and fmaxf return the other, the
// distance to both planes when they are equal.  So a marked box gives
// every slab test the bits the box itself gives.  A box of finite
// coordinates (RX_SCENE_FINITE_BOXES) has lo.x <= hi.x, so !(lo.x <= hi.x)
// reads the mark, and no child link is stored.
constexpr uint32_t IMP_LEAF_NAN = 0x7fc0f1a9u;
constexpr int IMP_MAX_DEPTH = 24;

__device__ __forceinline__ void imp_mark(float& lo, float& hi) {
    if (lo < hi) {
        const float t = lo;
        lo = hi;
        hi = t;
    } else {
        lo = __uint_as_float(IMP_LEAF_NAN);
    }
}

// Links of interior slot h's children (~slot for a leaf), from the marks of
// its record's child boxes.
__device__ __forceinline__ gv2i imp_links(gv4f a, gv4f b, gv4f c, int h) {
    const int l = 2 * h + 1, r = l + 1;
    return gv2i{a.x <= a.w ? l : ~l, b.z <= c.y ? r : ~r};
}

// Writes the three record words of a node's child boxes (left lo, hi;
// right lo, hi) at slot h, marking the leaf children.
__device__ __forceinline__ void imp_put(RX_GLOBAL gv4f* dst, long long h, float b[12], bool lleaf, bool rleaf) {
    if (lleaf) imp_mark(b[0], b[3]);
    if (rleaf) imp_mark(b[6], b[9]);
    dst[4 * h] = gv4f{b[0], b[1], b[2], b[3]};
    dst[4 * h + 1] = gv4f{b[4], b[5], b[6], b[7]};
    dst[4 * h + 2] = gv4f{b[8], b[9], b[10], b[11]};
}

// The link-free copy of job 0's BVH: one thread per heap slot h walks from
// the root along the bits of h + 1 below its leading one (0: left child,
// 1: right) through the node records' links; a slot below a leaf stays
// unused.  An interior node's slot gets its record's child boxes (and,
// with the eye-relative copies, those boxes minus the camera position,
// each marked after the subtraction), a leaf's slot its packed triangle
// range.
__global__ __launch_bounds__(BS) void k_imp_build(ws_view w) {
    const rx_job& J = w.jobs[0];
    const rx_scene& sc = J.scene;
    const f3 e = ld3(J.camera.pos);
    const long long slots = 2LL * w.imp_records + 1;
    for (long long h = (long long)blockIdx.x * BS + threadIdx.x; h < slots; h += (long long)gridDim.x * BS) {
        const unsigned hp = (unsigned)(h + 1);
        const int depth = 31 - __clz((int)hp);
        int link = 0;
        bool used = true;
        for (int k = depth - 1; k >= 0; --k) {
            if (link < 0) {
                used = false;
                break;
            }
            link = sc.nodes[link].link[(hp >> k) & 1u];
        }
        if (!used) continue;
        if (link < 0) {
            w.imp_leaf[h] = ~link;
            continue;
        }
        if (h >= w.imp_records) continue;  // deeper than the header's depth: never read
        const RX_GLOBAL gv4f* src = (const RX_GLOBAL gv4f*)(sc.nodes + link);
        const gv4f p0 = src[0], p1 = src[1], p2 = src[2];
        const gv2i ln = ((const RX_GLOBAL gv2i*)src)[6];
        float b[12] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w, p2.x, p2.y, p2.z, p2.w};
        if (w.eye_imp) {
            const float ec[3] = {e.x, e.y, e.z};
            float r[12];
            for (int q = 0; q < 12; ++q) r[q] = b[q] - ec[q % 3];
            imp_put((RX_GLOBAL gv4f*)w.eye_imp, h, r, ln.x < 0, ln.y < 0);
        }
        imp_put((RX_GLOBAL gv4f*)w.imp_nodes, h, b, ln.x < 0, ln.y < 0);
    }
}

// closest_prim() of a camera ray of job 0 from the eye-relative records
// (scalar loads: every lane reads the same record).
// With per-tile masks (tile_prims), only the primitives of the union of
// the active lanes' tiles' masks are tested, in ascending index order: a
// primitive outside a tile's mask misses every camera ray through the tile
// (k_prim_tiles), so the closest hit and its tie-break are the full scan's.
template <bool COUNT>
__device__ __forceinline__ void closest_prim_eye(const ws_view& w, f3 ro, f3 rd, int tile, float& prim_t,
                                                 int& prim_id, trav_count& cnt) {
    const RX_CONST rx_job* J = (const RX_CONST rx_job*)(w.jobs);
    const int np = J->scene.n_prims;
    const rx_prim* prims = J->scene.prims;
    const RX_CONST gv4f* ep = (const RX_CONST gv4f*)w.eye_prims;
    prim_t = 1e9f;
    prim_id = -1;
    auto test = [&](int k) {
        const gv4f r0 = ep[2 * k], r1 = ep[2 * k + 1];
        const float t = __float_as_int(r1.w) == EYE_PRIM_NONE ? prim_hit(cload(prims + k), ro, rd)
                                                              : prim_hit_rel(r0, r1, rd);
        if (COUNT) cnt.prims++;
        if (t > 0.001f && t < prim_t) {
            prim_t = t;
            prim_id = k;
        }
    };
    if (w.tile_prims) {
        // union over the distinct tiles of the active lanes (one or two: a
        // refill takes consecutive slots of a pixel-major queue)
        uint64_t m[PRIM_TILE_WORDS];
        for (int q = 0; q < PRIM_TILE_WORDS; ++q) m[q] = 0ull;
        bool pend = true;
        for (;;) {
            const unsigned long long b = __ballot(pend);
            if (b == 0ull) break;
            const int t = __builtin_amdgcn_readlane(tile, __builtin_ctzll(b));
            const RX_CONST uint64_t* tm = (const RX_CONST uint64_t*)w.tile_prims + PRIM_TILE_WORDS * (long long)t;
            for (int q = 0; q < PRIM_TILE_WORDS; ++q) m[q] |= tm[q];
            pend = pend && tile != t;
        }
        for (int q = 0; q < PRIM_TILE_WORDS; ++q) {
            for (uint64_t bits = m[q]; bits != 0ull; bits &= bits - 1ull) test(64 * q + __builtin_ctzll(bits));
        }
        return;
    }
    for (int k = 0; k < np; ++k) test(k);
}

// Per 8x8 tile of the (single) job's camera: the analytic primitives some
// camera ray through the tile may hit.  The tile's rays leave the eye e
// in directions camera_dir gives for pixel coordinates in [x0, x1] x
// [y0, y1] (the jitter stays in [0, 1]): the pyramid spanned by its four
// corner directions.  A sphere whose centre lies outside one of the
// pyramid's side planes by more than its radius plus a margin, or a cube
// whose corners all lie outside one by more than the margin, misses every
// such ray, and its bit stays clear; the same for the plane through e
// normal to the pyramid's mean direction (every ray of the tile has a
// positive component along it), which the side planes alone do not give
// for what lies behind the eye.  The margin, 1e-3 of the distance from
// e plus 1e-3, is orders of magnitude above the rounding of camera_dir's
// fp32 directions (~1e-7 relative) and of the fp32 sphere and slab tests at
// that distance, so a cleared primitive's fp32 test reports no hit past
// 0.001 for any of those rays.  Planes and triangle slots keep their bits.
// Computed in fp64 from the fp32 camera and primitive records.
__global__ __launch_bounds__(BS) void k_prim_tiles(ws_view w) {
    const rx_job& J = w.jobs[0];
    const rx_camera& cam = J.camera;
    const rx_scene& sc = J.scene;
    const int cols = w.tile_cols, rows = (cam.height + 7) >> 3;
    const int np = sc.n_prims;
    const double ex = cam.pos[0], ey = cam.pos[1], ez = cam.pos[2];
    for (int t = blockIdx.x * BS + threadIdx.x; t < cols * rows; t += gridDim.x * BS) {
        const int x0 = (t % cols) * 8, y0 = (t / cols) * 8;
        const int x1 = min(x0 + 8, cam.width), y1 = min(y0 + 8, cam.height);
        double dir[4][3];
        const int cx[4] = {x0, x1, x1, x0}, cy[4] = {y0, y0, y1, y1};
        for (int c = 0; c < 4; ++c) {
            const double u = (double)cx[c] / (double)cam.width * 2.0 - 1.0;
            const double v = 1.0 - (double)cy[c] / (double)cam.height * 2.0;
            const double a = u * (double)cam.aspect * (double)cam.fov_tan, b = v * (double)cam.fov_tan;
            for (int k = 0; k < 3; ++k)
                dir[c][k] = (double)cam.fwd[k] + (double)cam.right[k] * a + (double)cam.up[k] * b;
        }
        double mid[3];
        for (int k = 0; k < 3; ++k) mid[k] = dir[0][k] + dir[1][k] + dir[2][k] + dir[3][k];
        // inward unit normals of the four side planes (through e)
        double nrm[5][3];
        bool ok[5];
        for (int c = 0; c < 4; ++c) {
            const double* p = dir[c];
            const double* q = dir[(c + 1) & 3];
            double n[3] = {p[1] * q[2] - p[2] * q[1], p[2] * q[0] - p[0] * q[2], p[0] * q[1] - p[1] * q[0]};
            const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            ok[c] = len > 0.0;
            const double sg = (n[0] * mid[0] + n[1] * mid[1] + n[2] * mid[2]) < 0.0 ? -1.0 : 1.0;
            for (int k = 0; k < 3; ++k) nrm[c][k] = ok[c] ? sg * n[k] / len : 0.0;
        }
        // the fifth plane: through e, normal to the mean direction, used
        // when every corner direction points to its positive side
        {
            const double len = sqrt(mid[0] * mid[0] + mid[1] * mid[1] + mid[2] * mid[2]);
            bool front = len > 0.0;
            for (int c = 0; c < 4 && front; ++c)
                front = dir[c][0] * mid[0] + dir[c][1] * mid[1] + dir[c][2] * mid[2] > 0.0;
            ok[4] = front;
            for (int k = 0; k < 3; ++k) nrm[4][k] = front ? mid[k] / len : 0.0;
        }
        uint64_t m[PRIM_TILE_WORDS];
        for (int q = 0; q < PRIM_TILE_WORDS; ++q) m[q] = 0ull;
        for (int k = 0; k < np && k < PRIM_TILE_MAX; ++k) {
            const rx_prim pr = gload(sc.prims + k);
            bool keep = true;
            if (pr.type == RX_PRIM_SPHERE) {
                const double c[3] = {(double)pr.a[0] - ex, (double)pr.a[1] - ey, (double)pr.a[2] - ez};
                const double r = fabs((double)pr.a[3]);
                const double margin = 1e-3 * (sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) + r) + 1e-3;
                for (int f = 0; f < 5 && keep; ++f)
                    if (ok[f] && nrm[f][0] * c[0] + nrm[f][1] * c[1] + nrm[f][2] * c[2] < -(r + margin)) keep = false;
            } else if (pr.type == RX_PRIM_CUBE) {
                double lo[3], hi[3], far2 = 0.0;
                for (int a = 0; a < 3; ++a) {
                    const double l = (double)pr.a[a] - fabs((double)pr.b[a]), h = (double)pr.a[a] + fabs((double)pr.b[a]);
                    const double e = a == 0 ? ex : (a == 1 ? ey : ez);
                    lo[a] = l - e;
                    hi[a] = h - e;
                    const double fa = fmax(fabs(lo[a]), fabs(hi[a]));
                    far2 += fa * fa;
                }
                const double margin = 1e-3 * sqrt(far2) + 1e-3;
                for (int f = 0; f < 5 && keep; ++f) {
                    if (!ok[f]) continue;
                    double most = 0.0;  // the corner furthest along the normal
                    for (int a = 0; a < 3; ++a) most += fmax(nrm[f][a] * lo[a], nrm[f][a] * hi[a]);
                    if (most < -margin) keep = false;
                }
            }
            if (keep) m[k >> 6] |= 1ull << (k & 63);
        }
        for (int q = 0; q < PRIM_TILE_WORDS; ++q) w.tile_prims[(size_t)PRIM_TILE_WORDS * t + q] = m[q];
    }
}

// ---------------------------------------------------------------- treelet
// The top of the pass's BVH for the traversal kernels' LDS: the first
// TL_NODES interior nodes in breadth-first order, root in slot 0, each as
// its two child boxes and two links, where the link of an interior child
// that also has a slot becomes TL_REF + slot.  Level by level: one thread
// per node of the level; the level's interior children take the next free
// slots in level order (exclusive scan over the level's child counts).
// Links to leaves and to interior nodes past the last slot stay as they
// are.  One block; a level never holds more than TL_NODES < TLB nodes.
constexpr int TLB = 512;
static_assert(TL_NODES < TLB, "a treelet level must fit one block");
__global__ __launch_bounds__(TLB) void k_treelet(const rx_node* __restrict__ nodes,
                                                 const rx_bvh_header* __restrict__ header, ws_view w) {
    __shared__ int src[TL_NODES];
    __shared__ int scan[TLB];
    const int tid = threadIdx.x;
    const int root = gload(header).root_link;
    if (tid == 0) src[0] = root;
    __syncthreads();
    int lo = 0, hi = root >= 0 ? 1 : 0;
    while (lo < hi) {
        const int s = lo + tid;
        const bool mine = s < hi;
        gv4f a{0.0f, 0.0f, 0.0f, 0.0f}, b = a, c = a;
        gv2i d{-1, -1};
        if (mine) {
            const RX_GLOBAL gv4f* p = (const RX_GLOBAL gv4f*)(nodes + src[s]);
            a = p[0];
            b = p[1];
            c = p[2];
            d = ((const RX_GLOBAL gv2i*)p)[6];
        }
        const int kids = mine ? (d.x >= 0 ? 1 : 0) + (d.y >= 0 ? 1 : 0) : 0;
        scan[tid] = kids;
        __syncthreads();
        for (int off = 1; off < TLB; off <<= 1) {
            int v = tid >= off ? scan[tid - off] : 0;
            __syncthreads();
            scan[tid] += v;
            __syncthreads();
        }
        const int total = scan[TLB - 1];
        if (mine) {
            int k = hi + scan[tid] - kids;
            gv2i e = d;
            if (d.x >= 0) {
                if (k < TL_NODES) {
                    src[k] = d.x;
                    e.x = TL_REF + k;
                }
                ++k;
            }
            if (d.y >= 0 && k < TL_NODES) {
                src[k] = d.y;
                e.y = TL_REF + k;
            }
            RX_GLOBAL gv4f* o = (RX_GLOBAL gv4f*)w.tl_box + 3 * s;
            o[0] = a;
            o[1] = b;
            o[2] = c;
            ((RX_GLOBAL gv2i*)w.tl_link)[s] = e;
        }
        __syncthreads();
        lo = hi;
        hi = hi + total < TL_NODES ? hi + total : TL_NODES;
    }
    if (tid == 0) w.tl_count[0] = hi;
}

// A block's LDS copy of the treelet (TLN = 0: none).  Returns the slot
// count, 0 when the pass has no treelet.
// EYE: the boxes minus the eye of job 0 (the camera-ray traversal's
// eye-relative form, as the node copy k_eye_prep writes).
// IMP: the first TLN slots of the link-free copy (heap order: the top of
// the tree), boxes only; slot k of the copy is slot k in LDS.
template <int SBS, int TLN, bool EYE = false, bool IMP = false>
__device__ __forceinline__ int load_treelet(const ws_view& w, gv4f* box, gv2i* link) {
    static_assert(TLN == 0 || TLN == TL_NODES, "the treelet's links address TL_NODES slots");
    if (TLN == 0) return 0;
    if (IMP) {
        const int n = w.imp_records < TLN ? w.imp_records : TLN;
        const RX_GLOBAL gv4f* src = (const RX_GLOBAL gv4f*)(EYE ? w.eye_imp : w.imp_nodes);
        for (int k = threadIdx.x; k < 3 * n; k += SBS) box[k] = src[4 * (k / 3) + k % 3];
        __syncthreads();
        return n;
    }
    if (w.tl_nodes == nullptr) return 0;
    const int n = w.tl_count[0];
    const f3 e = EYE ? ld3(w.jobs[0].camera.pos) : mk(0.0f, 0.0f, 0.0f);
    for (int k = threadIdx.x; k < 3 * n; k += SBS) {
        gv4f v = ((const RX_GLOBAL gv4f*)w.tl_box)[k];
        if (EYE) v = v - eye_planes(k % 3, e);
        box[k] = v;
    }
    for (int k = threadIdx.x; k < n; k += SBS) link[k] = ((const RX_GLOBAL gv2i*)w.tl_link)[k];
    __syncthreads();
    return n;
}

// ---------------------------------------------------------------- extend
// Persistent closest-hit kernel with per-lane ray refill.  A fixed grid of
// waves runs until the queue is drained; whenever lanes of a wave have
// finished their rays, the wave takes as many new queue slots as it has
// idle lanes with one atomic (wave_fetch over the first FETCH_SEGS
// counters, zeroed before each extend launch)
// and hands them out in lane order.  Between refills each busy lane runs
// its traversal to the next leaf and intersects that leaf, so lanes are
// never held idle for a whole ray by the slowest ray of the wave.  Each
// ray's own sequence of pops, tests and pushes is the one of
// bvh_closest (rendering/taichi/intersect.py:81-152), so hits are
// unchanged.
// EYE: the camera-ray launch of a single-job render, on the eye-relative
// copies (k_eye_prep); its job record (`jobs`, w.jobs passed as a
// read-only, unaliased pointer) is then read with scalar loads.
// IMP: the pass traces the link-free copy of its (single) job's tree
// (imp_nodes, or eye_imp for the camera rays): node links are heap slots,
// a leaf's triangle range comes from imp_leaf; the same boxes, tests and
// visit order.
template <int SBS, int TLN, bool COUNT, bool PRE, bool EYE, bool IMP>
__global__ __launch_bounds__(SBS) __attribute__((amdgpu_waves_per_eu(EXTEND_WAVES_PER_EU))) void k_extend(
    ws_view w, const rx_job* __restrict__ jobs, int qi) {
    // the link-free copy's treelet keeps boxes only (48 of 56 bytes per
    // slot), which leaves room for one more LDS stack entry per lane
    constexpr int SL = TLN > 0 ? STACK_LDS_TL + (IMP ? 1 : 0) : STACK_LDS;
    __shared__ int2 stack_mem[SL * SBS];
    __shared__ gv4f tl_box_mem[TLN > 0 ? 3 * TLN : 1];
    __shared__ gv2i tl_link_mem[TLN > 0 && !IMP ? TLN : 1];
    lane_stack<SBS, SL> stk{stack_mem + threadIdx.x, w.spill + blockIdx.x * SBS + threadIdx.x, (int)gridDim.x * SBS};
    const int tl_n = load_treelet<SBS, TLN, EYE, IMP>(w, tl_box_mem, tl_link_mem);
    const RX_LDS gv4f* tl_box = (const RX_LDS gv4f*)tl_box_mem;
    const RX_LDS gv2i* tl_link = (const RX_LDS gv2i*)tl_link_mem;
    const RX_GLOBAL gv4f* imp = (const RX_GLOBAL gv4f*)(EYE ? w.eye_imp : w.imp_nodes);
    const int n = w.counts[qi];
    int32_t* fetch = w.fetch_seg;
    int seg = home_segment(SBS), tried = 0;
    const int cap = wave_cap(n, SBS);
    // the queue this bounce's shade fills and the next-event queue start empty
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        w.counts[1 - qi] = 0;
        w.counts[2] = 0;
    }
    const int lane = threadIdx.x & 63;
    // pixel-major camera rays refill whole waves: a wave then holds whole
    // pixels whose samples walk the tree together (2M-triangle scene,
    // camera-ray launch 6.18 ms refilling at 48 free lanes, 6.09 at 56,
    // 5.99 at 64; 6.52 at 32)
    const int refill_min = w.cam_gen ? 64 : REFILL_MIN;
    trav_count cnt{0, 0, 0};
    unsigned long long rays = 0;
    lane_rays jr;
    bool exhausted = false;
    int i = -1;  // queue slot of this lane's hit, -1 when idle
    int p = 0;
    const rx_scene* sc = nullptr;
    const rx_node* nodes = nullptr;
    const rx_tri* tris = nullptr;
    f3 ro = mk(0.0f, 0.0f, 0.0f), rd = ro, inv = ro;
    float prim_t = 1e9f, best_t = 1e9f, best_u = 0.0f, best_v = 0.0f;
    int prim_id = -1, best_i = -1, sp = 0, cur = 0;
    bool have = false;
    while (true) {
        bool idle = i < 0;
        int mine = exhausted ? n : wave_fetch(fetch, n, cap, idle, exhausted, seg, tried, refill_min);
        {
            if (idle && mine < n) {
                int tag = 0;
                if (w.cam_gen) {
                    p = pm_camera_ray<EYE>(w, jobs, mine, ro, rd, tag, i);
                } else {
                    const int rs = w.ray_slot ? (int)w.ray_slot[mine] : mine;
                    p = EYE ? get_ray_tag(w, qi, rs, ro, rd, tag) : get_ray(w, qi, rs, ro, rd);
                    i = w.hit_slot ? (int)w.hit_slot[mine] : mine;
                }
                // camera-ray launches and passes over the link-free copy
                // are single-job renders: job 0, its record in scalar loads
                const int jx = EYE || IMP ? 0 : job_of_path(w, p);
                sc = &jobs[jx].scene;
                rays++;
                if (!w.ray_per_path) jr.add(w, jx, 1);
                // primitive scan first (intersect.py:259-287), or its
                // result from k_prim_scan
                if (PRE) {
                    gv4f h = ((const RX_GLOBAL gv4f*)w.h_rec)[i];
                    prim_t = h.x;
                    prim_id = __float_as_int(h.y);
                } else if (EYE) {
                    closest_prim_eye<COUNT>(w, ro, rd, tag, prim_t, prim_id, cnt);
                } else {
                    closest_prim_lanes<COUNT>(w, jx, ro, rd, prim_t, prim_id, cnt);
                }
                best_t = 1e9f;
                best_i = -1;
                best_u = 0.0f;
                best_v = 0.0f;
                have = false;
                if (sc->n_tris > 0) {
                    nodes = EYE ? w.eye_nodes : sc->nodes;
                    tris = EYE ? w.eye_tris : sc->tris;
                    const rx_bvh_header hdr = EYE || IMP ? cload(EYE ? w.eye_header : sc->header)
                                                         : gload(sc->header);
                    inv = inv_dir(rd);
                    float rt;
                    const bool root = EYE ? slab_rel(inv, hdr.bmin[0], hdr.bmin[1], hdr.bmin[2], hdr.bmax[0],
                                                     hdr.bmax[1], hdr.bmax[2], best_t, &rt)
                                          : slab(ro, inv, hdr.bmin[0], hdr.bmin[1], hdr.bmin[2], hdr.bmax[0],
                                                 hdr.bmax[1], hdr.bmax[2], best_t, &rt);
                    if (root) {
                        have = true;
                        cur = IMP ? 0 : (TLN > 0 && tl_n > 0 && sc->nodes == w.tl_nodes ? TL_REF : hdr.root_link);
                        sp = 0;
                    }
                }
            }
        }
        if (__ballot(i >= 0) == 0ull) break;
        if (i >= 0) {
            if (have) {
                const int entered = __popcll(__ballot(true));
                for (int it = 0;; ++it) {
                    bool in = have && cur >= 0;
                    int n_in = __popcll(__ballot(in));
                    if (inner_done(n_in, entered, it)) break;
                    if (in) {
                        if (COUNT) cnt.pops++;
                        gv4f a, b, c;
                        gv2i dl;
                        if (IMP) {
                            if (TLN > 0 && cur < tl_n) {
                                a = tl_box[3 * cur];
                                b = tl_box[3 * cur + 1];
                                c = tl_box[3 * cur + 2];
                            } else {
                                const RX_GLOBAL gv4f* q = imp + 4 * (long long)cur;
                                a = q[0];
                                b = q[1];
                                c = q[2];
                            }
                            dl = imp_links(a, b, c, cur);
                        } else if (TLN > 0 && cur >= TL_REF) {
                            const int k = cur - TL_REF;
                            a = tl_box[3 * k];
                            b = tl_box[3 * k + 1];
                            c = tl_box[3 * k + 2];
                            dl = tl_link[k];
                        } else {
                            const RX_GLOBAL gv4f* q = (const RX_GLOBAL gv4f*)(nodes + cur);
                            a = q[0];
                            b = q[1];
                            c = q[2];
                            dl = ((const RX_GLOBAL gv2i*)q)[6];
                        }
                        const bool kids = EYE ? visit_boxes_rel(a, b, c, dl, inv, best_t, stk, sp, cur)
                                              : visit_boxes(a, b, c, dl, ro, inv, best_t, stk, sp, cur);
                        if (!kids) have = pop_next(stk, sp, best_t, cur);
                    }
                }
                if (have && cur < 0) {
                    if (COUNT) cnt.pops++;
                    int lw = IMP ? w.imp_leaf[~cur] : ~cur;
                    int start = lw >> 8, end = start + (lw & 255);
                    for (int t2 = start; t2 < end; ++t2) {
                        float u, v, t;
                        if (EYE) {
                            f3 sv, e1, e2, q;
                            load_tri_rel(tris, t2, sv, e1, e2, q);
                            t = tri_hit_rel(rd, sv, e1, e2, q, &u, &v);
                        } else {
                            f3 v0, e1, e2;
                            load_tri(tris, t2, v0, e1, e2);
                            t = tri_hit(ro, rd, v0, e1, e2, &u, &v);
                        }
                        if (COUNT) cnt.tris++;
                        if (0.001f < t && t < best_t) {
                            best_t = t;
                            best_i = t2;
                            best_u = u;
                            best_v = v;
                        }
                    }
                    have = pop_next(stk, sp, best_t, cur);
                }
            }
            if (!have) {
                // intersect.py:292: the BVH wins only if strictly closer
                float t = prim_t;
                int id = prim_id, kind = prim_id >= 0 ? RX_KIND_PRIM : RX_KIND_MISS;
                float u = 0.0f, v = 0.0f;
                if (0.001f < best_t && best_t < prim_t) {
                    t = best_t;
                    id = best_i;
                    kind = RX_KIND_BVH;
                    u = best_u;
                    v = best_v;
                }
                put_hit(w, i, t, kind, id, u, v);
                if (w.ray_per_path) w.p_rays[p] += 1;
                i = -1;
            }
        }
    }
    rx_counters* ctr = ctr_slot(w);
    wave_add(&ctr->rays, rays);
    if (!w.ray_per_path) jr.flush(w);
    if (COUNT) {
        wave_add(&ctr->node_pops, cnt.pops);
        wave_add(&ctr->tri_tests, cnt.tris);
        wave_add(&ctr->prim_tests, cnt.prims);
    }
}

// ---------------------------------------------------------------- shade helpers
// rendering/taichi/scatter.py:14-30 (cosine_hemisphere) and :33-38
__device__ __forceinline__ f3 diffuse_dir(f3 n, float r1, float r2) {
    float sp, cp;
    sincos_2pi(r1, &sp, &cp);
    float r = sqrtf(r2);
    float lx = cp * r, ly = sp * r, lz = sqrtf(fmaxf(0.0f, 1.0f - r2));
    f3 helper = fabsf(n.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    f3 b = normalize(cross(n, helper));
    f3 t = cross(b, n);
    f3 s = normalize((t * lx + b * ly) + n * lz);
    if (dot(s, s) < 1e-8f) s = n;
    return normalize(s);
}

__device__ __forceinline__ f3 reflect(f3 d, f3 n) { return d - n * (2.0f * dot(d, n)); }

// rendering/taichi/scatter.py:6-11 with cos(theta) = 1 - 2 xi directly
// (theta = acos(1 - 2 xi) there) and sin(theta) = sqrt(1 - cos^2).
__device__ __forceinline__ f3 unit_vector(float r1, float r2) {
    float z = 1.0f - 2.0f * r1;
    float st = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float sp, cp;
    sincos_2pi(r2, &sp, &cp);
    return mk(st * cp, st * sp, z);
}

// rendering/taichi/scatter.py:50-54
__device__ __forceinline__ f3 metal_dir(f3 d, f3 n, float rough, float r1, float r2) {
    return normalize(reflect(d, n) + unit_vector(r1, r2) * rough);
}

__device__ __forceinline__ float pow5(float x) {
    float x2 = x * x;
    return (x2 * x2) * x;
}

// rendering/taichi/scatter.py:40-49, :57-60, :71-81
__device__ __forceinline__ f3 dielectric_dir(f3 d, f3 n, float eta, float xi) {
    float cos_t = fminf(-dot(d, n), 1.0f);
    float sin_t = sqrtf(1.0f - cos_t * cos_t);
    bool cannot = eta * sin_t > 1.0f;
    float r0 = (1.0f - eta) / (1.0f + eta);
    r0 = r0 * r0;
    float sch = r0 + (1.0f - r0) * pow5(1.0f - cos_t);
    if (cannot || sch > xi) return reflect(d, n);
    f3 perp = (d + n * cos_t) * eta;
    f3 par = n * (-sqrtf(fabsf(1.0f - dot(perp, perp))));
    return perp + par;
}

// ---------------------------------------------------------------- shade
// 512-thread blocks: every block appends its continuing rays and its
// next-event requests with one atomic each on the queue counters, so
// larger blocks mean fewer same-address atomics (mesh: 4.28 ms of shading
// per 64-spp render with 256 threads, 3.28 with 512, 3.38 with 1024).
constexpr int SHADE_BS = 512;
// 8 waves per SIMD: 63 VGPRs and no VGPR spills with the shading in
// shade_hit (left to itself the compiler takes 116 VGPRs, 4 waves).
// Shading is a stream of dependent loads and stores, so the extra waves
// keep more of them in flight (mesh: 3.28 ms of shading per render at 4
// waves, 3.08 at 6, 2.96 at 8; 2M-triangle scene 1.22 -> 1.13 ms from 6
// to 8).
#ifndef RX_SHADE_WAVES_PER_EU
#define RX_SHADE_WAVES_PER_EU 8
#endif
// What shading one hit hands on: the continuation ray and the next-event
// request (cont / want_nee say whether there is one).
struct shade_out {
    bool cont, want_nee;
    f3 ro2, rd2, hit_p, nrm, alb, thr0;
    float scale;
};
// rendering/taichi/trace.py:154-234 for path p at bounce `depth`, whose ray
// (ro, rd) hit (t, kind, id, hu, hv) (a miss: t < 0 or t >= 1e9): sky or
// background on a miss, emission, or the material's next-event request
// and scattered ray with Russian roulette past depth 2.  thr and col are
// the path's throughput and colour, updated in place (the caller stores
// col); the throughput and the current IOR are stored in the path's
// slots, and at depth 0 the first-hit normal, albedo and depth.
__device__ __forceinline__ void shade_hit(const ws_view& w, const rx_job& J, int p, int pix, int frame, int depth,
                                          f3 ro, f3 rd, float t, int kind, int id, float hu, float hv, f3& thr,
                                          f3& col, shade_out& o) {
    const rx_scene& sc = J.scene;
    const rx_render_params& prm = J.params;
    const bool last = depth + 1 >= prm.max_depth;
    const bool nee_on = depth < prm.direct_light_max_depth && sc.n_lights > 0;
    o.cont = false;
    o.want_nee = false;
    o.scale = 1.0f;
    o.hit_p = o.nrm = o.alb = o.thr0 = o.ro2 = o.rd2 = mk(0.0f, 0.0f, 0.0f);
    if (t < 0.0f || t >= 1e9f) {
        // trace.py:154-159
        f3 sky;
        if (prm.use_sky) {
            float s = 0.5f * (rd.y + 1.0f);
            sky = mk(1.0f, 1.0f, 1.0f) * (1.0f - s) + mk(0.5f, 0.7f, 1.0f) * s;
        } else {
            sky = ld3(prm.bg);
        }
        col = col + thr * sky;
        // no first hit: depth 0, and k_accumulate takes normal and albedo 0
        if (depth == 0) w.fd[p] = 0.0f;
        return;
    }
    rx_material m;
    f3 nn, albedo;
    if (kind == RX_KIND_BVH) {
        const rx_tri_shade sh = gload(sc.tri_shade + id);
        float u = hu, v = hv;
        float ww = (1.0f - u) - v;
        nn = normalize((ld3(sh.n0) * ww + ld3(sh.n1) * u) + ld3(sh.n2) * v);
        m = gload(sc.materials + sh.mat);
        albedo = bvh_albedo(sc, m, id, sh.has_uv, u, v, ld3(m.albedo));
    } else {
        const rx_prim pr = gload(sc.prims + id);
        nn = prim_normal(pr, ro, rd, t);
        m = gload(sc.prim_materials + id);
        albedo = ld3(m.albedo);
    }
    // trace.py:161-165
    bool front = dot(nn, rd) < 0.0f;
    if (!front) nn = -nn;
    const f3 hit_p = ro + rd * t;
    o.hit_p = hit_p;
    if (depth == 0) {
        w.fn_x[p] = nn.x;
        w.fn_y[p] = nn.y;
        w.fn_z[p] = nn.z;
        w.fa_x[p] = albedo.x;
        w.fa_y[p] = albedo.y;
        w.fa_z[p] = albedo.z;
        w.fd[p] = t;
    }
    if (m.type == RX_MAT_EMISSIVE) {
        col = col + (thr * albedo) * m.emission;
        return;
    }
    u4 r0 = rng(prm, pix, frame, depth, 0u);
    u4 r1 = rng(prm, pix, frame, depth, 1u);
    f3 nd = rd;
    if (m.type == RX_MAT_DIFFUSE) {
        if (nee_on) {
            o.want_nee = true;
            o.scale = 1.0f;
        }
        o.thr0 = thr;
        nd = diffuse_dir(nn, u01(r1.x), u01(r1.y));
        thr = thr * albedo;
    } else if (m.type == RX_MAT_METAL) {
        nd = metal_dir(rd, nn, m.roughness, u01(r1.x), u01(r1.y));
        thr = thr * albedo;
    } else if (m.type == RX_MAT_GLOSSY) {
        if (nee_on) {
            o.want_nee = true;
            o.scale = m.roughness;
        }
        o.thr0 = thr;
        if (u01(r0.w) < m.roughness)
            nd = diffuse_dir(nn, u01(r1.x), u01(r1.y));
        else
            nd = metal_dir(rd, nn, m.roughness, u01(r1.x), u01(r1.y));
        thr = thr * albedo;
    } else if (m.type == RX_MAT_DIELECTRIC) {
        float cur = depth == 0 ? 1.0f : w.ior[p];
        float eta;
        if (front) {
            eta = cur / m.ior;
            cur = m.ior;
        } else {
            eta = cur / 1.0f;
            cur = 1.0f;
        }
        w.ior[p] = cur;
        nd = dielectric_dir(rd, nn, eta, u01(r0.w));
        thr = thr * albedo;
    }
    o.nrm = nn;
    o.alb = albedo;
    // trace.py:227-234
    bool alive = true;
    if (depth > 2) {
        float rr = fmaxf(thr.x, fmaxf(thr.y, thr.z));
        if (u01(r1.z) > rr)
            alive = false;
        else
            thr = thr / rr;
    }
    if (alive && !last) {
        o.cont = true;
        o.ro2 = hit_p + nn * 1e-4f;
        o.rd2 = nd;
    }
    w.thr_x[p] = thr.x;
    w.thr_y[p] = thr.y;
    w.thr_z[p] = thr.z;
    // the IOR a camera ray's path continues in (a dielectric wrote its own
    // above)
    if (depth == 0 && m.type != RX_MAT_DIELECTRIC) w.ior[p] = 1.0f;
}

// Shading of queue slot i (path p of job J) at bounce `depth`: the ray
// (a camera ray is recomputed from its pixel and sample), its hit, the
// path's throughput and colour, then shade_hit; stores the colour.
__device__ __forceinline__ void shade_slot(const ws_view& w, const rx_job& J, int jx, int i, int p, int depth,
                                           bool cam, f3 ro, f3 rd, shade_out& o) {
    int npix = J.camera.width * J.camera.height;
    int local = p - w.path_off[jx];
    int pix = slot_pixel(J.camera, local % npix), frame = J.first_sample + local / npix;
    if (cam) {
        // k_generate's camera ray of this pixel and sample
        ro = ld3(J.camera.pos);
        rd = camera_dir(J.camera, pix % J.camera.width, pix / J.camera.width, frame);
    }
    float t, hu, hv;
    int kind, id;
    get_hit(w, i, t, kind, id, hu, hv);
    // a camera ray's path starts with throughput 1 and colour 0
    f3 thr = depth == 0 ? mk(1.0f, 1.0f, 1.0f) : mk(w.thr_x[p], w.thr_y[p], w.thr_z[p]);
    f3 col = depth == 0 ? mk(0.0f, 0.0f, 0.0f) : mk(w.col_x[p], w.col_y[p], w.col_z[p]);
    shade_hit(w, J, p, pix, frame, depth, ro, rd, t, kind, id, hu, hv, thr, col, o);
    w.col_x[p] = col.x;
    w.col_y[p] = col.y;
    w.col_z[p] = col.z;
}

// `jobs` is w.jobs, passed on its own as a read-only, unaliased pointer:
// when every path of a wave belongs to one job, the job's record (camera,
// render parameters, scene pointers) is read with scalar loads at a
// wave-uniform address instead of one vector load per field and lane.
__global__ __launch_bounds__(SHADE_BS) __attribute__((amdgpu_waves_per_eu(RX_SHADE_WAVES_PER_EU))) void k_shade(
    ws_view w, const rx_job* __restrict__ jobs, int qi, int depth) {
    constexpr int BS = SHADE_BS;
    __shared__ int32_t app[2 * (BS / 64 + 1)];
    const int n = w.counts[qi];
    // the next bounce's extend hands out its queue from the first
    // FETCH_SEGS counters, this bounce's connect from the others
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < 2 * FETCH_SEGS; k += BS) w.fetch_seg[k * FETCH_STRIDE] = 0;
    const int qo = 1 - qi;
    const int stride = gridDim.x * BS;
    for (int base = xcd_block() * BS; base < n; base += stride) {
        int i = base + threadIdx.x;
        bool valid = i < n;
        int p = 0, jx = 0;
        shade_out o;
        o.cont = o.want_nee = false;
        f3 ro = mk(0.0f, 0.0f, 0.0f), rd = ro;
        const bool cam = depth == 0 && w.cam_rays;
        if (valid) {
            if (cam) {
                // primary slot i: path path_off + (i - q_off) of its job
                jx = w.n_jobs == 1 ? 0 : owner(w.q_off, w.n_jobs, i);
                p = w.path_off[jx] + (i - w.q_off[jx]);
            } else {
                p = get_ray(w, qi, i, ro, rd);
                jx = job_of_path(w, p);
            }
        }
        // Two inlined copies: the compiler spills registers in the second,
        // which runs only for waves that straddle two jobs of a batch.  One
        // copy in a loop over a wave's jobs spilled in the common case and
        // shaded at half the speed (2M-triangle scene 2.44 against 1.08 ms).
        const int j0 = __builtin_amdgcn_readfirstlane(jx);
        if (__ballot(valid && jx != j0) == 0ull) {
            if (valid) shade_slot(w, jobs[j0], j0, i, p, depth, cam, ro, rd, o);
        } else if (valid) {
            shade_slot(w, jobs[jx], jx, i, p, depth, cam, ro, rd, o);
        }
        int slot, es;
        block_append_pair<BS>(&w.counts[qo], o.cont, &w.counts[2], o.want_nee, app, slot, es);
        if (o.cont) {
            put_ray(w, qo, slot, p, o.ro2, o.rd2);
        }
        if (o.want_nee) {
            RX_GLOBAL gv4f* e = (RX_GLOBAL gv4f*)w.e_rec + 4 * (long long)es;
            e[0] = gv4f{o.hit_p.x, o.hit_p.y, o.hit_p.z, __int_as_float(p)};
            e[1] = gv4f{o.nrm.x, o.nrm.y, o.nrm.z, o.scale};
            e[2] = gv4f{o.alb.x, o.alb.y, o.alb.z, 0.0f};
            e[3] = gv4f{o.thr0.x, o.thr0.y, o.thr0.z, 0.0f};
        }
    }
}

// ---------------------------------------------------------------- connect
// rendering/taichi/trace.py:23-57 (_sphere_light_sample)
__device__ __forceinline__ void sphere_light_sample(f3 hp, f3 c, float radius, float xi1, float xi2, f3* dir,
                                                    float* t_light, float* solid) {
    f3 tc = c - hp;
    float dist2 = fmaxf(dot(tc, tc), 1e-6f);
    float dist = sqrtf(dist2);
    f3 axis = tc / dist;
    float sin2 = fminf(radius * radius / dist2, 1.0f);
    float cos_max = sqrtf(fmaxf(0.0f, 1.0f - sin2));
    float cos_t = 1.0f - xi1 * (1.0f - cos_max);
    float sin_t = sqrtf(fmaxf(0.0f, 1.0f - cos_t * cos_t));
    float sp, cp;
    sincos_2pi(xi2, &sp, &cp);
    f3 helper = fabsf(axis.x) > 0.9f ? mk(0.0f, 1.0f, 0.0f) : mk(1.0f, 0.0f, 0.0f);
    f3 b = normalize(cross(axis, helper));
    f3 t = cross(b, axis);
    f3 ld = normalize((t * (cp * sin_t) + b * (sp * sin_t)) + axis * cos_t);
    f3 oc = hp - c;
    float half_b = dot(oc, ld);
    float cc = dot(oc, oc) - radius * radius;
    float disc = fmaxf(half_b * half_b - cc, 0.0f);
    *t_light = -half_b - sqrtf(disc);
    *solid = 2.0f * (1.0f - cos_max);
    *dir = ld;
}

// rendering/taichi/trace.py:60-93 (direct_light_sample) for every queued
// next-event request, then colour += (throughput * direct) * scale
// (scale = 1 for diffuse, roughness for glossy, trace.py:199-212).
//
// Persistent like k_extend: lanes that finish a request take new ones
// (the second set of FETCH_SEGS counters, zeroed by shade).  A lane's
// request walks its lights in
// order (one light in mode "one", all of them in mode "all"); for each it
// samples the cone, casts the shadow ray if n.l > 0 and max_t > 0.001, and
// runs the any-hit traversal a leaf at a time between refills.  The
// per-request order of lights and the sum order of `direct` are the
// reference's.
template <int SBS, int TLN, bool COUNT>
__global__ __launch_bounds__(SBS) __attribute__((amdgpu_waves_per_eu(CONNECT_WAVES_PER_EU))) void k_connect(ws_view w, int depth) {
    constexpr int SL = TLN > 0 ? STACK_LDS_TL : STACK_LDS;
    __shared__ int2 stack_mem[SL * SBS];
    __shared__ gv4f tl_box_mem[TLN > 0 ? 3 * TLN : 1];
    __shared__ gv2i tl_link_mem[TLN > 0 ? TLN : 1];
    lane_stack<SBS, SL> stk{stack_mem + threadIdx.x, w.spill + blockIdx.x * SBS + threadIdx.x, (int)gridDim.x * SBS};
    const int tl_n = load_treelet<SBS, TLN>(w, tl_box_mem, tl_link_mem);
    const RX_LDS gv4f* tl_box = (const RX_LDS gv4f*)tl_box_mem;
    const RX_LDS gv2i* tl_link = (const RX_LDS gv2i*)tl_link_mem;
    const int n = w.counts[2];
    int32_t* fetch = w.fetch_seg + FETCH_SEGS * FETCH_STRIDE;
    int seg = home_segment(SBS), tried = 0;
    const int cap = wave_cap(n, SBS);
    const int lane = threadIdx.x & 63;
    trav_count cnt{0, 0, 0};
    unsigned long long shadow = 0;
    lane_rays jr;
    bool exhausted = false;
    int i = -1;  // request of this lane, -1 when idle
    // request state
    int p = 0, pix = 0, frame = 0, L = 0, li = 0, l_end = 0, cast = 0, jcur = 0;
    bool all = false;
    const rx_job* J = nullptr;
    f3 hp = mk(0.0f, 0.0f, 0.0f), nn = hp, alb = hp, direct = hp;
    u4 r0{0u, 0u, 0u, 0u};
    // shadow-ray state
    bool tracing = false, have = false, occ = false;
    int sp = 0, cur = 0;
    f3 so = hp, ld = hp, inv = hp;
    float max_t = 0.0f, ndotl = 0.0f, solid = 0.0f;
    while (true) {
        bool idle = i < 0;
        int mine = exhausted ? n : wave_fetch(fetch, n, cap, idle, exhausted, seg, tried);
        {
            if (idle && mine < n) {
                i = w.nee_slot ? (int)w.nee_slot[mine] : mine;
                const RX_GLOBAL gv4f* e = (const RX_GLOBAL gv4f*)w.e_rec + 4 * (long long)i;
                const gv4f e0 = e[0], e1 = e[1], e2 = e[2];
                p = __float_as_int(e0.w);
                int jx = job_of_path(w, p);
                jcur = jx;
                J = &w.jobs[jx];
                int npix = J->camera.width * J->camera.height;
                int local = p - w.path_off[jx];
                pix = slot_pixel(J->camera, local % npix);
                frame = J->first_sample + local / npix;
                L = J->scene.n_lights;
                all = J->params.direct_light_mode == 1;
                hp = mk(e0.x, e0.y, e0.z);
                nn = mk(e1.x, e1.y, e1.z);
                alb = mk(e2.x, e2.y, e2.z);
                r0 = rng(J->params, pix, frame, depth, 0u);
                int sel = (int)(u01(r0.x) * (float)L);
                if (sel > L - 1) sel = L - 1;
                li = all ? 0 : sel;
                l_end = all ? L : sel + 1;
                direct = mk(0.0f, 0.0f, 0.0f);
                cast = 0;
                tracing = false;
            }
        }
        if (__ballot(i >= 0) == 0ull) break;
        if (i >= 0) {
            // start the shadow ray of the next light that casts one
            while (!tracing && li < l_end) {
                float xi1, xi2;
                if (all) {
                    u4 rl = rng(J->params, pix, frame, depth, 2u + (uint32_t)li);
                    xi1 = u01(rl.x);
                    xi2 = u01(rl.y);
                } else {
                    xi1 = u01(r0.y);
                    xi2 = u01(r0.z);
                }
                const rx_light lt = gload(J->scene.lights + li);
                float t_light;
                sphere_light_sample(hp, ld3(lt.center), lt.radius, xi1, xi2, &ld, &t_light, &solid);
                ndotl = dot(nn, ld);
                max_t = t_light * 0.999f;
                if (ndotl > 0.0f && max_t > 0.001f) {
                    so = hp + nn * 1e-4f;
                    cast++;
                    tracing = true;
                    // primitive scan, then the BVH (intersect.py:222-246)
                    occ = any_prim_lanes<COUNT>(w, jcur, so, ld, max_t, cnt);
                    have = false;
                    if (!occ && J->scene.n_tris > 0) {
                        const rx_bvh_header hdr = gload(J->scene.header);
                        inv = inv_dir(ld);
                        float rt;
                        if (slab(so, inv, hdr.bmin[0], hdr.bmin[1], hdr.bmin[2], hdr.bmax[0], hdr.bmax[1],
                                 hdr.bmax[2], max_t, &rt)) {
                            have = true;
                            cur = TLN > 0 && tl_n > 0 && J->scene.nodes == w.tl_nodes ? TL_REF : hdr.root_link;
                            sp = 0;
                        }
                    }
                } else {
                    ++li;
                }
            }
            if (tracing && have && !occ) {
                const rx_node* nodes = J->scene.nodes;
                const rx_tri* tris = J->scene.tris;
                const int entered = __popcll(__ballot(true));
                for (int it = 0;; ++it) {
                    bool in = have && cur >= 0;
                    int n_in = __popcll(__ballot(in));
                    if (inner_done(n_in, entered, it)) break;
                    if (in) {
                        if (COUNT) cnt.pops++;
                        gv4f a, b, c;
                        gv2i dl;
                        if (TLN > 0 && cur >= TL_REF) {
                            const int k = cur - TL_REF;
                            a = tl_box[3 * k];
                            b = tl_box[3 * k + 1];
                            c = tl_box[3 * k + 2];
                            dl = tl_link[k];
                        } else {
                            const RX_GLOBAL gv4f* q = (const RX_GLOBAL gv4f*)(nodes + cur);
                            a = q[0];
                            b = q[1];
                            c = q[2];
                            dl = ((const RX_GLOBAL gv2i*)q)[6];
                        }
                        if (!visit_boxes(a, b, c, dl, so, inv, max_t, stk, sp, cur)) have = pop_next(stk, sp, max_t, cur);
                    }
                }
                if (have && cur < 0) {
                    if (COUNT) cnt.pops++;
                    int lw = ~cur;
                    int start = lw >> 8, end = start + (lw & 255);
                    for (int t2 = start; t2 < end; ++t2) {
                        f3 v0, e1, e2;
                        load_tri(tris, t2, v0, e1, e2);
                        float u, v;
                        float t = tri_hit(so, ld, v0, e1, e2, &u, &v);
                        if (COUNT) cnt.tris++;
                        if (0.001f < t && t < max_t) {
                            occ = true;
                            break;
                        }
                    }
                    if (!occ) have = pop_next(stk, sp, max_t, cur);
                }
            }
            if (tracing && (occ || !have)) {
                // shadow ray finished: unoccluded lights add their term
                if (!occ) {
                    const rx_light lt = gload(J->scene.lights + li);
                    float wgt = all ? 1.0f : (float)L;
                    direct = direct + ((((alb * ld3(lt.albedo)) * lt.emission) * ndotl) * solid) * wgt;
                }
                tracing = false;
                ++li;
            }
            if (!tracing && li >= l_end) {
                const RX_GLOBAL gv4f* e = (const RX_GLOBAL gv4f*)w.e_rec + 4 * (long long)i;
                const gv4f e3 = e[3];
                f3 thr = mk(e3.x, e3.y, e3.z);
                f3 add = (thr * direct) * ((const RX_GLOBAL float*)e)[7];
                w.col_x[p] = w.col_x[p] + add.x;
                w.col_y[p] = w.col_y[p] + add.y;
                w.col_z[p] = w.col_z[p] + add.z;
                if (w.ray_per_path)
                    w.p_rays[p] += cast;
                else
                    jr.add(w, jcur, (unsigned long long)cast);
                shadow += cast;
                i = -1;
            }
        }
    }
    rx_counters* ctr = ctr_slot(w);
    wave_add(&ctr->rays, shadow);
    wave_add(&ctr->shadow_rays, shadow);
    if (!w.ray_per_path) jr.flush(w);
    if (COUNT) {
        wave_add(&ctr->shadow_node_pops, cnt.pops);
        wave_add(&ctr->shadow_tri_tests, cnt.tris);
        wave_add(&ctr->shadow_prim_tests, cnt.prims);
    }
}

// ---------------------------------------------------------------- adaptive stop
// A path's colour sample as it is accumulated (trace.py:262-266 clamp).
__device__ __forceinline__ f3 clamped_sample(const ws_view& w, const rx_render_params& prm, long long p) {
    f3 smp = mk(w.col_x[p], w.col_y[p], w.col_z[p]);
    if (prm.sample_clamp > 0.0f) {
        float lim = prm.sample_clamp;
        smp = mk(fminf(fmaxf(smp.x, 0.0f), lim), fminf(fmaxf(smp.y, 0.0f), lim), fminf(fmaxf(smp.z, 0.0f), lim));
    }
    return smp;
}

__device__ __forceinline__ unsigned long long err_units(float d) {
    return __double2ull_rn((double)fminf(d, ERR_CAP) * ERR_SCALE);
}

// Error sums of rendering/ray_tracer.py:534 for every sample of this pass of
// the adaptive jobs: the change of each pixel's running mean (the same
// arithmetic as k_accumulate) summed over channels and pixels per sample
// index, into the job's partial slots.  Accumulators are only read.
__global__ __launch_bounds__(BS) void k_adapt_err(ws_view w) {
    const int stride = gridDim.x * BS;
    const int lane = threadIdx.x & 63;
    for (long long base = (long long)xcd_block() * BS; base < w.NP; base += stride) {
        long long q = base + threadIdx.x;
        int jx = 0, ns = 0, k = 0, npix = 1, first = 0;
        f3 c = mk(0.0f, 0.0f, 0.0f);
        const rx_job* J = nullptr;
        if (q < w.NP) {
            jx = w.n_jobs == 1 ? 0 : owner(w.pix_off, w.n_jobs, q);
            J = &w.jobs[jx];
            if (J->params.flags & RX_FLAG_ADAPTIVE) {
                ns = pass_samples(*J);
                npix = J->camera.width * J->camera.height;
                k = (int)(q - w.pix_off[jx]);
                first = J->first_sample;
                if (ns > 0) c = ld3(J->accum.color + 3 * slot_pixel(J->camera, k));
            }
        }
        // every lane walks the wave's longest sample range (wave-wide sums)
        int wave_ns = ns;
        for (int off = 32; off > 0; off >>= 1) wave_ns = max(wave_ns, __shfl_xor(wave_ns, off, 64));
        int j0 = __shfl(jx, 0, 64);
        bool uniform = __ballot(jx != j0) == 0ull;
        unsigned long long* part = w.err_part + ((size_t)(blockIdx.x & (EPART - 1)) * w.n_jobs) * w.err_stride;
        for (int s = 0; s < wave_ns; ++s) {
            unsigned long long units = 0;
            if (s < ns) {
                long long p = (long long)w.path_off[jx] + (long long)s * npix + k;
                f3 smp = clamped_sample(w, J->params, p);
                int frame = first + s;
                if (frame == 0) {
                    c = smp;
                } else {
                    float f = (float)frame, f1 = (float)(frame + 1);
                    f3 nc = (c * f + smp) / f1;
                    units = err_units(fabsf(nc.x - c.x)) + err_units(fabsf(nc.y - c.y)) + err_units(fabsf(nc.z - c.z));
                    c = nc;
                }
            }
            if (uniform) {
                unsigned long long t = wave_sum_u64(units);
                if (lane == 0 && t) atomicAdd(part + (size_t)j0 * w.err_stride + s, t);
            } else if (units) {
                atomicAdd(part + (size_t)jx * w.err_stride + s, units);
            }
        }
    }
}

// Samples of this pass each job accumulates (one thread per job): all of
// them, or for an adaptive job the samples up to the first n at which
// rendering/ray_tracer.py:531-541 stops: n < requested total,
// n >= min_samples, (n - min_samples) % check_interval == 0 and error <=
// threshold, the error after n >= 2 samples being the mean over pixels and
// channels of |mean_n - mean_(n-1)|.  Clears the error sums it read.
__global__ void k_decide(ws_view w) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= w.n_jobs) return;
    rx_job& J = w.jobs[j];
    int ns = pass_samples(J), eff = ns;
    if ((J.params.flags & RX_FLAG_ADAPTIVE) && ns > 0) {
        const rx_render_params& prm = J.params;
        const long long total = (long long)J.first_sample + J.samples_left;
        const double denom = 3.0 * (double)J.camera.width * (double)J.camera.height;
        bool stop = false;
        for (int s = 0; s < ns; ++s) {
            unsigned long long sum = 0;
            for (int e = 0; e < EPART; ++e) {
                unsigned long long* cell = w.err_part + ((size_t)e * w.n_jobs + j) * w.err_stride + s;
                sum += *cell;
                *cell = 0ull;
            }
            long long n = (long long)J.first_sample + s + 1;
            if (stop || n < 2) continue;
            double err = (double)sum / ERR_SCALE / denom;
            J.adaptive_error = (float)err;
            if (n < total && n >= prm.adaptive_min_samples &&
                (n - prm.adaptive_min_samples) % prm.adaptive_check_interval == 0 && err <= (double)prm.adaptive_threshold) {
                eff = s + 1;
                stop = true;
                J.adaptive_stopped = 1;
            }
        }
    }
    w.ns_eff[j] = eff;
}

// ---------------------------------------------------------------- accumulate
// rendering/taichi/trace.py:262-283: clamp, running mean of colour and the
// first-hit guides in sample order, display = sqrt(min(acc, 1)); the
// pass's rays are added to the job's count.
__global__ __launch_bounds__(BS) void k_accumulate(ws_view w) {
    const int stride = gridDim.x * BS;
    for (long long base = (long long)xcd_block() * BS; base < w.NP; base += stride) {
        long long q = base + threadIdx.x;
        int jx = 0;
        unsigned long long rays = 0;
        if (q < w.NP) {
            jx = w.n_jobs == 1 ? 0 : owner(w.pix_off, w.n_jobs, q);
            const rx_job& J = w.jobs[jx];
            const rx_render_params& prm = J.params;
            const rx_accum& acc = J.accum;
            int npix = J.camera.width * J.camera.height;
            int k = (int)(q - w.pix_off[jx]);  // slot within a sample
            int pix = slot_pixel(J.camera, k);
            int ns = w.ns_eff[jx];
            if (ns > 0) {
                f3 c = ld3(acc.color + 3 * pix), nn = ld3(acc.normal + 3 * pix), aa = ld3(acc.albedo + 3 * pix);
                float dd = acc.depth[pix];
                for (int s = 0; s < ns; ++s) {
                    long long p = (long long)w.path_off[jx] + (long long)s * npix + k;
                    int frame = J.first_sample + s;
                    f3 smp = clamped_sample(w, prm, p);
                    // a path's first-hit normal and albedo are written only
                    // when its camera ray hit something (depth > 0.001);
                    // otherwise they are 0
                    float ds = w.fd[p];
                    f3 ns_ = mk(0.0f, 0.0f, 0.0f), as_ = ns_;
                    if (ds > 0.0f) {
                        ns_ = mk(w.fn_x[p], w.fn_y[p], w.fn_z[p]);
                        as_ = mk(w.fa_x[p], w.fa_y[p], w.fa_z[p]);
                    }
                    if (frame == 0) {
                        c = smp;
                        nn = ns_;
                        aa = as_;
                        dd = ds;
                    } else {
                        float f = (float)frame, f1 = (float)(frame + 1);
                        c = (c * f + smp) / f1;
                        nn = (nn * f + ns_) / f1;
                        aa = (aa * f + as_) / f1;
                        dd = (dd * f + ds) / f1;
                    }
                    if (w.ray_per_path) rays += (unsigned long long)w.p_rays[p];
                }
                acc.color[3 * pix] = c.x;
                acc.color[3 * pix + 1] = c.y;
                acc.color[3 * pix + 2] = c.z;
                acc.normal[3 * pix] = nn.x;
                acc.normal[3 * pix + 1] = nn.y;
                acc.normal[3 * pix + 2] = nn.z;
                acc.albedo[3 * pix] = aa.x;
                acc.albedo[3 * pix + 1] = aa.y;
                acc.albedo[3 * pix + 2] = aa.z;
                acc.depth[pix] = dd;
                if (acc.display) {
                    acc.display[3 * pix] = sqrtf(fminf(c.x, 1.0f));
                    acc.display[3 * pix + 1] = sqrtf(fminf(c.y, 1.0f));
                    acc.display[3 * pix + 2] = sqrtf(fminf(c.z, 1.0f));
                }
            }
        }
        // per-job ray totals into this block's partial slot (folded by
        // k_advance): one atomic per wave when the wave is one job
        unsigned long long* part = w.ray_part + (size_t)(blockIdx.x & (NPART - 1)) * w.n_jobs;
        int j0 = __shfl(jx, 0, 64);
        bool uniform = __ballot(jx != j0) == 0ull;
        if (uniform) {
            unsigned long long s = wave_sum_u64(rays);
            if ((threadIdx.x & 63) == 0 && s) atomicAdd(part + j0, s);
        } else if (rays) {
            atomicAdd(part + jx, rays);
        }
    }
}

// ---------------------------------------------------------------- stage timing
struct stage_timer {
    bool on = false;
    hipStream_t s = nullptr;
    static constexpr int MAXE = 512;
    hipEvent_t ev[MAXE];
    int stage[MAXE / 2];
    int n = 0;
    int begin(int st) {
        if (!on) return RX_OK;
        if (n + 2 > MAXE) return rx::set_error(RX_ERR_INVALID, "too many profiled launches in one pass");
        RX_TRY_HIP(hipEventCreate(&ev[n]));
        RX_TRY_HIP(hipEventCreate(&ev[n + 1]));
        stage[n / 2] = st;
        RX_TRY_HIP(hipEventRecord(ev[n], s));
        return RX_OK;
    }
    int end() {
        if (!on) return RX_OK;
        RX_TRY_HIP(hipEventRecord(ev[n + 1], s));
        n += 2;
        return RX_OK;
    }
    int finish(rx_profile* prof) {
        if (!on) return RX_OK;
        RX_TRY_HIP(hipStreamSynchronize(s));
        float sum[7] = {0, 0, 0, 0, 0, 0, 0};
        int cnt[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int k = 0; k < n; k += 2) {
            float ms = 0.0f;
            RX_TRY_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            sum[stage[k / 2]] += ms;
            cnt[stage[k / 2]]++;
            (void)hipEventDestroy(ev[k]);
            (void)hipEventDestroy(ev[k + 1]);
        }
        n = 0;
        if (prof) {
            prof->ms_generate += sum[0];
            prof->ms_extend += sum[1];
            prof->ms_shade += sum[2];
            prof->ms_connect += sum[3];
            prof->ms_accumulate += sum[4];
            prof->n_extend += cnt[1];
            prof->n_connect += cnt[3];
            prof->ms_prim_scan += sum[5];
            prof->n_prim_scan += cnt[5];
            prof->ms_reorder += sum[6];
            prof->n_reorder += cnt[6];
        }
        return RX_OK;
    }
    ~stage_timer() {
        for (int k = 0; k < n; ++k) (void)hipEventDestroy(ev[k]);
    }
};

#define RX_STAGE(timer, st, launch)       \
    do {                                  \
        int rx_s_ = (timer).begin(st);    \
        if (rx_s_) return rx_s_;          \
        launch;                           \
        RX_CHECK_LAUNCH();                \
        rx_s_ = (timer).end();            \
        if (rx_s_) return rx_s_;          \
    } while (0)

// Host-side description of the launches of one pass.
struct pass_plan {
    int max_depth;           // bounces (max over jobs)
    int nee_depth;           // connect launched for depth < nee_depth
    bool deep;               // stack class
    bool count;              // traversal counters
    bool prim_prepass;       // closest-primitive scan as its own pass
    int trav_per_cu;         // resident traversal blocks per CU (0 = default)
    int max_spp;             // most samples per pass of any job
    bool adaptive;           // some job has RX_FLAG_ADAPTIVE
    bool sort_rays;          // reorder ray queues for k_extend (no job asks to keep the order)
    bool sort_nee;           // group next-event requests by the light they sample (a job
                             // samples one of several lights per request)
    // BVH whose top the traversal kernels keep in LDS: the one of the job
    // with the most paths per pass (null: no triangles)
    const rx_node* tl_nodes;
    const rx_bvh_header* tl_header;
};

// Scenes with more analytic primitives than this get their closest-
// primitive scan as a pre-pass over the ray queue (every lane of a wave
// scans together) instead of inside the persistent traversal kernel,
// where only the lanes taking a new ray run the loop.  Measured on a
// 67-primitive scene the inline scan is still faster (the traversal
// kernel waits on node fetches, not on arithmetic), hence the high bar;
// RX_FLAG_PRIM_PREPASS asks for the pre-pass at any count.
constexpr int PRIM_INLINE_MAX = 256;

// Passes of fewer path slots than this run the 64-thread traversal blocks
// without a treelet: building and loading the treelet, and filling a CU
// with two 1024-thread blocks, do not pay off on a few hundred thousand
// rays (a loop of one-sample passes over small frames).
constexpr long long SMALL_PASS_PATHS = 1 << 19;

// ---------------------------------------------------------------- reorder
// Secondary rays leave in scattered directions, so lanes of one wave walk
// unrelated subtrees.  Before k_extend of a bounce > 0 the queue is sorted
// by a 30-bit key that puts rays of one job with similar directions and
// nearby origins next to each other: the job (most significant, as many
// bits as the batch needs), the direction quantised to 8 levels per axis
// (9 bits), and a Morton code of the origin in the job's scene box with
// the bits left (7 per axis for a single job).  k_extend walks the rays in
// sorted order, reading each through the sorted slot list, and writes each
// hit back to its ray's own queue slot, so shading, path state and every
// result are those of the unsorted queue.
__device__ __forceinline__ uint32_t spread7(uint32_t v) {
    v &= 0x7fu;
    v = (v | (v << 8)) & 0x0000f00fu;
    v = (v | (v << 4)) & 0x000c30c3u;
    v = (v | (v << 2)) & 0x00249249u;
    return v;
}

// The single-job key: a two-point key (Meister et al., "On Ray Reordering
// Techniques for Faster GPU Ray Tracing", I3D 2020), the 6-D Morton code of
// the origin and of the point half the scene's largest extent along the
// ray, 5 bits per coordinate of the scene box (30 bits): rays that start
// near each other and head the same way sort together.  Measured against
// the direction-major key below: 2M-triangle scene 23.28-23.31 against
// 23.94-23.96 ms per render, mesh 33.02-33.20 against 33.04-33.40 (a
// quarter extent: 23.17 / 33.52, a whole extent: 23.57 / 33.00-33.07).
constexpr float KEY_REACH = 0.5f;
__device__ __forceinline__ uint32_t spread5_6d(uint32_t v) {
    uint32_t r = 0;
#pragma unroll
    for (int b = 0; b < 5; ++b) r |= ((v >> b) & 1u) << (6 * b);
    return r;
}
__device__ __forceinline__ uint32_t ray_key_two_point(f3 o, f3 d, const rx_bvh_header& h) {
    const float oc[3] = {o.x, o.y, o.z}, dc[3] = {d.x, d.y, d.z};
    float ext = 0.0f;
    for (int a = 0; a < 3; ++a) ext = fmaxf(ext, h.bmax[a] - h.bmin[a]);
    uint32_t key = 0;
    for (int a = 0; a < 3; ++a) {
        const float span = fmaxf(h.bmax[a] - h.bmin[a], 1e-20f);
        const float u0 = (oc[a] - h.bmin[a]) / span;
        const float u1 = (oc[a] + dc[a] * (KEY_REACH * ext) - h.bmin[a]) / span;
        const uint32_t q0 = (uint32_t)(fminf(fmaxf(u0, 0.0f), 1.0f) * 31.0f);
        const uint32_t q1 = (uint32_t)(fminf(fmaxf(u1, 0.0f), 1.0f) * 31.0f);
        key |= spread5_6d(q0) << (5 - a);
        key |= spread5_6d(q1) << (2 - a);
    }
    return key;
}

// The batch key: the job, the direction quantised to 8 levels per axis,
// then a Morton code of the origin with ob bits per axis (0..7).
__device__ __forceinline__ uint32_t ray_key(f3 o, f3 d, const rx_bvh_header& h, uint32_t job, int ob) {
    uint32_t q[3];
    const float oc[3] = {o.x, o.y, o.z};
    for (int a = 0; a < 3; ++a) {
        float u = (oc[a] - h.bmin[a]) / fmaxf(h.bmax[a] - h.bmin[a], 1e-20f);
        q[a] = (uint32_t)(fminf(fmaxf(u, 0.0f), 1.0f) * 127.0f) >> (7 - ob);
    }
    uint32_t dx = (uint32_t)fminf((d.x + 1.0f) * 4.0f, 7.0f);
    uint32_t dy = (uint32_t)fminf((d.y + 1.0f) * 4.0f, 7.0f);
    uint32_t dz = (uint32_t)fminf((d.z + 1.0f) * 4.0f, 7.0f);
    uint32_t dq = (dx << 6) | (dy << 3) | dz;
    uint32_t mo = (spread7(q[0]) << 2) | (spread7(q[1]) << 1) | spread7(q[2]);
    return (job << (9 + 3 * ob)) | (dq << (3 * ob)) | mo;
}

// Keys and the identity permutation of the slots of queue b, up to its
// length (read on the device).
__global__ void k_ray_keys(ws_view w, int b, int ob, uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const int count = w.counts[b];
    for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < count; i += (long long)gridDim.x * BS) {
        f3 o, d;
        const int p = get_ray(w, b, (int)i, o, d);
        const int jx = job_of_path(w, p);
        const rx_scene& sc = w.jobs[jx].scene;
        key[i] = sc.n_tris == 0 ? (uint32_t)jx << (9 + 3 * ob)
                 : w.n_jobs == 1 ? ray_key_two_point(o, d, gload(sc.header))
                                 : ray_key(o, d, gload(sc.header), (uint32_t)jx, ob);
        idx[i] = (uint32_t)i;
    }
}

// Keys and the identity permutation of the next-event requests, up to the
// queue's length: the job, the light the request samples in mode "one"
// (its index, capped at 7; 0 in mode "all"), then a Morton code of the
// hit point in the job's scene box, so a wave of k_connect casts shadow
// rays toward one light from nearby points.  k_connect draws the same
// light choice from the same Philox counter, so the order changes which
// lane casts a ray, never the ray.
__global__ void k_nee_keys(ws_view w, int depth, int ob, uint32_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const int count = w.counts[2];
    for (long long i = (long long)blockIdx.x * BS + threadIdx.x; i < count; i += (long long)gridDim.x * BS) {
        const gv4f e0 = ((const RX_GLOBAL gv4f*)w.e_rec)[4 * i];
        const int p = __float_as_int(e0.w);
        const int jx = job_of_path(w, p);
        const rx_job& J = w.jobs[jx];
        uint32_t light = 0;
        const int L = J.scene.n_lights;
        if (J.params.direct_light_mode != 1 && L > 1) {
            const int npix = J.camera.width * J.camera.height;
            const int local = p - w.path_off[jx];
            const u4 r0 = rng(J.params, slot_pixel(J.camera, local % npix), J.first_sample + local / npix, depth, 0u);
            int sel = (int)(u01(r0.x) * (float)L);
            sel = sel > L - 1 ? L - 1 : sel;
            light = (uint32_t)(sel < 7 ? sel : 7);
        }
        uint32_t mo = 0;
        if (J.scene.n_tris > 0) {
            const rx_bvh_header h = gload(J.scene.header);
            const float pc[3] = {e0.x, e0.y, e0.z};
            uint32_t q[3];
            for (int a = 0; a < 3; ++a) {
                const float u = (pc[a] - h.bmin[a]) / fmaxf(h.bmax[a] - h.bmin[a], 1e-20f);
                q[a] = (uint32_t)(fminf(fmaxf(u, 0.0f), 1.0f) * 127.0f) >> (7 - ob);
            }
            mo = (spread7(q[0]) << 2) | (spread7(q[1]) << 1) | spread7(q[2]);
        }
        key[i] = ((uint32_t)jx << (3 + 3 * ob)) | (light << (3 * ob)) | mo;
        idx[i] = (uint32_t)i;
    }
}

// Origin bits per axis of the next-event key for a batch of n_jobs, and
// the key's width: job bits, 3 light bits, 3 origin bits per axis.
inline int nee_key_job_bits(int n_jobs) {
    int jb = 0;
    while ((1 << jb) < n_jobs) ++jb;
    return jb;
}
inline int nee_key_origin_bits(int n_jobs) {
    int ob = (30 - 3 - nee_key_job_bits(n_jobs)) / 3;
    return ob > 7 ? 7 : (ob < 0 ? 0 : ob);
}
inline unsigned nee_key_bits(int n_jobs) {
    return (unsigned)(nee_key_job_bits(n_jobs) + 3 + 3 * nee_key_origin_bits(n_jobs));
}

// Queues shorter than this are traced in shade's order (the sort kernels
// see the length on the device and leave the identity permutation): a
// short queue is spread a few rays per wave over the whole grid, where
// the sort's four passes cost more than its order saves (2M-triangle
// scene, ms per render: 21.70-21.97 with 8192, 21.50 with 2^17, 21.42
// with 2^18 and with 2^19; the mesh within noise).  Passes of fewer path
// slots than SORT_MIN_PATHS are not reordered at all: for them the sort's
// launches cost more than they save (a loop of one-sample passes of small
// frames would pay them thousands of times).
constexpr int SORT_MIN_RAYS = 1 << 18;
constexpr long long SORT_MIN_PATHS = 1 << 20;
// Nor are passes whose largest BVH has fewer triangles than this: its
// nodes and triangles stay in the L1 and L2 caches whatever the ray order
// (a 36-triangle Cornell box: 0.46 ms of sorting against 0.74 ms of
// traversal per 16-spp 256^2 render).
constexpr int SORT_MIN_TRIS = 1 << 14;

// Origin bits per axis of the sort key for a batch of n_jobs, and the
// key's width: job bits, 9 direction bits, 3 origin bits per axis.
#ifndef RX_KEY_ORIGIN_MAX
#define RX_KEY_ORIGIN_MAX 7
#endif
inline int key_job_bits(int n_jobs) {
    int jb = 0;
    while ((1 << jb) < n_jobs) ++jb;
    return jb;
}
inline int key_origin_bits(int n_jobs) {
    int ob = (30 - 9 - key_job_bits(n_jobs)) / 3;
    return ob > RX_KEY_ORIGIN_MAX ? RX_KEY_ORIGIN_MAX : (ob < 0 ? 0 : ob);
}
inline unsigned key_bits(int n_jobs) { return (unsigned)(key_job_bits(n_jobs) + 9 + 3 * key_origin_bits(n_jobs)); }

// Every queue length stays on the device: the sort reads it there, so a
// pass never waits for the host and can be captured in a HIP graph.
// Deep trees: 64-thread blocks, the stack mostly in LDS, no treelet.
// Shallow trees: the large blocks that share one LDS treelet (measured on
// the mesh: 2074 Mrays/s without it, 2168 with it; the 2M-triangle scene
// is unchanged).
template <bool DEEP>
int enqueue_pass(const ws_view& w0, const pass_plan& pl, rx_counters* ctr, stage_timer& tm, hipStream_t s) {
    constexpr int EBS = DEEP ? BS_DEEP : BS_EXTEND;
    constexpr int CBS = DEEP ? BS_DEEP : BS_CONNECT;
    constexpr int TLN = DEEP ? 0 : TL_NODES;
    int ge = DEEP ? trav_grid(w0.P, true, pl.trav_per_cu)
                  : render_trav_grid(w0.P, BS_EXTEND, EXTEND_BLOCKS_PER_CU, pl.trav_per_cu);
    int gc = DEEP ? ge : render_trav_grid(w0.P, BS_CONNECT, CONNECT_BLOCKS_PER_CU, pl.trav_per_cu);
    ws_view w = w0;
    w.tl_nodes = TLN > 0 ? pl.tl_nodes : nullptr;
    w.ray_per_path = pl.adaptive ? 1 : 0;
    int gp = grid_for(w.P, BS, 16);
    RX_STAGE(tm, 0, (k_pass_begin<<<1, PB, 0, s>>>(w)));
    if (w.tl_nodes) RX_STAGE(tm, 0, (k_treelet<<<1, TLB, 0, s>>>(pl.tl_nodes, pl.tl_header, w)));
    // passes of several samples per pixel trace their primary rays
    // pixel-major, generated by k_extend itself (pm_camera_ray); later
    // bounces are sorted below
    const bool pm = pl.sort_rays && w.si_in && pl.max_spp > 1;
    // traced pixel-major and shaded from the camera: the sample-major
    // records would only be read by the primitive pre-pass
    w.cam_rays = pm && !pl.prim_prepass ? 1 : 0;
    // k_generate has nothing to write for a single job whose camera rays
    // are traced pixel-major and shaded from the camera, with rays counted
    // per wave and every path shaded at depth 0
    if (!(w.n_jobs == 1 && w.cam_rays && !w.ray_per_path && pl.max_depth >= 1))
        RX_STAGE(tm, 0, (k_generate<<<gp, BS, 0, s>>>(w)));
    for (int d = 0; d < pl.max_depth; ++d) {
        int qi = d & 1;
        ws_view we = w;
        if (d == 0 && pm) we.cam_gen = 1;
        if (d > 0 && pl.sort_rays && w.si_in && w.P >= SORT_MIN_PATHS) {
            RX_STAGE(tm, 6, (k_ray_keys<<<gp, BS, 0, s>>>(w, qi, key_origin_bits(w.n_jobs), w.sk_in, w.si_in)));
            int rs = tm.begin(6);
            if (rs) return rs;
            uint32_t* perm = nullptr;
            RX_TRY_HIP(enqueue_ray_sort(w.sk_in, w.si_in, w.sk_out, w.si_out, w.counts + qi, w.P, SORT_MIN_RAYS,
                                        key_bits(w.n_jobs), w.sort_scratch, s, nullptr, &perm));
            rs = tm.end();
            if (rs) return rs;
            // k_extend reads each ray through the permutation: cheaper
            // than gathering a sorted copy first (2.6 against 3.7 ms of
            // reordering per 64-spp mesh render, same traversal time)
            we.ray_slot = perm;
            we.hit_slot = perm;
        }
        // the camera rays of a single-job pass start at its eye: its
        // eye-relative copies serve them when the workspace holds them
        const bool eye = d == 0 && w.eye_header != nullptr;
        // the link-free copy of the pass's (single) tree, when built
        const bool imp = w.imp_nodes != nullptr;
#define RX_EXTEND(C, PR)                                                                         \
    (imp ? (eye ? (k_extend<EBS, TLN, C, PR, true, true><<<ge, EBS, 0, s>>>(we, we.jobs, qi))        \
                : (k_extend<EBS, TLN, C, PR, false, true><<<ge, EBS, 0, s>>>(we, we.jobs, qi)))      \
         : (eye ? (k_extend<EBS, TLN, C, PR, true, false><<<ge, EBS, 0, s>>>(we, we.jobs, qi))       \
                : (k_extend<EBS, TLN, C, PR, false, false><<<ge, EBS, 0, s>>>(we, we.jobs, qi))))
        if (pl.prim_prepass) {
            if (pl.count) {
                RX_STAGE(tm, 5, (k_prim_scan<true><<<gp, BS, 0, s>>>(w, qi)));
                RX_STAGE(tm, 1, RX_EXTEND(true, true));
            } else {
                RX_STAGE(tm, 5, (k_prim_scan<false><<<gp, BS, 0, s>>>(w, qi)));
                RX_STAGE(tm, 1, RX_EXTEND(false, true));
            }
        } else if (pl.count) {
            RX_STAGE(tm, 1, RX_EXTEND(true, false));
        } else {
            RX_STAGE(tm, 1, RX_EXTEND(false, false));
        }
#undef RX_EXTEND
        RX_STAGE(tm, 2, (k_shade<<<grid_for(w.P, SHADE_BS, 16 * BS / SHADE_BS), SHADE_BS, 0, s>>>(w, w.jobs, qi, d)));
        if (d < pl.nee_depth) {
            ws_view wc = w;
            if (pl.sort_rays && pl.sort_nee && w.si_in && w.P >= SORT_MIN_PATHS) {
                // the shadow rays, grouped by light and hit point (with one
                // light, shade's order is as coherent and costs no sort)
                RX_STAGE(tm, 6, (k_nee_keys<<<gp, BS, 0, s>>>(w, d, nee_key_origin_bits(w.n_jobs), w.sk_in, w.si_in)));
                int rs = tm.begin(6);
                if (rs) return rs;
                uint32_t* perm = nullptr;
                RX_TRY_HIP(enqueue_ray_sort(w.sk_in, w.si_in, w.sk_out, w.si_out, w.counts + 2, w.P, SORT_MIN_RAYS,
                                            nee_key_bits(w.n_jobs), w.sort_scratch, s, nullptr, &perm));
                rs = tm.end();
                if (rs) return rs;
                wc.nee_slot = perm;
            }
            if (pl.count)
                RX_STAGE(tm, 3, (k_connect<CBS, TLN, true><<<gc, CBS, 0, s>>>(wc, d)));
            else
                RX_STAGE(tm, 3, (k_connect<CBS, TLN, false><<<gc, CBS, 0, s>>>(wc, d)));
        }
    }
    if (pl.adaptive) RX_STAGE(tm, 4, (k_adapt_err<<<grid_for(w.NP, BS, 16), BS, 0, s>>>(w)));
    RX_STAGE(tm, 4, (k_decide<<<(w.n_jobs + 255) / 256, 256, 0, s>>>(w)));
    RX_STAGE(tm, 4, (k_accumulate<<<grid_for(w.NP, BS, 16), BS, 0, s>>>(w)));
    RX_STAGE(tm, 4, (k_advance<<<(w.n_jobs + 255) / 256, 256, 0, s>>>(w, ctr)));
    return RX_OK;
}

int enqueue(const ws_view& w, const pass_plan& pl, rx_counters* ctr, stage_timer& tm, hipStream_t s) {
    if (pl.deep) return enqueue_pass<true>(w, pl, ctr, tm, s);
    return enqueue_pass<false>(w, pl, ctr, tm, s);
}

int check_job(const rx_job& j, int idx) {
    const rx_scene& sc = j.scene;
    const rx_render_params& p = j.params;
    RX_REQUIRE(j.camera.width > 0 && j.camera.height > 0, RX_ERR_INVALID, "job %d: camera size must be positive",
               idx);
    RX_REQUIRE(j.first_sample >= 0 && j.samples_left >= 0, RX_ERR_INVALID, "job %d: negative sample range", idx);
    RX_REQUIRE(p.samples_per_pass > 0, RX_ERR_INVALID, "job %d: samples_per_pass must be positive", idx);
    RX_REQUIRE(p.max_depth >= 0, RX_ERR_INVALID, "job %d: max_depth must be >= 0", idx);
    RX_REQUIRE(p.trav_blocks_per_cu >= 0 && p.trav_blocks_per_cu <= TRAV_BLOCKS_PER_CU_DEEP, RX_ERR_INVALID,
               "job %d: trav_blocks_per_cu must be in 0..%d", idx, TRAV_BLOCKS_PER_CU_DEEP);
    if (p.flags & RX_FLAG_ADAPTIVE) {
        // rendering/ray_tracer.py:314-319
        RX_REQUIRE(p.adaptive_min_samples >= 1, RX_ERR_INVALID, "adaptive_min_samples must be at least 1");
        RX_REQUIRE(p.adaptive_threshold >= 0.0f, RX_ERR_INVALID, "adaptive_threshold must be non-negative");
        RX_REQUIRE(p.adaptive_check_interval >= 1, RX_ERR_INVALID, "adaptive_check_interval must be at least 1");
    }
    RX_REQUIRE(p.direct_light_mode == 0 || p.direct_light_mode == 1, RX_ERR_INVALID,
               "direct_light_mode must be 'one' or 'all'");
    RX_REQUIRE(j.accum.color && j.accum.normal && j.accum.albedo && j.accum.depth, RX_ERR_INVALID,
               "job %d: accumulation buffers must be given", idx);
    RX_REQUIRE(sc.n_prims >= 0 && sc.n_tris >= 0 && sc.n_lights >= 0, RX_ERR_INVALID, "job %d: negative scene size",
               idx);
    RX_REQUIRE(sc.n_prims == 0 || (sc.prims && sc.prim_materials), RX_ERR_INVALID,
               "job %d: scene has primitives but no primitive arrays", idx);
    RX_REQUIRE(sc.n_lights == 0 || sc.lights, RX_ERR_INVALID, "job %d: scene has lights but lights is NULL", idx);
    RX_REQUIRE(rx_scene_aligned(sc), RX_ERR_INVALID, "job %d: scene arrays must be 16-byte aligned", idx);
    RX_REQUIRE(sc.n_textures >= 0, RX_ERR_INVALID, "job %d: negative texture count", idx);
    RX_REQUIRE(sc.n_textures == 0 || (sc.textures && sc.texels && (sc.n_tris == 0 || sc.tri_uv)), RX_ERR_INVALID,
               "job %d: a textured scene needs textures, texels and per-triangle UVs", idx);
    if (sc.n_tris > 0) {
        RX_REQUIRE(sc.nodes && sc.header && sc.tris && sc.tri_shade && sc.materials, RX_ERR_INVALID,
                   "job %d: scene has triangles but no BVH/shading arrays", idx);
        RX_REQUIRE(sc.bvh_depth >= 0 && sc.bvh_depth <= RX_MAX_BVH_DEPTH, RX_ERR_CAPACITY, "BVH depth %d exceeds %d",
                   sc.bvh_depth, RX_MAX_BVH_DEPTH);
        // node links share their range with the treelet slots from TL_REF up
        RX_REQUIRE(sc.n_nodes >= 0 && sc.n_nodes < TL_REF, RX_ERR_CAPACITY, "job %d: BVH of %d nodes exceeds 2^30 - 1",
                   idx, sc.n_nodes);
    }
    return RX_OK;
}

// Offsets, totals and the launch plan of a job list.
int plan_jobs(const rx_job* jobs, int n, long long* P, long long* NP, pass_plan* pl, int32_t* path_off,
              int32_t* pix_off) {
    RX_REQUIRE(jobs && n >= 1 && n <= MAX_JOBS, RX_ERR_INVALID, "need 1..%d jobs", MAX_JOBS);
    long long p = 0, q = 0;
    pass_plan plan{0, 0, false, false, false, 0, 0, false, true, false, nullptr, nullptr};
    long long tl_paths = 0;
    int max_tris = 0;
    for (int i = 0; i < n; ++i) {
        int st = check_job(jobs[i], i);
        if (st) return st;
        const rx_job& j = jobs[i];
        long long npix = (long long)j.camera.width * j.camera.height;
        if (path_off) path_off[i] = (int32_t)p;
        if (pix_off) pix_off[i] = (int32_t)q;
        p += npix * j.params.samples_per_pass;
        q += npix;
        if (j.params.max_depth > plan.max_depth) plan.max_depth = j.params.max_depth;
        if (j.scene.n_lights > 0 && j.params.direct_light_max_depth > plan.nee_depth)
            plan.nee_depth = j.params.direct_light_max_depth;
        if (j.scene.n_tris > 0 && use_deep_stack(j.scene.bvh_depth)) plan.deep = true;
        if (j.params.flags & RX_FLAG_SMALL_BLOCKS) plan.deep = true;
        if (j.params.trav_blocks_per_cu > plan.trav_per_cu) plan.trav_per_cu = j.params.trav_blocks_per_cu;
        if (j.params.samples_per_pass > plan.max_spp) plan.max_spp = j.params.samples_per_pass;
        if (j.params.flags & RX_FLAG_ADAPTIVE) plan.adaptive = true;
        if (j.params.flags & RX_FLAG_COUNT_TRAFFIC) plan.count = true;
        if (j.params.flags & RX_FLAG_KEEP_QUEUE_ORDER) plan.sort_rays = false;
        if (j.scene.n_tris > max_tris) max_tris = j.scene.n_tris;
        if (j.scene.n_lights > 1 && j.params.direct_light_mode == 0) plan.sort_nee = true;
        if (j.scene.n_tris > 0 && npix * j.params.samples_per_pass > tl_paths) {
            tl_paths = npix * j.params.samples_per_pass;
            plan.tl_nodes = j.scene.nodes;
            plan.tl_header = j.scene.header;
        }
        if (j.scene.n_prims > 0 && (j.scene.n_prims > PRIM_INLINE_MAX || (j.params.flags & RX_FLAG_PRIM_PREPASS)))
            plan.prim_prepass = true;
    }
    RX_REQUIRE(p <= 0x7fffffffLL, RX_ERR_CAPACITY, "batch of %lld paths exceeds 2^31-1", p);
    if (p < SMALL_PASS_PATHS) plan.deep = true;
    if (max_tris < SORT_MIN_TRIS) plan.sort_rays = false;
    if (path_off) path_off[n] = (int32_t)p;
    if (pix_off) pix_off[n] = (int32_t)q;
    *P = p;
    *NP = q;
    *pl = plan;
    return RX_OK;
}

// sort: room for reordering secondary rays (the single-job entry point)
size_t workspace_bytes_for(int n_jobs, long long P, bool deep, int err_stride, bool sort = true) {
    rx_arena ar{nullptr, 0, 0};
    ws_view w;
    carve(ar, n_jobs, P, deep, err_stride, w, sort);
    return ar.off + 256;
}

int carve_into(void* d_workspace, size_t bytes, int n_jobs, long long P, long long NP, bool deep, int err_stride,
               ws_view& w, bool sort = true) {
    size_t need = workspace_bytes_for(n_jobs, P, deep, err_stride, sort);
    RX_REQUIRE(d_workspace && bytes >= need, RX_ERR_WORKSPACE, "render workspace too small: need %zu bytes, have %zu",
               need, bytes);
    rx_arena ar{reinterpret_cast<char*>(d_workspace), bytes, 0};
    ar.off = (256 - (reinterpret_cast<uintptr_t>(d_workspace) & 255)) & 255;
    carve(ar, n_jobs, P, deep, err_stride, w, sort);
    w.NP = NP;
    return RX_OK;
}

// Whether a render traces the link-free copy of the scene's tree: its
// boxes are finite (the scene says so), its root is interior, and the
// heap slots of its depth are at most four per node (a tree that is not
// far from balanced, as the Morton and median builds give).
bool imp_usable(const rx_scene& sc) {
    return sc.n_tris > 0 && (sc.flags & RX_SCENE_FINITE_BOXES) != 0 && sc.bvh_depth >= 1 &&
           sc.bvh_depth <= IMP_MAX_DEPTH && (2LL << sc.bvh_depth) - 1 <= 4LL * sc.n_nodes;
}

// The eye-relative copies of a scene (k_eye_prep), after a render's own
// workspace: the root record, the nodes, the triangles, 32 bytes per
// primitive.
// ... and, for 1..PRIM_TILE_MAX primitives, the per-tile primitive masks
// of a width x height camera.
// ... and, when imp_usable, the link-free copy of the tree in world space
// and minus the eye (64 bytes per interior heap slot each) and 4 bytes
// per heap slot for the leaves (the eye-relative node records are then
// not needed).
void take_eye(rx_arena& ar, const rx_scene& sc, int width, int height, ws_view& w) {
    const bool imp = imp_usable(sc);
    w.eye_header = ar.take<rx_bvh_header>(1);
    w.eye_nodes = sc.n_tris > 0 && !imp ? ar.take<rx_node>((size_t)sc.n_nodes) : nullptr;
    w.eye_tris = ar.take<rx_tri>((size_t)sc.n_tris);
    w.eye_prims = ar.take<gv4f>(2 * (size_t)sc.n_prims);
    w.tile_prims = nullptr;
    w.tile_cols = (width + 7) >> 3;
    if (sc.n_prims >= 1 && sc.n_prims <= PRIM_TILE_MAX) {
        const size_t tiles = (size_t)w.tile_cols * (size_t)((height + 7) >> 3);
        w.tile_prims = ar.take<uint64_t>(PRIM_TILE_WORDS * tiles);
    }
    w.imp_nodes = w.eye_imp = nullptr;
    w.imp_leaf = nullptr;
    w.imp_records = 0;
    if (imp) {
        w.imp_records = (1 << sc.bvh_depth) - 1;
        w.imp_nodes = ar.take<gv4f>(4 * (size_t)w.imp_records);
        w.eye_imp = ar.take<gv4f>(4 * (size_t)w.imp_records);
        w.imp_leaf = ar.take<int32_t>(2 * (size_t)w.imp_records + 1);
    }
}

size_t eye_bytes_for(const rx_scene& sc, int width, int height) {
    rx_arena ar{nullptr, 0, 0};
    ws_view w;
    take_eye(ar, sc, width, height, w);
    return ar.off + 256;
}

// Camera rays the call traces against the records the copies take: they
// pay for themselves when each record is read by many camera rays (and
// the copy is a single streaming pass).
bool eye_worthwhile(const rx_scene& sc, long long camera_rays) {
    const long long records = (sc.n_tris > 0 ? (long long)sc.n_nodes + sc.n_tris : 0) + sc.n_prims;
    return records > 0 && camera_rays >= 4 * records;
}

// Carves the copies from the workspace past its first `used` bytes; false
// (and no copies) when they do not fit.
bool carve_eye(void* d_workspace, size_t bytes, size_t used, const rx_scene& sc, int width, int height,
               ws_view& w) {
    if (bytes < used || bytes - used < eye_bytes_for(sc, width, height)) return false;
    rx_arena ar{reinterpret_cast<char*>(d_workspace), bytes, used};
    take_eye(ar, sc, width, height, w);
    return true;
}

}  // namespace

struct rx_batch {
    ws_view w;
    pass_plan plan;
    int n_jobs;
    bool use_graph;
    hipGraphExec_t exec;
    bool captured;
    rx_counters* graph_ctr;
    int graph_nodes[3];  // kernel, memset, other nodes of the captured pass
};

namespace {

// Node types of a captured pass (kernels, memory fills, anything else).
hipError_t count_graph_nodes(hipGraph_t g, int out[3]) {
    out[0] = out[1] = out[2] = 0;
    size_t n = 0;
    hipError_t e = hipGraphGetNodes(g, nullptr, &n);
    if (e != hipSuccess || n == 0) return e;
    hipGraphNode_t* nodes = (hipGraphNode_t*)malloc(sizeof(hipGraphNode_t) * n);
    if (!nodes) return hipErrorOutOfMemory;
    e = hipGraphGetNodes(g, nodes, &n);
    for (size_t k = 0; e == hipSuccess && k < n; ++k) {
        hipGraphNodeType t;
        e = hipGraphNodeGetType(nodes[k], &t);
        if (e == hipSuccess) out[t == hipGraphNodeTypeKernel ? 0 : (t == hipGraphNodeTypeMemset ? 1 : 2)]++;
    }
    free(nodes);
    return e;
}

}  // namespace

extern "C" int rx_render_workspace_size(int width, int height, int samples_per_pass, int bvh_depth, size_t* bytes) {
    RX_REQUIRE(bytes, RX_ERR_INVALID, "bytes is NULL");
    RX_REQUIRE(bvh_depth >= 0 && bvh_depth <= RX_MAX_BVH_DEPTH, RX_ERR_CAPACITY, "BVH depth %d exceeds %d",
               bvh_depth, RX_MAX_BVH_DEPTH);
    RX_REQUIRE(width > 0 && height > 0 && samples_per_pass > 0, RX_ERR_INVALID,
               "width, height and samples_per_pass must be positive");
    long long P = (long long)width * height * samples_per_pass;
    RX_REQUIRE(P <= 0x7fffffffLL, RX_ERR_CAPACITY, "pass of %lld paths exceeds 2^31-1", P);
    *bytes = workspace_bytes_for(1, P, use_deep_stack(bvh_depth) || P < SMALL_PASS_PATHS, samples_per_pass, true);
    return RX_OK;
}

extern "C" int rx_render_workspace_size_for_scene(const rx_scene* scene, int width, int height, int samples_per_pass,
                                                  size_t* bytes) {
    RX_REQUIRE(scene && bytes, RX_ERR_INVALID, "NULL argument");
    RX_REQUIRE(scene->n_nodes >= 0 && scene->n_tris >= 0 && scene->n_prims >= 0, RX_ERR_INVALID,
               "negative scene size");
    const int depth = scene->n_tris > 0 ? scene->bvh_depth : 0;
    const int st = rx_render_workspace_size(width, height, samples_per_pass, depth, bytes);
    if (st) return st;
    *bytes += eye_bytes_for(*scene, width, height);
    return RX_OK;
}

extern "C" int rx_render(const rx_scene* scene, const rx_camera* camera, const rx_render_params* params,
                         int first_sample, int n_samples, void* d_workspace, size_t workspace_bytes,
                         const rx_accum* accum, rx_counters* d_counters, rx_profile* profile, void* stream) {
    RX_REQUIRE(scene && camera && params && accum && d_counters, RX_ERR_INVALID, "NULL argument");
    rx_job j;
    j.scene = *scene;
    j.camera = *camera;
    j.params = *params;
    j.accum = *accum;
    j.first_sample = first_sample;
    j.samples_left = n_samples;
    j.rays = 0;
    j.adaptive_error = -1.0f;
    j.adaptive_stopped = 0;
    long long P = 0, NP = 0;
    pass_plan pl;
    int32_t poff[2], qoff[2];
    int st = plan_jobs(&j, 1, &P, &NP, &pl, poff, qoff);
    if (st) return st;
    j.path_offset = 0;
    j.pixel_offset = 0;
    if (n_samples == 0) return RX_OK;
    ws_view w;
    st = carve_into(d_workspace, workspace_bytes, 1, P, NP, pl.deep, pl.max_spp, w, true);
    if (st) return st;
    // eye-relative copies for the camera rays, when the workspace has room
    // past the render's own part (rx_render_workspace_size_for_scene)
    const long long camera_rays = (long long)camera->width * camera->height * n_samples;
    const bool eye = eye_worthwhile(j.scene, camera_rays) &&
                     carve_eye(d_workspace, workspace_bytes, workspace_bytes_for(1, P, pl.deep, pl.max_spp, true),
                               j.scene, camera->width, camera->height, w);
    hipStream_t s = rx_stream(stream);
    k_set_job<<<1, 1, 0, s>>>(w.jobs, j, w.path_off, w.pix_off, poff[1], qoff[1]);
    RX_CHECK_LAUNCH();
    k_clear_ray_part<<<1, 64, 0, s>>>(w);
    RX_CHECK_LAUNCH();
    if (eye) {
        long long recs = j.scene.n_tris > 0 ? (long long)j.scene.n_nodes : 0;
        if (j.scene.n_tris > recs) recs = j.scene.n_tris;
        if (j.scene.n_prims > recs) recs = j.scene.n_prims;
        k_eye_prep<<<grid_for(recs > 0 ? recs : 1, BS, 16), BS, 0, s>>>(w);
        RX_CHECK_LAUNCH();
        if (w.imp_nodes) {
            k_imp_build<<<grid_for(2LL * w.imp_records + 1, BS, 16), BS, 0, s>>>(w);
            RX_CHECK_LAUNCH();
        }
        if (w.tile_prims) {
            const long long tiles = (long long)w.tile_cols * ((camera->height + 7) >> 3);
            k_prim_tiles<<<grid_for(tiles, BS, 16), BS, 0, s>>>(w);
            RX_CHECK_LAUNCH();
        }
    }
    stage_timer tm;
    tm.on = (params->flags & RX_FLAG_PROFILE) != 0;
    tm.s = s;
    int spp = params->samples_per_pass;
    for (int done = 0; done < n_samples; done += spp) {
        st = enqueue(w, pl, d_counters, tm, s);
        if (st) return st;
        if (profile) profile->n_passes += 1;
        if (tm.on) {
            st = tm.finish(profile);
            if (st) return st;
        }
    }
    return RX_OK;
}

extern "C" int rx_radix_sort_scratch_size(long long cap, size_t* bytes) {
    RX_REQUIRE(bytes && cap >= 0 && cap <= 0x7fffffffLL, RX_ERR_INVALID, "capacity must be in 0..2^31-1");
    *bytes = sort_scratch_words(cap) * sizeof(uint32_t);
    return RX_OK;
}

extern "C" int rx_radix_sort_pairs(uint32_t* d_keys, uint32_t* d_vals, uint32_t* d_keys2, uint32_t* d_vals2,
                                   const int32_t* d_count, long long cap, int min_n, int key_bits, void* d_scratch,
                                   size_t scratch_bytes, void* stream) {
    RX_REQUIRE(d_keys && d_vals && d_keys2 && d_vals2 && d_count && d_scratch, RX_ERR_INVALID, "NULL argument");
    RX_REQUIRE(cap >= 1 && cap <= 0x7fffffffLL, RX_ERR_INVALID, "capacity must be in 1..2^31-1");
    RX_REQUIRE(key_bits >= 1 && key_bits <= 32, RX_ERR_INVALID, "key_bits must be in 1..32");
    RX_REQUIRE(scratch_bytes >= sort_scratch_words(cap) * sizeof(uint32_t), RX_ERR_WORKSPACE,
               "sort scratch too small");
    uint32_t *rk = nullptr, *rv = nullptr;
    RX_TRY_HIP(enqueue_ray_sort(d_keys, d_vals, d_keys2, d_vals2, d_count, cap, min_n < 0 ? 0 : min_n,
                                (unsigned)key_bits, (uint32_t*)d_scratch, rx_stream(stream), &rk, &rv));
    if (rk != d_keys) {
        k_sort_copy<<<grid_for(cap, 256, 4), 256, 0, rx_stream(stream)>>>(rk, rv, d_keys, d_vals, d_count, cap);
        RX_CHECK_LAUNCH();
    }
    return RX_OK;
}

extern "C" int rx_batch_workspace_size(const rx_job* jobs, int n_jobs, size_t* bytes) {
    RX_REQUIRE(bytes, RX_ERR_INVALID, "bytes is NULL");
    long long P = 0, NP = 0;
    pass_plan pl;
    int st = plan_jobs(jobs, n_jobs, &P, &NP, &pl, nullptr, nullptr);
    if (st) return st;
    *bytes = workspace_bytes_for(n_jobs, P, pl.deep, pl.max_spp);
    return RX_OK;
}

extern "C" int rx_batch_create(const rx_job* jobs, int n_jobs, void* d_workspace, size_t workspace_bytes,
                               int use_graph, rx_batch** out) {
    RX_REQUIRE(out, RX_ERR_INVALID, "out is NULL");
    *out = nullptr;
    RX_REQUIRE(jobs && n_jobs >= 1 && n_jobs <= MAX_JOBS, RX_ERR_INVALID, "need 1..%d jobs", MAX_JOBS);
    int32_t* poff = (int32_t*)malloc(sizeof(int32_t) * (n_jobs + 1));
    int32_t* qoff = (int32_t*)malloc(sizeof(int32_t) * (n_jobs + 1));
    rx_job* hj = (rx_job*)malloc(sizeof(rx_job) * n_jobs);
    if (!poff || !qoff || !hj) {
        free(poff);
        free(qoff);
        free(hj);
        return rx::set_error(RX_ERR_INVALID, "out of host memory");
    }
    long long P = 0, NP = 0;
    pass_plan pl;
    int st = plan_jobs(jobs, n_jobs, &P, &NP, &pl, poff, qoff);
    ws_view w;
    if (!st) st = carve_into(d_workspace, workspace_bytes, n_jobs, P, NP, pl.deep, pl.max_spp, w);
    if (!st) {
        for (int i = 0; i < n_jobs; ++i) {
            hj[i] = jobs[i];
            hj[i].path_offset = poff[i];
            hj[i].pixel_offset = qoff[i];
            hj[i].adaptive_error = -1.0f;
            hj[i].adaptive_stopped = 0;
        }
        hipError_t e = hipMemcpy(w.jobs, hj, sizeof(rx_job) * n_jobs, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(w.path_off, poff, sizeof(int32_t) * (n_jobs + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(w.pix_off, qoff, sizeof(int32_t) * (n_jobs + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) {
            k_clear_ray_part<<<grid_for((long long)NPART * n_jobs, 256, 4), 256, 0, nullptr>>>(w);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) st = rx::set_error(RX_ERR_HIP, "job upload failed: %s", hipGetErrorString(e));
    }
    free(poff);
    free(qoff);
    free(hj);
    if (st) return st;
    rx_batch* b = new (std::nothrow) rx_batch;
    RX_REQUIRE(b, RX_ERR_INVALID, "out of host memory");
    b->w = w;
    b->plan = pl;
    b->n_jobs = n_jobs;
    b->use_graph = use_graph != 0;
    b->exec = nullptr;
    b->captured = false;
    b->graph_ctr = nullptr;
    b->graph_nodes[0] = b->graph_nodes[1] = b->graph_nodes[2] = -1;
    *out = b;
    return RX_OK;
}

extern "C" int rx_batch_run(rx_batch* b, int passes, rx_counters* d_counters, rx_profile* profile, void* stream) {
    RX_REQUIRE(b && d_counters, RX_ERR_INVALID, "NULL argument");
    RX_REQUIRE(passes >= 0, RX_ERR_INVALID, "passes must be >= 0");
    hipStream_t s = rx_stream(stream);
    bool prof = profile != nullptr;
    if (b->use_graph && !prof) {
        if (b->captured && b->graph_ctr != d_counters) {
            (void)hipGraphExecDestroy(b->exec);
            b->captured = false;
        }
        if (!b->captured) {
            hipStream_t cap;
            RX_TRY_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
            stage_timer off;
            off.on = false;
            hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
            int st = RX_OK;
            if (e == hipSuccess) st = enqueue(b->w, b->plan, d_counters, off, cap);
            hipGraph_t g = nullptr;
            hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(cap, &g) : e;
            (void)hipStreamDestroy(cap);
            if (st || e != hipSuccess || e2 != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (st) return st;
                return rx::set_error(RX_ERR_HIP, "graph capture failed: %s / %s", hipGetErrorString(e),
                                     hipGetErrorString(e2));
            }
            hipError_t e3 = count_graph_nodes(g, b->graph_nodes);
            if (e3 == hipSuccess) e3 = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e3 != hipSuccess)
                return rx::set_error(RX_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(e3));
            b->captured = true;
            b->graph_ctr = d_counters;
        }
        for (int k = 0; k < passes; ++k) RX_TRY_HIP(hipGraphLaunch(b->exec, s));
        return RX_OK;
    }
    stage_timer tm;
    tm.on = prof;
    tm.s = s;
    for (int k = 0; k < passes; ++k) {
        int st = enqueue(b->w, b->plan, d_counters, tm, s);
        if (st) return st;
        if (profile) profile->n_passes += 1;
        if (tm.on) {
            st = tm.finish(profile);
            if (st) return st;
        }
    }
    return RX_OK;
}

extern "C" int rx_batch_graph_nodes(const rx_batch* b, int* kernels, int* memsets, int* others) {
    RX_REQUIRE(b && kernels && memsets && others, RX_ERR_INVALID, "NULL argument");
    *kernels = b->graph_nodes[0];
    *memsets = b->graph_nodes[1];
    *others = b->graph_nodes[2];
    return RX_OK;
}

extern "C" int rx_batch_read_jobs(rx_batch* b, rx_job* jobs_out, void* stream) {
    RX_REQUIRE(b && jobs_out, RX_ERR_INVALID, "NULL argument");
    hipStream_t s = rx_stream(stream);
    RX_TRY_HIP(hipStreamSynchronize(s));
    RX_TRY_HIP(hipMemcpy(jobs_out, b->w.jobs, sizeof(rx_job) * b->n_jobs, hipMemcpyDeviceToHost));
    return RX_OK;
}

extern "C" int rx_batch_destroy(rx_batch* b) {
    if (!b) return RX_OK;
    if (b->captured) (void)hipGraphExecDestroy(b->exec);
    delete b;
    return RX_OK;
}

// ---------------------------------------------------------------- no-allocation render
// rx_render with the workspace in a library-owned device array instead of
// a caller buffer, for callers that cannot allocate (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).  The array is
// a file-scope device symbol: each GPU holds its own copy until the
// process ends.  Its device address comes from hipGetSymbolAddress, once
// per device on that device's first call (the array's name is only a host
// stand-in), and only that address is handed to the kernels.  The kernels
// clear every counter, queue length and partial sum they read before use
// (k_set_job, k_clear_ray_part, k_pass_begin), so the array's previous
// contents do not matter; calls that overlap on one GPU share it and race.
static __device__ __align__(256) unsigned char scratch[RX_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_ptr[MAX_DEVICES] = {nullptr};

int scratch_address(void** out) {
    int dev = 0;
    RX_TRY_HIP(hipGetDevice(&dev));
    RX_REQUIRE(dev >= 0 && dev < MAX_DEVICES, RX_ERR_INVALID, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void* p = nullptr;
        RX_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RX_REQUIRE(p, RX_ERR_HIP, "no device address for the no-allocation scratch array");
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return RX_OK;
}

}  // namespace

extern "C" int rx_render_noalloc(const rx_scene* scene, const rx_camera* camera, const rx_render_params* params,
                                 int first_sample, int n_samples, const rx_accum* accum, rx_counters* d_counters,
                                 void* stream) {
    RX_REQUIRE(scene && camera && params && accum && d_counters, RX_ERR_INVALID, "NULL argument");
    RX_REQUIRE(!(params->flags & RX_FLAG_PROFILE), RX_ERR_INVALID,
               "rx_render_noalloc does not time stages (RX_FLAG_PROFILE)");
    // the workspace this pass size needs, from the same plan rx_render makes
    rx_job j;
    j.scene = *scene;
    j.camera = *camera;
    j.params = *params;
    j.accum = *accum;
    j.first_sample = first_sample;
    j.samples_left = n_samples;
    long long P = 0, NP = 0;
    pass_plan pl;
    int st = plan_jobs(&j, 1, &P, &NP, &pl, nullptr, nullptr);
    if (st) return st;
    const size_t need = workspace_bytes_for(1, P, pl.deep, pl.max_spp, true);
    RX_REQUIRE(need <= RX_NOALLOC_SCRATCH_BYTES, RX_ERR_CAPACITY,
               "a pass of %lld paths needs %zu bytes of workspace; rx_render_noalloc has %zu (use fewer "
               "samples_per_pass)", P, need, (size_t)RX_NOALLOC_SCRATCH_BYTES);
    void* ptr = nullptr;
    st = scratch_address(&ptr);
    if (st) return st;
    return rx_render(scene, camera, params, first_sample, n_samples, ptr, RX_NOALLOC_SCRATCH_BYTES, accum,
                     d_counters, nullptr, stream);
}
