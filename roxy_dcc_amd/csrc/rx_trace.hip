// rx_trace.hip — batched closest-hit and any-hit queries over ray arrays:
// rendering/taichi/intersect.py:249-309 (scene_intersect) and :222-246
// (scene_occluded), one ray per lane, short traversal stack in LDS with
// a global spill column for deeper entries.
#include "rx_launch.h"

namespace {

using namespace rx;

template <int BS, bool COUNT>
__global__ __launch_bounds__(BS) void k_closest(rx_scene sc, const float* __restrict__ rays, int n,
                                                float* __restrict__ t_out, int32_t* __restrict__ id_out,
                                                int32_t* __restrict__ kind_out, float* __restrict__ u_out,
                                                float* __restrict__ v_out, rx_counters* __restrict__ ctr,
                                                int2* __restrict__ spill) {
    __shared__ int2 stack_mem[STACK_LDS * BS];
    lane_stack<BS, STACK_LDS> stk{stack_mem + threadIdx.x, spill + blockIdx.x * BS + threadIdx.x,
                                  (int)gridDim.x * BS};
    rx_bvh_header hdr;
    if (sc.n_tris > 0) hdr = gload(sc.header);
    trav_count cnt{0, 0, 0};
    const int stride = gridDim.x * BS;
    for (int base = xcd_block() * BS; base < n; base += stride) {
        int i = base + threadIdx.x;
        if (i < n) {
            f3 ro = mk(rays[i], rays[n + i], rays[2 * n + i]);
            f3 rd = mk(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
            hit_rec h = scene_closest<lane_stack<BS, STACK_LDS>, COUNT>(sc, hdr, ro, rd, stk, cnt);
            if (t_out) t_out[i] = h.t;
            if (id_out) id_out[i] = h.id;
            if (kind_out) kind_out[i] = h.kind;
            if (u_out) u_out[i] = h.u;
            if (v_out) v_out[i] = h.v;
        }
    }
    if (COUNT) {
        wave_add(&ctr->node_pops, cnt.pops);
        wave_add(&ctr->tri_tests, cnt.tris);
        wave_add(&ctr->prim_tests, cnt.prims);
    }
}

template <int BS, bool COUNT>
__global__ __launch_bounds__(BS) void k_any(rx_scene sc, const float* __restrict__ rays,
                                            const float* __restrict__ tmax, int n, int32_t* __restrict__ occ,
                                            rx_counters* __restrict__ ctr, int2* __restrict__ spill) {
    __shared__ int2 stack_mem[STACK_LDS * BS];
    lane_stack<BS, STACK_LDS> stk{stack_mem + threadIdx.x, spill + blockIdx.x * BS + threadIdx.x,
                                  (int)gridDim.x * BS};
    rx_bvh_header hdr;
    if (sc.n_tris > 0) hdr = gload(sc.header);
    trav_count cnt{0, 0, 0};
    const int stride = gridDim.x * BS;
    for (int base = xcd_block() * BS; base < n; base += stride) {
        int i = base + threadIdx.x;
        if (i < n) {
            f3 ro = mk(rays[i], rays[n + i], rays[2 * n + i]);
            f3 rd = mk(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
            occ[i] = scene_occluded<lane_stack<BS, STACK_LDS>, COUNT>(sc, hdr, ro, rd, tmax[i], stk, cnt) ? 1 : 0;
        }
    }
    if (COUNT) {
        wave_add(&ctr->shadow_node_pops, cnt.pops);
        wave_add(&ctr->shadow_tri_tests, cnt.tris);
        wave_add(&ctr->shadow_prim_tests, cnt.prims);
    }
}

int check_scene(const rx_scene* sc) {
    RX_REQUIRE(sc, RX_ERR_INVALID, "scene is NULL");
    RX_REQUIRE(sc->n_prims >= 0 && sc->n_tris >= 0 && sc->n_lights >= 0, RX_ERR_INVALID, "negative scene size");
    RX_REQUIRE(sc->n_prims == 0 || sc->prims, RX_ERR_INVALID, "scene has primitives but prims is NULL");
    RX_REQUIRE(rx_scene_aligned(*sc), RX_ERR_INVALID, "scene arrays must be 16-byte aligned");
    if (sc->n_tris > 0) {
        RX_REQUIRE(sc->nodes && sc->header && sc->tris, RX_ERR_INVALID, "scene has triangles but no BVH");
        RX_REQUIRE(sc->bvh_depth >= 0 && sc->bvh_depth <= RX_MAX_BVH_DEPTH, RX_ERR_CAPACITY,
                   "BVH depth %d exceeds %d", sc->bvh_depth, RX_MAX_BVH_DEPTH);
    }
    return RX_OK;
}

}  // namespace

namespace rx {

int query_spill(bool deep, int2** out) {
    static int2* buf[64] = {nullptr};
    static size_t cap[64] = {0};
    int dev = 0;
    RX_TRY_HIP(hipGetDevice(&dev));
    RX_REQUIRE(dev >= 0 && dev < 64, RX_ERR_INVALID, "device index %d out of range", dev);
    size_t need = trav_spill_entries(deep) * sizeof(int2);
    if (cap[dev] < need) {
        if (buf[dev]) RX_TRY_HIP(hipFree(buf[dev]));
        buf[dev] = nullptr;
        cap[dev] = 0;
        RX_TRY_HIP(hipMalloc(reinterpret_cast<void**>(&buf[dev]), need));
        cap[dev] = need;
    }
    *out = buf[dev];
    return RX_OK;
}

}  // namespace rx

extern "C" int rx_closest_hit(const rx_scene* scene, const float* d_rays, int n, float* d_t, int32_t* d_id,
                              int32_t* d_kind, float* d_u, float* d_v, rx_counters* d_counters, void* stream) {
    int st = check_scene(scene);
    if (st) return st;
    RX_REQUIRE(n >= 0, RX_ERR_INVALID, "n must be >= 0");
    if (n == 0) return RX_OK;
    RX_REQUIRE(d_rays, RX_ERR_INVALID, "rays is NULL");
    hipStream_t s = rx_stream(stream);
    bool deep = scene->n_tris > 0 && use_deep_stack(scene->bvh_depth);
    int2* spill = nullptr;
    st = query_spill(deep, &spill);
    if (st) return st;
    int g = trav_grid(n, deep);
    if (deep) {
        if (d_counters)
            k_closest<BS_DEEP, true><<<g, BS_DEEP, 0, s>>>(*scene, d_rays, n, d_t, d_id, d_kind, d_u, d_v,
                                                           d_counters, spill);
        else
            k_closest<BS_DEEP, false><<<g, BS_DEEP, 0, s>>>(*scene, d_rays, n, d_t, d_id, d_kind, d_u, d_v,
                                                            d_counters, spill);
    } else {
        if (d_counters)
            k_closest<BS_SHALLOW, true><<<g, BS_SHALLOW, 0, s>>>(*scene, d_rays, n, d_t, d_id, d_kind, d_u, d_v,
                                                                 d_counters, spill);
        else
            k_closest<BS_SHALLOW, false><<<g, BS_SHALLOW, 0, s>>>(*scene, d_rays, n, d_t, d_id, d_kind, d_u, d_v,
                                                                  d_counters, spill);
    }
    RX_CHECK_LAUNCH();
    return RX_OK;
}

extern "C" int rx_any_hit(const rx_scene* scene, const float* d_rays, const float* d_tmax, int n,
                          int32_t* d_occluded, rx_counters* d_counters, void* stream) {
    int st = check_scene(scene);
    if (st) return st;
    RX_REQUIRE(n >= 0, RX_ERR_INVALID, "n must be >= 0");
    if (n == 0) return RX_OK;
    RX_REQUIRE(d_rays && d_tmax && d_occluded, RX_ERR_INVALID, "NULL argument");
    hipStream_t s = rx_stream(stream);
    bool deep = scene->n_tris > 0 && use_deep_stack(scene->bvh_depth);
    int2* spill = nullptr;
    st = query_spill(deep, &spill);
    if (st) return st;
    int g = trav_grid(n, deep);
    if (deep) {
        if (d_counters)
            k_any<BS_DEEP, true><<<g, BS_DEEP, 0, s>>>(*scene, d_rays, d_tmax, n, d_occluded, d_counters, spill);
        else
            k_any<BS_DEEP, false><<<g, BS_DEEP, 0, s>>>(*scene, d_rays, d_tmax, n, d_occluded, d_counters, spill);
    } else {
        if (d_counters)
            k_any<BS_SHALLOW, true><<<g, BS_SHALLOW, 0, s>>>(*scene, d_rays, d_tmax, n, d_occluded, d_counters,
                                                             spill);
        else
            k_any<BS_SHALLOW, false><<<g, BS_SHALLOW, 0, s>>>(*scene, d_rays, d_tmax, n, d_occluded, d_counters,
                                                              spill);
    }
    RX_CHECK_LAUNCH();
    return RX_OK;
}
