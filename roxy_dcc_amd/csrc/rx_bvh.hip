// rx_bvh.hip — GPU construction of the reference's Morton BVH.
//
// rendering/taichi/bvh_builder.py builds the tree on the host: padded
// triangle boxes and centroids (:363-367), 10-bit Morton codes of the
// normalised centroids (:416-443), a stable argsort (:426), then a
// recursive count-median split (:445-478) whose shape depends only on the
// triangle count.  Here every step is a kernel:
//   prep      per-triangle padded box and centroid           (exact fp32 ops)
//   reduce    centroid bounds                                 (min/max, exact)
//   morton    codes with the reference's divide/clip/truncate
//   sort      stable LSD radix sort of (code, index)          (rocPRIM)
//   emit      one launch per tree level: node ids in closed form
//   bounds    one launch per level, deepest first: leaf boxes from their
//             triangles, interior boxes from their children
//   pack      64-byte two-child nodes + root header
// Node ids, links, leaf ranges and boxes equal the reference's bit for bit.
#include <rocprim/device/device_radix_sort.hpp>

#include "rx_common.h"
#include "rx_device.h"

namespace {

constexpr int BS = 256;
constexpr int REDUCE_BLOCKS = 1024;

// Leaves of the count-median tree over c items with leaves of at most L.
// At depth k every node holds floor(c/2^k) or ceil(c/2^k) items (c mod 2^k
// of them the ceiling).  At the first depth k* where floor(c/2^k*) <= L all
// nodes exist; those holding L+1 (only possible when floor == L) split once
// more into two leaves.
__host__ __device__ inline long long median_leaves(long long c, int L) {
    if (c <= L) return 1;
    for (int k = 1;; ++k) {
        long long q = c >> k;
        if (q <= L) {
            long long r = c & ((1LL << k) - 1);
            return (1LL << k) + (q == L ? r : 0);
        }
    }
}

__host__ __device__ inline long long median_size(long long c, int L) { return 2 * median_leaves(c, L) - 1; }

__host__ inline int median_depth(long long c, int L) {
    if (c <= L) return 0;
    for (int k = 1;; ++k) {
        long long q = c >> k;
        if (q <= L) {
            long long r = c & ((1LL << k) - 1);
            return k + ((q == L && r > 0) ? 1 : 0);
        }
    }
}

__global__ void k_prep(const float* __restrict__ v0, const float* __restrict__ v1, const float* __restrict__ v2,
                       int n, float* __restrict__ blo, float* __restrict__ bhi, float* __restrict__ cen) {
    const float eps = 1e-4f;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float x0 = v0[3 * i + a], x1 = v1[3 * i + a], x2 = v2[3 * i + a];
            blo[3 * i + a] = fminf(fminf(x0, x1), x2) - eps;
            bhi[3 * i + a] = fmaxf(fmaxf(x0, x1), x2) + eps;
            cen[3 * i + a] = ((x0 + x1) + x2) / 3.0f;
        }
    }
}

__global__ void k_cen_bounds_partial(const float* __restrict__ cen, int n, float* __restrict__ part) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float x = cen[3 * i + a];
            lo[a] = fminf(lo[a], x);
            hi[a] = fmaxf(hi[a], x);
        }
    }
    __shared__ float s[6][BS];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ void k_cen_bounds_final(const float* __restrict__ part, int nb, float* __restrict__ out) {
    __shared__ float s[6][BS];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += BS) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            v[a] = fminf(v[a], part[6 * b + a]);
            v[3 + a] = fmaxf(v[3 + a], part[6 * b + 3 + a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 6; ++a) s[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

// bvh_builder.py:437-443
__device__ __forceinline__ uint32_t expand_bits10(uint32_t v) {
    v &= 0x000003ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// bvh_builder.py:416-434: ((c - mins) / extent) clipped to [0, 1], times
// 1023 in fp32, truncated; extent <= 1e-12 replaced by 1.
__global__ void k_morton(const float* __restrict__ cen, int n, const float* __restrict__ bounds,
                         uint32_t* __restrict__ codes, int32_t* __restrict__ idx) {
    float mn[3], ext[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        mn[a] = bounds[a];
        float e = bounds[3 + a] - bounds[a];
        ext[a] = (e <= 1e-12f) ? 1.0f : e;
    }
    const float scale = 1023.0f;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        uint32_t q[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float x = (cen[3 * i + a] - mn[a]) / ext[a];
            x = fminf(fmaxf(x, 0.0f), 1.0f);
            q[a] = (uint32_t)(x * scale);
        }
        codes[i] = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
        idx[i] = i;
    }
}

__global__ void k_root(int3* lv, int n) { lv[0] = make_int3(0, n, 0); }

// One tree level of the preorder emission (bvh_builder.py:445-478).  Entry
// j of level k is (start, count, node id); count 0 marks an empty slot.
// Its children go to slots 2j and 2j+1 of level k+1.
__global__ void k_emit_level(const int3* __restrict__ lv, int n_slots, int3* __restrict__ next, int L,
                             int32_t* __restrict__ left, int32_t* __restrict__ right,
                             int32_t* __restrict__ tri_start, int32_t* __restrict__ tri_count) {
    int j = blockIdx.x * BS + threadIdx.x;
    if (j >= n_slots) return;
    int3 e = lv[j];
    int start = e.x, count = e.y, id = e.z;
    if (count == 0) {
        if (next) {
            next[2 * j] = make_int3(0, 0, -1);
            next[2 * j + 1] = make_int3(0, 0, -1);
        }
        return;
    }
    if (count <= L) {
        left[id] = -1;
        right[id] = -1;
        tri_start[id] = start;
        tri_count[id] = count;
        if (next) {
            next[2 * j] = make_int3(0, 0, -1);
            next[2 * j + 1] = make_int3(0, 0, -1);
        }
    } else {
        int lc = count / 2;
        int lid = id + 1;
        int rid = id + 1 + (int)median_size(lc, L);
        left[id] = lid;
        right[id] = rid;
        tri_start[id] = -1;
        tri_count[id] = 0;
        next[2 * j] = make_int3(start, lc, lid);
        next[2 * j + 1] = make_int3(start + lc, count - lc, rid);
    }
}

// Boxes of one level, deepest level first (bvh_builder.py:452, :480-486,
// :546-550).  Leaves: min/max of their triangles' padded boxes; interior
// nodes: min/max of their children, which the previous launch finished.
__global__ void k_bounds_level(const int3* __restrict__ lv, int n_slots, int L, const int32_t* __restrict__ order,
                               const float* __restrict__ blo, const float* __restrict__ bhi,
                               const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                               float* __restrict__ amin, float* __restrict__ amax) {
    int j = blockIdx.x * BS + threadIdx.x;
    if (j >= n_slots) return;
    int3 e = lv[j];
    int start = e.x, count = e.y, id = e.z;
    if (count == 0) return;
    float lo[3], hi[3];
    if (count <= L) {
        lo[0] = lo[1] = lo[2] = INFINITY;
        hi[0] = hi[1] = hi[2] = -INFINITY;
        for (int i = start; i < start + count; ++i) {
            int t = order[i];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = fminf(lo[a], blo[3 * t + a]);
                hi[a] = fmaxf(hi[a], bhi[3 * t + a]);
            }
        }
    } else {
        int l = left[id], r = right[id];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = fminf(amin[3 * l + a], amin[3 * r + a]);
            hi[a] = fmaxf(amax[3 * l + a], amax[3 * r + a]);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        amin[3 * id + a] = lo[a];
        amax[3 * id + a] = hi[a];
    }
}

__device__ __forceinline__ int child_link(int c, const int32_t* __restrict__ tri_start,
                                          const int32_t* __restrict__ tri_count) {
    int cnt = tri_count[c];
    if (cnt > 0) return ~((tri_start[c] << 8) | cnt);
    return c;
}

__global__ void k_pack(const float* __restrict__ amin, const float* __restrict__ amax,
                       const int32_t* __restrict__ left, const int32_t* __restrict__ right,
                       const int32_t* __restrict__ tri_start, const int32_t* __restrict__ tri_count, int n_nodes,
                       int depth, rx_node* __restrict__ nodes, rx_bvh_header* __restrict__ hdr) {
    int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n_nodes) return;
    rx_node nd;
    int l = left[i], r = right[i];
    if (l >= 0) {
        for (int a = 0; a < 3; ++a) {
            nd.box[a] = amin[3 * l + a];
            nd.box[3 + a] = amax[3 * l + a];
            nd.box[6 + a] = amin[3 * r + a];
            nd.box[9 + a] = amax[3 * r + a];
        }
        nd.link[0] = child_link(l, tri_start, tri_count);
        nd.link[1] = child_link(r, tri_start, tri_count);
        nd.node[0] = l;
        nd.node[1] = r;
    } else {
        for (int a = 0; a < 3; ++a) {
            nd.box[a] = amin[3 * i + a];
            nd.box[3 + a] = amax[3 * i + a];
            nd.box[6 + a] = 0.0f;
            nd.box[9 + a] = 0.0f;
        }
        nd.link[0] = child_link(i, tri_start, tri_count);
        nd.link[1] = -1;
        nd.node[0] = -1;
        nd.node[1] = -1;
    }
    nodes[i] = nd;
    if (i == 0) {
        rx_bvh_header h;
        for (int a = 0; a < 3; ++a) {
            h.bmin[a] = amin[a];
            h.bmax[a] = amax[a];
        }
        h.root_link = child_link(0, tri_start, tri_count);
        h.depth = depth;
        *hdr = h;
    }
}

__global__ void k_pack_tris(const int32_t* __restrict__ order, int n, const float* __restrict__ v0,
                            const float* __restrict__ v1, const float* __restrict__ v2,
                            const float* __restrict__ n0, const float* __restrict__ n1,
                            const float* __restrict__ n2, const int32_t* __restrict__ mat,
                            const int32_t* __restrict__ has_uv, rx_tri* __restrict__ tris,
                            rx_tri_shade* __restrict__ shade) {
    using namespace rx;
    int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    int s = order[i];
    f3 a = ld3(v0 + 3 * s), b = ld3(v1 + 3 * s), c = ld3(v2 + 3 * s);
    f3 e1 = b - a, e2 = c - a;
    rx_tri t;
    t.v0[0] = a.x; t.v0[1] = a.y; t.v0[2] = a.z;
    t.e1[0] = e1.x; t.e1[1] = e1.y; t.e1[2] = e1.z;
    t.e2[0] = e2.x; t.e2[1] = e2.y; t.e2[2] = e2.z;
    t.pad[0] = t.pad[1] = t.pad[2] = 0.0f;
    tris[i] = t;
    if (shade) {
        rx_tri_shade sh;
        f3 m0, m1, m2;
        if (n0) {
            m0 = ld3(n0 + 3 * s);
            m1 = ld3(n1 + 3 * s);
            m2 = ld3(n2 + 3 * s);
        } else {
            m0 = m1 = m2 = normalize(cross(e1, e2));
        }
        sh.n0[0] = m0.x; sh.n0[1] = m0.y; sh.n0[2] = m0.z;
        sh.n1[0] = m1.x; sh.n1[1] = m1.y; sh.n1[2] = m1.z;
        sh.n2[0] = m2.x; sh.n2[1] = m2.y; sh.n2[2] = m2.z;
        sh.mat = mat ? mat[s] : 0;
        sh.has_uv = has_uv ? has_uv[s] : 0;
        sh.pad = 0;
        shade[i] = sh;
    }
}

struct build_ws {
    float* blo;
    float* bhi;
    float* cen;
    float* part;
    float* cb;
    uint32_t* codes_in;
    uint32_t* codes_out;
    int32_t* idx_in;
    int3* levels;
    void* sort_tmp;
    size_t sort_bytes;
};

int carve(rx_arena& ar, int n, int L, build_ws& w) {
    int depth = median_depth(n, L);
    w.blo = ar.take<float>(3 * (size_t)n);
    w.bhi = ar.take<float>(3 * (size_t)n);
    w.cen = ar.take<float>(3 * (size_t)n);
    w.part = ar.take<float>(6 * REDUCE_BLOCKS);
    w.cb = ar.take<float>(8);
    w.codes_in = ar.take<uint32_t>(n);
    w.codes_out = ar.take<uint32_t>(n);
    w.idx_in = ar.take<int32_t>(n);
    w.levels = ar.take<int3>(((size_t)2 << depth));
    size_t sb = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, sb, (uint32_t*)nullptr, (uint32_t*)nullptr, (int32_t*)nullptr,
                                             (int32_t*)nullptr, (unsigned int)n, 0u, 30u);
    if (e != hipSuccess) return rx::set_error(RX_ERR_HIP, "radix sort size query failed: %s", hipGetErrorString(e));
    w.sort_bytes = sb;
    w.sort_tmp = ar.take_bytes(sb);
    return RX_OK;
}

int grid_for(int n) {
    int g = (n + BS - 1) / BS;
    int cap = rx::device_cu_count() * 8;
    return g < 1 ? 1 : (g > cap ? cap : g);
}

}  // namespace

extern "C" int rx_bvh_morton_shape(int n_tris, int max_leaf, int* n_nodes, int* depth) {
    RX_REQUIRE(n_tris >= 0, RX_ERR_INVALID, "n_tris must be >= 0");
    RX_REQUIRE(max_leaf >= 1 && max_leaf <= RX_MAX_LEAF_TRIS, RX_ERR_INVALID, "max_leaf must be in [1, %d]",
               RX_MAX_LEAF_TRIS);
    if (n_tris == 0) {
        if (n_nodes) *n_nodes = 0;
        if (depth) *depth = 0;
        return RX_OK;
    }
    if (n_nodes) *n_nodes = (int)median_size(n_tris, max_leaf);
    if (depth) *depth = median_depth(n_tris, max_leaf);
    return RX_OK;
}

extern "C" int rx_bvh_workspace_size(int n_tris, int max_leaf, size_t* bytes) {
    RX_REQUIRE(bytes, RX_ERR_INVALID, "bytes is NULL");
    RX_REQUIRE(n_tris >= 0 && n_tris <= RX_MAX_TRIANGLES, RX_ERR_CAPACITY, "BVH has %d triangles, limit is %d",
               n_tris, RX_MAX_TRIANGLES);
    RX_REQUIRE(max_leaf >= 1 && max_leaf <= RX_MAX_LEAF_TRIS, RX_ERR_INVALID, "max_leaf must be in [1, %d]",
               RX_MAX_LEAF_TRIS);
    rx_arena ar{nullptr, 0, 0};
    build_ws w;
    int st = carve(ar, n_tris > 0 ? n_tris : 1, max_leaf, w);
    if (st) return st;
    *bytes = ar.off + 256;
    return RX_OK;
}

extern "C" int rx_bvh_build_morton(const float* d_v0, const float* d_v1, const float* d_v2, int n_tris, int max_leaf,
                                   void* d_workspace, size_t workspace_bytes, const rx_bvh_ref* ref,
                                   rx_node* d_nodes, rx_bvh_header* d_header, int32_t* d_order, void* stream) {
    RX_REQUIRE(n_tris >= 1 && n_tris <= RX_MAX_TRIANGLES, RX_ERR_CAPACITY,
               "BVH has %d triangles, need 1..%d", n_tris, RX_MAX_TRIANGLES);
    RX_REQUIRE(max_leaf >= 1 && max_leaf <= RX_MAX_LEAF_TRIS, RX_ERR_INVALID, "max_leaf must be in [1, %d]",
               RX_MAX_LEAF_TRIS);
    RX_REQUIRE(d_v0 && d_v1 && d_v2 && ref && d_nodes && d_header && d_order, RX_ERR_INVALID, "NULL argument");
    RX_REQUIRE(ref->aabb_min && ref->aabb_max && ref->left && ref->right && ref->tri_start && ref->tri_count,
               RX_ERR_INVALID, "NULL reference-layout array");
    int depth = median_depth(n_tris, max_leaf);
    RX_REQUIRE(depth <= RX_MAX_BVH_DEPTH, RX_ERR_CAPACITY, "BVH depth %d exceeds %d", depth, RX_MAX_BVH_DEPTH);
    int n_nodes = (int)median_size(n_tris, max_leaf);

    rx_arena ar{reinterpret_cast<char*>(d_workspace), workspace_bytes, 0};
    ar.off = (256 - (reinterpret_cast<uintptr_t>(d_workspace) & 255)) & 255;
    build_ws w;
    int st = carve(ar, n_tris, max_leaf, w);
    if (st) return st;
    RX_REQUIRE(ar.off <= workspace_bytes, RX_ERR_WORKSPACE, "BVH workspace too small: need %zu bytes, have %zu",
               ar.off, workspace_bytes);
    hipStream_t s = rx_stream(stream);
    int g = grid_for(n_tris);

    k_prep<<<g, BS, 0, s>>>(d_v0, d_v1, d_v2, n_tris, w.blo, w.bhi, w.cen);
    RX_CHECK_LAUNCH();
    int nb = g < REDUCE_BLOCKS ? g : REDUCE_BLOCKS;
    k_cen_bounds_partial<<<nb, BS, 0, s>>>(w.cen, n_tris, w.part);
    RX_CHECK_LAUNCH();
    k_cen_bounds_final<<<1, BS, 0, s>>>(w.part, nb, w.cb);
    RX_CHECK_LAUNCH();
    k_morton<<<g, BS, 0, s>>>(w.cen, n_tris, w.cb, w.codes_in, w.idx_in);
    RX_CHECK_LAUNCH();
    size_t sb = w.sort_bytes;
    RX_TRY_HIP(rocprim::radix_sort_pairs(w.sort_tmp, sb, w.codes_in, w.codes_out, w.idx_in, d_order,
                                         (unsigned int)n_tris, 0u, 30u, s));

    k_root<<<1, 1, 0, s>>>(w.levels, n_tris);
    RX_CHECK_LAUNCH();
    for (int k = 0; k <= depth; ++k) {
        int slots = 1 << k;
        int3* lv = w.levels + (slots - 1);
        int3* next = (k < depth) ? w.levels + (2 * slots - 1) : nullptr;
        k_emit_level<<<(slots + BS - 1) / BS, BS, 0, s>>>(lv, slots, next, max_leaf, ref->left, ref->right,
                                                          ref->tri_start, ref->tri_count);
        RX_CHECK_LAUNCH();
    }
    for (int k = depth; k >= 0; --k) {
        int slots = 1 << k;
        int3* lv = w.levels + (slots - 1);
        k_bounds_level<<<(slots + BS - 1) / BS, BS, 0, s>>>(lv, slots, max_leaf, d_order, w.blo, w.bhi, ref->left,
                                                            ref->right, ref->aabb_min, ref->aabb_max);
        RX_CHECK_LAUNCH();
    }
    k_pack<<<(n_nodes + BS - 1) / BS, BS, 0, s>>>(ref->aabb_min, ref->aabb_max, ref->left, ref->right, ref->tri_start,
                                                  ref->tri_count, n_nodes, depth, d_nodes, d_header);
    RX_CHECK_LAUNCH();
    return RX_OK;
}

extern "C" int rx_bvh_pack(const rx_bvh_ref* ref, int n_nodes, int depth, rx_node* d_nodes, rx_bvh_header* d_header,
                           void* stream) {
    RX_REQUIRE(ref && d_nodes && d_header, RX_ERR_INVALID, "NULL argument");
    RX_REQUIRE(n_nodes >= 1, RX_ERR_INVALID, "n_nodes must be >= 1");
    RX_REQUIRE(depth >= 0 && depth <= RX_MAX_BVH_DEPTH, RX_ERR_CAPACITY, "BVH depth %d exceeds %d", depth,
               RX_MAX_BVH_DEPTH);
    k_pack<<<(n_nodes + BS - 1) / BS, BS, 0, rx_stream(stream)>>>(ref->aabb_min, ref->aabb_max, ref->left,
                                                                  ref->right, ref->tri_start, ref->tri_count,
                                                                  n_nodes, depth, d_nodes, d_header);
    RX_CHECK_LAUNCH();
    return RX_OK;
}

// Per-corner UVs in leaf order, 6 floats per triangle.
__global__ void k_pack_uvs(const int32_t* __restrict__ order, int n, const float* __restrict__ uv0,
                           const float* __restrict__ uv1, const float* __restrict__ uv2, float* __restrict__ out) {
    int i = blockIdx.x * BS + threadIdx.x;
    if (i >= n) return;
    int s = order[i];
    float* o = out + 6 * (long long)i;
    o[0] = uv0[2 * s];
    o[1] = uv0[2 * s + 1];
    o[2] = uv1[2 * s];
    o[3] = uv1[2 * s + 1];
    o[4] = uv2[2 * s];
    o[5] = uv2[2 * s + 1];
}

extern "C" int rx_bvh_pack_uvs(const int32_t* d_order, int n_tris, const float* d_uv0, const float* d_uv1,
                               const float* d_uv2, float* d_tri_uv, void* stream) {
    RX_REQUIRE(n_tris >= 0 && n_tris <= RX_MAX_TRIANGLES, RX_ERR_CAPACITY, "BVH has %d triangles, limit is %d",
               n_tris, RX_MAX_TRIANGLES);
    if (n_tris == 0) return RX_OK;
    RX_REQUIRE(d_order && d_uv0 && d_uv1 && d_uv2 && d_tri_uv, RX_ERR_INVALID, "NULL argument");
    k_pack_uvs<<<(n_tris + BS - 1) / BS, BS, 0, rx_stream(stream)>>>(d_order, n_tris, d_uv0, d_uv1, d_uv2, d_tri_uv);
    RX_CHECK_LAUNCH();
    return RX_OK;
}

extern "C" int rx_bvh_pack_triangles(const int32_t* d_order, int n_tris, const float* d_v0, const float* d_v1,
                                     const float* d_v2, const float* d_n0, const float* d_n1, const float* d_n2,
                                     const int32_t* d_mat_idx, const int32_t* d_has_uv, rx_tri* d_tris,
                                     rx_tri_shade* d_shade, void* stream) {
    RX_REQUIRE(n_tris >= 0 && n_tris <= RX_MAX_TRIANGLES, RX_ERR_CAPACITY, "BVH has %d triangles, limit is %d",
               n_tris, RX_MAX_TRIANGLES);
    if (n_tris == 0) return RX_OK;
    RX_REQUIRE(d_order && d_v0 && d_v1 && d_v2 && d_tris, RX_ERR_INVALID, "NULL argument");
    RX_REQUIRE((d_n0 && d_n1 && d_n2) || (!d_n0 && !d_n1 && !d_n2), RX_ERR_INVALID,
               "normals must be all given or all NULL");
    k_pack_tris<<<(n_tris + BS - 1) / BS, BS, 0, rx_stream(stream)>>>(d_order, n_tris, d_v0, d_v1, d_v2, d_n0, d_n1,
                                                                      d_n2, d_mat_idx, d_has_uv, d_tris, d_shade);
    RX_CHECK_LAUNCH();
    return RX_OK;
}
