// rx_bvh_object.hip — GPU construction of core/bvh.py's BVHNode tree.
//
// core/bvh.py:2-33 builds an object-linked tree on the host.  A node's box
// is the union of its triangles' Triangle.local_bounds (scene/mesh.py:68-76:
// fp64 min/max of the three vertices, -/+ 1e-4).  A node holding more than
// max_leaf_size triangles (the root's argument; every other node gets the
// default 4, since the recursion does not forward it, :12-13) at depth 20
// or less sorts its triangles stably by the fp64 centroid
// ((v0 + v1) + v2) * (1 / 3) (scene/mesh.py:78-79, core/vectors.py:112-113)
// along the longest box axis (ties: x, then y, :41-48) and gives the first
// len // 2 to the left child.
//
// The split position depends only on the count, so the topology — preorder
// node ids, each node's range of positions, which nodes are leaves — is
// laid out on the host.  The device does what depends on the geometry, one
// tree level at a time:
//   prep     fp64 boxes and centroids per triangle, identity order
//   bounds   segmented min/max over each node's range (rocPRIM), the box
//            rounded to fp32 into the reference-layout arrays, the split
//            axis of every node that splits
//   sort     stable sort of every splitting node's range by its centroid
//            coordinate: two stable LSD radix passes, first by the
//            centroid key, then by the node's first position
//   pack     the 64-byte nodes and header of the Morton builder
// Output: the reference-layout arrays of fields.py:50-56 in preorder (the
// fp64 boxes rounded to fp32), the packed nodes, and the leaf order.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_segmented_reduce.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <map>
#include <utility>
#include <vector>

#include "rx_common.h"

namespace {

constexpr int BS = 256;
constexpr int OM_CHILD_LEAF = 4;   // core/bvh.py:2 default, used below the root
constexpr int OM_MAX_DEPTH = 20;   // core/bvh.py:5: deeper nodes are leaves

struct om_node {
    int32_t start, count, depth, left, right, leaf;
};

inline bool om_is_leaf(long long count, int depth, int max_leaf) {
    int limit = depth == 0 ? max_leaf : OM_CHILD_LEAF;
    return count <= limit || depth > OM_MAX_DEPTH;
}

// Preorder layout of the tree over n triangles: the recursion of
// core/bvh.py:2-13 on counts alone, left subtree first.
void om_layout(int n, int max_leaf, std::vector<om_node>& out) {
    struct item {
        int32_t start, count, depth, parent, side;
    };
    out.clear();
    std::vector<item> stack;
    stack.push_back({0, n, 0, -1, 0});
    while (!stack.empty()) {
        item it = stack.back();
        stack.pop_back();
        int id = (int)out.size();
        bool leaf = om_is_leaf(it.count, it.depth, max_leaf);
        out.push_back({it.start, it.count, it.depth, -1, -1, leaf ? 1 : 0});
        if (it.parent >= 0) {
            if (it.side == 0)
                out[it.parent].left = id;
            else
                out[it.parent].right = id;
        }
        if (!leaf) {
            int mid = it.count / 2;
            stack.push_back({it.start + mid, it.count - mid, it.depth + 1, id, 1});
            stack.push_back({it.start, mid, it.depth + 1, id, 0});
        }
    }
}

// Node count, depth and largest leaf of the subtree over `count`
// triangles rooted at `depth`; at a given depth the counts are the floor
// or ceiling of n / 2^depth, so memoising on (count, depth) keeps this
// to a few entries per level.
struct om_shape_t {
    long long nodes;
    int depth;
    long long max_leaf_count;
};

om_shape_t om_shape(long long count, int depth, int max_leaf, std::map<std::pair<long long, int>, om_shape_t>& memo) {
    auto key = std::make_pair(count, depth);
    auto f = memo.find(key);
    if (f != memo.end()) return f->second;
    om_shape_t r;
    if (om_is_leaf(count, depth, max_leaf)) {
        r = {1, depth, count};
    } else {
        om_shape_t a = om_shape(count / 2, depth + 1, max_leaf, memo);
        om_shape_t b = om_shape(count - count / 2, depth + 1, max_leaf, memo);
        r = {1 + a.nodes + b.nodes, a.depth > b.depth ? a.depth : b.depth,
             a.max_leaf_count > b.max_leaf_count ? a.max_leaf_count : b.max_leaf_count};
    }
    memo[key] = r;
    return r;
}

om_shape_t shape_of(long long n, int max_leaf) {
    std::map<std::pair<long long, int>, om_shape_t> memo;
    return om_shape(n, 0, max_leaf, memo);
}

struct box6 {
    double lo[3];
    double hi[3];
};

// Identity of box_union: the reduction starts from it.
constexpr box6 EMPTY_BOX{{INFINITY, INFINITY, INFINITY}, {-INFINITY, -INFINITY, -INFINITY}};

struct box_union {
    __host__ __device__ box6 operator()(const box6& a, const box6& b) const {
        box6 r;
        for (int k = 0; k < 3; ++k) {
            r.lo[k] = b.lo[k] < a.lo[k] ? b.lo[k] : a.lo[k];
            r.hi[k] = b.hi[k] > a.hi[k] ? b.hi[k] : a.hi[k];
        }
        return r;
    }
};

// Box of the triangle at a position of the current order.
struct box_at {
    const int32_t* perm;
    const double* lo;
    const double* hi;
    __host__ __device__ box6 operator()(int pos) const {
        int t = perm[pos];
        box6 b;
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = lo[3 * t + k];
            b.hi[k] = hi[3 * t + k];
        }
        return b;
    }
};

// Per triangle: fp64 box (scene/mesh.py:68-76) and centroid
// ((v0 + v1) + v2) * (1 / 3); the order starts as the identity.
__global__ void k_om_prep(const double* __restrict__ v0, const double* __restrict__ v1,
                          const double* __restrict__ v2, int n, double* __restrict__ lo, double* __restrict__ hi,
                          double* __restrict__ cen, int32_t* __restrict__ perm) {
    const double eps = 1e-4;
    const double third = 1.0 / 3.0;
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        for (int k = 0; k < 3; ++k) {
            double a = v0[3 * i + k], b = v1[3 * i + k], c = v2[3 * i + k];
            double mn = a, mx = a;
            if (b < mn) mn = b;
            if (c < mn) mn = c;
            if (b > mx) mx = b;
            if (c > mx) mx = c;
            lo[3 * i + k] = mn - eps;
            hi[3 * i + k] = mx + eps;
            cen[3 * i + k] = ((a + b) + c) * third;
        }
        perm[i] = i;
    }
}

// The level's boxes into the reference layout (rounded to fp32) and the
// split axis of the nodes that split: longest extent, ties to x then y
// (core/bvh.py:41-48).
__global__ void k_om_boxes(const box6* __restrict__ boxes, const int32_t* __restrict__ node, const int32_t* __restrict__ split,
                           int n_seg, float* __restrict__ amin, float* __restrict__ amax, int32_t* __restrict__ axis) {
    int s = blockIdx.x * BS + threadIdx.x;
    if (s >= n_seg) return;
    const box6 b = boxes[s];
    const int id = node[s];
    for (int k = 0; k < 3; ++k) {
        amin[3 * id + k] = __double2float_rn(b.lo[k]);
        amax[3 * id + k] = __double2float_rn(b.hi[k]);
    }
    int ax = 0;
    if (split[s]) {
        double ex = b.hi[0] - b.lo[0], ey = b.hi[1] - b.lo[1], ez = b.hi[2] - b.lo[2];
        if (ex >= ey && ex >= ez)
            ax = 0;
        else if (ey >= ez)
            ax = 1;
        else
            ax = 2;
    }
    axis[s] = ax;
}

// Ascending unsigned key with the order of the doubles; -0.0 is mapped
// to +0.0 first (Python's sort compares them equal, so their relative
// order must stay).
__device__ __forceinline__ unsigned long long order_key(double x) {
    if (x == 0.0) x = 0.0;
    unsigned long long b = (unsigned long long)__double_as_longlong(x);
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

// Sort keys of every position for one level.  A position inside a node
// that splits gets (centroid key along the node's axis, node's first
// position); any other position (inside a leaf) gets (0, its own
// position), so a stable sort by the centroid key and then by the second
// key reorders each splitting node's range by centroid, ties in their
// current order, and leaves everything else where it is.
__global__ void k_om_keys(const int32_t* __restrict__ perm, const double* __restrict__ cen, int n,
                          const int32_t* __restrict__ seg_begin, const int32_t* __restrict__ seg_end,
                          const int32_t* __restrict__ split, const int32_t* __restrict__ axis, int n_seg,
                          unsigned long long* __restrict__ key, unsigned long long* __restrict__ val) {
    for (int pos = blockIdx.x * BS + threadIdx.x; pos < n; pos += gridDim.x * BS) {
        // last segment starting at or before pos
        int lo = 0, hi = n_seg - 1, s = -1;
        while (lo <= hi) {
            int mid = (lo + hi) >> 1;
            if (seg_begin[mid] <= pos) {
                s = mid;
                lo = mid + 1;
            } else {
                hi = mid - 1;
            }
        }
        const int t = perm[pos];
        const bool inside = s >= 0 && pos < seg_end[s] && split[s];
        const unsigned second = inside ? (unsigned)seg_begin[s] : (unsigned)pos;
        key[pos] = inside ? order_key(cen[3 * t + axis[s]]) : 0ull;
        val[pos] = ((unsigned long long)second << 32) | (unsigned)t;
    }
}

__global__ void k_om_unpack(const unsigned long long* __restrict__ val, int n, uint32_t* __restrict__ key2,
                            int32_t* __restrict__ tri) {
    for (int i = blockIdx.x * BS + threadIdx.x; i < n; i += gridDim.x * BS) {
        unsigned long long v = val[i];
        key2[i] = (uint32_t)(v >> 32);
        tri[i] = (int32_t)(uint32_t)v;
    }
}

struct om_ws {
    double *lo, *hi, *cen;
    int32_t *perm, *perm2;
    unsigned long long *key_a, *key_b, *val_a, *val_b;
    uint32_t *key2_a, *key2_b;
    int32_t *seg_begin, *seg_end, *seg_node, *seg_split, *seg_axis;
    box6* seg_box;
    void* tmp;
    size_t tmp_bytes;
};

int bits_for(int n) {
    int b = 1;
    while (b < 31 && (1LL << b) < (long long)n) ++b;
    return b;
}

int om_carve(rx_arena& ar, int n, long long n_nodes, long long max_level, om_ws& w) {
    size_t N = (size_t)n;
    w.lo = ar.take<double>(3 * N);
    w.hi = ar.take<double>(3 * N);
    w.cen = ar.take<double>(3 * N);
    w.perm = ar.take<int32_t>(N);
    w.perm2 = ar.take<int32_t>(N);
    w.key_a = ar.take<unsigned long long>(N);
    w.key_b = ar.take<unsigned long long>(N);
    w.val_a = ar.take<unsigned long long>(N);
    w.val_b = ar.take<unsigned long long>(N);
    w.key2_a = ar.take<uint32_t>(N);
    w.key2_b = ar.take<uint32_t>(N);
    w.seg_begin = ar.take<int32_t>((size_t)n_nodes);
    w.seg_end = ar.take<int32_t>((size_t)n_nodes);
    w.seg_node = ar.take<int32_t>((size_t)n_nodes);
    w.seg_split = ar.take<int32_t>((size_t)n_nodes);
    w.seg_axis = ar.take<int32_t>((size_t)n_nodes);
    w.seg_box = ar.take<box6>((size_t)n_nodes);
    size_t a = 0, b = 0, c = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, a, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                             (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                             (unsigned int)n, 0u, 64u);
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(nullptr, b, (uint32_t*)nullptr, (uint32_t*)nullptr, (int32_t*)nullptr,
                                      (int32_t*)nullptr, (unsigned int)n, 0u, (unsigned)bits_for(n));
    if (e == hipSuccess) {
        auto in = rocprim::make_transform_iterator(rocprim::make_counting_iterator<int>(0), box_at{nullptr, nullptr, nullptr});
        e = rocprim::segmented_reduce(nullptr, c, in, (box6*)nullptr, (unsigned int)max_level, (int32_t*)nullptr,
                                      (int32_t*)nullptr, box_union(), EMPTY_BOX, (hipStream_t)0);
    }
    if (e != hipSuccess) return rx::set_error(RX_ERR_HIP, "rocPRIM size query failed: %s", hipGetErrorString(e));
    size_t m = a > b ? a : b;
    w.tmp_bytes = m > c ? m : c;
    w.tmp = ar.take_bytes(w.tmp_bytes);
    return RX_OK;
}

int grid_for(int n) {
    int g = (n + BS - 1) / BS;
    int cap = rx::device_cu_count() * 8;
    return g < 1 ? 1 : (g > cap ? cap : g);
}

int check_args(int n_tris, int max_leaf) {
    RX_REQUIRE(n_tris >= 0 && n_tris <= RX_MAX_TRIANGLES, RX_ERR_CAPACITY, "BVH has %d triangles, limit is %d",
               n_tris, RX_MAX_TRIANGLES);
    RX_REQUIRE(max_leaf >= 1 && max_leaf <= RX_MAX_LEAF_TRIS, RX_ERR_INVALID, "max_leaf must be in [1, %d]",
               RX_MAX_LEAF_TRIS);
    return RX_OK;
}

}  // namespace

extern "C" int rx_bvh_object_median_shape(int n_tris, int max_leaf, int* n_nodes, int* depth) {
    RX_REQUIRE(n_tris >= 0, RX_ERR_INVALID, "n_tris must be >= 0");
    RX_REQUIRE(max_leaf >= 1 && max_leaf <= RX_MAX_LEAF_TRIS, RX_ERR_INVALID, "max_leaf must be in [1, %d]",
               RX_MAX_LEAF_TRIS);
    if (n_tris == 0) {
        if (n_nodes) *n_nodes = 0;
        if (depth) *depth = 0;
        return RX_OK;
    }
    om_shape_t sh = shape_of(n_tris, max_leaf);
    RX_REQUIRE(sh.nodes <= 0x7fffffffLL, RX_ERR_CAPACITY, "tree of %lld nodes exceeds 2^31-1", sh.nodes);
    if (n_nodes) *n_nodes = (int)sh.nodes;
    if (depth) *depth = sh.depth;
    return RX_OK;
}

extern "C" int rx_bvh_object_median_workspace_size(int n_tris, int max_leaf, size_t* bytes) {
    RX_REQUIRE(bytes, RX_ERR_INVALID, "bytes is NULL");
    int st = check_args(n_tris, max_leaf);
    if (st) return st;
    int n = n_tris > 0 ? n_tris : 1;
    om_shape_t sh = shape_of(n, max_leaf);
    rx_arena ar{nullptr, 0, 0};
    om_ws w;
    st = om_carve(ar, n, sh.nodes, sh.nodes, w);
    if (st) return st;
    *bytes = ar.off + 256;
    return RX_OK;
}

extern "C" int rx_bvh_build_object_median(const double* d_v0, const double* d_v1, const double* d_v2, int n_tris,
                                          int max_leaf, void* d_workspace, size_t workspace_bytes,
                                          const rx_bvh_ref* ref, rx_node* d_nodes, rx_bvh_header* d_header,
                                          int32_t* d_order, void* stream) {
    int st = check_args(n_tris, max_leaf);
    if (st) return st;
    RX_REQUIRE(n_tris >= 1, RX_ERR_INVALID, "core/bvh.py builds over at least one triangle");
    RX_REQUIRE(d_v0 && d_v1 && d_v2 && ref && d_nodes && d_header && d_order, RX_ERR_INVALID, "NULL argument");
    RX_REQUIRE(ref->aabb_min && ref->aabb_max && ref->left && ref->right && ref->tri_start && ref->tri_count,
               RX_ERR_INVALID, "NULL reference-layout array");

    // topology on the host
    std::vector<om_node> nodes;
    om_layout(n_tris, max_leaf, nodes);
    const int n_nodes = (int)nodes.size();
    int depth = 0;
    for (const om_node& d : nodes) {
        if (d.depth > depth) depth = d.depth;
        RX_REQUIRE(!d.leaf || d.count <= RX_MAX_LEAF_TRIS, RX_ERR_CAPACITY, "BVH leaf of %d triangles exceeds %d",
                   d.count, RX_MAX_LEAF_TRIS);
    }
    RX_REQUIRE(depth <= RX_MAX_BVH_DEPTH, RX_ERR_CAPACITY, "BVH depth %d exceeds %d", depth, RX_MAX_BVH_DEPTH);
    // nodes grouped by depth, preorder (= ascending start) within a depth
    std::vector<int> lvl_off(depth + 2, 0);
    for (const om_node& d : nodes) lvl_off[d.depth + 1]++;
    for (int k = 0; k <= depth; ++k) lvl_off[k + 1] += lvl_off[k];
    std::vector<int32_t> h_begin(n_nodes), h_end(n_nodes), h_node(n_nodes), h_split(n_nodes);
    std::vector<int32_t> h_left(n_nodes), h_right(n_nodes), h_tstart(n_nodes), h_tcount(n_nodes);
    {
        std::vector<int> fill(lvl_off.begin(), lvl_off.end() - 1);
        for (int i = 0; i < n_nodes; ++i) {
            const om_node& d = nodes[i];
            int s = fill[d.depth]++;
            h_begin[s] = d.start;
            h_end[s] = d.start + d.count;
            h_node[s] = i;
            h_split[s] = d.leaf ? 0 : 1;
            h_left[i] = d.left;
            h_right[i] = d.right;
            h_tstart[i] = d.leaf ? d.start : -1;
            h_tcount[i] = d.leaf ? d.count : 0;
        }
    }
    int max_level = 0;
    for (int k = 0; k <= depth; ++k)
        if (lvl_off[k + 1] - lvl_off[k] > max_level) max_level = lvl_off[k + 1] - lvl_off[k];

    rx_arena ar{reinterpret_cast<char*>(d_workspace), workspace_bytes, 0};
    ar.off = (256 - (reinterpret_cast<uintptr_t>(d_workspace) & 255)) & 255;
    om_ws w;
    st = om_carve(ar, n_tris, n_nodes, n_nodes, w);
    if (st) return st;
    RX_REQUIRE(d_workspace && ar.off <= workspace_bytes, RX_ERR_WORKSPACE,
               "BVH workspace too small: need %zu bytes, have %zu", ar.off, workspace_bytes);
    hipStream_t s = rx_stream(stream);
    const size_t nb = (size_t)n_nodes * sizeof(int32_t);
    RX_TRY_HIP(hipMemcpyAsync(w.seg_begin, h_begin.data(), nb, hipMemcpyHostToDevice, s));
    RX_TRY_HIP(hipMemcpyAsync(w.seg_end, h_end.data(), nb, hipMemcpyHostToDevice, s));
    RX_TRY_HIP(hipMemcpyAsync(w.seg_node, h_node.data(), nb, hipMemcpyHostToDevice, s));
    RX_TRY_HIP(hipMemcpyAsync(w.seg_split, h_split.data(), nb, hipMemcpyHostToDevice, s));
    RX_TRY_HIP(hipMemcpyAsync(ref->left, h_left.data(), nb, hipMemcpyHostToDevice, s));
    RX_TRY_HIP(hipMemcpyAsync(ref->right, h_right.data(), nb, hipMemcpyHostToDevice, s));
    RX_TRY_HIP(hipMemcpyAsync(ref->tri_start, h_tstart.data(), nb, hipMemcpyHostToDevice, s));
    RX_TRY_HIP(hipMemcpyAsync(ref->tri_count, h_tcount.data(), nb, hipMemcpyHostToDevice, s));
    // the host tables go out of scope on every return path below
    RX_TRY_HIP(hipStreamSynchronize(s));

    const int g = grid_for(n_tris);
    k_om_prep<<<g, BS, 0, s>>>(d_v0, d_v1, d_v2, n_tris, w.lo, w.hi, w.cen, w.perm);
    RX_CHECK_LAUNCH();
    const unsigned bits2 = (unsigned)bits_for(n_tris);
    for (int k = 0; k <= depth; ++k) {
        const int off = lvl_off[k], cnt = lvl_off[k + 1] - off;
        int32_t* begin = w.seg_begin + off;
        int32_t* end = w.seg_end + off;
        int32_t* node = w.seg_node + off;
        int32_t* split = w.seg_split + off;
        int32_t* axis = w.seg_axis + off;
        box6* boxes = w.seg_box + off;
        auto in = rocprim::make_transform_iterator(rocprim::make_counting_iterator<int>(0),
                                                   box_at{w.perm, w.lo, w.hi});
        size_t tb = w.tmp_bytes;
        RX_TRY_HIP(rocprim::segmented_reduce(w.tmp, tb, in, boxes, (unsigned int)cnt, begin, end, box_union(),
                                             EMPTY_BOX, s));
        k_om_boxes<<<(cnt + BS - 1) / BS, BS, 0, s>>>(boxes, node, split, cnt, ref->aabb_min, ref->aabb_max, axis);
        RX_CHECK_LAUNCH();
        bool any_split = false;
        for (int i = off; i < off + cnt && !any_split; ++i) any_split = h_split[i] != 0;
        if (!any_split) continue;
        k_om_keys<<<g, BS, 0, s>>>(w.perm, w.cen, n_tris, begin, end, split, axis, cnt, w.key_a, w.val_a);
        RX_CHECK_LAUNCH();
        tb = w.tmp_bytes;
        RX_TRY_HIP(rocprim::radix_sort_pairs(w.tmp, tb, w.key_a, w.key_b, w.val_a, w.val_b, (unsigned int)n_tris, 0u,
                                             64u, s));
        k_om_unpack<<<g, BS, 0, s>>>(w.val_b, n_tris, w.key2_a, w.perm2);
        RX_CHECK_LAUNCH();
        tb = w.tmp_bytes;
        RX_TRY_HIP(rocprim::radix_sort_pairs(w.tmp, tb, w.key2_a, w.key2_b, w.perm2, w.perm, (unsigned int)n_tris, 0u,
                                             bits2, s));
    }
    RX_TRY_HIP(hipMemcpyAsync(d_order, w.perm, (size_t)n_tris * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    st = rx_bvh_pack(ref, n_nodes, depth, d_nodes, d_header, stream);
    if (st) return st;
    // callers free the workspace and vertex arrays once this returns
    RX_TRY_HIP(hipStreamSynchronize(s));
    return RX_OK;
}
