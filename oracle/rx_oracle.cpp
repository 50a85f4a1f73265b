// rx_oracle.cpp — CPU restatement of the reference render path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
// compared against; only tests/, __graft_entry__.smoke() and bench.py's
// CPU-baseline leg may load it.  It is written independently of the HIP
// sources (no shared headers) and follows the reference's Taichi code
// statement by statement:
//   rendering/taichi/intersect.py:17-309   slab test, Moeller-Trumbore,
//                                          stack BVH closest/any hit,
//                                          primitive scan, scene queries
//   scene/primitives.py:90-215             sphere/plane/cube intersectors
//   rendering/taichi/trace.py:23-283       light sampling, NEE, bounce loop,
//                                          accumulation
//   rendering/taichi/scatter.py:6-81       BSDF sampling
//   rendering/taichi/camera.py:4-30        jittered pinhole rays
//   rendering/taichi/sky.py:4-7            sky gradient
// over the reference's own data layout (fields.py: per-node AABB, left,
// right, tri_start, tri_count; triangles as v0, v1, v2).
//
// Randomness: the reference draws ti.random() (Taichi-internal, not
// reproducible).  Here every draw is a Philox4x32-10 output keyed by
// (seed) and counted by (pixel, sample, bounce, draw) — the same
// assignment the GPU path documents — so both produce identical sample
// paths.  Floating point: compiled with -ffp-contract=off; the fp32
// instance evaluates every expression in fp32 in the reference's operand
// order.  The fp64 instance (the RMSE oracle) runs the same estimator in
// double with libm sin/cos.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>

#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

typedef struct orc_scene {
    // BVH in the reference layout (rendering/taichi/fields.py:50-79)
    int n_nodes;
    const float* aabb_min;
    const float* aabb_max;
    const int32_t* left;
    const int32_t* right;
    const int32_t* tri_start;
    const int32_t* tri_count;
    int n_tris;
    const float* v0;
    const float* v1;
    const float* v2;
    const float* n0;
    const float* n1;
    const float* n2;
    const int32_t* mat_idx;
    const int32_t* mat_type;
    const float* mat_albedo;
    const float* mat_roughness;
    const float* mat_ior;
    const float* mat_emission;
    // primitive slots (fields.py:22-40)
    int n_prims;
    const int32_t* obj_type;
    const float* obj_center;
    const float* obj_radius;
    const float* obj_normal;
    const float* obj_offset;
    const float* obj_extra;
    const float* obj_v0;
    const float* obj_v1;
    const float* obj_v2;
    const int32_t* obj_mat_type;
    const float* obj_albedo;
    const float* obj_roughness;
    const float* obj_ior;
    const float* obj_emission;
    // lights (fields.py:43-47)
    int n_lights;
    const float* light_center;
    const float* light_radius;
    const float* light_albedo;
    const float* light_emission;
    // textures (fields.py:65-67 uvs in BVH order, :78 material texture,
    // :82-88 cache); mat_texture NULL = no textures
    const int32_t* mat_texture;
    const int32_t* has_uv;
    const float* uv0;
    const float* uv1;
    const float* uv2;
    int n_textures;
    const int32_t* tex_width;
    const int32_t* tex_height;
    const int32_t* tex_flip_v;
    const long long* tex_offset;
    const unsigned char* texels;
} orc_scene;

typedef struct orc_camera {
    float pos[3], fwd[3], right[3], up[3];
    float fov_tan, aspect;
    int32_t width, height;
} orc_camera;

typedef struct orc_params {
    int32_t max_depth, use_sky, direct_light_mode, direct_light_max_depth;
    float sample_clamp;
    float bg[3];
    uint32_t seed_lo, seed_hi;
} orc_params;

typedef struct orc_counts {
    unsigned long long rays, shadow_rays, node_pops, tri_tests, prim_tests, shadow_node_pops, shadow_tri_tests,
        shadow_prim_tests;
} orc_counts;

}  // extern "C"

namespace {

// ------------------------------------------------------------- Philox4x32-10
// Salmon et al., SC'11.  Round: multiply words 0 and 2 by the two round
// constants; new state = (hi(c2*M1) ^ c1 ^ k0, lo(c2*M1), hi(c0*M0) ^ c3 ^ k1,
// lo(c0*M0)); the key then advances by the two Weyl constants.
struct ctr4 {
    uint32_t v[4];
};

ctr4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1) {
    uint32_t x[4] = {c0, c1, c2, c3};
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * x[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * x[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ x[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ x[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        x[0] = n0;
        x[1] = n1;
        x[2] = n2;
        x[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    ctr4 r;
    memcpy(r.v, x, sizeof(x));
    return r;
}

inline float unit_float(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// ------------------------------------------------------------- sin/cos
// sin/cos(2*pi*u), u in [0,1): quadrant from 4u, fold to [0, pi/4], Taylor.
void sincos_turns_f32(float u, float* so, float* co) {
    float a = u * 4.0f;
    float quadrant = floorf(a);
    float frac = a - quadrant;
    bool mirrored = frac > 0.5f;
    float g = mirrored ? 1.0f - frac : frac;
    float x = g * 1.57079637050628662109375f;
    float x2 = x * x;
    float s = x + x * x2 * (-1.66666666666666666667e-1f +
                            x2 * (8.33333333333333333333e-3f +
                                  x2 * (-1.98412698412698412698e-4f + x2 * 2.75573192239858906526e-6f)));
    float c = 1.0f + x2 * (-0.5f + x2 * (4.16666666666666666667e-2f +
                                         x2 * (-1.38888888888888888889e-3f +
                                               x2 * (2.48015873015873015873e-5f +
                                                     x2 * -2.75573192239858906526e-7f))));
    if (mirrored) std::swap(s, c);
    switch ((int)quadrant & 3) {
        case 0: *so = s; *co = c; break;
        case 1: *so = c; *co = -s; break;
        case 2: *so = -s; *co = -c; break;
        default: *so = -c; *co = s; break;
    }
}

// fp64 sine for the jitter hash: Cody-Waite reduction by pi/2 in three
// 30-bit pieces, Taylor series on [-pi/4, pi/4].
double hash_sine(double x) {
    const double P1 = 1.570796325802803, P2 = 9.920935791635221e-10, P3 = 5.170182981794105e-19;
    double k = nearbyint(x * 6.36619772367581382433e-01);
    double r = ((x - k * P1) - k * P2) - k * P3;
    double r2 = r * r;
    double s = r + r * r2 * (-1.0 / 6.0 + r2 * (1.0 / 120.0 + r2 * (-1.0 / 5040.0 + r2 * (1.0 / 362880.0 +
                 r2 * (-1.0 / 39916800.0 + r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0 +
                 r2 * (1.0 / 355687428096000.0))))))));
    double c = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24.0 + r2 * (-1.0 / 720.0 + r2 * (1.0 / 40320.0 +
                 r2 * (-1.0 / 3628800.0 + r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0 +
                 r2 * (1.0 / 20922789888000.0 + r2 * (-1.0 / 6402373705728000.0)))))))));
    switch (((long long)k) & 3) {
        case 0: return s;
        case 1: return c;
        case 2: return -s;
        default: return -c;
    }
}

// ------------------------------------------------------------- vectors
template <class R>
struct V {
    R x, y, z;
};
template <class R> inline V<R> v3(R x, R y, R z) { return V<R>{x, y, z}; }
template <class R> inline V<R> add(V<R> a, V<R> b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V<R> sub(V<R> a, V<R> b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V<R> scl(V<R> a, R s) { return v3(a.x * s, a.y * s, a.z * s); }
template <class R> inline V<R> hmul(V<R> a, V<R> b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> inline V<R> divs(V<R> a, R s) { return v3(a.x / s, a.y / s, a.z / s); }
template <class R> inline V<R> neg(V<R> a) { return v3(-a.x, -a.y, -a.z); }
template <class R> inline R dot(V<R> a, V<R> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class R> inline V<R> cross(V<R> a, V<R> b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> inline V<R> unit(V<R> a) {
    R inv = R(1) / std::sqrt(dot(a, a));
    return scl(a, inv);
}
template <class R> inline V<R> at(const float* p, long i) { return v3<R>((R)p[3 * i], (R)p[3 * i + 1], (R)p[3 * i + 2]); }

template <class R> void sincos_turns(R u, R* s, R* c);
template <> void sincos_turns<float>(float u, float* s, float* c) { sincos_turns_f32(u, s, c); }
template <> void sincos_turns<double>(double u, double* s, double* c) {
    double ph = 6.283185307179586476925286766559 * u;
    *s = sin(ph);
    *c = cos(ph);
}

// ------------------------------------------------------------- intersection
// rendering/taichi/intersect.py:17-25
template <class R> inline V<R> safe_inverse(V<R> d) {
    R big = (R)1e30f, tiny = (R)1e-10f;
    return v3<R>(std::fabs(d.x) > tiny ? R(1) / d.x : big, std::fabs(d.y) > tiny ? R(1) / d.y : big,
                 std::fabs(d.z) > tiny ? R(1) / d.z : big);
}

// rendering/taichi/intersect.py:28-43
template <class R>
inline bool box_hit(V<R> ro, V<R> inv, V<R> lo, V<R> hi, R limit, R* t_enter) {
    V<R> a = hmul(sub(lo, ro), inv), b = hmul(sub(hi, ro), inv);
    V<R> t1 = v3(std::fmin(a.x, b.x), std::fmin(a.y, b.y), std::fmin(a.z, b.z));
    V<R> t2 = v3(std::fmax(a.x, b.x), std::fmax(a.y, b.y), std::fmax(a.z, b.z));
    R te = std::fmax(t1.x, std::fmax(t1.y, t1.z));
    R tx = std::fmin(t2.x, std::fmin(t2.y, t2.z));
    *t_enter = te;
    return tx >= te && tx >= (R)0.001f && te <= limit;
}

// rendering/taichi/intersect.py:57-78 (e1, e2 formed from the vertices here)
template <class R>
inline R triangle(V<R> ro, V<R> rd, V<R> a, V<R> b, V<R> c, R* uo, R* vo) {
    V<R> e1 = sub(b, a), e2 = sub(c, a);
    V<R> h = cross(rd, e2);
    R det = dot(e1, h);
    R t = -1, u = 0, v = 0;
    if (std::fabs(det) > (R)1e-8f) {
        R f = R(1) / det;
        V<R> s = sub(ro, a);
        u = f * dot(s, h);
        if (u >= 0 && u <= 1) {
            V<R> q = cross(s, e1);
            v = f * dot(rd, q);
            if (v >= 0 && u + v <= 1) {
                R tc = f * dot(e2, q);
                if (tc > (R)0.001f) t = tc;
            }
        }
    }
    *uo = u;
    *vo = v;
    return t;
}

// scene/primitives.py:90-105, :145-154, :187-215
template <class R> inline R sphere(V<R> ro, V<R> rd, V<R> c, R r) {
    V<R> oc = sub(ro, c);
    R h = dot(rd, oc);
    R cc = dot(oc, oc) - r * r;
    R disc = h * h - cc;
    R t = -1;
    if (disc >= 0) {
        R sq = std::sqrt(disc);
        t = -h - sq;
        if (t < (R)0.001f) t = -h + sq;
        if (t < (R)0.001f) t = -1;
    }
    return t;
}
template <class R> inline R plane(V<R> ro, V<R> rd, V<R> n, R off) {
    R den = dot(n, rd);
    R t = -1;
    if (std::fabs(den) >= (R)1e-6f) {
        R tc = -(dot(n, ro) + off) / den;
        if (tc >= (R)0.001f) t = tc;
    }
    return t;
}
template <class R> inline R cube(V<R> ro, V<R> rd, V<R> c, V<R> e) {
    V<R> lo = sub(c, e), hi = add(c, e);
    V<R> a = v3((lo.x - ro.x) / rd.x, (lo.y - ro.y) / rd.y, (lo.z - ro.z) / rd.z);
    V<R> b = v3((hi.x - ro.x) / rd.x, (hi.y - ro.y) / rd.y, (hi.z - ro.z) / rd.z);
    // ti.min/ti.max on NaN return the other operand, as fmin/fmax do
    R te = std::fmax(std::fmin(a.x, b.x), std::fmax(std::fmin(a.y, b.y), std::fmin(a.z, b.z)));
    R tx = std::fmin(std::fmax(a.x, b.x), std::fmin(std::fmax(a.y, b.y), std::fmax(a.z, b.z)));
    R t = -1;
    if (tx >= te && tx >= (R)0.001f) t = te >= (R)0.001f ? te : tx;
    return t;
}
template <class R> inline V<R> cube_normal(V<R> p, V<R> c, V<R> e) {
    V<R> d = v3((p.x - c.x) / e.x, (p.y - c.y) / e.y, (p.z - c.z) / e.z);
    R ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    if (ax >= ay && ax >= az) return v3<R>(d.x > 0 ? 1 : -1, 0, 0);
    if (ay >= az) return v3<R>(0, d.y > 0 ? 1 : -1, 0);
    return v3<R>(0, 0, d.z > 0 ? 1 : -1);
}

template <class R> R object_t(const orc_scene& s, int i, V<R> ro, V<R> rd) {
    switch (s.obj_type[i]) {
        case 0: return sphere(ro, rd, at<R>(s.obj_center, i), (R)s.obj_radius[i]);
        case 1: return plane(ro, rd, at<R>(s.obj_normal, i), (R)s.obj_offset[i]);
        case 2: return cube(ro, rd, at<R>(s.obj_center, i), at<R>(s.obj_extra, i));
        case 4: {
            R u, v;
            return triangle(ro, rd, at<R>(s.obj_v0, i), at<R>(s.obj_v1, i), at<R>(s.obj_v2, i), &u, &v);
        }
        default: return -1;
    }
}

template <class R> V<R> object_normal(const orc_scene& s, int i, V<R> ro, V<R> rd, R t) {
    switch (s.obj_type[i]) {
        case 0: return unit(sub(add(ro, scl(rd, t)), at<R>(s.obj_center, i)));
        case 1: return at<R>(s.obj_normal, i);
        case 2: return cube_normal(add(ro, scl(rd, t)), at<R>(s.obj_center, i), at<R>(s.obj_extra, i));
        case 4: {
            V<R> a = at<R>(s.obj_v0, i), b = at<R>(s.obj_v1, i), c = at<R>(s.obj_v2, i);
            return unit(cross(sub(b, a), sub(c, a)));
        }
        default: return v3<R>(0, 0, 0);
    }
}

struct tally {
    unsigned long long pops = 0, tris = 0, prims = 0;
};

// rendering/taichi/intersect.py:81-152.  The stack mirrors the reference:
// entries (node, t_enter), pop skipped when t_enter > closest.
template <class R>
R bvh_closest(const orc_scene& s, V<R> ro, V<R> rd, int* idx, R* uo, R* vo, tally& tl) {
    int stack[128];
    R stack_t[128];
    int sp = -1;
    V<R> inv = safe_inverse(rd);
    R best = (R)1e9f, bu = 0, bv = 0;
    int bi = -1;
    R rt;
    if (box_hit(ro, inv, at<R>(s.aabb_min, 0), at<R>(s.aabb_max, 0), best, &rt)) {
        stack[++sp] = 0;
        stack_t[sp] = rt;
    }
    while (sp >= 0) {
        int node = stack[sp];
        R nt = stack_t[sp];
        --sp;
        if (nt > best) continue;
        tl.pops++;
        if (s.tri_count[node] > 0) {
            for (int i = s.tri_start[node]; i < s.tri_start[node] + s.tri_count[node]; ++i) {
                R u, v;
                R t = triangle(ro, rd, at<R>(s.v0, i), at<R>(s.v1, i), at<R>(s.v2, i), &u, &v);
                tl.tris++;
                if ((R)0.001f < t && t < best) {
                    best = t;
                    bi = i;
                    bu = u;
                    bv = v;
                }
            }
        } else {
            int l = s.left[node], r = s.right[node];
            R lt, rt2;
            bool lh = box_hit(ro, inv, at<R>(s.aabb_min, l), at<R>(s.aabb_max, l), best, &lt);
            bool rh = box_hit(ro, inv, at<R>(s.aabb_min, r), at<R>(s.aabb_max, r), best, &rt2);
            if (lh && rh) {
                if (lt < rt2) {
                    stack[++sp] = r; stack_t[sp] = rt2;
                    stack[++sp] = l; stack_t[sp] = lt;
                } else {
                    stack[++sp] = l; stack_t[sp] = lt;
                    stack[++sp] = r; stack_t[sp] = rt2;
                }
            } else if (lh) {
                stack[++sp] = l; stack_t[sp] = lt;
            } else if (rh) {
                stack[++sp] = r; stack_t[sp] = rt2;
            }
        }
    }
    *idx = bi;
    *uo = bu;
    *vo = bv;
    return best;
}

// rendering/taichi/intersect.py:155-219 (stops at the first occluder)
template <class R>
bool bvh_blocked(const orc_scene& s, V<R> ro, V<R> rd, R max_t, tally& tl) {
    int stack[128];
    R stack_t[128];
    int sp = -1;
    V<R> inv = safe_inverse(rd);
    R rt;
    if (box_hit(ro, inv, at<R>(s.aabb_min, 0), at<R>(s.aabb_max, 0), max_t, &rt)) {
        stack[++sp] = 0;
        stack_t[sp] = rt;
    }
    while (sp >= 0) {
        int node = stack[sp];
        R nt = stack_t[sp];
        --sp;
        if (nt > max_t) continue;
        tl.pops++;
        if (s.tri_count[node] > 0) {
            for (int i = s.tri_start[node]; i < s.tri_start[node] + s.tri_count[node]; ++i) {
                R u, v;
                R t = triangle(ro, rd, at<R>(s.v0, i), at<R>(s.v1, i), at<R>(s.v2, i), &u, &v);
                tl.tris++;
                if ((R)0.001f < t && t < max_t) return true;
            }
        } else {
            int l = s.left[node], r = s.right[node];
            R lt, rt2;
            bool lh = box_hit(ro, inv, at<R>(s.aabb_min, l), at<R>(s.aabb_max, l), max_t, &lt);
            bool rh = box_hit(ro, inv, at<R>(s.aabb_min, r), at<R>(s.aabb_max, r), max_t, &rt2);
            if (lh && rh) {
                if (lt < rt2) {
                    stack[++sp] = r; stack_t[sp] = rt2;
                    stack[++sp] = l; stack_t[sp] = lt;
                } else {
                    stack[++sp] = l; stack_t[sp] = lt;
                    stack[++sp] = r; stack_t[sp] = rt2;
                }
            } else if (lh) {
                stack[++sp] = l; stack_t[sp] = lt;
            } else if (rh) {
                stack[++sp] = r; stack_t[sp] = rt2;
            }
        }
    }
    return false;
}

template <class R> struct hit {
    R t;
    int idx, kind;  // kind: 0 miss, 1 primitive, 2 BVH triangle
    R u, v;
};

// rendering/taichi/intersect.py:249-309 (normal computed by the caller)
template <class R> hit<R> closest(const orc_scene& s, V<R> ro, V<R> rd, tally& tl) {
    hit<R> h{(R)1e9f, -1, 0, 0, 0};
    for (int i = 0; i < s.n_prims; ++i) {
        R t = object_t(s, i, ro, rd);
        tl.prims++;
        if (t > (R)0.001f && t < h.t) {
            h.t = t;
            h.idx = i;
            h.kind = 1;
        }
    }
    if (s.n_tris > 0) {
        int bi;
        R u, v;
        R t = bvh_closest(s, ro, rd, &bi, &u, &v, tl);
        if ((R)0.001f < t && t < h.t) {
            h.t = t;
            h.idx = bi;
            h.kind = 2;
            h.u = u;
            h.v = v;
        }
    }
    return h;
}

// rendering/taichi/intersect.py:222-246
template <class R> bool blocked(const orc_scene& s, V<R> ro, V<R> rd, R max_t, tally& tl) {
    for (int i = 0; i < s.n_prims; ++i) {
        R t = object_t(s, i, ro, rd);
        tl.prims++;
        if ((R)0.001f < t && t < max_t) return true;
    }
    if (s.n_tris > 0) return bvh_blocked(s, ro, rd, max_t, tl);
    return false;
}

// ------------------------------------------------------------- sampling
// rendering/taichi/camera.py:9-30
inline float hash01(int x, int y, float salt) {
    float seed = ((float)x * 127.1f + (float)y * 311.7f) + salt * 74.7f;
    float sv = (float)hash_sine((double)seed);
    float h = sv * 43758.5453123f;
    return h - floorf(h);
}

template <class R> V<R> camera_ray(const orc_camera& cam, int px, int py, int frame) {
    float i = (float)(frame + 1);
    float jx = hash01(px, py, 0.0f) + i * 0.7548776662466927f;
    jx = jx - floorf(jx);
    float jy = hash01(px, py, 1.0f) + i * 0.5698402909980532f;
    jy = jy - floorf(jy);
    R u = ((R)px + (R)jx) / (R)cam.width * (R)2 - (R)1;
    R v = (R)1 - ((R)py + (R)jy) / (R)cam.height * (R)2;
    R ndc_x = u * (R)cam.aspect * (R)cam.fov_tan;
    R ndc_y = v * (R)cam.fov_tan;
    V<R> f = v3<R>(cam.fwd[0], cam.fwd[1], cam.fwd[2]);
    V<R> r = v3<R>(cam.right[0], cam.right[1], cam.right[2]);
    V<R> up = v3<R>(cam.up[0], cam.up[1], cam.up[2]);
    return unit(add(add(f, scl(r, ndc_x)), scl(up, ndc_y)));
}

template <class R> struct draws {
    const orc_params* p;
    int pix, frame, depth;
    R get(uint32_t call, int word) const {
        ctr4 c = philox((uint32_t)pix, (uint32_t)frame, (uint32_t)depth, call, p->seed_lo, p->seed_hi);
        return (R)unit_float(c.v[word]);
    }
};

// rendering/taichi/scatter.py:14-38
template <class R> V<R> cosine_dir(V<R> n, R r1, R r2) {
    R sp, cp;
    sincos_turns<R>(r1, &sp, &cp);
    R r = std::sqrt(r2);
    R lx = cp * r, ly = sp * r, lz = std::sqrt(std::fmax((R)0, (R)1 - r2));
    V<R> helper = std::fabs(n.x) > (R)0.9f ? v3<R>(0, 1, 0) : v3<R>(1, 0, 0);
    V<R> b = unit(cross(n, helper));
    V<R> t = cross(b, n);
    V<R> s = unit(add(add(scl(t, lx), scl(b, ly)), scl(n, lz)));
    if (dot(s, s) < (R)1e-8f) s = n;
    return unit(s);
}

template <class R> V<R> mirror(V<R> d, V<R> n) { return sub(d, scl(n, (R)2 * dot(d, n))); }

// rendering/taichi/scatter.py:6-11 with cos(theta) = 1 - 2 xi and
// sin(theta) = sqrt(1 - cos^2) in place of acos/sin/cos.
template <class R> V<R> sphere_dir(R r1, R r2) {
    R z = (R)1 - (R)2 * r1;
    R st = std::sqrt(std::fmax((R)0, (R)1 - z * z));
    R sp, cp;
    sincos_turns<R>(r2, &sp, &cp);
    return v3(st * cp, st * sp, z);
}

template <class R> V<R> fuzzy_mirror(V<R> d, V<R> n, R rough, R r1, R r2) {
    return unit(add(mirror(d, n), scl(sphere_dir(r1, r2), rough)));
}

template <class R> R fifth(R x) {
    R x2 = x * x;
    return (x2 * x2) * x;
}

// rendering/taichi/scatter.py:40-81
template <class R> V<R> glass_dir(V<R> d, V<R> n, R eta, R xi) {
    R ct = std::fmin(-dot(d, n), (R)1);
    R st = std::sqrt((R)1 - ct * ct);
    bool tir = eta * st > (R)1;
    R r0 = ((R)1 - eta) / ((R)1 + eta);
    r0 = r0 * r0;
    R fres = r0 + ((R)1 - r0) * fifth((R)1 - ct);
    if (tir || fres > xi) return mirror(d, n);
    V<R> perp = scl(add(d, scl(n, ct)), eta);
    V<R> par = scl(n, -std::sqrt(std::fabs((R)1 - dot(perp, perp))));
    return add(perp, par);
}

// rendering/taichi/trace.py:23-57
template <class R>
void light_sample(V<R> hp, V<R> c, R radius, R xi1, R xi2, V<R>* dir, R* t_light, R* solid) {
    V<R> tc = sub(c, hp);
    R d2 = std::fmax(dot(tc, tc), (R)1e-6f);
    R d = std::sqrt(d2);
    V<R> axis = divs(tc, d);
    R sin2 = std::fmin(radius * radius / d2, (R)1);
    R cmax = std::sqrt(std::fmax((R)0, (R)1 - sin2));
    R ct = (R)1 - xi1 * ((R)1 - cmax);
    R st = std::sqrt(std::fmax((R)0, (R)1 - ct * ct));
    R sp, cp;
    sincos_turns<R>(xi2, &sp, &cp);
    V<R> helper = std::fabs(axis.x) > (R)0.9f ? v3<R>(0, 1, 0) : v3<R>(1, 0, 0);
    V<R> b = unit(cross(axis, helper));
    V<R> t = cross(b, axis);
    V<R> ld = unit(add(add(scl(t, cp * st), scl(b, sp * st)), scl(axis, ct)));
    V<R> oc = sub(hp, c);
    R hb = dot(oc, ld);
    R cc = dot(oc, oc) - radius * radius;
    R disc = std::fmax(hb * hb - cc, (R)0);
    *t_light = -hb - std::sqrt(disc);
    *solid = (R)2 * ((R)1 - cmax);
    *dir = ld;
}

// rendering/taichi/trace.py:60-93
template <class R>
V<R> direct_light(const orc_scene& s, const orc_params& p, const draws<R>& dr, V<R> hp, V<R> n, V<R> alb,
                  orc_counts& cnt, tally& tl) {
    V<R> direct = v3<R>(0, 0, 0);
    int L = s.n_lights;
    if (L <= 0) return direct;
    int sel = (int)(dr.get(0, 0) * (R)L);
    if (sel > L - 1) sel = L - 1;
    bool all = p.direct_light_mode == 1;
    for (int i = 0; i < L; ++i) {
        if (!(all || i == sel)) continue;
        R xi1 = all ? dr.get(2u + (uint32_t)i, 0) : dr.get(0, 1);
        R xi2 = all ? dr.get(2u + (uint32_t)i, 1) : dr.get(0, 2);
        V<R> ld;
        R t_light, solid;
        light_sample(hp, at<R>(s.light_center, i), (R)s.light_radius[i], xi1, xi2, &ld, &t_light, &solid);
        R ndotl = dot(n, ld);
        if (ndotl > 0) {
            R max_t = t_light * (R)0.999f;
            if (max_t > (R)0.001f) {
                V<R> so = add(hp, scl(n, (R)1e-4f));
                cnt.rays++;
                cnt.shadow_rays++;
                if (!blocked(s, so, ld, max_t, tl)) {
                    R w = all ? (R)1 : (R)L;
                    V<R> term = scl(scl(scl(hmul(alb, at<R>(s.light_albedo, i)), (R)s.light_emission[i]), ndotl),
                                    solid);
                    direct = add(direct, scl(term, w));
                }
            }
        }
    }
    return direct;
}

template <class R> struct path_out {
    V<R> color, normal, albedo;
    R depth;
};

// rendering/taichi/trace.py:139-236
// rendering/taichi/trace.py:96-126: repeat, optional v flip, bilinear of
// u8 texels / 255; a zero-size slot samples as white.
template <class R>
V<R> sample_texture(const orc_scene& s, int t, R tu, R tv) {
    int W = s.tex_width[t], H = s.tex_height[t];
    if (W <= 0 || H <= 0) return v3<R>(1, 1, 1);
    R u = tu - std::floor(tu);
    R v = tv - std::floor(tv);
    if (s.tex_flip_v[t] != 0) v = (R)1 - v;
    R x = u * (R)(W - 1), y = v * (R)(H - 1);
    int x0 = (int)std::floor(x), y0 = (int)std::floor(y);
    int x1 = std::min(x0 + 1, W - 1), y1 = std::min(y0 + 1, H - 1);
    R fx = x - (R)x0, fy = y - (R)y0;
    auto texel = [&](int yy, int xx) {
        const unsigned char* q = s.texels + s.tex_offset[t] + 3 * ((long long)yy * W + xx);
        return v3<R>((R)q[0] / (R)255, (R)q[1] / (R)255, (R)q[2] / (R)255);
    };
    V<R> c0 = add(scl(texel(y0, x0), (R)1 - fx), scl(texel(y0, x1), fx));
    V<R> c1 = add(scl(texel(y1, x0), (R)1 - fx), scl(texel(y1, x1), fx));
    return add(scl(c0, (R)1 - fy), scl(c1, fy));
}

template <class R>
path_out<R> trace(const orc_scene& s, const orc_params& p, int pix, int frame, V<R> ro, V<R> rd, orc_counts& cnt,
                  tally& tl, tally& tl_shadow) {
    V<R> color = v3<R>(0, 0, 0), thr = v3<R>(1, 1, 1);
    R cur_ior = 1;
    path_out<R> out{v3<R>(0, 0, 0), v3<R>(0, 0, 0), v3<R>(0, 0, 0), 0};
    for (int depth = 0; depth < p.max_depth; ++depth) {
        cnt.rays++;
        hit<R> h = closest(s, ro, rd, tl);
        if (h.t < 0 || h.t >= (R)1e9f) {
            V<R> sky;
            if (p.use_sky) {
                R t = (R)0.5f * (rd.y + (R)1);
                sky = add(scl(v3<R>(1, 1, 1), (R)1 - t), scl(v3<R>((R)0.5f, (R)0.7f, (R)1), t));
            } else {
                sky = v3<R>(p.bg[0], p.bg[1], p.bg[2]);
            }
            color = add(color, hmul(thr, sky));
            break;
        }
        V<R> n;
        int mtype;
        V<R> alb;
        R rough, ior, emis;
        if (h.kind == 2) {
            R w = ((R)1 - h.u) - h.v;
            long i = h.idx;
            n = unit(add(add(scl(at<R>(s.n0, i), w), scl(at<R>(s.n1, i), h.u)), scl(at<R>(s.n2, i), h.v)));
            int m = s.mat_idx[i];
            mtype = s.mat_type[m];
            alb = at<R>(s.mat_albedo, m);
            // trace.py:129-136 (_bvh_albedo)
            if (s.mat_texture && s.mat_texture[m] >= 0 && s.has_uv[i] != 0 && s.mat_texture[m] < s.n_textures) {
                R tu = (s.uv0[2 * i] * w + s.uv1[2 * i] * h.u) + s.uv2[2 * i] * h.v;
                R tv = (s.uv0[2 * i + 1] * w + s.uv1[2 * i + 1] * h.u) + s.uv2[2 * i + 1] * h.v;
                alb = hmul(alb, sample_texture<R>(s, s.mat_texture[m], tu, tv));
            }
            rough = s.mat_roughness[m];
            ior = s.mat_ior[m];
            emis = s.mat_emission[m];
        } else {
            n = object_normal(s, h.idx, ro, rd, h.t);
            int i = h.idx;
            mtype = s.obj_mat_type[i];
            alb = at<R>(s.obj_albedo, i);
            rough = s.obj_roughness[i];
            ior = s.obj_ior[i];
            emis = s.obj_emission[i];
        }
        bool front = dot(n, rd) < 0;
        if (!front) n = neg(n);
        V<R> hp = add(ro, scl(rd, h.t));
        if (depth == 0) {
            out.normal = n;
            out.albedo = alb;
            out.depth = h.t;
        }
        draws<R> dr{&p, pix, frame, depth};
        bool nee = depth < p.direct_light_max_depth;
        if (mtype == 3) {
            color = add(color, scl(hmul(thr, alb), emis));
            break;
        } else if (mtype == 0) {
            if (nee && s.n_lights > 0) color = add(color, hmul(thr, direct_light(s, p, dr, hp, n, alb, cnt, tl_shadow)));
            rd = cosine_dir(n, dr.get(1, 0), dr.get(1, 1));
            thr = hmul(thr, alb);
        } else if (mtype == 1) {
            rd = fuzzy_mirror(rd, n, rough, dr.get(1, 0), dr.get(1, 1));
            thr = hmul(thr, alb);
        } else if (mtype == 4) {
            if (nee && s.n_lights > 0)
                color = add(color, scl(hmul(thr, direct_light(s, p, dr, hp, n, alb, cnt, tl_shadow)), rough));
            if (dr.get(0, 3) < rough)
                rd = cosine_dir(n, dr.get(1, 0), dr.get(1, 1));
            else
                rd = fuzzy_mirror(rd, n, rough, dr.get(1, 0), dr.get(1, 1));
            thr = hmul(thr, alb);
        } else if (mtype == 2) {
            R eta;
            if (front) {
                eta = cur_ior / ior;
                cur_ior = ior;
            } else {
                eta = cur_ior / (R)1;
                cur_ior = 1;
            }
            rd = glass_dir(rd, n, eta, dr.get(0, 3));
            thr = hmul(thr, alb);
        }
        if (depth > 2) {
            R rr = std::fmax(thr.x, std::fmax(thr.y, thr.z));
            if (dr.get(1, 2) > rr) break;
            thr = divs(thr, rr);
        }
        ro = add(hp, scl(n, (R)1e-4f));
    }
    out.color = color;
    return out;
}

template <class R>
void render(const orc_scene& s, const orc_camera& cam, const orc_params& p, int first, int count, float* acc_c,
            float* acc_n, float* acc_a, float* acc_d, double* acc_c64, orc_counts* total, int threads,
            const int32_t* pixel_list, int n_list, double* acc_sq = nullptr) {
    int npix = cam.width * cam.height;
    int nwork = pixel_list ? n_list : npix;
    unsigned long long rays = 0, shadow = 0, pops = 0, tris = 0, prims = 0, spops = 0, stris = 0, sprims = 0;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads) \
    reduction(+ : rays, shadow, pops, tris, prims, spops, stris, sprims)
    for (int w = 0; w < nwork; ++w) {
        int pix = pixel_list ? pixel_list[w] : w;
        int x = pix % cam.width, y = pix / cam.width;
        orc_counts cnt;
        memset(&cnt, 0, sizeof(cnt));
        tally tl, ts;
        V<R> c = v3<R>(0, 0, 0), nn = c, aa = c;
        R dd = 0;
        if (first > 0) {
            if (acc_c64) c = v3<R>((R)acc_c64[3 * w], (R)acc_c64[3 * w + 1], (R)acc_c64[3 * w + 2]);
            else c = at<R>(acc_c, w);
            nn = at<R>(acc_n, w);
            aa = at<R>(acc_a, w);
            dd = (R)acc_d[w];
        }
        for (int k = 0; k < count; ++k) {
            int frame = first + k;
            V<R> rd = camera_ray<R>(cam, x, y, frame);
            V<R> ro = v3<R>(cam.pos[0], cam.pos[1], cam.pos[2]);
            path_out<R> po = trace<R>(s, p, pix, frame, ro, rd, cnt, tl, ts);
            V<R> smp = po.color;
            if (p.sample_clamp > 0) {
                R lim = p.sample_clamp;
                smp = v3(std::fmin(std::fmax(smp.x, (R)0), lim), std::fmin(std::fmax(smp.y, (R)0), lim),
                         std::fmin(std::fmax(smp.z, (R)0), lim));
            }
            if (acc_sq) {
                // second moments of the accumulated samples (per-pixel variance)
                acc_sq[3 * w] += (double)smp.x * (double)smp.x;
                acc_sq[3 * w + 1] += (double)smp.y * (double)smp.y;
                acc_sq[3 * w + 2] += (double)smp.z * (double)smp.z;
            }
            // rendering/taichi/trace.py:266-281
            if (frame == 0) {
                c = smp;
                nn = po.normal;
                aa = po.albedo;
                dd = po.depth;
            } else {
                R f = (R)frame, f1 = (R)(frame + 1);
                c = divs(add(scl(c, f), smp), f1);
                nn = divs(add(scl(nn, f), po.normal), f1);
                aa = divs(add(scl(aa, f), po.albedo), f1);
                dd = (dd * f + po.depth) / f1;
            }
        }
        if (acc_c64) {
            acc_c64[3 * w] = (double)c.x;
            acc_c64[3 * w + 1] = (double)c.y;
            acc_c64[3 * w + 2] = (double)c.z;
        }
        if (acc_c) {
            acc_c[3 * w] = (float)c.x;
            acc_c[3 * w + 1] = (float)c.y;
            acc_c[3 * w + 2] = (float)c.z;
        }
        acc_n[3 * w] = (float)nn.x;
        acc_n[3 * w + 1] = (float)nn.y;
        acc_n[3 * w + 2] = (float)nn.z;
        acc_a[3 * w] = (float)aa.x;
        acc_a[3 * w + 1] = (float)aa.y;
        acc_a[3 * w + 2] = (float)aa.z;
        acc_d[w] = (float)dd;
        rays += cnt.rays;
        shadow += cnt.shadow_rays;
        pops += tl.pops;
        tris += tl.tris;
        prims += tl.prims;
        spops += ts.pops;
        stris += ts.tris;
        sprims += ts.prims;
    }
    if (total) {
        total->rays += rays;
        total->shadow_rays += shadow;
        total->node_pops += pops;
        total->tri_tests += tris;
        total->prim_tests += prims;
        total->shadow_node_pops += spops;
        total->shadow_tri_tests += stris;
        total->shadow_prim_tests += sprims;
    }
}

}  // namespace

extern "C" {

int orc_version(void) { return 1; }

void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    ctr4 c = philox(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1]);
    memcpy(out, c.v, sizeof(c.v));
}

void orc_sincos_turns(const float* u, int n, float* s, float* c) {
    for (int i = 0; i < n; ++i) sincos_turns_f32(u[i], &s[i], &c[i]);
}

void orc_hash_sine(const double* x, int n, double* out) {
    for (int i = 0; i < n; ++i) out[i] = hash_sine(x[i]);
}

void orc_camera_rays(const orc_camera* cam, int frame, float* dirs) {
    for (int y = 0; y < cam->height; ++y)
        for (int x = 0; x < cam->width; ++x) {
            V<float> d = camera_ray<float>(*cam, x, y, frame);
            long i = (long)y * cam->width + x;
            dirs[3 * i] = d.x;
            dirs[3 * i + 1] = d.y;
            dirs[3 * i + 2] = d.z;
        }
}

// The direction samplers of rendering/taichi/scatter.py evaluated on given
// uniforms (fp32), for distribution tests of the restatement:
//   kind 0  cosine_hemisphere(n) / diffuse_scatter       (:14-38), u1 = r1, u2 = r2
//   kind 1  random_unit_vector                            (:6-11),  u1, u2
//   kind 2  metal_scatter(d, n, roughness = param)        (:50-54), u1, u2
//   kind 3  dielectric_scatter(d, n, eta = param), xi = u1 (:71-81)
// d and n are 3 floats; out is count x 3.
void orc_scatter_dirs(int kind, const float* d, const float* n, float param, const float* u1, const float* u2,
                      int count, float* out) {
    V<float> dd = v3(d[0], d[1], d[2]), nn = v3(n[0], n[1], n[2]);
    for (int i = 0; i < count; ++i) {
        V<float> r = v3(0.0f, 0.0f, 0.0f);
        if (kind == 0) r = cosine_dir<float>(nn, u1[i], u2[i]);
        if (kind == 1) r = sphere_dir<float>(u1[i], u2[i]);
        if (kind == 2) r = fuzzy_mirror<float>(dd, nn, param, u1[i], u2[i]);
        if (kind == 3) r = glass_dir<float>(dd, nn, param, u1[i]);
        out[3 * i] = r.x;
        out[3 * i + 1] = r.y;
        out[3 * i + 2] = r.z;
    }
}

// Closest hit for n rays (SoA like the GPU entry point).
void orc_closest_hit(const orc_scene* s, const float* rays, int n, float* t, int32_t* idx, int32_t* kind, float* u,
                     float* v, orc_counts* total, int threads) {
    unsigned long long pops = 0, tris = 0, prims = 0;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads) reduction(+ : pops, tris, prims)
    for (int i = 0; i < n; ++i) {
        tally tl;
        V<float> ro = v3(rays[i], rays[n + i], rays[2 * n + i]);
        V<float> rd = v3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        hit<float> h = closest<float>(*s, ro, rd, tl);
        t[i] = h.t;
        idx[i] = h.idx;
        kind[i] = h.kind;
        u[i] = h.u;
        v[i] = h.v;
        pops += tl.pops;
        tris += tl.tris;
        prims += tl.prims;
    }
    if (total) {
        total->node_pops += pops;
        total->tri_tests += tris;
        total->prim_tests += prims;
    }
}

void orc_any_hit(const orc_scene* s, const float* rays, const float* tmax, int n, int32_t* occ, int threads) {
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (int i = 0; i < n; ++i) {
        tally tl;
        V<float> ro = v3(rays[i], rays[n + i], rays[2 * n + i]);
        V<float> rd = v3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        occ[i] = blocked<float>(*s, ro, rd, tmax[i], tl) ? 1 : 0;
    }
}

// Shaded hit normals (oriented against the ray) for n rays whose closest
// hits are known — the normal trace.py:161-163 shades with.
void orc_hit_normals(const orc_scene* s, const float* rays, int n, const float* t, const int32_t* idx,
                     const int32_t* kind, const float* u, const float* v, float* nrm) {
    for (int i = 0; i < n; ++i) {
        V<float> ro = v3(rays[i], rays[n + i], rays[2 * n + i]);
        V<float> rd = v3(rays[3 * n + i], rays[4 * n + i], rays[5 * n + i]);
        V<float> nn = v3(0.0f, 0.0f, 0.0f);
        if (kind[i] == 2) {
            float w = (1.0f - u[i]) - v[i];
            long k = idx[i];
            nn = unit(add(add(scl(at<float>(s->n0, k), w), scl(at<float>(s->n1, k), u[i])), scl(at<float>(s->n2, k), v[i])));
        } else if (kind[i] == 1) {
            nn = object_normal<float>(*s, idx[i], ro, rd, t[i]);
        }
        if (kind[i] != 0 && !(dot(nn, rd) < 0.0f)) nn = neg(nn);
        nrm[3 * i] = nn.x;
        nrm[3 * i + 1] = nn.y;
        nrm[3 * i + 2] = nn.z;
    }
}

// fp32 render of samples [first, first+count) folded into the running
// means (arrays of npix, or of n_list pixels when pixel_list is given).
void orc_render_f32(const orc_scene* s, const orc_camera* cam, const orc_params* p, int first, int count,
                    float* color, float* normal, float* albedo, float* depth, orc_counts* total, int threads,
                    const int32_t* pixel_list, int n_list) {
    render<float>(*s, *cam, *p, first, count, color, normal, albedo, depth, nullptr, total, threads, pixel_list,
                  n_list);
}

// fp64 instance of the same estimator (the RMSE oracle); colour running
// mean kept in double.
void orc_render_f64(const orc_scene* s, const orc_camera* cam, const orc_params* p, int first, int count,
                    double* color, float* normal, float* albedo, float* depth, orc_counts* total, int threads,
                    const int32_t* pixel_list, int n_list) {
    render<double>(*s, *cam, *p, first, count, nullptr, normal, albedo, depth, color, total, threads, pixel_list,
                   n_list);
}

// orc_render_f64 that also adds each sample's squared colour to sq (n x 3,
// summed over calls): with the running mean, the per-pixel sample variance.
void orc_render_f64_moments(const orc_scene* s, const orc_camera* cam, const orc_params* p, int first, int count,
                            double* color, float* normal, float* albedo, float* depth, orc_counts* total,
                            int threads, const int32_t* pixel_list, int n_list, double* sq) {
    render<double>(*s, *cam, *p, first, count, nullptr, normal, albedo, depth, color, total, threads, pixel_list,
                   n_list, sq);
}

}  // extern "C"
