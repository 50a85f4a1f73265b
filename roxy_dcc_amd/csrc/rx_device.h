// rx_device.h — device math, sampling and BVH traversal for gfx950.
//
// Floating-point contract: every expression here is evaluated in IEEE fp32
// (fp64 for the jitter hash) with no FMA contraction (the library is
// compiled with -ffp-contract=off and the pragma below), correctly rounded
// division and sqrt, and the operand order the reference's Taichi code
// spells out (dot = (x*x' + y*y') + z*z', left-to-right products).  That
// is what makes closest-hit ids and t bit-identical to a CPU restatement
// that follows the same order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "roxy_hip.h"

#pragma clang fp contract(off)

// Address-space qualifiers: pointers read out of scene records are generic
// to the compiler, and generic (flat) accesses wait on both the LDS and
// the vector-memory counters.  Casting the hot traversal data to these
// makes the loads global_load_* and the LDS stack ds_*.
#define RX_GLOBAL __attribute__((address_space(1)))
#define RX_LDS __attribute__((address_space(3)))
// Constant address space: loads at a wave-uniform address through it are
// scalar loads (s_load, scalar cache), which do not use the vector memory
// path.  Only for data no kernel writes while the reading kernel runs.
#define RX_CONST __attribute__((address_space(4)))
typedef float gv4f __attribute__((ext_vector_type(4)));
typedef int gv4i __attribute__((ext_vector_type(4)));
typedef int gv2i __attribute__((ext_vector_type(2)));

// Loads a record at a wave-uniform address with scalar loads.
template <class T>
__device__ __forceinline__ T cload(const T* p) {
    static_assert(sizeof(T) % 16 == 0, "scene records are whole 16-byte words");
    constexpr int N = sizeof(T) / 16;
    gv4f tmp[N];
    const RX_CONST gv4f* q = (const RX_CONST gv4f*)p;
#pragma unroll
    for (int i = 0; i < N; ++i) tmp[i] = q[i];
    T out;
    __builtin_memcpy(&out, tmp, sizeof(T));
    return out;
}

// Loads a scene record (a multiple of 16 bytes, 16-byte aligned: the entry
// points check the array alignment) as 16-byte global loads into a local
// copy, so later field reads do not go back through a generic pointer.
template <class T>
__device__ __forceinline__ T gload(const T* p) {
    static_assert(sizeof(T) % 16 == 0, "scene records are whole 16-byte words");
    constexpr int N = sizeof(T) / 16;
    gv4f tmp[N];
    const RX_GLOBAL gv4f* q = (const RX_GLOBAL gv4f*)p;
#pragma unroll
    for (int i = 0; i < N; ++i) tmp[i] = q[i];
    T out;
    __builtin_memcpy(&out, tmp, sizeof(T));
    return out;
}

namespace rx {

// ---------------------------------------------------------------- vectors
struct f3 {
    float x, y, z;
};

__device__ __forceinline__ f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot(f3 a, f3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// v * (1 / |v|): one correctly rounded divide, then three products.
__device__ __forceinline__ f3 normalize(f3 v) {
    float inv = 1.0f / sqrtf(dot(v, v));
    return v * inv;
}
__device__ __forceinline__ float fract(float x) { return x - floorf(x); }
__device__ __forceinline__ f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

// ---------------------------------------------------------------- Philox4x32-10
// Counter-based generator of Salmon, Moraes, Dror and Shaw, "Parallel random
// numbers: as easy as 1, 2, 3" (SC'11): ten rounds of two 32x32->64
// multiplies with Weyl-sequence key bumps.
struct u4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox4x32_10(u4 c, uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = M0 * c.x, hi0 = __umulhi(M0, c.x);
        uint32_t lo1 = M1 * c.z, hi1 = __umulhi(M1, c.z);
        c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Top 24 bits -> [0, 1), exactly representable.
__device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// ---------------------------------------------------------------- sin/cos
// sin and cos of 2*pi*u for u in [0, 1): quadrant split on 4u (exact),
// fold to [0, pi/4] and Taylor polynomials whose truncation error is far
// below half an ulp there.  Plain fp32 arithmetic, so a CPU build of the
// same steps gives the same bits.
__device__ __forceinline__ void sincos_2pi(float u, float* s_out, float* c_out) {
    float u4x = u * 4.0f;
    float q = floorf(u4x);
    float f = u4x - q;  // [0, 1) exact
    bool swap = f > 0.5f;
    float g = swap ? (1.0f - f) : f;           // [0, 0.5] exact
    float x = g * 1.57079637050628662109375f;  // pi/2 rounded to fp32
    float x2 = x * x;
    float s = x + x * x2 * (-1.66666666666666666667e-1f +
                            x2 * (8.33333333333333333333e-3f +
                                  x2 * (-1.98412698412698412698e-4f + x2 * 2.75573192239858906526e-6f)));
    float c = 1.0f + x2 * (-0.5f + x2 * (4.16666666666666666667e-2f +
                                         x2 * (-1.38888888888888888889e-3f +
                                               x2 * (2.48015873015873015873e-5f +
                                                     x2 * -2.75573192239858906526e-7f))));
    if (swap) {
        float t = s;
        s = c;
        c = t;
    }
    int qi = (int)q & 3;
    float so, co;
    if (qi == 0) {
        so = s;
        co = c;
    } else if (qi == 1) {
        so = c;
        co = -s;
    } else if (qi == 2) {
        so = -s;
        co = -c;
    } else {
        so = -c;
        co = s;
    }
    *s_out = so;
    *c_out = co;
}

// fp64 sine for the camera jitter hash: three-piece Cody-Waite reduction
// by pi/2 (30-bit pieces, exact products for |k| < 2^23) and Taylor series
// on [-pi/4, pi/4].
__device__ __forceinline__ double sin_hash_d(double x) {
    const double P1 = 1.570796325802803;       // pi/2, first 30 bits
    const double P2 = 9.920935791635221e-10;   // next 30 bits
    const double P3 = 5.170182981794105e-19;   // remainder
    double k = rint(x * 6.36619772367581382433e-01);
    double r = ((x - k * P1) - k * P2) - k * P3;
    double r2 = r * r;
    double s = r + r * r2 * (-1.0 / 6.0 + r2 * (1.0 / 120.0 + r2 * (-1.0 / 5040.0 + r2 * (1.0 / 362880.0 +
                 r2 * (-1.0 / 39916800.0 + r2 * (1.0 / 6227020800.0 + r2 * (-1.0 / 1307674368000.0 +
                 r2 * (1.0 / 355687428096000.0))))))));
    double c = 1.0 + r2 * (-0.5 + r2 * (1.0 / 24.0 + r2 * (-1.0 / 720.0 + r2 * (1.0 / 40320.0 +
                 r2 * (-1.0 / 3628800.0 + r2 * (1.0 / 479001600.0 + r2 * (-1.0 / 87178291200.0 +
                 r2 * (1.0 / 20922789888000.0 + r2 * (-1.0 / 6402373705728000.0)))))))));
    long long qi = ((long long)k) & 3;
    if (qi == 0) return s;
    if (qi == 1) return c;
    if (qi == 2) return -s;
    return -c;
}

// rendering/taichi/camera.py:9-20: fract(sin(127.1 x + 311.7 y + 74.7 s) * 43758.5453123)
__device__ __forceinline__ float hash_float(int px, int py, float salt) {
    float seed = ((float)px * 127.1f + (float)py * 311.7f) + salt * 74.7f;
    float sv = (float)sin_hash_d((double)seed);
    return fract(sv * 43758.5453123f);
}

// rendering/taichi/camera.py:23-30 with the R2 jitter of :16-20.
__device__ __forceinline__ f3 camera_dir(const rx_camera& cam, int px, int py, int frame) {
    float i = (float)(frame + 1);
    float jx = fract(hash_float(px, py, 0.0f) + i * 0.7548776662466927f);
    float jy = fract(hash_float(px, py, 1.0f) + i * 0.5698402909980532f);
    float u = ((float)px + jx) / (float)cam.width * 2.0f - 1.0f;
    float v = 1.0f - ((float)py + jy) / (float)cam.height * 2.0f;
    float ndc_x = u * cam.aspect * cam.fov_tan;
    float ndc_y = v * cam.fov_tan;
    f3 d = (ld3(cam.fwd) + ld3(cam.right) * ndc_x) + ld3(cam.up) * ndc_y;
    return normalize(d);
}

// ---------------------------------------------------------------- textures
// rendering/taichi/trace.py:96-126 (_sample_texture): repeat, optional v
// flip, bilinear between the four texels around (u (W-1), v (H-1)), u8 /
// 255; a texture of zero size samples as white.
__device__ __forceinline__ f3 sample_texture(const rx_texture& tx, const unsigned char* __restrict__ texels, float tu,
                                             float tv) {
    if (tx.width <= 0 || tx.height <= 0) return mk(1.0f, 1.0f, 1.0f);
    float u = tu - floorf(tu);
    float v = tv - floorf(tv);
    if (tx.flip_v != 0) v = 1.0f - v;
    float x = u * (float)(tx.width - 1);
    float y = v * (float)(tx.height - 1);
    int x0 = (int)floorf(x), y0 = (int)floorf(y);
    int x1 = min(x0 + 1, tx.width - 1), y1 = min(y0 + 1, tx.height - 1);
    float fx = x - (float)x0, fy = y - (float)y0;
    const unsigned char* base = texels + tx.offset;
    auto texel = [&](int yy, int xx) {
        const unsigned char* t = base + 3 * ((long long)yy * tx.width + xx);
        return mk((float)t[0] / 255.0f, (float)t[1] / 255.0f, (float)t[2] / 255.0f);
    };
    f3 c00 = texel(y0, x0), c10 = texel(y0, x1), c01 = texel(y1, x0), c11 = texel(y1, x1);
    f3 c0 = c00 * (1.0f - fx) + c10 * fx;
    f3 c1 = c01 * (1.0f - fx) + c11 * fx;
    return c0 * (1.0f - fy) + c1 * fy;
}

// rendering/taichi/trace.py:129-136 (_bvh_albedo): the material's albedo
// times its texture at the hit's interpolated UV, for triangles with UVs.
__device__ __forceinline__ f3 bvh_albedo(const rx_scene& sc, const rx_material& m, int tri, int has_uv, float bu,
                                         float bv, f3 base) {
    if (m.texture < 0 || has_uv == 0) return base;
    if (m.texture >= sc.n_textures) return base;  // an empty cache slot samples as white
    const float* t = sc.tri_uv + 6 * (long long)tri;
    float w = (1.0f - bu) - bv;
    float tu = (t[0] * w + t[2] * bu) + t[4] * bv;
    float tv = (t[1] * w + t[3] * bu) + t[5] * bv;
    return base * sample_texture(gload(sc.textures + m.texture), sc.texels, tu, tv);
}

// ---------------------------------------------------------------- intersection
// rendering/taichi/intersect.py:17-25
__device__ __forceinline__ f3 inv_dir(f3 d) {
    return mk(fabsf(d.x) > 1e-10f ? 1.0f / d.x : 1e30f, fabsf(d.y) > 1e-10f ? 1.0f / d.y : 1e30f,
              fabsf(d.z) > 1e-10f ? 1.0f / d.z : 1e30f);
}

// rendering/taichi/intersect.py:28-43.  lo/hi point at xyz triples.
__device__ __forceinline__ bool slab(f3 ro, f3 inv, float lx, float ly, float lz, float hx, float hy, float hz,
                                     float limit, float* t_enter_out) {
    float tx0 = (lx - ro.x) * inv.x, tx1 = (hx - ro.x) * inv.x;
    float ty0 = (ly - ro.y) * inv.y, ty1 = (hy - ro.y) * inv.y;
    float tz0 = (lz - ro.z) * inv.z, tz1 = (hz - ro.z) * inv.z;
    float t_enter = fmaxf(fminf(tx0, tx1), fmaxf(fminf(ty0, ty1), fminf(tz0, tz1)));
    float t_exit = fminf(fmaxf(tx0, tx1), fminf(fmaxf(ty0, ty1), fmaxf(tz0, tz1)));
    *t_enter_out = t_enter;
    return t_exit >= t_enter && t_exit >= 0.001f && t_enter <= limit;
}

// Moeller-Trumbore of rendering/taichi/intersect.py:57-78 on (v0, e1, e2).
__device__ __forceinline__ float tri_hit(f3 ro, f3 rd, f3 v0, f3 e1, f3 e2, float* uo, float* vo) {
    f3 h = cross(rd, e2);
    float a = dot(e1, h);
    float t = -1.0f, u = 0.0f, v = 0.0f;
    if (fabsf(a) > 1e-8f) {
        float f = 1.0f / a;
        f3 s = ro - v0;
        u = f * dot(s, h);
        if (u >= 0.0f && u <= 1.0f) {
            f3 q = cross(s, e1);
            v = f * dot(rd, q);
            if (v >= 0.0f && u + v <= 1.0f) {
                float tc = f * dot(e2, q);
                if (tc > 0.001f) t = tc;
            }
        }
    }
    *uo = u;
    *vo = v;
    return t;
}

// scene/primitives.py:90-105
__device__ __forceinline__ float sphere_hit(f3 ro, f3 rd, f3 c, float r) {
    f3 oc = ro - c;
    float h = dot(rd, oc);
    float cc = dot(oc, oc) - r * r;
    float disc = h * h - cc;
    float t = -1.0f;
    if (disc >= 0.0f) {
        float sq = sqrtf(disc);
        t = -h - sq;
        if (t < 0.001f) t = -h + sq;
        if (t < 0.001f) t = -1.0f;
    }
    return t;
}

// scene/primitives.py:145-154
__device__ __forceinline__ float plane_hit(f3 ro, f3 rd, f3 n, float off) {
    float denom = dot(n, rd);
    float t = -1.0f;
    if (fabsf(denom) >= 1e-6f) {
        float tc = -(dot(n, ro) + off) / denom;
        if (tc >= 0.001f) t = tc;
    }
    return t;
}

// scene/primitives.py:187-201 (raw division: inf/NaN resolved by min/max)
__device__ __forceinline__ float cube_hit(f3 ro, f3 rd, f3 c, f3 e) {
    f3 lo = c - e, hi = c + e;
    float tx0 = (lo.x - ro.x) / rd.x, tx1 = (hi.x - ro.x) / rd.x;
    float ty0 = (lo.y - ro.y) / rd.y, ty1 = (hi.y - ro.y) / rd.y;
    float tz0 = (lo.z - ro.z) / rd.z, tz1 = (hi.z - ro.z) / rd.z;
    float t_enter = fmaxf(fminf(tx0, tx1), fmaxf(fminf(ty0, ty1), fminf(tz0, tz1)));
    float t_exit = fminf(fmaxf(tx0, tx1), fminf(fmaxf(ty0, ty1), fmaxf(tz0, tz1)));
    float t = -1.0f;
    if (t_exit >= t_enter && t_exit >= 0.001f) t = t_enter >= 0.001f ? t_enter : t_exit;
    return t;
}

// ---------------------------------------------------------------- eye-relative tests
// Camera rays all start at the eye e.  Their tests subtract e from every box
// plane, triangle vertex and primitive centre, and some terms are built from
// those differences alone; those are the same fp32 values for every camera
// ray, so a pass computes them once (rx_render.hip: k_eye_prep) and the
// camera-ray traversal reads them.  Each function below performs the rest of
// its counterpart's operations above in the same order, so it returns
// bit-identical results for a ray that starts at e.

// slab() with each plane given as plane - e.
__device__ __forceinline__ bool slab_rel(f3 inv, float lx, float ly, float lz, float hx, float hy, float hz,
                                         float limit, float* t_enter_out) {
    float tx0 = lx * inv.x, tx1 = hx * inv.x;
    float ty0 = ly * inv.y, ty1 = hy * inv.y;
    float tz0 = lz * inv.z, tz1 = hz * inv.z;
    float t_enter = fmaxf(fminf(tx0, tx1), fmaxf(fminf(ty0, ty1), fminf(tz0, tz1)));
    float t_exit = fminf(fmaxf(tx0, tx1), fminf(fmaxf(ty0, ty1), fmaxf(tz0, tz1)));
    *t_enter_out = t_enter;
    return t_exit >= t_enter && t_exit >= 0.001f && t_enter <= limit;
}

// tri_hit() with s = e - v0 and q = cross(s, e1) given.
__device__ __forceinline__ float tri_hit_rel(f3 rd, f3 s, f3 e1, f3 e2, f3 q, float* uo, float* vo) {
    f3 h = cross(rd, e2);
    float a = dot(e1, h);
    float t = -1.0f, u = 0.0f, v = 0.0f;
    if (fabsf(a) > 1e-8f) {
        float f = 1.0f / a;
        u = f * dot(s, h);
        if (u >= 0.0f && u <= 1.0f) {
            v = f * dot(rd, q);
            if (v >= 0.0f && u + v <= 1.0f) {
                float tc = f * dot(e2, q);
                if (tc > 0.001f) t = tc;
            }
        }
    }
    *uo = u;
    *vo = v;
    return t;
}

// sphere_hit() with oc = e - c and cc = dot(oc, oc) - r * r given.
__device__ __forceinline__ float sphere_hit_rel(f3 rd, f3 oc, float cc) {
    float h = dot(rd, oc);
    float disc = h * h - cc;
    float t = -1.0f;
    if (disc >= 0.0f) {
        float sq = sqrtf(disc);
        t = -h - sq;
        if (t < 0.001f) t = -h + sq;
        if (t < 0.001f) t = -1.0f;
    }
    return t;
}

// plane_hit() with k = dot(n, e) + off given.
__device__ __forceinline__ float plane_hit_rel(f3 rd, f3 n, float k) {
    float denom = dot(n, rd);
    float t = -1.0f;
    if (fabsf(denom) >= 1e-6f) {
        float tc = -k / denom;
        if (tc >= 0.001f) t = tc;
    }
    return t;
}

// cube_hit() with lo = (c - half) - e and hi = (c + half) - e given.
__device__ __forceinline__ float cube_hit_rel(f3 rd, f3 lo, f3 hi) {
    float tx0 = lo.x / rd.x, tx1 = hi.x / rd.x;
    float ty0 = lo.y / rd.y, ty1 = hi.y / rd.y;
    float tz0 = lo.z / rd.z, tz1 = hi.z / rd.z;
    float t_enter = fmaxf(fminf(tx0, tx1), fmaxf(fminf(ty0, ty1), fminf(tz0, tz1)));
    float t_exit = fminf(fmaxf(tx0, tx1), fminf(fmaxf(ty0, ty1), fmaxf(tz0, tz1)));
    float t = -1.0f;
    if (t_exit >= t_enter && t_exit >= 0.001f) t = t_enter >= 0.001f ? t_enter : t_exit;
    return t;
}

// The eye-relative record of primitive p (32 bytes, written by k_eye_prep):
// sphere {e - c, dot(e - c, e - c) - r * r}; plane {n, dot(n, e) + off};
// cube {lo - e, hi.x - e.x}, {hi.y - e.y, hi.z - e.z}; the type's bits in
// the last word.  Other types keep no eye-relative terms (EYE_PRIM_NONE)
// and are tested from their own record.
constexpr int EYE_PRIM_NONE = -1;
__device__ __forceinline__ void eye_prim(const rx_prim& p, f3 e, gv4f& r0, gv4f& r1) {
    f3 a = mk(p.a[0], p.a[1], p.a[2]);
    r0 = gv4f{0.0f, 0.0f, 0.0f, 0.0f};
    r1 = gv4f{0.0f, 0.0f, 0.0f, __int_as_float(p.type)};
    switch (p.type) {
        case RX_PRIM_SPHERE: {
            f3 oc = e - a;
            float r = p.a[3];
            r0 = gv4f{oc.x, oc.y, oc.z, dot(oc, oc) - r * r};
            break;
        }
        case RX_PRIM_PLANE:
            r0 = gv4f{a.x, a.y, a.z, dot(a, e) + p.a[3]};
            break;
        case RX_PRIM_CUBE: {
            f3 h = mk(p.b[0], p.b[1], p.b[2]);
            f3 lo = a - h, hi = a + h;
            r0 = gv4f{lo.x - e.x, lo.y - e.y, lo.z - e.z, hi.x - e.x};
            r1 = gv4f{hi.y - e.y, hi.z - e.z, 0.0f, __int_as_float(p.type)};
            break;
        }
        default:
            r1.w = __int_as_float(EYE_PRIM_NONE);
            break;
    }
}

// prim_hit() of a camera ray from the eye-relative record (r0, r1) of a
// type that has one (the caller tests the others from their own record).
__device__ __forceinline__ float prim_hit_rel(gv4f r0, gv4f r1, f3 rd) {
    switch (__float_as_int(r1.w)) {
        case RX_PRIM_SPHERE:
            return sphere_hit_rel(rd, mk(r0.x, r0.y, r0.z), r0.w);
        case RX_PRIM_PLANE:
            return plane_hit_rel(rd, mk(r0.x, r0.y, r0.z), r0.w);
        case RX_PRIM_CUBE:
            return cube_hit_rel(rd, mk(r0.x, r0.y, r0.z), mk(r0.w, r1.x, r1.y));
        default:
            return -1.0f;
    }
}

// scene/primitives.py:203-215
__device__ __forceinline__ f3 cube_normal(f3 p, f3 c, f3 e) {
    f3 d = mk((p.x - c.x) / e.x, (p.y - c.y) / e.y, (p.z - c.z) / e.z);
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    if (ax >= ay && ax >= az) return mk(d.x > 0.0f ? 1.0f : -1.0f, 0.0f, 0.0f);
    if (ay >= az) return mk(0.0f, d.y > 0.0f ? 1.0f : -1.0f, 0.0f);
    return mk(0.0f, 0.0f, d.z > 0.0f ? 1.0f : -1.0f);
}

__device__ __forceinline__ float prim_hit(const rx_prim& p, f3 ro, f3 rd) {
    f3 a = mk(p.a[0], p.a[1], p.a[2]);
    switch (p.type) {
        case RX_PRIM_SPHERE:
            return sphere_hit(ro, rd, a, p.a[3]);
        case RX_PRIM_PLANE:
            return plane_hit(ro, rd, a, p.a[3]);
        case RX_PRIM_CUBE:
            return cube_hit(ro, rd, a, mk(p.b[0], p.b[1], p.b[2]));
        case RX_PRIM_TRIANGLE: {
            f3 v1 = mk(p.a[3], p.b[0], p.b[1]), v2 = mk(p.b[2], p.b[3], p.c[0]);
            float u, v;
            return tri_hit(ro, rd, a, v1 - a, v2 - a, &u, &v);
        }
        default:
            return -1.0f;
    }
}

// Outward normal of primitive p at distance t (rendering/taichi/intersect.py:265-282).
__device__ __forceinline__ f3 prim_normal(const rx_prim& p, f3 ro, f3 rd, float t) {
    f3 a = mk(p.a[0], p.a[1], p.a[2]);
    switch (p.type) {
        case RX_PRIM_SPHERE:
            return normalize((ro + rd * t) - a);
        case RX_PRIM_PLANE:
            return a;
        case RX_PRIM_CUBE:
            return cube_normal(ro + rd * t, a, mk(p.b[0], p.b[1], p.b[2]));
        case RX_PRIM_TRIANGLE: {
            f3 v1 = mk(p.a[3], p.b[0], p.b[1]), v2 = mk(p.b[2], p.b[3], p.c[0]);
            return normalize(cross(v1 - a, v2 - a));
        }
        default:
            return mk(0.0f, 0.0f, 0.0f);
    }
}

// ---------------------------------------------------------------- traversal
// Per-lane traversal counters (summed per wave by the caller).
struct trav_count {
    uint32_t pops, tris, prims;
};

// A lane's traversal stack of (link, float bits of t_enter) pairs.  The
// first SL entries live in LDS at lds[k * STRIDE] (STRIDE = threads per
// block, so the 64 lanes of a wave sit on distinct banks whatever their
// depths); deeper entries go to the lane's column of a global spill area,
// spill[(k - SL) * spill_stride].  It is one LIFO either way, so the visit
// order does not depend on SL.  SL = 0 keeps every entry in LDS.
//
// Entries are one 64-bit word (link low, t bits high) and the two parts are
// typed by address space, so the LDS part compiles to ds_read/ds_write_b64
// and the spill to global loads/stores (a generic pointer selected between
// them would become flat accesses, which wait on both the LDS and the
// vector-memory counters).
template <int STRIDE, int SL>
struct lane_stack {
    RX_LDS uint64_t* lds;
    RX_GLOBAL uint64_t* spill;
    int spill_stride;
    __device__ __forceinline__ lane_stack(int2* lds_slot, int2* spill_slot, int stride)
        : lds((RX_LDS uint64_t*)lds_slot), spill((RX_GLOBAL uint64_t*)spill_slot), spill_stride(stride) {}
    __device__ __forceinline__ void put(int k, int link, float t) {
        uint64_t v = (uint64_t)(uint32_t)link | ((uint64_t)__float_as_uint(t) << 32);
        if (SL == 0 || k < SL)
            lds[k * STRIDE] = v;
        else
            spill[(k - SL) * spill_stride] = v;
    }
    __device__ __forceinline__ int2 get(int k) const {
        uint64_t v;
        if (SL == 0 || k < SL)
            v = lds[k * STRIDE];
        else
            v = spill[(k - SL) * spill_stride];
        return make_int2((int)(uint32_t)v, (int)(uint32_t)(v >> 32));
    }
};

// Both child boxes of an interior node, far child pushed first (equal
// entry t: left pushed first, so the right child pops first) —
// rendering/taichi/intersect.py:120-150.  The reference then pops the
// child it pushed last at once; that entry's t_enter passed the same
// limit and nothing has lowered the limit since, so its pop is never
// skipped.  The near child is therefore returned in `next` instead of
// being written to the stack and read back: same visit sequence, one
// stack round trip fewer per node.  Returns false when neither child box
// is hit (the caller pops).
template <class Stack>
__device__ __forceinline__ bool take_children(bool lh, bool rh, float lt, float rt, gv2i d, Stack& stk, int& sp,
                                              int& next) {
    if (lh && rh) {
        if (lt < rt) {
            stk.put(sp++, d.y, rt);
            next = d.x;
        } else {
            stk.put(sp++, d.x, lt);
            next = d.y;
        }
        return true;
    }
    if (lh) {
        next = d.x;
        return true;
    }
    if (rh) {
        next = d.y;
        return true;
    }
    return false;
}

template <class Stack>
__device__ __forceinline__ bool visit_boxes(gv4f a, gv4f b, gv4f c, gv2i d, f3 ro, f3 inv, float limit, Stack& stk,
                                            int& sp, int& next) {
    float lt, rt;
    bool lh = slab(ro, inv, a.x, a.y, a.z, a.w, b.x, b.y, limit, &lt);
    bool rh = slab(ro, inv, b.z, b.w, c.x, c.y, c.z, c.w, limit, &rt);
    return take_children(lh, rh, lt, rt, d, stk, sp, next);
}

// visit_boxes() of a camera ray on a node whose planes are given minus the
// eye (the camera-ray traversal's eye-relative node copy).
template <class Stack>
__device__ __forceinline__ bool visit_boxes_rel(gv4f a, gv4f b, gv4f c, gv2i d, f3 inv, float limit, Stack& stk,
                                                int& sp, int& next) {
    float lt, rt;
    bool lh = slab_rel(inv, a.x, a.y, a.z, a.w, b.x, b.y, limit, &lt);
    bool rh = slab_rel(inv, b.z, b.w, c.x, c.y, c.z, c.w, limit, &rt);
    return take_children(lh, rh, lt, rt, d, stk, sp, next);
}

// The planes of the three 16-byte words of a node's child boxes (minima
// x y z, maxima x y z, twice): the eye's coordinates in that order, so
// word - eye_planes(k % 3, e) is word k minus the eye.
__device__ __forceinline__ gv4f eye_planes(int part, f3 e) {
    return part == 0 ? gv4f{e.x, e.y, e.z, e.x} : (part == 1 ? gv4f{e.y, e.z, e.x, e.y} : gv4f{e.z, e.x, e.y, e.z});
}

template <class Stack>
__device__ __forceinline__ bool visit_interior(const rx_node* __restrict__ nodes, int link, f3 ro, f3 inv,
                                               float limit, Stack& stk, int& sp, int& next) {
    // the two child boxes and the two links: 56 of the record's 64 bytes
    const RX_GLOBAL gv4f* p = (const RX_GLOBAL gv4f*)(nodes + link);
    return visit_boxes(p[0], p[1], p[2], ((const RX_GLOBAL gv2i*)p)[6], ro, inv, limit, stk, sp, next);
}

// Treelet links: node indices stay below TL_REF; TL_REF + k names slot k of
// a block's LDS copy of the top of the tree (at most TL_NODES slots): the
// node's two child boxes as three 16-byte words at box[3 k ..] and its
// links, already rewritten to slots where the child has one, at link[k].
// The same boxes, links and tests as the node record, so the visit
// sequence is visit_interior's.
#ifndef RX_TL_NODES
#define RX_TL_NODES 511
#endif
constexpr int TL_NODES = RX_TL_NODES;
constexpr int TL_REF = 1 << 30;

template <class Stack>
__device__ __forceinline__ bool visit_treelet(const RX_LDS gv4f* box, const RX_LDS gv2i* link, int k, f3 ro, f3 inv,
                                              float limit, Stack& stk, int& sp, int& next) {
    return visit_boxes(box[3 * k], box[3 * k + 1], box[3 * k + 2], link[k], ro, inv, limit, stk, sp, next);
}

// Pops the next entry whose t_enter does not exceed `limit` (entries
// beyond it are discarded, intersect.py:107-108).  Returns false when the
// stack runs empty.
template <class Stack>
__device__ __forceinline__ bool pop_next(const Stack& stk, int& sp, float limit, int& link) {
    while (sp > 0) {
        int2 e = stk.get(--sp);
        if (__int_as_float(e.y) > limit) continue;
        link = e.x;
        return true;
    }
    return false;
}

__device__ __forceinline__ void load_tri(const rx_tri* __restrict__ tris, int i, f3& v0, f3& e1, f3& e2) {
    // v0, e1, e2: the first 36 of the record's 48 bytes
    const RX_GLOBAL gv4f* tp = (const RX_GLOBAL gv4f*)(tris + i);
    gv4f p0 = tp[0], p1 = tp[1];
    float p2 = ((const RX_GLOBAL float*)tp)[8];
    v0 = mk(p0.x, p0.y, p0.z);
    e1 = mk(p0.w, p1.x, p1.y);
    e2 = mk(p1.z, p1.w, p2);
}

// An eye-relative triangle record {s = e - v0, e1, e2, q = cross(s, e1)}.
__device__ __forceinline__ void load_tri_rel(const rx_tri* __restrict__ tris, int i, f3& s, f3& e1, f3& e2, f3& q) {
    const RX_GLOBAL gv4f* tp = (const RX_GLOBAL gv4f*)(tris + i);
    gv4f p0 = tp[0], p1 = tp[1], p2 = tp[2];
    s = mk(p0.x, p0.y, p0.z);
    e1 = mk(p0.w, p1.x, p1.y);
    e2 = mk(p1.z, p1.w, p2.x);
    q = mk(p2.y, p2.z, p2.w);
}

// Closest hit over the BVH: rendering/taichi/intersect.py:81-152.  Visit
// order, pruning and acceptance follow the reference exactly: pop, skip if
// the stored entry t exceeds the current closest t, test a leaf's
// triangles in ascending order with 0.001 < t < closest, or test both
// child boxes against the current closest and push them far child first.
// The loop is arranged "while-while": a lane runs interior nodes until it
// holds a leaf, then the wave's leaves are intersected together.  That
// changes only how lanes interleave, never a lane's own sequence of pops,
// tests and pushes.
template <class Stack, bool COUNT>
__device__ __forceinline__ void bvh_closest(const rx_node* __restrict__ nodes, const rx_tri* __restrict__ tris,
                                            const rx_bvh_header& hdr, f3 ro, f3 rd, Stack stk, float& best_t,
                                            int& best_i, float& best_u, float& best_v, trav_count& cnt) {
    f3 inv = inv_dir(rd);
    best_t = 1e9f;
    best_i = -1;
    best_u = 0.0f;
    best_v = 0.0f;
    float rt;
    if (!slab(ro, inv, hdr.bmin[0], hdr.bmin[1], hdr.bmin[2], hdr.bmax[0], hdr.bmax[1], hdr.bmax[2], best_t, &rt))
        return;
    // the root is pushed and popped at once; its t_enter <= 1e9 = closest
    int sp = 0;
    int cur = hdr.root_link;
    bool have = true;
    while (have) {
        while (have && cur >= 0) {
            if (COUNT) cnt.pops++;
            if (!visit_interior(nodes, cur, ro, inv, best_t, stk, sp, cur)) have = pop_next(stk, sp, best_t, cur);
        }
        if (have) {
            if (COUNT) cnt.pops++;
            int w = ~cur;
            int start = w >> 8, end = start + (w & 255);
            for (int i = start; i < end; ++i) {
                f3 v0, e1, e2;
                load_tri(tris, i, v0, e1, e2);
                float u, v;
                float t = tri_hit(ro, rd, v0, e1, e2, &u, &v);
                if (COUNT) cnt.tris++;
                if (0.001f < t && t < best_t) {
                    best_t = t;
                    best_i = i;
                    best_u = u;
                    best_v = v;
                }
            }
            have = pop_next(stk, sp, best_t, cur);
        }
    }
}

// Any hit before max_t: rendering/taichi/intersect.py:155-219 (returns on
// the first occluder; the boolean does not depend on the visit order).
template <class Stack, bool COUNT>
__device__ __forceinline__ bool bvh_occluded(const rx_node* __restrict__ nodes, const rx_tri* __restrict__ tris,
                                             const rx_bvh_header& hdr, f3 ro, f3 rd, float max_t, Stack stk,
                                             trav_count& cnt) {
    f3 inv = inv_dir(rd);
    float rt;
    if (!slab(ro, inv, hdr.bmin[0], hdr.bmin[1], hdr.bmin[2], hdr.bmax[0], hdr.bmax[1], hdr.bmax[2], max_t, &rt))
        return false;
    int sp = 0;
    int cur = hdr.root_link;
    bool have = rt <= max_t;
    while (have) {
        while (have && cur >= 0) {
            if (COUNT) cnt.pops++;
            if (!visit_interior(nodes, cur, ro, inv, max_t, stk, sp, cur)) have = pop_next(stk, sp, max_t, cur);
        }
        if (have) {
            if (COUNT) cnt.pops++;
            int w = ~cur;
            int start = w >> 8, end = start + (w & 255);
            for (int i = start; i < end; ++i) {
                f3 v0, e1, e2;
                load_tri(tris, i, v0, e1, e2);
                float u, v;
                float t = tri_hit(ro, rd, v0, e1, e2, &u, &v);
                if (COUNT) cnt.tris++;
                if (0.001f < t && t < max_t) return true;
            }
            have = pop_next(stk, sp, max_t, cur);
        }
    }
    return false;
}

// Result of scene_intersect (rendering/taichi/intersect.py:249-309) before
// shading: the analytic primitives are scanned first with strict
// "closer than" acceptance, then the BVH wins only if strictly closer.
struct hit_rec {
    float t;
    int id;
    int kind;
    float u, v;
};

template <class Stack, bool COUNT>
__device__ __forceinline__ hit_rec scene_closest(const rx_scene& sc, const rx_bvh_header& hdr, f3 ro, f3 rd,
                                                 Stack stk, trav_count& cnt) {
    hit_rec h{1e9f, -1, RX_KIND_MISS, 0.0f, 0.0f};
    for (int i = 0; i < sc.n_prims; ++i) {
        float t = prim_hit(gload(sc.prims + i), ro, rd);
        if (COUNT) cnt.prims++;
        if (t > 0.001f && t < h.t) {
            h.t = t;
            h.id = i;
            h.kind = RX_KIND_PRIM;
        }
    }
    if (sc.n_tris > 0) {
        float t, u, v;
        int idx;
        bvh_closest<Stack, COUNT>(sc.nodes, sc.tris, hdr, ro, rd, stk, t, idx, u, v, cnt);
        if (0.001f < t && t < h.t) {
            h.t = t;
            h.id = idx;
            h.kind = RX_KIND_BVH;
            h.u = u;
            h.v = v;
        }
    }
    return h;
}

// rendering/taichi/intersect.py:222-246
template <class Stack, bool COUNT>
__device__ __forceinline__ bool scene_occluded(const rx_scene& sc, const rx_bvh_header& hdr, f3 ro, f3 rd,
                                               float max_t, Stack stk, trav_count& cnt) {
    for (int i = 0; i < sc.n_prims; ++i) {
        float t = prim_hit(gload(sc.prims + i), ro, rd);
        if (COUNT) cnt.prims++;
        if (0.001f < t && t < max_t) return true;
    }
    if (sc.n_tris > 0) return bvh_occluded<Stack, COUNT>(sc.nodes, sc.tris, hdr, ro, rd, max_t, stk, cnt);
    return false;
}

// ---------------------------------------------------------------- wave helpers
// Sum of a per-lane value over the 64 lanes of a wave.  Call only where
// every lane of the wave is active (kernels keep out-of-range lanes alive
// and idle instead of returning early).
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Adds a per-lane count to *dst with one atomic per wave.
__device__ __forceinline__ void wave_add(unsigned long long* dst, unsigned long long x) {
    unsigned long long s = wave_sum_u64(x);
    if ((threadIdx.x & 63) == 0 && s != 0ull) atomicAdd(dst, s);
}

}  // namespace rx
