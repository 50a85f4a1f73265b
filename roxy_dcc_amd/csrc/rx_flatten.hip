// rx_flatten.hip — host-side scene flattening of indexed meshes:
// scene/mesh.py:397-430 (indexed_triangle_arrays) with the transforms of
// :547-568, as rendering/taichi/extractor.py:174-190 calls it for every
// IndexedMesh object.  Native C++ on the host cores (the arrays feed the
// BVH build and upload that follow), parallel over triangles.
//
// fp64 throughout, one fixed operation order, no contraction (the library
// is built with -ffp-contract=off):
//   p'_j = ((x m_j0 + y m_j1) + z m_j2) + m_j3,   j = 0..2
//   w    = ((x m_30 + y m_31) + z m_32) + m_33,   p' /= w where |w| > 1e-12
//   n'_j = (n_x N_j0 + n_y N_j1) + n_z N_j2,      then n' / |n'| where |n'| > 1e-12
//   face = normalise(cross(v1 - v0, v2 - v0)) on the untransformed corners,
//          used for corners without a normal index
// then a cast to fp32.  The reference writes the products as numpy matmuls,
// whose rounding is that of the BLAS it links; this order is the one the
// Python restatement (roxy_dcc_amd/world.py) spells out elementwise.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <thread>
#include <vector>

#include "rx_common.h"

namespace {

struct xf {
    double m[16];
    bool on;
};

inline void load_xf(const double* m16, xf& x) {
    x.on = m16 != nullptr;
    if (x.on) memcpy(x.m, m16, sizeof(x.m));
}

inline void xf_point(const xf& x, const double* p, double* o) {
    if (!x.on) {
        o[0] = p[0];
        o[1] = p[1];
        o[2] = p[2];
        return;
    }
    const double* m = x.m;
    double r[3];
    for (int j = 0; j < 3; ++j) r[j] = ((p[0] * m[4 * j] + p[1] * m[4 * j + 1]) + p[2] * m[4 * j + 2]) + m[4 * j + 3];
    double w = ((p[0] * m[12] + p[1] * m[13]) + p[2] * m[14]) + m[15];
    if (fabs(w) > 1e-12) {
        for (int j = 0; j < 3; ++j) r[j] = r[j] / w;
    }
    o[0] = r[0];
    o[1] = r[1];
    o[2] = r[2];
}

inline void xf_vector(const xf& x, const double* v, double* o) {
    if (!x.on) {
        o[0] = v[0];
        o[1] = v[1];
        o[2] = v[2];
        return;
    }
    const double* m = x.m;
    for (int j = 0; j < 3; ++j) o[j] = (v[0] * m[4 * j] + v[1] * m[4 * j + 1]) + v[2] * m[4 * j + 2];
}

inline void normalise(double* v) {
    double len = sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    if (len > 1e-12) {
        v[0] = v[0] / len;
        v[1] = v[1] / len;
        v[2] = v[2] / len;
    }
}

struct job {
    const double* pos;
    int n_pos;
    const int32_t* tri_pos;
    const double* nrm;
    int n_nrm;
    const int32_t* tri_nrm;
    const double* uv;
    int n_uv;
    const int32_t* tri_uv;
    xf m, nm;
    float *v[3], *n[3], *t[3];
    int32_t* has_uv;
};

void run(const job& J, long long lo, long long hi) {
    for (long long i = lo; i < hi; ++i) {
        double c[3][3];
        for (int k = 0; k < 3; ++k) memcpy(c[k], J.pos + 3 * (long long)J.tri_pos[3 * i + k], sizeof(c[k]));
        // face normal from the untransformed corners
        double e1[3] = {c[1][0] - c[0][0], c[1][1] - c[0][1], c[1][2] - c[0][2]};
        double e2[3] = {c[2][0] - c[0][0], c[2][1] - c[0][1], c[2][2] - c[0][2]};
        double face[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        normalise(face);
        bool per_vertex = J.nrm && J.tri_nrm && J.tri_nrm[3 * i] >= 0 && J.tri_nrm[3 * i + 1] >= 0 &&
                          J.tri_nrm[3 * i + 2] >= 0;
        bool has_uv = J.uv && J.tri_uv && J.tri_uv[3 * i] >= 0 && J.tri_uv[3 * i + 1] >= 0 && J.tri_uv[3 * i + 2] >= 0;
        for (int k = 0; k < 3; ++k) {
            double p[3], nv[3], n_in[3];
            xf_point(J.m, c[k], p);
            if (per_vertex)
                memcpy(n_in, J.nrm + 3 * (long long)J.tri_nrm[3 * i + k], sizeof(n_in));
            else
                memcpy(n_in, face, sizeof(n_in));
            xf_vector(J.nm, n_in, nv);
            normalise(nv);
            for (int j = 0; j < 3; ++j) {
                J.v[k][3 * i + j] = (float)p[j];
                J.n[k][3 * i + j] = (float)nv[j];
            }
            if (has_uv) {
                const double* u = J.uv + 2 * (long long)J.tri_uv[3 * i + k];
                J.t[k][2 * i] = (float)u[0];
                J.t[k][2 * i + 1] = (float)u[1];
            } else {
                J.t[k][2 * i] = 0.0f;
                J.t[k][2 * i + 1] = 0.0f;
            }
        }
        J.has_uv[i] = has_uv ? 1 : 0;
    }
}

}  // namespace

extern "C" int rx_flatten_indexed(const double* positions, int n_positions, const int32_t* tri_pos_idx, int n_tris,
                                  const double* normals, int n_normals, const int32_t* tri_normal_idx,
                                  const double* uvs, int n_uvs, const int32_t* tri_uv_idx, const double* matrix,
                                  const double* normal_matrix, float* v0, float* v1, float* v2, float* n0, float* n1,
                                  float* n2, float* uv0, float* uv1, float* uv2, int32_t* has_uv, int threads) {
    RX_REQUIRE(n_tris >= 0 && n_positions >= 0 && n_normals >= 0 && n_uvs >= 0, RX_ERR_INVALID,
               "negative array size");
    if (n_tris == 0) return RX_OK;
    RX_REQUIRE(positions && tri_pos_idx, RX_ERR_INVALID, "positions and triangle indices are required");
    RX_REQUIRE(v0 && v1 && v2 && n0 && n1 && n2 && uv0 && uv1 && uv2 && has_uv, RX_ERR_INVALID, "NULL output");
    // indices must address the arrays they index (a negative normal or uv
    // index means "none" for that triangle)
    for (long long i = 0; i < 3LL * n_tris; ++i) {
        RX_REQUIRE(tri_pos_idx[i] >= 0 && tri_pos_idx[i] < n_positions, RX_ERR_INVALID,
                   "position index %d out of range at %lld", tri_pos_idx[i], i);
        if (normals && tri_normal_idx)
            RX_REQUIRE(tri_normal_idx[i] < n_normals, RX_ERR_INVALID, "normal index out of range at %lld", i);
        if (uvs && tri_uv_idx) RX_REQUIRE(tri_uv_idx[i] < n_uvs, RX_ERR_INVALID, "uv index out of range at %lld", i);
    }
    job J;
    J.pos = positions;
    J.n_pos = n_positions;
    J.tri_pos = tri_pos_idx;
    J.nrm = normals;
    J.n_nrm = n_normals;
    J.tri_nrm = tri_normal_idx;
    J.uv = uvs;
    J.n_uv = n_uvs;
    J.tri_uv = tri_uv_idx;
    load_xf(matrix, J.m);
    load_xf(normal_matrix, J.nm);
    J.v[0] = v0;
    J.v[1] = v1;
    J.v[2] = v2;
    J.n[0] = n0;
    J.n[1] = n1;
    J.n[2] = n2;
    J.t[0] = uv0;
    J.t[1] = uv1;
    J.t[2] = uv2;
    J.has_uv = has_uv;
    int hw = (int)std::thread::hardware_concurrency();
    int nt = threads > 0 ? threads : (hw > 0 ? hw : 1);
    long long per_thread = 16384;
    nt = (int)std::max(1LL, std::min<long long>(nt, (n_tris + per_thread - 1) / per_thread));
    if (nt == 1) {
        run(J, 0, n_tris);
        return RX_OK;
    }
    std::vector<std::thread> pool;
    long long chunk = (n_tris + nt - 1) / nt;
    for (int k = 0; k < nt; ++k) {
        long long lo = k * chunk, hi = std::min<long long>(n_tris, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back(run, std::cref(J), lo, hi);
    }
    for (auto& t : pool) t.join();
    return RX_OK;
}
