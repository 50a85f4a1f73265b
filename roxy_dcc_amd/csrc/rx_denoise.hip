// rx_denoise.hip — guided bilateral ("edge-aware") filter of
// rendering/denoise.py:23-104, generalised to dilated taps for a-trous
// passes (dilation 1 is the reference filter).
//
// Tiling: a block owns a 16x16 tile of outputs on one residue lattice of
// the dilation d, i.e. pixels (ry + d*ky, rx + d*kx).  Every tap of such an
// output, (y + d*dy, x + d*dx), is again a lattice point, so the block's
// whole neighbourhood is a compact (16+2r)^2 tile of lattice points staged
// once in LDS (colour, normal, albedo, depth: 10 floats a point), with the
// image-border reflection applied to each staged point.  Each input pixel
// is read from HBM once per pass; the (2r+1)^2 taps come from LDS.
#include "rx_common.h"
#include "rx_device.h"

namespace {

constexpr int T = 16;         // tile edge (outputs)
constexpr int MAX_R = 8;      // largest radius the LDS tile supports
constexpr int MAX_E = T + 2 * MAX_R;

// numpy.pad(mode="reflect") index map: mirror about the edge pixels
// without repeating them, period 2(n-1).
__device__ __forceinline__ int reflect_idx(int i, int n) {
    if (n == 1) return 0;
    int period = 2 * (n - 1);
    i %= period;
    if (i < 0) i += period;
    return i < n ? i : period - i;
}

struct dn_params {
    int W, H, r, d;
    int has_n, has_a, has_d;
    float inv_c, inv_n, inv_a, inv_d;  // divisors 2*sigma^2 (fp32)
    float amount, one_minus_amount;
    int lat_w, lat_h;  // lattice tiles per residue in x and y
};

__global__ __launch_bounds__(T* T) void k_denoise(const float* __restrict__ col, const float* __restrict__ nrm,
                                                  const float* __restrict__ alb, const float* __restrict__ dep,
                                                  dn_params p, float* __restrict__ out) {
    extern __shared__ float lds[];
    const int E = T + 2 * p.r;
    const int EE = E * E;
    float* sc = lds;            // 3*EE colour
    float* sn = sc + 3 * EE;    // 3*EE normal
    float* sa = sn + 3 * EE;    // 3*EE albedo
    float* sd = sa + 3 * EE;    // EE depth

    // block -> (residue ry, rx; lattice tile ty0, tx0)
    int b = blockIdx.x;
    int tiles = p.lat_w * p.lat_h;
    int res = b / tiles, tb = b % tiles;
    int ry = res / p.d, rx = res % p.d;
    int ky0 = (tb / p.lat_w) * T, kx0 = (tb % p.lat_w) * T;

    for (int k = threadIdx.x; k < EE; k += T * T) {
        int ly = k / E, lx = k % E;
        int y = reflect_idx(ry + (ky0 + ly - p.r) * p.d, p.H);
        int x = reflect_idx(rx + (kx0 + lx - p.r) * p.d, p.W);
        int src = y * p.W + x;
        sc[3 * k] = col[3 * src];
        sc[3 * k + 1] = col[3 * src + 1];
        sc[3 * k + 2] = col[3 * src + 2];
        if (p.has_n) {
            sn[3 * k] = nrm[3 * src];
            sn[3 * k + 1] = nrm[3 * src + 1];
            sn[3 * k + 2] = nrm[3 * src + 2];
        }
        if (p.has_a) {
            sa[3 * k] = alb[3 * src];
            sa[3 * k + 1] = alb[3 * src + 1];
            sa[3 * k + 2] = alb[3 * src + 2];
        }
        if (p.has_d) sd[k] = dep[src];
    }
    __syncthreads();

    int ty = threadIdx.x / T, tx = threadIdx.x % T;
    int y = ry + (ky0 + ty) * p.d, x = rx + (kx0 + tx) * p.d;
    if (y >= p.H || x >= p.W) return;
    int c = (ty + p.r) * E + (tx + p.r);
    float c0 = sc[3 * c], c1 = sc[3 * c + 1], c2 = sc[3 * c + 2];
    float n0 = 0, n1 = 0, n2 = 0, a0 = 0, a1 = 0, a2 = 0, d0 = 0;
    if (p.has_n) {
        n0 = sn[3 * c];
        n1 = sn[3 * c + 1];
        n2 = sn[3 * c + 2];
    }
    if (p.has_a) {
        a0 = sa[3 * c];
        a1 = sa[3 * c + 1];
        a2 = sa[3 * c + 2];
    }
    if (p.has_d) d0 = sd[c];
    float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, wsum = 0.0f;
    for (int dy = -p.r; dy <= p.r; ++dy) {
        for (int dx = -p.r; dx <= p.r; ++dx) {
            int k = (ty + p.r + dy) * E + (tx + p.r + dx);
            float q0 = sc[3 * k], q1 = sc[3 * k + 1], q2 = sc[3 * k + 2];
            float e0 = q0 - c0, e1 = q1 - c1, e2 = q2 - c2;
            float cd2 = (e0 * e0 + e1 * e1) + e2 * e2;
            float sw = (float)(1.0 / (1.0 + (double)(dx * dx) + (double)(dy * dy)));
            float wgt = sw * expf(-cd2 / p.inv_c);
            if (p.has_n) {
                float g0 = sn[3 * k] - n0, g1 = sn[3 * k + 1] - n1, g2 = sn[3 * k + 2] - n2;
                wgt *= expf(-((g0 * g0 + g1 * g1) + g2 * g2) / p.inv_n);
            }
            if (p.has_a) {
                float g0 = sa[3 * k] - a0, g1 = sa[3 * k + 1] - a1, g2 = sa[3 * k + 2] - a2;
                wgt *= expf(-((g0 * g0 + g1 * g1) + g2 * g2) / p.inv_a);
            }
            if (p.has_d) {
                float g0 = sd[k] - d0;
                wgt *= expf(-(g0 * g0) / p.inv_d);
            }
            acc0 += q0 * wgt;
            acc1 += q1 * wgt;
            acc2 += q2 * wgt;
            wsum += wgt;
        }
    }
    float den = fmaxf(wsum, 1e-8f);
    int o = 3 * (y * p.W + x);
    out[o] = c0 * p.one_minus_amount + (acc0 / den) * p.amount;
    out[o + 1] = c1 * p.one_minus_amount + (acc1 / den) * p.amount;
    out[o + 2] = c2 * p.one_minus_amount + (acc2 / den) * p.amount;
}

__global__ void k_copy(const float* __restrict__ a, float* __restrict__ b, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) b[i] = a[i];
}

// rendering/denoise.py:107-108
__global__ void k_display(const float* __restrict__ a, float* __restrict__ b, long long n) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        b[i] = sqrtf(fminf(fmaxf(a[i], 0.0f), 1.0f));
}

float two_sigma2(float sigma) {
    double s2 = (double)sigma * (double)sigma;
    if (s2 < 1e-8) s2 = 1e-8;
    return (float)(2.0 * s2);
}

}  // namespace

extern "C" int rx_denoise(const float* d_color, const float* d_normal, const float* d_albedo, const float* d_depth,
                          int width, int height, int radius, int dilation, float sigma_color, float sigma_normal,
                          float sigma_albedo, float sigma_depth, float amount, float* d_out, void* stream) {
    RX_REQUIRE(width > 0 && height > 0, RX_ERR_INVALID, "image size must be positive");
    RX_REQUIRE(d_color && d_out, RX_ERR_INVALID, "NULL argument");
    RX_REQUIRE(d_color != d_out, RX_ERR_INVALID, "output must not alias the input colour");
    RX_REQUIRE(dilation >= 1, RX_ERR_INVALID, "dilation must be >= 1");
    hipStream_t s = rx_stream(stream);
    long long n = 3LL * width * height;
    if (radius <= 0 || amount <= 0.0f) {
        // denoise.py:41-42 returns the input unchanged
        k_copy<<<(int)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096), 256, 0, s>>>(d_color, d_out, n);
        RX_CHECK_LAUNCH();
        return RX_OK;
    }
    RX_REQUIRE(radius <= MAX_R, RX_ERR_INVALID, "radius %d exceeds %d", radius, MAX_R);
    dn_params p;
    p.W = width;
    p.H = height;
    p.r = radius;
    p.d = dilation;
    p.has_n = d_normal != nullptr;
    p.has_a = d_albedo != nullptr;
    p.has_d = d_depth != nullptr;
    p.inv_c = two_sigma2(sigma_color);
    p.inv_n = two_sigma2(sigma_normal);
    p.inv_a = two_sigma2(sigma_albedo);
    p.inv_d = two_sigma2(sigma_depth);
    float am = amount < 0.0f ? 0.0f : (amount > 1.0f ? 1.0f : amount);
    p.amount = am;
    p.one_minus_amount = (float)(1.0 - (double)am);
    // lattice extent of residue 0 (the longest) in tiles
    int lat_y = (height + dilation - 1) / dilation, lat_x = (width + dilation - 1) / dilation;
    p.lat_h = (lat_y + T - 1) / T;
    p.lat_w = (lat_x + T - 1) / T;
    long long blocks = (long long)dilation * dilation * p.lat_h * p.lat_w;
    RX_REQUIRE(blocks <= 0x7fffffffLL, RX_ERR_CAPACITY, "image too large");
    int E = T + 2 * radius;
    size_t lds = sizeof(float) * 10 * (size_t)E * E;
    k_denoise<<<(int)blocks, T * T, lds, s>>>(d_color, d_normal, d_albedo, d_depth, p, d_out);
    RX_CHECK_LAUNCH();
    return RX_OK;
}

extern "C" int rx_display(const float* d_linear, int n, float* d_out, void* stream) {
    RX_REQUIRE(n >= 0, RX_ERR_INVALID, "n must be >= 0");
    if (n == 0) return RX_OK;
    RX_REQUIRE(d_linear && d_out, RX_ERR_INVALID, "NULL argument");
    long long g = ((long long)n + 255) / 256;
    if (g > 4096) g = 4096;
    k_display<<<(int)g, 256, 0, rx_stream(stream)>>>(d_linear, d_out, n);
    RX_CHECK_LAUNCH();
    return RX_OK;
}
