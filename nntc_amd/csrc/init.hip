// init.hip: the latents' initial state and their grids, on gfx950 (src/init.cu, src/cpu/cpu_init.cpp).
//
//   - level 0 starts at zero, or at 2 luma - 1 on channel 0 under --init0 luma (k_init_level0_cont);
//   - level 1 starts at the block means of source channel j mapped by 2 mean - 1 (k_init_level1_box), with a float
//     accumulator as the reference's;
//   - a grid is fitted per channel over the base and every chain plane together, by min/max or by the 0.1 % / 99.9 %
//     order statistics of the values, and every value is snapped to the nearest grid value (level1_index/value).
//     The order statistics come from a radix select on the device: four passes of 8-bit digit histograms over the
//     order-preserving unsigned image of each float, which lands on exactly the element a full sort would put at
//     that rank;
//   - the residual seed: the centre residual of the whole chain in fp32, its covariance over the base plane summed in
//     the reference's k_cov_partial order (SITE_BLOCKS workgroups of 256 lanes, grid-stride, then a halving tree -
//     the order the reference CPU backend replays because the leading eigenvector amplifies it, src/cpu/cpu_init.cpp
//     :9-15), the host's Jacobi eigensolver, and the projection scaled by its peak over the chain.

#include "arm.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace nntc
{

namespace
{

__global__ __launch_bounds__(256) void k_init_level0_cont(const float* src, float* v0, uint8_t* k0, long long n, int nout,
                                                          int c0, float lr, float lg, float lb, int luma0)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float* s = src + i * nout;
    const float luma = lr * s[0] + lg * s[1] + lb * s[2];
    for (int c = 0; c < c0; c++)
    {
        v0[i * c0 + c] = (c == 0 && luma0) ? 2.0f * luma - 1.0f : 0.0f;
        k0[i * c0 + c] = 0;
    }
}

// The footprint of level-1 texel (x, y): [x w0 / w1, (x + 1) w0 / w1) widened to one pixel and cut at the edge.
NNTC_HD void box_footprint(int x, int y, int w0, int h0, int w1, int h1, int& xa, int& xb, int& ya, int& yb)
{
    xa = (int)((long long)x * w0 / w1);
    xb = (int)((long long)(x + 1) * w0 / w1);
    ya = (int)((long long)y * h0 / h1);
    yb = (int)((long long)(y + 1) * h0 / h1);
    if (xb <= xa)
        xb = xa + 1;
    if (yb <= ya)
        yb = ya + 1;
    xb = xb > w0 ? w0 : xb;
    yb = yb > h0 ? h0 : yb;
}

__global__ __launch_bounds__(256) void k_init_level1_box(Planes planes, const float* src_all, float* out_all, int nout,
                                                         int c1, long long total)
{
    const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total)
        return;
    int p = 0;
    while (p + 1 < planes.n && gi >= planes.p[p + 1].off1)
        p++;
    const PlaneGeom& g = planes.p[p];
    const long long i = gi - g.off1;
    const float* src = src_all + g.off0 * nout;
    const int x = (int)(i % g.w1), y = (int)(i / g.w1);
    int xa, xb, ya, yb;
    box_footprint(x, y, g.w0, g.h0, g.w1, g.h1, xa, xb, ya, yb);
    float acc[MAX_CH];
    for (int c = 0; c < c1; c++)
        acc[c] = 0.0f;
    int n = 0;
    for (int sy = ya; sy < yb; sy++)
        for (int sx = xa; sx < xb; sx++)
        {
            const float* s = src + ((size_t)sy * g.w0 + sx) * nout;
            for (int c = 0; c < c1; c++)
                acc[c] += c < nout ? s[c] : 0.0f;
            n++;
        }
    const float inv = n > 0 ? 1.0f / (float)n : 0.0f;
    for (int c = 0; c < c1; c++)
        out_all[gi * c1 + c] = c < nout ? 2.0f * acc[c] * inv - 1.0f : 0.0f;
}

// Value i snapped onto its channel's grid: v = level1_value(level1_index(in)), k = the index. `in` may be `v`.
__global__ __launch_bounds__(256) void k_quantise(const float* in, float* v, uint8_t* k, long long n, int channels,
                                                  int bits, const float* lo, const float* hi)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int c = (int)(i % channels);
    const int q = level1_index(lo[c], hi[c], bits, in[i]);
    v[i] = level1_value(lo[c], hi[c], bits, q);
    k[i] = (uint8_t)q;
}

// Level-1 values that are not the grid value of the index beside them, counted per workgroup (one atomic each): the
// check of src/main.cpp:2696-2711 (verify_level1_on_grid, src/export.cpp:70-84) over the planes as the device holds them.
struct GridPair
{
    float lo[4], hi[4];
};

__global__ __launch_bounds__(256) void k_off_grid(const float* v, const uint8_t* k, long long n, int channels, int bits,
                                                  GridPair g, unsigned int* off)
{
    __shared__ unsigned int s_off;
    if (threadIdx.x == 0)
        s_off = 0;
    __syncthreads();
    unsigned int mine = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    {
        const int c = (int)(i % channels);
        mine += v[i] != level1_value(g.lo[c], g.hi[c], bits, k[i]) ? 1u : 0u;
    }
    if (mine)
        atomicAdd(&s_off, mine);
    __syncthreads();
    if (threadIdx.x == 0 && s_off)
        atomicAdd(off, s_off);
}

// Per-channel min/max over n interleaved values: partial per workgroup.
__global__ __launch_bounds__(256) void k_minmax(const float* v, long long n, int channels, float* out)
{
    __shared__ float slo[256 * MAX_CH], shi[256 * MAX_CH];
    float lo[MAX_CH], hi[MAX_CH];
    for (int c = 0; c < MAX_CH; c++)
    {
        lo[c] = 1e30f;
        hi[c] = -1e30f;
    }
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t * channels < n; t += stride)
        for (int c = 0; c < channels; c++)
        {
            const float x = v[t * channels + c];
            lo[c] = x < lo[c] ? x : lo[c];
            hi[c] = hi[c] < x ? x : hi[c];
        }
    for (int c = 0; c < MAX_CH; c++)
    {
        slo[threadIdx.x * MAX_CH + c] = lo[c];
        shi[threadIdx.x * MAX_CH + c] = hi[c];
    }
    __syncthreads();
    if (threadIdx.x < MAX_CH)
    {
        const int c = threadIdx.x;
        float l = 1e30f, h = -1e30f;
        for (int t = 0; t < 256; t++)
        {
            const float a = slo[t * MAX_CH + c], b = shi[t * MAX_CH + c];
            l = a < l ? a : l;
            h = h < b ? b : h;
        }
        out[blockIdx.x * 2 * MAX_CH + c] = l;
        out[blockIdx.x * 2 * MAX_CH + MAX_CH + c] = h;
    }
}

// The order-preserving unsigned image of a float: negative floats flip every bit, non-negative ones the sign bit.
__device__ __forceinline__ unsigned int order_key(float f)
{
    const unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// One radix-select pass: for every channel and both ranks, the histogram of digit `shift` over the keys whose higher
// digits equal that rank's prefix so far. hist is channels x 2 x 256 counters.
__global__ __launch_bounds__(256) void k_radix_hist(const float* v, long long n, int channels, int shift,
                                                    const unsigned int* prefix, unsigned int mask, unsigned int* hist)
{
    __shared__ unsigned int h[MAX_CH * 2 * 256];
    for (int i = threadIdx.x; i < MAX_CH * 2 * 256; i += blockDim.x)
        h[i] = 0;
    __syncthreads();
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    {
        const int c = (int)(i % channels);
        const unsigned int key = order_key(v[i]);
        const unsigned int digit = (key >> shift) & 0xFFu;
        for (int r = 0; r < 2; r++)
            if ((key & mask) == prefix[c * 2 + r])
                atomicAdd(&h[(c * 2 + r) * 256 + digit], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < channels * 2 * 256; i += blockDim.x)
        if (h[i])
            atomicAdd(&hist[i], h[i]);
}

// k_cov_partial: entries e of the nout + nout^2 sums over `texels` rows of `nout` floats, lane x of workgroup b taking
// rows b 256 + x, then every SITE_BLOCKS 256 further, and the 256 lanes halved into one. Workgroup (b, g) takes one
// row of the table - the nout means (g = 0) or the nout products of column g - 1 - so a lane reads its rows once for
// nout sums instead of once per sum; each sum keeps its own order.
__global__ __launch_bounds__(256) void k_cov_partial(const float* data, long long texels, int nout, double* partial)
{
    __shared__ double lane[MAX_NOUT][256];
    const int b = blockIdx.x, g = blockIdx.y, x = threadIdx.x;
    const int entries = nout + nout * nout;
    const int r = g - 1;   // the products' first column; -1 for the means
    const long long stride = (long long)SITE_BLOCKS * 256;
    double acc[MAX_NOUT];
#pragma unroll
    for (int s = 0; s < MAX_NOUT; s++)
        acc[s] = 0.0;
    for (long long row = (long long)b * 256; row < texels; row += stride)
    {
        const long long t = row + x;
        if (t < texels)
        {
            const float* d = data + t * nout;
            const double a = r < 0 ? 0.0 : (double)d[r];
#pragma unroll
            for (int s = 0; s < MAX_NOUT; s++)
                if (s < nout)
                    acc[s] += r < 0 ? (double)d[s] : a * (double)d[s];
        }
    }
#pragma unroll
    for (int s = 0; s < MAX_NOUT; s++)
        lane[s][x] = acc[s];
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1)
    {
        if (x < half)
            for (int s = 0; s < nout; s++)
                lane[s][x] += lane[s][x + half];
        __syncthreads();
    }
    if (x < nout)
        partial[(size_t)b * entries + (r < 0 ? x : nout + r * nout + x)] = lane[x][0];
}

struct ProjDir
{
    double e[MAX_NOUT];
    double mu[MAX_NOUT];
};

__device__ __forceinline__ double proj_value(const float* resid, long long t, int nout, const ProjDir& dir)
{
    double acc = 0.0;
    for (int r = 0; r < nout; r++)
        acc += dir.e[r] * ((double)resid[t * nout + r] - dir.mu[r]);
    return acc;
}

__global__ __launch_bounds__(256) void k_proj_peak(const float* resid, long long texels, int nout, ProjDir dir,
                                                   double* out)
{
    __shared__ double red[256];
    double acc = 0.0;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < texels; t += stride)
    {
        const double v = fabs(proj_value(resid, t, nout, dir));
        acc = v > acc ? v : acc;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1)
    {
        if (threadIdx.x < half)
        {
            const double o = red[threadIdx.x + half];
            red[threadIdx.x] = o > red[threadIdx.x] ? o : red[threadIdx.x];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        out[blockIdx.x] = red[0];
}

// The seeded channel: the scaled projection clamped to [-1, 1], continuous under --l0 bc8; in the palette mode
// snapped to index round((value + 1) / 2 levels) of the channel's palette (src/init.cu:1106-1130).
__global__ __launch_bounds__(256) void k_seed_channel(const float* resid, long long texels, int nout, ProjDir dir,
                                                      double inv, int c0, int channel, float* v0, uint8_t* k0,
                                                      const float* palette, int levels)
{
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= texels)
        return;
    const double p = proj_value(resid, t, nout, dir) * inv;
    const float value = (float)(p < -1.0 ? -1.0 : (p > 1.0 ? 1.0 : p));
    if (palette)
    {
        int k = (int)lroundf((value + 1.0f) * 0.5f * (float)levels);
        k = k < 0 ? 0 : (k > levels ? levels : k);
        v0[t * c0 + channel] = palette[channel * 16 + k];
        k0[t * c0 + channel] = (uint8_t)k;
    }
    else
        v0[t * c0 + channel] = value;
}

// The palette mode's divisor: |projection| / peak binned over PROJ_HIST_BINS uniform bins of [0, 1] on every plane,
// 64-bit counts (order-free), src/init.cu:1088-1101.
__global__ __launch_bounds__(256) void k_proj_hist(const float* resid, long long texels, int nout, ProjDir dir,
                                                   double inv_peak, int bins, unsigned long long* hist)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < texels; t += stride)
    {
        const double v = fabs(proj_value(resid, t, nout, dir)) * inv_peak;
        int b = (int)(v * (double)bins);
        b = b < 0 ? 0 : (b >= bins ? bins - 1 : b);
        atomicAdd(hist + b, 1ull);
    }
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        p[i] = 0ull;
}

// The fast form's covariance partials (nout <= 12): one pass over the rows, each lane accumulating the sums and the
// upper triangle of the products of its rows in fp64 registers, then wave trees and the four waves in order; block b
// writes its entries (the full matrix, mirrored) where k_cov_partial would. Every row is read once, where the
// reference-order kernel reads each row once per matrix row.
constexpr int COV_FAST_NOUT = 12;

__global__ __launch_bounds__(256) void k_cov_partial_fast(const float* data, long long texels, int nout, double* partial)
{
    constexpr int N = COV_FAST_NOUT, NU = N * (N + 1) / 2;
    __shared__ double s_w[4][N + NU];
    const int b = blockIdx.x;
    double sum[N], prod[NU];
#pragma unroll
    for (int i = 0; i < N; i++)
        sum[i] = 0.0;
#pragma unroll
    for (int e = 0; e < NU; e++)
        prod[e] = 0.0;
    for (long long t = (long long)b * 256 + threadIdx.x; t < texels; t += (long long)gridDim.x * 256)
    {
        const float* d = data + t * nout;
        double v[N];
#pragma unroll
        for (int i = 0; i < N; i++)
        {
            v[i] = 0.0;
            if (i < nout)   // uniform: only the row's own values are read
                v[i] = (double)d[i];
        }
        int e = 0;
#pragma unroll
        for (int i = 0; i < N; i++)
        {
            sum[i] += v[i];
#pragma unroll
            for (int j = i; j < N; j++, e++)
                prod[e] = fma(v[i], v[j], prod[e]);
        }
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
#pragma unroll
    for (int i = 0; i < N; i++)
    {
        double x = sum[i];
        for (int off = 32; off > 0; off >>= 1)
            x += __shfl_xor(x, off);
        if (lane == 0)
            s_w[wave][i] = x;
    }
#pragma unroll
    for (int e = 0; e < NU; e++)
    {
        double x = prod[e];
        for (int off = 32; off > 0; off >>= 1)
            x += __shfl_xor(x, off);
        if (lane == 0)
            s_w[wave][N + e] = x;
    }
    __syncthreads();
    const int entries = nout + nout * nout;
    for (int q = threadIdx.x; q < entries; q += 256)
    {
        int k;
        if (q < nout)
            k = q;
        else
        {
            const int r = (q - nout) / nout, c = (q - nout) % nout, i = r < c ? r : c, j = r < c ? c : r;
            k = N + i * N - i * (i - 1) / 2 + (j - i);
        }
        partial[(size_t)b * entries + q] = ((s_w[0][k] + s_w[1][k]) + s_w[2][k]) + s_w[3][k];
    }
}

// acc[e] (from zero when `first`) plus weight x partial[b entries + e] for b = 0 .. SITE_BLOCKS - 1, in that order: the
// host fold of each plane's workgroup partials into the raw sums, kept on the device so the chain's planes need one
// readback.
__global__ __launch_bounds__(256) void k_cov_fold(const double* partial, int entries, double weight, int first,
                                                  double* acc)
{
    for (int e = threadIdx.x; e < entries; e += blockDim.x)
    {
        double a = first ? 0.0 : acc[e];
        for (int b = 0; b < SITE_BLOCKS; b++)
            a += weight * partial[(size_t)b * entries + e];
        acc[e] = a;
    }
}

// ---- the source chain built on the device from the textures' bytes (nntc_hip_device_create_rgb8)

struct Rgb8Textures
{
    const uint8_t* t[NNTC_MAX_NOUT / 3];
};

// The packed base plane (src/image.h:298-331, src/main.cpp:1865-1884): texel (x, y) of the padded wp x hp plane takes
// the source pixel (min(x, width - 1), min(y, height - 1)) of every texture, each byte times the float 1/255.
__global__ __launch_bounds__(256) void k_pack_base(Rgb8Textures tex, int textures, int width, int height, int wp,
                                                   long long n, float* base)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int y = (int)(i / wp), x = (int)(i - (long long)y * wp);
    const int sx = x < width ? x : width - 1, sy = y < height ? y : height - 1;
    const size_t at = ((size_t)sy * width + sx) * 3;
    const float inv = 1.0f / 255.0f;
    const int nout = 3 * textures;
    for (int t = 0; t < textures; t++)
        for (int c = 0; c < 3; c++)
            base[(size_t)i * nout + 3 * t + c] = (float)tex.t[t][at + c] * inv;
}

// image_box (src/image.h:343-366): 0.25 ((a + b) + (c + d)) of the 2x2 block, the last row or column repeated on a
// one-wide plane; one lane per output value.
__global__ __launch_bounds__(256) void k_box_halve(const float* in, int w, int h, int nc, float* out, int ow, int oh)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)ow * oh * nc)
        return;
    const int k = (int)(i % nc);
    const long long px = i / nc;
    const int y = (int)(px / ow), x = (int)(px - (long long)y * ow);
    const int xa = 2 * x < w ? 2 * x : w - 1, xb = 2 * x + 1 < w ? 2 * x + 1 : w - 1;
    const int ya = 2 * y < h ? 2 * y : h - 1, yb = 2 * y + 1 < h ? 2 * y + 1 : h - 1;
    const float a = in[((size_t)ya * w + xa) * nc + k], b = in[((size_t)ya * w + xb) * nc + k];
    const float cc = in[((size_t)yb * w + xa) * nc + k], d = in[((size_t)yb * w + xb) * nc + k];
    out[i] = 0.25f * ((a + b) + (cc + d));
}

// differs[s] bit e (e < s): channel s and channel e of the base plane differ somewhere. Each lane ORs its pixels'
// comparisons, each workgroup folds its lanes' masks and adds them with one atomic OR per channel.
__global__ __launch_bounds__(256) void k_channel_differs(const float* base, long long n, int nc, unsigned int* differs)
{
    __shared__ unsigned int s_m[NNTC_MAX_NOUT];
    if (threadIdx.x < NNTC_MAX_NOUT)
        s_m[threadIdx.x] = 0u;
    __syncthreads();
    unsigned int m[NNTC_MAX_NOUT];
    for (int s = 0; s < NNTC_MAX_NOUT; s++)
        m[s] = 0u;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    {
        const float* v = base + (size_t)i * nc;
        for (int s = 1; s < NNTC_MAX_NOUT; s++)
            for (int e = 0; e < s; e++)
                if (s < nc && v[s] != v[e])
                    m[s] |= 1u << e;
    }
    for (int s = 1; s < NNTC_MAX_NOUT; s++)
        if (m[s])
            atomicOr(&s_m[s], m[s]);
    __syncthreads();
    if (threadIdx.x < NNTC_MAX_NOUT && s_m[threadIdx.x])
        atomicOr(&differs[threadIdx.x], s_m[threadIdx.x]);
}

}   // namespace

void init_level0_kernel(Ctx& c, const float* luma, int luma_channel0)
{
    hipLaunchKernelGGL(k_init_level0_cont, dim3(blocks_for(c.n0_total, 256)), dim3(256), 0, c.stream,
                       (const float*)c.src.p, c.v0.p, c.k0.p, c.n0_total, c.nout, c.c0, luma[0], luma[1], luma[2],
                       luma_channel0);
    HIP_OK(hipGetLastError());
}

// The pca init's block means: every source channel's mean over level-1 texel t's footprint (box_footprint), with a
// float accumulator added in footprint order, times the reciprocal of the count (src/init.cu:747-777).
__global__ __launch_bounds__(256) void k_block_means(const float* src, int w0, int h0, int w1, int h1, int nout,
                                                     float* means)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)w1 * h1)
        return;
    const int x = (int)(i % w1), y = (int)(i / w1);
    int xa, xb, ya, yb;
    box_footprint(x, y, w0, h0, w1, h1, xa, xb, ya, yb);
    float acc[MAX_NOUT];
    for (int c = 0; c < nout; c++)
        acc[c] = 0.0f;
    int n = 0;
    for (int sy = ya; sy < yb; sy++)
        for (int sx = xa; sx < xb; sx++)
        {
            const float* s = src + ((size_t)sy * w0 + sx) * nout;
            for (int c = 0; c < nout; c++)
                acc[c] += s[c];
            n++;
        }
    const float inv = n > 0 ? 1.0f / (float)n : 0.0f;
    for (int c = 0; c < nout; c++)
        means[i * nout + c] = acc[c] * inv;
}

void block_means(Ctx& c, int plane, float* out)
{
    const PlaneGeom& g = c.g(plane);
    hipLaunchKernelGGL(k_block_means, dim3(blocks_for(g.n1, 256)), dim3(256), 0, c.stream,
                       (const float*)c.src.p + g.off0 * c.nout, g.w0, g.h0, g.w1, g.h1, c.nout, out);
    HIP_OK(hipGetLastError());
}

void init_level1_box(Ctx& c)
{
    hipLaunchKernelGGL(k_init_level1_box, dim3(blocks_for(c.n1_total, 256)), dim3(256), 0, c.stream, c.planes,
                       (const float*)c.src.p, c.v1.p, c.nout, c.c1, c.n1_total);
    HIP_OK(hipGetLastError());
}

// Snap every value of a latent onto the grid the device holds; `in_device` is the continuous plane (may be the
// plane itself). plane < 0 means every plane of the chain.
void quantise_values(Ctx& c, bool level0, const float* in_device, int plane)
{
    const int ch = level0 ? c.c0 : c.c1;
    const int bits = level0 ? 8 : c.md.bits1;
    float* v = level0 ? c.v0.p : c.v1.p;
    uint8_t* k = level0 ? c.k0.p : c.k1.p;
    long long begin = 0, n = level0 ? c.n0_total : c.n1_total;
    if (plane >= 0)
    {
        begin = level0 ? c.g(plane).off0 : c.g(plane).off1;
        n = level0 ? c.g(plane).n0 : c.g(plane).n1;
    }
    begin *= ch;
    n *= ch;
    const float* lo = c.grid.p + (level0 ? 0 : 8);
    const float* hi = c.grid.p + (level0 ? 4 : 12);
    hipLaunchKernelGGL(k_quantise, dim3(blocks_for(n, 256)), dim3(256), 0, c.stream, in_device + begin, v + begin,
                       k + begin, n, ch, bits, lo, hi);
    HIP_OK(hipGetLastError());
}

unsigned long long level1_off_grid(Ctx& c, const float* lo, const float* hi)
{
    GridPair g {};
    for (int q = 0; q < c.c1 && q < 4; q++)
    {
        g.lo[q] = lo[q];
        g.hi[q] = hi[q];
    }
    const long long n = c.n1_total * c.c1;
    unsigned int* off = c.work_u.p + MAX_CH * 2 * 256 + 32;   // past the radix prefixes in the histograms' tail
    HIP_OK(hipMemsetAsync(off, 0, sizeof(unsigned int), c.stream));
    const unsigned grid = (unsigned)std::min<long long>(blocks_for(n, 256), 2048);
    hipLaunchKernelGGL(k_off_grid, dim3(grid > 0 ? grid : 1), dim3(256), 0, c.stream, c.v1.p, c.k1.p, n, c.c1,
                       c.md.bits1, g, off);
    HIP_OK(hipGetLastError());
    return *read_back(c, (const unsigned int*)off, 1);
}

// The per-channel range of a latent over every plane (src/cpu/cpu_init.cpp:338-370).
void fit_range(Ctx& c, bool level0, int policy, float* lo, float* hi)
{
    const int ch = level0 ? c.c0 : c.c1;
    const float* v = level0 ? c.v0.p : c.v1.p;
    const long long n = (level0 ? c.n0_total : c.n1_total) * ch;
    for (int q = 0; q < ch; q++)
    {
        lo[q] = 1e30f;
        hi[q] = -1e30f;
    }
    if (n == 0)
        return;
    if (policy == NNTC_RANGE_MINMAX)
    {
        const int blocks = 256;
        float* out = c.work_f.p;
        hipLaunchKernelGGL(k_minmax, dim3(blocks), dim3(256), 0, c.stream, v, n, ch, out);
        HIP_OK(hipGetLastError());
        const float* h = read_back(c, (const float*)out, (size_t)blocks * 2 * MAX_CH);
        for (int b = 0; b < blocks; b++)
            for (int q = 0; q < ch; q++)
            {
                lo[q] = std::min(lo[q], h[(size_t)b * 2 * MAX_CH + q]);
                hi[q] = std::max(hi[q], h[(size_t)b * 2 * MAX_CH + MAX_CH + q]);
            }
        return;
    }
    // pct: ranks a = round(0.001 last), b = round(0.999 last) of each channel's sorted values.
    const long long per = n / ch;
    const long long last = per - 1;
    const unsigned long long rank0 = (unsigned long long)(0.001 * (double)last + 0.5);
    const unsigned long long rank1 = (unsigned long long)(0.999 * (double)last + 0.5);
    unsigned int prefix[MAX_CH * 2] = {};
    unsigned long long want[MAX_CH * 2];
    for (int q = 0; q < ch; q++)
    {
        want[q * 2] = rank0;
        want[q * 2 + 1] = rank1;
    }
    unsigned int mask = 0;
    unsigned int* d_hist = c.work_u.p;
    unsigned int* d_prefix = c.work_u.p + MAX_CH * 2 * 256;
    for (int shift = 24; shift >= 0; shift -= 8)
    {
        HIP_OK(hipMemsetAsync(d_hist, 0, sizeof(unsigned int) * MAX_CH * 2 * 256, c.stream));
        upload_staged(c, d_prefix, prefix, MAX_CH * 2);
        const unsigned blocks = (unsigned)std::min<long long>(1024, (n + 255) / 256);
        hipLaunchKernelGGL(k_radix_hist, dim3(blocks), dim3(256), 0, c.stream, v, n, ch, shift,
                           (const unsigned int*)d_prefix, mask, d_hist);
        HIP_OK(hipGetLastError());
        const unsigned int* hist = read_back(c, (const unsigned int*)d_hist, (size_t)ch * 2 * 256);
        for (int i = 0; i < ch * 2; i++)
        {
            unsigned long long cum = 0;
            int digit = 255;
            for (int d = 0; d < 256; d++)
            {
                const unsigned long long cnt = hist[(size_t)i * 256 + d];
                if (want[i] < cum + cnt)
                {
                    digit = d;
                    break;
                }
                cum += cnt;
            }
            want[i] -= cum;
            prefix[i] |= (unsigned int)digit << shift;
        }
        mask |= 0xFFu << shift;
    }
    for (int q = 0; q < ch; q++)
        for (int r = 0; r < 2; r++)
        {
            const unsigned int key = prefix[q * 2 + r];
            const unsigned int u = (key & 0x80000000u) ? (key & 0x7FFFFFFFu) : ~key;
            float f;
            std::memcpy(&f, &u, sizeof(f));
            (r == 0 ? lo : hi)[q] = f;
        }
}

void cov_accumulate(Ctx& c, const float* data, long long texels, double weight, bool first)
{
    const int entries = c.nout + c.nout * c.nout;
    if (c.form != NNTC_FORM_EXACT && c.nout <= COV_FAST_NOUT)
        hipLaunchKernelGGL(k_cov_partial_fast, dim3(SITE_BLOCKS), dim3(256), 0, c.stream, data, texels, c.nout,
                           c.cov_partial.p);
    else
        hipLaunchKernelGGL(k_cov_partial, dim3(SITE_BLOCKS, c.nout + 1), dim3(256), 0, c.stream, data, texels, c.nout,
                           c.cov_partial.p);
    hipLaunchKernelGGL(k_cov_fold, dim3(1), dim3(256), 0, c.stream, (const double*)c.cov_partial.p, entries, weight,
                       first ? 1 : 0, c.cov_partial.p + (size_t)SITE_BLOCKS * entries);
    HIP_OK(hipGetLastError());
}

void cov_read(Ctx& c, std::vector<double>& raw)
{
    const int entries = c.nout + c.nout * c.nout;
    const double* h = read_back(c, (const double*)c.cov_partial.p + (size_t)SITE_BLOCKS * entries, (size_t)entries);
    raw.assign(h, h + entries);
}

double proj_peak_chain(Ctx& c, const float* resid, const double* e, const double* mu)
{
    ProjDir dir;
    for (int r = 0; r < MAX_NOUT; r++)
    {
        dir.e[r] = r < c.nout ? e[r] : 0.0;
        dir.mu[r] = r < c.nout ? mu[r] : 0.0;
    }
    // Plane p's workgroup peaks at out[256 p ..]: every plane's launch first, one readback, the maximum on the host.
    double* out = c.cov_partial.p;
    std::vector<unsigned> blocks((size_t)c.planes.n);
    for (int p = 0; p < c.planes.n; p++)
    {
        const long long texels = c.g(p).n0;
        blocks[(size_t)p] = (unsigned)std::max<long long>(1, std::min<long long>(256, (texels + 255) / 256));
        hipLaunchKernelGGL(k_proj_peak, dim3(blocks[(size_t)p]), dim3(256), 0, c.stream, resid + c.g(p).off0 * c.nout,
                           texels, c.nout, dir, out + (size_t)p * 256);
    }
    HIP_OK(hipGetLastError());
    const double* h = read_back(c, (const double*)out, (size_t)c.planes.n * 256);
    double peak = 0.0;
    for (int p = 0; p < c.planes.n; p++)
        for (unsigned b = 0; b < blocks[(size_t)p]; b++)
            peak = std::max(peak, h[(size_t)p * 256 + b]);
    return peak;
}

void seed_channel(Ctx& c, int plane, const float* resid, const double* e, const double* mu, double inv, int channel)
{
    ProjDir dir;
    for (int r = 0; r < MAX_NOUT; r++)
    {
        dir.e[r] = r < c.nout ? e[r] : 0.0;
        dir.mu[r] = r < c.nout ? mu[r] : 0.0;
    }
    const PlaneGeom& g = c.g(plane);
    const bool palette = !c.md.l0_bc8;
    hipLaunchKernelGGL(k_seed_channel, dim3(blocks_for(g.n0, 256)), dim3(256), 0, c.stream, resid, g.n0, c.nout, dir,
                       inv, c.c0, channel, c.v0.p + g.off0 * c.c0, c.k0.p + g.off0 * c.c0,
                       palette ? (const float*)(c.grid.p + 16) : nullptr, (1 << c.md.bits0[channel]) - 1);
    HIP_OK(hipGetLastError());
}

double proj_percentile_chain(Ctx& c, const float* resid, const double* e, const double* mu, double peak)
{
    constexpr int BINS = 4096;                 // PROJ_HIST_BINS (src/model.h:312)
    constexpr double PERCENTILE = 0.99;        // INIT0_PERCENTILE (src/model.h:311)
    ProjDir dir;
    for (int r = 0; r < MAX_NOUT; r++)
    {
        dir.e[r] = r < c.nout ? e[r] : 0.0;
        dir.mu[r] = r < c.nout ? mu[r] : 0.0;
    }
    if (c.cov_partial.n < (size_t)BINS)
        throw Refused { "init0: the histogram does not fit the covariance partials" };
    unsigned long long* hist = reinterpret_cast<unsigned long long*>(c.cov_partial.p);
    hipLaunchKernelGGL(k_zero_u64, dim3(BINS / 256), dim3(256), 0, c.stream, hist, BINS);
    unsigned long long total = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        hipLaunchKernelGGL(k_proj_hist, dim3(SITE_BLOCKS), dim3(256), 0, c.stream, resid + c.g(p).off0 * c.nout,
                           c.g(p).n0, c.nout, dir, 1.0 / peak, BINS, hist);
        total += (unsigned long long)c.g(p).n0;
    }
    HIP_OK(hipGetLastError());
    const unsigned long long* h = read_back(c, (const unsigned long long*)hist, (size_t)BINS);
    const unsigned long long want = (unsigned long long)(PERCENTILE * (double)total);
    unsigned long long cum = 0;
    int bin = BINS - 1;
    for (int b = 0; b < BINS; b++)
    {
        cum += h[b];
        if (cum >= want)
        {
            bin = b;
            break;
        }
    }
    const double divisor = peak * (double)(bin + 1) / (double)BINS;
    return divisor > 0.0 ? divisor : peak;
}

}   // namespace nntc

namespace nntc
{

void build_source_chain(Ctx& c, const uint8_t* const* rgb8_device, int width, int height)
{
    Rgb8Textures tex {};
    const int textures = c.nout / 3;
    for (int t = 0; t < textures; t++)
        tex.t[t] = rgb8_device[t];
    const PlaneGeom& g0 = c.g(0);
    hipLaunchKernelGGL(k_pack_base, dim3(blocks_for(g0.n0, 256)), dim3(256), 0, c.stream, tex, textures, width, height,
                       g0.w0, g0.n0, c.src.p);
    for (int p = 1; p < c.planes.n; p++)
    {
        const PlaneGeom &a = c.g(p - 1), &b = c.g(p);
        hipLaunchKernelGGL(k_box_halve, dim3(blocks_for(b.n0 * c.nout, 256)), dim3(256), 0, c.stream,
                           (const float*)(c.src.p + a.off0 * c.nout), a.w0, a.h0, c.nout, c.src.p + b.off0 * c.nout,
                           b.w0, b.h0);
    }
    HIP_OK(hipGetLastError());
}

void channel_copies(Ctx& c, unsigned int* copies)
{
    unsigned int* d = c.work_u.p;   // the radix-select histograms' area, free outside a grid fit
    HIP_OK(hipMemsetAsync(d, 0, NNTC_MAX_NOUT * sizeof(unsigned int), c.stream));
    hipLaunchKernelGGL(k_channel_differs, dim3(1024), dim3(256), 0, c.stream, (const float*)c.src.p, c.g(0).n0, c.nout,
                       d);
    HIP_OK(hipGetLastError());
    const unsigned int* h = read_back(c, (const unsigned int*)d, NNTC_MAX_NOUT);
    for (int s = 0; s < c.nout; s++)
        copies[s] = ~h[s] & ((1u << s) - 1u);
}

}   // namespace nntc
