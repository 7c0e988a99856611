// nntc_common.h: the sampling rule, the feature order, the grid math and the BC4 palette, for the gfx950 kernels and
// the arm's host code alike.
//
// Every function here restates a rule the reference fixes, operation for operation, because the encode's promise is
// that the value the solver measured and the value the consumer samples are the same number:
//   - the bilinear clamp-to-edge addressing from pixel coordinates and the four-tap blend (src/sample.cuh:43-72,
//     src/cpu/cpu_sample.h:36-66);
//   - phi = [c_j][s_i][s_i c_j] (src/sample.cuh:74-85);
//   - the K = 1, 2, 4 subtexel offsets turned a quarter per texel of a 2x2 cell (src/sample.cuh:87-128);
//   - level1_value / level1_index / replicate_index (src/model.h:141-199);
//   - the BC4 eight-value palette entry and the plane value of a decoded byte (src/refine_bc.cu:102-131).
// All of it is compiled with -ffp-contract=off: the reference's host build and its CUDA build both forbid fused
// multiply-adds (CMakeLists.txt:183-187, 206-208), and a fused add here would move last bits against it.

#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define NNTC_HD __host__ __device__ __forceinline__

namespace nntc
{

constexpr int MAX_PLANES = 13;
constexpr int MAX_CH = 4;
constexpr int MAX_NIN = 24;
constexpr int MAX_NOUT = 18;
constexpr int STENCIL_BLOCKS = 5;
constexpr int MAX_MONO = 15;          // 1 + 4 + 4*5/2, the quadratic form's monomials at the channel cap
constexpr int CHUNK = 4096;           // the reference CPU backend's fixed reduction chunk (src/cpu/cpu_constants.h:103)
constexpr int SITE_BLOCKS = 512;      // the decoder accumulation's fixed block count (src/cpu/cpu_constants.h:41)
constexpr int LS_TILE = 256;          // the decoder accumulation's tile (src/cpu/cpu_solve_decoder.cpp:55)

// ---------------------------------------------------------------------------------------------------------------------
// Sampling (src/cpu/cpu_sample.h:27-141)
// ---------------------------------------------------------------------------------------------------------------------

struct Tap
{
    int x0, x1, y0, y1;
    float fx, fy;
};

NNTC_HD int clamp_index(int i, int n)
{
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// The hardware's bilinear addressing with clamp-to-edge, for a site given as a pixel of the w0 x h0 plane being
// fitted, into a w x h plane: exact in float when the two extents are equal.
// One axis of it: p + d itself on the pixel's own grid, else ((p + 0.5 + d) n) / n0 - 0.5 in fp32. When n and n0 are
// both powers of two, (p + 0.5 + d) n is exact and so is its quotient by n0, a power-of-two scaling of p + 0.5 + d: a
// scaling by 2^(log2 n - log2 n0) (an exponent adjustment) gives the division's bits without its instruction sequence.
NNTC_HD float tap_coord(int p, float d, int n0, int n)
{
    if (n == n0)
        return (float)p + d;
    const float a = (float)p + 0.5f + d;
    if ((n & (n - 1)) == 0 && (n0 & (n0 - 1)) == 0)
        return ldexpf(a, __builtin_ctz((unsigned)n) - __builtin_ctz((unsigned)n0)) - 0.5f;
    return (a * (float)n) / (float)n0 - 0.5f;
}

NNTC_HD Tap tap_pixel(int px, int py, float dx, float dy, int w0, int h0, int w, int h)
{
    const float x = tap_coord(px, dx, w0, w);
    const float y = tap_coord(py, dy, h0, h);
    const int x0 = (int)floorf(x);
    const int y0 = (int)floorf(y);
    Tap t;
    t.fx = x - (float)x0;
    t.fy = y - (float)y0;
    t.x0 = clamp_index(x0, w);
    t.x1 = clamp_index(x0 + 1, w);
    t.y0 = clamp_index(y0, h);
    t.y1 = clamp_index(y0 + 1, h);
    return t;
}

// The four clamped taps blended channel by channel, in fp32, in the order w00 a + w10 b + w01 e + w11 f.
NNTC_HD void sample_bilinear(const float* plane, int w, int c, const Tap& t, float* out)
{
    const float* a = plane + ((size_t)t.y0 * w + t.x0) * c;
    const float* b = plane + ((size_t)t.y0 * w + t.x1) * c;
    const float* e = plane + ((size_t)t.y1 * w + t.x0) * c;
    const float* f = plane + ((size_t)t.y1 * w + t.x1) * c;
    const float w00 = (1.0f - t.fx) * (1.0f - t.fy), w10 = t.fx * (1.0f - t.fy);
    const float w01 = (1.0f - t.fx) * t.fy, w11 = t.fx * t.fy;
    for (int k = 0; k < c; k++)
        out[k] = w00 * a[k] + w10 * b[k] + w01 * e[k] + w11 * f[k];
}

// phi(s, c): [c_j][s_i][s_i c_j], i outer and j inner in the product block.
NNTC_HD int build_phi(const float* s, int c0, const float* c, int c1, float* phi)
{
    int k = 0;
    for (int j = 0; j < c1; j++)
        phi[k++] = c[j];
    for (int i = 0; i < c0; i++)
        phi[k++] = s[i];
    for (int i = 0; i < c0; i++)
        for (int j = 0; j < c1; j++)
            phi[k++] = s[i] * c[j];
    return k;
}

// Fractional site j of pixel (tx, ty): the K = 4 rotated grid {(-3/8,-1/8),(-1/8,3/8),(1/8,-3/8),(3/8,1/8)} (or the
// K = 2 pair, or K = 1's (1/4, 1/4)), rotated a quarter turn per texel of the 2x2 cell, n = 2 (ty & 1) + (tx & 1).
NNTC_HD void subtexel_offset(int k_count, int j, int tx, int ty, float& dx, float& dy)
{
    if (k_count == 2)
    {
        dx = j == 0 ? -0.25f : 0.25f;
        dy = j == 0 ? 0.25f : -0.25f;
        return;
    }
    float ox = 0.25f, oy = 0.25f;
    if (k_count == 4)
    {
        ox = j == 0 ? -0.375f : (j == 1 ? -0.125f : (j == 2 ? 0.125f : 0.375f));
        oy = j == 0 ? -0.125f : (j == 1 ? 0.375f : (j == 2 ? -0.375f : 0.125f));
    }
    const int n = (ty & 1) * 2 + (tx & 1);
    for (int r = 0; r < n; r++)
    {
        const float t = ox;
        ox = -oy;
        oy = t;
    }
    dx = ox;
    dy = oy;
}

// Site j of pixel (px, py): j = 0 the centre, 1..K the fractional sites.
NNTC_HD void site_delta(int k_count, int j, int px, int py, float& dx, float& dy)
{
    dx = 0.0f;
    dy = 0.0f;
    if (j > 0)
        subtexel_offset(k_count, j - 1, px, py, dx, dy);
}

// The decoder's input phi and the target at one site, in fp32: level 1 sampled bilinearly; level 0 and the source
// read at the texel for the centre and sampled bilinearly at a fractional site (src/cpu/cpu_sample.h:115-139).
NNTC_HD void site_row(const float* v0, const float* v1, const float* src, int w0, int h0, int w1, int h1, int c0,
                      int c1, int nout, int k_count, int j, int px, int py, float* phi, float* target)
{
    float dx, dy;
    site_delta(k_count, j, px, py, dx, dy);
    float s[MAX_CH], c[MAX_CH];
    const Tap t1 = tap_pixel(px, py, dx, dy, w0, h0, w1, h1);
    sample_bilinear(v1, w1, c1, t1, c);
    if (j == 0)
    {
        const size_t idx = (size_t)py * w0 + px;
        for (int i = 0; i < c0; i++)
            s[i] = v0[idx * c0 + i];
        for (int i = 0; i < nout; i++)
            target[i] = src[idx * nout + i];
    }
    else
    {
        const Tap t0 = tap_pixel(px, py, dx, dy, w0, h0, w0, h0);
        sample_bilinear(v0, w0, c0, t0, s);
        sample_bilinear(src, w0, nout, t0, target);
    }
    build_phi(s, c0, c, c1, phi);
}

// The objective's own bilinear rule in double, from normalised coordinates (src/objective.cu:104-118).
NNTC_HD void obj_sample(const float* plane, int w, int h, int c, double u, double v, double* out)
{
    const double x = u * (double)w - 0.5, y = v * (double)h - 0.5;
    const double xf = floor(x), yf = floor(y);
    const double fx = x - xf, fy = y - yf;
    const int x0 = clamp_index((int)xf, w), x1 = clamp_index((int)xf + 1, w);
    const int y0 = clamp_index((int)yf, h), y1 = clamp_index((int)yf + 1, h);
    const float* a = plane + ((size_t)y0 * w + x0) * c;
    const float* b = plane + ((size_t)y0 * w + x1) * c;
    const float* e = plane + ((size_t)y1 * w + x0) * c;
    const float* f = plane + ((size_t)y1 * w + x1) * c;
    const double w00 = (1.0 - fx) * (1.0 - fy), w10 = fx * (1.0 - fy), w01 = (1.0 - fx) * fy, w11 = fx * fy;
    for (int k = 0; k < c; k++)
        out[k] = w00 * (double)a[k] + w10 * (double)b[k] + w01 * (double)e[k] + w11 * (double)f[k];
}

// The same two blends with the channel loop unrolled to a compile-time bound and guarded by the runtime count, so the
// output array is indexed by constants and stays in registers. The arithmetic per channel is unchanged.
template <int MAXC>
NNTC_HD void sample_bilinear_r(const float* plane, int w, int c, const Tap& t, float* out)
{
    const float* a = plane + ((size_t)t.y0 * w + t.x0) * c;
    const float* b = plane + ((size_t)t.y0 * w + t.x1) * c;
    const float* e = plane + ((size_t)t.y1 * w + t.x0) * c;
    const float* f = plane + ((size_t)t.y1 * w + t.x1) * c;
    const float w00 = (1.0f - t.fx) * (1.0f - t.fy), w10 = t.fx * (1.0f - t.fy);
    const float w01 = (1.0f - t.fx) * t.fy, w11 = t.fx * t.fy;
#pragma unroll
    for (int k = 0; k < MAXC; k++)
        if (k < c)
            out[k] = w00 * a[k] + w10 * b[k] + w01 * e[k] + w11 * f[k];
}

// Channels [0, c) of one texel of a plane of c-channel rows into out[], whole float4 (or float2) loads where the rows
// allow: a chain array starts 256-byte aligned and its planes start at texel offsets, so with c a multiple of four (of
// two) every row is 16-byte (8-byte) aligned. Only the load width changes, not a value.
template <int MAXC>
NNTC_HD void load_texel(const float* p, int c, float* out)
{
    if ((c & 3) == 0)
    {
#pragma unroll
        for (int k = 0; k < MAXC / 4; k++)
            if (4 * k < c)
            {
                const float4 v = reinterpret_cast<const float4*>(p)[k];
                out[4 * k] = v.x;
                out[4 * k + 1] = v.y;
                out[4 * k + 2] = v.z;
                out[4 * k + 3] = v.w;
            }
    }
    else if ((c & 1) == 0)
    {
#pragma unroll
        for (int k = 0; k < MAXC / 2; k++)
            if (2 * k < c)
            {
                const float2 v = reinterpret_cast<const float2*>(p)[k];
                out[2 * k] = v.x;
                out[2 * k + 1] = v.y;
            }
    }
    else
    {
#pragma unroll
        for (int k = 0; k < MAXC; k++)
            if (k < c)
                out[k] = p[k];
    }
}

// sample_bilinear_r with each tap's channels read by load_texel; the blend is the same expression, channel by channel
// (FUSED: with its last three products fused into the sum, for the fast form).
template <int MAXC, bool FUSED = false>
NNTC_HD void sample_bilinear_v(const float* plane, int w, int c, const Tap& t, float* out)
{
    float a[MAXC], b[MAXC], e[MAXC], f[MAXC];
    // 32-bit offsets: a plane holds fewer than 2^31 values (checked when the device model is created).
    const unsigned int r0 = (unsigned int)(t.y0 * w), r1 = (unsigned int)(t.y1 * w), uc = (unsigned int)c;
    load_texel<MAXC>(plane + (r0 + (unsigned int)t.x0) * uc, c, a);
    load_texel<MAXC>(plane + (r0 + (unsigned int)t.x1) * uc, c, b);
    load_texel<MAXC>(plane + (r1 + (unsigned int)t.x0) * uc, c, e);
    load_texel<MAXC>(plane + (r1 + (unsigned int)t.x1) * uc, c, f);
    const float w00 = (1.0f - t.fx) * (1.0f - t.fy), w10 = t.fx * (1.0f - t.fy);
    const float w01 = (1.0f - t.fx) * t.fy, w11 = t.fx * t.fy;
#pragma unroll
    for (int k = 0; k < MAXC; k++)
        if (k < c)
        {
            if (FUSED)
                out[k] = fmaf(w11, f[k], fmaf(w01, e[k], fmaf(w10, b[k], w00 * a[k])));
            else
                out[k] = w00 * a[k] + w10 * b[k] + w01 * e[k] + w11 * f[k];
        }
}

template <int MAXC>
NNTC_HD void obj_sample_r(const float* plane, int w, int h, int c, double u, double v, double* out)
{
    const double x = u * (double)w - 0.5, y = v * (double)h - 0.5;
    const double xf = floor(x), yf = floor(y);
    const double fx = x - xf, fy = y - yf;
    const int x0 = clamp_index((int)xf, w), x1 = clamp_index((int)xf + 1, w);
    const int y0 = clamp_index((int)yf, h), y1 = clamp_index((int)yf + 1, h);
    const float* a = plane + ((size_t)y0 * w + x0) * c;
    const float* b = plane + ((size_t)y0 * w + x1) * c;
    const float* e = plane + ((size_t)y1 * w + x0) * c;
    const float* f = plane + ((size_t)y1 * w + x1) * c;
    const double w00 = (1.0 - fx) * (1.0 - fy), w10 = fx * (1.0 - fy), w01 = (1.0 - fx) * fy, w11 = fx * fy;
#pragma unroll
    for (int k = 0; k < MAXC; k++)
        if (k < c)
            out[k] = w00 * (double)a[k] + w10 * (double)b[k] + w01 * (double)e[k] + w11 * (double)f[k];
}

// One output of the decoder, b_r + W_r phi, with phi = [c_j][s_i][s_i c_j] formed on the fly (in double when T is
// double, in float when float), the terms added in phi's order.
template <class T>
NNTC_HD T decode_row(const float* row, float bias, const T* s, int c0, const T* c, int c1)
{
    T acc = (T)bias;
#pragma unroll
    for (int q = 0; q < MAX_CH; q++)
        if (q < c1)
            acc += (T)row[q] * c[q];
#pragma unroll
    for (int i = 0; i < MAX_CH; i++)
        if (i < c0)
            acc += (T)row[c1 + i] * s[i];
#pragma unroll
    for (int i = 0; i < MAX_CH; i++)
#pragma unroll
        for (int q = 0; q < MAX_CH; q++)
            if (i < c0 && q < c1)
                acc += (T)row[c1 + c0 + i * c1 + q] * (s[i] * c[q]);
    return acc;
}

// ---------------------------------------------------------------------------------------------------------------------
// The grid (src/model.h:141-199)
// ---------------------------------------------------------------------------------------------------------------------

// The index in the byte's top bits replicated into the low bits.
NNTC_HD int replicate_index(int k, int bits)
{
    unsigned b8 = 0;
    for (int sh = 8 - bits; sh > -bits; sh -= bits)
        b8 |= sh >= 0 ? ((unsigned)k << sh) : ((unsigned)k >> -sh);
    return (int)(b8 & 0xFFu);
}

// value = lo + rep(k) / 255 (hi - lo), in double with the multiply and the add rounded separately, rounded to float
// once.
NNTC_HD float level1_value(float lo, float hi, int bits, int k)
{
    const double span = (double)hi - (double)lo;
    const double frac = (double)replicate_index(k, bits) / 255.0;
    return (float)((double)lo + frac * span);
}

// The index of the nearest grid value; a tie keeps the lower index.
NNTC_HD int level1_index(float lo, float hi, int bits, float value)
{
    const double span = (double)hi - (double)lo;
    const int levels = (1 << bits) - 1;
    if (!(span > 0.0))
        return 0;
    double x = ((double)value - (double)lo) / span * (double)levels;
    x = x < 0.0 ? 0.0 : (x > (double)levels ? (double)levels : x);
    int k = (int)(x + 0.5);
    k = k < 0 ? 0 : (k > levels ? levels : k);
    int best = k;
    double best_d = fabs((double)value - (double)level1_value(lo, hi, bits, k));
    for (int d = -1; d <= 1; d += 2)
    {
        const int n = k + d;
        if (n < 0 || n > levels)
            continue;
        const double dist = fabs((double)value - (double)level1_value(lo, hi, bits, n));
        if (dist < best_d || (dist == best_d && n < best))
        {
            best_d = dist;
            best = n;
        }
    }
    return best;
}

// ---------------------------------------------------------------------------------------------------------------------
// BC4 (shared/bc_pack.h:62-79, src/refine_bc.cu:102-131)
// ---------------------------------------------------------------------------------------------------------------------

// The eight-value palette (r0 > r1) as the weight on the high endpoint: p = a r0 + (1 - a) r1.
NNTC_HD double bc4_weight8(int sel)
{
    return sel == 0 ? 1.0 : (sel == 1 ? 0.0 : (double)(8 - sel) / 7.0);
}

// Palette entry `sel` of a block with endpoints r0, r1, in float, in either mode.
NNTC_HD float bc4_entry(int r0, int r1, int sel)
{
    if (sel == 0)
        return (float)r0;
    if (sel == 1)
        return (float)r1;
    if (r0 > r1)
        return (float)((8 - sel) * r0 + (sel - 1) * r1) / 7.0f;
    if (sel >= 6)
        return sel == 6 ? 0.0f : 255.0f;
    return (float)((6 - sel) * r0 + (sel - 1) * r1) / 5.0f;
}

// The eight-value-mode entry the refinement's own arithmetic uses (src/refine_bc.cu:111-118).
NNTC_HD float bc4_entry8(int r0, int r1, int sel)
{
    if (sel == 0)
        return (float)r0;
    if (sel == 1)
        return (float)r1;
    return (float)((8 - sel) * r0 + (sel - 1) * r1) / 7.0f;
}

// The plane value a decoded byte stands for: lo + byte / 255 (hi - lo), multiply and add rounded apart.
NNTC_HD float bc0_value(float lo, float hi, float byte)
{
    const double frac = (double)byte / 255.0;
    const double span = (double)hi - (double)lo;
    return (float)((double)lo + frac * span);
}

// The byte's nearest 8-bit index, with the float round trip (int)(byte + 0.5f).
NNTC_HD uint8_t bc0_index(float byte)
{
    return (uint8_t)(byte < 0.0f ? 0 : (byte > 255.0f ? 255 : (int)(byte + 0.5f)));
}

// ---------------------------------------------------------------------------------------------------------------------
// Stencil geometry (src/cpu/cpu_constants.h:47-68)
// ---------------------------------------------------------------------------------------------------------------------

// The five blocks a texel owns: itself, right, down-left, down, down-right.
NNTC_HD void stencil_offset(int slot, int& dx, int& dy)
{
    dx = slot == 1 ? 1 : (slot == 2 ? -1 : (slot == 4 ? 1 : 0));
    dy = slot >= 2 ? 1 : 0;
}

NNTC_HD int stencil_slot(int dx, int dy)
{
    for (int s = 0; s < STENCIL_BLOCKS; s++)
    {
        int ox, oy;
        stencil_offset(s, ox, oy);
        if (ox == dx && oy == dy)
            return s;
    }
    return -1;
}

NNTC_HD int stencil_monomials(int c)
{
    return 1 + c + c * (c + 1) / 2;
}

// ---------------------------------------------------------------------------------------------------------------------
// The plane table every batched launch carries: one launch covers every stored plane of the chain.
// ---------------------------------------------------------------------------------------------------------------------

struct PlaneGeom
{
    int w0, h0, w1, h1;
    long long n0, n1;       // texels of level 0 and of level 1
    long long off0, off1;   // this plane's first texel in the chain-concatenated level-0 / level-1 arrays
    long long nb, offb;     // BC blocks ((w0 + 3) / 4) ((h0 + 3) / 4) and the first one
};

// Exact division of 0 <= n < 2^31 by a divisor 1 <= d < 2^31 chosen on the host: q = (n m) >> s with s = 31 + ceil(log2
// d) and m = ceil(2^s / d). With e = m d - 2^s < d <= 2^(s - 31), n m / 2^s exceeds n / d by n e / (d 2^s) < 1 / d,
// so the floor is exact; n m < 2^31 2^33 fits in 64 bits. A multiply and a shift in place of a 64-bit division.
struct Div31
{
    unsigned long long m = 1ull << 31;
    int s = 31;
    NNTC_HD Div31() {}
    NNTC_HD explicit Div31(unsigned int d)
    {
        int l = 0;
        while ((1ull << l) < (unsigned long long)d)
            l++;
        s = 31 + l;
        m = ((1ull << s) + d - 1) / d;
    }
    NNTC_HD unsigned int div(unsigned int n) const { return (unsigned int)(((unsigned long long)n * m) >> s); }
};

// Division by a workgroup-uniform divisor d of numerators below 2^16: q = floor(n m / 2^32) with m = ceil(2^32 / d)
// is exact, since the error n (m d - 2^32) / (d 2^32) stays below 1 / d for n, d < 2^16. A multiply-high in place of the
// compiler's general 32-bit division sequence in the staging loops' index arithmetic.
struct Div16
{
    unsigned int d, m;
    __device__ explicit Div16(unsigned int dd) : d(dd), m(dd > 1 ? (unsigned int)((0x100000000ull + dd - 1) / dd) : 0u) {}
    __device__ unsigned int operator()(unsigned int n) const { return d > 1 ? __umulhi(n, m) : n; }
};

struct Planes
{
    int n;
    PlaneGeom p[MAX_PLANES];
};

// A 1-D index space that is the concatenation of one range per plane: `count[p]` items starting at `start[p]`.
struct Ranges
{
    int n;
    long long start[MAX_PLANES + 1];
};

// acc plus t[0] + t[1] + ... + t[n-1], added in index order (the exact form's serial folds); sixteen values are read
// ahead of their sixteen dependent additions so the chain waits on the adds alone.
NNTC_HD double fold_serial16(const double* t, int n, double acc)
{
    int i = 0;
    for (; i + 16 <= n; i += 16)
    {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++)
            v[u] = t[i + u];
#pragma unroll
        for (int u = 0; u < 16; u++)
            acc += v[u];
    }
    for (; i < n; i++)
        acc += t[i];
    return acc;
}

// The workgroup's place in an index range split into eight contiguous parts, one per XCD. The dispatcher deals
// workgroups to the chip's eight XCDs round-robin (workgroup b to XCD b % 8), and each XCD has its own L2: with the
// identity mapping, neighbouring workgroups - which read the same cache lines of their neighbours' data - land on
// eight different L2s, and each fetches the lines from HBM. XCD x takes the x-th part (its workgroups b = x + 8 i, in
// order of i). A bijection of [0, n) for every n.
__device__ __forceinline__ unsigned int xcd_block(unsigned int b, unsigned int n)
{
    const unsigned int q = n / 8, r = n % 8, x = b % 8, i = b / 8;
    return x * q + (x < r ? x : r) + i;
}

// The plane whose range holds global index g (ranges are contiguous and increasing).
NNTC_HD int range_plane(const Ranges& r, long long g)
{
    int p = 0;
    while (p + 1 < r.n && g >= r.start[p + 1])
        p++;
    return p;
}

// ---------------------------------------------------------------------------------------------------------------------
// The stencil's layout in device memory. Entry e (slot * c * c + i * c + j of a texel's five c x c blocks) of every
// texel of the chain is one run of `sn` doubles, so a wavefront reading one entry of 64 texels reads one contiguous
// span. Within a run the texels are grouped by plane, then by colour (x & 1, y & 1) in the order the quantised sweeps
// visit them, then row-major inside the colour: a sweep over one colour reads each entry as a contiguous run, and
// the neighbours a texel reads (all of other colours) sit at consecutive positions for consecutive texels.
// ---------------------------------------------------------------------------------------------------------------------

struct StencilMap
{
    long long sn;                       // texels of the chain: the stride between entries
    long long cpos[MAX_PLANES][4];      // the first position of colour (cy * 2 + cx) of each plane
};

// Texels of colour (cx, cy) in a w x h plane.
NNTC_HD long long colour_count(int w, int h, int colour)
{
    const int cx = colour & 1, cy = colour >> 1;
    return (long long)((w - cx + 1) / 2) * ((h - cy + 1) / 2);
}

// The position of texel (tx, ty) of plane p (width w) within each entry's run.
NNTC_HD long long stencil_pos(const StencilMap& m, int p, int w, int tx, int ty)
{
    const int cx = tx & 1, cy = ty & 1;
    const int cw = (w - cx + 1) >> 1;
    return m.cpos[p][cy * 2 + cx] + (long long)(ty >> 1) * cw + (tx >> 1);
}

}   // namespace nntc
