// adjoint.h: the site geometry and the decoder's linearisations shared by the assembly, the decode adjoint and the
// matrix-free level-0 operator.
//
// The decode at a site is out_r = b_r + W_r . phi(s0, s1), phi = [s1_q][s0_i][s0_i s1_q] (src/sample.cuh:74-85), with
// s0 and s1 the bilinear samples of the two latents at the site (src/sample.cuh:43-72, 87-178). It is linear in each
// latent's sample when the other is held:
//   d out_r / d s0_i = W[r][C1 + i] + sum_q s1_q W[r][C1 + C0 + i C1 + q]     (level0_row)
//   d out_r / d s1_q = W[r][q]      + sum_i s0_i W[r][C1 + C0 + i C1 + q]     (level1_row)
// and each sample is a fixed linear map of its latent's texels (the four clamped taps), so the adjoint of the decode is
// these rows transposed, then the taps' weights scattered back - here gathered, texel by texel, over the sites whose taps
// reach it (tap_weight), in a fixed order.

#pragma once

#include "nntc_common.h"

namespace nntc
{

// The pixels [lo, hi] of an n0-wide plane whose sites can tap texel ti of an n1-wide one (src/solve_level1.cu:250-257).
NNTC_HD void footprint_range(int ti, int n0, int n1, int& lo, int& hi)
{
    const double r = (double)n0 / (double)n1;
    lo = (int)ceil(((double)ti - 0.5) * r - 1.0);
    hi = (int)ceil(((double)ti + 1.5) * r) - 1;
    lo = lo < 0 ? 0 : lo;
    hi = hi > n0 - 1 ? n0 - 1 : hi;
}

// The weight a site's taps (x0, x1, f) give texel x along one axis: both taps of a clamped pair add up at one texel.
__device__ __forceinline__ float tap_weight(int x, int x0, int x1, float f)
{
    return (x0 == x ? 1.0f - f : 0.0f) + (x1 == x ? f : 0.0f);
}

// The taps of site j of pixel (px, py) into a w x h plane; level 0's centre site reads its own texel with weight 1.
__device__ __forceinline__ Tap site_taps(int k, int j, int px, int py, int w0, int h0, int w, int h, bool nearest_centre)
{
    if (nearest_centre && j == 0)
    {
        Tap t;
        t.x0 = t.x1 = px;
        t.y0 = t.y1 = py;
        t.fx = t.fy = 0.0f;
        return t;
    }
    float dx, dy;
    site_delta(k, j, px, py, dx, dy);
    return tap_pixel(px, py, dx, dy, w0, h0, w, h);
}

// Row r of d out / d s0 at a site whose level-1 sample is cbar, as the fast site terms form it (fused multiply-adds
// in q order).
template <int C0, int C1>
__device__ __forceinline__ void level0_row(const float* w, int nin, int r, const float* cbar, float* m)
{
    const float* row = w + r * nin;
#pragma unroll
    for (int i = 0; i < C0; i++)
    {
        float qq = row[C1 + i];
#pragma unroll
        for (int q = 0; q < C1; q++)
            qq = fmaf(cbar[q], row[C1 + C0 + i * C1 + q], qq);
        m[i] = qq;
    }
}

// The same rows in any precision T (the decode adjoint's fp64 form for gradient checks).
template <class T, int C0, int C1>
__device__ __forceinline__ void level0_row_t(const T* w, int nin, int r, const T* cbar, T* m)
{
    const T* row = w + r * nin;
#pragma unroll
    for (int i = 0; i < C0; i++)
    {
        T qq = row[C1 + i];
#pragma unroll
        for (int q = 0; q < C1; q++)
            qq += cbar[q] * row[C1 + C0 + i * C1 + q];
        m[i] = qq;
    }
}

template <class T, int C0, int C1>
__device__ __forceinline__ void level1_row_t(const T* w, int nin, int r, const T* sbar, T* g)
{
    const T* row = w + r * nin;
#pragma unroll
    for (int q = 0; q < C1; q++)
    {
        T gg = row[q];
#pragma unroll
        for (int i = 0; i < C0; i++)
            gg += sbar[i] * row[C1 + C0 + i * C1 + q];
        g[q] = gg;
    }
}

// The bilinear blend of C channels of a plane at a site's taps, in precision T, weights formed in fp32 as the model's
// addressing forms them (w00 a + w10 b + w01 e + w11 f).
template <class T, int C>
__device__ __forceinline__ void blend_t(const T* plane, int w, const Tap& t, T* out)
{
    const T w00 = (T)((1.0f - t.fx) * (1.0f - t.fy)), w10 = (T)(t.fx * (1.0f - t.fy));
    const T w01 = (T)((1.0f - t.fx) * t.fy), w11 = (T)(t.fx * t.fy);
    const T* a = plane + ((size_t)t.y0 * w + t.x0) * C;
    const T* b = plane + ((size_t)t.y0 * w + t.x1) * C;
    const T* e = plane + ((size_t)t.y1 * w + t.x0) * C;
    const T* f = plane + ((size_t)t.y1 * w + t.x1) * C;
#pragma unroll
    for (int i = 0; i < C; i++)
        out[i] = w00 * a[i] + w10 * b[i] + w01 * e[i] + w11 * f[i];
}

}   // namespace nntc
