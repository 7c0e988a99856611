// tile_stage.h: the fast form's LDS-staged pixel tiles, shared by the tiled objective and the tiled decoder Gram.
//
// A 256-lane workgroup owns a TILE_W x TILE_H tile of one plane's pixels, one pixel (its 1 + K sites) per lane. It
// first copies, with coalesced row reads, the tile and its one-pixel halo of the source and level 0, and the level-1
// texels the tile's sites can tap, into LDS; every tap of every site is then an LDS read, and each input byte leaves
// HBM once per tile (plus the halo) instead of once per tap. Sampling follows the fp32 hardware bilinear rule of the
// assembly (tap_pixel, src/sample.cuh:43-72), so the tiled kernels see the model the latent solves see.

#pragma once

#include "arm.h"

namespace nntc
{

constexpr int TILE_W = OBJ_TW, TILE_H = OBJ_TH;
constexpr int TILE_L0_PIX = (TILE_W + 2) * (TILE_H + 2);
constexpr int TILE_L1_W = TILE_W + 6, TILE_L1_H = TILE_H + 6;   // bounds the level-1 box when w1 <= w0, h1 <= h0
// Floats per staged level-0 pixel: the nout source values, then C0, padded to an odd count so that lanes reading
// neighbouring pixels land in distinct LDS banks.
__host__ __device__ constexpr int tile_l0_stride(int nout, int c0)
{
    return (nout + c0) | 1;
}
template <int C0>
constexpr int tile_l0_floats()
{
    return TILE_L0_PIX * tile_l0_stride(MAX_NOUT, C0);
}
constexpr int TILE_L1_FLOATS = TILE_L1_W * TILE_L1_H * MAX_CH;

// The level-1 texels the taps of sites at pixel columns [p0, p0 + n) can reach (sub-texel offsets stay within
// +-3/8), with a margin of one texel each side for rounding.
__device__ __forceinline__ void level1_box(int p0, int n, int w0, int w1, int& lo, int& hi)
{
    const float xmin = (((float)p0 + 0.125f) * (float)w1) / (float)w0 - 0.5f;
    const float xmax = (((float)(p0 + n) - 0.125f) * (float)w1) / (float)w0 - 0.5f;
    lo = max(0, (int)floorf(xmin) - 1);
    hi = min(w1 - 1, (int)floorf(xmax) + 2);
}

struct TileGeom
{
    int w0, h0, w1, h1, nout;
    int tx0, ty0;            // the tile's first pixel
    int bx0, by0, bw, bh;    // the staged level-0 box (tile plus halo, clipped to the plane)
    int lx0, ly0, lw, lh;    // the staged level-1 box
    int P;                   // floats per staged level-0 pixel: nout source values, then C0 (tile_l0_stride)
};

// Stage tile t (row-major over the plane's tiles) of plane g. The caller synchronises before reading.
template <int C0, int C1>
__device__ __forceinline__ TileGeom stage_tile(const PlaneGeom& g, int t, int nout, const float* src_chain,
                                               const float* v0_chain, const float* v1_chain, float* s_l0,
                                               float* s_l1)
{
    TileGeom tg;
    tg.w0 = g.w0;
    tg.h0 = g.h0;
    tg.w1 = g.w1;
    tg.h1 = g.h1;
    tg.nout = nout;
    const int tiles_x = (g.w0 + TILE_W - 1) / TILE_W;
    tg.tx0 = (t % tiles_x) * TILE_W;
    tg.ty0 = (t / tiles_x) * TILE_H;
    tg.bx0 = max(0, tg.tx0 - 1);
    tg.by0 = max(0, tg.ty0 - 1);
    tg.bw = min(g.w0 - 1, tg.tx0 + TILE_W) - tg.bx0 + 1;
    tg.bh = min(g.h0 - 1, tg.ty0 + TILE_H) - tg.by0 + 1;
    tg.P = tile_l0_stride(nout, C0);
    const int tid = threadIdx.x, nthreads = blockDim.x;
    const float* src = src_chain + g.off0 * nout;
    const float* v0 = v0_chain + g.off0 * C0;
    // (The staged counts stay below 2^16: (TILE_W + 2) (TILE_H + 2) pixels of at most MAX_NOUT + MAX_CH floats.)
    const Div16 by_row(tg.bw * nout), by_nout(nout), by_row0(tg.bw * C0);
    // Eight loads in flight per lane before their LDS stores.
    const int n_src = tg.bw * tg.bh * nout;
    for (int i0 = tid; i0 < n_src; i0 += 8 * nthreads)
    {
        float v[8];
        int at[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
        {
            const int i = i0 + u * nthreads;
            at[u] = -1;
            if (i < n_src)
            {
                const int row = (int)by_row((unsigned)i), rem = i - row * tg.bw * nout;
                const int pix = (int)by_nout((unsigned)rem), r = rem - pix * nout;
                v[u] = src[((size_t)(tg.by0 + row) * g.w0 + tg.bx0) * nout + rem];
                at[u] = (row * tg.bw + pix) * tg.P + r;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
            if (at[u] >= 0)
                s_l0[at[u]] = v[u];
    }
    for (int i = tid; i < tg.bw * tg.bh * C0; i += nthreads)
    {
        const int row = (int)by_row0((unsigned)i), rem = i - row * tg.bw * C0;
        const int pix = rem / C0, q = rem - pix * C0;
        s_l0[(row * tg.bw + pix) * tg.P + nout + q] = v0[((size_t)(tg.by0 + row) * g.w0 + tg.bx0) * C0 + rem];
    }
    int lx1, ly1;
    level1_box(tg.tx0, TILE_W, g.w0, g.w1, tg.lx0, lx1);
    level1_box(tg.ty0, TILE_H, g.h0, g.h1, tg.ly0, ly1);
    tg.lw = lx1 - tg.lx0 + 1;
    tg.lh = ly1 - tg.ly0 + 1;
    const float* v1 = v1_chain + g.off1 * C1;
    const Div16 by_row1(tg.lw * C1);
    for (int i = tid; i < tg.lw * tg.lh * C1; i += nthreads)
    {
        const int row = (int)by_row1((unsigned)i), rem = i - row * tg.lw * C1;
        s_l1[row * tg.lw * C1 + rem] = v1[((size_t)(tg.ly0 + row) * g.w1 + tg.lx0) * C1 + rem];
    }
    return tg;
}

// Site j of pixel (px, py) of the staged tile: phi = [c (C1)][s (C0)][s_i c_q] and the target (nout values), fp32.
// The blends are w00 a + w10 b + w01 e + w11 f with the last three products fused into the sum (these tiles serve the
// fast form only), and the LDS offsets use 24-bit multiplies (every offset is far below 2^24).
template <int C0, int C1>
__device__ __forceinline__ void tile_site(const TileGeom& tg, const float* s_l0, const float* s_l1, int k, int j,
                                          int px, int py, float* phi, float* target)
{
    float dx, dy;
    site_delta(k, j, px, py, dx, dy);
    // Level 1 from its staged box (a tap outside the box cannot occur; the clamp keeps LDS reads in bounds).
    const Tap t1 = tap_pixel(px, py, dx, dy, tg.w0, tg.h0, tg.w1, tg.h1);
    {
        const int lw = tg.lw, lh = tg.lh;
        const int xa = min(max(t1.x0 - tg.lx0, 0), lw - 1), xb = min(max(t1.x1 - tg.lx0, 0), lw - 1);
        const int ya = min(max(t1.y0 - tg.ly0, 0), lh - 1), yb = min(max(t1.y1 - tg.ly0, 0), lh - 1);
        const float w00 = (1.0f - t1.fx) * (1.0f - t1.fy), w10 = t1.fx * (1.0f - t1.fy);
        const float w01 = (1.0f - t1.fx) * t1.fy, w11 = t1.fx * t1.fy;
        const int ra = __mul24(ya, lw), rb = __mul24(yb, lw);
        const float* la = s_l1 + (ra + xa) * C1;
        const float* lb = s_l1 + (ra + xb) * C1;
        const float* le = s_l1 + (rb + xa) * C1;
        const float* lf = s_l1 + (rb + xb) * C1;
#pragma unroll
        for (int q = 0; q < C1; q++)
            phi[q] = fmaf(w11, lf[q], fmaf(w01, le[q], fmaf(w10, lb[q], w00 * la[q])));
    }
    const int nout = tg.nout, P = tg.P;
    if (j == 0)
    {
        const float* pp = s_l0 + __mul24(__mul24(py - tg.by0, tg.bw) + (px - tg.bx0), P);
#pragma unroll
        for (int i = 0; i < C0; i++)
            phi[C1 + i] = pp[nout + i];
#pragma unroll
        for (int r = 0; r < MAX_NOUT; r++)
            if (r < nout)
                target[r] = pp[r];
    }
    else
    {
        const Tap t0 = tap_pixel(px, py, dx, dy, tg.w0, tg.h0, tg.w0, tg.h0);
        const int ra = __mul24(t0.y0 - tg.by0, tg.bw), rb = __mul24(t0.y1 - tg.by0, tg.bw);
        const float* pa = s_l0 + __mul24(ra + (t0.x0 - tg.bx0), P);
        const float* pb = s_l0 + __mul24(ra + (t0.x1 - tg.bx0), P);
        const float* pe = s_l0 + __mul24(rb + (t0.x0 - tg.bx0), P);
        const float* pf = s_l0 + __mul24(rb + (t0.x1 - tg.bx0), P);
        const float w00 = (1.0f - t0.fx) * (1.0f - t0.fy), w10 = t0.fx * (1.0f - t0.fy);
        const float w01 = (1.0f - t0.fx) * t0.fy, w11 = t0.fx * t0.fy;
#pragma unroll
        for (int i = 0; i < C0; i++)
            phi[C1 + i] = fmaf(w11, pf[nout + i], fmaf(w01, pe[nout + i], fmaf(w10, pb[nout + i], w00 * pa[nout + i])));
#pragma unroll
        for (int r = 0; r < MAX_NOUT; r++)
            if (r < nout)
                target[r] = fmaf(w11, pf[r], fmaf(w01, pe[r], fmaf(w10, pb[r], w00 * pa[r])));
    }
#pragma unroll
    for (int i = 0; i < C0; i++)
#pragma unroll
        for (int q = 0; q < C1; q++)
            phi[C1 + C0 + i * C1 + q] = phi[C1 + i] * phi[q];
}

}   // namespace nntc
