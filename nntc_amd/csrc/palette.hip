// palette.hip: --l0 palette, level 0 as a 1-4 bit index per channel on a fixed palette (src/solve_level0.cu,
// src/init.cu:389-428).
//
// Block (c) in the palette mode holds level 1 and the decoder and moves each level-0 texel to the palette state that
// minimises E over the sites its value reaches. With the other texels held, those sites' outputs are affine in the
// texel's own C0 values x (out = p + alpha Q x at a site where the texel carries weight alpha), so E over them is the
// quadratic A0 + 2 x.A1 + x.A2.x, accumulated in fp64 from fp32 site terms, and the argmin over the palette is exact:
// a joint enumeration of every state when the channels' bits add up to at most 8, coordinate sweeps over each
// channel's alphabet otherwise, a tie keeping the current state. Texels of one colour (x & 1, y & 1) share no site, so
// a colour pass moves all of them at once; the four passes run in order.
//
// The gfx950 layout: one lane per texel of a colour, every plane of the chain in one launch per colour (the plane of
// a lane from the colour's Ranges, as the quantised sweeps find theirs), the decoder and palette read at uniform
// addresses. The per-site arithmetic is the reference's operation for operation (src/cpu/cpu_solve_level0.cpp, which
// the CUDA kernel matches): the level-1 sample, the source sample, the tap weights; the site's alpha summed over the
// taps that fall on the texel before any other tap's level-0 value is read (so a lane never reads a texel of its own
// colour, which another lane of the pass may be writing); p and Q's columns in fp32 in the reference's order; A0, A1
// and A2 in fp64. Every form runs this one kernel: the moves are integer decisions and must be the reference's.

#include "arm.h"

namespace nntc
{

namespace
{

constexpr int SEARCH_THREADS = 128;

struct Bits4
{
    int b[MAX_CH];
};

struct SearchArgs
{
    Planes planes;
    Ranges colour;          // per plane: the texels of this colour
    int colour_index;
    const float* v1;        // chain-concatenated planes
    const float* src;
    const float* weights;
    const float* bias;
    const float* cw;
    const float* palette;   // 4 x 16, channel-major
    float* v0;
    uint8_t* k0;
    unsigned int* moved;    // per plane, four colour counters
    int c0, c1, nin, nout, k, joint, sweeps;
    int bits[MAX_CH];
};

__device__ __forceinline__ double search_energy(double a0, const double* a1, const double* a2, const float* x, int c0)
{
    double e = a0;
    for (int i = 0; i < c0; i++)
    {
        e += 2.0 * (double)x[i] * a1[i];
        for (int j = 0; j < c0; j++)
            e += (double)x[i] * (double)x[j] * a2[i * c0 + j];
    }
    return e;
}

__global__ __launch_bounds__(SEARCH_THREADS) void k_level0_search(SearchArgs a)
{
    const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= a.colour.start[a.colour.n])
        return;
    const int p = range_plane(a.colour, gi);
    const PlaneGeom& g = a.planes.p[p];
    const int w0 = g.w0, h0 = g.h0, w1 = g.w1, h1 = g.h1;
    const int cx = a.colour_index & 1, cy = a.colour_index >> 1;
    const long long ci = gi - a.colour.start[p];
    const int cw0 = (w0 - cx + 1) / 2;
    const int tx = cx + 2 * (int)(ci % cw0), ty = cy + 2 * (int)(ci / cw0);
    const int c0 = a.c0, c1 = a.c1, nin = a.nin, nout = a.nout;
    const float* v0 = a.v0 + g.off0 * c0;
    const float* v1 = a.v1 + g.off1 * c1;
    const float* src = a.src + g.off0 * nout;
    const long long t = (long long)ty * w0 + tx;

    double a0 = 0.0, a1[MAX_CH], a2[MAX_CH * MAX_CH];
    for (int i = 0; i < c0; i++)
        a1[i] = 0.0;
    for (int i = 0; i < c0 * c0; i++)
        a2[i] = 0.0;
    for (int oy = -1; oy <= 1; oy++)
        for (int ox = -1; ox <= 1; ox++)
        {
            const int px = tx + ox, py = ty + oy;
            if (px < 0 || py < 0 || px >= w0 || py >= h0)
                continue;
            const int first = (ox == 0 && oy == 0) ? 0 : 1;   // only the texel's own pixel has its centre site
            for (int j = first; j <= a.k; j++)
            {
                float sdx, sdy;
                site_delta(a.k, j, px, py, sdx, sdy);
                float alpha = 0.0f, beta[MAX_CH], cbar[MAX_CH], target[MAX_NOUT];
                sample_bilinear(v1, w1, c1, tap_pixel(px, py, sdx, sdy, w0, h0, w1, h1), cbar);
                for (int i = 0; i < c0; i++)
                    beta[i] = 0.0f;
                if (j == 0)
                {
                    alpha = 1.0f;
                    for (int r = 0; r < nout; r++)
                        target[r] = src[t * nout + r];
                }
                else
                {
                    const Tap t0 = tap_pixel(px, py, sdx, sdy, w0, h0, w0, h0);
                    sample_bilinear(src, w0, nout, t0, target);
                    const int qx[4] = { t0.x0, t0.x1, t0.x0, t0.x1 };
                    const int qy[4] = { t0.y0, t0.y0, t0.y1, t0.y1 };
                    const float qw[4] = { (1.0f - t0.fx) * (1.0f - t0.fy), t0.fx * (1.0f - t0.fy),
                                          (1.0f - t0.fx) * t0.fy, t0.fx * t0.fy };
                    for (int q = 0; q < 4; q++)
                        if (qx[q] == tx && qy[q] == ty)
                            alpha += qw[q];
                    if (!(alpha > 0.0f))
                        continue;   // the site does not see the texel: the same constant in every state
                    for (int q = 0; q < 4; q++)
                        if (!(qx[q] == tx && qy[q] == ty))
                        {
                            const float* nv = v0 + ((long long)qy[q] * w0 + qx[q]) * c0;
                            for (int i = 0; i < c0; i++)
                                beta[i] += qw[q] * nv[i];
                        }
                }
                // p (the decode with the texel's channels at zero) and Q's columns at the site's level-1 sample.
                float qcol[MAX_NOUT * MAX_CH], resid[MAX_NOUT];
                for (int r = 0; r < nout; r++)
                {
                    const float* row = a.weights + (size_t)r * nin;
                    float out = a.bias[r];
                    for (int q = 0; q < c1; q++)
                        out += row[q] * cbar[q];
                    for (int i = 0; i < c0; i++)
                    {
                        float slope = row[c1 + i];
                        for (int q = 0; q < c1; q++)
                            slope += row[c1 + c0 + i * c1 + q] * cbar[q];
                        qcol[r * c0 + i] = slope;
                        out += slope * beta[i];
                    }
                    resid[r] = out - target[r];
                }
                for (int r = 0; r < nout; r++)
                    a0 += (double)a.cw[r] * (double)resid[r] * (double)resid[r];
                for (int i = 0; i < c0; i++)
                {
                    double gsum = 0.0;
                    for (int r = 0; r < nout; r++)
                        gsum += (double)a.cw[r] * (double)qcol[r * c0 + i] * (double)resid[r];
                    a1[i] += (double)alpha * gsum;
                    for (int i2 = 0; i2 < c0; i2++)
                    {
                        double s = 0.0;
                        for (int r = 0; r < nout; r++)
                            s += (double)a.cw[r] * (double)qcol[r * c0 + i] * (double)qcol[r * c0 + i2];
                        a2[i * c0 + i2] += (double)alpha * (double)alpha * s;
                    }
                }
            }
        }

    // The argmin over the palette from the current state (a tie keeps it).
    int cur[MAX_CH], best[MAX_CH], levels[MAX_CH];
    float x[MAX_CH];
    uint8_t* k0 = a.k0 + g.off0 * c0;
    for (int i = 0; i < c0; i++)
    {
        levels[i] = (1 << a.bits[i]) - 1;
        cur[i] = (int)k0[t * c0 + i];
        best[i] = cur[i];
        x[i] = a.palette[i * 16 + cur[i]];
    }
    double best_e = search_energy(a0, a1, a2, x, c0);
    if (a.joint)
    {
        int total = 1;
        for (int i = 0; i < c0; i++)
            total *= levels[i] + 1;
        for (int n = 0; n < total; n++)
        {
            int rest = n;
            bool same = true;
            int state[MAX_CH];
            for (int i = 0; i < c0; i++)
            {
                state[i] = rest % (levels[i] + 1);
                rest /= levels[i] + 1;
                same = same && state[i] == cur[i];
                x[i] = a.palette[i * 16 + state[i]];
            }
            if (same)
                continue;
            const double e = search_energy(a0, a1, a2, x, c0);
            if (e < best_e)
            {
                best_e = e;
                for (int i = 0; i < c0; i++)
                    best[i] = state[i];
            }
        }
    }
    else
    {
        for (int sw = 0; sw < a.sweeps; sw++)
        {
            bool changed = false;
            for (int i = 0; i < c0; i++)
            {
                for (int q = 0; q < c0; q++)
                    x[q] = a.palette[q * 16 + best[q]];
                const int hold = best[i];
                for (int kk = 0; kk <= levels[i]; kk++)
                {
                    if (kk == best[i])
                        continue;
                    x[i] = a.palette[i * 16 + kk];
                    const double e = search_energy(a0, a1, a2, x, c0);
                    if (e < best_e)
                    {
                        best_e = e;
                        best[i] = kk;
                    }
                }
                changed = changed || best[i] != hold;
            }
            if (!changed)
                break;
        }
    }
    bool moved = false;
    for (int i = 0; i < c0; i++)
        moved = moved || best[i] != cur[i];
    if (moved)
    {
        float* v0w = a.v0 + g.off0 * c0;
        for (int i = 0; i < c0; i++)
        {
            v0w[t * c0 + i] = a.palette[i * 16 + best[i]];
            k0[t * c0 + i] = (uint8_t)best[i];
        }
        atomicAdd(a.moved + p * 4 + a.colour_index, 1u);
    }
}

// The palette mode's init: channel 0 at 2 luma - 1 under --init0 luma and every other channel at 0, each snapped to
// the nearest palette VALUE (a tie keeps the lower index), src/init.cu:389-428.
__global__ __launch_bounds__(256) void k_init_level0_palette(const float* src, float* v0, uint8_t* k0, long long n,
                                                             int nout, int c0, float lr, float lg, float lb,
                                                             int luma0, const float* palette, Bits4 bits)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float* s = src + i * nout;
    const float luma = lr * s[0] + lg * s[1] + lb * s[2];
    for (int c = 0; c < c0; c++)
    {
        const float value = (c == 0 && luma0) ? 2.0f * luma - 1.0f : 0.0f;
        const int levels = (1 << bits.b[c]) - 1;
        int k = 0;
        float best = fabsf(palette[c * 16] - value);
        for (int j = 1; j <= levels; j++)
        {
            const float dist = fabsf(palette[c * 16 + j] - value);
            if (dist < best)
            {
                best = dist;
                k = j;
            }
        }
        v0[i * c0 + c] = palette[c * 16 + k];
        k0[i * c0 + c] = (uint8_t)k;
    }
}

__global__ void k_zero_u32(unsigned int* p, int n)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        p[i] = 0u;
}

}   // namespace

void init_level0_palette(Ctx& c, const float* luma, int luma_channel0)
{
    Bits4 bits {};
    for (int q = 0; q < c.c0; q++)
        bits.b[q] = c.md.bits0[q];
    hipLaunchKernelGGL(k_init_level0_palette, dim3(blocks_for(c.n0_total, 256)), dim3(256), 0, c.stream,
                       (const float*)c.src.p, c.v0.p, c.k0.p, c.n0_total, c.nout, c.c0, luma[0], luma[1], luma[2],
                       luma_channel0, (const float*)(c.grid.p + 16), bits);
    HIP_OK(hipGetLastError());
}

void solve_level0_all(Ctx& c, int k, int sweeps, nntc_level0_report* rep)
{
    SearchArgs a {};
    a.planes = c.planes;
    a.v1 = c.v1.p;
    a.src = c.src.p;
    a.weights = c.weights.p;
    a.bias = c.bias.p;
    a.cw = c.cw.p;
    a.palette = c.grid.p + 16;
    a.v0 = c.v0.p;
    a.k0 = c.k0.p;
    a.c0 = c.c0;
    a.c1 = c.c1;
    a.nin = c.nin;
    a.nout = c.nout;
    a.k = k;
    a.sweeps = sweeps;
    long long states = 1;
    int total_bits = 0;
    for (int q = 0; q < c.c0; q++)
    {
        a.bits[q] = c.md.bits0[q];
        states *= 1ll << c.md.bits0[q];
        total_bits += c.md.bits0[q];
    }
    a.joint = total_bits <= 8 ? 1 : 0;
    // Four counters a plane, in the head of the sweeps' moved-index counters (the two are never live together).
    a.moved = c.counters.p;
    const int ncount = c.planes.n * 4;
    if ((size_t)ncount > c.counters.n)
        throw Refused { "solve_level0_all: the counters hold too few planes" };
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c.stream, a.moved, ncount);
    HIP_OK(hipEventRecord(c.ev0, c.stream));
    for (int col = 0; col < 4; col++)
    {
        const int cx = col & 1, cy = col >> 1;
        a.colour.n = c.planes.n;
        long long acc = 0;
        for (int p = 0; p < c.planes.n; p++)
        {
            a.colour.start[p] = acc;
            acc += (long long)((c.g(p).w0 - cx + 1) / 2) * ((c.g(p).h0 - cy + 1) / 2);
        }
        a.colour.start[c.planes.n] = acc;
        a.colour_index = col;
        if (acc == 0)
            continue;
        KernelTimer kt(c.stream, "k_level0_search", 0.0, 0.0);
        hipLaunchKernelGGL(k_level0_search, dim3(blocks_for(acc, SEARCH_THREADS)), dim3(SEARCH_THREADS), 0, c.stream,
                           a);
    }
    HIP_OK(hipGetLastError());
    const float ms = c.ms_since(c.ev0);
    const unsigned int* h = read_back(c, (const unsigned int*)a.moved, (size_t)ncount);
    // One launch per colour covers every plane, so the block's time is reported on the base plane.
    for (int p = 0; p < c.planes.n; p++)
    {
        nntc_level0_report& r = rep[p];
        r = nntc_level0_report {};
        r.joint = a.joint;
        r.states = states;
        r.ms = p == 0 ? ms : 0.0;
        for (int col = 0; col < 4; col++)
        {
            r.moved[col] = h[p * 4 + col];
            r.moved_total += r.moved[col];
        }
    }
}

}   // namespace nntc
