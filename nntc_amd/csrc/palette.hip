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
    unsigned int* moved;    // per colour, MOVED_SLOTS counters a plane MOVED_STRIDE apart, colours moved_stride apart
    size_t moved_stride;
    int c0, c1, nin, nout, k, joint, sweeps;
    int bits[MAX_CH];
};

template <int C0>
__device__ __forceinline__ double search_energy(double a0, const double (&a1)[C0], const double (&a2)[C0 * C0],
                                                const float (&x)[C0])
{
    double e = a0;
#pragma unroll
    for (int i = 0; i < C0; i++)
    {
        e += 2.0 * (double)x[i] * a1[i];
#pragma unroll
        for (int j = 0; j < C0; j++)
            e += (double)x[i] * (double)x[j] * a2[i * C0 + j];
    }
    return e;
}

// One lane per texel of the colour. The channel counts are compile-time so that the quadratic's fp64 terms (A1, A2 and
// a site's partial sums g, S) stay in registers; the outputs are visited in an outer loop - for each output its
// decode, its two slopes and its residual, then its term of A0, g and S - so no per-site array of outputs is kept.
// Each of those sums still receives its terms in increasing output order, the reference's order.
template <int C0, int C1>
__device__ __forceinline__ bool search_texel(const SearchArgs& a, long long gi, int p)
{
    const PlaneGeom& g = a.planes.p[p];
    const int w0 = g.w0, h0 = g.h0, w1 = g.w1, h1 = g.h1;
    const int cx = a.colour_index & 1, cy = a.colour_index >> 1;
    const long long ci = gi - a.colour.start[p];
    const int cw0 = (w0 - cx + 1) / 2;
    const int tx = cx + 2 * (int)(ci % cw0), ty = cy + 2 * (int)(ci / cw0);
    const int nin = a.nin, nout = a.nout;
    const float* __restrict__ v0 = a.v0 + g.off0 * C0;
    const float* __restrict__ v1 = a.v1 + g.off1 * C1;
    const float* __restrict__ src = a.src + g.off0 * nout;
    const float* __restrict__ wts = a.weights;
    const float* __restrict__ bias = a.bias;
    const float* __restrict__ cwt = a.cw;
    const long long t = (long long)ty * w0 + tx;

    double a0 = 0.0, a1[C0], a2[C0 * C0];
#pragma unroll
    for (int i = 0; i < C0; i++)
        a1[i] = 0.0;
#pragma unroll
    for (int i = 0; i < C0 * C0; i++)
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
                float alpha = 1.0f, beta[C0];
#pragma unroll
                for (int i = 0; i < C0; i++)
                    beta[i] = 0.0f;
                // The source taps of a fractional site: the texel's own weight among them decides whether the site
                // sees the texel at all; a site that does not is the same constant in every state and is skipped.
                const float* s00 = src + t * nout;
                const float* s10 = s00;
                const float* s01 = s00;
                const float* s11 = s00;
                float q00 = 1.0f, q10 = 0.0f, q01 = 0.0f, q11 = 0.0f;
                if (j > 0)
                {
                    const Tap t0 = tap_pixel(px, py, sdx, sdy, w0, h0, w0, h0);
                    const int qx[4] = { t0.x0, t0.x1, t0.x0, t0.x1 };
                    const int qy[4] = { t0.y0, t0.y0, t0.y1, t0.y1 };
                    const float qw[4] = { (1.0f - t0.fx) * (1.0f - t0.fy), t0.fx * (1.0f - t0.fy),
                                          (1.0f - t0.fx) * t0.fy, t0.fx * t0.fy };
                    alpha = 0.0f;
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (qx[q] == tx && qy[q] == ty)
                            alpha += qw[q];
                    if (!(alpha > 0.0f))
                        continue;
                    // Every neighbour value is read before any is summed; the texel's own taps are not read, so a
                    // lane never reads a texel of its own colour, which another lane of the pass may be writing.
#pragma unroll
                    for (int q = 0; q < 4; q++)
                        if (!(qx[q] == tx && qy[q] == ty))
                        {
                            const float* nv = v0 + ((long long)qy[q] * w0 + qx[q]) * C0;
#pragma unroll
                            for (int i = 0; i < C0; i++)
                                beta[i] += qw[q] * nv[i];
                        }
                    s00 = src + ((long long)t0.y0 * w0 + t0.x0) * nout;
                    s10 = src + ((long long)t0.y0 * w0 + t0.x1) * nout;
                    s01 = src + ((long long)t0.y1 * w0 + t0.x0) * nout;
                    s11 = src + ((long long)t0.y1 * w0 + t0.x1) * nout;
                    q00 = qw[0];
                    q10 = qw[1];
                    q01 = qw[2];
                    q11 = qw[3];
                }
                float cbar[C1];
                {
                    const Tap t1 = tap_pixel(px, py, sdx, sdy, w0, h0, w1, h1);
                    const float* b00 = v1 + ((long long)t1.y0 * w1 + t1.x0) * C1;
                    const float* b10 = v1 + ((long long)t1.y0 * w1 + t1.x1) * C1;
                    const float* b01 = v1 + ((long long)t1.y1 * w1 + t1.x0) * C1;
                    const float* b11 = v1 + ((long long)t1.y1 * w1 + t1.x1) * C1;
                    const float w00 = (1.0f - t1.fx) * (1.0f - t1.fy), w10 = t1.fx * (1.0f - t1.fy);
                    const float w01 = (1.0f - t1.fx) * t1.fy, w11 = t1.fx * t1.fy;
#pragma unroll
                    for (int q = 0; q < C1; q++)
                        cbar[q] = w00 * b00[q] + w10 * b10[q] + w01 * b01[q] + w11 * b11[q];
                }
                double gs[C0], ss[C0 * C0];
#pragma unroll
                for (int i = 0; i < C0; i++)
                    gs[i] = 0.0;
#pragma unroll
                for (int i = 0; i < C0 * C0; i++)
                    ss[i] = 0.0;
                for (int r = 0; r < nout; r++)
                {
                    // p (the decode with the texel's channels at zero) and Q's column for output r
                    const float* row = wts + (size_t)r * nin;
                    float out = bias[r];
#pragma unroll
                    for (int q = 0; q < C1; q++)
                        out += row[q] * cbar[q];
                    float qc[C0];
#pragma unroll
                    for (int i = 0; i < C0; i++)
                    {
                        float slope = row[C1 + i];
#pragma unroll
                        for (int q = 0; q < C1; q++)
                            slope += row[C1 + C0 + i * C1 + q] * cbar[q];
                        qc[i] = slope;
                        out += slope * beta[i];
                    }
                    const float target = j == 0 ? s00[r] : q00 * s00[r] + q10 * s10[r] + q01 * s01[r] + q11 * s11[r];
                    const float resid = out - target;
                    const double cr = (double)cwt[r];
                    a0 += cr * (double)resid * (double)resid;
#pragma unroll
                    for (int i = 0; i < C0; i++)
                    {
                        gs[i] += cr * (double)qc[i] * (double)resid;
#pragma unroll
                        for (int i2 = 0; i2 < C0; i2++)
                            ss[i * C0 + i2] += cr * (double)qc[i] * (double)qc[i2];
                    }
                }
#pragma unroll
                for (int i = 0; i < C0; i++)
                {
                    a1[i] += (double)alpha * gs[i];
#pragma unroll
                    for (int i2 = 0; i2 < C0; i2++)
                        a2[i * C0 + i2] += (double)alpha * (double)alpha * ss[i * C0 + i2];
                }
            }
        }

    // The argmin over the palette from the current state (a tie keeps it).
    int cur[C0], best[C0], levels[C0];
    float x[C0];
    uint8_t* k0 = a.k0 + g.off0 * C0;
#pragma unroll
    for (int i = 0; i < C0; i++)
    {
        levels[i] = (1 << a.bits[i]) - 1;
        cur[i] = (int)k0[t * C0 + i];
        best[i] = cur[i];
        x[i] = a.palette[i * 16 + cur[i]];
    }
    double best_e = search_energy<C0>(a0, a1, a2, x);
    if (a.joint)
    {
        int total = 1;
#pragma unroll
        for (int i = 0; i < C0; i++)
            total *= levels[i] + 1;
        for (int n = 0; n < total; n++)
        {
            int rest = n;
            bool same = true;
            int state[C0];
#pragma unroll
            for (int i = 0; i < C0; i++)
            {
                state[i] = rest % (levels[i] + 1);
                rest /= levels[i] + 1;
                same = same && state[i] == cur[i];
                x[i] = a.palette[i * 16 + state[i]];
            }
            if (same)
                continue;
            const double e = search_energy<C0>(a0, a1, a2, x);
            if (e < best_e)
            {
                best_e = e;
#pragma unroll
                for (int i = 0; i < C0; i++)
                    best[i] = state[i];
            }
        }
    }
    else
    {
        for (int sw = 0; sw < a.sweeps; sw++)
        {
            bool changed = false;
#pragma unroll
            for (int i = 0; i < C0; i++)
            {
#pragma unroll
                for (int q = 0; q < C0; q++)
                    x[q] = a.palette[q * 16 + best[q]];
                const int hold = best[i];
                for (int kk = 0; kk <= levels[i]; kk++)
                {
                    if (kk == best[i])
                        continue;
                    x[i] = a.palette[i * 16 + kk];
                    const double e = search_energy<C0>(a0, a1, a2, x);
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
#pragma unroll
    for (int i = 0; i < C0; i++)
        moved = moved || best[i] != cur[i];
    if (moved)
    {
        float* v0w = a.v0 + g.off0 * C0;
#pragma unroll
        for (int i = 0; i < C0; i++)
        {
            v0w[t * C0 + i] = a.palette[i * 16 + best[i]];
            k0[t * C0 + i] = (uint8_t)best[i];
        }
    }
    return moved;
}

// The moved texels are counted as the quantised sweeps count theirs: the lanes of the workgroup's first plane summed
// by shuffles and LDS into one atomic a workgroup, a lane of another plane (a workgroup across a plane boundary)
// adding its own, the workgroups spread over MOVED_SLOTS counters (plane, colour) on separate cache lines - a count
// per moved texel on one address had the launches queue on it.
template <int C0, int C1>
__global__ __launch_bounds__(SEARCH_THREADS) void k_level0_search(SearchArgs a)
{
    const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = a.colour.start[a.colour.n];
    const bool live = gi < total;
    const int p = range_plane(a.colour, live ? gi : total - 1);
    const unsigned int moved = live && search_texel<C0, C1>(a, gi, p) ? 1u : 0u;
    __shared__ unsigned int s_wave[SEARCH_THREADS / 64];
    const int pw = range_plane(a.colour, (long long)blockIdx.x * blockDim.x);
    unsigned int same = p == pw ? moved : 0u;
    for (int off = 32; off > 0; off >>= 1)
        same += __shfl_xor(same, off);
    if (threadIdx.x % 64 == 0)
        s_wave[threadIdx.x / 64] = same;
    unsigned int* group = a.moved + (size_t)a.colour_index * a.moved_stride;
    if (p != pw && moved)
        atomicAdd(group + ((size_t)p * MOVED_SLOTS + blockIdx.x % MOVED_SLOTS) * MOVED_STRIDE, 1u);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned int sum = 0;
        for (int w = 0; w < SEARCH_THREADS / 64; w++)
            sum += s_wave[w];
        if (sum)
            atomicAdd(group + ((size_t)pw * MOVED_SLOTS + blockIdx.x % MOVED_SLOTS) * MOVED_STRIDE, sum);
    }
}

// Each (plane, colour)'s slots summed into out[plane * 4 + colour].
__global__ void k_fold_moved(const unsigned int* slots, size_t stride, int planes, unsigned int* out)
{
    const int i = (int)threadIdx.x;
    if (i >= planes * 4)
        return;
    const int p = i / 4, col = i % 4;
    unsigned int sum = 0;
    for (int sl = 0; sl < MOVED_SLOTS; sl++)
        sum += slots[(size_t)col * stride + ((size_t)p * MOVED_SLOTS + sl) * MOVED_STRIDE];
    out[i] = sum;
}

template <int C0>
void launch_search_c1(int c1, unsigned grid, hipStream_t s, const SearchArgs& a)
{
    switch (c1)
    {
    case 1: hipLaunchKernelGGL((k_level0_search<C0, 1>), dim3(grid), dim3(SEARCH_THREADS), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_level0_search<C0, 2>), dim3(grid), dim3(SEARCH_THREADS), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_level0_search<C0, 3>), dim3(grid), dim3(SEARCH_THREADS), 0, s, a); break;
    default: hipLaunchKernelGGL((k_level0_search<C0, 4>), dim3(grid), dim3(SEARCH_THREADS), 0, s, a); break;
    }
}

void launch_search(int c0, int c1, unsigned grid, hipStream_t s, const SearchArgs& a)
{
    switch (c0)
    {
    case 1: launch_search_c1<1>(c1, grid, s, a); break;
    case 2: launch_search_c1<2>(c1, grid, s, a); break;
    case 3: launch_search_c1<3>(c1, grid, s, a); break;
    default: launch_search_c1<4>(c1, grid, s, a); break;
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
    // The slots of the four colours, then the folded counts (four a plane) after them.
    const size_t stride = (size_t)c.planes.n * MOVED_SLOTS * MOVED_STRIDE;
    const int ncount = c.planes.n * 4;
    if (c.pal_counters.n < 4 * stride + (size_t)ncount)
        c.pal_counters.alloc(4 * stride + (size_t)ncount);
    a.moved = c.pal_counters.p;
    a.moved_stride = stride;
    hipLaunchKernelGGL(k_zero_u32, dim3((unsigned)blocks_for((long long)(4 * stride), 256)), dim3(256), 0, c.stream,
                       a.moved, (int)(4 * stride));
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
        // Algorithmic bytes: each texel's own source, values and indices read once and its values and indices
        // written once, and the level-1 planes read once. Flops, nominal: every candidate site of the texel's 3 x 3
        // footprint (9K + 1 of them) decoded in fp32 (nout x nin multiply-adds) and folded into the quadratic in fp64
        // (nout (1 + C0 + C0^2) multiply-adds); a site whose taps miss the texel stops before its decode.
        const double n_col = (double)acc;
        const double bytes = n_col * (4.0 * c.nout + 2.0 * (4.0 * c.c0 + c.c0)) + 4.0 * (double)c.n1_total * c.c1;
        const double flops = n_col * (9.0 * k + 1.0) * 2.0 * c.nout * (c.nin + 1.0 + c.c0 + (double)c.c0 * c.c0);
        KernelTimer kt(c.stream, "k_level0_search", bytes, flops);
        launch_search(c.c0, c.c1, (unsigned)blocks_for(acc, SEARCH_THREADS), c.stream, a);
    }
    HIP_OK(hipGetLastError());
    unsigned int* folded = c.pal_counters.p + 4 * stride;
    hipLaunchKernelGGL(k_fold_moved, dim3(1), dim3(4 * MAX_PLANES), 0, c.stream, (const unsigned int*)a.moved, stride,
                       c.planes.n, folded);
    HIP_OK(hipGetLastError());
    const float ms = c.ms_since(c.ev0);
    const unsigned int* h = read_back(c, (const unsigned int*)folded, (size_t)ncount);
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
