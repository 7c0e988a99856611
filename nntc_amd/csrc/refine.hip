// refine.hip: level 0's BC4/BC5 pack and its refinement by analysis by synthesis, on gfx950 (src/refine_bc.cu,
// src/cpu/cpu_refine_bc.cpp, shared/bc_pack.h, src/export.cpp:294-363).
//
//   seed    every 4x4 block of every channel packed stb_dxt-style from the 8-bit index plane: the block's largest and
//           smallest byte as the endpoints (eight-value mode), each texel's nearest palette entry as its selector, one
//           thread per block-channel (integer, bit-exact);
//   decode  every texel's plane value, index and correction from its block's bytes;
//   refine  one workgroup per 4x4 block of one colour of the block grid (blocks two apart share no site): the block's
//           exact quadratic gathered from block (c')'s stencil with every texel outside it held, each channel's
//           endpoints by a least-squares step and nine integer candidates, then two Gauss-Seidel passes of exact
//           8-way selector argmins, every step taken only on a strict decrease. The parallel parts (the gather, the
//           row products, the updates) are one thread per texel or row; every sum that feeds a decision is folded by
//           one thread over the sixteen texels in increasing order, as the reference CPU backend folds it;
//   pack    the endpoints and selectors written as the .dds's 8-byte BC4 blocks, and the pack's squared error in byte
//           units against the index plane the host mirrors (a report figure).

#include "arm.h"

#include <cmath>

namespace nntc
{

namespace
{

struct BcArgs
{
    Planes planes;
    int c0;
    const uint8_t* k0m;     // the index mirror (seed, error)
    uint8_t* ep;
    uint8_t* sel;
    float* v0;
    uint8_t* k0;
    double* delta;
    const float* prev;
    const double* stencil;
    const double* grad;
    const float* lo0;
    const float* hi0;
    double* taken;          // per block
    uint8_t* blocks;
    double* se;             // per block: packing error
};

// Block b (global) of plane p and channel c, seeded: bc_pack_level at 8 bits (shared/bc_pack.h:101-133, 196-224).
__global__ __launch_bounds__(256) void k_bc_seed(BcArgs a, long long total)
{
    const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total * a.c0)
        return;
    const long long gb = gi / a.c0;
    const int c = (int)(gi % a.c0);
    int p = 0;
    while (p + 1 < a.planes.n && gb >= a.planes.p[p + 1].offb)
        p++;
    const PlaneGeom& g = a.planes.p[p];
    const long long b = gb - g.offb;
    const int bx = (g.w0 + 3) / 4;
    const int x = (int)(b % bx), y = (int)(b / bx);
    const uint8_t* k0 = a.k0m + g.off0 * a.c0;
    int v[16];
    int kmin = 255, kmax = 0;
    for (int t = 0; t < 16; t++)
    {
        const int px = min(g.w0 - 1, x * 4 + t % 4), py = min(g.h0 - 1, y * 4 + t / 4);
        v[t] = k0[((size_t)py * g.w0 + px) * a.c0 + c];
        kmin = min(kmin, v[t]);
        kmax = max(kmax, v[t]);
    }
    const long long slot = gb * a.c0 + c;
    const int r0 = kmax, r1 = kmin;   // lround(k * 255.0 / 255) is k itself at 8 bits
    a.ep[slot * 2] = (uint8_t)r0;
    a.ep[slot * 2 + 1] = (uint8_t)r1;
    if (r0 == r1)
    {
        for (int t = 0; t < 16; t++)
            a.sel[slot * 16 + t] = 0;
        return;
    }
    float pal[8];
    for (int i = 0; i < 8; i++)
        pal[i] = bc4_entry(r0, r1, i);
    for (int t = 0; t < 16; t++)
    {
        const float target = (float)v[t] * 255.0f / 255.0f;
        int best = 0;
        float best_d = 1e30f;
        for (int i = 0; i < 8; i++)
        {
            const float d = fabsf(pal[i] - target);
            if (d < best_d)
            {
                best_d = d;
                best = i;
            }
        }
        a.sel[slot * 16 + t] = (uint8_t)best;
    }
}

// k_bc_decode: texel ti's plane value, index and correction from its block's bytes.
__global__ __launch_bounds__(256) void k_bc_decode(BcArgs a, long long total)
{
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= total)
        return;
    int p = 0;
    while (p + 1 < a.planes.n && gt >= a.planes.p[p + 1].off0)
        p++;
    const PlaneGeom& g = a.planes.p[p];
    const long long ti = gt - g.off0;
    const int px = (int)(ti % g.w0), py = (int)(ti / g.w0);
    const int bx = (g.w0 + 3) / 4;
    const long long gb = g.offb + (long long)(py / 4) * bx + (px / 4);
    const int t = (py % 4) * 4 + (px % 4);
    for (int c = 0; c < a.c0; c++)
    {
        const long long b = gb * a.c0 + c;
        const float byte = bc4_entry8(a.ep[b * 2], a.ep[b * 2 + 1], a.sel[b * 16 + t]);
        const float value = bc0_value(a.lo0[c], a.hi0[c], byte);
        const long long idx = gt * a.c0 + c;
        a.v0[idx] = value;
        a.k0[idx] = bc0_index(byte);
        a.delta[idx] = (double)value - (double)a.prev[idx];
    }
}

constexpr int RB_THREADS = 64;
constexpr int RB_MAXN = 16 * MAX_CH;

// k_bc_refine: one block of one colour per workgroup (src/cpu/cpu_refine_bc.cpp:138-384).
__global__ __launch_bounds__(RB_THREADS) void k_bc_refine(BcArgs a, Ranges colour, int colour_index)
{
    __shared__ double a2[RB_MAXN * RB_MAXN];
    __shared__ double a1[RB_MAXN], dv[RB_MAXN], uv[RB_MAXN], dlt[16];
    __shared__ double t5[5][16];
    __shared__ int s_ep[2 * MAX_CH];
    __shared__ unsigned char s_sel[16 * MAX_CH];
    __shared__ int s_new[3];          // new0, new1, accept
    __shared__ double s_move;
    __shared__ int s_which;
    __shared__ int s_i;

    const long long gi = blockIdx.x;
    const int p = range_plane(colour, gi);
    const PlaneGeom& g = a.planes.p[p];
    const long long ci = gi - colour.start[p];
    const int bxn = (g.w0 + 3) / 4;
    const int cx = colour_index & 1, cy = colour_index >> 1;
    const int cbx = (bxn - cx + 1) / 2;
    const int gx = cx + 2 * (int)(ci % cbx);
    const int gy = cy + 2 * (int)(ci / cbx);
    const int w0 = g.w0, h0 = g.h0, c0 = a.c0;
    const int n = 16 * c0;
    const int tid = threadIdx.x;
    const long long base = g.off0;
    const double* stencil = a.stencil;
    const size_t bstride = (size_t)(STENCIL_BLOCKS * c0 * c0);

    for (int i = tid; i < n * n; i += RB_THREADS)
        a2[i] = 0.0;
    for (int i = tid; i < n; i += RB_THREADS)
    {
        a1[i] = 0.0;
        dv[i] = 0.0;
        uv[i] = 0.0;
    }
    __syncthreads();

    // ---- the block's quadratic, one thread per texel; every entry of a2 receives at most one addition.
    if (tid < 16)
    {
        const int px = gx * 4 + (tid & 3), py = gy * 4 + (tid >> 2);
        if (px < w0 && py < h0)
        {
            const long long ti = base + (long long)py * w0 + px;
            const double* own = stencil + ti * bstride;
            double acc[MAX_CH];
            for (int i = 0; i < c0; i++)
                acc[i] = a.grad[ti * c0 + i];
            for (int slot = 0; slot < STENCIL_BLOCKS; slot++)
            {
                int dx, dy;
                stencil_offset(slot, dx, dy);
                const int fx = px + dx, fy = py + dy;
                if (fx >= 0 && fx < w0 && fy >= 0 && fy < h0)
                {
                    const double* blk = own + (size_t)slot * c0 * c0;
                    const int lx = fx - gx * 4, ly = fy - gy * 4;
                    if (lx >= 0 && lx < 4 && ly >= 0 && ly < 4)
                    {
                        const int ft = ly * 4 + lx;
                        for (int i = 0; i < c0; i++)
                            for (int j = 0; j < c0; j++)
                            {
                                a2[(tid * c0 + i) * n + (ft * c0 + j)] += blk[i * c0 + j];
                                if (slot != 0)
                                    a2[(ft * c0 + j) * n + (tid * c0 + i)] += blk[i * c0 + j];
                            }
                    }
                    else
                    {
                        const double* xn = a.delta + (base + (long long)fy * w0 + fx) * c0;
                        for (int i = 0; i < c0; i++)
                            for (int j = 0; j < c0; j++)
                                acc[i] += blk[i * c0 + j] * xn[j];
                    }
                }
                if (slot == 0)
                    continue;
                const int ax = px - dx, ay = py - dy;
                const int lx = ax - gx * 4, ly = ay - gy * 4;
                if (ax < 0 || ax >= w0 || ay < 0 || ay >= h0 || (lx >= 0 && lx < 4 && ly >= 0 && ly < 4))
                    continue;
                const long long nt = base + (long long)ay * w0 + ax;
                const double* blk = stencil + nt * bstride + (size_t)slot * c0 * c0;
                const double* xn = a.delta + nt * c0;
                for (int i = 0; i < c0; i++)
                    for (int j = 0; j < c0; j++)
                        acc[i] += blk[j * c0 + i] * xn[j];
            }
            for (int i = 0; i < c0; i++)
            {
                a1[tid * c0 + i] = acc[i];
                dv[tid * c0 + i] = (double)a.v0[ti * c0 + i] - (double)a.prev[ti * c0 + i];
            }
        }
    }
    __syncthreads();

    const long long gb = g.offb + (long long)gy * bxn + gx;
    if (tid < n)
    {
        double s = 0.0;
        for (int j = 0; j < n; j++)
            s += a2[tid * n + j] * dv[j];
        uv[tid] = s;
    }
    if (tid < 2 * c0)
        s_ep[tid] = a.ep[(gb * c0 + tid / 2) * 2 + (tid & 1)];
    if (tid < n)
        s_sel[tid] = a.sel[gb * c0 * 16 + tid];
    __syncthreads();

    double taken = 0.0;   // thread 0's running total
    const float* lo0 = a.lo0;
    const float* hi0 = a.hi0;

    // ---- (i) the endpoints of each channel
    for (int c = 0; c < c0; c++)
    {
        const double step = ((double)hi0[c] - (double)lo0[c]) / 255.0;
        if (tid < 16)
        {
            const int idx = tid * c0 + c;
            const double aw = bc4_weight8(s_sel[c * 16 + tid]);
            const double pp = step * aw;
            const double qq = step * (1.0 - aw);
            const double b = a1[idx] + uv[idx];
            double ap = 0.0, aq = 0.0;
            for (int j = 0; j < 16; j++)
            {
                const double e = a2[idx * n + (j * c0 + c)];
                const double aj = bc4_weight8(s_sel[c * 16 + j]);
                ap += e * step * aj;
                aq += e * step * (1.0 - aj);
            }
            t5[0][tid] = b * pp;
            t5[1][tid] = b * qq;
            t5[2][tid] = pp * ap;
            t5[3][tid] = pp * aq;
            t5[4][tid] = qq * aq;
        }
        __syncthreads();
        if (tid == 0)
        {
            double bp = 0.0, bq = 0.0, pap = 0.0, paq = 0.0, qaq = 0.0;
            for (int t = 0; t < 16; t++)
            {
                bp += t5[0][t];
                bq += t5[1][t];
                pap += t5[2][t];
                paq += t5[3][t];
                qaq += t5[4][t];
            }
            const int r0 = s_ep[2 * c], r1 = s_ep[2 * c + 1];
            const double det = pap * qaq - paq * paq;
            double d0 = 0.0, d1 = 0.0;
            if (det > 1e-12 * pap * qaq)
            {
                d0 = (paq * bq - qaq * bp) / det;
                d1 = (paq * bp - pap * bq) / det;
            }
            const double v0r = (double)r0 + d0, v1r = (double)r1 + d1;
            const int c0r = (int)floor(v0r + 0.5), c1r = (int)floor(v1r + 0.5);
            double best = 0.0;
            int best0 = r0, best1 = r1;
            for (int e0 = -1; e0 <= 1; e0++)
                for (int e1 = -1; e1 <= 1; e1++)
                {
                    const int a0 = c0r + e0 < 0 ? 0 : (c0r + e0 > 255 ? 255 : c0r + e0);
                    const int a1i = c1r + e1 < 0 ? 0 : (c1r + e1 > 255 ? 255 : c1r + e1);
                    if (a0 <= a1i)
                        continue;
                    const double f0 = (double)(a0 - r0), f1 = (double)(a1i - r1);
                    const double de =
                        2.0 * (bp * f0 + bq * f1) + pap * f0 * f0 + 2.0 * paq * f0 * f1 + qaq * f1 * f1;
                    if (de < best)
                    {
                        best = de;
                        best0 = a0;
                        best1 = a1i;
                    }
                }
            s_new[0] = best0;
            s_new[1] = best1;
            s_new[2] = (best0 == s_ep[2 * c] && best1 == s_ep[2 * c + 1]) ? 0 : 1;
        }
        __syncthreads();
        if (!s_new[2])
            continue;
        const int new0 = s_new[0], new1 = s_new[1];
        // The candidate weighed again in the arithmetic the plane is written in.
        if (tid < 16)
        {
            const float now = bc0_value(lo0[c], hi0[c], bc4_entry8(s_ep[2 * c], s_ep[2 * c + 1], s_sel[c * 16 + tid]));
            const float next = bc0_value(lo0[c], hi0[c], bc4_entry8(new0, new1, s_sel[c * 16 + tid]));
            dlt[tid] = (double)next - (double)now;
        }
        __syncthreads();
        if (tid < 16)
        {
            const int idx = tid * c0 + c;
            double quad = 0.0;
            for (int j = 0; j < 16; j++)
                quad += a2[idx * n + (j * c0 + c)] * dlt[j];
            t5[0][tid] = dlt[tid] * (2.0 * (a1[idx] + uv[idx]) + quad);
        }
        __syncthreads();
        if (tid == 0)
        {
            double exact = 0.0;
            for (int t = 0; t < 16; t++)
                exact += t5[0][t];
            s_new[2] = exact < 0.0 ? 1 : 0;
            if (exact < 0.0)
                taken += exact;
        }
        __syncthreads();
        if (s_new[2])
        {
            if (tid < n)
            {
                double add = 0.0;
                for (int j = 0; j < 16; j++)
                    add += a2[tid * n + (j * c0 + c)] * dlt[j];
                uv[tid] += add;
            }
            __syncthreads();
            if (tid == 0)
            {
                s_ep[2 * c] = new0;
                s_ep[2 * c + 1] = new1;
            }
            if (tid < 16)
                dv[tid * c0 + c] += dlt[tid];
        }
        __syncthreads();
    }

    // ---- (ii) the selectors, two Gauss-Seidel passes of exact 8-way argmins
    for (int pass = 0; pass < 2; pass++)
        for (int c = 0; c < c0; c++)
            for (int t = 0; t < 16; t++)
            {
                if (tid == 0)
                {
                    const int i = t * c0 + c;
                    const double b = a1[i] + uv[i], aa = a2[i * n + i];
                    const int r0 = s_ep[2 * c], r1 = s_ep[2 * c + 1];
                    const double now = (double)bc0_value(lo0[c], hi0[c], bc4_entry8(r0, r1, s_sel[c * 16 + t]));
                    double best = 0.0, move = 0.0;
                    int which = s_sel[c * 16 + t];
                    for (int kk = 0; kk < 8; kk++)
                    {
                        const double d = (double)bc0_value(lo0[c], hi0[c], bc4_entry8(r0, r1, kk)) - now;
                        const double de = 2.0 * b * d + aa * d * d;
                        if (de < best)
                        {
                            best = de;
                            move = d;
                            which = kk;
                        }
                    }
                    s_i = best < 0.0 ? i : -1;
                    s_move = move;
                    s_which = which;
                    if (best < 0.0)
                        taken += best;
                }
                __syncthreads();
                const int i = s_i;
                if (i >= 0)
                {
                    if (tid < n)
                        uv[tid] += a2[tid * n + i] * s_move;
                    __syncthreads();
                    if (tid == 0)
                    {
                        s_sel[c * 16 + t] = (unsigned char)s_which;
                        dv[i] += s_move;
                    }
                }
                __syncthreads();
            }

    // ---- what the block leaves behind
    if (tid < 2 * c0)
        a.ep[(gb * c0 + tid / 2) * 2 + (tid & 1)] = (uint8_t)s_ep[tid];
    if (tid < n)
        a.sel[gb * c0 * 16 + tid] = s_sel[tid];
    if (tid < 16)
    {
        const int px = gx * 4 + (tid & 3), py = gy * 4 + (tid >> 2);
        if (px < w0 && py < h0)
        {
            const long long ti = base + (long long)py * w0 + px;
            for (int c = 0; c < c0; c++)
            {
                const float byte = bc4_entry8(s_ep[2 * c], s_ep[2 * c + 1], s_sel[c * 16 + tid]);
                const float value = bc0_value(lo0[c], hi0[c], byte);
                const long long idx = ti * c0 + c;
                a.v0[idx] = value;
                a.k0[idx] = bc0_index(byte);
                a.delta[idx] = (double)value - (double)a.prev[idx];
            }
        }
    }
    if (tid == 0)
        a.taken[gb] = taken;
}

// The .dds blocks of every file of every plane (bc4_pack, shared/bc_pack.h:51-60), and each block's squared error in
// byte units against the mirror's exact index values over the texels inside the plane
// (src/export.cpp:283-292).
__global__ __launch_bounds__(256) void k_bc_pack(BcArgs a, long long total)
{
    const long long gb = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gb >= total)
        return;
    int p = 0;
    while (p + 1 < a.planes.n && gb >= a.planes.p[p + 1].offb)
        p++;
    const PlaneGeom& g = a.planes.p[p];
    const long long b = gb - g.offb;
    const int bx = (g.w0 + 3) / 4;
    const int x = (int)(b % bx), y = (int)(b / bx);
    const int files = a.c0 <= 2 ? 1 : 2;
    double se = 0.0;
    for (int f = 0; f < files; f++)
    {
        const int nc = f == 0 ? (a.c0 == 1 ? 1 : 2) : a.c0 - 2;
        const int bpb = nc == 1 ? 8 : 16;
        uint8_t* file = a.blocks + g.offb * 32 + (long long)f * g.nb * 16;
        for (int cc = 0; cc < nc; cc++)
        {
            const int channel = 2 * f + cc;
            const long long slot = gb * a.c0 + channel;
            const int r0 = a.ep[slot * 2], r1 = a.ep[slot * 2 + 1];
            unsigned long long bits = 0;
            for (int t = 0; t < 16; t++)
                bits |= (unsigned long long)(a.sel[slot * 16 + t] & 7) << (3 * t);
            uint8_t* out = file + b * bpb + cc * 8;
            out[0] = (uint8_t)r0;
            out[1] = (uint8_t)r1;
            for (int i = 0; i < 6; i++)
                out[2 + i] = (uint8_t)(bits >> (8 * i));
            for (int t = 0; t < 16; t++)
            {
                const int px = x * 4 + t % 4, py = y * 4 + t / 4;
                if (px >= g.w0 || py >= g.h0)
                    continue;
                const float dec = bc4_entry(r0, r1, a.sel[slot * 16 + t] & 7);
                const int k = a.k0m[(g.off0 + (long long)py * g.w0 + px) * a.c0 + channel];
                const double e = (double)dec - (double)k * 255.0 / 255.0;
                se += e * e;
            }
        }
    }
    a.se[gb] = se;
}

BcArgs bc_args(Ctx& c)
{
    BcArgs a;
    a.planes = c.planes;
    a.c0 = c.c0;
    a.k0m = c.k0_mirror.p;
    a.ep = c.bc_ep.p;
    a.sel = c.bc_sel.p;
    a.v0 = c.v0.p;
    a.k0 = c.k0.p;
    a.delta = c.lat0.x.p;
    a.prev = c.lat0.prev.p;
    a.stencil = c.lat0.stencil.p;
    a.grad = c.lat0.grad.p;
    a.lo0 = c.grid.p;
    a.hi0 = c.grid.p + 4;
    a.taken = c.bc_taken.p;
    a.blocks = c.bc_blocks.p;
    a.se = c.bc_se.p;
    return a;
}

}   // namespace

void bc_seed(Ctx& c)
{
    BcArgs a = bc_args(c);
    hipLaunchKernelGGL(k_bc_seed, dim3(blocks_for(c.nb_total * c.c0, 256)), dim3(256), 0, c.stream, a, c.nb_total);
    HIP_OK(hipGetLastError());
}

void bc_decode(Ctx& c)
{
    BcArgs a = bc_args(c);
    hipLaunchKernelGGL(k_bc_decode, dim3(blocks_for(c.n0_total, 256)), dim3(256), 0, c.stream, a, c.n0_total);
    HIP_OK(hipGetLastError());
}

// One pass: the four colours over every plane's block grid, then the pass figures folded in the reference's order:
// per plane, per colour, the colour's blocks in chunks of 64 (a chunk's decreases summed in block order, the chunks
// added in order), the colours added in order (src/cpu/cpu_refine_bc.cpp:398-447).
void bc_refine_pass(Ctx& c, double* taken, double* improved)
{
    BcArgs a = bc_args(c);
    Ranges colour[4];
    for (int col = 0; col < 4; col++)
    {
        const int cx = col & 1, cy = col >> 1;
        colour[col].n = c.planes.n;
        long long acc = 0;
        for (int p = 0; p < c.planes.n; p++)
        {
            colour[col].start[p] = acc;
            const int bx = (c.g(p).w0 + 3) / 4, by = (c.g(p).h0 + 3) / 4;
            acc += (long long)((bx - cx + 1) / 2) * ((by - cy + 1) / 2);
        }
        colour[col].start[c.planes.n] = acc;
    }
    for (int col = 0; col < 4; col++)
    {
        const long long n = colour[col].start[colour[col].n];
        if (n == 0)
            continue;
        // Each block reads its sixteen texels' stencil blocks, gradients, planes and corrections once.
        const double c0 = c.c0;
        KernelTimer kt(c.stream, "k_bc_refine", (double)n * 16.0 * (40.0 * c0 * c0 + 8.0 * c0 + 8.0 * c0 + 9.0 * c0),
                       0.0);
        hipLaunchKernelGGL(k_bc_refine, dim3((unsigned)n), dim3(RB_THREADS), 0, c.stream, a, colour[col], col);
        HIP_OK(hipGetLastError());
    }
    std::vector<double> t((size_t)c.nb_total);
    HIP_OK(hipMemcpyAsync(t.data(), c.bc_taken.p, t.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    const size_t chunk = 64;
    for (int p = 0; p < c.planes.n; p++)
    {
        const PlaneGeom& g = c.g(p);
        const int bx = (g.w0 + 3) / 4, by = (g.h0 + 3) / 4;
        double tk = 0.0, im = 0.0;
        for (int col = 0; col < 4; col++)
        {
            const int cx = col & 1, cy = col >> 1;
            const int cbx = (bx - cx + 1) / 2;
            const size_t count = (size_t)cbx * (size_t)((by - cy + 1) / 2);
            if (count == 0)
                continue;
            double ct = 0.0, cm = 0.0;
            for (size_t b0 = 0; b0 < count; b0 += chunk)
            {
                double pt = 0.0, pm = 0.0;
                for (size_t b = b0; b < b0 + chunk && b < count; b++)
                {
                    const int gx = cx + 2 * (int)(b % (size_t)cbx);
                    const int gy = cy + 2 * (int)(b / (size_t)cbx);
                    const double v = t[(size_t)(g.offb + (long long)gy * bx + gx)];
                    if (v < 0.0)
                    {
                        pt += v;
                        pm += 1.0;
                    }
                }
                ct += pt;
                cm += pm;
            }
            tk += ct;
            im += cm;
        }
        taken[p] = tk;
        improved[p] = im;
    }
}

void bc_pack_blocks(Ctx& c, double& se, size_t& n)
{
    BcArgs a = bc_args(c);
    hipLaunchKernelGGL(k_bc_pack, dim3(blocks_for(c.nb_total, 256)), dim3(256), 0, c.stream, a, c.nb_total);
    HIP_OK(hipGetLastError());
    std::vector<double> h((size_t)c.nb_total);
    HIP_OK(hipMemcpyAsync(h.data(), c.bc_se.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    for (double v : h)
        se += v;
    for (int p = 0; p < c.planes.n; p++)
        n += (size_t)c.g(p).n0 * (size_t)c.c0;
}

}   // namespace nntc
