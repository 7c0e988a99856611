// decoder.hip: block (a)'s normal equations, A = sum omega v v^T (upper triangle) and rhs = sum omega v t^T over every
// site of every plane, v = [phi; 1] (src/solve_decoder.cu:87-180, src/cpu/cpu_solve_decoder.cpp:125-197).
//
// Plain form (the reference's own structure, which is also its summation order): a launch of SITE_BLOCKS x planes
// workgroups of LS_TILE threads. Workgroup b of plane p walks the tiles b, b + SITE_BLOCKS, ... of LS_TILE sites; each
// thread builds one site's row (phi, the constant one, the target; a site past the plane's end is a row of zeros) into
// an LDS tile stored entry-major, then every normal-equation entry is one thread's serial double dot product over the
// tile, added into that thread's accumulator. The accumulators land in column b of the plane's partials, and a second
// kernel folds, per entry, the planes in order and inside each plane the workgroups in order, each plane's sum scaled
// by its site weight. That is the reference CPU backend's order exactly, so A and rhs are its bits.
//
// Roofline: the useful flops are 2 S (mm (mm + 1) / 2 + nout mm) for S sites (4.19 GF at 4 x 1024^2); the fp64 dot
// products run on the vector ALUs, so the bound is fp64 peak. The MFMA split-precision form replaces this tile
// product when the exact form is not asked for.

#include "arm.h"

namespace nntc
{

namespace
{

struct LsArgs
{
    Planes planes;
    const float* src;
    const float* v0;
    const float* v1;
    int c0, c1, nin, nout, k;
    int entries, tri, mm;
    double* partial;   // (plane * entries + e) * SITE_BLOCKS + b
};

// The two staged rows whose dot product entry e is: the triangle's (i, j) row by row, then rhs (target c, feature i).
__device__ __forceinline__ void entry_rows(int e, int tri, int mm, int& ra, int& rb)
{
    if (e < tri)
    {
        int i = 0, rest = e;
        while (rest >= mm - i)
            rest -= mm - i++;
        ra = i;
        rb = i + rest;
    }
    else
    {
        const int e2 = e - tri;
        ra = mm + e2 / mm;
        rb = e2 % mm;
    }
}

constexpr int MAX_ENTRIES = 25 * 26 / 2 + MAX_NOUT * 25;   // 775 at the caps
constexpr int PER_THREAD = (MAX_ENTRIES + LS_TILE - 1) / LS_TILE;

__global__ __launch_bounds__(LS_TILE) void k_ls_accumulate(LsArgs a)
{
    __shared__ float tile[(MAX_NIN + 1 + MAX_NOUT) * LS_TILE];
    const int b = blockIdx.x;
    const int p = blockIdx.y;
    const PlaneGeom& g = a.planes.p[p];
    const int tid = threadIdx.x;
    const int per_pixel = 1 + a.k;
    const long long sites = g.n0 * per_pixel;
    const long long tiles = (sites + LS_TILE - 1) / LS_TILE;
    const float* src = a.src + g.off0 * a.nout;
    const float* v0 = a.v0 + g.off0 * a.c0;
    const float* v1 = a.v1 + g.off1 * a.c1;

    int ra[PER_THREAD], rb[PER_THREAD];
    double acc[PER_THREAD];
    for (int q = 0; q < PER_THREAD; q++)
    {
        acc[q] = 0.0;
        const int e = tid + q * LS_TILE;
        if (e < a.entries)
            entry_rows(e, a.tri, a.mm, ra[q], rb[q]);
        else
            ra[q] = rb[q] = 0;
    }

    for (long long t = b; t < tiles; t += SITE_BLOCKS)
    {
        const long long si = t * LS_TILE + tid;
        float phi[MAX_NIN], target[MAX_NOUT];
        for (int i = 0; i < a.nin; i++)
            phi[i] = 0.0f;
        for (int c = 0; c < a.nout; c++)
            target[c] = 0.0f;
        float one = 0.0f;
        if (si < sites)
        {
            const long long pixel = si / per_pixel;
            const int j = (int)(si % per_pixel);
            const int px = (int)(pixel % g.w0), py = (int)(pixel / g.w0);
            site_row(v0, v1, src, g.w0, g.h0, g.w1, g.h1, a.c0, a.c1, a.nout, a.k, j, px, py, phi, target);
            one = 1.0f;
        }
        for (int i = 0; i < a.nin; i++)
            tile[i * LS_TILE + tid] = phi[i];
        tile[a.nin * LS_TILE + tid] = one;
        for (int c = 0; c < a.nout; c++)
            tile[(a.mm + c) * LS_TILE + tid] = target[c];
        __syncthreads();
        for (int q = 0; q < PER_THREAD; q++)
        {
            if (tid + q * LS_TILE >= a.entries)
                break;
            const float* x = tile + ra[q] * LS_TILE;
            const float* y = tile + rb[q] * LS_TILE;
            double sum = 0.0;
            for (int s = 0; s < LS_TILE; s++)
                sum += (double)x[s] * (double)y[s];
            acc[q] += sum;
        }
        __syncthreads();
    }
    for (int q = 0; q < PER_THREAD; q++)
    {
        const int e = tid + q * LS_TILE;
        if (e < a.entries)
            a.partial[((size_t)p * a.entries + e) * SITE_BLOCKS + b] = acc[q];
    }
}

// Per entry: the planes in index order, inside each plane the SITE_BLOCKS partials in index order, each plane's sum
// times its own site weight.
__global__ __launch_bounds__(64) void k_ls_reduce(const double* partial, const double* omega, int planes, int entries,
                                                  double* out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= entries)
        return;
    double total = 0.0;
    for (int p = 0; p < planes; p++)
    {
        const double* row = partial + ((size_t)p * entries + e) * SITE_BLOCKS;
        double sum = 0.0;
        for (int b = 0; b < SITE_BLOCKS; b++)
            sum += row[b];
        total += omega[p] * sum;
    }
    out[e] = total;
}

}   // namespace

void ls_accumulate(Ctx& c, int k, const double* omega, std::vector<double>& acc)
{
    LsArgs a;
    a.planes = c.planes;
    a.src = c.src.p;
    a.v0 = c.v0.p;
    a.v1 = c.v1.p;
    a.c0 = c.c0;
    a.c1 = c.c1;
    a.nin = c.nin;
    a.nout = c.nout;
    a.k = k;
    a.mm = c.nin + 1;
    a.tri = a.mm * (a.mm + 1) / 2;
    a.entries = a.tri + c.nout * a.mm;
    a.partial = c.ls_partial.p;
    if ((size_t)c.planes.n * a.entries * SITE_BLOCKS > c.ls_partial.n)
        throw Refused { "decoder: partials exceed the workspace" };
    {
        const double n0 = (double)c.n0_total, n1 = (double)c.n1_total, s = n0 * (1 + k);
        const double bytes = 4.0 * n0 * (c.nout + c.c0) + 4.0 * n1 * c.c1;
        const double flops = 2.0 * s * ((double)a.mm * (a.mm + 1) / 2.0 + (double)c.nout * a.mm);
        KernelTimer kt(c.stream, "k_ls_accumulate", bytes, flops);
        hipLaunchKernelGGL(k_ls_accumulate, dim3(SITE_BLOCKS, c.planes.n), dim3(LS_TILE), 0, c.stream, a);
    }
    HIP_OK(hipGetLastError());
    // omega travels through the scalar workspace's tail: nplanes doubles uploaded with the launch.
    double* d_omega = c.ls_sum.p + a.entries;
    HIP_OK(hipMemcpyAsync(d_omega, omega, sizeof(double) * c.planes.n, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_ls_reduce, dim3(blocks_for(a.entries, 64)), dim3(64), 0, c.stream, (const double*)a.partial,
                       (const double*)d_omega, c.planes.n, a.entries, c.ls_sum.p);
    HIP_OK(hipGetLastError());
    acc.resize((size_t)a.entries);
    HIP_OK(hipMemcpyAsync(acc.data(), c.ls_sum.p, sizeof(double) * a.entries, hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
}

}   // namespace nntc
