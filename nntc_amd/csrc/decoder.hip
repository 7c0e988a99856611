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
// products run on the vector ALUs, so the bound is fp64 peak. The fast form replaces this with the matrix-core Gram
// below (k_gram_tiles).

#include "arm.h"
#include "tile_stage.h"

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

// Per entry (one wavefront each): the planes in index order, inside each plane the SITE_BLOCKS partials in index
// order, each plane's sum times its own site weight. The wavefront stages a plane's partials in LDS with coalesced
// reads and lane 0 adds them in order, sixteen read ahead of their additions.
__global__ __launch_bounds__(64) void k_ls_reduce(const double* partial, const double* omega, int planes, int entries,
                                                  double* out)
{
    __shared__ double row[SITE_BLOCKS];
    const int e = blockIdx.x, lane = threadIdx.x;
    double total = 0.0;
    for (int p = 0; p < planes; p++)
    {
        const double* src = partial + ((size_t)p * entries + e) * SITE_BLOCKS;
        for (int b = lane; b < SITE_BLOCKS; b += 64)
            row[b] = src[b];
        __syncthreads();
        if (lane == 0)
            total += omega[p] * fold_serial16(row, SITE_BLOCKS, 0.0);
        __syncthreads();
    }
    if (lane == 0)
        out[e] = total;
}

// ---------------------------------------------------------------------------------------------------------------------
// The fast form: the Gram on the matrix cores, BF16 hi/lo split products accumulated in fp32
// ---------------------------------------------------------------------------------------------------------------------
//
// With u = [phi; 1; t] (D = mm + nout values per site), every normal-equation entry is an entry of the Gram
// G = sum u u^T: A = G[0:mm, 0:mm], rhs = G[mm + c, i]. One 256-lane workgroup per pixel tile (tile_stage.h) builds the
// u of its 1 + K sites per pixel, splits each value into bf16 hi + lo (x - hi is exact in fp32, so hi + lo carries 16
// of x's 24 mantissa bits), stores them feature-major in LDS, and each wavefront multiplies its 64 sites' slice on the
// matrix cores: G_tile += hi hi^T + hi lo^T + lo hi^T, v_mfma_f32_32x32x16_bf16 per 32 x 32 block of G and 16
// sites, accumulated in fp32 over the wavefront's 320 sites. The four wavefronts' blocks are added in double, in order,
// and the tile's entries written; a second kernel folds the tiles of each plane in order, times the plane's weight.
// Roofline: 2 S D^2 MFMA flops per pass (three split products), bounded by the feature pass that feeds them.

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short short8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

struct GramArgs
{
    Planes planes;
    Ranges tiles;          // plane p owns tiles [start[p], start[p+1])
    const float* src;
    const float* v0;
    const float* v1;
    int nout, k, mm, tri, entries;
    double* partial;       // e * tiles + tile (entry-major)
};

// x as bf16 rounded to nearest even (the hardware conversion; finite inputs), and the bf16 of what it leaves.
__device__ __forceinline__ void split_bf16(float x, unsigned short& hi, unsigned short& lo)
{
    hi = __builtin_bit_cast(unsigned short, (__bf16)x);
    const float rest = x - __uint_as_float((unsigned int)hi << 16);
    lo = __builtin_bit_cast(unsigned short, (__bf16)rest);
}

// Site s of feature f in the hi / lo tables. A matrix-core operand read takes one 16-byte group of eight sites from
// each of 32 features; with rows of 512 bytes every such group would sit in the same LDS banks, so the groups of a row
// are XOR-swizzled by the feature and the 32 reads spread over the banks.
__device__ __forceinline__ int gram_slot(int f, int s)
{
    return f * 256 + ((((s >> 3) ^ (f & 31)) << 3) | (s & 7));
}

template <int C0, int C1, int DB>
__global__ __launch_bounds__(256) void k_gram_tiles(GramArgs a)
{
    constexpr int NIN = C1 + C0 + C0 * C1, MM = NIN + 1;
    constexpr int DP = 32 * DB;                 // features padded to whole 32-blocks
    constexpr int NT = DB * (DB + 1) / 2;       // the blocks of G's upper block triangle
    __shared__ float s_l0[tile_l0_floats<C0>()];
    __shared__ float s_l1[TILE_L1_FLOATS];
    __shared__ __align__(16) unsigned short s_hi[DP * 256], s_lo[DP * 256];   // [feature][site], gram_slot order
    __shared__ double s_g[DP * DP];

    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int p = range_plane(a.tiles, tile);
    const int nout = a.nout, tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const TileGeom tg = stage_tile<C0, C1>(a.planes.p[p], tile - (int)a.tiles.start[p], nout, a.src, a.v0, a.v1,
                                           s_l0, s_l1);
    for (int i = tid; i < DP * DP; i += 256)
        s_g[i] = 0.0;
    __syncthreads();

    const int px = tg.tx0 + tid % TILE_W, py = tg.ty0 + tid / TILE_W;
    const bool live = px < tg.w0 && py < tg.h0;
    floatx16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int r = 0; r < 16; r++)
            acc[t][r] = 0.0f;

    // Rows past the last target (the padding of G to whole 32-blocks) stay zero in every round.
    const int D = MM + nout;
#pragma unroll
    for (int f = 0; f < DP; f++)
        if (f >= D)
        {
            s_hi[gram_slot(f, tid)] = 0;
            s_lo[gram_slot(f, tid)] = 0;
        }
    for (int j = 0; j <= a.k; j++)
    {
        // ---- this round's site of every lane, as a column of bf16 hi and lo values (zeros past the plane)
        {
            float phi[NIN], target[MAX_NOUT];
            if (live)
                tile_site<C0, C1>(tg, s_l0, s_l1, a.k, j, px, py, phi, target);
#pragma unroll
            for (int f = 0; f < DP; f++)
            {
                if (f >= D)
                    continue;
                float x = 0.0f;
                if (live)
                {
                    if (f < NIN)
                        x = phi[f];
                    else if (f == NIN)
                        x = 1.0f;
                    else if (f - MM < nout && f - MM < MAX_NOUT)
                        x = target[f - MM];
                }
                unsigned short hi, lo;
                split_bf16(x, hi, lo);
                s_hi[gram_slot(f, tid)] = hi;
                s_lo[gram_slot(f, tid)] = lo;
            }
        }
        __syncthreads();
        // ---- the wavefront's 64 sites, 16 per product; lane (feature l % 32, sites 8 (l / 32) .. + 7)
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
        {
            const int site0 = 64 * wave + 16 * ks + 8 * (lane / 32);
            bf16x8 hv[DB], lv[DB];
#pragma unroll
            for (int b = 0; b < DB; b++)
            {
                const int f = 32 * b + lane % 32;
                hv[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const short8*>(s_hi + gram_slot(f, site0)));
                lv[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const short8*>(s_lo + gram_slot(f, site0)));
            }
            int t = 0;
#pragma unroll
            for (int bi = 0; bi < DB; bi++)
#pragma unroll
                for (int bj = bi; bj < DB; bj++, t++)
                {
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hv[bi], hv[bj], acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hv[bi], lv[bj], acc[t], 0, 0, 0);
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lv[bi], hv[bj], acc[t], 0, 0, 0);
                }
        }
        __syncthreads();
    }

    // ---- the four wavefronts' blocks added into G in double, wavefront by wavefront; lane l, value r holds
    // row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32 of its 32 x 32 block
    for (int w = 0; w < 4; w++)
    {
        if (wave == w)
        {
            int t = 0;
#pragma unroll
            for (int bi = 0; bi < DB; bi++)
#pragma unroll
                for (int bj = bi; bj < DB; bj++, t++)
#pragma unroll
                    for (int r = 0; r < 16; r++)
                    {
                        const int m = 32 * bi + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
                        const int n = 32 * bj + lane % 32;
                        s_g[m * DP + n] += (double)acc[t][r];
                    }
        }
        __syncthreads();
    }
    // ---- the tile's entries: A's upper triangle row by row, then rhs (target c, feature i); G's lower half is the
    // mirror of the blocks above
    for (int e = tid; e < a.entries; e += 256)
    {
        int i, jj;
        if (e < a.tri)
        {
            int rest = e;
            i = 0;
            while (rest >= MM - i)
                rest -= MM - i++;
            jj = i + rest;
        }
        else
        {
            const int e2 = e - a.tri;
            i = e2 % MM;           // feature
            jj = MM + e2 / MM;     // target, always in a block at or right of the feature's
        }
        const int bi = i / 32, bj = jj / 32;
        // Entry-major (entry e of every tile contiguous), so the reduction reads each entry's column coalesced.
        a.partial[(size_t)e * a.tiles.start[a.tiles.n] + tile] = bi <= bj ? s_g[i * DP + jj] : s_g[jj * DP + i];
    }
}

// One 256-lane workgroup per entry: each plane's tiles summed by the lanes (strided) and a tree, times the plane's
// weight, the planes added in order.
__global__ __launch_bounds__(256) void k_gram_reduce(const double* partial, const double* omega, Ranges tiles,
                                                     int entries, double* out)
{
    __shared__ double part[4];
    const int e = blockIdx.x, tid = threadIdx.x;
    double total = 0.0;
    for (int p = 0; p < tiles.n; p++)
    {
        double sum = 0.0;
        for (long long t = tiles.start[p] + tid; t < tiles.start[p + 1]; t += 256)
            sum += partial[(size_t)e * tiles.start[tiles.n] + t];
        for (int off = 32; off > 0; off >>= 1)
            sum += __shfl_xor(sum, off);
        if (tid % 64 == 0)
            part[tid / 64] = sum;
        __syncthreads();
        total += omega[p] * (part[0] + part[1] + part[2] + part[3]);
        __syncthreads();
    }
    if (tid == 0)
        out[e] = total;
}

template <int C0, int C1>
void launch_gram_tiles_db(int db, unsigned grid, hipStream_t s, const GramArgs& a)
{
    if (db == 1)
        hipLaunchKernelGGL((k_gram_tiles<C0, C1, 1>), dim3(grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((k_gram_tiles<C0, C1, 2>), dim3(grid), dim3(256), 0, s, a);
}

template <int C0>
void launch_gram_tiles_c1(int c1, int db, unsigned grid, hipStream_t s, const GramArgs& a)
{
    switch (c1)
    {
    case 1: launch_gram_tiles_db<C0, 1>(db, grid, s, a); break;
    case 2: launch_gram_tiles_db<C0, 2>(db, grid, s, a); break;
    case 3: launch_gram_tiles_db<C0, 3>(db, grid, s, a); break;
    default: launch_gram_tiles_db<C0, 4>(db, grid, s, a); break;
    }
}

void launch_gram_tiles(int c0, int c1, int db, unsigned grid, hipStream_t s, const GramArgs& a)
{
    switch (c0)
    {
    case 1: launch_gram_tiles_c1<1>(c1, db, grid, s, a); break;
    case 2: launch_gram_tiles_c1<2>(c1, db, grid, s, a); break;
    case 3: launch_gram_tiles_c1<3>(c1, db, grid, s, a); break;
    default: launch_gram_tiles_c1<4>(c1, db, grid, s, a); break;
    }
}

// The fast form's accumulation: tiled MFMA Gram, then the per-plane weighted fold.
void gram_accumulate(Ctx& c, int k, const double* omega, std::vector<double>& acc)
{
    GramArgs a;
    a.planes = c.planes;
    a.tiles.n = c.planes.n;
    long long total = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        const PlaneGeom& g = c.g(p);
        if (g.w1 > g.w0 || g.h1 > g.h0)
            throw Refused { "decoder: the tiled Gram needs level 1 no larger than level 0" };
        a.tiles.start[p] = total;
        total += objective_tiles(g);
    }
    a.tiles.start[c.planes.n] = total;
    a.src = c.src.p;
    a.v0 = c.v0.p;
    a.v1 = c.v1.p;
    a.nout = c.nout;
    a.k = k;
    a.mm = c.nin + 1;
    a.tri = a.mm * (a.mm + 1) / 2;
    a.entries = a.tri + c.nout * a.mm;
    a.partial = c.ls_partial.p;
    if ((size_t)total * a.entries > c.ls_partial.n)
        throw Refused { "decoder: tile partials exceed the workspace" };
    const int dims = a.mm + c.nout;
    const int db = dims <= 32 ? 1 : 2;
    if (dims > 64)
        throw Refused { "decoder: more than 64 Gram features" };
    {
        const double n0 = (double)c.n0_total, n1 = (double)c.n1_total, s = n0 * (1 + k);
        const double bytes = 4.0 * n0 * (c.nout + c.c0) + 4.0 * n1 * c.c1;
        const double flops = 2.0 * s * ((double)a.mm * (a.mm + 1) / 2.0 + (double)c.nout * a.mm);
        KernelTimer kt(c.stream, "k_gram_tiles", bytes, flops);
        launch_gram_tiles(c.c0, c.c1, db, (unsigned)total, c.stream, a);
    }
    HIP_OK(hipGetLastError());
    double* d_omega = c.ls_sum.p + a.entries;
    upload_staged(c, d_omega, omega, (size_t)c.planes.n);
    hipLaunchKernelGGL(k_gram_reduce, dim3(a.entries), dim3(256), 0, c.stream, (const double*)a.partial,
                       (const double*)d_omega, a.tiles, a.entries, c.ls_sum.p);
    HIP_OK(hipGetLastError());
    const double* h = read_back(c, (const double*)c.ls_sum.p, (size_t)a.entries);
    acc.assign(h, h + a.entries);
}

}   // namespace

void ls_accumulate(Ctx& c, int k, const double* omega, std::vector<double>& acc)
{
    if (c.form != NNTC_FORM_EXACT)
    {
        gram_accumulate(c, k, omega, acc);
        return;
    }
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
    upload_staged(c, d_omega, omega, (size_t)c.planes.n);
    hipLaunchKernelGGL(k_ls_reduce, dim3(a.entries), dim3(64), 0, c.stream, (const double*)a.partial,
                       (const double*)d_omega, c.planes.n, a.entries, c.ls_sum.p);
    HIP_OK(hipGetLastError());
    const double* h = read_back(c, (const double*)c.ls_sum.p, (size_t)a.entries);
    acc.assign(h, h + a.entries);
}

}   // namespace nntc
