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

// The plain form's fold of the same partials: one 256-lane workgroup per entry, each lane's strided blocks of every
// plane (times the plane's weight) summed in registers, then a tree.
__global__ __launch_bounds__(256) void k_ls_reduce_tree(const double* partial, const double* omega, int planes,
                                                        int entries, double* out)
{
    __shared__ double s_w[4];
    const int e = blockIdx.x;
    double acc = 0.0;
    for (int p = 0; p < planes; p++)
    {
        const double* src = partial + ((size_t)p * entries + e) * SITE_BLOCKS;
        double part = 0.0;
        for (int b = threadIdx.x; b < SITE_BLOCKS; b += 256)
            part += src[b];
        acc += omega[p] * part;
    }
    for (int off = 32; off > 0; off >>= 1)
        acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0)
        s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
        out[e] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
}

// ---------------------------------------------------------------------------------------------------------------------
// The fast form: the Gram on the matrix cores, BF16 hi/lo split products accumulated in fp32
// ---------------------------------------------------------------------------------------------------------------------
//
// With u = [phi; 1; t] (D = mm + nout values per site), every normal-equation entry is an entry of the Gram
// G = sum u u^T: A = G[0:mm, 0:mm], rhs = G[mm + c, i]. One 256-lane workgroup per pixel tile (tile_stage.h) builds the
// u of its 1 + K sites per pixel and splits each value exactly into three bf16 parts, x = h + m + l (h = bf16(x),
// m = bf16(x - h), l = bf16(x - h - m); each difference is exact in fp32, so the three parts carry all 24 of x's
// mantissa bits). The parts are stored feature-major in LDS and each wavefront multiplies its 64 sites' slice on the
// matrix cores, v_mfma_f32_32x32x16_bf16 per 32 x 32 block of G and 16 sites, keeping every product down to 2^-16 of
// h h^T: h h^T + h m^T + m h^T + m m^T + h l^T + l h^T (the products of two bf16 values are exact; the matrix core adds
// them in fp32). On a diagonal block the mirrored pairs are one product and its transpose, so a diagonal block takes
// four products, A += h h^T + m m^T and S += h m^T + h l^T, and G = A + S + S^T at the end. The leading part's fp32
// accumulator is flushed into fp64 registers after every 64 sites (one round of the wavefront), so fp32 rounding acts
// on at most 64 of its terms. The four wavefronts' blocks are added in double, in order, and the tile's entries written; a second kernel
// folds the tiles of each plane, times the plane's weight. (With two parts and three products - 16 bits of each
// feature - the decoder's small coefficients moved by up to 2 % of their size from the fp64 solve; with three parts and
// the per-round flush they stay within 1e-4.)
// Roofline: 2 S D^2 per product; the useful flops are the upper triangle and rhs, 2 S (mm (mm + 1) / 2 + nout mm).

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

// x = h + m + l in bf16, each part the nearest-even bf16 of what the parts before it leave (the hardware conversion;
// finite inputs). x - h and x - h - m are exact in fp32.
__device__ __forceinline__ void split3_bf16(float x, unsigned short& h, unsigned short& m, unsigned short& l)
{
    h = __builtin_bit_cast(unsigned short, (__bf16)x);
    const float r1 = x - __uint_as_float((unsigned int)h << 16);
    m = __builtin_bit_cast(unsigned short, (__bf16)r1);
    const float r2 = r1 - __uint_as_float((unsigned int)m << 16);
    l = __builtin_bit_cast(unsigned short, (__bf16)r2);
}

// Orders a wavefront's LDS accesses before and after it (writes then reads of other lanes' words, and back), without
// waiting for the rest of the workgroup.
__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Site s of feature f in a part's table. A matrix-core operand read takes one 16-byte group of eight sites from
// each of 32 features; with rows of 512 bytes every such group would sit in the same LDS banks, so the groups of a row
// are XOR-swizzled by the feature and the 32 reads spread over the banks.
__device__ __forceinline__ int gram_slot(int f, int s)
{
    return f * 256 + ((((s >> 3) ^ (f & 31)) << 3) | (s & 7));
}

// The workgroup's LDS (dynamic, sized by the launch): the three part tables [feature][site] (gram_slot order; after the
// last round the first 8 DP^2 bytes hold G in double, the tables being no longer read then), then the staged level 1,
// then the staged level 0 at nout + C0 floats a pixel (odd), so a workload of 12 outputs keeps the one-block kernel
// within half the CU's LDS and two workgroups share a CU.
template <int DB>
constexpr size_t gram_table_bytes()
{
    return (size_t)3 * 32 * DB * 256 * sizeof(unsigned short);
}
inline size_t gram_lds_bytes(int db, int nout, int c0)
{
    return (size_t)3 * 32 * db * 256 * sizeof(unsigned short) + (size_t)TILE_L1_FLOATS * sizeof(float) +
           (size_t)TILE_L0_PIX * tile_l0_stride(nout, c0) * sizeof(float);
}

// Two workgroups a CU for the one-block form (registers held to 256 a lane); the two-block form's tables take most of
// the LDS, one workgroup a CU.
template <int C0, int C1, int DB>
__global__ __launch_bounds__(256, DB == 1 ? 2 : 1) void k_gram_tiles(GramArgs a)
{
    constexpr int NIN = C1 + C0 + C0 * C1, MM = NIN + 1;
    constexpr int DP = 32 * DB;                 // features padded to whole 32-blocks
    constexpr int NT = DB * (DB + 1) / 2;       // the blocks of G's upper block triangle
    constexpr int TABLE = DP * 256;             // bf16 values per part table
    extern __shared__ __align__(16) unsigned char s_gram_lds[];
    static_assert(DP * DP * sizeof(double) <= gram_table_bytes<DB>(), "G fits over the tables");
    unsigned short* s_h = reinterpret_cast<unsigned short*>(s_gram_lds);
    unsigned short* s_m = s_h + TABLE;
    unsigned short* s_lo = s_h + 2 * TABLE;
    double* s_g = reinterpret_cast<double*>(s_gram_lds);
    float* s_l1 = reinterpret_cast<float*>(s_gram_lds + gram_table_bytes<DB>());
    float* s_l0 = s_l1 + TILE_L1_FLOATS;

    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int p = range_plane(a.tiles, tile);
    const int nout = a.nout, tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const TileGeom tg = stage_tile<C0, C1>(a.planes.p[p], tile - (int)a.tiles.start[p], nout, a.src, a.v0, a.v1,
                                           s_l0, s_l1);
    const int px = tg.tx0 + tid % TILE_W, py = tg.ty0 + tid / TILE_W;
    const bool live = px < tg.w0 && py < tg.h0;
    // acc[t]: block t's h h^T + m m^T (diagonal) or all six products (off the diagonal), flushed into acc64 (fp64)
    // after every round; sacc[b]: diagonal block b's h m^T + h l^T, 2^-8 of the block's size, kept in fp32 over the
    // tile's 320 sites (its rounding then sits 2^-8 below the flushed part's).
    floatx16 acc[NT], sacc[DB];
    double acc64[NT][16];
#pragma unroll
    for (int t = 0; t < NT; t++)
#pragma unroll
        for (int r = 0; r < 16; r++)
            acc64[t][r] = 0.0;
#pragma unroll
    for (int b = 0; b < DB; b++)
#pragma unroll
        for (int r = 0; r < 16; r++)
            sacc[b][r] = 0.0f;

    __syncthreads();   // the staged tile, written by every lane, is read by every lane's sites below
    for (int j = 0; j <= a.k; j++)
    {
        // ---- this round's site of every lane, as a column of its three bf16 parts (zeros past the plane)
        {
            float phi[NIN], target[MAX_NOUT];
            if (live)
                tile_site<C0, C1>(tg, s_l0, s_l1, a.k, j, px, py, phi, target);
            // Every feature slot is written every round (a padding slot or a lane past the plane writes zeros), its
            // value picked by selects on compile-time slot numbers: no branch per feature.
#pragma unroll
            for (int f = 0; f < DP; f++)
            {
                float x;
                if (f < NIN)
                    x = phi[f];
                else if (f == NIN)
                    x = 1.0f;
                else if (f - MM < MAX_NOUT)
                    x = f - MM < nout ? target[f - MM] : 0.0f;
                else
                    x = 0.0f;
                x = live ? x : 0.0f;
                unsigned short h, m, l;
                split3_bf16(x, h, m, l);
                s_h[gram_slot(f, tid)] = h;
                s_m[gram_slot(f, tid)] = m;
                s_lo[gram_slot(f, tid)] = l;
            }
        }
        // A wavefront reads back only the site columns its own lanes wrote (sites 64 wave .. 64 wave + 63), so the
        // rounds need no workgroup barrier: a wavefront-scope fence orders its LDS writes before its reads (and, at
        // the end of the round, its reads before the next round's writes), and the four wavefronts run their rounds
        // out of step, one's feature pass beside another's products.
        wave_lds_fence();
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                acc[t][r] = 0.0f;
        // ---- the wavefront's 64 sites, 16 per product; lane (feature l % 32, sites 8 (l / 32) .. + 7)
#pragma unroll
        for (int ks = 0; ks < 4; ks++)
        {
            const int site0 = 64 * wave + 16 * ks + 8 * (lane / 32);
            bf16x8 hv[DB], mv[DB], lv[DB];
#pragma unroll
            for (int b = 0; b < DB; b++)
            {
                const int slot = gram_slot(32 * b + lane % 32, site0);
                hv[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const short8*>(s_h + slot));
                mv[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const short8*>(s_m + slot));
                lv[b] = __builtin_bit_cast(bf16x8, *reinterpret_cast<const short8*>(s_lo + slot));
            }
            int t = 0;
#pragma unroll
            for (int bi = 0; bi < DB; bi++)
#pragma unroll
                for (int bj = bi; bj < DB; bj++, t++)
                {
                    if (bi == bj)
                    {
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hv[bi], hv[bj], acc[t], 0, 0, 0);
                        sacc[bi] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hv[bi], mv[bj], sacc[bi], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(mv[bi], mv[bj], acc[t], 0, 0, 0);
                        sacc[bi] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hv[bi], lv[bj], sacc[bi], 0, 0, 0);
                    }
                    else
                    {
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hv[bi], hv[bj], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hv[bi], mv[bj], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(mv[bi], hv[bj], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(mv[bi], mv[bj], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(hv[bi], lv[bj], acc[t], 0, 0, 0);
                        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(lv[bi], hv[bj], acc[t], 0, 0, 0);
                    }
                }
        }
        // ---- the round's 64 sites into the fp64 sums
#pragma unroll
        for (int t = 0; t < NT; t++)
#pragma unroll
            for (int r = 0; r < 16; r++)
                acc64[t][r] += (double)acc[t][r];
        wave_lds_fence();
    }
    __syncthreads();   // G below overlays every wavefront's tables

    // ---- the four wavefronts' blocks added into G in double, wavefront by wavefront; lane l, value r holds
    // row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32 of its 32 x 32 block. A diagonal block adds S at its place,
    // then (a second sweep, so that no two lanes of one wavefront add to one entry at once) S^T at the mirrored place.
    for (int i = tid; i < DP * DP; i += 256)
        s_g[i] = 0.0;
    __syncthreads();
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
                        s_g[m * DP + n] += bi == bj ? acc64[t][r] + (double)sacc[bi][r] : acc64[t][r];
                    }
        }
        __syncthreads();
    }
    for (int w = 0; w < 4; w++)
    {
        if (wave == w)
        {
#pragma unroll
            for (int b = 0; b < DB; b++)
#pragma unroll
                for (int r = 0; r < 16; r++)
                {
                    const int m = 32 * b + 8 * (r / 4) + 4 * (lane / 32) + (r % 4);
                    const int n = 32 * b + lane % 32;
                    s_g[n * DP + m] += (double)sacc[b][r];
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
    const size_t lds = gram_lds_bytes(db, a.nout, C0);
    if (db == 1)
        hipLaunchKernelGGL((k_gram_tiles<C0, C1, 1>), dim3(grid), dim3(256), lds, s, a);
    else
        hipLaunchKernelGGL((k_gram_tiles<C0, C1, 2>), dim3(grid), dim3(256), lds, s, a);
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
void gram_accumulate(Ctx& c, int k, const double* omega, std::vector<double>& acc, bool to_host)
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
    if (!to_host)
        return;
    const double* h = read_back(c, (const double*)c.ls_sum.p, (size_t)a.entries);
    acc.assign(h, h + a.entries);
}

// ---------------------------------------------------------------------------------------------------------------------
// The fast form's decoder solve on the device: fp64 Cholesky with one step of iterative refinement
// ---------------------------------------------------------------------------------------------------------------------
//
// The reference solves [A + ridge I | rhs] by Gauss-Jordan on the host for each rung of its ridge ladder and accepts
// the first rung whose candidate does not raise the objective's quadratic Q beyond its rounding bound
// (src/solve_decoder.cu:267-458). The fast form keeps the ladder, the rounding to float and the acceptance rule, and
// solves each rung on the device from the fp64 Gram it accumulated: one 256-lane workgroup factors A + ridge I = L L^T
// (right-looking, a column per step), solves every output's system by forward and back substitution, and refines each
// solution once against the same fp64 system (r = rhs - (A + ridge I) x, x += (L L^T)^-1 r), which removes the
// factorisation's own rounding. A rung whose factorisation meets a pivot that is not positive and finite is skipped,
// as the reference skips a singular elimination.
//
// The split-bf16 Gram is accurate to about 1e-6 of its entries, far coarser than the first rung's ridge of 1e-9 of
// the mean diagonal. Where a feature is (nearly) a combination of the features before it - a level-0 channel that
// holds one nonzero value everywhere is a multiple of the bias - the part of the system the ridge is there to settle
// is below that accuracy, and the Cholesky factors of the rounded Gram settle it arbitrarily. The first rung
// therefore measures each feature's Schur complement (its pivot less the ridge) against its own diagonal entry, and
// a feature below DS_RESOLVE of it hands the call back to the host (status 4), which redoes it from the fp64 Gram.
// Exactly zero features (a level-0 plane at 0, whose products are exact zeros in any precision) are not referred.

constexpr int DS_MM = NNTC_MAX_NIN + 1, DS_NOUT = NNTC_MAX_NOUT;
constexpr double DS_RESOLVE = 1e-4;

struct DecoderSolveArgs
{
    const double* acc;         // tri + nout mm: A's upper triangle row by row, then rhs (output c, feature i)
    const double* x_prev;      // nout x mm: the incoming decoder (W row, then b), when had_previous
    const float* cw;
    int mm, nout, tri, had_previous, norm_ok;
    double inv_norm, accept_rise;
    double ladder[3];
    float* x_out;              // nout x mm: the decoder the call leaves (accepted, or the constant fallback)
    int* status;               // rung taken (0..2), 3 when every rung was refused, 4 when the Gram cannot resolve the
                               // system (see DS_RESOLVE); [1]: 1 when x_out holds a decoder
};

// sum_c cw_c Q_c(x) and its rounding bound, Q_c(x) = x_c^T A x_c - 2 x_c . rhs_c (x_c at x + c DS_MM); lanes < nout
// take an output each and every lane adds them in output order.
__device__ void ds_qform(const DecoderSolveArgs& a, const double* sA, const double* sR, const double* x, double* s_q,
                         double* s_e, double& q, double& err)
{
    const int mm = a.mm, tid = threadIdx.x;
    if (tid < a.nout)
    {
        const double* xc = x + tid * DS_MM;
        double qc = 0.0, ec = 0.0;
        for (int i = 0; i < mm; i++)
        {
            for (int j = 0; j < mm; j++)
            {
                const double v = sA[i * DS_MM + j];
                qc += xc[i] * v * xc[j];
                ec += fabs(xc[i]) * fabs(v) * fabs(xc[j]);
            }
            qc -= 2.0 * xc[i] * sR[tid * DS_MM + i];
            ec += 2.0 * fabs(xc[i]) * fabs(sR[tid * DS_MM + i]);
        }
        s_q[tid] = (double)a.cw[tid] * qc;
        s_e[tid] = (double)a.cw[tid] * ec;
    }
    __syncthreads();
    double qq = 0.0, ee = 0.0;
    for (int c = 0; c < a.nout; c++)
    {
        qq += s_q[c];
        ee += s_e[c];
    }
    q = qq;
    err = 2.220446049250313e-16 * (double)mm * ee * a.inv_norm;
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_decoder_solve(DecoderSolveArgs a)
{
    __shared__ double sA[DS_MM * DS_MM], sL[DS_MM * DS_MM], sR[DS_NOUT * DS_MM], sX[DS_NOUT * DS_MM];
    __shared__ double sP[DS_NOUT * DS_MM], s_q[DS_NOUT], s_e[DS_NOUT];
    __shared__ int s_ok, s_weak;
    const int mm = a.mm, nout = a.nout, tid = threadIdx.x;
    for (int e = tid; e < mm * mm; e += 256)
    {
        const int i = e / mm, j = e % mm, lo = i < j ? i : j, hi = i < j ? j : i;
        sA[i * DS_MM + j] = a.acc[lo * mm - lo * (lo - 1) / 2 + (hi - lo)];
    }
    for (int e = tid; e < nout * mm; e += 256)
    {
        sR[(e / mm) * DS_MM + e % mm] = a.acc[a.tri + e];
        sP[(e / mm) * DS_MM + e % mm] = a.had_previous ? a.x_prev[e] : 0.0;
    }
    __syncthreads();
    double trace = 0.0;
    for (int i = 0; i < mm; i++)
        trace += sA[i * DS_MM + i];
    const double mean_diag = trace > 0.0 ? trace / mm : 1.0;
    double q_prev = 0.0, err_prev = 0.0;
    if (a.had_previous)
        ds_qform(a, sA, sR, sP, s_q, s_e, q_prev, err_prev);
    int taken = 3;
    for (int step = 0; step < 3 && taken == 3; step++)
    {
        // Every lane has read the previous rung's s_ok and sL before they are written again.
        __syncthreads();
        const double ridge = a.ladder[step] * mean_diag;
        for (int e = tid; e < mm * mm; e += 256)
        {
            const int i = e / mm, j = e % mm;
            sL[i * DS_MM + j] = sA[i * DS_MM + j] + (i == j ? ridge : 0.0);
        }
        if (tid == 0)
        {
            s_ok = 1;
            s_weak = 0;
        }
        __syncthreads();
        // ---- L L^T = A + ridge I in the lower triangle of sL (right-looking: column j, then the trailing update)
        for (int j = 0; j < mm; j++)
        {
            if (tid == 0)
            {
                const double d = sL[j * DS_MM + j];
                if (!(d > 0.0) || !isfinite(d))
                    s_ok = 0;
                const double diag = sA[j * DS_MM + j];
                if (step == 0 && diag > 0.0 && !(d - ridge >= DS_RESOLVE * diag))
                    s_weak = 1;
                sL[j * DS_MM + j] = d > 0.0 ? sqrt(d) : 1.0;
            }
            __syncthreads();
            const double piv = sL[j * DS_MM + j];
            if (tid > j && tid < mm)
                sL[tid * DS_MM + j] /= piv;
            __syncthreads();
            const int rest = mm - j - 1;
            for (int e = tid; e < rest * rest; e += 256)
            {
                const int i = j + 1 + e / rest, k = j + 1 + e % rest;
                if (k <= i)
                    sL[i * DS_MM + k] -= sL[i * DS_MM + j] * sL[k * DS_MM + j];
            }
            __syncthreads();
        }
        if (s_weak)
        {
            taken = 4;
            break;
        }
        if (!s_ok)
            continue;
        // ---- each output's solve, refined once: lanes < nout take an output each
        if (tid < nout)
        {
            double* x = sX + tid * DS_MM;
            double r[DS_MM];
            for (int i = 0; i < mm; i++)
                r[i] = sR[tid * DS_MM + i];
            for (int pass = 0; pass < 2; pass++)
            {
                // forward: L y = r, then back: L^T z = y (z overwrites r)
                for (int i = 0; i < mm; i++)
                {
                    double v = r[i];
                    for (int k = 0; k < i; k++)
                        v -= sL[i * DS_MM + k] * r[k];
                    r[i] = v / sL[i * DS_MM + i];
                }
                for (int i = mm - 1; i >= 0; i--)
                {
                    double v = r[i];
                    for (int k = i + 1; k < mm; k++)
                        v -= sL[k * DS_MM + i] * r[k];
                    r[i] = v / sL[i * DS_MM + i];
                }
                if (pass == 0)
                {
                    for (int i = 0; i < mm; i++)
                        x[i] = r[i];
                    // the residual of the solve against the fp64 system
                    for (int i = 0; i < mm; i++)
                    {
                        double v = sR[tid * DS_MM + i] - ridge * x[i];
                        for (int k = 0; k < mm; k++)
                            v -= sA[i * DS_MM + k] * x[k];
                        r[i] = v;
                    }
                }
                else
                    for (int i = 0; i < mm; i++)
                        x[i] = (double)(float)(x[i] + r[i]);
            }
        }
        __syncthreads();
        int finite = 1;
        for (int e = 0; e < nout * mm; e++)
            finite = finite && isfinite(sX[(e / mm) * DS_MM + e % mm]);
        if (!finite)
            continue;
        double q_cand = 0.0, err_cand = 0.0;
        ds_qform(a, sA, sR, sX, s_q, s_e, q_cand, err_cand);
        if (!a.had_previous || (a.norm_ok && (q_cand - q_prev) * a.inv_norm + err_prev + err_cand <= a.accept_rise))
            taken = step;
    }
    int wrote = 0;
    if (taken < 3)
    {
        for (int e = tid; e < nout * mm; e += 256)
            a.x_out[e] = (float)sX[(e / mm) * DS_MM + e % mm];
        wrote = 1;
    }
    else if (taken == 3 && !a.had_previous)
    {
        // the constant fallback: every weight zero, each bias the mean target (src/solve_decoder.cu:440-452)
        const double count = sA[(mm - 1) * DS_MM + (mm - 1)];
        for (int e = tid; e < nout * mm; e += 256)
        {
            const int cc = e / mm, i = e % mm;
            a.x_out[e] = i == mm - 1 ? (count > 0.0 ? (float)(sR[cc * DS_MM + mm - 1] / count) : 0.0f) : 0.0f;
        }
        wrote = 1;
    }
    if (tid == 0)
    {
        a.status[0] = taken;
        a.status[1] = wrote;
    }
}

}   // namespace

void ls_accumulate(Ctx& c, int k, const double* omega, std::vector<double>& acc, bool to_host)
{
    if (c.form == NNTC_FORM_FAST)
    {
        gram_accumulate(c, k, omega, acc, to_host);
        return;
    }
    // The exact and the plain form: fp64 dots per site block; the exact form folds the blocks in the reference's
    // order, the plain form by trees.
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
    if (c.form == NNTC_FORM_EXACT)
        hipLaunchKernelGGL(k_ls_reduce, dim3(a.entries), dim3(64), 0, c.stream, (const double*)a.partial,
                           (const double*)d_omega, c.planes.n, a.entries, c.ls_sum.p);
    else
        hipLaunchKernelGGL(k_ls_reduce_tree, dim3(a.entries), dim3(256), 0, c.stream, (const double*)a.partial,
                           (const double*)d_omega, c.planes.n, a.entries, c.ls_sum.p);
    HIP_OK(hipGetLastError());
    if (!to_host)
        return;
    const double* h = read_back(c, (const double*)c.ls_sum.p, (size_t)a.entries);
    acc.assign(h, h + a.entries);
}

void decoder_solve_device(Ctx& c, const nntc_model_state* st, double inv_norm, bool norm_ok, const double ladder[3],
                          double accept_rise, int& taken, bool& wrote, std::vector<float>& x)
{
    const int mm = c.nin + 1, nout = c.nout;
    DecoderSolveArgs a;
    a.acc = c.ls_sum.p;
    a.mm = mm;
    a.nout = nout;
    a.tri = mm * (mm + 1) / 2;
    a.cw = c.cw.p;
    a.had_previous = st->has_decoder != 0;
    a.norm_ok = norm_ok ? 1 : 0;
    a.inv_norm = inv_norm;
    a.accept_rise = accept_rise;
    for (int i = 0; i < 3; i++)
        a.ladder[i] = ladder[i];
    double* xp = c.dec_io.p;
    a.x_prev = xp;
    a.x_out = reinterpret_cast<float*>(xp + (size_t)NNTC_MAX_NOUT * DS_MM);
    a.status = reinterpret_cast<int*>(a.x_out + (size_t)NNTC_MAX_NOUT * DS_MM);
    if (a.had_previous)
    {
        std::vector<double> prev((size_t)nout * mm);
        for (int cc = 0; cc < nout; cc++)
        {
            for (int i = 0; i < c.nin; i++)
                prev[(size_t)cc * mm + i] = (double)st->w[(size_t)cc * c.nin + i];
            prev[(size_t)cc * mm + c.nin] = (double)st->b[cc];
        }
        upload_staged(c, xp, prev.data(), prev.size());
    }
    hipLaunchKernelGGL(k_decoder_solve, dim3(1), dim3(256), 0, c.stream, a);
    HIP_OK(hipGetLastError());
    // One readback: the decoder (nout x mm floats) and the two status words after it.
    const size_t nx = (size_t)nout * mm;
    const float* h = read_back(c, a.x_out, (size_t)NNTC_MAX_NOUT * DS_MM + 2);
    x.assign(h, h + nx);
    const int* stat = reinterpret_cast<const int*>(h + (size_t)NNTC_MAX_NOUT * DS_MM);
    taken = stat[0];
    wrote = stat[1] != 0;
}

}   // namespace nntc
