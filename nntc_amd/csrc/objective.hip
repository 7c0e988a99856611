// objective.hip: the objective E, the decode of a stored level to bytes, and the centre residual, on gfx950.
//
// E (src/objective.cu:123-279, src/cpu/cpu_objective.cpp:112-291) is the cw-weighted squared error of the decoded
// material over every site of every plane - the centre and K fractional sites per pixel - plus the unweighted centre
// and fractional sums. One launch covers every plane.
//
// Exact form: the site space of each plane is cut into chunks of CHUNK consecutive sites, the reference CPU backend's
// own reduction unit, one wavefront each. Its 64 lanes evaluate 64 sites at a time in parallel (the double sampling
// rule, phi, the 3T x nin decode in double) and stage every site's per-output terms in LDS; then lanes 0, 1, 2 fold
// the weighted, centre and fractional terms serially in site-then-output order - the same additions in the same order
// as the reference's chunk loop, so each chunk's three sums are its bits. The host folds the chunk sums in chunk order
// (src/cpu/pool.h:314-324).
// Fast form: LDS-staged pixel tiles, fp32 sampling and decode, double sums per tile (k_objective_tiles below).
//
// Roofline: per pass the algorithmic bytes are the three planes read once, 4N(nout + c0) + 4 N1 c1; the flops are
// S 2 (4 c1 + 4 (c0 + nout) K / (K + 1) + c0 c1 + nout nin) for S = N (1 + K) sites, about 41 flop/byte at nout 12:
// compute-side of the ridge, so this kernel's fraction is reported against the vector peak (78.6 TFLOP/s, fp64 or
// fp32 without packed operations).

#include "arm.h"
#include "tile_stage.h"

#include <mutex>

namespace nntc
{

namespace
{

constexpr int OBJ_WAVE = 64;

// One site's error vector e[r] = b_r + W_r phi - target_r, in double, by the objective's double sampling rule.
__device__ __forceinline__ void objective_site(const float* v0, const float* v1, const float* src, int w0, int h0,
                                               int w1, int h1, int c0, int c1, int nin, int nout,
                                               const float* weights, const float* bias, int k_count, long long si,
                                               double* e, int& j_out)
{
    const int per_pixel = 1 + k_count;
    const long long pixel = si / per_pixel;
    const int j = (int)(si % per_pixel);
    const int px = (int)(pixel % w0), py = (int)(pixel / w0);
    float dx = 0.0f, dy = 0.0f;
    if (j > 0)
        subtexel_offset(k_count, j - 1, px, py, dx, dy);
    const double u = ((double)px + 0.5 + (double)dx) / (double)w0;
    const double v = ((double)py + 0.5 + (double)dy) / (double)h0;
    // Every array below is indexed by compile-time constants (loops unrolled to the caps, guarded by the runtime
    // counts), so it lives in registers; the operations and their order are the reference's.
    double sv[MAX_CH], c[MAX_CH], target[MAX_NOUT];
    obj_sample_r<MAX_CH>(v1, w1, h1, c1, u, v, c);
    if (j == 0)
    {
#pragma unroll
        for (int i = 0; i < MAX_CH; i++)
            if (i < c0)
                sv[i] = (double)v0[pixel * c0 + i];
#pragma unroll
        for (int i = 0; i < MAX_NOUT; i++)
            if (i < nout)
                target[i] = (double)src[pixel * nout + i];
    }
    else
    {
        obj_sample_r<MAX_CH>(v0, w0, h0, c0, u, v, sv);
        obj_sample_r<MAX_NOUT>(src, w0, h0, nout, u, v, target);
    }
#pragma unroll
    for (int r = 0; r < MAX_NOUT; r++)
        if (r < nout)
            e[r] = decode_row<double>(weights + (size_t)r * nin, bias[r], sv, c0, c, c1) - target[r];
    j_out = j;
}

struct ObjArgs
{
    Planes planes;
    Ranges chunks;            // reduction units (chunks or tiles): plane p owns [start[p], start[p+1])
    const float* src;
    const float* v0;
    const float* v1;
    const float* weights;
    const float* bias;
    const float* cw;
    int c0, c1, nin, nout, k;
    double* partial;          // 3 per chunk
};

__global__ __launch_bounds__(OBJ_WAVE) void k_objective(ObjArgs a)
{
    __shared__ double term[3][OBJ_WAVE * MAX_NOUT];
    const long long chunk = blockIdx.x;
    const int p = range_plane(a.chunks, chunk);
    const PlaneGeom& g = a.planes.p[p];
    const long long sites = g.n0 * (long long)(1 + a.k);
    const long long begin = (chunk - a.chunks.start[p]) * (long long)CHUNK;
    const long long end = begin + CHUNK < sites ? begin + CHUNK : sites;
    const int lane = threadIdx.x;
    const float* src = a.src + g.off0 * a.nout;
    const float* v0 = a.v0 + g.off0 * a.c0;
    const float* v1 = a.v1 + g.off1 * a.c1;
    const int nout = a.nout;

    double acc = 0.0;   // lane 0 weighted, lane 1 centre, lane 2 fractional
    for (long long base = begin; base < end; base += OBJ_WAVE)
    {
        const long long si = base + lane;
        const int count = (int)(end - base < OBJ_WAVE ? end - base : OBJ_WAVE);
        if (lane < count)
        {
            double e[MAX_NOUT];
            int j;
            objective_site(v0, v1, src, g.w0, g.h0, g.w1, g.h1, a.c0, a.c1, a.nin, nout, a.weights, a.bias, a.k, si,
                           e, j);
#pragma unroll
            for (int r = 0; r < MAX_NOUT; r++)
                if (r < nout)
                {
                    const double ee = e[r] * e[r];
                    const double we = (double)a.cw[r] * e[r] * e[r];
                    term[0][lane * nout + r] = we;
                    term[1][lane * nout + r] = j == 0 ? ee : 0.0;
                    term[2][lane * nout + r] = j == 0 ? 0.0 : ee;
                }
        }
        __syncthreads();
        {
            // The reference's chunk loop in its own order: site by site, output by output. Adding the 0.0 a lane's
            // array holds where the reference skips the add leaves a non-negative sum's bits unchanged.
            if (lane < 3)
            {
                // Sixteen terms read ahead of their sixteen dependent additions, so the chain waits on adds only.
                const double* t = term[lane];
                const int n = count * nout;
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
            }
        }
        __syncthreads();
    }
    if (lane < 3)
        a.partial[chunk * 3 + lane] = acc;
}

// ---------------------------------------------------------------------------------------------------------------------
// The fast form: LDS-staged tiles (tile_stage.h), the decode in fp32 with fused multiply-adds, the tile's sums in double
// ---------------------------------------------------------------------------------------------------------------------

template <int C0, int C1>
__global__ __launch_bounds__(256) void k_objective_tiles(ObjArgs a)
{
    constexpr int NIN = C1 + C0 + C0 * C1;
    __shared__ float s_l0[tile_l0_floats<C0>()];
    __shared__ float s_l1[TILE_L1_FLOATS];
    __shared__ double part[4][3];

    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int p = range_plane(a.chunks, tile);
    const int nout = a.nout, tid = threadIdx.x;
    const TileGeom tg = stage_tile<C0, C1>(a.planes.p[p], tile - (int)a.chunks.start[p], nout, a.src, a.v0, a.v1,
                                           s_l0, s_l1);
    // The decoder is the same for every lane: its rows are read at uniform addresses, which the compiler serves with
    // scalar loads into scalar registers, each a direct operand of the lanes' multiply-adds (no LDS copy, no vector
    // register per weight).
    const float* __restrict__ wts = a.weights;
    const float* __restrict__ bias = a.bias;
    const float* __restrict__ cw = a.cw;
    __syncthreads();

    // ---- one pixel per lane
    const int px = tg.tx0 + tid % TILE_W, py = tg.ty0 + tid / TILE_W;
    float acc_w = 0.0f, acc_c = 0.0f, acc_f = 0.0f;
    if (px < tg.w0 && py < tg.h0)
        for (int j = 0; j <= a.k; j++)
        {
            float phi[NIN], target[MAX_NOUT];
            tile_site<C0, C1>(tg, s_l0, s_l1, a.k, j, px, py, phi, target);
            float we = 0.0f, ee = 0.0f;
#pragma unroll
            for (int r = 0; r < MAX_NOUT; r++)
                if (r < nout)
                {
                    float out = bias[r];
#pragma unroll
                    for (int i = 0; i < NIN; i++)
                        out = fmaf(wts[r * NIN + i], phi[i], out);
                    const float e = out - target[r];
                    we = fmaf(cw[r] * e, e, we);
                    ee = fmaf(e, e, ee);
                }
            acc_w += we;
            if (j == 0)
                acc_c += ee;
            else
                acc_f += ee;
        }

    // ---- the tile's three sums: lane values in double, wavefront trees, the four wavefronts in order
    double s[3] = { (double)acc_w, (double)acc_c, (double)acc_f };
#pragma unroll
    for (int q = 0; q < 3; q++)
        for (int off = 32; off > 0; off >>= 1)
            s[q] += __shfl_xor(s[q], off);
    if (tid % 64 == 0)
        for (int q = 0; q < 3; q++)
            part[tid / 64][q] = s[q];
    __syncthreads();
    if (tid < 3)
        a.partial[(size_t)tile * 3 + tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
}

template <int C0>
void launch_objective_tiles_c1(int c1, unsigned grid, hipStream_t s, const ObjArgs& a)
{
    switch (c1)
    {
    case 1: hipLaunchKernelGGL((k_objective_tiles<C0, 1>), dim3(grid), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_objective_tiles<C0, 2>), dim3(grid), dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_objective_tiles<C0, 3>), dim3(grid), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_objective_tiles<C0, 4>), dim3(grid), dim3(256), 0, s, a); break;
    }
}

void launch_objective_tiles(int c0, int c1, unsigned grid, hipStream_t s, const ObjArgs& a)
{
    switch (c0)
    {
    case 1: launch_objective_tiles_c1<1>(c1, grid, s, a); break;
    case 2: launch_objective_tiles_c1<2>(c1, grid, s, a); break;
    case 3: launch_objective_tiles_c1<3>(c1, grid, s, a); break;
    default: launch_objective_tiles_c1<4>(c1, grid, s, a); break;
    }
}

struct DecodeArgs
{
    const float* v0;
    const float* v1;
    const float* src;
    const float* weights;
    const float* bias;
    int w0, h0, w1, h1, c0, c1, nin, nout;
    uint8_t* out;     // bytes when decoding
    float* resid;     // out - target when forming the residual
};

// One pixel decoded at its centre in the decoder's own fp32 arithmetic (src/objective.cu:53-75): saturated and
// rounded half away from zero to a byte, or, for the residual (src/init.cu:1024-1045), left unsaturated and
// differenced against the target.
// Pixel idx's features at its centre, and output r of the decoder on them (before saturation).
__device__ __forceinline__ void decode_phi(const DecodeArgs& a, long long idx, float* phi)
{
    const int px = (int)(idx % a.w0), py = (int)(idx / a.w0);
    float s[MAX_CH], c[MAX_CH];
    for (int i = 0; i < a.c0; i++)
        s[i] = a.v0[idx * a.c0 + i];
    const Tap t = tap_pixel(px, py, 0.0f, 0.0f, a.w0, a.h0, a.w1, a.h1);
    sample_bilinear(a.v1, a.w1, a.c1, t, c);
    build_phi(s, a.c0, c, a.c1, phi);
}

__device__ __forceinline__ float decode_row(const DecodeArgs& a, const float* phi, int r)
{
    float acc = a.bias[r];
    const float* row = a.weights + (size_t)r * a.nin;
    for (int i = 0; i < a.nin; i++)
        acc += row[i] * phi[i];
    return acc;
}

// The decoded byte of a pre-saturation output: saturated, then rounded half away from zero.
__device__ __forceinline__ uint8_t decode_byte(float acc)
{
    acc = acc < 0.0f ? 0.0f : (acc > 1.0f ? 1.0f : acc);
    return (uint8_t)(int)roundf(acc * 255.0f);
}

template <bool RESIDUAL>
__global__ __launch_bounds__(256) void k_decode(DecodeArgs a)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)a.w0 * a.h0)
        return;
    float phi[MAX_NIN];
    decode_phi(a, idx, phi);
    for (int r = 0; r < a.nout; r++)
    {
        const float acc = decode_row(a, phi, r);
        if (RESIDUAL)
            a.resid[idx * a.nout + r] = acc - a.src[idx * a.nout + r];
        else
            a.out[idx * a.nout + r] = decode_byte(acc);
    }
}

// Each texture's squared byte error of the decoded base plane against the source over the top-left width x height
// (image_psnr_texture, src/image.h:387-407): the source value saturated and rounded half away from zero to a byte, the
// difference squared, summed as integers - exact in any order, so the sums are the reference's double sums' values.
// One atomic per wavefront and texture.
// Integer sums, so any order gives the same total: each lane strides over the pixels, the workgroup's sums meet in
// shuffles and LDS, and a workgroup adds them into one of SSE_SLOTS counters per texture on separate cache lines (one
// counter per texture took tens of thousands of same-address atomics).
__global__ __launch_bounds__(256) void k_texture_sse(const uint8_t* rgb8, const float* src, int w0, int nout,
                                                     int width, int height, unsigned long long* sse)
{
    __shared__ unsigned long long s_part[4][NNTC_MAX_NOUT / 3];
    const int textures = nout / 3;
    const long long n = (long long)width * height;
    unsigned long long s[NNTC_MAX_NOUT / 3];
#pragma unroll
    for (int t = 0; t < NNTC_MAX_NOUT / 3; t++)
        s[t] = 0;
    for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long long)gridDim.x * blockDim.x)
    {
        const int x = (int)(idx % width), y = (int)(idx / width);
        const size_t base = ((size_t)y * w0 + x) * nout;
#pragma unroll
        for (int t = 0; t < NNTC_MAX_NOUT / 3; t++)
            if (t < textures)
                for (int ch = 0; ch < 3; ch++)
                {
                    const float f = src[base + 3 * t + ch];
                    const int a = (int)lroundf(fminf(1.0f, fmaxf(0.0f, f)) * 255.0f);
                    const int d = a - (int)rgb8[base + 3 * t + ch];
                    s[t] += (unsigned long long)(d * d);
                }
    }
#pragma unroll
    for (int t = 0; t < NNTC_MAX_NOUT / 3; t++)
    {
        for (int off = 32; off > 0; off >>= 1)
            s[t] += __shfl_xor(s[t], off);
        if (threadIdx.x % 64 == 0)
            s_part[threadIdx.x / 64][t] = s[t];
    }
    __syncthreads();
    if ((int)threadIdx.x < textures)
    {
        const int t = threadIdx.x;
        const unsigned long long sum = s_part[0][t] + s_part[1][t] + s_part[2][t] + s_part[3][t];
        if (sum)
            atomicAdd(sse + ((size_t)t * SSE_SLOTS + blockIdx.x % SSE_SLOTS) * SSE_STRIDE, sum);
    }
}

// ---- the allocation-free texture error sums (nntc_hip_texture_sse_noalloc)
//
// The decode and the error sums in one pass, so no decoded plane is stored: every lane decodes its pixels of the
// top-left width x height at their centres (decode_phi / decode_row / decode_byte: k_decode's arithmetic, so the bytes
// are decode_plane's) and adds each texture's squared byte error; a workgroup's sums go to SSE_SLOTS counters per
// texture in the file-scope scratch below, which a second kernel folds into the caller's output.

// The scratch: SSE_SLOTS counters, SSE_STRIDE apart, for each of at most NNTC_MAX_NOUT / 3 textures - 48 KiB, held on
// every device the library runs on until the process ends.
constexpr int NOALLOC_TEXTURES = NNTC_MAX_NOUT / 3;
constexpr size_t NOALLOC_SCRATCH = (size_t)NOALLOC_TEXTURES * SSE_SLOTS * SSE_STRIDE;
__device__ __align__(256) unsigned long long scratch[NOALLOC_SCRATCH];

__global__ __launch_bounds__(256) void k_zero_u64(unsigned long long* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0ull;
}

__global__ __launch_bounds__(256) void k_decode_sse(DecodeArgs a, int width, int height, unsigned long long* slots)
{
    __shared__ unsigned long long s_part[4][NOALLOC_TEXTURES];
    const int textures = a.nout / 3;
    const long long n = (long long)width * height;
    unsigned long long s[NOALLOC_TEXTURES];
#pragma unroll
    for (int t = 0; t < NOALLOC_TEXTURES; t++)
        s[t] = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    {
        const int x = (int)(i % width), y = (int)(i / width);
        const long long idx = (long long)y * a.w0 + x;
        float phi[MAX_NIN];
        decode_phi(a, idx, phi);
#pragma unroll
        for (int t = 0; t < NOALLOC_TEXTURES; t++)
            if (t < textures)
                for (int ch = 0; ch < 3; ch++)
                {
                    const int r = 3 * t + ch;
                    const float f = a.src[idx * a.nout + r];
                    const int want = (int)lroundf(fminf(1.0f, fmaxf(0.0f, f)) * 255.0f);
                    const int d = want - (int)decode_byte(decode_row(a, phi, r));
                    s[t] += (unsigned long long)(d * d);
                }
    }
#pragma unroll
    for (int t = 0; t < NOALLOC_TEXTURES; t++)
    {
        for (int off = 32; off > 0; off >>= 1)
            s[t] += __shfl_xor(s[t], off);
        if (threadIdx.x % 64 == 0)
            s_part[threadIdx.x / 64][t] = s[t];
    }
    __syncthreads();
    if ((int)threadIdx.x < textures)
    {
        const int t = threadIdx.x;
        const unsigned long long sum = s_part[0][t] + s_part[1][t] + s_part[2][t] + s_part[3][t];
        if (sum)
            atomicAdd(slots + ((size_t)t * SSE_SLOTS + blockIdx.x % SSE_SLOTS) * SSE_STRIDE, sum);
    }
}

__global__ void k_fold_sse(const unsigned long long* slots, int textures, unsigned long long* out)
{
    const int t = threadIdx.x;
    if (t >= textures)
        return;
    unsigned long long sum = 0;
    for (int k = 0; k < SSE_SLOTS; k++)
        sum += slots[((size_t)t * SSE_SLOTS + k) * SSE_STRIDE];
    out[t] = sum;
}

}   // namespace

void texture_sse_pass(Ctx& c, int width, int height, unsigned long long* sse_out)
{
    const PlaneGeom& g = c.g(0);
    decode_pass(c, 0, c.rgb8.p);
    const int textures = c.nout / 3;
    HIP_OK(hipMemsetAsync(c.tex_sse.p, 0, c.tex_sse.bytes(), c.stream));
    const unsigned grid = (unsigned)std::min<long long>(1024, blocks_for((long long)width * height, 256));
    hipLaunchKernelGGL(k_texture_sse, dim3(grid), dim3(256), 0, c.stream, c.rgb8.p, c.src.p, g.w0, c.nout, width,
                       height, c.tex_sse.p);
    HIP_OK(hipGetLastError());
    const unsigned long long* slots = read_back(c, (const unsigned long long*)c.tex_sse.p,
                                                (size_t)textures * SSE_SLOTS * SSE_STRIDE);
    for (int t = 0; t < textures; t++)
    {
        unsigned long long sum = 0;
        for (int s = 0; s < SSE_SLOTS; s++)
            sum += slots[((size_t)t * SSE_SLOTS + s) * SSE_STRIDE];
        sse_out[t] = sum;
    }
}

namespace
{

// Each plane's three sums from the reduction units' partials (unit u's at partial[3u .. 3u + 2]). Serially (the exact
// form): lanes 0-2 each add one sum's partials in unit order from zero (src/cpu/pool.h:314-324). Otherwise, in this
// kernel's fixed order: lane t adds units t, t + 256, ..., then a tree over the lanes.
__global__ __launch_bounds__(256) void k_objective_fold(const double* partial, Ranges units, int serial, double* out)
{
    const int p = blockIdx.x, t = threadIdx.x;
    const long long a = units.start[p], b = units.start[p + 1];
    if (serial)
    {
        if (t < 3)
        {
            double acc = 0.0;
            long long u = a;
            for (; u + 16 <= b; u += 16)
            {
                double v[16];
#pragma unroll
                for (int j = 0; j < 16; j++)
                    v[j] = partial[(u + j) * 3 + t];
#pragma unroll
                for (int j = 0; j < 16; j++)
                    acc += v[j];
            }
            for (; u < b; u++)
                acc += partial[u * 3 + t];
            out[p * 3 + t] = acc;
        }
        return;
    }
    __shared__ double s[3][256];
    double w = 0.0, ce = 0.0, fr = 0.0;
    for (long long u = a + t; u < b; u += 256)
    {
        w += partial[u * 3];
        ce += partial[u * 3 + 1];
        fr += partial[u * 3 + 2];
    }
    s[0][t] = w;
    s[1][t] = ce;
    s[2][t] = fr;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1)
    {
        if (t < h)
        {
            s[0][t] += s[0][t + h];
            s[1][t] += s[1][t + h];
            s[2][t] += s[2][t + h];
        }
        __syncthreads();
    }
    if (t < 3)
        out[p * 3 + t] = s[t][0];
}

}   // namespace

void objective_pass(Ctx& c, int k, double* sums_out)
{
    const bool exact = c.form == NNTC_FORM_EXACT;
    ObjArgs a;
    a.planes = c.planes;
    a.chunks.n = c.planes.n;
    // The reduction units: chunks of CHUNK sites (exact form) or OBJ_TW x OBJ_TH pixel tiles (fast form), per plane.
    long long total = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        const PlaneGeom& g = c.g(p);
        a.chunks.start[p] = total;
        if (exact)
            total += (g.n0 * (1 + k) + CHUNK - 1) / CHUNK;
        else
        {
            if (g.w1 > g.w0 || g.h1 > g.h0)
                throw Refused { "objective: the tiled form needs level 1 no larger than level 0" };
            total += objective_tiles(g);
        }
    }
    a.chunks.start[c.planes.n] = total;
    if ((size_t)total * 3 > c.partial.n)
        throw Refused { "objective: reduction partials exceed the workspace" };
    a.src = c.src.p;
    a.v0 = c.v0.p;
    a.v1 = c.v1.p;
    a.weights = c.weights.p;
    a.bias = c.bias.p;
    a.cw = c.cw.p;
    a.c0 = c.c0;
    a.c1 = c.c1;
    a.nin = c.nin;
    a.nout = c.nout;
    a.k = k;
    a.partial = c.partial.p;
    {
        // Algorithmic work of one pass: the three planes read once, and the per-site flops of the decode.
        const double n0 = (double)c.n0_total, n1 = (double)c.n1_total, s = n0 * (1 + k);
        const double bytes = 4.0 * n0 * (c.nout + c.c0) + 4.0 * n1 * c.c1;
        const double flops = s * 2.0 *
                             (4.0 * c.c1 + 4.0 * (c.c0 + c.nout) * k / (k + 1.0) + c.c0 * c.c1 + (double)c.nout * c.nin);
        KernelTimer kt(c.stream, exact ? "k_objective" : "k_objective_tiles", bytes, flops);
        if (exact)
            hipLaunchKernelGGL(k_objective, dim3((unsigned)total), dim3(OBJ_WAVE), 0, c.stream, a);
        else
            launch_objective_tiles(c.c0, c.c1, (unsigned)total, c.stream, a);
    }
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(k_objective_fold, dim3((unsigned)c.planes.n), dim3(256), 0, c.stream, (const double*)c.partial.p,
                       a.chunks, exact ? 1 : 0, c.obj.p);
    HIP_OK(hipGetLastError());
    const double* sums = read_back(c, (const double*)c.obj.p, (size_t)c.planes.n * 3);
    for (int i = 0; i < c.planes.n * 3; i++)
        sums_out[i] = sums[i];
}

static DecodeArgs decode_args(Ctx& c, int plane)
{
    const PlaneGeom& g = c.g(plane);
    DecodeArgs a;
    a.v0 = c.v0.p + g.off0 * c.c0;
    a.v1 = c.v1.p + g.off1 * c.c1;
    a.src = c.src.p + g.off0 * c.nout;
    a.weights = c.weights.p;
    a.bias = c.bias.p;
    a.w0 = g.w0;
    a.h0 = g.h0;
    a.w1 = g.w1;
    a.h1 = g.h1;
    a.c0 = c.c0;
    a.c1 = c.c1;
    a.nin = c.nin;
    a.nout = c.nout;
    a.out = nullptr;
    a.resid = nullptr;
    return a;
}

// The scratch's device address on each device, taken once per device when a context is first created on it (with that
// device current). The array's name is a host-side stand-in; kernels take this address.
namespace
{
constexpr int NOALLOC_DEVICES = 64;
std::mutex g_noalloc_lock;
unsigned long long* g_noalloc_scratch[NOALLOC_DEVICES] = {};
}   // namespace

void noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_DEVICES)
        throw Refused { "texture_sse_noalloc: device index beyond the scratch table" };
    std::lock_guard<std::mutex> hold(g_noalloc_lock);
    if (g_noalloc_scratch[device])
        return;
    void* ptr = nullptr;
    HIP_OK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_noalloc_scratch[device] = static_cast<unsigned long long*>(ptr);
}

void texture_sse_noalloc_pass(Ctx& c, int width, int height, unsigned long long* sse_device, hipStream_t stream)
{
    const int textures = c.nout / 3;
    if (textures > NOALLOC_TEXTURES || (size_t)textures * SSE_SLOTS * SSE_STRIDE > NOALLOC_SCRATCH)
        throw Refused { "texture_sse_noalloc: more textures than the scratch holds (6)" };
    if (width < 0 || height < 0 || width > c.g(0).w0 || height > c.g(0).h0)
        throw Refused { "texture_sse_noalloc: the extent exceeds the base plane" };
    if (!sse_device)
        throw Refused { "texture_sse_noalloc: no output" };
    unsigned long long* slots = nullptr;
    {
        std::lock_guard<std::mutex> hold(g_noalloc_lock);
        slots = c.device >= 0 && c.device < NOALLOC_DEVICES ? g_noalloc_scratch[c.device] : nullptr;
    }
    if (!slots)
        throw Refused { "texture_sse_noalloc: the scratch of this context's device was not set up" };
    const int n = (int)((size_t)textures * SSE_SLOTS * SSE_STRIDE);
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)blocks_for(n, 256)), dim3(256), 0, stream, slots, n);
    DecodeArgs a = decode_args(c, 0);
    const unsigned grid = (unsigned)std::max<long long>(1, std::min<long long>(1024, blocks_for((long long)width * height,
                                                                                                 256)));
    hipLaunchKernelGGL(k_decode_sse, dim3(grid), dim3(256), 0, stream, a, width, height, slots);
    hipLaunchKernelGGL(k_fold_sse, dim3(1), dim3(64), 0, stream, (const unsigned long long*)slots, textures,
                       sse_device);
    HIP_OK(hipGetLastError());
}

void decode_pass(Ctx& c, int plane, uint8_t* out)
{
    DecodeArgs a = decode_args(c, plane);
    a.out = out;
    hipLaunchKernelGGL(k_decode<false>, dim3(blocks_for(c.g(plane).n0, 256)), dim3(256), 0, c.stream, a);
    HIP_OK(hipGetLastError());
}

void residual_pass(Ctx& c, int plane, float* resid)
{
    DecodeArgs a = decode_args(c, plane);
    a.resid = resid;
    hipLaunchKernelGGL(k_decode<true>, dim3(blocks_for(c.g(plane).n0, 256)), dim3(256), 0, c.stream, a);
    HIP_OK(hipGetLastError());
}

}   // namespace nntc
