// objective.hip: the objective E, the decode of a stored level to bytes, and the centre residual, on gfx950.
//
// E (src/objective.cu:123-279, src/cpu/cpu_objective.cpp:112-291) is the cw-weighted squared error of the decoded
// material over every site of every plane - the centre and K fractional sites per pixel - in double, plus the
// unweighted centre and fractional sums. One launch covers every plane: the site space of the chain is cut into
// chunks of CHUNK consecutive sites of one plane, the reference CPU backend's own reduction unit.
//
// Exact form: each chunk is one wavefront. Its 64 lanes evaluate 64 sites at a time in parallel (sampling, phi, the
// 3T x nin decode in double) and stage every site's per-output terms in LDS; then lanes 0, 1, 2 fold the weighted,
// centre and fractional terms serially in site-then-output order - the same additions in the same order as the
// reference's chunk loop, so each chunk's three sums are its bits. The host folds the chunk sums in chunk order
// (src/cpu/pool.h:314-324).
// Fast form: one 256-lane workgroup per chunk, each lane summing its own sites, then trees; last bits differ.
//
// Roofline: per pass the algorithmic bytes are the three planes read once, 4N(nout + c0) + 4 N1 c1; the flops are
// S 2 (4 c1 + 4 (c0 + nout) K / (K + 1) + c0 c1 + nout nin) for S = N (1 + K) sites, about 41 flop/byte at nout 12:
// compute-side of the fp64 ridge, so this kernel's fraction is reported against fp64 peak.

#include "arm.h"

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
    Ranges chunks;            // chunk index space: plane p owns chunks [start[p], start[p+1])
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

// The fast form: one 256-lane workgroup per chunk, each lane summing its own sites (16 of them, strided so a wavefront
// reads consecutive sites), then the wavefront trees and the four wavefronts' partials added in order. Sixteen
// wavefronts per SIMD over a launch instead of three, and no serial fold.
constexpr int OBJ_FAST_THREADS = 256;

__global__ __launch_bounds__(OBJ_FAST_THREADS) void k_objective_fast(ObjArgs a)
{
    __shared__ double part[OBJ_FAST_THREADS / 64][3];
    const long long chunk = blockIdx.x;
    const int p = range_plane(a.chunks, chunk);
    const PlaneGeom& g = a.planes.p[p];
    const long long sites = g.n0 * (long long)(1 + a.k);
    const long long begin = (chunk - a.chunks.start[p]) * (long long)CHUNK;
    const long long end = begin + CHUNK < sites ? begin + CHUNK : sites;
    const float* src = a.src + g.off0 * a.nout;
    const float* v0 = a.v0 + g.off0 * a.c0;
    const float* v1 = a.v1 + g.off1 * a.c1;
    const int nout = a.nout;
    double s[3] = { 0.0, 0.0, 0.0 };
    for (long long si = begin + threadIdx.x; si < end; si += OBJ_FAST_THREADS)
    {
        double e[MAX_NOUT];
        int j;
        objective_site(v0, v1, src, g.w0, g.h0, g.w1, g.h1, a.c0, a.c1, a.nin, nout, a.weights, a.bias, a.k, si, e, j);
        double we = 0.0, ee = 0.0;
#pragma unroll
        for (int r = 0; r < MAX_NOUT; r++)
            if (r < nout)
            {
                we += (double)a.cw[r] * e[r] * e[r];
                ee += e[r] * e[r];
            }
        s[0] += we;
        s[j == 0 ? 1 : 2] += ee;
    }
#pragma unroll
    for (int q = 0; q < 3; q++)
        for (int off = 32; off > 0; off >>= 1)
            s[q] += __shfl_xor(s[q], off);
    const int wave = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0)
        for (int q = 0; q < 3; q++)
            part[wave][q] = s[q];
    __syncthreads();
    if (threadIdx.x < 3)
    {
        double t = 0.0;
        for (int w = 0; w < OBJ_FAST_THREADS / 64; w++)
            t += part[w][threadIdx.x];
        a.partial[chunk * 3 + threadIdx.x] = t;
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
template <bool RESIDUAL>
__global__ __launch_bounds__(256) void k_decode(DecodeArgs a)
{
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)a.w0 * a.h0)
        return;
    const int px = (int)(idx % a.w0), py = (int)(idx / a.w0);
    float s[MAX_CH], c[MAX_CH], phi[MAX_NIN];
    for (int i = 0; i < a.c0; i++)
        s[i] = a.v0[idx * a.c0 + i];
    const Tap t = tap_pixel(px, py, 0.0f, 0.0f, a.w0, a.h0, a.w1, a.h1);
    sample_bilinear(a.v1, a.w1, a.c1, t, c);
    build_phi(s, a.c0, c, a.c1, phi);
    for (int r = 0; r < a.nout; r++)
    {
        float acc = a.bias[r];
        const float* row = a.weights + (size_t)r * a.nin;
        for (int i = 0; i < a.nin; i++)
            acc += row[i] * phi[i];
        if (RESIDUAL)
            a.resid[idx * a.nout + r] = acc - a.src[idx * a.nout + r];
        else
        {
            acc = acc < 0.0f ? 0.0f : (acc > 1.0f ? 1.0f : acc);
            a.out[idx * a.nout + r] = (uint8_t)(int)roundf(acc * 255.0f);
        }
    }
}

}   // namespace

void objective_pass(Ctx& c, int k, double* sums_out)
{
    ObjArgs a;
    a.planes = c.planes;
    a.chunks.n = c.planes.n;
    long long total = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        a.chunks.start[p] = total;
        total += (c.g(p).n0 * (1 + k) + CHUNK - 1) / CHUNK;
    }
    a.chunks.start[c.planes.n] = total;
    if ((size_t)total * 3 > c.partial.n)
        throw Refused { "objective: chunk partials exceed the workspace" };
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
        KernelTimer kt(c.stream, "k_objective", bytes, flops);
        if (c.form == NNTC_FORM_EXACT)
            hipLaunchKernelGGL(k_objective, dim3((unsigned)total), dim3(OBJ_WAVE), 0, c.stream, a);
        else
            hipLaunchKernelGGL(k_objective_fast, dim3((unsigned)total), dim3(OBJ_FAST_THREADS), 0, c.stream, a);
    }
    HIP_OK(hipGetLastError());
    std::vector<double> part((size_t)total * 3);
    HIP_OK(hipMemcpyAsync(part.data(), c.partial.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    // The chunk sums folded in chunk order, from zero (src/cpu/pool.h:314-324).
    for (int p = 0; p < c.planes.n; p++)
    {
        double w = 0.0, ce = 0.0, fr = 0.0;
        for (long long ch = a.chunks.start[p]; ch < a.chunks.start[p + 1]; ch++)
        {
            w += part[(size_t)ch * 3];
            ce += part[(size_t)ch * 3 + 1];
            fr += part[(size_t)ch * 3 + 2];
        }
        sums_out[3 * p] = w;
        sums_out[3 * p + 1] = ce;
        sums_out[3 * p + 2] = fr;
    }
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
