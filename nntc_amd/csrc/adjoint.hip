// adjoint.hip: the decode at every sample site of one plane and its hand-written adjoint, on gfx950.
//
// The forward (k_decode_sites) is the objective's model (src/sample.cuh:43-178, src/objective.cu:104-118 in the fp32
// addressing of the assembly): at site j of pixel (px, py) the level-0 sample s0 (the pixel's own texel at the centre
// site, else the bilinear blend of its four clamped taps), the level-1 sample s1, phi = [s1][s0][s0 s1] and
// out = W phi + b. Sites are numbered s = (py w0 + px)(1 + K) + j, as the assembly numbers them.
//
// The adjoint (k_adjoint_v0 / k_adjoint_v1 / k_adjoint_w) takes a cotangent g (one per site and output) and returns
// every input's gradient of sum_s,r g_sr out_sr:
//   dW_rf = sum_s g_sr phi_sf,   db_r = sum_s g_sr;
//   ds0 = sum_r g_r (d out_r / d s0), ds1 likewise (adjoint.h's rows);
//   dv0_t = sum over the sites whose level-0 taps reach texel t of their weight there times ds0 (the transpose of the
//   bilinear sampling), dv1 likewise over level 1's taps.
// Each texel gathers its own sites (its 3 x 3 pixels for level 0, footprint_range's pixels for level 1) in pixel-row,
// pixel, site order, so the result does not depend on scheduling; dW and db are summed per chunk of 128 sites and the
// chunks folded in order. The kernels are instantiated in fp32 and fp64: the fp64 form is what a gradient check compares
// against finite differences; the fp32 form's level-0 half is, with the decoder's rows cached per site, the matrix-free
// operator of the level-0 conjugate gradients (latent.hip).

#include "arm.h"
#include "adjoint.h"

#include <functional>
#include <vector>

namespace nntc
{

int guarded_call(const std::function<void()>& f);

namespace
{

struct SitesArgs
{
    int w0, h0, w1, h1, nout, k, nin;
    const void* v0;
    const void* v1;
    const void* w;
    const void* b;
    const void* g;
    void* out;
    void* dv0;
    void* dv1;
    void* partial;     // dW / db partials: nout (nin + 1) per chunk of ADJ_CHUNK sites
    void* dw;
    void* db;
    long long sites;
};

constexpr int ADJ_CHUNK = 128;

// Site s's two samples in precision T.
template <class T, int C0, int C1>
__device__ __forceinline__ void site_samples(const SitesArgs& a, int px, int py, int j, T* s0, T* s1)
{
    const T* v0 = static_cast<const T*>(a.v0);
    const T* v1 = static_cast<const T*>(a.v1);
    const Tap t1 = site_taps(a.k, j, px, py, a.w0, a.h0, a.w1, a.h1, false);
    blend_t<T, C1>(v1, a.w1, t1, s1);
    const Tap t0 = site_taps(a.k, j, px, py, a.w0, a.h0, a.w0, a.h0, true);
    blend_t<T, C0>(v0, a.w0, t0, s0);
}

template <class T, int C0, int C1>
__device__ __forceinline__ void site_phi(const T* s0, const T* s1, T* phi)
{
    int n = 0;
#pragma unroll
    for (int q = 0; q < C1; q++)
        phi[n++] = s1[q];
#pragma unroll
    for (int i = 0; i < C0; i++)
        phi[n++] = s0[i];
#pragma unroll
    for (int i = 0; i < C0; i++)
#pragma unroll
        for (int q = 0; q < C1; q++)
            phi[n++] = s0[i] * s1[q];
}

template <class T, int C0, int C1>
__global__ __launch_bounds__(256) void k_decode_sites(SitesArgs a)
{
    constexpr int NIN = C1 + C0 + C0 * C1;
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.sites)
        return;
    const int per = 1 + a.k, j = (int)(s % per);
    const long long pix = s / per;
    const int px = (int)(pix % a.w0), py = (int)(pix / a.w0);
    T s0[C0], s1[C1], phi[NIN];
    site_samples<T, C0, C1>(a, px, py, j, s0, s1);
    site_phi<T, C0, C1>(s0, s1, phi);
    const T* w = static_cast<const T*>(a.w);
    const T* b = static_cast<const T*>(a.b);
    T* out = static_cast<T*>(a.out);
    for (int r = 0; r < a.nout; r++)
    {
        T acc = b[r];
#pragma unroll
        for (int f = 0; f < NIN; f++)
            acc += w[r * NIN + f] * phi[f];
        out[s * a.nout + r] = acc;
    }
}

// dv0 at level-0 texel (tx, ty): its 3 x 3 pixels' sites, each weighted by its taps' weight at the texel.
template <class T, int C0, int C1>
__global__ __launch_bounds__(256) void k_adjoint_v0(SitesArgs a)
{
    constexpr int NIN = C1 + C0 + C0 * C1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)a.w0 * a.h0)
        return;
    const int tx = (int)(t % a.w0), ty = (int)(t / a.w0);
    const T* w = static_cast<const T*>(a.w);
    const T* g = static_cast<const T*>(a.g);
    const T* v1 = static_cast<const T*>(a.v1);
    const int per = 1 + a.k;
    T acc[C0];
#pragma unroll
    for (int i = 0; i < C0; i++)
        acc[i] = (T)0;
    int xl, xh, yl, yh;
    footprint_range(tx, a.w0, a.w0, xl, xh);
    footprint_range(ty, a.h0, a.h0, yl, yh);
    for (int py = yl; py <= yh; py++)
        for (int px = xl; px <= xh; px++)
            for (int j = 0; j < per; j++)
            {
                const Tap t0 = site_taps(a.k, j, px, py, a.w0, a.h0, a.w0, a.h0, true);
                const float wt = tap_weight(tx, t0.x0, t0.x1, t0.fx) * tap_weight(ty, t0.y0, t0.y1, t0.fy);
                if (wt == 0.0f)
                    continue;
                T s1[C1];
                const Tap t1 = site_taps(a.k, j, px, py, a.w0, a.h0, a.w1, a.h1, false);
                blend_t<T, C1>(v1, a.w1, t1, s1);
                const long long s = ((long long)py * a.w0 + px) * per + j;
                T ds0[C0];
#pragma unroll
                for (int i = 0; i < C0; i++)
                    ds0[i] = (T)0;
                for (int r = 0; r < a.nout; r++)
                {
                    T m[C0];
                    level0_row_t<T, C0, C1>(w, NIN, r, s1, m);
                    const T gr = g[s * a.nout + r];
#pragma unroll
                    for (int i = 0; i < C0; i++)
                        ds0[i] += gr * m[i];
                }
#pragma unroll
                for (int i = 0; i < C0; i++)
                    acc[i] += (T)wt * ds0[i];
            }
    T* dv0 = static_cast<T*>(a.dv0);
#pragma unroll
    for (int i = 0; i < C0; i++)
        dv0[t * C0 + i] = acc[i];
}

// dv1 at level-1 texel (tx, ty): the sites of the level-0 pixels whose level-1 taps can reach it.
template <class T, int C0, int C1>
__global__ __launch_bounds__(256) void k_adjoint_v1(SitesArgs a)
{
    constexpr int NIN = C1 + C0 + C0 * C1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)a.w1 * a.h1)
        return;
    const int tx = (int)(t % a.w1), ty = (int)(t / a.w1);
    const T* w = static_cast<const T*>(a.w);
    const T* g = static_cast<const T*>(a.g);
    const T* v0 = static_cast<const T*>(a.v0);
    const int per = 1 + a.k;
    T acc[C1];
#pragma unroll
    for (int q = 0; q < C1; q++)
        acc[q] = (T)0;
    int xl, xh, yl, yh;
    footprint_range(tx, a.w0, a.w1, xl, xh);
    footprint_range(ty, a.h0, a.h1, yl, yh);
    for (int py = yl; py <= yh; py++)
        for (int px = xl; px <= xh; px++)
            for (int j = 0; j < per; j++)
            {
                const Tap t1 = site_taps(a.k, j, px, py, a.w0, a.h0, a.w1, a.h1, false);
                const float wt = tap_weight(tx, t1.x0, t1.x1, t1.fx) * tap_weight(ty, t1.y0, t1.y1, t1.fy);
                if (wt == 0.0f)
                    continue;
                T s0[C0];
                const Tap t0 = site_taps(a.k, j, px, py, a.w0, a.h0, a.w0, a.h0, true);
                blend_t<T, C0>(v0, a.w0, t0, s0);
                const long long s = ((long long)py * a.w0 + px) * per + j;
                T ds1[C1];
#pragma unroll
                for (int q = 0; q < C1; q++)
                    ds1[q] = (T)0;
                for (int r = 0; r < a.nout; r++)
                {
                    T gq[C1];
                    level1_row_t<T, C0, C1>(w, NIN, r, s0, gq);
                    const T gr = g[s * a.nout + r];
#pragma unroll
                    for (int q = 0; q < C1; q++)
                        ds1[q] += gr * gq[q];
                }
#pragma unroll
                for (int q = 0; q < C1; q++)
                    acc[q] += (T)wt * ds1[q];
            }
    T* dv1 = static_cast<T*>(a.dv1);
#pragma unroll
    for (int q = 0; q < C1; q++)
        dv1[t * C1 + q] = acc[q];
}

// dW and db per chunk of ADJ_CHUNK sites: each lane builds one site's features into LDS, then lane e adds entry e
// (output r, feature f; f = nin is the bias) over the chunk's sites in order.
template <class T, int C0, int C1>
__global__ __launch_bounds__(ADJ_CHUNK) void k_adjoint_w(SitesArgs a)
{
    constexpr int NIN = C1 + C0 + C0 * C1;
    __shared__ T s_phi[ADJ_CHUNK][NIN];
    __shared__ T s_g[ADJ_CHUNK][NNTC_MAX_NOUT];
    const long long s0i = (long long)blockIdx.x * ADJ_CHUNK;
    const long long s = s0i + threadIdx.x;
    const T* g = static_cast<const T*>(a.g);
    if (s < a.sites)
    {
        const int per = 1 + a.k, j = (int)(s % per);
        const long long pix = s / per;
        const int px = (int)(pix % a.w0), py = (int)(pix / a.w0);
        T sa[C0], sb[C1];
        site_samples<T, C0, C1>(a, px, py, j, sa, sb);
        site_phi<T, C0, C1>(sa, sb, s_phi[threadIdx.x]);
        for (int r = 0; r < a.nout; r++)
            s_g[threadIdx.x][r] = g[s * a.nout + r];
    }
    __syncthreads();
    const int entries = a.nout * (NIN + 1);
    const long long n = a.sites - s0i < ADJ_CHUNK ? a.sites - s0i : ADJ_CHUNK;
    T* part = static_cast<T*>(a.partial) + (size_t)blockIdx.x * entries;
    for (int e = threadIdx.x; e < entries; e += ADJ_CHUNK)
    {
        const int r = e / (NIN + 1), f = e % (NIN + 1);
        T acc = (T)0;
        for (int i = 0; i < n; i++)
            acc += s_g[i][r] * (f < NIN ? s_phi[i][f] : (T)1);
        part[e] = acc;
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_adjoint_w_fold(SitesArgs a, long long chunks)
{
    const int nin = a.nin, entries = a.nout * (nin + 1);
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= entries)
        return;
    const T* part = static_cast<const T*>(a.partial);
    T acc = (T)0;
    for (long long c = 0; c < chunks; c++)
        acc += part[(size_t)c * entries + e];
    const int r = e / (nin + 1), f = e % (nin + 1);
    if (f < nin)
        static_cast<T*>(a.dw)[r * nin + f] = acc;
    else
        static_cast<T*>(a.db)[r] = acc;
}

template <class T, int C0, int C1>
void run_sites(SitesArgs& a, bool adjoint, hipStream_t st)
{
    const unsigned gs = blocks_for(a.sites, 256);
    if (!adjoint)
    {
        hipLaunchKernelGGL((k_decode_sites<T, C0, C1>), dim3(gs), dim3(256), 0, st, a);
        HIP_OK(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL((k_adjoint_v0<T, C0, C1>), dim3(blocks_for((long long)a.w0 * a.h0, 256)), dim3(256), 0, st, a);
    hipLaunchKernelGGL((k_adjoint_v1<T, C0, C1>), dim3(blocks_for((long long)a.w1 * a.h1, 256)), dim3(256), 0, st, a);
    const long long chunks = (a.sites + ADJ_CHUNK - 1) / ADJ_CHUNK;
    hipLaunchKernelGGL((k_adjoint_w<T, C0, C1>), dim3((unsigned)chunks), dim3(ADJ_CHUNK), 0, st, a);
    hipLaunchKernelGGL((k_adjoint_w_fold<T>), dim3(blocks_for((long long)a.nout * (a.nin + 1), 256)), dim3(256), 0, st,
                       a, chunks);
    HIP_OK(hipGetLastError());
}

template <class T, int C0>
void run_sites_c1(int c1, SitesArgs& a, bool adjoint, hipStream_t st)
{
    switch (c1)
    {
    case 1: run_sites<T, C0, 1>(a, adjoint, st); break;
    case 2: run_sites<T, C0, 2>(a, adjoint, st); break;
    case 3: run_sites<T, C0, 3>(a, adjoint, st); break;
    default: run_sites<T, C0, 4>(a, adjoint, st); break;
    }
}

template <class T>
void run_sites_t(int c0, int c1, SitesArgs& a, bool adjoint, hipStream_t st)
{
    switch (c0)
    {
    case 1: run_sites_c1<T, 1>(c1, a, adjoint, st); break;
    case 2: run_sites_c1<T, 2>(c1, a, adjoint, st); break;
    case 3: run_sites_c1<T, 3>(c1, a, adjoint, st); break;
    default: run_sites_c1<T, 4>(c1, a, adjoint, st); break;
    }
}

SitesArgs sites_args(const nntc_sites_desc* d)
{
    if (!d || d->w0 < 1 || d->h0 < 1 || d->w1 < 1 || d->h1 < 1 || d->w1 > d->w0 || d->h1 > d->h0 || d->c0 < 1 ||
        d->c0 > MAX_CH || d->c1 < 1 || d->c1 > MAX_CH || d->nout < 1 || d->nout > MAX_NOUT || d->k < 0 || d->k > 4)
        throw Refused { "decode sites: descriptor out of range" };
    if ((long long)d->w0 * d->h0 * (1 + d->k) * d->nout >= (1ll << 31))
        throw Refused { "decode sites: the plane's outputs must number fewer than 2^31" };
    SitesArgs a {};
    a.w0 = d->w0;
    a.h0 = d->h0;
    a.w1 = d->w1;
    a.h1 = d->h1;
    a.nout = d->nout;
    a.k = d->k;
    a.nin = d->c1 + d->c0 + d->c0 * d->c1;
    a.sites = (long long)d->w0 * d->h0 * (1 + d->k);
    return a;
}

}   // namespace

}   // namespace nntc

using namespace nntc;

int nntc_hip_decode_sites(const nntc_sites_desc* d, int fp64, const void* v0, const void* v1, const void* w,
                          const void* b, void* out, void* stream)
{
    return guarded_call([&] {
        SitesArgs a = sites_args(d);
        if (!v0 || !v1 || !w || !b || !out)
            throw Refused { "decode_sites: a null array" };
        a.v0 = v0;
        a.v1 = v1;
        a.w = w;
        a.b = b;
        a.out = out;
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (fp64)
            run_sites_t<double>(d->c0, d->c1, a, false, st);
        else
            run_sites_t<float>(d->c0, d->c1, a, false, st);
        HIP_OK(hipStreamSynchronize(st));
    });
}

int nntc_hip_decode_adjoint(const nntc_sites_desc* d, int fp64, const void* v0, const void* v1, const void* w,
                            const void* g, void* dv0, void* dv1, void* dw, void* db, void* stream)
{
    return guarded_call([&] {
        SitesArgs a = sites_args(d);
        if (!v0 || !v1 || !w || !g || !dv0 || !dv1 || !dw || !db)
            throw Refused { "decode_adjoint: a null array" };
        a.v0 = v0;
        a.v1 = v1;
        a.w = w;
        a.g = g;
        a.dv0 = dv0;
        a.dv1 = dv1;
        a.dw = dw;
        a.db = db;
        const long long chunks = (a.sites + ADJ_CHUNK - 1) / ADJ_CHUNK;
        const size_t bytes = (size_t)chunks * a.nout * (a.nin + 1) * (fp64 ? sizeof(double) : sizeof(float));
        DevBuf<unsigned char> partial;
        partial.alloc(bytes);
        a.partial = partial.p;
        hipStream_t st = static_cast<hipStream_t>(stream);
        if (fp64)
            run_sites_t<double>(d->c0, d->c1, a, true, st);
        else
            run_sites_t<float>(d->c0, d->c1, a, true, st);
        HIP_OK(hipStreamSynchronize(st));
    });
}
