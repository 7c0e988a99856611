// latent.hip: blocks (b) and (c') - one latent plane as one sparse weighted least squares per stored plane, on gfx950
// (src/solve_level1.cu, src/cpu/cpu_solve_level1.cpp).
//
// With the other latent and the decoder held, E is an exact quadratic in the unknown plane. Its 9-point block stencil
// is assembled by a GATHER, one texel per thread, from monomial sums of the other latent's sample expanded against
// fixed C x C matrices built once per call on the host (build_form_matrices). The proximal step
// (H + lambda I) delta = -grad, lambda = ridge * mean diag H, is solved by conjugate gradients preconditioned by the
// block Jacobi of the stencil, every plane of the chain in lockstep with a plane dropping out at the probe where its
// relative residual is under the bar. Once the grids are frozen the same stencil drives four-colour Gauss-Seidel sweeps
// whose per-value step is the exact argmin over the grid.
//
// Every per-texel routine here - the gather, the block inverse, the mat-vec (neighbours tested, not clamped), the
// sweep's channel-by-channel argmin - is the reference's arithmetic in the reference's order, one texel per thread.
// The reductions are the fixed-order folds of the reference CPU backend in the exact form: one wavefront per chunk of
// CHUNK texels folds the chunk's values serially (lane order through readlane), and one thread per plane folds the
// chunk sums in chunk order. The fast form folds each chunk with a wavefront tree.
//
// Roofline (HBM): assembly writes 8 (5 c^2 + c) bytes per texel and reads the three planes; a CG iteration streams
// the stencil (40 c^2 B per texel), the preconditioner (4 c^2) and about six fp64 vectors (48 c); a sweep colour reads
// the stencil of its texels and their neighbours' corrections.

#include "arm.h"
#include "adjoint.h"
#include "tile_stage.h"

#include <cmath>
#include <functional>
#include <type_traits>

namespace nntc
{

namespace
{

constexpr int CG_PROBE = 4;
constexpr int CG_MAX_ITERATIONS = 200;
constexpr double CG_TOLERANCE = 1e-10;        // the exact form (the reference's bar)
constexpr double CG_TOLERANCE_FP32 = 1e-6;     // the plain and the fast form: fp32 vectors

// ---------------------------------------------------------------------------------------------------------------------
// Host: the fixed matrices of the quadratic form (src/cpu/cpu_solve_level1.cpp:103-187)
// ---------------------------------------------------------------------------------------------------------------------

// X^T C X = P_0 + sum_i y_i P_i + sum_{i <= j} y_i y_j P_ij for X(y) = X0 + sum_i y_i Y_i.
void build_form_matrices(const Ctx& c, const std::vector<double>& x0, const std::vector<double>& y, int terms, int dim,
                         std::vector<double>& p, const std::vector<float>& cw)
{
    const int nout = c.nout;
    const int nmono = 1 + terms + terms * (terms + 1) / 2;
    p.assign((size_t)nmono * dim * dim, 0.0);
    auto product = [&](const double* x, const double* z, double* out) {
        for (int u = 0; u < dim; u++)
            for (int v = 0; v < dim; v++)
            {
                double s = 0.0;
                for (int r = 0; r < nout; r++)
                    s += (double)cw[(size_t)r] * x[(size_t)r * dim + u] * z[(size_t)r * dim + v];
                out[u * dim + v] += s;
            }
    };
    int slot = 0;
    product(x0.data(), x0.data(), p.data());
    slot++;
    for (int i = 0; i < terms; i++, slot++)
    {
        const double* y_i = y.data() + (size_t)i * nout * dim;
        product(y_i, x0.data(), p.data() + (size_t)slot * dim * dim);
        product(x0.data(), y_i, p.data() + (size_t)slot * dim * dim);
    }
    for (int i = 0; i < terms; i++)
        for (int j = i; j < terms; j++, slot++)
        {
            const double* y_i = y.data() + (size_t)i * nout * dim;
            const double* y_j = y.data() + (size_t)j * nout * dim;
            product(y_i, y_j, p.data() + (size_t)slot * dim * dim);
            if (i != j)
                product(y_j, y_i, p.data() + (size_t)slot * dim * dim);
        }
}

// ---------------------------------------------------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------------------------------------------------

__device__ __forceinline__ double lane_value(double v, int l)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

// acc plus the values of lanes 0 .. count-1 in lane order: a serial fold that every lane computes identically.
__device__ __forceinline__ double fold_lanes(double acc, double v, int count)
{
    if (count == 64)
    {
#pragma unroll
        for (int l = 0; l < 64; l++)
            acc += lane_value(v, l);
    }
    else
        for (int l = 0; l < count; l++)
            acc += lane_value(v, l);
    return acc;
}

// The fast form's wavefront tree: every lane ends holding the sum of all 64.
__device__ __forceinline__ double tree_lanes(double v)
{
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off);
    return v;
}

struct LatentView
{
    Planes planes;
    Ranges texels;        // per plane, this latent's texels
    bool level0;
    int c;                // channels
    int bits;
    int levels;
    double* stencil;      // entry e of the texel at position s: stencil[e * sm.sn + s]
    float* stencil32;     // the fast form's fp32 copy, same layout
    StencilMap sm;
    double* grad;
    float* precond;
    double* x;
    double* r;
    double* z;
    double* p;
    double* ap;
    float* prev;
    float* v;             // the latent's plane values (chain-concatenated)
    uint8_t* k;           // its indices
    const float* lo;      // grid lo/hi of this latent, 4 each
    const float* hi;
    double* scalars;      // nplanes * SC_COUNT
    int pw[MAX_PLANES];   // per plane latent width
    int ph[MAX_PLANES];   // per plane latent height
};

__device__ __forceinline__ long long view_offset(const LatentView& L, int p)
{
    return L.texels.start[p];
}

// ---------------------------------------------------------------------------------------------------------------------
// Assembly (src/cpu/cpu_solve_level1.cpp:203-468)
// ---------------------------------------------------------------------------------------------------------------------

// The assembly in two passes. A site's contribution to a texel's sums is w_self * tw_e * mono(site) into the owned
// slots and w_self * acc(site) into the gradient, where mono is the monomials of the OTHER latent's sample at the site
// and acc = sum_r cw_r G_ra res_r is the site's residual pulled back through the decoder - neither depends on the
// texel. So:
//   1. k_site_terms: one thread per site computes the other latent's fp32 sample (from which the gather rebuilds the
//      monomials exactly: a product of two floats is exact in double) and the fp64 acc vector, once, into the site
//      buffer;
//   2. k_gather: one thread per texel walks its footprint's sites in the reference's order (pixel row, pixel, site),
//      recomputes each site's taps (fp32 addressing only) and adds the stored terms. The sums and the gradient receive
//      the same operands in the same order as the reference's per-texel loop, so the stencil is its bits. The kernel
//      is instantiated per (channels, sample channels), so its five-slot accumulators live in registers.
// The reference's gather evaluated every site about four times over (each site reaches four level-1 texels) and
// carried its accumulators in scratch; this form reads each site's 8 c_other + 8 c bytes instead.

struct AssembleArgs
{
    LatentView L;
    const float* src;
    const float* v0;
    const float* v1;
    const float* weights;
    const float* bias;
    const float* cw;
    const double* mono_p;
    float* site_f;          // per site: the other latent's sample (cs floats)
    double* site_d;         // per site: acc (c doubles)
    float* site_d32;        // the fast form's acc (c floats, fp32 sums), over the same buffer
    const float* site_target;   // the fast form's per-site source sample (nout floats), or null
    Ranges sites;           // per plane: its sites, w0 h0 (1 + K)
    int nmono, c0, c1, nin, nout, k;
    Div31 by_sites;                 // division by the sites per pixel, 1 + K
    Div31 by_w0[MAX_PLANES];        // division by each plane's width
    const double* tab;              // the fast gathers' slot weights of an interior texel (k_gather_table)
};

// (footprint_range: adjoint.h)

// The monomials of the form in the one order the matrices and the sums use: 1, s_i, then s_i s_j (i <= j).
template <int CS>
__device__ __forceinline__ void monomials(const float* s, double* mono)
{
    int n = 0;
    mono[n++] = 1.0;
#pragma unroll
    for (int i = 0; i < CS; i++)
        mono[n++] = (double)s[i];
#pragma unroll
    for (int i = 0; i < CS; i++)
#pragma unroll
        for (int j = i; j < CS; j++)
            mono[n++] = (double)s[i] * (double)s[j];
}

// One site's terms for the level-1 (LEVEL0 false) or the level-0 (LEVEL0 true) assembly, instantiated per channel
// counts so every per-site array is indexed by constants and lives in registers; the outputs run to the cap under a
// runtime guard. The sums per output (the residual, then each channel's acc) keep the reference's operand order
// (src/cpu/cpu_solve_level1.cpp:262-331).
//
// FAST (the fast form): the same terms with every multiply-add fused and the per-site sums in fp32 (a site's sums run
// over at most MAX_NOUT outputs), stored as doubles; the samples blend with fused multiply-adds.
template <bool LEVEL0, int C0, int C1, bool DIRECT, bool FAST>
__global__ __launch_bounds__(256) void k_site_terms(AssembleArgs a)
{
    const long long gs = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (gs >= a.sites.start[a.sites.n])
        return;
    const int p = range_plane(a.sites, gs);
    const PlaneGeom& g = a.L.planes.p[p];
    // A plane's sites number fewer than 2^31 (checked on the host).
    const unsigned int si = (unsigned int)(gs - a.sites.start[p]);
    const unsigned int per_pixel = (unsigned int)(1 + a.k);
    const unsigned int pixel = a.by_sites.div(si);
    const int j = (int)(si - pixel * per_pixel);
    const int w0 = g.w0, h0 = g.h0, w1 = g.w1, h1 = g.h1;
    const unsigned int pyu = a.by_w0[p].div(pixel);
    const int px = (int)(pixel - pyu * (unsigned int)w0), py = (int)pyu;
    const int nout = a.nout, nin = a.nin;
    const float* src = a.src + g.off0 * nout;
    const float* v0 = a.v0 + g.off0 * C0;
    const float* v1 = a.v1 + g.off1 * C1;
    float dx, dy;
    site_delta(a.k, j, px, py, dx, dy);
    const Tap tp = tap_pixel(px, py, dx, dy, w0, h0, w1, h1);
    float sbar[C0], cbar[C1], target[MAX_NOUT];
    sample_bilinear_v<C1, FAST>(v1, w1, C1, tp, cbar);
    if (j == 0)
    {
        const unsigned int pi = (unsigned int)py * (unsigned int)w0 + (unsigned int)px;
        load_texel<C0>(v0 + pi * C0, C0, sbar);
        if (!FAST)
            load_texel<MAX_NOUT>(src + pi * (unsigned int)nout, nout, target);
    }
    else
    {
        const Tap t0 = tap_pixel(px, py, dx, dy, w0, h0, w0, h0);
        sample_bilinear_v<C0, FAST>(v0, w0, C0, t0, sbar);
        if (!FAST)
            sample_bilinear_v<MAX_NOUT, FAST>(src, w0, nout, t0, target);
    }
    // The fast form's source samples are fixed for the encode: read them from the per-site table.
    if (FAST)
        load_texel<MAX_NOUT>(a.site_target + gs * nout, nout, target);
    constexpr int C = LEVEL0 ? C0 : C1;   // the solved latent's channels
    constexpr int NH = C * (C + 1) / 2;
    if constexpr (FAST)
    {
        float accf[C], hqf[NH];
#pragma unroll
        for (int aa = 0; aa < C; aa++)
            accf[aa] = 0.0f;
#pragma unroll
        for (int e = 0; e < NH; e++)
            hqf[e] = 0.0f;
#pragma unroll
        for (int r = 0; r < MAX_NOUT; r++)
        {
            if (r >= nout)
                continue;
            const float* row = a.weights + (size_t)r * nin;
            float out = a.bias[r];
            float m[C];
            if (!LEVEL0)
            {
#pragma unroll
                for (int i = 0; i < C0; i++)
                    out = fmaf(row[C1 + i], sbar[i], out);
#pragma unroll
                for (int q = 0; q < C1; q++)
                {
                    float gg = row[q];
#pragma unroll
                    for (int i = 0; i < C0; i++)
                        gg = fmaf(sbar[i], row[C1 + C0 + i * C1 + q], gg);
                    m[q] = gg;
                    out = fmaf(gg, cbar[q], out);
                }
            }
            else
            {
#pragma unroll
                for (int q = 0; q < C1; q++)
                    out = fmaf(row[q], cbar[q], out);
#pragma unroll
                for (int i = 0; i < C0; i++)
                {
                    float qq = row[C1 + i];
#pragma unroll
                    for (int q = 0; q < C1; q++)
                        qq = fmaf(cbar[q], row[C1 + C0 + i * C1 + q], qq);
                    m[i] = qq;
                    out = fmaf(qq, sbar[i], out);
                }
            }
            const float wr = a.cw[r];
            const float t = wr * (out - target[r]);
#pragma unroll
            for (int aa = 0; aa < C; aa++)
                accf[aa] = fmaf(m[aa], t, accf[aa]);
            if (DIRECT)
            {
                int e = 0;
#pragma unroll
                for (int i = 0; i < C; i++)
                {
                    const float wm = wr * m[i];
#pragma unroll
                    for (int jj = i; jj < C; jj++, e++)
                        hqf[e] = fmaf(wm, m[jj], hqf[e]);
                }
            }
        }
        if (DIRECT)
        {
#pragma unroll
            for (int e = 0; e < NH; e++)
                a.site_f[gs * NH + e] = hqf[e];
        }
        else if (!LEVEL0)
        {
#pragma unroll
            for (int i = 0; i < C0; i++)
                a.site_f[gs * C0 + i] = sbar[i];
        }
        else
        {
#pragma unroll
            for (int q = 0; q < C1; q++)
                a.site_f[gs * C1 + q] = cbar[q];
        }
#pragma unroll
        for (int aa = 0; aa < C; aa++)
            a.site_d32[gs * C + aa] = accf[aa];
        return;
    }
    double acc[C], hq[NH];
#pragma unroll
    for (int aa = 0; aa < C; aa++)
        acc[aa] = 0.0;
#pragma unroll
    for (int e = 0; e < NH; e++)
        hq[e] = 0.0;
    // Per output r: the row's matrix entries for the solved latent (G = A + sum_i sbar_i M_i for level 1,
    // Q = B + sum_q cbar_q N_q for level 0), the fp32 residual, then each channel's cw_r G_ra res_r added in r order.
#pragma unroll
    for (int r = 0; r < MAX_NOUT; r++)
    {
        if (r >= nout)
            continue;
        const float* row = a.weights + (size_t)r * nin;
        float out = a.bias[r];
        float m[C];
        if (!LEVEL0)
        {
#pragma unroll
            for (int i = 0; i < C0; i++)
                out += row[C1 + i] * sbar[i];
#pragma unroll
            for (int q = 0; q < C1; q++)
            {
                float gg = row[q];
#pragma unroll
                for (int i = 0; i < C0; i++)
                    gg += sbar[i] * row[C1 + C0 + i * C1 + q];
                m[q] = gg;
                out += gg * cbar[q];
            }
        }
        else
        {
#pragma unroll
            for (int q = 0; q < C1; q++)
                out += row[q] * cbar[q];
#pragma unroll
            for (int i = 0; i < C0; i++)
            {
                float qq = row[C1 + i];
#pragma unroll
                for (int q = 0; q < C1; q++)
                    qq += cbar[q] * row[C1 + C0 + i * C1 + q];
                m[i] = qq;
                out += qq * sbar[i];
            }
        }
        const float res = out - target[r];
#pragma unroll
        for (int aa = 0; aa < C; aa++)
            acc[aa] += (double)a.cw[r] * (double)m[aa] * (double)res;
        if (DIRECT)
        {
            int e = 0;
#pragma unroll
            for (int i = 0; i < C; i++)
#pragma unroll
                for (int jj = i; jj < C; jj++, e++)
                    hq[e] += (double)a.cw[r] * (double)m[i] * (double)m[jj];
        }
    }
    if (DIRECT)
    {
        // The site's own block Q^T diag(cw) Q, upper triangle, in place of the sample it is a function of.
#pragma unroll
        for (int e = 0; e < NH; e++)
            a.site_f[gs * NH + e] = (float)hq[e];
    }
    else if (!LEVEL0)
    {
#pragma unroll
        for (int i = 0; i < C0; i++)
            a.site_f[gs * C0 + i] = sbar[i];
    }
    else
    {
#pragma unroll
        for (int q = 0; q < C1; q++)
            a.site_f[gs * C1 + q] = cbar[q];
    }
#pragma unroll
    for (int aa = 0; aa < C; aa++)
        a.site_d[gs * C + aa] = acc[aa];
}

// The fast form's per-site source samples: site j of pixel (px, py) reads the source at the pixel (j = 0) or blends it
// with fused multiply-adds at the site's taps - what the fast site terms used to sample each time, fixed while K is.
__global__ __launch_bounds__(256) void k_site_target(AssembleArgs a, const float* src_chain, float* out)
{
    const long long gs = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (gs >= a.sites.start[a.sites.n])
        return;
    const int p = range_plane(a.sites, gs);
    const PlaneGeom& g = a.L.planes.p[p];
    const unsigned int si = (unsigned int)(gs - a.sites.start[p]);
    const unsigned int per_pixel = (unsigned int)(1 + a.k);
    const unsigned int pixel = a.by_sites.div(si);
    const int j = (int)(si - pixel * per_pixel);
    const unsigned int pyu = a.by_w0[p].div(pixel);
    const int px = (int)(pixel - pyu * (unsigned int)g.w0), py = (int)pyu;
    const int nout = a.nout;
    const float* src = src_chain + g.off0 * nout;
    float target[MAX_NOUT];
    if (j == 0)
        load_texel<MAX_NOUT>(src + ((unsigned int)py * (unsigned int)g.w0 + (unsigned int)px) * (unsigned int)nout,
                             nout, target);
    else
    {
        float dx, dy;
        site_delta(a.k, j, px, py, dx, dy);
        const Tap t0 = tap_pixel(px, py, dx, dy, g.w0, g.h0, g.w0, g.h0);
        sample_bilinear_v<MAX_NOUT, true>(src, g.w0, nout, t0, target);
    }
    for (int r = 0; r < nout; r++)
        out[gs * nout + r] = target[r];
}

template <bool LEVEL0, int C0, bool DIRECT, bool FAST>
void launch_site_terms_c1(int c1, dim3 grid, hipStream_t s, const AssembleArgs& a)
{
    switch (c1)
    {
    case 1: hipLaunchKernelGGL((k_site_terms<LEVEL0, C0, 1, DIRECT, FAST>), grid, dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_site_terms<LEVEL0, C0, 2, DIRECT, FAST>), grid, dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_site_terms<LEVEL0, C0, 3, DIRECT, FAST>), grid, dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_site_terms<LEVEL0, C0, 4, DIRECT, FAST>), grid, dim3(256), 0, s, a); break;
    }
}

template <bool LEVEL0, bool DIRECT, bool FAST>
void launch_site_terms(int c0, int c1, dim3 grid, hipStream_t s, const AssembleArgs& a)
{
    switch (c0)
    {
    case 1: launch_site_terms_c1<LEVEL0, 1, DIRECT, FAST>(c1, grid, s, a); break;
    case 2: launch_site_terms_c1<LEVEL0, 2, DIRECT, FAST>(c1, grid, s, a); break;
    case 3: launch_site_terms_c1<LEVEL0, 3, DIRECT, FAST>(c1, grid, s, a); break;
    default: launch_site_terms_c1<LEVEL0, 4, DIRECT, FAST>(c1, grid, s, a); break;
    }
}

// Merge four taps by texel index, so a clamped duplicate adds its weight to the one texel.
__device__ __forceinline__ int merge_taps(const int* cx, const int* cy, const float* cwt, int* nx, int* ny, float* tw)
{
    int taps = 0;
    for (int q = 0; q < 4; q++)
    {
        int found = -1;
        for (int e = 0; e < taps; e++)
            if (nx[e] == cx[q] && ny[e] == cy[q])
                found = e;
        if (found >= 0)
            tw[found] += cwt[q];
        else
        {
            nx[taps] = cx[q];
            ny[taps] = cy[q];
            tw[taps] = cwt[q];
            taps++;
        }
    }
    return taps;
}

// Texel ti's five blocks and gradient in the exact form. LEVEL0 false: a level-1 texel, every site whose level-1 taps
// reach it, C = c1, CS = c0. LEVEL0 true: a level-0 texel (block (c')), the 3x3-pixel neighbourhood with the centre site
// read nearest as one tap of weight 1, C = c0, CS = c1. The lane visits the footprint's sites in the reference's order
// (pixel row, pixel, site), so every sum receives the reference's operands in the reference's order.
template <bool LEVEL0, int C, int CS>
__global__ __launch_bounds__(128) void k_gather(AssembleArgs a)
{
    constexpr int NM = 1 + CS + CS * (CS + 1) / 2;
    const LatentView& L = a.L;
    const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = L.texels.start[L.texels.n];
    if (gi >= total)
        return;
    const int p = range_plane(L.texels, gi);
    const PlaneGeom& g = L.planes.p[p];
    const long long ti = gi - L.texels.start[p];
    const int w0 = g.w0, h0 = g.h0, w1 = g.w1, h1 = g.h1;
    const int tw_ = LEVEL0 ? w0 : w1;
    const int tx = (int)(ti % tw_), ty = (int)(ti / tw_);
    const int per_pixel = 1 + a.k;
    const long long site0 = a.sites.start[p];

    double sums[STENCIL_BLOCKS][NM];
    double gacc[C];
#pragma unroll
    for (int s = 0; s < STENCIL_BLOCKS; s++)
#pragma unroll
        for (int i = 0; i < NM; i++)
            sums[s][i] = 0.0;
#pragma unroll
    for (int i = 0; i < C; i++)
        gacc[i] = 0.0;

    int px_lo, px_hi, py_lo, py_hi;
    if (LEVEL0)
    {
        footprint_range(tx, w0, w0, px_lo, px_hi);
        footprint_range(ty, h0, h0, py_lo, py_hi);
    }
    else
    {
        footprint_range(tx, w0, w1, px_lo, px_hi);
        footprint_range(ty, h0, h1, py_lo, py_hi);
    }
    // The sites in the reference's order (pixel row, pixel, site), stepped without a division.
    const int nxp = px_hi - px_lo + 1;
    const int visits = nxp * (py_hi - py_lo + 1) * per_pixel;
    int px = px_lo, py = py_lo, j = 0;
    for (int v = 0; v < visits; v++)
    {
        if (v > 0 && ++j > a.k)
        {
            j = 0;
            if (++px > px_hi)
            {
                px = px_lo;
                ++py;
            }
        }
        float dx, dy;
        site_delta(a.k, j, px, py, dx, dy);
        int nx[4], ny[4];
        float tw[4];
        int taps;
        if (LEVEL0 && j == 0)
        {
            nx[0] = px;
            ny[0] = py;
            tw[0] = 1.0f;
            taps = 1;
        }
        else
        {
            const Tap t = LEVEL0 ? tap_pixel(px, py, dx, dy, w0, h0, w0, h0)
                                 : tap_pixel(px, py, dx, dy, w0, h0, w1, h1);
            const int cx[4] = { t.x0, t.x1, t.x0, t.x1 };
            const int cy[4] = { t.y0, t.y0, t.y1, t.y1 };
            const float cwt[4] = { (1.0f - t.fx) * (1.0f - t.fy), t.fx * (1.0f - t.fy), (1.0f - t.fx) * t.fy,
                                   t.fx * t.fy };
            taps = merge_taps(cx, cy, cwt, nx, ny, tw);
        }
        int self = -1;
        for (int e = 0; e < taps; e++)
            if (nx[e] == tx && ny[e] == ty)
                self = e;
        if (self < 0)
            continue;
        const float w_self = tw[self];
        const long long s = site0 + ((long long)py * w0 + px) * per_pixel + j;
        float sv[CS];
#pragma unroll
        for (int i = 0; i < CS; i++)
            sv[i] = a.site_f[s * CS + i];
        double mono[NM];
        monomials<CS>(sv, mono);
#pragma unroll
        for (int aa = 0; aa < C; aa++)
            gacc[aa] += (double)w_self * a.site_d[s * C + aa];
        for (int e = 0; e < taps; e++)
        {
            const int slot = stencil_slot(nx[e] - tx, ny[e] - ty);
            if (slot < 0)
                continue;
            const double scale = (double)w_self * (double)tw[e];
#pragma unroll
            for (int sl = 0; sl < STENCIL_BLOCKS; sl++)
                if (sl == slot)
#pragma unroll
                    for (int i = 0; i < NM; i++)
                        sums[sl][i] += scale * mono[i];
        }
    }

    // The five owned blocks expanded from the monomial sums; a zero sum is skipped.
    const long long gtx = L.texels.start[p] + ti;
    const long long sn = L.sm.sn;
    double* out_h = L.stencil + stencil_pos(L.sm, p, tw_, tx, ty);
#pragma unroll
    for (int slot = 0; slot < STENCIL_BLOCKS; slot++)
    {
        double blk[C * C];
#pragma unroll
        for (int i = 0; i < C * C; i++)
            blk[i] = 0.0;
#pragma unroll
        for (int i = 0; i < NM; i++)
        {
            const double sv = sums[slot][i];
            if (sv == 0.0)
                continue;
            const double* mat = a.mono_p + (size_t)i * C * C;
#pragma unroll
            for (int e = 0; e < C * C; e++)
                blk[e] += sv * mat[e];
        }
#pragma unroll
        for (int i = 0; i < C * C; i++)
            out_h[(slot * C * C + i) * sn] = blk[i];
    }
#pragma unroll
    for (int i = 0; i < C; i++)
        L.grad[gtx * C + i] = gacc[i];
}

template <bool LEVEL0, int C>
void launch_gather_cs(int cs, dim3 grid, hipStream_t s, const AssembleArgs& a)
{
    switch (cs)
    {
    case 1: hipLaunchKernelGGL((k_gather<LEVEL0, C, 1>), grid, dim3(128), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_gather<LEVEL0, C, 2>), grid, dim3(128), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_gather<LEVEL0, C, 3>), grid, dim3(128), 0, s, a); break;
    default: hipLaunchKernelGGL((k_gather<LEVEL0, C, 4>), grid, dim3(128), 0, s, a); break;
    }
}

template <bool LEVEL0>
void launch_gather(int c, int cs, long long texels, hipStream_t s, const AssembleArgs& a)
{
    const dim3 grid(blocks_for(texels, 128));
    switch (c)
    {
    case 1: launch_gather_cs<LEVEL0, 1>(cs, grid, s, a); break;
    case 2: launch_gather_cs<LEVEL0, 2>(cs, grid, s, a); break;
    case 3: launch_gather_cs<LEVEL0, 3>(cs, grid, s, a); break;
    default: launch_gather_cs<LEVEL0, 4>(cs, grid, s, a); break;
    }
}

// The fast form's view of a site's taps: a tap texel's weight separates into a column weight times a row weight (both
// taps of a clamped pair add up at the one texel), so the weights of a texel's five owned slots are products of five
// column and row weights - no merge of the four taps and no test per tap and slot.
// (tap_weight: adjoint.h)

// w_self (the site's weight at texel (tx, ty)) times its weight at each owned slot's texel; false when the texel holds no
// weight of the site.
__device__ __forceinline__ bool slot_scales(const Tap& t, int tx, int ty, double (&scale)[STENCIL_BLOCKS], float& w_self)
{
    const float wxm = tap_weight(tx - 1, t.x0, t.x1, t.fx), wx0 = tap_weight(tx, t.x0, t.x1, t.fx),
                wxp = tap_weight(tx + 1, t.x0, t.x1, t.fx);
    const float wy0 = tap_weight(ty, t.y0, t.y1, t.fy), wyp = tap_weight(ty + 1, t.y0, t.y1, t.fy);
    w_self = wx0 * wy0;
    if (w_self == 0.0f)
        return false;
    const double ws = (double)w_self;
    scale[0] = ws * (double)(wx0 * wy0);
    scale[1] = ws * (double)(wxp * wy0);
    scale[2] = ws * (double)(wxm * wyp);
    scale[3] = ws * (double)(wx0 * wyp);
    scale[4] = ws * (double)(wxp * wyp);
    return true;
}

// (site_taps: adjoint.h)

// One DPP move of a double (both halves) under the data-parallel-primitive control CTRL, within the row.
template <int CTRL>
__device__ __forceinline__ double dpp_double(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// The sum over an aligned group of G lanes (G a power of two up to 16), on every lane of the group: pairs of lanes
// (quad_perm [1,0,3,2], then [2,3,0,1]), then halves of eight (row_half_mirror) and of sixteen (row_mirror). Each level
// adds two partial sums that every lane of the pair holds in the other order, and addition commutes, so every lane of
// the group ends with the same bits.
template <int G>
__device__ __forceinline__ double group_sum(double x)
{
    if constexpr (G >= 2)
        x += dpp_double<0xB1>(x);
    if constexpr (G >= 4)
        x += dpp_double<0x4E>(x);
    if constexpr (G >= 8)
        x += dpp_double<0x141>(x);
    if constexpr (G >= 16)
        x += dpp_double<0x140>(x);
    return x;
}

// Lane sub's walk over a footprint's sites in the order (pixel row, pixel, site j), taking every G-th site of the
// order: the stride of G sites is split once into (rows, pixels, sites) steps, so the walk has no division.
struct SiteWalk
{
    int j, px, py, px_lo, px_hi, nxp, per_pixel, dj, dpx, dpy;
    __device__ __forceinline__ void init(int sub, int g, int lo_x, int hi_x, int lo_y, int pp)
    {
        px_lo = lo_x;
        px_hi = hi_x;
        nxp = hi_x - lo_x + 1;
        per_pixel = pp;
        j = sub % pp;
        const int pix = sub / pp;
        px = lo_x + pix % nxp;
        py = lo_y + pix / nxp;
        dj = g % pp;
        const int dpix = g / pp;
        dpx = dpix % nxp;
        dpy = dpix / nxp;
    }
    __device__ __forceinline__ void next()
    {
        j += dj;
        int cpx = dpx;
        if (j >= per_pixel)
        {
            j -= per_pixel;
            cpx++;
        }
        px += cpx;
        py += dpy;
        if (px > px_hi)
        {
            px -= nxp;
            py++;
        }
    }
};

constexpr int GATHER_TAB_STRIDE = 2 + STENCIL_BLOCKS;   // an interior list's entry: code, w_self, five scales

// The fast form's level-1 gather: G lanes per texel take the footprint's sites in turn (site j fastest, then pixel,
// then pixel row), each adding its five slots' terms branch-free with fused multiply-adds; the group's sums meet in a
// fixed butterfly (deterministic).
template <int C, int CS, int G>
__global__ __launch_bounds__(128) void k_gather_fast1(AssembleArgs a)
{
    constexpr int NM = 1 + CS + CS * (CS + 1) / 2;
    const LatentView& L = a.L;
    const long long gt = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const long long gi = gt / G;
    const int sub = (int)(gt % G);
    const long long total = L.texels.start[L.texels.n];
    // Every lane of a group must reach the shuffles, so a lane past the end idles instead of leaving.
    const bool live = gi < total;
    const int p = range_plane(L.texels, live ? gi : total - 1);
    const PlaneGeom& g = L.planes.p[p];
    const long long ti = (live ? gi : total - 1) - L.texels.start[p];
    const int w0 = g.w0, h0 = g.h0, w1 = g.w1, h1 = g.h1;
    const int tx = (int)(ti % w1), ty = (int)(ti / w1);
    const int per_pixel = 1 + a.k;
    const long long site0 = a.sites.start[p];

    double sums[STENCIL_BLOCKS][NM], gacc[C];
#pragma unroll
    for (int s = 0; s < STENCIL_BLOCKS; s++)
#pragma unroll
        for (int i = 0; i < NM; i++)
            sums[s][i] = 0.0;
#pragma unroll
    for (int i = 0; i < C; i++)
        gacc[i] = 0.0;

    int px_lo, px_hi, py_lo, py_hi;
    footprint_range(tx, w0, w1, px_lo, px_hi);
    footprint_range(ty, h0, h1, py_lo, py_hi);
    const int visits = live ? (px_hi - px_lo + 1) * (py_hi - py_lo + 1) * per_pixel : 0;
    SiteWalk wk;
    wk.init(sub, G, px_lo, px_hi, py_lo, per_pixel);
    for (int v = sub; v < visits; v += G, wk.next())
    {
        const int j = wk.j, px = wk.px, py = wk.py;
        const Tap t = site_taps(a.k, j, px, py, w0, h0, w1, h1, false);
        double scale[STENCIL_BLOCKS];
        float w_self;
        if (slot_scales(t, tx, ty, scale, w_self))
        {
            const long long s = site0 + ((long long)py * w0 + px) * per_pixel + j;
            float sv[CS];
#pragma unroll
            for (int i = 0; i < CS; i++)
                sv[i] = a.site_f[s * CS + i];
            double mono[NM];
            monomials<CS>(sv, mono);
#pragma unroll
            for (int aa = 0; aa < C; aa++)
                gacc[aa] = fma((double)w_self, (double)a.site_d32[s * C + aa], gacc[aa]);
#pragma unroll
            for (int sl = 0; sl < STENCIL_BLOCKS; sl++)
#pragma unroll
                for (int i = 0; i < NM; i++)
                    sums[sl][i] = fma(scale[sl], mono[i], sums[sl][i]);
        }
    }

#pragma unroll
    for (int s = 0; s < STENCIL_BLOCKS; s++)
#pragma unroll
        for (int i = 0; i < NM; i++)
            sums[s][i] = group_sum<G>(sums[s][i]);
#pragma unroll
    for (int i = 0; i < C; i++)
        gacc[i] = group_sum<G>(gacc[i]);
    if (!live)
        return;

    // The five owned blocks expanded from the monomial sums, which the butterfly left on every lane of the group: lane
    // sub writes entries sub, sub + G, ... of each block.
    const long long gtx = L.texels.start[p] + ti;
    const long long sn = L.sm.sn;
    const long long opos = stencil_pos(L.sm, p, w1, tx, ty);
    float* out_f = L.stencil32 + opos;
    for (int e = sub; e < C * C; e += G)
    {
        double mat[NM];
#pragma unroll
        for (int i = 0; i < NM; i++)
            mat[i] = a.mono_p[(size_t)i * C * C + e];
#pragma unroll
        for (int slot = 0; slot < STENCIL_BLOCKS; slot++)
        {
            double v = 0.0;
#pragma unroll
            for (int i = 0; i < NM; i++)
                v = fma(sums[slot][i], mat[i], v);
            out_f[(slot * C * C + e) * sn] = (float)v;
        }
    }
    for (int ch = sub; ch < C; ch += G)
    {
        // gacc[ch], picked by selects (a register array indexed by a lane's value would live in scratch).
        double gv = gacc[0];
#pragma unroll
        for (int i = 1; i < C; i++)
            gv = ch == i ? gacc[i] : gv;
        L.grad[gtx * C + ch] = gv;
    }
}

template <int C, int G>
void launch_gather_fast1(int cs, long long texels, hipStream_t s, const AssembleArgs& a)
{
    const dim3 grid(blocks_for(texels * G, 128));
    switch (cs)
    {
    case 1: hipLaunchKernelGGL((k_gather_fast1<C, 1, G>), grid, dim3(128), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_gather_fast1<C, 2, G>), grid, dim3(128), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_gather_fast1<C, 3, G>), grid, dim3(128), 0, s, a); break;
    default: hipLaunchKernelGGL((k_gather_fast1<C, 4, G>), grid, dim3(128), 0, s, a); break;
    }
}

// The fast level-0 gather's weights of an interior texel: away from the plane's edges no tap is clamped, and a site's
// taps relative to the texel (so its weight w_self and its five slot scales) depend only on the site's place in the
// 3x3-pixel footprint and on the pixel parity that turns its sub-texel offset - the same bits for every such texel,
// since p + d and its floor are exact. So there are four classes, the texel's (x, y) parity. Class c lists, in the walk's
// order (pixel row, pixel, site), the sites that put weight on the texel: entry e = [code, w_self, five scales] with
// code = 64 oy + 8 ox + j (the pixel's offset from the footprint's corner, and the site), and the list's length at
// tab[GATHER_TAB_MAX * GATHER_TAB_STRIDE + c]. Computed by the functions the general path calls, on a texel in the
// middle of a plane large enough that nothing clamps.
constexpr int GATHER_TAB_MAX = 4 * 9 * 5;   // four classes of up to 45 sites (K <= 4)

__global__ __launch_bounds__(64) void k_gather_table(int k, double* tab)
{
    const int cls = threadIdx.x;
    if (cls >= 4)
        return;
    const int pp = 1 + k;
    const int tx = 16 + (cls & 1), ty = 16 + (cls >> 1);
    int n = 0;
    for (int oy = 0; oy < 3; oy++)
        for (int ox = 0; ox < 3; ox++)
            for (int j = 0; j < pp; j++)
            {
                const Tap t = site_taps(k, j, tx - 1 + ox, ty - 1 + oy, 64, 64, 64, 64, true);
                double scale[STENCIL_BLOCKS];
                float w_self;
                if (!slot_scales(t, tx, ty, scale, w_self))
                    continue;
                double* out = tab + (size_t)(cls * 45 + n) * GATHER_TAB_STRIDE;
                out[0] = (double)(64 * oy + 8 * ox + j);
                out[1] = (double)w_self;
                for (int q = 0; q < STENCIL_BLOCKS; q++)
                    out[2 + q] = scale[q];
                n++;
            }
    tab[GATHER_TAB_MAX * GATHER_TAB_STRIDE + cls] = (double)n;
}

// The fast form's level-0 gather over LDS tiles (C <= 3: the footprint's terms fill 61 KB at C = 3, and four
// channels' would pass the 64 KB a workgroup's static LDS holds), one texel per lane: a workgroup takes a GT_W x GT_H
// tile of texels, copies its footprint's site terms (the pixels of the tile and a one-pixel ring: rows of consecutive
// sites, so the copy is coalesced) into LDS, and each lane walks its texel's 3x3 pixels' sites there, in the order
// (pixel row, pixel, site), interior texels taking their sites and weights from their class's list and the others
// from their taps. The per-site sums are fp32 values (k_site_terms' fast form), so LDS holds them as floats without
// loss.
constexpr int GT_W = 32, GT_H = 8;

template <int C>
__global__ __launch_bounds__(256) void k_gather_tile0(AssembleArgs a, Ranges tiles)
{
    constexpr int NH = C * (C + 1) / 2;
    constexpr int BOX = (GT_W + 2) * (GT_H + 2) * 5;   // the footprint's sites at K <= 4
    __shared__ float s_f[BOX * NH];
    __shared__ float s_d[BOX * C];
    const LatentView& L = a.L;
    const int pp = 1 + a.k;
    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int p = range_plane(tiles, tile);
    const PlaneGeom& g = L.planes.p[p];
    const int w0 = g.w0, h0 = g.h0;
    const int t = tile - (int)tiles.start[p];
    const int tiles_x = (w0 + GT_W - 1) / GT_W;
    const int tx0 = (t % tiles_x) * GT_W, ty0 = (t / tiles_x) * GT_H;
    const int bx0 = max(0, tx0 - 1), by0 = max(0, ty0 - 1);
    const int bw = min(w0 - 1, tx0 + GT_W) - bx0 + 1, bh = min(h0 - 1, ty0 + GT_H) - by0 + 1;
    const int row_sites = bw * pp;
    const long long site0 = a.sites.start[p];
    {
        // Box row r holds sites [((by0 + r) w0 + bx0) pp, + row_sites) of the plane, at LDS site r row_sites, each
        // sum in its own array (component q of site s at q BOX + s): the lanes of a wavefront walk sites ten apart
        // (two texels of pp = 5 sites), which the site-major layout (six floats a site) put on a quarter of the
        // banks. Sixteen loads in flight per lane before their LDS stores.
        constexpr int U = 16;
        const Div16 by_rowf(row_sites * NH), by_rowd(row_sites * C);
        const float* gf = a.site_f + (site0 + ((long long)by0 * w0 + bx0) * pp) * NH;
        const float* gd = a.site_d32 + (site0 + ((long long)by0 * w0 + bx0) * pp) * C;
        const long long stride_f = (long long)w0 * pp * NH, stride_d = (long long)w0 * pp * C;
        const int nf = bh * row_sites * NH, nd = bh * row_sites * C;
        for (int i0 = threadIdx.x; i0 < nf || i0 < nd; i0 += U * 256)
        {
            float vf[U], vd[U];
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const int i = i0 + u * 256;
                if (i < nf)
                {
                    const int r = (int)by_rowf((unsigned)i);
                    vf[u] = gf[r * stride_f + (i - r * row_sites * NH)];
                }
                if (i < nd)
                {
                    const int r = (int)by_rowd((unsigned)i);
                    vd[u] = gd[r * stride_d + (i - r * row_sites * C)];
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++)
            {
                const int i = i0 + u * 256;
                if (i < nf)
                    s_f[(i % NH) * BOX + i / NH] = vf[u];
                if (i < nd)
                    s_d[(i % C) * BOX + i / C] = vd[u];
            }
        }
    }
    __syncthreads();

    // Wavefront w takes the texels of parity class c = w ((x, y) parity (c & 1, c >> 1)): 16 columns x 4 rows of the
    // tile. One class per wavefront makes the interior list the same for all its lanes (scalar reads), and the class's
    // texels are the contiguous ones of the stencil's colour-major layout.
    const int cls = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    const int tx = tx0 + 2 * (lane & 15) + (cls & 1), ty = ty0 + 2 * (lane >> 4) + (cls >> 1);
    if (tx >= w0 || ty >= h0)
        return;
    double hs[STENCIL_BLOCKS][NH], gacc[C];
#pragma unroll
    for (int sl = 0; sl < STENCIL_BLOCKS; sl++)
#pragma unroll
        for (int e = 0; e < NH; e++)
            hs[sl][e] = 0.0;
#pragma unroll
    for (int i = 0; i < C; i++)
        gacc[i] = 0.0;
    auto add = [&](int site, double ws, const double* scale) {
        const float* f = s_f + site;
        const float* d = s_d + site;
#pragma unroll
        for (int aa = 0; aa < C; aa++)
            gacc[aa] = fma(ws, (double)d[aa * BOX], gacc[aa]);
#pragma unroll
        for (int sl = 0; sl < STENCIL_BLOCKS; sl++)
#pragma unroll
            for (int q = 0; q < NH; q++)
                hs[sl][q] = fma(scale[sl], (double)f[q * BOX], hs[sl][q]);
    };
    if (tx >= 2 && tx <= w0 - 3 && ty >= 2 && ty <= h0 - 3)
    {
        const double* ct = a.tab + cls * 45 * GATHER_TAB_STRIDE;
        const int n = (int)a.tab[GATHER_TAB_MAX * GATHER_TAB_STRIDE + cls];
        const int corner = ((ty - 1 - by0) * bw + (tx - 1 - bx0)) * pp;
        for (int v = 0; v < n; v++)
        {
            const double* e = ct + v * GATHER_TAB_STRIDE;
            const int code = (int)e[0];
            add(corner + ((code >> 6) * bw + ((code >> 3) & 7)) * pp + (code & 7), e[1], e + 2);
        }
    }
    else
    {
        int px_lo, px_hi, py_lo, py_hi;
        footprint_range(tx, w0, w0, px_lo, px_hi);
        footprint_range(ty, h0, h0, py_lo, py_hi);
        for (int py = py_lo; py <= py_hi; py++)
            for (int px = px_lo; px <= px_hi; px++)
                for (int j = 0; j < pp; j++)
                {
                    const Tap tp = site_taps(a.k, j, px, py, w0, h0, w0, h0, true);
                    double scale[STENCIL_BLOCKS];
                    float w_self;
                    if (slot_scales(tp, tx, ty, scale, w_self))
                        add(((py - by0) * bw + (px - bx0)) * pp + j, (double)w_self, scale);
                }
    }

    // The five owned blocks, entry (i, jj) from the triangle's (min, max) element, and the gradient.
    const long long sn = L.sm.sn;
    const long long opos = stencil_pos(L.sm, p, w0, tx, ty);
    float* out_f = L.stencil32 + opos;
#pragma unroll
    for (int i = 0; i < C; i++)
#pragma unroll
        for (int jj = 0; jj < C; jj++)
        {
            const int lo = i < jj ? i : jj, hi = i < jj ? jj : i;
            const int tri = lo * C - lo * (lo - 1) / 2 + (hi - lo);
            const int e = i * C + jj;
#pragma unroll
            for (int sl = 0; sl < STENCIL_BLOCKS; sl++)
                out_f[(sl * C * C + e) * sn] = (float)hs[sl][tri];
        }
    const long long gtx = L.texels.start[p] + (long long)ty * w0 + tx;
#pragma unroll
    for (int ch = 0; ch < C; ch++)
        L.grad[gtx * C + ch] = gacc[ch];
}

// The fast form's level-0 gather: each site stores its own block Q^T diag(cw) Q (upper triangle, fp32) in place of
// the level-1 sample, so a texel adds scale x that block into each owned slot - C (C + 1) / 2 values per tap instead of
// the 1 + c1 + c1 (c1 + 1) / 2 monomials and their expansion. G lanes share a texel's 3x3-pixel footprint (45 sites at
// K = 4) and meet in the butterfly; the blocks are symmetric, so each slot's block is written from its triangle.
template <int C, int G>
__global__ __launch_bounds__(128) void k_gather_direct(AssembleArgs a)
{
    constexpr int NH = C * (C + 1) / 2;
    const LatentView& L = a.L;
    const long long gt = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const long long gi = gt / G;
    const int sub = (int)(gt % G);
    const long long total = L.texels.start[L.texels.n];
    // Every lane of a group must reach the butterfly, so a lane past the end idles instead of leaving.
    const bool live = gi < total;
    const int p = range_plane(L.texels, live ? gi : total - 1);
    const PlaneGeom& g = L.planes.p[p];
    const long long ti = (live ? gi : total - 1) - L.texels.start[p];
    const int w0 = g.w0, h0 = g.h0;
    const int tx = (int)(ti % w0), ty = (int)(ti / w0);
    const int per_pixel = 1 + a.k;
    const long long site0 = a.sites.start[p];

    double hs[STENCIL_BLOCKS][NH], gacc[C];
#pragma unroll
    for (int sl = 0; sl < STENCIL_BLOCKS; sl++)
#pragma unroll
        for (int e = 0; e < NH; e++)
            hs[sl][e] = 0.0;
#pragma unroll
    for (int i = 0; i < C; i++)
        gacc[i] = 0.0;

    int px_lo, px_hi, py_lo, py_hi;
    footprint_range(tx, w0, w0, px_lo, px_hi);
    footprint_range(ty, h0, h0, py_lo, py_hi);
    const int visits = live ? (px_hi - px_lo + 1) * (py_hi - py_lo + 1) * per_pixel : 0;
    SiteWalk wk;
    wk.init(sub, G, px_lo, px_hi, py_lo, per_pixel);
    for (int v = sub; v < visits; v += G, wk.next())
    {
        const int j = wk.j, px = wk.px, py = wk.py;
        const Tap t = site_taps(a.k, j, px, py, w0, h0, w0, h0, true);
        double scale[STENCIL_BLOCKS];
        float w_self;
        if (!slot_scales(t, tx, ty, scale, w_self))
            continue;
        const long long s = site0 + ((long long)py * w0 + px) * per_pixel + j;
        float hv[NH];
#pragma unroll
        for (int e = 0; e < NH; e++)
            hv[e] = a.site_f[s * NH + e];
#pragma unroll
        for (int aa = 0; aa < C; aa++)
            gacc[aa] = fma((double)w_self, (double)a.site_d32[s * C + aa], gacc[aa]);
#pragma unroll
        for (int sl = 0; sl < STENCIL_BLOCKS; sl++)
#pragma unroll
            for (int q = 0; q < NH; q++)
                hs[sl][q] = fma(scale[sl], (double)hv[q], hs[sl][q]);
    }

#pragma unroll
    for (int sl = 0; sl < STENCIL_BLOCKS; sl++)
#pragma unroll
        for (int q = 0; q < NH; q++)
            hs[sl][q] = group_sum<G>(hs[sl][q]);
#pragma unroll
    for (int i = 0; i < C; i++)
        gacc[i] = group_sum<G>(gacc[i]);
    if (!live)
        return;

    // Lane sub writes entries sub, sub + G, ... of each slot's C x C block, entry (i, jj) from the triangle's
    // (min, max) element, picked by selects (a register array indexed by a lane's value would live in scratch).
    const long long sn = L.sm.sn;
    const long long opos = stencil_pos(L.sm, p, w0, tx, ty);
    float* out_f = L.stencil32 + opos;
    for (int e = sub; e < C * C; e += G)
    {
        const int i = e / C, jj = e % C;
        const int lo = i < jj ? i : jj, hi = i < jj ? jj : i;
        const int tri = lo * C - lo * (lo - 1) / 2 + (hi - lo);
#pragma unroll
        for (int sl = 0; sl < STENCIL_BLOCKS; sl++)
        {
            double v = hs[sl][0];
#pragma unroll
            for (int q = 1; q < NH; q++)
                v = tri == q ? hs[sl][q] : v;
            out_f[(sl * C * C + e) * sn] = (float)v;
        }
    }
    for (int ch = sub; ch < C; ch += G)
    {
        double gv = gacc[0];
#pragma unroll
        for (int i = 1; i < C; i++)
            gv = ch == i ? gacc[i] : gv;
        L.grad[(L.texels.start[p] + ti) * C + ch] = gv;
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// Fixed-order reductions
// ---------------------------------------------------------------------------------------------------------------------

// The chunks of CHUNK texels of every plane, one wavefront each. Chunk index space: plane p owns
// [cstart[p], cstart[p+1]).
struct ChunkArgs
{
    LatentView L;
    Ranges chunks;
    double* partial;
    int exact;
};

__device__ __forceinline__ void chunk_bounds(const ChunkArgs& a, long long chunk, int& p, long long& vb, long long& ve)
{
    p = range_plane(a.chunks, chunk);
    const long long texels = a.L.texels.start[p + 1] - a.L.texels.start[p];
    const long long tb = (chunk - a.chunks.start[p]) * CHUNK;
    const long long te = tb + CHUNK < texels ? tb + CHUNK : texels;
    vb = (a.L.texels.start[p] + tb) * a.L.c;
    ve = (a.L.texels.start[p] + te) * a.L.c;
}

// The fast form's diagonal statistics per chunk: one 256-lane workgroup, each lane's strided values, then trees.
__global__ __launch_bounds__(256) void k_diag_fast(ChunkArgs a)
{
    __shared__ double part[4][3];
    int p;
    long long vb, ve;
    chunk_bounds(a, blockIdx.x, p, vb, ve);
    const int c = a.L.c, w = a.L.pw[p];
    const long long t0 = a.L.texels.start[p];
    double sum = 0.0, lo = 1e300, hi = -1e300;
    // Texel by texel (32-bit position arithmetic: a plane's texels number fewer than 2^31), each texel's C diagonal
    // entries in turn.
    const int tb = (int)(vb / c - t0), te = (int)(ve / c - t0);
    for (int t = tb + (int)threadIdx.x; t < te; t += 256)
    {
        const int ty = t / w, tx = t - ty * w;
        const long long pos = stencil_pos(a.L.sm, p, w, tx, ty);
        for (int q = 0; q < c; q++)
        {
            const double x = a.L.stencil32[(q * c + q) * a.L.sm.sn + pos];
            sum += x;
            lo = x < lo ? x : lo;
            hi = x > hi ? x : hi;
        }
    }
    for (int off = 32; off > 0; off >>= 1)
    {
        sum += __shfl_xor(sum, off);
        const double ol = __shfl_xor(lo, off), oh = __shfl_xor(hi, off);
        lo = ol < lo ? ol : lo;
        hi = oh > hi ? oh : hi;
    }
    if (threadIdx.x % 64 == 0)
    {
        part[threadIdx.x / 64][0] = sum;
        part[threadIdx.x / 64][1] = lo;
        part[threadIdx.x / 64][2] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double s = 0.0, l = 1e300, h = -1e300;
        for (int wv = 0; wv < 4; wv++)
        {
            s += part[wv][0];
            l = part[wv][1] < l ? part[wv][1] : l;
            h = part[wv][2] > h ? part[wv][2] : h;
        }
        a.partial[blockIdx.x * 3] = s;
        a.partial[blockIdx.x * 3 + 1] = l;
        a.partial[blockIdx.x * 3 + 2] = h;
    }
}

// The diagonal of H over the chunk's texels: its sum (serial, texel then channel), smallest and largest.
__global__ __launch_bounds__(64) void k_diag_chunks(ChunkArgs a)
{
    int p;
    long long vb, ve;
    chunk_bounds(a, blockIdx.x, p, vb, ve);
    const int c = a.L.c;
    const long long tb = vb / c, te = ve / c;
    const int lane = threadIdx.x;
    double sum = 0.0, lo = 1e300, hi = -1e300;
    constexpr int ROWS = 8;
    __shared__ double buf[64 * ROWS];
    const long long end = te * c;
    for (long long base = tb * c; base < end; base += 64 * ROWS)
    {
        const int n = (int)(end - base < 64 * ROWS ? end - base : 64 * ROWS);
        double x[ROWS];
#pragma unroll
        for (int u = 0; u < ROWS; u++)
        {
            const long long i = base + u * 64 + lane;
            x[u] = 0.0;
            if (i < end)
            {
                const long long t = i / c - a.L.texels.start[p];
                const int q = (int)(i % c);
                const int w = a.L.pw[p];
                const long long s = stencil_pos(a.L.sm, p, w, (int)(t % w), (int)(t / w));
                x[u] = a.L.stencil[(q * c + q) * a.L.sm.sn + s];
            }
        }
#pragma unroll
        for (int u = 0; u < ROWS; u++)
        {
            const long long row = base + u * 64;
            const int count = row >= end ? 0 : (int)(end - row < 64 ? end - row : 64);
            if (count == 0)
                continue;
            // Minimum and maximum are exact in any order; the sum is folded serially below.
            double l = lane < count ? x[u] : 1e300, h = lane < count ? x[u] : -1e300;
            for (int off = 32; off > 0; off >>= 1)
            {
                const double ol = __shfl_xor(l, off), oh = __shfl_xor(h, off);
                l = ol < l ? ol : l;
                h = oh > h ? oh : h;
            }
            lo = l < lo ? l : lo;
            hi = h > hi ? h : hi;
        }
        // Staged through LDS so lane 0 folds the batch in index order (texel then channel).
#pragma unroll
        for (int u = 0; u < ROWS; u++)
            buf[u * 64 + lane] = x[u];
        __syncthreads();
        if (lane == 0)
            sum = fold_serial16(buf, n, sum);
        __syncthreads();
    }
    if (lane == 0)
    {
        a.partial[blockIdx.x * 3] = sum;
        a.partial[blockIdx.x * 3 + 1] = lo;
        a.partial[blockIdx.x * 3 + 2] = hi;
    }
}

// Per plane: the chunk sums in chunk order, then lambda = ridge * sum / unknowns.
// acc + t[0] + t[stride] + ... (n terms) in index order, sixteen loads issued ahead of their dependent additions.
__device__ __forceinline__ double fold_strided(const double* t, long long n, int stride, double acc)
{
    long long i = 0;
    for (; i + 16 <= n; i += 16)
    {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++)
            v[u] = t[(i + u) * stride];
#pragma unroll
        for (int u = 0; u < 16; u++)
            acc += v[u];
    }
    for (; i < n; i++)
        acc += t[i * stride];
    return acc;
}

// The workgroup's (256 lanes) minimum, maximum or sum of one value per lane, through LDS; every lane gets it.
__device__ __forceinline__ double block_reduce256(double v, int op, double* scratch)
{
    scratch[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1)
    {
        if ((int)threadIdx.x < h)
        {
            const double o = scratch[threadIdx.x + h], m = scratch[threadIdx.x];
            scratch[threadIdx.x] = op == 0 ? (o < m ? o : m) : (op == 1 ? (o > m ? o : m) : m + o);
        }
        __syncthreads();
    }
    const double r = scratch[0];
    __syncthreads();
    return r;
}

// One 256-lane workgroup per plane: the extremes over the plane's chunks in any order (they are exact), the sum in chunk
// order from zero (the exact form) or in this kernel's fixed order (the fast form).
__global__ __launch_bounds__(256) void k_diag_fold(ChunkArgs a, double ridge)
{
    __shared__ double scratch[256];
    const int p = blockIdx.x;
    const long long b = a.chunks.start[p], e = a.chunks.start[p + 1];
    double sum = 0.0, lo = 1e300, hi = -1e300;
    for (long long ch = b + threadIdx.x; ch < e; ch += 256)
    {
        if (!a.exact)
            sum += a.partial[ch * 3];
        const double l = a.partial[ch * 3 + 1], h = a.partial[ch * 3 + 2];
        lo = l < lo ? l : lo;
        hi = h > hi ? h : hi;
    }
    lo = block_reduce256(lo, 0, scratch);
    hi = block_reduce256(hi, 1, scratch);
    sum = a.exact ? 0.0 : block_reduce256(sum, 2, scratch);
    if (threadIdx.x != 0)
        return;
    if (a.exact)
        sum = fold_strided(a.partial + b * 3, e - b, 3, 0.0);
    double* s = a.L.scalars + (size_t)p * SC_COUNT;
    const long long unknowns = (a.L.texels.start[p + 1] - a.L.texels.start[p]) * a.L.c;
    s[SC_DIAG_SUM] = sum;
    s[SC_DIAG_MIN] = lo;
    s[SC_DIAG_MAX] = hi;
    s[SC_LAMBDA] = unknowns > 0 ? ridge * sum / (double)unknowns : 0.0;
}

// The exact form's two inner products over the chunk's values, sum a_i b_i and sum c_i d_i, each folded serially in
// index order: the products staged through LDS, lanes 0 and 1 adding them.
__global__ __launch_bounds__(64) void k_dot_chunks(ChunkArgs a, const double* x0, const double* y0, const double* x1,
                                                   const double* y1)
{
    int p;
    long long vb, ve;
    chunk_bounds(a, blockIdx.x, p, vb, ve);
    if (a.L.scalars[(size_t)p * SC_COUNT + SC_ACTIVE] == 0.0)
        return;
    const int lane = threadIdx.x;
    double acc = 0.0;   // lane 0 folds the first product, lane 1 the second
    constexpr int ROWS = 8;   // 512 values staged per batch: one memory latency per batch
    __shared__ double buf[2][64 * ROWS];
    for (long long base = vb; base < ve; base += 64 * ROWS)
    {
        const int n = (int)(ve - base < 64 * ROWS ? ve - base : 64 * ROWS);
#pragma unroll
        for (int u = 0; u < ROWS; u++)
        {
            const long long i = base + u * 64 + lane;
            buf[0][u * 64 + lane] = i < ve ? x0[i] * y0[i] : 0.0;
            buf[1][u * 64 + lane] = (x1 && i < ve) ? x1[i] * y1[i] : 0.0;
        }
        __syncthreads();
        if (lane < 2)
            acc = fold_serial16(buf[lane], n, acc);
        __syncthreads();
    }
    if (lane < 2)
        a.partial[blockIdx.x * 2 + lane] = acc;
}

// The fast form's two inner products: one 256-lane workgroup per chunk, each lane's own strided values, then trees.
constexpr int DOT_FAST_THREADS = 256;

template <class V>
__global__ __launch_bounds__(DOT_FAST_THREADS) void k_dot_fast(ChunkArgs a, const V* x0, const V* y0, const V* x1,
                                                               const V* y1)
{
    __shared__ double part[DOT_FAST_THREADS / 64][2];
    int p;
    long long vb, ve;
    chunk_bounds(a, blockIdx.x, p, vb, ve);
    if (a.L.scalars[(size_t)p * SC_COUNT + SC_ACTIVE] == 0.0)
        return;
    double s0 = 0.0, s1 = 0.0;
    for (long long i = vb + threadIdx.x; i < ve; i += DOT_FAST_THREADS)
    {
        s0 += (double)x0[i] * (double)y0[i];
        if (x1)
            s1 += (double)x1[i] * (double)y1[i];
    }
    s0 = tree_lanes(s0);
    s1 = tree_lanes(s1);
    const int wave = threadIdx.x / 64;
    if (threadIdx.x % 64 == 0)
    {
        part[wave][0] = s0;
        part[wave][1] = s1;
    }
    __syncthreads();
    if (threadIdx.x < 2)
    {
        double t = 0.0;
        for (int w = 0; w < DOT_FAST_THREADS / 64; w++)
            t += part[w][threadIdx.x];
        a.partial[blockIdx.x * 2 + threadIdx.x] = t;
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// Preconditioner, mat-vec and the CG vector updates (src/cpu/cpu_solve_level1.cpp:476-604, 777-894)
// ---------------------------------------------------------------------------------------------------------------------

__device__ bool invert_block(double* a, int n, double* inv)
{
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            inv[i * n + j] = i == j ? 1.0 : 0.0;
    for (int k = 0; k < n; k++)
    {
        int pivot = k;
        double best = fabs(a[k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(a[i * n + k]) > best)
            {
                best = fabs(a[i * n + k]);
                pivot = i;
            }
        if (!(best > 0.0))
            return false;
        if (pivot != k)
            for (int j = 0; j < n; j++)
            {
                double t = a[k * n + j];
                a[k * n + j] = a[pivot * n + j];
                a[pivot * n + j] = t;
                t = inv[k * n + j];
                inv[k * n + j] = inv[pivot * n + j];
                inv[pivot * n + j] = t;
            }
        const double s = 1.0 / a[k * n + k];
        for (int j = 0; j < n; j++)
        {
            a[k * n + j] *= s;
            inv[k * n + j] *= s;
        }
        for (int i = 0; i < n; i++)
        {
            if (i == k)
                continue;
            const double f = a[i * n + k];
            if (f == 0.0)
                continue;
            for (int j = 0; j < n; j++)
            {
                a[i * n + j] -= f * a[k * n + j];
                inv[i * n + j] -= f * inv[k * n + j];
            }
        }
    }
    return true;
}

template <class V>
__device__ __forceinline__ void apply_precond(const float* minv, const V* r, int c, V* z, long long t)
{
    const float* a = minv + t * (c * c);
    const V* x = r + t * c;
    V* o = z + t * c;
    for (int i = 0; i < c; i++)
    {
        double acc = 0.0;
        for (int j = 0; j < c; j++)
            acc += (double)a[i * c + j] * (double)x[j];
        o[i] = (V)acc;
    }
}

// The fast form's conjugate-gradient vectors r, z, p and Ap are fp32 (x, the solution, stays fp64), held in the first
// half of the fp64 buffers; every product and sum on them is formed in fp64.
template <bool FAST>
using CgVec = typename std::conditional<FAST, float, double>::type;
template <bool FAST>
__device__ __forceinline__ CgVec<FAST>* cg_vec(double* v)
{
    return reinterpret_cast<CgVec<FAST>*>(v);
}

// The start of the iteration, per texel: the block inverse of (H_tt + lambda I) kept as floats (the identity where
// it is singular), delta = 0, r = -grad, z = Minv r, p = z.
template <bool FAST>
__global__ __launch_bounds__(256) void k_cg_start(LatentView L)
{
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= L.texels.start[L.texels.n])
        return;
    const int p = range_plane(L.texels, gt);
    const double lambda = L.scalars[(size_t)p * SC_COUNT + SC_LAMBDA];
    const int c = L.c;
    const long long t = gt - L.texels.start[p];
    const int w = L.pw[p];
    const long long pos = stencil_pos(L.sm, p, w, (int)(t % w), (int)(t / w));
    const long long sn = L.sm.sn;
    double a[MAX_CH * MAX_CH], inv[MAX_CH * MAX_CH];
    // The fast form's gathers build only the fp32 stencil.
    for (int i = 0; i < c; i++)
        for (int j = 0; j < c; j++)
            a[i * c + j] = (FAST ? (double)L.stencil32[(i * c + j) * sn + pos] : L.stencil[(i * c + j) * sn + pos]) +
                           (i == j ? lambda : 0.0);
    if (!invert_block(a, c, inv))
        for (int i = 0; i < c; i++)
            for (int j = 0; j < c; j++)
                inv[i * c + j] = i == j ? 1.0 : 0.0;
    float* o = L.precond + gt * (c * c);
    for (int i = 0; i < c * c; i++)
        o[i] = (float)inv[i];
    CgVec<FAST>* r = cg_vec<FAST>(L.r);
    CgVec<FAST>* z = cg_vec<FAST>(L.z);
    CgVec<FAST>* pv = cg_vec<FAST>(L.p);
    for (int i = 0; i < c; i++)
    {
        L.x[gt * c + i] = 0.0;
        r[gt * c + i] = (CgVec<FAST>)(-L.grad[gt * c + i]);
    }
    apply_precond(L.precond, r, c, z, gt);
    for (int i = 0; i < c; i++)
        pv[gt * c + i] = z[gt * c + i];
}

// y_t = (H_tt + lambda I) x_t + the eight neighbours' blocks, the four the texel does not own read as the transpose of
// the neighbour's; a neighbour outside the plane is absent, not clamped.
template <int C>
__global__ __launch_bounds__(256) void k_matvec(LatentView L)
{
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= L.texels.start[L.texels.n])
        return;
    const int p = range_plane(L.texels, gt);
    const double* s = L.scalars + (size_t)p * SC_COUNT;
    if (s[SC_ACTIVE] == 0.0)
        return;
    const double lambda = s[SC_LAMBDA];
    const long long t = gt - L.texels.start[p];
    const int w = L.pw[p], h = L.ph[p];
    const int tx = (int)(t % w), ty = (int)(t / w);
    const long long base = L.texels.start[p];
    const long long sn = L.sm.sn;
    const double* S = L.stencil;
    const long long pos = stencil_pos(L.sm, p, w, tx, ty);
    const double* x = L.p;
    double acc[C], xt[C];
#pragma unroll
    for (int j = 0; j < C; j++)
        xt[j] = x[gt * C + j];
#pragma unroll
    for (int i = 0; i < C; i++)
    {
        double sacc = lambda * xt[i];
#pragma unroll
        for (int j = 0; j < C; j++)
            sacc += S[(i * C + j) * sn + pos] * xt[j];
        acc[i] = sacc;
    }
#pragma unroll
    for (int slot = 1; slot < STENCIL_BLOCKS; slot++)
    {
        int dx, dy;
        stencil_offset(slot, dx, dy);
        const int fx = tx + dx, fy = ty + dy;
        if (fx >= 0 && fx < w && fy >= 0 && fy < h)
        {
            const double* xn = x + (base + (long long)fy * w + fx) * C;
            double xv[C];
#pragma unroll
            for (int j = 0; j < C; j++)
                xv[j] = xn[j];
#pragma unroll
            for (int i = 0; i < C; i++)
#pragma unroll
                for (int j = 0; j < C; j++)
                    acc[i] += S[(slot * C * C + i * C + j) * sn + pos] * xv[j];
        }
        const int bx = tx - dx, by = ty - dy;
        if (bx >= 0 && bx < w && by >= 0 && by < h)
        {
            const long long npos = stencil_pos(L.sm, p, w, bx, by);
            const double* xn = x + (base + (long long)by * w + bx) * C;
            double xv[C];
#pragma unroll
            for (int j = 0; j < C; j++)
                xv[j] = xn[j];
#pragma unroll
            for (int i = 0; i < C; i++)
#pragma unroll
                for (int j = 0; j < C; j++)
                    acc[i] += S[(slot * C * C + j * C + i) * sn + npos] * xv[j];
        }
    }
#pragma unroll
    for (int i = 0; i < C; i++)
        L.ap[gt * C + i] = acc[i];
}

// Per plane after the curvature fold: p.Ap, then alpha = rz / pAp (zero on non-positive curvature).
__global__ void k_fold_alpha(ChunkArgs a)
{
    const int p = threadIdx.x;
    if (p >= a.chunks.n)
        return;
    double* s = a.L.scalars + (size_t)p * SC_COUNT;
    if (s[SC_ACTIVE] == 0.0)
        return;
    double pap = 0.0;
    for (long long ch = a.chunks.start[p]; ch < a.chunks.start[p + 1]; ch++)
        pap += a.partial[ch * 2];
    s[SC_PAP] = pap;
    s[SC_ALPHA] = pap > 0.0 ? s[SC_RZ] / pap : 0.0;
}

// x += alpha p, r += -alpha Ap, z = Minv r, per texel.
__global__ __launch_bounds__(256) void k_cg_update(LatentView L)
{
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= L.texels.start[L.texels.n])
        return;
    const int p = range_plane(L.texels, gt);
    const double* s = L.scalars + (size_t)p * SC_COUNT;
    if (s[SC_ACTIVE] == 0.0)
        return;
    const double alpha = s[SC_ALPHA];
    const double plus = 1.0, minus = -1.0;
    const int c = L.c;
    for (int i = 0; i < c; i++)
    {
        const long long e = gt * c + i;
        L.x[e] += plus * alpha * L.p[e];
        L.r[e] += minus * alpha * L.ap[e];
    }
    apply_precond(L.precond, L.r, c, L.z, gt);
}

// Per plane: |r|^2 and r.z, beta, the relative residual; at a probe iteration a plane under the bar stops.
__global__ void k_fold_beta(ChunkArgs a, int probe)
{
    const int p = threadIdx.x;
    if (p >= a.chunks.n)
        return;
    double* s = a.L.scalars + (size_t)p * SC_COUNT;
    if (s[SC_ACTIVE] == 0.0)
        return;
    double rr = 0.0, rz_new = 0.0;
    for (long long ch = a.chunks.start[p]; ch < a.chunks.start[p + 1]; ch++)
    {
        rr += a.partial[ch * 2];
        rz_new += a.partial[ch * 2 + 1];
    }
    s[SC_RR] = rr;
    s[SC_RZ_NEW] = rz_new;
    const double rz = s[SC_RZ];
    s[SC_BETA] = rz != 0.0 ? rz_new / rz : 0.0;
    s[SC_RZ] = rz_new;
    const double rhs2 = s[SC_RHS2];
    s[SC_RESID] = rhs2 > 0.0 ? sqrt((rr > 0.0 ? rr : 0.0) / rhs2) : 0.0;
}

// p = z + beta p; then, at a probe iteration, a plane whose residual is under the bar (or NaN) stops.
__global__ __launch_bounds__(256) void k_xpby(LatentView L)
{
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= L.texels.start[L.texels.n])
        return;
    const int p = range_plane(L.texels, gt);
    const double* s = L.scalars + (size_t)p * SC_COUNT;
    if (s[SC_ACTIVE] == 0.0)
        return;
    const double beta = s[SC_BETA];
    for (int i = 0; i < L.c; i++)
    {
        const long long e = gt * L.c + i;
        L.p[e] = L.z[e] + beta * L.p[e];
    }
}

__global__ void k_probe(LatentView L, int nplanes)
{
    const int p = threadIdx.x;
    if (p >= nplanes)
        return;
    double* s = L.scalars + (size_t)p * SC_COUNT;
    if (s[SC_ACTIVE] != 0.0 && !(s[SC_RESID] >= CG_TOLERANCE))
        s[SC_ACTIVE] = 0.0;
}

// The start's two products, folded per plane: |r|^2 -> rhs2, r.z -> rz; every plane becomes active.
__global__ void k_fold_start(ChunkArgs a)
{
    const int p = threadIdx.x;
    if (p >= a.chunks.n)
        return;
    double* s = a.L.scalars + (size_t)p * SC_COUNT;
    double rr = 0.0, rz = 0.0;
    for (long long ch = a.chunks.start[p]; ch < a.chunks.start[p + 1]; ch++)
    {
        rr += a.partial[ch * 2];
        rz += a.partial[ch * 2 + 1];
    }
    s[SC_RHS2] = rr;
    s[SC_RZ] = rz;
}

// ---------------------------------------------------------------------------------------------------------------------
// The correction and the movement probe (src/cpu/cpu_solve_level1.cpp:896-965)
// ---------------------------------------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_add_correction(LatentView L)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.texels.start[L.texels.n] * L.c)
        return;
    L.v[e] = (float)((double)L.v[e] + L.x[e]);
}

// Per chunk: each channel's smallest and largest value (exact in any order).
__global__ __launch_bounds__(64) void k_range_chunks(ChunkArgs a)
{
    int p;
    long long vb, ve;
    chunk_bounds(a, blockIdx.x, p, vb, ve);
    const int c = a.L.c, lane = threadIdx.x;
    double lo[MAX_CH], hi[MAX_CH];
    for (int q = 0; q < MAX_CH; q++)
    {
        lo[q] = 1e300;
        hi[q] = -1e300;
    }
    const int n = (int)(ve - vb);   // vb is a whole texel's first value: value o is channel o % c
    for (int o = lane; o < n; o += 64)
    {
        const int q = o % c;
        const double val = (double)a.L.v[vb + o];
        lo[q] = val < lo[q] ? val : lo[q];
        hi[q] = val > hi[q] ? val : hi[q];
    }
    for (int q = 0; q < c; q++)
        for (int off = 32; off > 0; off >>= 1)
        {
            const double l = __shfl_xor(lo[q], off), h = __shfl_xor(hi[q], off);
            lo[q] = l < lo[q] ? l : lo[q];
            hi[q] = h > hi[q] ? h : hi[q];
        }
    if (lane == 0)
        for (int q = 0; q < MAX_CH; q++)
        {
            a.partial[blockIdx.x * 8 + q] = lo[q];
            a.partial[blockIdx.x * 8 + 4 + q] = hi[q];
        }
}

// One 256-lane workgroup per plane: each channel's extremes over the plane's chunks (exact in any order).
__global__ __launch_bounds__(256) void k_range_fold(ChunkArgs a)
{
    __shared__ double scratch[256];
    const int p = blockIdx.x;
    double* s = a.L.scalars + (size_t)p * SC_COUNT;
    for (int q = 0; q < a.L.c; q++)
    {
        double lo = 1e300, hi = -1e300;
        for (long long ch = a.chunks.start[p] + threadIdx.x; ch < a.chunks.start[p + 1]; ch += 256)
        {
            const double l = a.partial[ch * 8 + q], h = a.partial[ch * 8 + 4 + q];
            lo = l < lo ? l : lo;
            hi = h > hi ? h : hi;
        }
        lo = block_reduce256(lo, 0, scratch);
        hi = block_reduce256(hi, 1, scratch);
        if (threadIdx.x == 0)
        {
            s[SC_LO + q] = lo;
            s[SC_HI + q] = hi;
        }
    }
}

// Per chunk, serially over its values: the count that moved by more than half a grid step of the plane's range, the
// correction's square norm and the plane's.
__global__ __launch_bounds__(64) void k_movement_chunks(ChunkArgs a)
{
    int p;
    long long vb, ve;
    chunk_bounds(a, blockIdx.x, p, vb, ve);
    const int c = a.L.c, lane = threadIdx.x;
    const double* s = a.L.scalars + (size_t)p * SC_COUNT;
    double moved = 0.0, d2 = 0.0, p2 = 0.0;
    if (!a.exact)
    {
        // The fast form: each lane's own sums over the chunk (32-bit offsets within it), one tree at the end.
        double step[MAX_CH];
        for (int q = 0; q < c; q++)
            step[q] = (s[SC_HI + q] - s[SC_LO + q]) / (double)a.L.levels;
        const int n = (int)(ve - vb);
        for (int o = lane; o < n; o += 64)
        {
            const int q = o % c;   // vb is a whole texel's first value
            double st = step[0];
            for (int k = 1; k < c; k++)
                st = q == k ? step[k] : st;
            const double x = a.L.x[vb + o];
            const double v = (double)a.L.v[vb + o];
            moved += fabs(x) > 0.5 * st ? 1.0 : 0.0;
            d2 += x * x;
            p2 += v * v;
        }
        moved = tree_lanes(moved);
        d2 = tree_lanes(d2);
        p2 = tree_lanes(p2);
        if (lane == 0)
        {
            a.partial[blockIdx.x * 3] = moved;
            a.partial[blockIdx.x * 3 + 1] = d2;
            a.partial[blockIdx.x * 3 + 2] = p2;
        }
        return;
    }
    for (long long base = vb; base < ve; base += 64)
    {
        const long long i = base + lane;
        const int count = (int)(ve - base < 64 ? ve - base : 64);
        double m = 0.0, xx = 0.0, vv = 0.0;
        if (lane < count)
        {
            const int q = (int)(i % c);
            const double step = (s[SC_HI + q] - s[SC_LO + q]) / (double)a.L.levels;
            const double x = a.L.x[i];
            m = fabs(x) > 0.5 * step ? 1.0 : 0.0;
            xx = x * x;
            vv = (double)a.L.v[i] * (double)a.L.v[i];
        }
        if (a.exact)
        {
            moved = fold_lanes(moved, m, count);
            d2 = fold_lanes(d2, xx, count);
            p2 = fold_lanes(p2, vv, count);
        }
        else
        {
            moved += tree_lanes(m);
            d2 += tree_lanes(xx);
            p2 += tree_lanes(vv);
        }
    }
    if (lane == 0)
    {
        a.partial[blockIdx.x * 3] = moved;
        a.partial[blockIdx.x * 3 + 1] = d2;
        a.partial[blockIdx.x * 3 + 2] = p2;
    }
}

// One workgroup per plane: the three sums over the plane's chunks, in chunk order from zero (the exact form; lanes 0-2
// take one sum each) or in this kernel's fixed order (the fast form).
__global__ __launch_bounds__(256) void k_movement_fold(ChunkArgs a)
{
    __shared__ double scratch[256];
    const int p = blockIdx.x;
    const long long b = a.chunks.start[p], e = a.chunks.start[p + 1];
    double* s = a.L.scalars + (size_t)p * SC_COUNT;
    double moved = 0.0, d2 = 0.0, p2 = 0.0;
    if (a.exact)
    {
        if (threadIdx.x < 3)
            s[SC_MOVED + threadIdx.x] = fold_strided(a.partial + b * 3 + threadIdx.x, e - b, 3, 0.0);
        return;
    }
    for (long long ch = b + threadIdx.x; ch < e; ch += 256)
    {
        moved += a.partial[ch * 3];
        d2 += a.partial[ch * 3 + 1];
        p2 += a.partial[ch * 3 + 2];
    }
    moved = block_reduce256(moved, 2, scratch);
    d2 = block_reduce256(d2, 2, scratch);
    p2 = block_reduce256(p2, 2, scratch);
    if (threadIdx.x != 0)
        return;
    s[SC_MOVED] = moved;
    s[SC_DELTA2] = d2;
    s[SC_PLANE2] = p2;
}

// ---------------------------------------------------------------------------------------------------------------------
// The quantised sweeps (src/cpu/cpu_solve_level1.cpp:610-674, 988-1054)
// ---------------------------------------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_copy_prev(LatentView L)
{
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L.texels.start[L.texels.n] * L.c)
        return;
    L.prev[e] = L.v[e];
    L.x[e] = 0.0;
}

template <int C, class T>
__device__ __forceinline__ unsigned int sweep_texel(const LatentView& L, const Ranges& colour, int colour_index,
                                                    long long gi, int p);

// The stencil the sweeps and the fast mat-vec read: fp64, or the fast form's fp32 copy.
template <class T>
__device__ __forceinline__ const T* stencil_of(const LatentView& L);
template <>
__device__ __forceinline__ const double* stencil_of<double>(const LatentView& L)
{
    return L.stencil;
}
template <>
__device__ __forceinline__ const float* stencil_of<float>(const LatentView& L)
{
    return L.stencil32;
}

// One colour of one sweep: every texel (tx, ty) of the colour, the neighbours' share of (H delta)_t once, then its
// channels in turn, each one's exact grid argmin using the values the earlier channels have just taken.
template <int C, class T>
__device__ __forceinline__ void quant_sweep(LatentView L, Ranges colour, int colour_index, unsigned int* changed)
{
    const long long gi = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const bool live = gi < colour.start[colour.n];
    const int p = range_plane(colour, live ? gi : colour.start[colour.n] - 1);
    const unsigned int moved = live ? sweep_texel<C, T>(L, colour, colour_index, gi, p) : 0u;
    // One atomic per workgroup: the lanes of the plane of the workgroup's first texel are summed by shuffles and LDS;
    // a lane of another plane (a workgroup across a plane boundary) adds its own. The workgroups spread their adds
    // over MOVED_SLOTS counters per plane on separate cache lines, so they do not queue on one address.
    __shared__ unsigned int s_wave[4];
    const long long g0 = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x;
    const int pw = range_plane(colour, g0);
    unsigned int same = p == pw ? moved : 0u;
    for (int off = 32; off > 0; off >>= 1)
        same += __shfl_xor(same, off);
    if (threadIdx.x % 64 == 0)
        s_wave[threadIdx.x / 64] = same;
    if (p != pw && moved)
        atomicAdd(changed + ((size_t)p * MOVED_SLOTS + blockIdx.x % MOVED_SLOTS) * MOVED_STRIDE, moved);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const unsigned int sum = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        if (sum)
            atomicAdd(changed + ((size_t)pw * MOVED_SLOTS + blockIdx.x % MOVED_SLOTS) * MOVED_STRIDE, sum);
    }
}

template <int C, class T>
__global__ __launch_bounds__(256) void k_quant_sweep(LatentView L, Ranges colour, int colour_index,
                                                     unsigned int* changed)
{
    quant_sweep<C, T>(L, colour, colour_index, changed);
}

// Lane q of a quad's double, to every lane of the quad (DPP quad_perm [q, q, q, q]).
template <int Q>
__device__ __forceinline__ double quad_bcast(double x)
{
    constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// The fast form's sweep at C >= 3 (level 1): four lanes per texel, lane i of the quad summing row i of the neighbours'
// share (the groups in order, the columns in order: row i of the one-lane form's sums), so a colour of the small
// level-1 planes runs four times the wavefronts, each with a quarter of the loads. The channels then move in turn:
// lane q takes channel q and hands its correction to the quad by a DPP broadcast. With C = 3 the fourth lane mirrors
// row 2 and writes nothing.
template <int C>
__global__ __launch_bounds__(256) void k_quant_sweep_rows(LatentView L, Ranges colour, int colour_index,
                                                          unsigned int* changed)
{
    static_assert(C >= 3 && C <= 4, "four lanes hold at most four channel rows");
    const long long gq = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const long long total = colour.start[colour.n];
    const long long gi = gq >> 2;
    const int lane_row = (int)(gq & 3), i = lane_row < C ? lane_row : C - 1;
    // Every lane of a quad reaches the broadcasts: a quad past the end works on the last texel and writes nothing.
    const bool live = gi < total;
    const long long gt_i = live ? gi : total - 1;
    const int p = range_plane(colour, gt_i);
    const long long ci = gt_i - colour.start[p];
    const int w = L.pw[p], h = L.ph[p];
    const int cx = colour_index & 1, cy = colour_index >> 1;
    const int cw1 = (w - cx + 1) / 2;
    const int tx = cx + 2 * (int)(ci % cw1);
    const int ty = cy + 2 * (int)(ci / cw1);
    const long long base = L.texels.start[p];
    const long long t = base + (long long)ty * w + tx;
    const double lambda = L.scalars[(size_t)p * SC_COUNT + SC_LAMBDA];
    const long long sn = L.sm.sn;
    const float* S = L.stencil32;
    const long long pos = L.sm.cpos[p][colour_index] + ci;
    float* delta = reinterpret_cast<float*>(L.x);

    // Row i of each neighbour group's block and the neighbour's corrections, all loads first.
    float sr[8][C], xv[8][C];
    double keep[8];
#pragma unroll
    for (int g = 0; g < 8; g++)
    {
        const int slot = 1 + g / 2;
        int dx, dy;
        stencil_offset(slot, dx, dy);
        if (g % 2 == 0)
        {
            const int fx = tx + dx, fy = ty + dy;
            const bool ok = fx >= 0 && fx < w && fy >= 0 && fy < h;
            keep[g] = ok ? 1.0 : 0.0;
            const float* xn = delta + (ok ? base + (long long)fy * w + fx : t) * C;
#pragma unroll
            for (int j = 0; j < C; j++)
            {
                xv[g][j] = xn[j];
                sr[g][j] = S[(slot * C * C + i * C + j) * sn + pos];
            }
        }
        else
        {
            const int bx = tx - dx, by = ty - dy;
            const bool ok = bx >= 0 && bx < w && by >= 0 && by < h;
            keep[g] = ok ? 1.0 : 0.0;
            const int bxs = ok ? bx : tx, bys = ok ? by : ty;
            const int ncx = bxs & 1, ncy = bys & 1;
            const long long npos = L.sm.cpos[p][ncy * 2 + ncx] + (long long)(bys >> 1) * ((w - ncx + 1) >> 1) + (bxs >> 1);
            const float* xn = delta + (base + (long long)bys * w + bxs) * C;
#pragma unroll
            for (int j = 0; j < C; j++)
            {
                xv[g][j] = xn[j];
                sr[g][j] = S[(slot * C * C + j * C + i) * sn + npos];
            }
        }
    }
    float own[C];
    double dl[C];
#pragma unroll
    for (int j = 0; j < C; j++)
    {
        own[j] = S[(i * C + j) * sn + pos];
        dl[j] = (double)delta[t * C + j];
    }
    const double gr = L.grad[t * C + i];
    const float pv = L.prev[t * C + i];
    const int kold = L.k[t * C + i];
    double nb = 0.0;
#pragma unroll
    for (int g = 0; g < 8; g++)
#pragma unroll
        for (int j = 0; j < C; j++)
        {
            const double term = (double)sr[g][j] * (double)xv[g][j];
            nb += term * keep[g];
        }

    // The channels in order, each seeing the ones before it already moved.
    unsigned int moved = 0;
    const float l = L.lo[i], span = L.hi[i] - L.lo[i];
    auto step = [&](auto qc) {
        constexpr int q = decltype(qc)::value;
        double mine = dl[q];
        if (lane_row == q && live)
        {
            const double aq = (double)own[q] + lambda;
            if (aq > 0.0 && span > 0.0f)
            {
                double rest = nb;
#pragma unroll
                for (int j = 0; j < C; j++)
                    if (j != q)
                        rest += (double)own[j] * dl[j];
                const double bq = gr + rest;
                const double want = (double)pv - bq / aq;
                const int kk = level1_index(l, L.hi[i], L.bits, (float)want);
                const float value = level1_value(l, L.hi[i], L.bits, kk);
                const long long idx = t * C + q;
                if (kold != kk)
                {
                    L.k[idx] = (uint8_t)kk;
                    moved++;
                }
                L.v[idx] = value;
                mine = (double)value - (double)pv;
                delta[idx] = (float)mine;
            }
        }
        dl[q] = quad_bcast<q>(mine);
    };
    step(std::integral_constant<int, 0>());
    step(std::integral_constant<int, 1>());
    step(std::integral_constant<int, 2>());
    if constexpr (C == 4)
        step(std::integral_constant<int, 3>());

    // One atomic per workgroup, as in the one-lane form (a texel's moves are its quad's lanes' moves).
    __shared__ unsigned int s_wave[4];
    const long long g0 = ((long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x) >> 2;
    const int pw = range_plane(colour, g0 < total ? g0 : total - 1);
    unsigned int same = p == pw ? moved : 0u;
    for (int off = 32; off > 0; off >>= 1)
        same += __shfl_xor(same, off);
    if (threadIdx.x % 64 == 0)
        s_wave[threadIdx.x / 64] = same;
    if (p != pw && moved)
        atomicAdd(changed + ((size_t)p * MOVED_SLOTS + blockIdx.x % MOVED_SLOTS) * MOVED_STRIDE, moved);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        const unsigned int sum = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        if (sum)
            atomicAdd(changed + ((size_t)pw * MOVED_SLOTS + blockIdx.x % MOVED_SLOTS) * MOVED_STRIDE, sum);
    }
}

// One neighbour group of a texel's sweep: the neighbour's correction, the stencil block that couples it to the texel
// (entry (i, j) at s[i * C + j], transposed for a backward neighbour, whose block it is) and a 0/1 weight. Every load is
// unconditional: a neighbour outside the plane reads the texel's own entries (a valid address, finite values) and
// weighs its terms by an opaque 0.0 (1.0 inside). x * 1.0 is x, and x * 0.0 is a zero that leaves the sum unchanged
// (the sum starts at +0.0 and so is never -0.0); the weight is opaque so that the compiler cannot turn it back into a
// branch around the loads, which would serialise them.
template <int C, class T = double, class V = double>
struct SweepNeighbour
{
    double x[C], keep;
    T s[C * C];   // the stencil entries as stored (converted where they are used)

    // Group g: slot 1 + g / 2, forward for even g, backward for odd g; V is the neighbours' vector type.
    __device__ __forceinline__ void load(const LatentView& L, const T* S0, long long sn, const V* delta0,
                                         const long long (&cbase)[4], int g, int tx, int ty, int w, int h,
                                         long long base, long long t, long long pos)
    {
        // The addresses hang on an opaque zero made here, so the loads cannot be issued before this point.
        long long z = 0;
        asm volatile("" : "+v"(z));
        const T* S = S0 + z;
        const V* delta = delta0 + z;
        int dx, dy;
        stencil_offset(1 + g / 2, dx, dy);
        const int slot = 1 + g / 2;
        if (g % 2 == 0)
        {
            const int fx = tx + dx, fy = ty + dy;
            const bool ok = fx >= 0 && fx < w && fy >= 0 && fy < h;
            keep = ok ? 1.0 : 0.0;
            asm volatile("" : "+v"(keep));
            const V* xn = delta + (ok ? base + (long long)fy * w + fx : t) * C;
#pragma unroll
            for (int j = 0; j < C; j++)
                x[j] = (double)xn[j];
#pragma unroll
            for (int e = 0; e < C * C; e++)
                s[e] = S[(slot * C * C + e) * sn + pos];
        }
        else
        {
            const int bx = tx - dx, by = ty - dy;
            const bool ok = bx >= 0 && bx < w && by >= 0 && by < h;
            keep = ok ? 1.0 : 0.0;
            asm volatile("" : "+v"(keep));
            const int bxs = ok ? bx : tx, bys = ok ? by : ty;
            const int ncx = bxs & 1, ncy = bys & 1;
            const long long nbase = ncy ? (ncx ? cbase[3] : cbase[2]) : (ncx ? cbase[1] : cbase[0]);
            const long long npos = nbase + (long long)(bys >> 1) * ((w - ncx + 1) >> 1) + (bxs >> 1);
            const V* xn = delta + (base + (long long)bys * w + bxs) * C;
#pragma unroll
            for (int j = 0; j < C; j++)
                x[j] = (double)xn[j];
#pragma unroll
            for (int i = 0; i < C; i++)
#pragma unroll
                for (int j = 0; j < C; j++)
                    s[i * C + j] = S[(slot * C * C + j * C + i) * sn + npos];
        }
    }

    __device__ __forceinline__ void add_to(double (&nb)[C]) const
    {
#pragma unroll
        for (int i = 0; i < C; i++)
#pragma unroll
            for (int j = 0; j < C; j++)
            {
                const double term = (double)s[i * C + j] * x[j];
                nb[i] += term * keep;
            }
    }
};

// One texel of the sweep's colour: returns 1 per channel whose index changed.
template <int C, class T>
__device__ __forceinline__ unsigned int sweep_texel(const LatentView& L, const Ranges& colour, int colour_index,
                                                    long long gi, int p)
{
    const long long ci = gi - colour.start[p];
    const int w = L.pw[p], h = L.ph[p];
    const int cx = colour_index & 1, cy = colour_index >> 1;
    const int cw1 = (w - cx + 1) / 2;
    const int tx = cx + 2 * (int)(ci % cw1);
    const int ty = cy + 2 * (int)(ci / cw1);
    const long long base = L.texels.start[p];
    const long long t = base + (long long)ty * w + tx;
    const double lambda = L.scalars[(size_t)p * SC_COUNT + SC_LAMBDA];
    const long long sn = L.sm.sn;
    const T* S = stencil_of<T>(L);
    const long long pos = L.sm.cpos[p][colour_index] + ci;
    // The corrections in the stencil's type: fp64, or fp32 in the fast form (in the first half of the fp64 buffer,
    // which k_copy_prev has zeroed).
    T* delta = reinterpret_cast<T*>(L.x);

    // The plane's four colour bases in the stencil's runs, read once: a neighbour's position is then arithmetic and
    // selects, not a dependent load (opaque, so the selects are not folded back into an indexed load).
    long long cbase[4];
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
        cbase[k] = L.sm.cpos[p][k];
        asm volatile("" : "+v"(cbase[k]));
    }

    // The eight neighbour groups (slots 1..4, forward then backward) pipelined: group g's loads are issued D groups
    // before its terms are added, so about D groups are in flight per lane - enough, over the wavefronts a SIMD holds,
    // to cover the memory latency, without holding every group's values at once. The order is pinned by data: the
    // loads wait on an opaque value made when the group starts, and the sums are made opaque after each group's
    // terms, and volatile asm statements keep their order. The terms reach each sum slot by slot, forward then
    // backward, channel by channel, as in a loop over the neighbours.
    constexpr int D = C <= 2 ? 4 : 2;
    SweepNeighbour<C, T, T> q[8];
    double nb[C], dl[C], own[C * C], gr[C];
    float pvs[C];
    uint8_t kold[C];
#pragma unroll
    for (int i = 0; i < C; i++)
        nb[i] = 0.0;
#pragma unroll
    for (int g = 0; g < 8 + D; g++)
    {
        if (g < 8)
            q[g].load(L, S, sn, delta, cbase, g, tx, ty, w, h, base, t, pos);
        if (g == 8)
        {
            // The texel's own values, behind the last groups.
#pragma unroll
            for (int j = 0; j < C; j++)
            {
                kold[j] = L.k[t * C + j];
                dl[j] = (double)delta[t * C + j];
                gr[j] = L.grad[t * C + j];
                pvs[j] = L.prev[t * C + j];
            }
#pragma unroll
            for (int e = 0; e < C * C; e++)
                own[e] = (double)S[e * sn + pos];
        }
        if (g >= D)
        {
            q[g - D].add_to(nb);
            // The sums are made opaque here, so the group's terms are added before the next group's loads issue.
#pragma unroll
            for (int i = 0; i < C; i++)
                asm volatile("" : "+v"(nb[i]));
        }
    }

    // The texel's own channels in order, each seeing the ones before it already moved (the deltas kept in registers
    // are the values the reference reads back from its plane).
    unsigned int moved = 0;
#pragma unroll
    for (int q = 0; q < C; q++)
    {
        const long long idx = t * C + q;
        const float l = L.lo[q], span = L.hi[q] - L.lo[q];
        const double aq = own[q * C + q] + lambda;
        if (!(aq > 0.0) || !(span > 0.0f))
            continue;
        double rest = nb[q];
#pragma unroll
        for (int j = 0; j < C; j++)
            if (j != q)
                rest += own[q * C + j] * dl[j];
        const double bq = gr[q] + rest;
        const float pv = pvs[q];
        const double want = (double)pv - bq / aq;
        const int kk = level1_index(l, L.hi[q], L.bits, (float)want);
        const float value = level1_value(l, L.hi[q], L.bits, kk);
        if ((int)kold[q] != kk)
        {
            L.k[idx] = (uint8_t)kk;
            moved++;
        }
        L.v[idx] = value;
        dl[q] = (double)value - (double)pv;
        delta[idx] = (T)dl[q];
    }
    return moved;
}

// ---------------------------------------------------------------------------------------------------------------------
// The fast form's conjugate gradients: the inner products inside the vector launches
// ---------------------------------------------------------------------------------------------------------------------
//
// The mat-vec and the update each produce the iteration's inner products as workgroup partials over chunks of
// FAST_CG_CHUNK texels of one plane, and each fold reads a plane's partials in one fixed order. Where a plane has at
// most 256 chunks (level 1: one L2 load per lane), the consumers fold them themselves: every workgroup of the update
// folds p.Ap into alpha, and every workgroup of p = z + beta p folds |r|^2 and r.z into beta, the residual and, at a
// probe iteration, the stopping bar (the plane's first chunk records the scalars; r.z before the update is kept apart
// from the r.z being written, so no workgroup reads a scalar another rewrites). Where a plane has more chunks, a
// one-workgroup-per-plane launch folds each before its consumer. The two inner-product launches are gone, and with them
// two passes over the vectors. (A plane's last workgroup folding in place would need an agent-scope release and acquire
// per workgroup, which on a chip of eight L2 caches writes back and invalidates them.)

constexpr int FAST_CG_CHUNK = 256;

struct FastCg
{
    LatentView L;
    Ranges chunks;            // per plane: its chunks of FAST_CG_CHUNK texels
    double* partial;          // two per chunk: the update's |r|^2 and r.z
    double* pap;              // one per chunk: the mat-vec's p.Ap
    int probe;                // the fold after the update applies the stopping bar
    double tol;               // the stopping bar: relative residual below tol (the plain and the fast form: 1e-6)
};

// The workgroup's chunk: its plane and its texels [t0, t1) (chain indices).
__device__ __forceinline__ int fast_chunk(const FastCg& a, long long& t0, long long& t1)
{
    const int p = range_plane(a.chunks, xcd_block(blockIdx.x, gridDim.x));
    const long long texels = a.L.texels.start[p + 1] - a.L.texels.start[p];
    const long long b = ((long long)xcd_block(blockIdx.x, gridDim.x) - a.chunks.start[p]) * FAST_CG_CHUNK;
    t0 = a.L.texels.start[p] + b;
    t1 = a.L.texels.start[p] + (b + FAST_CG_CHUNK < texels ? b + FAST_CG_CHUNK : texels);
    return p;
}

// The sum over the 256-lane workgroup: each wave's tree, then the four waves in order (every lane gets the total).
__device__ __forceinline__ double block_sum256(double v, double* s4)
{
    v = tree_lanes(v);
    __syncthreads();
    if (threadIdx.x % 64 == 0)
        s4[threadIdx.x / 64] = v;
    __syncthreads();
    return ((s4[0] + s4[1]) + s4[2]) + s4[3];
}

// Lane t's sum of the values at chunks t, t + 256, ... of [c0, c1) (chunk ch at base[ch * stride]), added in that
// order with four loads in flight.
__device__ __forceinline__ double lane_fold256(const double* base, int stride, long long c0, long long c1)
{
    double v = 0.0;
    long long ch = c0 + threadIdx.x;
    for (; ch + 3 * 256 < c1; ch += 4 * 256)
    {
        const double v0 = base[(size_t)ch * stride], v1 = base[(size_t)(ch + 256) * stride];
        const double v2 = base[(size_t)(ch + 512) * stride], v3 = base[(size_t)(ch + 768) * stride];
        v += v0;
        v += v1;
        v += v2;
        v += v3;
    }
    for (; ch < c1; ch += 256)
        v += base[(size_t)ch * stride];
    return v;
}

// The plane's partials (component k) folded by one workgroup: lane l adds chunks l, l + 256, ... in order, then the
// workgroup's sum - a fixed order, the same every run.
__device__ __forceinline__ double fold_partials(const FastCg& a, int p, int k, double* s4)
{
    return block_sum256(lane_fold256(a.partial + k, 2, a.chunks.start[p], a.chunks.start[p + 1]), s4);
}

// The plane's p.Ap, folded in the same order by the workgroup that asks.
__device__ __forceinline__ double fold_pap(const FastCg& a, int p, double* s4)
{
    return block_sum256(lane_fold256(a.pap, 1, a.chunks.start[p], a.chunks.start[p + 1]), s4);
}

// p = z + beta p on the fast form's fp32 vectors.
__global__ __launch_bounds__(256) void k_xpby_fast(LatentView L)
{
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= L.texels.start[L.texels.n])
        return;
    const int p = range_plane(L.texels, gt);
    const double* s = L.scalars + (size_t)p * SC_COUNT;
    if (s[SC_ACTIVE] == 0.0)
        return;
    const double beta = s[SC_BETA];
    float* pv = cg_vec<true>(L.p);
    const float* z = cg_vec<true>(L.z);
    for (int i = 0; i < L.c; i++)
    {
        const long long e = gt * L.c + i;
        pv[e] = (float)((double)z[e] + beta * (double)pv[e]);
    }
}

// Ap for the chunk's texels (the mat-vec of k_matvec, its neighbour loads pipelined as in the quantised sweep), and the
// chunk's p.Ap.
template <int C>
__global__ __launch_bounds__(256) void k_matvec_fast(FastCg a)
{
    __shared__ double s4[4];
    long long t0, t1;
    const int p = fast_chunk(a, t0, t1);
    const LatentView& L = a.L;
    double* sc = L.scalars + (size_t)p * SC_COUNT;
    if (sc[SC_ACTIVE] == 0.0)
        return;   // the whole workgroup: one plane
    const long long gt = t0 + threadIdx.x;
    double pap = 0.0;
    if (gt < t1)
    {
        const double lambda = sc[SC_LAMBDA];
        const long long t = gt - L.texels.start[p];
        const int w = L.pw[p], h = L.ph[p];
        const int tx = (int)(t % w), ty = (int)(t / w);
        const long long base = L.texels.start[p];
        const long long sn = L.sm.sn;
        const float* S = L.stencil32;   // the fast form's fp32 copy: half the mat-vec's dominant bytes
        const long long pos = stencil_pos(L.sm, p, w, tx, ty);
        long long cbase[4];
#pragma unroll
        for (int k = 0; k < 4; k++)
        {
            cbase[k] = L.sm.cpos[p][k];
            asm volatile("" : "+v"(cbase[k]));
        }
        const float* pf = cg_vec<true>(L.p);
        double xt[C], acc[C];
#pragma unroll
        for (int j = 0; j < C; j++)
            xt[j] = (double)pf[gt * C + j];
#pragma unroll
        for (int i = 0; i < C; i++)
        {
            double sacc = lambda * xt[i];
#pragma unroll
            for (int j = 0; j < C; j++)
                sacc += (double)S[(i * C + j) * sn + pos] * xt[j];
            acc[i] = sacc;
        }
        constexpr int D = C <= 2 ? 4 : 2;
        SweepNeighbour<C, float, float> q[8];
#pragma unroll
        for (int g = 0; g < 8 + D; g++)
        {
            if (g < 8)
                q[g].load(L, S, sn, pf, cbase, g, tx, ty, w, h, base, gt, pos);
            if (g >= D)
            {
                q[g - D].add_to(acc);
#pragma unroll
                for (int i = 0; i < C; i++)
                    asm volatile("" : "+v"(acc[i]));
            }
        }
#pragma unroll
        for (int i = 0; i < C; i++)
        {
            cg_vec<true>(L.ap)[gt * C + i] = (float)acc[i];
            pap += xt[i] * acc[i];
        }
    }
    pap = block_sum256(pap, s4);
    if (threadIdx.x == 0)
        a.pap[xcd_block(blockIdx.x, gridDim.x)] = pap;
}

// The mat-vec at C >= 3 (level 1) with four lanes per texel, lane i of the quad forming row i of Ap (the own block's row,
// then the neighbour groups in order, as the one-lane form orders row i): a 1024-lane workgroup per chunk of 256
// texels, so the chunk partials are laid out as the other fast CG kernels expect.
template <int C>
__global__ __launch_bounds__(1024) void k_matvec_fast_rows(FastCg a)
{
    static_assert(C >= 3 && C <= 4, "four lanes hold at most four rows");
    __shared__ double s16[16];
    long long t0, t1;
    const int p = fast_chunk(a, t0, t1);
    const LatentView& L = a.L;
    double* sc = L.scalars + (size_t)p * SC_COUNT;
    if (sc[SC_ACTIVE] == 0.0)
        return;   // the whole workgroup: one plane
    const long long gt = t0 + (threadIdx.x >> 2);
    const int lane_row = (int)(threadIdx.x & 3), i = lane_row < C ? lane_row : C - 1;
    double pap = 0.0;
    if (gt < t1)
    {
        const double lambda = sc[SC_LAMBDA];
        const long long t = gt - L.texels.start[p];
        const int w = L.pw[p], h = L.ph[p];
        const int tx = (int)(t % w), ty = (int)(t / w);
        const long long base = L.texels.start[p];
        const long long sn = L.sm.sn;
        const float* S = L.stencil32;
        const long long pos = stencil_pos(L.sm, p, w, tx, ty);
        const float* pf = cg_vec<true>(L.p);
        float sr[8][C], xv[8][C];
        double keep[8];
#pragma unroll
        for (int g = 0; g < 8; g++)
        {
            const int slot = 1 + g / 2;
            int dx, dy;
            stencil_offset(slot, dx, dy);
            if (g % 2 == 0)
            {
                const int fx = tx + dx, fy = ty + dy;
                const bool ok = fx >= 0 && fx < w && fy >= 0 && fy < h;
                keep[g] = ok ? 1.0 : 0.0;
                const float* xn = pf + (ok ? base + (long long)fy * w + fx : gt) * C;
#pragma unroll
                for (int j = 0; j < C; j++)
                {
                    xv[g][j] = xn[j];
                    sr[g][j] = S[(slot * C * C + i * C + j) * sn + pos];
                }
            }
            else
            {
                const int bx = tx - dx, by = ty - dy;
                const bool ok = bx >= 0 && bx < w && by >= 0 && by < h;
                keep[g] = ok ? 1.0 : 0.0;
                const int bxs = ok ? bx : tx, bys = ok ? by : ty;
                const int ncx = bxs & 1, ncy = bys & 1;
                const long long npos =
                    L.sm.cpos[p][ncy * 2 + ncx] + (long long)(bys >> 1) * ((w - ncx + 1) >> 1) + (bxs >> 1);
                const float* xn = pf + (base + (long long)bys * w + bxs) * C;
#pragma unroll
                for (int j = 0; j < C; j++)
                {
                    xv[g][j] = xn[j];
                    sr[g][j] = S[(slot * C * C + j * C + i) * sn + npos];
                }
            }
        }
        double xt[C];
        float own[C];
#pragma unroll
        for (int j = 0; j < C; j++)
        {
            xt[j] = (double)pf[gt * C + j];
            own[j] = S[(i * C + j) * sn + pos];
        }
        double acc = lambda * xt[i];
#pragma unroll
        for (int j = 0; j < C; j++)
            acc += (double)own[j] * xt[j];
#pragma unroll
        for (int g = 0; g < 8; g++)
#pragma unroll
            for (int j = 0; j < C; j++)
            {
                const double term = (double)sr[g][j] * (double)xv[g][j];
                acc += term * keep[g];
            }
        if (lane_row < C)
        {
            cg_vec<true>(L.ap)[gt * C + i] = (float)acc;
            pap = xt[i] * acc;
        }
    }
    // The workgroup's sum: each wave's tree, then the sixteen waves in order.
    pap = tree_lanes(pap);
    if (threadIdx.x % 64 == 0)
        s16[threadIdx.x / 64] = pap;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double v = 0.0;
        for (int k = 0; k < 16; k++)
            v += s16[k];
        a.pap[xcd_block(blockIdx.x, gridDim.x)] = v;
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// The matrix-free level-0 operator: (H + lambda I) p from the decode's adjoint instead of the assembled stencil
// ---------------------------------------------------------------------------------------------------------------------
//
// Level 0's system is H = J^T C J over the sites (C = diag(cw)): at a site the decode is affine in the level-0 sample,
// d out_r / d s0 = m_r (adjoint.h's level0_row of the site's level-1 sample), and the sample is the bilinear map of
// its taps. So (H p)_t = sum over the sites s whose taps reach t of w_st y_s, y_s = sum_r cw_r m_r (m_r . p(s)) with
// p(s) the bilinear sample of p at s - the decode linearised at the site (J p), weighted, and taken back through the
// decode's adjoint (J^T), without forming H. One 256-lane workgroup per MF_TW x MF_TH tile of texels: it stages p over
// the tile and a two-texel halo, level 1 over the level-1 texels the halo's sites tap, and the decoder; every site of
// the tile's pixels and their one-pixel halo forms its y_s into LDS; each lane then gathers its texel's 3 x 3 pixels'
// sites in pixel-row, pixel, site order (the assembly's order) and adds lambda p_t. Bytes per launch: p read once
// (about 1.5 times with the halo), Ap written once, level 1 from the cache; the decoder's rows are recomputed per site
// (C0 C1 + 2 C0 + 1 multiply-adds per output), so the kernel is bound by the vector ALUs, not HBM - the assembled
// stencil it replaces is 5 C0^2 fp32 values per texel, read on every iteration. p . Ap is summed per chunk afterwards
// (k_pap_chunks), so the fast conjugate gradients' folds read the partials they read from the assembled mat-vec.
constexpr int MF_TW = 32, MF_TH = 8;
constexpr int MF_PW = MF_TW + 4, MF_PH = MF_TH + 4;              // p: the tile and a two-texel halo
constexpr int MF_SW = MF_TW + 2, MF_SH = MF_TH + 2;              // sites: the tile's pixels and a one-pixel halo
constexpr int MF_L1W = MF_SW + 6, MF_L1H = MF_SH + 6;            // level 1 tapped by those sites (w1 <= w0)

struct MatFreeArgs
{
    FastCg fc;
    Ranges tiles;                // per plane: its tiles
    const float* v1;             // level 1 of the chain
    long long off1[MAX_PLANES];  // each plane's first level-1 texel
    const float* weights;
    const float* cw;
    const double* tab;           // the interior classes' site lists (k_gather_table) for this K
    int nout, nin, k;
};

template <int C0, int C1>
__global__ __launch_bounds__(256) void k_adjoint_matvec0(MatFreeArgs a)
{
    // A lane's sites of the tile's pixels and halo: at most ceil(MF_SW MF_SH 5 / 256).
    constexpr int SPL = (MF_SW * MF_SH * 5 + 255) / 256;
    __shared__ float s_w[NNTC_MAX_NOUT * NNTC_MAX_NIN + NNTC_MAX_NOUT];
    __shared__ float s_p[MF_PW * MF_PH * C0];
    __shared__ float s_l1[MF_L1W * MF_L1H * C1];
    __shared__ float s_y[MF_SW * MF_SH * 5 * C0];
    __shared__ int s_code[GATHER_TAB_MAX];       // the interior classes' site lists (k_gather_table): code, weight
    __shared__ float s_wt[GATHER_TAB_MAX];
    __shared__ int s_len[4];
    const int tile = (int)xcd_block(blockIdx.x, gridDim.x);
    const int p = range_plane(a.tiles, tile);
    const LatentView& L = a.fc.L;
    const double* sc = L.scalars + (size_t)p * SC_COUNT;
    if (sc[SC_ACTIVE] == 0.0)
        return;   // the whole workgroup: one plane
    const PlaneGeom& g = L.planes.p[p];
    const int w0 = g.w0, h0 = g.h0, w1 = g.w1, h1 = g.h1;
    const int tiles_x = (w0 + MF_TW - 1) / MF_TW;
    const int tl = tile - (int)a.tiles.start[p];
    const int tx0 = (tl % tiles_x) * MF_TW, ty0 = (tl / tiles_x) * MF_TH;
    const int tid = threadIdx.x, nout = a.nout, nin = a.nin, per = 1 + a.k;
    const float* pf = cg_vec<true>(L.p) + L.texels.start[p] * C0;

    // ---- stage the decoder, the interior site lists, p (clamped into the plane) and level 1
    for (int e = tid; e < nout * nin; e += 256)
        s_w[e] = a.weights[e];
    if (tid < nout)
        s_w[NNTC_MAX_NOUT * NNTC_MAX_NIN + tid] = a.cw[tid];
    for (int e = tid; e < GATHER_TAB_MAX; e += 256)
    {
        s_code[e] = (int)a.tab[e * GATHER_TAB_STRIDE];
        s_wt[e] = (float)a.tab[e * GATHER_TAB_STRIDE + 1];
    }
    if (tid < 4)
        s_len[tid] = (int)a.tab[GATHER_TAB_MAX * GATHER_TAB_STRIDE + tid];
    for (int e = tid; e < MF_PW * MF_PH * C0; e += 256)
    {
        const int i = e % C0, b = e / C0;
        const int x = clamp_index(tx0 - 2 + b % MF_PW, w0), y = clamp_index(ty0 - 2 + b / MF_PW, h0);
        s_p[e] = pf[((long long)y * w0 + x) * C0 + i];
    }
    int lx0, lx1, ly0, ly1;
    level1_box(tx0 - 1, MF_SW, w0, w1, lx0, lx1);
    level1_box(ty0 - 1, MF_SH, h0, h1, ly0, ly1);
    const int lw = min(lx1 - lx0 + 1, MF_L1W), lh = min(ly1 - ly0 + 1, MF_L1H);
    const float* v1 = a.v1 + a.off1[p] * C1;
    for (int e = tid; e < lw * lh * C1; e += 256)
    {
        const int q = e % C1, b = e / C1;
        s_l1[e] = v1[((long long)(ly0 + b / lw) * w1 + (lx0 + b % lw)) * C1 + q];
    }
    __syncthreads();

    // ---- y_s for every site of the tile's pixels and their halo. Each lane takes sites tid, tid + 256, ...: first
    // their samples of p and of level 1, then the decoder's rows one output at a time (each row read once a lane, for
    // all its sites).
    const int nsites = MF_SW * MF_SH * per;
    float bp[SPL][C0], cbar[SPL][C1], y[SPL][C0];
    bool live[SPL];
#pragma unroll
    for (int u = 0; u < SPL; u++)
    {
        const int q = tid + 256 * u;
        const int j = q % per, hp = q / per;
        const int px = tx0 - 1 + hp % MF_SW, py = ty0 - 1 + hp / MF_SW;
        live[u] = q < nsites && px >= 0 && px < w0 && py >= 0 && py < h0;
#pragma unroll
        for (int i = 0; i < C0; i++)
        {
            y[u][i] = 0.0f;
            bp[u][i] = 0.0f;
        }
#pragma unroll
        for (int c = 0; c < C1; c++)
            cbar[u][c] = 0.0f;
        if (!live[u])
            continue;
        {
            const Tap t0 = site_taps(a.k, j, px, py, w0, h0, w0, h0, true);
            const int xa = t0.x0 - (tx0 - 2), xb = t0.x1 - (tx0 - 2);
            const int ya = t0.y0 - (ty0 - 2), yb = t0.y1 - (ty0 - 2);
            const float w00 = (1.0f - t0.fx) * (1.0f - t0.fy), w10 = t0.fx * (1.0f - t0.fy);
            const float w01 = (1.0f - t0.fx) * t0.fy, w11 = t0.fx * t0.fy;
#pragma unroll
            for (int i = 0; i < C0; i++)
                bp[u][i] = fmaf(w11, s_p[(yb * MF_PW + xb) * C0 + i],
                                fmaf(w01, s_p[(yb * MF_PW + xa) * C0 + i],
                                     fmaf(w10, s_p[(ya * MF_PW + xb) * C0 + i], w00 * s_p[(ya * MF_PW + xa) * C0 + i])));
        }
        {
            const Tap t1 = site_taps(a.k, j, px, py, w0, h0, w1, h1, false);
            const int xa = min(max(t1.x0 - lx0, 0), lw - 1), xb = min(max(t1.x1 - lx0, 0), lw - 1);
            const int ya = min(max(t1.y0 - ly0, 0), lh - 1), yb = min(max(t1.y1 - ly0, 0), lh - 1);
            const float w00 = (1.0f - t1.fx) * (1.0f - t1.fy), w10 = t1.fx * (1.0f - t1.fy);
            const float w01 = (1.0f - t1.fx) * t1.fy, w11 = t1.fx * t1.fy;
#pragma unroll
            for (int c = 0; c < C1; c++)
                cbar[u][c] = fmaf(w11, s_l1[(yb * lw + xb) * C1 + c],
                                  fmaf(w01, s_l1[(yb * lw + xa) * C1 + c],
                                       fmaf(w10, s_l1[(ya * lw + xb) * C1 + c], w00 * s_l1[(ya * lw + xa) * C1 + c])));
        }
    }
    const float* cw = s_w + NNTC_MAX_NOUT * NNTC_MAX_NIN;
    for (int r = 0; r < nout; r++)
    {
        const float* row = s_w + r * nin;
        float rb[C0], rq[C0][C1];
#pragma unroll
        for (int i = 0; i < C0; i++)
        {
            rb[i] = row[C1 + i];
#pragma unroll
            for (int q = 0; q < C1; q++)
                rq[i][q] = row[C1 + C0 + i * C1 + q];
        }
        const float wr = cw[r];
#pragma unroll
        for (int u = 0; u < SPL; u++)
        {
            // d out_r / d s0 at the site (level0_row's arithmetic), then the site's share of J^T C J p.
            float m[C0];
#pragma unroll
            for (int i = 0; i < C0; i++)
            {
                float qq = rb[i];
#pragma unroll
                for (int q = 0; q < C1; q++)
                    qq = fmaf(cbar[u][q], rq[i][q], qq);
                m[i] = qq;
            }
            float d = 0.0f;
#pragma unroll
            for (int i = 0; i < C0; i++)
                d = fmaf(m[i], bp[u][i], d);
            const float t = wr * d;
#pragma unroll
            for (int i = 0; i < C0; i++)
                y[u][i] = fmaf(t, m[i], y[u][i]);
        }
    }
#pragma unroll
    for (int u = 0; u < SPL; u++)
    {
        const int q = tid + 256 * u;
        if (q < MF_SW * MF_SH * 5)
#pragma unroll
            for (int i = 0; i < C0; i++)
                s_y[q * C0 + i] = y[u][i];
    }
    __syncthreads();

    // ---- each lane's texel: its sites' y weighted by their taps' weight there (an interior texel's sites and weights
    // from its class's list, the others from their taps), then lambda p
    const int tx = tx0 + tid % MF_TW, ty = ty0 + tid / MF_TW;
    if (tx >= w0 || ty >= h0)
        return;
    float acc[C0];
    const float lambda = (float)sc[SC_LAMBDA];
#pragma unroll
    for (int i = 0; i < C0; i++)
        acc[i] = lambda * s_p[((ty - ty0 + 2) * MF_PW + (tx - tx0 + 2)) * C0 + i];
    if (tx >= 2 && tx <= w0 - 3 && ty >= 2 && ty <= h0 - 3)
    {
        const int cls = (tx & 1) + 2 * (ty & 1);
        const int n = s_len[cls];
        const int corner = ((ty - ty0) * MF_SW + (tx - tx0)) * per;   // pixel (tx - 1, ty - 1) in the site box
        for (int v = 0; v < n; v++)
        {
            const int code = s_code[cls * 45 + v];
            const float wt = s_wt[cls * 45 + v];
            const int q = corner + ((code >> 6) * MF_SW + ((code >> 3) & 7)) * per + (code & 7);
#pragma unroll
            for (int i = 0; i < C0; i++)
                acc[i] = fmaf(wt, s_y[q * C0 + i], acc[i]);
        }
    }
    else
    {
        int xl, xh, yl, yh;
        footprint_range(tx, w0, w0, xl, xh);
        footprint_range(ty, h0, h0, yl, yh);
        for (int py = yl; py <= yh; py++)
            for (int px = xl; px <= xh; px++)
                for (int j = 0; j < per; j++)
                {
                    const Tap t0 = site_taps(a.k, j, px, py, w0, h0, w0, h0, true);
                    const float wt = tap_weight(tx, t0.x0, t0.x1, t0.fx) * tap_weight(ty, t0.y0, t0.y1, t0.fy);
                    if (wt == 0.0f)
                        continue;
                    const int q = ((py - (ty0 - 1)) * MF_SW + (px - (tx0 - 1))) * per + j;
#pragma unroll
                    for (int i = 0; i < C0; i++)
                        acc[i] = fmaf(wt, s_y[q * C0 + i], acc[i]);
                }
    }
    float* ap = cg_vec<true>(L.ap) + (L.texels.start[p] + (long long)ty * w0 + tx) * C0;
#pragma unroll
    for (int i = 0; i < C0; i++)
        ap[i] = acc[i];
}

// p . Ap per chunk of FAST_CG_CHUNK texels of the active planes, into the partials the fast folds read.
__global__ __launch_bounds__(256) void k_pap_chunks(FastCg a)
{
    __shared__ double s4[4];
    long long t0, t1;
    const int p = fast_chunk(a, t0, t1);
    const LatentView& L = a.L;
    if (L.scalars[(size_t)p * SC_COUNT + SC_ACTIVE] == 0.0)
        return;
    const float* pv = cg_vec<true>(L.p);
    const float* ap = cg_vec<true>(L.ap);
    double v = 0.0;
    for (long long e = t0 * L.c + threadIdx.x; e < t1 * L.c; e += 256)
        v += (double)pv[e] * (double)ap[e];
    v = block_sum256(v, s4);
    if (threadIdx.x == 0)
        a.pap[xcd_block(blockIdx.x, gridDim.x)] = v;
}

// Per plane (one workgroup each): pAp, then alpha = rz / pAp (zero on non-positive curvature).
__global__ __launch_bounds__(256) void k_fold_alpha_fast(FastCg a)
{
    __shared__ double s4[4];
    const int p = blockIdx.x;
    double* sc = a.L.scalars + (size_t)p * SC_COUNT;
    if (sc[SC_ACTIVE] == 0.0)
        return;
    const double pap = fold_pap(a, p, s4);
    if (threadIdx.x == 0)
    {
        sc[SC_PAP] = pap;
        sc[SC_ALPHA] = pap > 0.0 ? sc[SC_RZ] / pap : 0.0;
    }
}

// x += alpha p, r -= alpha Ap, z = Minv r for the chunk's texels, and the chunk's |r|^2 and r.z. With FOLD the
// workgroup first folds the plane's pAp and alpha itself (the plane's first chunk records them), as
// k_fold_alpha_fast does.
template <bool FOLD>
__global__ __launch_bounds__(256) void k_cg_update_fast(FastCg a)
{
    __shared__ double s4[4];
    long long t0, t1;
    const int p = fast_chunk(a, t0, t1);
    const LatentView& L = a.L;
    double* sc = L.scalars + (size_t)p * SC_COUNT;
    if (sc[SC_ACTIVE] == 0.0)
        return;
    double alpha;
    if constexpr (FOLD)
    {
        const double pap = fold_pap(a, p, s4);
        alpha = pap > 0.0 ? sc[SC_RZ] / pap : 0.0;
        if (threadIdx.x == 0 && xcd_block(blockIdx.x, gridDim.x) == a.chunks.start[p])
        {
            sc[SC_PAP] = pap;
            sc[SC_ALPHA] = alpha;
            sc[SC_RZ_PREV] = sc[SC_RZ];
        }
    }
    else
        alpha = sc[SC_ALPHA];
    const long long gt = t0 + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (gt < t1)
    {
        const int c = L.c;
        float* r = cg_vec<true>(L.r);
        float* z = cg_vec<true>(L.z);
        const float* pv = cg_vec<true>(L.p);
        const float* ap = cg_vec<true>(L.ap);
        for (int i = 0; i < c; i++)
        {
            const long long e = gt * c + i;
            L.x[e] += alpha * (double)pv[e];
            r[e] = (float)((double)r[e] - alpha * (double)ap[e]);
        }
        apply_precond(L.precond, r, c, z, gt);
        for (int i = 0; i < c; i++)
        {
            const double ri = r[gt * c + i];
            rr += ri * ri;
            rz += ri * (double)z[gt * c + i];
        }
    }
    rr = block_sum256(rr, s4);
    rz = block_sum256(rz, s4);
    if (threadIdx.x == 0)
    {
        a.partial[(size_t)xcd_block(blockIdx.x, gridDim.x) * 2] = rr;
        a.partial[(size_t)xcd_block(blockIdx.x, gridDim.x) * 2 + 1] = rz;
    }
}

// k_cg_update_fast<false> with one wavefront per chunk and four consecutive texels a lane, so that every vector is
// moved in 16-byte pieces: r, p, Ap and z as float4s, x as double2s, the block inverses as float4s (a lane's four
// texels hold 4 C floats of each fp32 vector, 4 C doubles of x and 4 C^2 floats of inverses, whole numbers of 16-byte
// pieces when the chunk's first texel is a multiple of 4). Each element's arithmetic is k_cg_update_fast's; the
// chunk's |r|^2 and r.z add a lane's texels in order, then the wavefront's lanes by a tree. A chunk that does not start
// on a multiple of 4 (only where a small mip plane has an odd extent) takes the element path.
template <int C>
__global__ __launch_bounds__(256) void k_cg_update_vec(FastCg a)
{
    const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
    const long long chunk = (long long)xcd_block(blockIdx.x, gridDim.x) * 4 + wave;
    if (chunk >= a.chunks.start[a.chunks.n])
        return;   // a whole wavefront: nothing below synchronises beyond it
    const LatentView& L = a.L;
    const int p = range_plane(a.chunks, chunk);
    const double* sc = L.scalars + (size_t)p * SC_COUNT;
    if (sc[SC_ACTIVE] == 0.0)
        return;
    const double alpha = sc[SC_ALPHA];
    const long long texels = L.texels.start[p + 1] - L.texels.start[p];
    const long long off = (chunk - a.chunks.start[p]) * FAST_CG_CHUNK;
    const long long t0 = L.texels.start[p] + off;
    const long long t1 = L.texels.start[p] + (off + FAST_CG_CHUNK < texels ? off + FAST_CG_CHUNK : texels);
    const long long tb = t0 + 4 * lane;
    float* r = cg_vec<true>(L.r);
    float* z = cg_vec<true>(L.z);
    const float* pv = cg_vec<true>(L.p);
    const float* ap = cg_vec<true>(L.ap);
    double rr = 0.0, rz = 0.0;
    if (tb + 4 <= t1 && (t0 & 3) == 0)
    {
        constexpr int NF = 4 * C, NM = 4 * C * C;
        float rv[NF], pp[NF], av[NF], mv[NM], zv[NF];
        double xv[NF];
#pragma unroll
        for (int q = 0; q < C; q++)
        {
            const float4 r4 = reinterpret_cast<const float4*>(r + tb * C)[q];
            const float4 p4 = reinterpret_cast<const float4*>(pv + tb * C)[q];
            const float4 a4 = reinterpret_cast<const float4*>(ap + tb * C)[q];
            rv[4 * q] = r4.x; rv[4 * q + 1] = r4.y; rv[4 * q + 2] = r4.z; rv[4 * q + 3] = r4.w;
            pp[4 * q] = p4.x; pp[4 * q + 1] = p4.y; pp[4 * q + 2] = p4.z; pp[4 * q + 3] = p4.w;
            av[4 * q] = a4.x; av[4 * q + 1] = a4.y; av[4 * q + 2] = a4.z; av[4 * q + 3] = a4.w;
        }
#pragma unroll
        for (int q = 0; q < 2 * C; q++)
        {
            const double2 x2 = reinterpret_cast<const double2*>(L.x + tb * C)[q];
            xv[2 * q] = x2.x;
            xv[2 * q + 1] = x2.y;
        }
#pragma unroll
        for (int q = 0; q < C * C; q++)
        {
            const float4 m4 = reinterpret_cast<const float4*>(L.precond + tb * C * C)[q];
            mv[4 * q] = m4.x; mv[4 * q + 1] = m4.y; mv[4 * q + 2] = m4.z; mv[4 * q + 3] = m4.w;
        }
#pragma unroll
        for (int e = 0; e < NF; e++)
        {
            xv[e] += alpha * (double)pp[e];
            rv[e] = (float)((double)rv[e] - alpha * (double)av[e]);
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
#pragma unroll
            for (int i = 0; i < C; i++)
            {
                double acc = 0.0;
#pragma unroll
                for (int j = 0; j < C; j++)
                    acc += (double)mv[u * C * C + i * C + j] * (double)rv[u * C + j];
                zv[u * C + i] = (float)acc;
            }
#pragma unroll
            for (int i = 0; i < C; i++)
            {
                const double ri = rv[u * C + i];
                rr += ri * ri;
                rz += ri * (double)zv[u * C + i];
            }
        }
#pragma unroll
        for (int q = 0; q < C; q++)
        {
            reinterpret_cast<float4*>(r + tb * C)[q] = make_float4(rv[4 * q], rv[4 * q + 1], rv[4 * q + 2], rv[4 * q + 3]);
            reinterpret_cast<float4*>(z + tb * C)[q] = make_float4(zv[4 * q], zv[4 * q + 1], zv[4 * q + 2], zv[4 * q + 3]);
        }
#pragma unroll
        for (int q = 0; q < 2 * C; q++)
            reinterpret_cast<double2*>(L.x + tb * C)[q] = make_double2(xv[2 * q], xv[2 * q + 1]);
    }
    else
    {
        for (long long gt = tb; gt < tb + 4 && gt < t1; gt++)
        {
            for (int i = 0; i < C; i++)
            {
                const long long e = gt * C + i;
                L.x[e] += alpha * (double)pv[e];
                r[e] = (float)((double)r[e] - alpha * (double)ap[e]);
            }
            apply_precond(L.precond, r, C, z, gt);
            for (int i = 0; i < C; i++)
            {
                const double ri = r[gt * C + i];
                rr += ri * ri;
                rz += ri * (double)z[gt * C + i];
            }
        }
    }
    rr = tree_lanes(rr);
    rz = tree_lanes(rz);
    if (lane == 0)
    {
        a.partial[(size_t)chunk * 2] = rr;
        a.partial[(size_t)chunk * 2 + 1] = rz;
    }
}

// k_xpby_fast over the flat vector, four elements (a float4) a lane; a piece that straddles two planes, or the tail,
// takes the element path. Each element's arithmetic is k_xpby_fast's.
__global__ __launch_bounds__(256) void k_xpby_vec(LatentView L)
{
    const long long n = L.texels.start[L.texels.n] * L.c;
    const long long e0 = ((long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x) * 4;
    if (e0 >= n)
        return;
    float* pv = cg_vec<true>(L.p);
    const float* z = cg_vec<true>(L.z);
    const int pa = range_plane(L.texels, e0 / L.c);
    const int pb = range_plane(L.texels, (e0 + 3 < n ? e0 + 3 : n - 1) / L.c);
    if (pa == pb && e0 + 4 <= n)
    {
        const double* s = L.scalars + (size_t)pa * SC_COUNT;
        if (s[SC_ACTIVE] == 0.0)
            return;
        const double beta = s[SC_BETA];
        const float4 z4 = reinterpret_cast<const float4*>(z)[e0 / 4];
        float4 p4 = reinterpret_cast<float4*>(pv)[e0 / 4];
        p4.x = (float)((double)z4.x + beta * (double)p4.x);
        p4.y = (float)((double)z4.y + beta * (double)p4.y);
        p4.z = (float)((double)z4.z + beta * (double)p4.z);
        p4.w = (float)((double)z4.w + beta * (double)p4.w);
        reinterpret_cast<float4*>(pv)[e0 / 4] = p4;
        return;
    }
    for (long long e = e0; e < e0 + 4 && e < n; e++)
    {
        const double* s = L.scalars + (size_t)range_plane(L.texels, e / L.c) * SC_COUNT;
        if (s[SC_ACTIVE] == 0.0)
            continue;
        pv[e] = (float)((double)z[e] + s[SC_BETA] * (double)pv[e]);
    }
}

// Per plane (one workgroup each): |r|^2, r.z, beta, the relative residual; at a probe iteration a plane under the bar
// (or NaN) stops.
__global__ __launch_bounds__(256) void k_fold_beta_fast(FastCg a)
{
    __shared__ double s4[4];
    const int p = blockIdx.x;
    double* sc = a.L.scalars + (size_t)p * SC_COUNT;
    if (sc[SC_ACTIVE] == 0.0)
        return;
    const double rr_t = fold_partials(a, p, 0, s4);
    const double rz_t = fold_partials(a, p, 1, s4);
    if (threadIdx.x == 0)
    {
        sc[SC_RR] = rr_t;
        sc[SC_RZ_NEW] = rz_t;
        const double rz_old = sc[SC_RZ];
        sc[SC_BETA] = rz_old != 0.0 ? rz_t / rz_old : 0.0;
        sc[SC_RZ] = rz_t;
        const double rhs2 = sc[SC_RHS2];
        sc[SC_RESID] = rhs2 > 0.0 ? sqrt((rr_t > 0.0 ? rr_t : 0.0) / rhs2) : 0.0;
        if (a.probe && !(sc[SC_RESID] >= a.tol))
            sc[SC_ACTIVE] = 0.0;
    }
}

// k_fold_beta_fast then p = z + beta p, for the chunk's texels, with every workgroup folding the plane's |r|^2 and r.z
// itself (the plane's first chunk records the scalars). r.z before the update comes from SC_RZ_PREV, which no
// workgroup of this launch writes. Whether the plane still iterates is read once per workgroup: the first chunk may stop
// the plane during the launch, and every other workgroup then reaches the same stop from the same fold.
__global__ __launch_bounds__(256) void k_xpby_fold_fast(FastCg a)
{
    __shared__ double s4[4];
    __shared__ int s_active;
    long long t0, t1;
    const int p = fast_chunk(a, t0, t1);
    const LatentView& L = a.L;
    double* sc = L.scalars + (size_t)p * SC_COUNT;
    if (threadIdx.x == 0)
        s_active = sc[SC_ACTIVE] != 0.0;
    __syncthreads();
    if (!s_active)
        return;
    const double rr_t = fold_partials(a, p, 0, s4);
    const double rz_t = fold_partials(a, p, 1, s4);
    const double rz_old = sc[SC_RZ_PREV];
    const double beta = rz_old != 0.0 ? rz_t / rz_old : 0.0;
    const double rhs2 = sc[SC_RHS2];
    const double resid = rhs2 > 0.0 ? sqrt((rr_t > 0.0 ? rr_t : 0.0) / rhs2) : 0.0;
    const bool stop = a.probe && !(resid >= a.tol);
    if (threadIdx.x == 0 && xcd_block(blockIdx.x, gridDim.x) == a.chunks.start[p])
    {
        sc[SC_RR] = rr_t;
        sc[SC_RZ_NEW] = rz_t;
        sc[SC_BETA] = beta;
        sc[SC_RZ] = rz_t;
        sc[SC_RESID] = resid;
        if (stop)
            sc[SC_ACTIVE] = 0.0;
    }
    const long long gt = t0 + threadIdx.x;
    if (stop || gt >= t1)
        return;
    float* pv = cg_vec<true>(L.p);
    const float* z = cg_vec<true>(L.z);
    for (int i = 0; i < L.c; i++)
    {
        const long long e = gt * L.c + i;
        pv[e] = (float)((double)z[e] + beta * (double)pv[e]);
    }
}

}   // namespace

// ---------------------------------------------------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------------------------------------------------

static LatentView make_view(Ctx& c, bool level0)
{
    Latent& lt = c.latent(level0);
    LatentView L;
    L.planes = c.planes;
    L.texels.n = c.planes.n;
    long long acc = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        L.texels.start[p] = acc;
        acc += level0 ? c.g(p).n0 : c.g(p).n1;
        L.pw[p] = level0 ? c.g(p).w0 : c.g(p).w1;
        L.ph[p] = level0 ? c.g(p).h0 : c.g(p).h1;
    }
    L.texels.start[c.planes.n] = acc;
    L.level0 = level0;
    L.c = lt.c;
    L.bits = lt.bits;
    L.levels = (1 << lt.bits) - 1;
    L.stencil = lt.stencil.p;
    L.stencil32 = lt.stencil32.p;
    L.sm = stencil_map(c, level0);
    L.grad = lt.grad.p;
    L.precond = lt.precond.p;
    L.x = lt.x.p;
    L.r = lt.r.p;
    L.z = lt.z.p;
    L.p = lt.p.p;
    L.ap = lt.ap.p;
    L.prev = lt.prev.p;
    L.v = level0 ? c.v0.p : c.v1.p;
    L.k = level0 ? c.k0.p : c.k1.p;
    L.lo = c.grid.p + (level0 ? 0 : 8);
    L.hi = c.grid.p + (level0 ? 4 : 12);
    L.scalars = c.scalars.p;
    return L;
}

static ChunkArgs make_chunks(Ctx& c, const LatentView& L)
{
    ChunkArgs a;
    a.L = L;
    a.chunks.n = c.planes.n;
    long long acc = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        a.chunks.start[p] = acc;
        const long long texels = L.texels.start[p + 1] - L.texels.start[p];
        acc += (texels + CHUNK - 1) / CHUNK;
    }
    a.chunks.start[c.planes.n] = acc;
    a.partial = c.partial.p;
    a.exact = c.form == NNTC_FORM_EXACT ? 1 : 0;
    if ((size_t)acc * 8 > c.partial.n)
        throw Refused { "latent: chunk partials exceed the workspace" };
    return a;
}

void build_block_matrices(Ctx& c, bool level0)
{
    const int c0 = c.c0, c1 = c.c1, nout = c.nout, nin = c.nin;
    std::vector<float> cw(c.md.cw, c.md.cw + nout);
    std::vector<double> mono;
    if (!level0)
    {
        // Block (b): A and the sc-blocks M_i of the decoder, G(s) = A + sum_i s_i M_i.
        std::vector<double> a((size_t)nout * c1), mi((size_t)MAX_CH * nout * c1, 0.0);
        for (int r = 0; r < nout; r++)
        {
            const float* row = c.h_w.data() + (size_t)r * nin;
            for (int q = 0; q < c1; q++)
                a[(size_t)r * c1 + q] = (double)row[q];
            for (int i = 0; i < c0; i++)
                for (int q = 0; q < c1; q++)
                    mi[((size_t)i * nout + r) * c1 + q] = (double)row[c1 + c0 + i * c1 + q];
        }
        build_form_matrices(c, a, mi, c0, c1, mono, cw);
    }
    else
    {
        // Block (c'): B and the sc-columns N_q, Q(c) = B + sum_q c_q N_q.
        std::vector<double> b((size_t)nout * c0), nq((size_t)MAX_CH * nout * c0, 0.0);
        for (int r = 0; r < nout; r++)
        {
            const float* row = c.h_w.data() + (size_t)r * nin;
            for (int i = 0; i < c0; i++)
                b[(size_t)r * c0 + i] = (double)row[c1 + i];
            for (int q = 0; q < c1; q++)
                for (int i = 0; i < c0; i++)
                    nq[((size_t)q * nout + r) * c0 + i] = (double)row[c1 + c0 + i * c1 + q];
        }
        build_form_matrices(c, b, nq, c1, c0, mono, cw);
    }
    Latent& lt = c.latent(level0);
    const size_t n = mono.size() < lt.mono.n ? mono.size() : lt.mono.n;
    HIP_OK(hipMemcpyAsync(lt.mono.p, mono.data(), n * sizeof(double), hipMemcpyHostToDevice, c.stream));
}

void assemble(Ctx& c, bool level0, int k, double ridge)
{
    LatentView L = make_view(c, level0);
    AssembleArgs a;
    a.L = L;
    a.src = c.src.p;
    a.v0 = c.v0.p;
    a.v1 = c.v1.p;
    a.weights = c.weights.p;
    a.bias = c.bias.p;
    a.cw = c.cw.p;
    a.mono_p = c.latent(level0).mono.p;
    a.nmono = c.latent(level0).nmono;
    a.c0 = c.c0;
    a.c1 = c.c1;
    a.nin = c.nin;
    a.nout = c.nout;
    a.k = k;
    a.site_f = c.site_f.p;
    a.site_d = c.site_d.p;
    a.site_d32 = reinterpret_cast<float*>(c.site_d.p);
    a.sites.n = c.planes.n;
    c.assembled_k = k;
    a.by_sites = Div31((unsigned int)(1 + k));
    long long acc = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        a.sites.start[p] = acc;
        acc += c.g(p).n0 * (1 + k);
        a.by_w0[p] = Div31((unsigned int)c.g(p).w0);
    }
    a.sites.start[c.planes.n] = acc;
    const int cs = level0 ? c.c1 : c.c0;   // the other latent's channels: the sample each site stores
    // Bytes of a site's pulled-back residual and of a stencil entry: fp64 in the exact form, fp32 in the others.
    const double sd = c.form == NNTC_FORM_EXACT ? 8.0 : 4.0, se = sd;
    // The fast form's level 0 stores each site's own block (its triangle) instead of the sample.
    const bool direct = level0 && c.form != NNTC_FORM_EXACT;
    const int per_site_f = direct ? L.c * (L.c + 1) / 2 : cs;
    if (k > 4 || (size_t)acc * (size_t)per_site_f > c.site_f.n || (size_t)acc * (size_t)L.c > c.site_d.n)
        throw Refused { "assembly: the site buffer holds K <= 4" };
    const long long n = L.texels.start[L.texels.n];
    {
        // Pass 1 reads both latents once and the source (the exact form: the source plane once; the others: each
        // site's entry of the per-site source table) and writes (4 per_site_f + sd c) bytes per site: the sample or
        // the site's own block, and the residual's pull-back in fp64 (the exact form) or fp32.
        const double in = 4.0 * c.n0_total * c.c0 + 4.0 * c.n1_total * c.c1 +
                          (c.form == NNTC_FORM_EXACT ? 4.0 * c.n0_total : 4.0 * (double)acc) * c.nout;
        KernelTimer kt(c.stream, level0 ? "k_site_terms0" : "k_site_terms1",
                       in + (double)acc * (4.0 * per_site_f + sd * L.c), 0.0);
        const dim3 grid(blocks_for(acc, 256));
        a.site_target = c.form != NNTC_FORM_EXACT ? site_targets(c, k) : nullptr;
        if (level0 && direct)
            launch_site_terms<true, true, true>(c.c0, c.c1, grid, c.stream, a);
        else if (level0)
            launch_site_terms<true, false, false>(c.c0, c.c1, grid, c.stream, a);
        else if (c.form != NNTC_FORM_EXACT)
            launch_site_terms<false, false, true>(c.c0, c.c1, grid, c.stream, a);
        else
            launch_site_terms<false, false, false>(c.c0, c.c1, grid, c.stream, a);
    }
    HIP_OK(hipGetLastError());
    {
        // Pass 2 reads every site's terms once per texel it reaches (4 for level 1's sites, about 3.4 for level 0's)
        // and writes the stencil (fp64 in the exact form, the fp32 copy in the others) and the fp64 gradient; the
        // algorithmic traffic counts each site once.
        const double out = n * (se * 5.0 * L.c * L.c + 8.0 * L.c);
        KernelTimer kt(c.stream, level0 ? "k_gather0" : "k_gather1", (double)acc * (4.0 * per_site_f + sd * L.c) + out,
                       0.0);
        // The exact form gathers a texel per lane in the reference's order; the fast form spreads a level-1 texel's
        // footprint (125 sites at K = 4) over eight lanes and level 0's per-site blocks (45 sites) over two.
        if (direct && L.c <= 3)
        {
            Ranges tiles;
            tiles.n = c.planes.n;
            long long nt = 0;
            for (int p = 0; p < c.planes.n; p++)
            {
                tiles.start[p] = nt;
                nt += (long long)((c.g(p).w0 + GT_W - 1) / GT_W) * ((c.g(p).h0 + GT_H - 1) / GT_H);
            }
            tiles.start[c.planes.n] = nt;
            if (c.gather_tab_k != k)
            {
                hipLaunchKernelGGL(k_gather_table, dim3(1), dim3(64), 0, c.stream, k, c.gather_tab.p);
                c.gather_tab_k = k;
            }
            a.tab = c.gather_tab.p;
            if (L.c == 1)
                hipLaunchKernelGGL(k_gather_tile0<1>, dim3((unsigned)nt), dim3(256), 0, c.stream, a, tiles);
            else if (L.c == 2)
                hipLaunchKernelGGL(k_gather_tile0<2>, dim3((unsigned)nt), dim3(256), 0, c.stream, a, tiles);
            else
                hipLaunchKernelGGL(k_gather_tile0<3>, dim3((unsigned)nt), dim3(256), 0, c.stream, a, tiles);
        }
        else if (direct)
        {
            constexpr int G0 = 2;
            const dim3 grid(blocks_for(n * G0, 128));
            switch (L.c)
            {
            case 1: hipLaunchKernelGGL((k_gather_direct<1, G0>), grid, dim3(128), 0, c.stream, a); break;
            case 2: hipLaunchKernelGGL((k_gather_direct<2, G0>), grid, dim3(128), 0, c.stream, a); break;
            case 3: hipLaunchKernelGGL((k_gather_direct<3, G0>), grid, dim3(128), 0, c.stream, a); break;
            default: hipLaunchKernelGGL((k_gather_direct<4, G0>), grid, dim3(128), 0, c.stream, a); break;
            }
        }
        else if (c.form == NNTC_FORM_EXACT)
        {
            if (level0)
                launch_gather<true>(L.c, cs, n, c.stream, a);
            else
                launch_gather<false>(L.c, cs, n, c.stream, a);
        }
        else
        {
            switch (L.c)
            {
            case 1: launch_gather_fast1<1, 8>(cs, n, c.stream, a); break;
            case 2: launch_gather_fast1<2, 8>(cs, n, c.stream, a); break;
            case 3: launch_gather_fast1<3, 8>(cs, n, c.stream, a); break;
            default: launch_gather_fast1<4, 8>(cs, n, c.stream, a); break;
            }
        }
    }
    HIP_OK(hipGetLastError());
    ChunkArgs ch = make_chunks(c, L);
    const long long nch = ch.chunks.start[ch.chunks.n];
    {
        KernelTimer kt(c.stream, "k_diag", 8.0 * (double)n * L.c, 0.0);
        if (ch.exact)
            hipLaunchKernelGGL(k_diag_chunks, dim3((unsigned)nch), dim3(64), 0, c.stream, ch);
        else
            hipLaunchKernelGGL(k_diag_fast, dim3((unsigned)nch), dim3(256), 0, c.stream, ch);
    }
    hipLaunchKernelGGL(k_diag_fold, dim3((unsigned)c.planes.n), dim3(256), 0, c.stream, ch, ridge);
    HIP_OK(hipGetLastError());
}

namespace
{

// Every plane's conjugate gradients marked active.
__global__ __launch_bounds__(64) void k_mark_active(double* scalars, int np)
{
    if ((int)threadIdx.x < np)
        scalars[(size_t)threadIdx.x * SC_COUNT + SC_ACTIVE] = 1.0;
}

}   // namespace

namespace
{

template <int C0>
void launch_adjoint_matvec0_c1(int c1, unsigned grid, hipStream_t s, const MatFreeArgs& a)
{
    switch (c1)
    {
    case 1: hipLaunchKernelGGL((k_adjoint_matvec0<C0, 1>), dim3(grid), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_adjoint_matvec0<C0, 2>), dim3(grid), dim3(256), 0, s, a); break;
    case 3: hipLaunchKernelGGL((k_adjoint_matvec0<C0, 3>), dim3(grid), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_adjoint_matvec0<C0, 4>), dim3(grid), dim3(256), 0, s, a); break;
    }
}

void launch_adjoint_matvec0(int c0, int c1, unsigned grid, hipStream_t s, const MatFreeArgs& a)
{
    switch (c0)
    {
    case 1: launch_adjoint_matvec0_c1<1>(c1, grid, s, a); break;
    case 2: launch_adjoint_matvec0_c1<2>(c1, grid, s, a); break;
    case 3: launch_adjoint_matvec0_c1<3>(c1, grid, s, a); break;
    default: launch_adjoint_matvec0_c1<4>(c1, grid, s, a); break;
    }
}

// Whether the fast form's level-0 conjugate gradients run on the matrix-free operator: by default on a base plane of
// 4096^2 texels and more, where the assembled stencil is several GB read on every iteration.
bool matrix_free_on(const Ctx& c)
{
    if (c.form != NNTC_FORM_FAST)
        return false;
    if (c.matrix_free >= 0)
        return c.matrix_free != 0;
    return (long long)c.g(0).w0 * c.g(0).h0 >= 4096ll * 4096ll;
}

// The fast conjugate gradients' launch arguments: the chunks of FAST_CG_CHUNK texels per plane and their partials.
FastCg make_fast_cg(Ctx& c, const LatentView& L)
{
    const int np = c.planes.n;
    FastCg fc {};
    fc.L = L;
    fc.tol = CG_TOLERANCE_FP32;
    fc.chunks.n = np;
    long long acc = 0;
    for (int p = 0; p < np; p++)
    {
        fc.chunks.start[p] = acc;
        acc += (L.texels.start[p + 1] - L.texels.start[p] + FAST_CG_CHUNK - 1) / FAST_CG_CHUNK;
    }
    fc.chunks.start[np] = acc;
    if ((size_t)acc * 3 > c.cg_partial.n)
        throw Refused { "latent: CG partials exceed the workspace" };
    fc.partial = c.cg_partial.p;
    fc.pap = c.cg_partial.p + 2 * acc;
    return fc;
}

MatFreeArgs matfree_args(Ctx& c, const FastCg& fc, unsigned& grid)
{
    MatFreeArgs a {};
    a.fc = fc;
    a.tiles.n = c.planes.n;
    long long nt = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        a.tiles.start[p] = nt;
        nt += (long long)((c.g(p).w0 + MF_TW - 1) / MF_TW) * ((c.g(p).h0 + MF_TH - 1) / MF_TH);
        a.off1[p] = c.g(p).off1;
        if (c.g(p).w1 > c.g(p).w0 || c.g(p).h1 > c.g(p).h0)
            throw Refused { "matrix-free operator: level 1 larger than level 0" };
    }
    a.tiles.start[c.planes.n] = nt;
    a.v1 = c.v1.p;
    a.weights = c.weights.p;
    a.cw = c.cw.p;
    a.nout = c.nout;
    a.nin = c.nin;
    a.k = c.assembled_k;
    if (a.k < 0 || a.k > 4)
        throw Refused { "matrix-free operator: the system was not assembled" };
    if (c.gather_tab_k != a.k)
    {
        hipLaunchKernelGGL(k_gather_table, dim3(1), dim3(64), 0, c.stream, a.k, c.gather_tab.p);
        c.gather_tab_k = a.k;
    }
    a.tab = c.gather_tab.p;
    grid = (unsigned)nt;
    return a;
}

// The launch arguments a captured graph was made from, as bytes (values and buffer addresses).
struct GraphKey
{
    std::vector<unsigned char> bytes;
    template <class T>
    GraphKey& add(const T& v)
    {
        const unsigned char* b = reinterpret_cast<const unsigned char*>(&v);
        bytes.insert(bytes.end(), b, b + sizeof(T));
        return *this;
    }
};

// The context's graph for `key`, captured from `body`'s launches on the context's stream when it has none yet. The
// capture is thread-local, so contexts driven by other host threads keep launching while one captures.
hipGraphExec_t period_graph(Ctx& c, const std::vector<unsigned char>& key, const std::function<void()>& body)
{
    for (const Ctx::CgGraph& g : c.cg_graphs)
        if (g.key == key)
            return g.exec;
    hipGraph_t graph = nullptr;
    HIP_OK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try
    {
        body();
    }
    catch (...)
    {
        if (hipStreamEndCapture(c.stream, &graph) == hipSuccess && graph)
            (void)hipGraphDestroy(graph);
        throw;
    }
    HIP_OK(hipStreamEndCapture(c.stream, &graph));
    hipGraphExec_t exec = nullptr;
    const hipError_t made = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_OK(made);
    c.cg_graphs.push_back(Ctx::CgGraph { key, exec });
    c.graph_captures++;
    return exec;
}

}   // namespace

void cg_solve(Ctx& c, bool level0, std::vector<int>& iterations)
{
    LatentView L = make_view(c, level0);
    ChunkArgs ch = make_chunks(c, L);
    const long long n = L.texels.start[L.texels.n];
    const long long nch = ch.chunks.start[ch.chunks.n];
    const int np = c.planes.n;
    const unsigned tb = blocks_for(n, 256);
    // Every plane starts active.
    hipLaunchKernelGGL(k_mark_active, dim3(1), dim3(64), 0, c.stream, c.scalars.p, np);
    HIP_OK(hipGetLastError());
    // Two inner products per chunk, into the chunk partials: one wavefront per chunk, folded serially in index order
    // (the exact form) or by a wavefront tree (the fast form).
    auto dot = [&](const double* x0, const double* y0, const double* x1, const double* y1, double bytes) {
        KernelTimer kt(c.stream, level0 ? "k_dot0" : "k_dot1", bytes, 0.0);
        if (ch.exact)
            hipLaunchKernelGGL(k_dot_chunks, dim3((unsigned)nch), dim3(64), 0, c.stream, ch, x0, y0, x1, y1);
        else
            hipLaunchKernelGGL(k_dot_fast<double>, dim3((unsigned)nch), dim3(DOT_FAST_THREADS), 0, c.stream, ch, x0,
                               y0, x1, y1);
    };
    if (c.form == NNTC_FORM_EXACT)
    {
        hipLaunchKernelGGL(k_cg_start<false>, dim3(tb), dim3(256), 0, c.stream, L);
        dot(L.r, L.r, L.r, L.z, 16.0 * (double)n * L.c);
    }
    else
    {
        // The fast form's r and z are fp32 (CgVec).
        hipLaunchKernelGGL(k_cg_start<true>, dim3(tb), dim3(256), 0, c.stream, L);
        KernelTimer kt(c.stream, level0 ? "k_dot0" : "k_dot1", 8.0 * (double)n * L.c, 0.0);
        const float *r = reinterpret_cast<const float*>(L.r), *z = reinterpret_cast<const float*>(L.z);
        hipLaunchKernelGGL(k_dot_fast<float>, dim3((unsigned)nch), dim3(DOT_FAST_THREADS), 0, c.stream, ch, r, r, r, z);
    }
    hipLaunchKernelGGL(k_fold_start, dim3(1), dim3(64), 0, c.stream, ch);
    HIP_OK(hipGetLastError());

    iterations.assign((size_t)np, CG_MAX_ITERATIONS);
    std::vector<int> done((size_t)np, 0);
    // Algorithmic bytes per launch at full activity: the stencil (40 c^2 per texel) and the vectors each kernel streams.
    const double cc = L.c, nv = (double)n * cc;
    const double b_matvec = (double)n * 40.0 * cc * cc + 16.0 * nv;
    const double b_update = (double)n * 4.0 * cc * cc + 56.0 * nv;
    // The fast form: the inner products inside the mat-vec and the update, folded per plane by a workgroup each.
    FastCg fc = make_fast_cg(c, L);
    long long most_chunks = 0;
    for (int p = 0; p < np; p++)
        most_chunks = std::max(most_chunks, fc.chunks.start[p + 1] - fc.chunks.start[p]);
    const bool fold_in_update = most_chunks <= 256;
    fc.partial = c.cg_partial.p;
    const unsigned fgrid = (unsigned)fc.chunks.start[np];
    // The fast form's level 0 on the matrix-free operator (the decode's adjoint) in place of the assembled stencil.
    const bool mf = level0 && !ch.exact && matrix_free_on(c);
    MatFreeArgs mfa {};
    unsigned mf_grid = 0;
    double mf_bytes = 0.0, mf_flops = 0.0;
    if (mf)
    {
        mfa = matfree_args(c, fc, mf_grid);
        // p read (with the halo, about 1.5 times) and Ap written; the flops of y_s per site and the gather per texel.
        const double sites = (double)c.n0_total * (1 + mfa.k);
        mf_bytes = 8.0 * nv;
        mf_flops = sites * c.nout * 2.0 * (c.c0 * c.c1 + 2.0 * c.c0 + 1.0) + sites * 3.4 * 2.0 * c.c0;
    }
    // One iteration of the fast form's conjugate gradients: the operator (and p.Ap), alpha, the update (with |r|^2 and
    // r.z), beta and p = z + beta p; the fold after the update applies the stopping bar on every CG_PROBE-th.
    auto fast_iteration = [&](int it) {
        fc.probe = (it + 1) % CG_PROBE == 0;
        // The fast form's bytes: the fp32 stencil (20 c^2 per texel), p read and Ap written (4 c each); the update
        // reads the preconditioner, x, p, r, Ap and writes x, r, z; p = z + beta p reads two and writes one.
        if (mf)
        {
            {
                KernelTimer kt(c.stream, "k_adjoint_matvec0", mf_bytes, mf_flops);
                launch_adjoint_matvec0(c.c0, c.c1, mf_grid, c.stream, mfa);
            }
            hipLaunchKernelGGL(k_pap_chunks, dim3(fgrid), dim3(256), 0, c.stream, fc);
        }
        else
        {
            KernelTimer kt(c.stream, level0 ? "k_matvec0" : "k_matvec1", (double)n * 20.0 * cc * cc + 8.0 * nv,
                           (double)n * (2.0 * 9.0 * cc * cc));
            switch (L.c)
            {
            case 1: hipLaunchKernelGGL(k_matvec_fast<1>, dim3(fgrid), dim3(256), 0, c.stream, fc); break;
            case 2: hipLaunchKernelGGL(k_matvec_fast<2>, dim3(fgrid), dim3(256), 0, c.stream, fc); break;
            case 3: hipLaunchKernelGGL(k_matvec_fast_rows<3>, dim3(fgrid), dim3(1024), 0, c.stream, fc); break;
            default: hipLaunchKernelGGL(k_matvec_fast_rows<4>, dim3(fgrid), dim3(1024), 0, c.stream, fc); break;
            }
        }
        if (!fold_in_update)
            hipLaunchKernelGGL(k_fold_alpha_fast, dim3(np), dim3(256), 0, c.stream, fc);
        {
            // Per element: x (fp64) read and written, p and Ap read, r read and written, z written (36 bytes); per
            // texel its C x C block of the preconditioner.
            KernelTimer kt(c.stream, level0 ? "k_cg_update0" : "k_cg_update1", (double)n * 4.0 * cc * cc + 36.0 * nv,
                           nv * (4.0 + 2.0 * cc));
            if (fold_in_update)
                hipLaunchKernelGGL(k_cg_update_fast<true>, dim3(fgrid), dim3(256), 0, c.stream, fc);
            else
            {
                // One wavefront per chunk, four chunks a workgroup.
                const dim3 vgrid((fgrid + 3) / 4);
                switch (L.c)
                {
                case 1: hipLaunchKernelGGL(k_cg_update_vec<1>, vgrid, dim3(256), 0, c.stream, fc); break;
                case 2: hipLaunchKernelGGL(k_cg_update_vec<2>, vgrid, dim3(256), 0, c.stream, fc); break;
                case 3: hipLaunchKernelGGL(k_cg_update_vec<3>, vgrid, dim3(256), 0, c.stream, fc); break;
                default: hipLaunchKernelGGL(k_cg_update_vec<4>, vgrid, dim3(256), 0, c.stream, fc); break;
                }
            }
        }
        if (!fold_in_update)
            hipLaunchKernelGGL(k_fold_beta_fast, dim3(np), dim3(256), 0, c.stream, fc);
        {
            KernelTimer kt(c.stream, level0 ? "k_xpby0" : "k_xpby1", 12.0 * nv, 2.0 * nv);
            if (fold_in_update)
                hipLaunchKernelGGL(k_xpby_fold_fast, dim3(fgrid), dim3(256), 0, c.stream, fc);
            else
                hipLaunchKernelGGL(k_xpby_vec, dim3(blocks_for((n * L.c + 3) / 4, 256)), dim3(256), 0, c.stream, L);
        }
        HIP_OK(hipGetLastError());
    };
    // The fast form runs each probe period - CG_PROBE iterations, the same launches on the same buffers every time -
    // as one HIP graph, captured on the first call that meets these launch arguments and replayed by every later
    // period and call. Not while the profiler times one of its kernels (its events are host-side per launch).
    hipGraphExec_t period = nullptr;
    if (c.form == NNTC_FORM_FAST && c.graphs != 0 && !profiled(mf ? "k_adjoint_matvec0" : (level0 ? "k_matvec0" : "k_matvec1")) &&
        !profiled(level0 ? "k_cg_update0" : "k_cg_update1") && !profiled(level0 ? "k_xpby0" : "k_xpby1"))
    {
        GraphKey key;
        key.add(level0).add(mf).add(fold_in_update).add(fgrid).add(tb).add(mf_grid).add(mfa.k).add(np);
        key.add(L.x).add(L.r).add(L.z).add(L.p).add(L.ap).add(L.precond).add(L.stencil32).add(L.scalars);
        key.add(fc.partial).add(fc.pap).add(mfa.tab).add(mfa.v1).add(mfa.weights).add(mfa.cw);
        for (int p = 0; p <= np; p++)
            key.add(L.texels.start[p]).add(fc.chunks.start[p]);
        period = period_graph(c, key.bytes, [&] {
            for (int i = 0; i < CG_PROBE; i++)
                fast_iteration(i);
        });
    }
    for (int it = 0; it < CG_MAX_ITERATIONS; it++)
    {
        if (!ch.exact)
        {
            if (period)
            {
                HIP_OK(hipGraphLaunch(period, c.stream));
                c.graph_launches++;
                it += CG_PROBE - 1;
            }
            else
                fast_iteration(it);
        }
        else
        {
            {
                KernelTimer kt(c.stream, level0 ? "k_matvec0" : "k_matvec1", b_matvec, (double)n * (2.0 * 9.0 * cc * cc));
                switch (L.c)
                {
                case 1: hipLaunchKernelGGL(k_matvec<1>, dim3(tb), dim3(256), 0, c.stream, L); break;
                case 2: hipLaunchKernelGGL(k_matvec<2>, dim3(tb), dim3(256), 0, c.stream, L); break;
                case 3: hipLaunchKernelGGL(k_matvec<3>, dim3(tb), dim3(256), 0, c.stream, L); break;
                default: hipLaunchKernelGGL(k_matvec<4>, dim3(tb), dim3(256), 0, c.stream, L); break;
                }
            }
            dot(L.p, L.ap, nullptr, nullptr, 16.0 * nv);
            hipLaunchKernelGGL(k_fold_alpha, dim3(1), dim3(64), 0, c.stream, ch);
            {
                KernelTimer kt(c.stream, level0 ? "k_cg_update0" : "k_cg_update1", b_update, nv * (4.0 + 2.0 * cc));
                hipLaunchKernelGGL(k_cg_update, dim3(tb), dim3(256), 0, c.stream, L);
            }
            dot(L.r, L.r, L.r, L.z, 16.0 * nv);
            hipLaunchKernelGGL(k_fold_beta, dim3(1), dim3(64), 0, c.stream, ch, 0);
            {
                KernelTimer kt(c.stream, level0 ? "k_xpby0" : "k_xpby1", 24.0 * nv, 2.0 * nv);
                hipLaunchKernelGGL(k_xpby, dim3(tb), dim3(256), 0, c.stream, L);
            }
            HIP_OK(hipGetLastError());
        }
        if ((it + 1) % CG_PROBE != 0)
            continue;
        if (ch.exact)
            hipLaunchKernelGGL(k_probe, dim3(1), dim3(64), 0, c.stream, L, np);
        HIP_OK(hipGetLastError());
        HIP_OK(hipMemcpyAsync(c.host_scalars, c.scalars.p, sizeof(double) * SC_COUNT * np, hipMemcpyDeviceToHost,
                              c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
        bool any = false;
        for (int p = 0; p < np; p++)
        {
            if (!done[(size_t)p] && c.host_scalars[(size_t)p * SC_COUNT + SC_ACTIVE] == 0.0)
            {
                done[(size_t)p] = 1;
                iterations[(size_t)p] = it + 1;
            }
            any = any || !done[(size_t)p];
        }
        if (!any)
            break;
    }
}

void level0_apply(Ctx& c, int mode, const double* p_host, double* out_host)
{
    LatentView L = make_view(c, true);
    const long long n = L.texels.start[L.texels.n];
    const size_t nv = (size_t)n * L.c;
    const int np = c.planes.n;
    hipLaunchKernelGGL(k_mark_active, dim3(1), dim3(64), 0, c.stream, c.scalars.p, np);
    const bool fast = c.form != NNTC_FORM_EXACT;
    if (mode == 1 && c.form != NNTC_FORM_FAST)
        throw Refused { "level0_apply: the matrix-free operator is the fast form's" };
    if (fast)
    {
        std::vector<float> pf(nv);
        for (size_t i = 0; i < nv; i++)
            pf[i] = (float)p_host[i];
        upload_staged(c, reinterpret_cast<float*>(L.p), pf.data(), nv);
        FastCg fc = make_fast_cg(c, L);
        const unsigned fgrid = (unsigned)fc.chunks.start[np];
        if (mode == 1)
        {
            unsigned grid = 0;
            MatFreeArgs a = matfree_args(c, fc, grid);
            launch_adjoint_matvec0(c.c0, c.c1, grid, c.stream, a);
        }
        else
            switch (L.c)
            {
            case 1: hipLaunchKernelGGL(k_matvec_fast<1>, dim3(fgrid), dim3(256), 0, c.stream, fc); break;
            case 2: hipLaunchKernelGGL(k_matvec_fast<2>, dim3(fgrid), dim3(256), 0, c.stream, fc); break;
            case 3: hipLaunchKernelGGL(k_matvec_fast_rows<3>, dim3(fgrid), dim3(1024), 0, c.stream, fc); break;
            default: hipLaunchKernelGGL(k_matvec_fast_rows<4>, dim3(fgrid), dim3(1024), 0, c.stream, fc); break;
            }
        HIP_OK(hipGetLastError());
        const float* h = read_back(c, (const float*)reinterpret_cast<float*>(L.ap), nv);
        for (size_t i = 0; i < nv; i++)
            out_host[i] = (double)h[i];
        return;
    }
    upload_staged(c, L.p, p_host, nv);
    const unsigned tb = blocks_for(n, 256);
    switch (L.c)
    {
    case 1: hipLaunchKernelGGL(k_matvec<1>, dim3(tb), dim3(256), 0, c.stream, L); break;
    case 2: hipLaunchKernelGGL(k_matvec<2>, dim3(tb), dim3(256), 0, c.stream, L); break;
    case 3: hipLaunchKernelGGL(k_matvec<3>, dim3(tb), dim3(256), 0, c.stream, L); break;
    default: hipLaunchKernelGGL(k_matvec<4>, dim3(tb), dim3(256), 0, c.stream, L); break;
    }
    HIP_OK(hipGetLastError());
    const double* h = read_back(c, (const double*)L.ap, nv);
    for (size_t i = 0; i < nv; i++)
        out_host[i] = h[i];
}

void correction_and_probe(Ctx& c, bool level0)
{
    LatentView L = make_view(c, level0);
    ChunkArgs ch = make_chunks(c, L);
    const long long n = L.texels.start[L.texels.n];
    const long long nch = ch.chunks.start[ch.chunks.n];
    hipLaunchKernelGGL(k_add_correction, dim3(blocks_for(n * L.c, 256)), dim3(256), 0, c.stream, L);
    hipLaunchKernelGGL(k_range_chunks, dim3((unsigned)nch), dim3(64), 0, c.stream, ch);
    hipLaunchKernelGGL(k_range_fold, dim3((unsigned)c.planes.n), dim3(256), 0, c.stream, ch);
    hipLaunchKernelGGL(k_movement_chunks, dim3((unsigned)nch), dim3(64), 0, c.stream, ch);
    hipLaunchKernelGGL(k_movement_fold, dim3((unsigned)c.planes.n), dim3(256), 0, c.stream, ch);
    HIP_OK(hipGetLastError());
}

void copy_prev_zero_delta(Ctx& c, bool level0)
{
    LatentView L = make_view(c, level0);
    const long long n = L.texels.start[L.texels.n];
    hipLaunchKernelGGL(k_copy_prev, dim3(blocks_for(n * L.c, 256)), dim3(256), 0, c.stream, L);
    HIP_OK(hipGetLastError());
}

template <class T>
static void launch_sweep(int ch, dim3 grid, hipStream_t s, const LatentView& L, const Ranges& colour, int col,
                         unsigned int* counters)
{
    switch (ch)
    {
    case 1: hipLaunchKernelGGL((k_quant_sweep<1, T>), grid, dim3(256), 0, s, L, colour, col, counters); break;
    case 2: hipLaunchKernelGGL((k_quant_sweep<2, T>), grid, dim3(256), 0, s, L, colour, col, counters); break;
    case 3: hipLaunchKernelGGL((k_quant_sweep<3, T>), grid, dim3(256), 0, s, L, colour, col, counters); break;
    default: hipLaunchKernelGGL((k_quant_sweep<4, T>), grid, dim3(256), 0, s, L, colour, col, counters); break;
    }
}

const float* site_targets(Ctx& c, int k)
{
    AssembleArgs a {};
    a.L.planes = c.planes;
    a.nout = c.nout;
    a.k = k;
    a.sites.n = c.planes.n;
    a.by_sites = Div31((unsigned int)(1 + k));
    long long acc = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        a.sites.start[p] = acc;
        acc += c.g(p).n0 * (1 + k);
        a.by_w0[p] = Div31((unsigned int)c.g(p).w0);
    }
    a.sites.start[c.planes.n] = acc;
    if (c.site_target.n < (size_t)acc * c.nout)
    {
        c.site_target.alloc((size_t)acc * c.nout);
        c.site_target_k = -1;
    }
    if (c.site_target_k != k)
    {
        hipLaunchKernelGGL(k_site_target, dim3(blocks_for(acc, 256)), dim3(256), 0, c.stream, a, c.src.p,
                           c.site_target.p);
        HIP_OK(hipGetLastError());
        c.site_target_k = k;
    }
    return c.site_target.p;
}

// One colour of one sweep in the fast form at C >= 3: one lane a texel, or four (one channel row each). The two forms
// take the same decisions from the same sums and write the same bits.
static void launch_sweep_fast(int lanes, long long n, hipStream_t s, const LatentView& L, const Ranges& colour, int col,
                              unsigned int* counters)
{
    if (lanes == 1)
    {
        launch_sweep<float>(L.c, dim3(blocks_for(n, 256)), s, L, colour, col, counters);
        return;
    }
    const dim3 grid4(blocks_for(4 * n, 256));
    if (L.c == 3)
        hipLaunchKernelGGL(k_quant_sweep_rows<3>, grid4, dim3(256), 0, s, L, colour, col, counters);
    else
        hipLaunchKernelGGL(k_quant_sweep_rows<4>, grid4, dim3(256), 0, s, L, colour, col, counters);
}

// The lanes a texel of the fast sweeps at C >= 3, from the launch-shape tuner: the first call of a shape (latent,
// channels, texel count to a power of two) times one sweep of each form on the call's own state, restoring the values,
// indices and corrections it moves after each trial, and every later call of that shape on the device takes the
// faster. The forms write the same bits, so the choice changes the time and nothing else.
static int sweep_lanes_for(Ctx& c, const LatentView& L, const Ranges colour[4])
{
    if (c.sweep_lanes > 0)
        return c.sweep_lanes;
    const long long n = L.texels.start[L.texels.n];
    const TuneKey key { c.device, TUNE_SWEEP_LANES, L.c, ilog2_ceil(n) };
    const int known = tuned_choice(key);
    if (known > 0)
        return known;
    const size_t nv = (size_t)n * L.c;
    const size_t bytes = nv * (sizeof(float) + sizeof(uint8_t) + sizeof(double));
    DevBuf<uint8_t> save;
    save.alloc(bytes);
    uint8_t* sv = save.p;
    uint8_t* sk = sv + nv * sizeof(float);
    uint8_t* sx = sk + nv;
    HIP_OK(hipMemcpyAsync(sv, L.v, nv * sizeof(float), hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(sk, L.k, nv, hipMemcpyDeviceToDevice, c.stream));
    HIP_OK(hipMemcpyAsync(sx, L.x, nv * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    const int forms[2] = { 1, 4 };
    double ms[2];
    for (int f = 0; f < 2; f++)
    {
        HIP_OK(hipEventRecord(c.ev0, c.stream));
        for (int col = 0; col < 4; col++)
        {
            const long long nc = colour[col].start[colour[col].n];
            if (nc > 0)
                launch_sweep_fast(forms[f], nc, c.stream, L, colour[col], col, c.counters.p);
        }
        HIP_OK(hipGetLastError());
        ms[f] = c.ms_since(c.ev0);
        HIP_OK(hipMemcpyAsync(L.v, sv, nv * sizeof(float), hipMemcpyDeviceToDevice, c.stream));
        HIP_OK(hipMemcpyAsync(L.k, sk, nv, hipMemcpyDeviceToDevice, c.stream));
        HIP_OK(hipMemcpyAsync(L.x, sx, nv * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    }
    HIP_OK(hipStreamSynchronize(c.stream));
    save.release();
    const int choice = ms[0] <= ms[1] ? forms[0] : forms[1];
    record_choice(key, choice, forms, ms, 2);
    return choice;
}

void quant_sweeps(Ctx& c, bool level0, int sweeps)
{
    LatentView L = make_view(c, level0);
    Ranges colour[4];
    for (int col = 0; col < 4; col++)
    {
        const int cx = col & 1, cy = col >> 1;
        colour[col].n = c.planes.n;
        long long acc = 0;
        for (int p = 0; p < c.planes.n; p++)
        {
            colour[col].start[p] = acc;
            const int cw1 = (L.pw[p] - cx + 1) / 2, ch1 = (L.ph[p] - cy + 1) / 2;
            acc += (long long)cw1 * ch1;
        }
        colour[col].start[c.planes.n] = acc;
    }
    // The plain form takes one lane a texel; the fast form at three or four channels the tuner's launch form.
    const bool fast_rows = c.form == NNTC_FORM_FAST && L.c >= 3;
    const int lanes = fast_rows && sweeps > 0 ? sweep_lanes_for(c, L, colour) : 1;
    // The counters start at zero after any trials of the tuner.
    HIP_OK(hipMemsetAsync(c.counters.p, 0, c.counters.bytes(), c.stream));
    for (int sw = 0; sw < sweeps; sw++)
        for (int col = 0; col < 4; col++)
        {
            const long long n = colour[col].start[colour[col].n];
            if (n == 0)
                continue;
            // A colour's texels read their own stencil blocks and their four backward neighbours' (the transposes
            // of the blocks those own towards the texel: nine C x C blocks in all, each read by this texel alone in
            // this launch), nine corrections (their own and their eight neighbours'), 8-byte values or 4-byte in the
            // fast form, and their gradient, plane and index.
            const double cc = L.c, sb = c.form == NNTC_FORM_EXACT ? 8.0 : 4.0;
            KernelTimer kt(c.stream, level0 ? "k_quant_sweep0" : "k_quant_sweep1",
                           (double)n * (9.0 * sb * cc * cc + 8.0 * cc + 9.0 * sb * cc + 9.0 * cc), 0.0);
            // The exact form reads the fp64 stencil; the fast form its fp32 copy (half the bytes).
            if (c.form == NNTC_FORM_EXACT)
                launch_sweep<double>(L.c, dim3(blocks_for(n, 256)), c.stream, L, colour[col], col, c.counters.p);
            else if (fast_rows)
                launch_sweep_fast(lanes, n, c.stream, L, colour[col], col, c.counters.p);
            else
                launch_sweep<float>(L.c, dim3(blocks_for(n, 256)), c.stream, L, colour[col], col, c.counters.p);
        }
    HIP_OK(hipGetLastError());
}

}   // namespace nntc
