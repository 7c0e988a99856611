// twins.hip: the --check twins of the seam (src/backend.h:124, :136) for the gfx950 arm.
//
// The reference keeps two host-side computations beside each arm so that its --check mode can test the device's
// answers against something that shares nothing with them but the model's definition: E by brute force
// (objective_check_host, src/objective.cu:317-380) and block (b)'s level-1 system solved densely
// (solve_level1_dense_host, src/solve_level1.cu:1502-1688). They are checkers, not a path: the encode never calls
// them, and they run only on planes small enough for a dense (texels * C1)^2 system. Both read the arm's current
// planes back from the device and work in fp64 on the host.
//
// Written from the model's definition: every site is the centre of a pixel or one of its K rotated sub-texel
// positions; level 1 (and level 0 and the source away from the centre) is read there by the bilinear clamp rule; the
// decode is out = W phi + b with phi = [z1, z0, z0 (x) z1]. The dense system is E's Gauss-Newton quadratic in the
// level-1 plane: at a site, out is affine in the level-1 sample with Jacobian G = W_a + sum_i z0_i W_sc(i), and the
// sample is a weighted sum of at most four texels (clamped taps that land on one texel add their weights), so each
// site adds w_e w_f G^T diag(cw) G to the (e, f) block of H and w_e G^T diag(cw) r to texel e's gradient.

#include "arm.h"

#include <cmath>
#include <vector>

namespace nntc
{

namespace
{

// The bilinear clamp rule at position (u, v) in [0, 1]^2 of a w x h plane: x = u w - 0.5, taps floor(x) and
// floor(x) + 1 clamped into the plane, weights 1 - frac and frac; the same in y. Returns up to four distinct texels
// and their weights (a texel two clamped taps land on carries both weights).
int bilinear_taps(double u, double v, int w, int h, long long* texel, double* weight)
{
    const double x = u * (double)w - 0.5, y = v * (double)h - 0.5;
    const double fx = std::floor(x), fy = std::floor(y);
    const double ax = x - fx, ay = y - fy;
    auto clamp = [](long long i, int n) { return i < 0 ? 0ll : (i > n - 1 ? (long long)(n - 1) : i); };
    const long long xs[2] = { clamp((long long)fx, w), clamp((long long)fx + 1, w) };
    const long long ys[2] = { clamp((long long)fy, h), clamp((long long)fy + 1, h) };
    const double wx[2] = { 1.0 - ax, ax }, wy[2] = { 1.0 - ay, ay };
    int n = 0;
    for (int b = 0; b < 2; b++)
        for (int a = 0; a < 2; a++)
        {
            const long long t = ys[b] * w + xs[a];
            const double wt = wx[a] * wy[b];
            int at = -1;
            for (int e = 0; e < n; e++)
                if (texel[e] == t)
                    at = e;
            if (at >= 0)
                weight[at] += wt;
            else
            {
                texel[n] = t;
                weight[n++] = wt;
            }
        }
    return n;
}

void sample(const std::vector<float>& plane, int channels, int n, const long long* texel, const double* weight,
            double* out)
{
    for (int q = 0; q < channels; q++)
        out[q] = 0.0;
    for (int e = 0; e < n; e++)
        for (int q = 0; q < channels; q++)
            out[q] += weight[e] * (double)plane[(size_t)texel[e] * channels + q];
}

template <class T>
std::vector<T> fetch(Ctx& c, const T* device, size_t count)
{
    std::vector<T> out(count);
    if (count)
        HIP_OK(hipMemcpy(out.data(), device, count * sizeof(T), hipMemcpyDeviceToHost));
    return out;
}

// One site of plane geometry g: its position in [0, 1]^2 (pixel (px, py), site j of 0..K).
void site_position(const PlaneGeom& g, int k, int px, int py, int j, double& u, double& v)
{
    float dx = 0.0f, dy = 0.0f;
    if (j > 0)
        subtexel_offset(k, j - 1, px, py, dx, dy);
    u = ((double)px + 0.5 + (double)dx) / (double)g.w0;
    v = ((double)py + 0.5 + (double)dy) / (double)g.h0;
}

}   // namespace

void objective_check_host(Ctx& c, const nntc_model_state* st, int k, double* acc)
{
    HIP_OK(hipStreamSynchronize(c.stream));
    const int c0 = c.c0, c1 = c.c1, nout = c.nout, nin = c.nin;
    for (int p = 0; p < c.planes.n; p++)
    {
        const PlaneGeom& g = c.g(p);
        const std::vector<float> v0 = fetch(c, c.v0.p + g.off0 * c0, (size_t)g.n0 * c0);
        const std::vector<float> v1 = fetch(c, c.v1.p + g.off1 * c1, (size_t)g.n1 * c1);
        const std::vector<float> src = fetch(c, c.src.p + g.off0 * nout, (size_t)g.n0 * nout);
        double weighted = 0.0, centre = 0.0, fractional = 0.0;
        for (int py = 0; py < g.h0; py++)
            for (int px = 0; px < g.w0; px++)
                for (int j = 0; j <= k; j++)
                {
                    double u, v;
                    site_position(g, k, px, py, j, u, v);
                    long long tx[4];
                    double tw[4];
                    double z0[NNTC_MAX_CHANNELS], z1[NNTC_MAX_CHANNELS], target[NNTC_MAX_NOUT];
                    sample(v1, c1, bilinear_taps(u, v, g.w1, g.h1, tx, tw), tx, tw, z1);
                    if (j == 0)
                    {
                        // The centre reads level 0 and the source at the pixel itself.
                        tx[0] = (long long)py * g.w0 + px;
                        tw[0] = 1.0;
                        sample(v0, c0, 1, tx, tw, z0);
                        sample(src, nout, 1, tx, tw, target);
                    }
                    else
                    {
                        const int n = bilinear_taps(u, v, g.w0, g.h0, tx, tw);
                        sample(v0, c0, n, tx, tw, z0);
                        sample(src, nout, n, tx, tw, target);
                    }
                    double phi[NNTC_MAX_NIN];
                    int f = 0;
                    for (int q = 0; q < c1; q++)
                        phi[f++] = z1[q];
                    for (int i = 0; i < c0; i++)
                        phi[f++] = z0[i];
                    for (int i = 0; i < c0; i++)
                        for (int q = 0; q < c1; q++)
                            phi[f++] = z0[i] * z1[q];
                    for (int r = 0; r < nout; r++)
                    {
                        double out = (double)st->b[r];
                        for (int i = 0; i < nin; i++)
                            out += (double)st->w[r * nin + i] * phi[i];
                        const double e = out - target[r];
                        weighted += (double)c.md.cw[r] * e * e;
                        (j == 0 ? centre : fractional) += e * e;
                    }
                }
        acc[3 * p] = weighted;
        acc[3 * p + 1] = centre;
        acc[3 * p + 2] = fractional;
    }
}

bool solve_level1_dense_host(Ctx& c, const nntc_model_state* st, int k, int plane, double lambda,
                             const float* c_prev, std::vector<double>& delta, int max_unknowns)
{
    const PlaneGeom& g = c.g(plane);
    const int c0 = c.c0, c1 = c.c1, nout = c.nout, nin = c.nin;
    const long long nunk = g.n1 * c1;
    if (nunk > max_unknowns)
        return false;
    HIP_OK(hipStreamSynchronize(c.stream));
    const std::vector<float> v0 = fetch(c, c.v0.p + g.off0 * c0, (size_t)g.n0 * c0);
    const std::vector<float> src = fetch(c, c.src.p + g.off0 * nout, (size_t)g.n0 * nout);
    const std::vector<float> v1(c_prev, c_prev + nunk);
    const size_t n = (size_t)nunk;
    std::vector<double> h(n * n, 0.0), grad(n, 0.0);
    for (int py = 0; py < g.h0; py++)
        for (int px = 0; px < g.w0; px++)
            for (int j = 0; j <= k; j++)
            {
                double u, v;
                site_position(g, k, px, py, j, u, v);
                long long t1[4], t0[4];
                double w1[4], w0[4];
                const int n1 = bilinear_taps(u, v, g.w1, g.h1, t1, w1);
                double z0[NNTC_MAX_CHANNELS], z1[NNTC_MAX_CHANNELS], target[NNTC_MAX_NOUT];
                sample(v1, c1, n1, t1, w1, z1);
                int n0 = 1;
                if (j == 0)
                {
                    t0[0] = (long long)py * g.w0 + px;
                    w0[0] = 1.0;
                }
                else
                    n0 = bilinear_taps(u, v, g.w0, g.h0, t0, w0);
                sample(v0, c0, n0, t0, w0, z0);
                sample(src, nout, n0, t0, w0, target);
                // G (nout x C1): d out / d z1; r: the residual at the current level-1 sample.
                double gm[NNTC_MAX_NOUT][NNTC_MAX_CHANNELS], res[NNTC_MAX_NOUT];
                for (int r = 0; r < nout; r++)
                {
                    const float* row = st->w + (size_t)r * nin;
                    double out = (double)st->b[r];
                    for (int i = 0; i < c0; i++)
                        out += (double)row[c1 + i] * z0[i];
                    for (int q = 0; q < c1; q++)
                    {
                        double d = (double)row[q];
                        for (int i = 0; i < c0; i++)
                            d += z0[i] * (double)row[c1 + c0 + i * c1 + q];
                        gm[r][q] = d;
                        out += d * z1[q];
                    }
                    res[r] = out - target[r];
                }
                // M = G^T diag(cw) G and m = G^T diag(cw) r, then scattered by the taps' weights.
                double mm[NNTC_MAX_CHANNELS][NNTC_MAX_CHANNELS], mv[NNTC_MAX_CHANNELS];
                for (int a = 0; a < c1; a++)
                {
                    mv[a] = 0.0;
                    for (int r = 0; r < nout; r++)
                        mv[a] += (double)c.md.cw[r] * gm[r][a] * res[r];
                    for (int b = 0; b < c1; b++)
                    {
                        mm[a][b] = 0.0;
                        for (int r = 0; r < nout; r++)
                            mm[a][b] += (double)c.md.cw[r] * gm[r][a] * gm[r][b];
                    }
                }
                for (int e = 0; e < n1; e++)
                {
                    for (int a = 0; a < c1; a++)
                        grad[(size_t)t1[e] * c1 + a] += w1[e] * mv[a];
                    for (int f = 0; f < n1; f++)
                        for (int a = 0; a < c1; a++)
                            for (int b = 0; b < c1; b++)
                                h[((size_t)t1[e] * c1 + a) * n + (size_t)t1[f] * c1 + b] += w1[e] * w1[f] * mm[a][b];
                }
            }
    // (H + lambda I) delta = -grad by Cholesky: H is a sum of Gram matrices, so H + lambda I is positive definite for
    // lambda > 0; a pivot that is not positive and finite (lambda = 0 on a singular H) reports no solution.
    for (size_t i = 0; i < n; i++)
        h[i * n + i] += lambda;
    for (size_t j = 0; j < n; j++)
    {
        double d = h[j * n + j];
        for (size_t q = 0; q < j; q++)
            d -= h[j * n + q] * h[j * n + q];
        if (!(d > 0.0) || !std::isfinite(d))
            return false;
        const double l = std::sqrt(d);
        h[j * n + j] = l;
        for (size_t i = j + 1; i < n; i++)
        {
            double s = h[i * n + j];
            for (size_t q = 0; q < j; q++)
                s -= h[i * n + q] * h[j * n + q];
            h[i * n + j] = s / l;
        }
    }
    delta.assign(n, 0.0);
    for (size_t i = 0; i < n; i++)
    {
        double s = -grad[i];
        for (size_t q = 0; q < i; q++)
            s -= h[i * n + q] * delta[q];
        delta[i] = s / h[i * n + i];
    }
    for (size_t i = n; i-- > 0;)
    {
        double s = delta[i];
        for (size_t q = i + 1; q < n; q++)
            s -= h[q * n + i] * delta[q];
        delta[i] = s / h[i * n + i];
    }
    return true;
}

}   // namespace nntc
