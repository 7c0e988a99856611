// ref_shim.cpp: the reference CPU backend behind the C ABI of include/nntc_hip.h, as the test oracle.
//
// TEST INFRASTRUCTURE ONLY. This file is compiled by oracle/Makefile against the reference's own sources where they
// lie (src/backend.cpp, src/cpu/*.cpp, src/export.cpp, src/mips.cpp) into oracle/_ref/libnntc_ref.so. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it. It exports every nntc_hip_* entry point under the
// prefix nntc_ref_ with the same arguments, so one call sequence drives both arms and their outputs are compared.
//
// Each entry point converts the plain descriptor/state to the reference's Model (src/model.h:56-114), calls the be::
// forwarder of the same name with Backend::Cpu (src/backend.h:86-150), and writes the Model's mutated fields back.

#include "nntc_hip.h"

#include "backend.h"
#include "image.h"
#include "model.h"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace
{

struct RefCtx
{
    Model m;
    be::Device* d = nullptr;
    Image base;   // the source's base level, for the texture error sums
    struct Saved  // state_snapshot's copy
    {
        std::vector<std::vector<float>> v0, v1;
        std::vector<std::vector<uint8_t>> k0, k1, ep, sel, mirror;
        bool held = false;
    } saved;
};

thread_local std::string g_error;

void state_to_model(const nntc_model_state* s, Model& m)
{
    if (!s)
        return;
    const int nin = m.dec.nin, nout = m.nout;
    if (s->has_decoder)
    {
        m.dec.w.assign(s->w, s->w + (size_t)nout * nin);
        m.dec.b.assign(s->b, s->b + nout);
    }
    else
    {
        m.dec.w.clear();
        m.dec.b.clear();
    }
    if (s->has_grid0)
    {
        m.lo0.assign(s->lo0, s->lo0 + m.c0);
        m.hi0.assign(s->hi0, s->hi0 + m.c0);
    }
    else
    {
        m.lo0.clear();
        m.hi0.clear();
    }
    if (s->has_grid1)
    {
        m.lo1.assign(s->lo1, s->lo1 + m.c1);
        m.hi1.assign(s->hi1, s->hi1 + m.c1);
    }
    else
    {
        m.lo1.clear();
        m.hi1.clear();
    }
    // The palette mode's palette: one list per level-0 channel, 2^bits0[c] entries.
    m.palette0.clear();
    if (s->has_palette0)
        for (int c = 0; c < m.c0; c++)
            m.palette0.emplace_back(s->palette0[c], s->palette0[c] + (1 << m.bits0[(size_t)c]));
}

void model_to_state(const Model& m, nntc_model_state* s)
{
    if (!s)
        return;
    s->has_decoder = !m.dec.w.empty() && !m.dec.b.empty();
    if (s->has_decoder)
    {
        std::memcpy(s->w, m.dec.w.data(), m.dec.w.size() * sizeof(float));
        std::memcpy(s->b, m.dec.b.data(), m.dec.b.size() * sizeof(float));
    }
    s->has_grid0 = !m.lo0.empty();
    for (size_t c = 0; c < m.lo0.size() && c < 4; c++)
    {
        s->lo0[c] = m.lo0[c];
        s->hi0[c] = m.hi0[c];
    }
    s->has_grid1 = !m.lo1.empty();
    for (size_t c = 0; c < m.lo1.size() && c < 4; c++)
    {
        s->lo1[c] = m.lo1[c];
        s->hi1[c] = m.hi1[c];
    }
}

const char* policy_name(int p)
{
    return p == NNTC_RANGE_PCT ? "pct" : "minmax";
}

void copy_l1(const Level1Report& a, nntc_level1_report& b)
{
    b.iterations = a.iterations;
    b.residual = a.residual;
    b.lambda = a.lambda;
    b.min_diag = a.min_diag;
    b.max_diag = a.max_diag;
    b.ms_assemble = a.ms_assemble;
    b.ms_precond = a.ms_precond;
    b.ms_cg = a.ms_cg;
    b.moved_values = a.moved_values;
    b.delta2 = a.delta2;
    b.plane2 = a.plane2;
    for (int c = 0; c < 4; c++)
    {
        b.lo[c] = a.lo[c];
        b.hi[c] = a.hi[c];
    }
    b.quantised = a.quantised;
    b.sweeps = a.sweeps;
    b.moved = a.moved;
    b.ms_sweeps = a.ms_sweeps;
}

template <class T>
int download(const std::vector<T>& v, T* out, size_t n, size_t* count)
{
    if (count)
        *count = v.size();
    const size_t m = v.size() < n ? v.size() : n;
    if (out && m)
        std::memcpy(out, v.data(), m * sizeof(T));
    return 0;
}

}   // namespace

extern "C"
{

const char* nntc_ref_last_error(void)
{
    return g_error.c_str();
}

int nntc_ref_set_threads(int threads)
{
    be::backend_select(be::Backend::Cpu, threads);
    return 0;
}

int nntc_ref_device_select(int index, nntc_device_info* info)
{
    DeviceInfo d;
    const bool ok = be::device_select(index, d);
    if (info)
    {
        std::snprintf(info->name, sizeof(info->name), "%s", d.name.c_str());
        info->major = d.major;
        info->minor = d.minor;
        std::snprintf(info->error, sizeof(info->error), "%s", d.error.c_str());
    }
    return ok ? 1 : 0;
}

nntc_hip_ctx* nntc_ref_device_create(const nntc_model_desc* md, const float* const* source_chain)
{
    RefCtx* r = new RefCtx();
    Model& m = r->m;
    m.textures = md->textures;
    m.nout = md->nout;
    m.c0 = md->c0;
    m.c1 = md->c1;
    m.bits0.assign(md->bits0, md->bits0 + md->c0);
    m.bits1 = md->bits1;
    m.bc0 = md->bc0 != 0;
    m.l0_bc8 = md->l0_bc8 != 0;
    m.init0_residual = md->init0_residual != 0;
    m.init0_texture = md->init0_texture != 0;
    m.init0_chain = md->init0_chain != 0;
    m.dec.nin = feature_count(m.c0, m.c1);
    m.dec.nout = m.nout;
    m.planes.resize((size_t)md->nplanes);
    for (int i = 0; i < md->nplanes; i++)
    {
        m.planes[i].w0 = md->planes[i].w0;
        m.planes[i].h0 = md->planes[i].h0;
        m.planes[i].w1 = md->planes[i].w1;
        m.planes[i].h1 = md->planes[i].h1;
    }
    m.width = m.planes[0].w0;
    m.height = m.planes[0].h0;
    m.source_width = m.width;
    m.source_height = m.height;
    m.cw.assign(md->cw, md->cw + m.nout);
    if (m.l0_bc8)
    {
        m.lo0.assign((size_t)m.c0, -1.0f);
        m.hi0.assign((size_t)m.c0, 1.0f);
    }
    else
        build_level0_palette(m.bits0, m.c0, m.bc0, m.palette0);
    std::vector<Image> chain((size_t)md->nplanes);
    for (int i = 0; i < md->nplanes; i++)
    {
        chain[i].w = m.planes[i].w0;
        chain[i].h = m.planes[i].h0;
        chain[i].nc = m.nout;
        chain[i].v.assign(source_chain[i], source_chain[i] + (size_t)chain[i].w * chain[i].h * m.nout);
    }
    r->base = chain[0];
    r->d = be::device_create(m, chain);
    return reinterpret_cast<nntc_hip_ctx*>(r);
}

void nntc_ref_device_destroy(nntc_hip_ctx* ctx)
{
    RefCtx* r = reinterpret_cast<RefCtx*>(ctx);
    if (!r)
        return;
    be::device_destroy(r->d);
    delete r;
}

size_t nntc_ref_device_memory(const nntc_hip_ctx* ctx)
{
    const RefCtx* r = reinterpret_cast<const RefCtx*>(ctx);
    return be::device_memory(r->d);
}

int nntc_ref_set_form(nntc_hip_ctx*, int)
{
    return 0;
}

int nntc_ref_stencil_monomials(int c0)
{
    return be::stencil_monomials(c0);
}

// Grids the reference keeps in the Model are the model's at creation until a call fits them; the state handed in
// replaces them, so a test can set any grid it likes.
#define REF_CTX RefCtx* r = reinterpret_cast<RefCtx*>(ctx); Model& m = r->m

int nntc_ref_init_level0(nntc_hip_ctx* ctx, nntc_model_state* st, const float luma[3], int luma_channel0)
{
    REF_CTX;
    state_to_model(st, m);
    be::init_level0(r->d, m, luma, luma_channel0 != 0);
    model_to_state(m, st);
    return 0;
}

int nntc_ref_init_level1(nntc_hip_ctx* ctx, nntc_model_state* st, int pca, nntc_level1_init_report* rep)
{
    REF_CTX;
    state_to_model(st, m);
    Level1InitReport lr;
    be::init_level1(r->d, m, pca != 0, lr);
    model_to_state(m, st);
    if (rep)
    {
        rep->pca = lr.pca;
        rep->components = lr.components;
        for (int i = 0; i < 4; i++)
        {
            rep->eigenvalue[i] = lr.eigenvalue[i];
            rep->peak[i] = lr.peak[i];
        }
        rep->total_variance = lr.total_variance;
    }
    return 0;
}

int nntc_ref_init0_residual_channel(nntc_hip_ctx* ctx, nntc_model_state* st, int channel, const double* pw,
                                    nntc_init0_report* rep)
{
    REF_CTX;
    state_to_model(st, m);
    std::vector<double> w(pw, pw + m.planes.size());
    Init0ChannelReport cr;
    be::init0_residual_channel(r->d, m, channel, w, cr);
    model_to_state(m, st);
    if (rep)
    {
        rep->eigenvalue = cr.eigenvalue;
        rep->total_variance = cr.total_variance;
        rep->share = cr.share;
        rep->peak = cr.peak;
        rep->divisor = cr.divisor;
        for (int i = 0; i < NNTC_MAX_NOUT; i++)
            rep->dir[i] = cr.dir[i];
        rep->nout = cr.nout;
        rep->texture = cr.texture;
    }
    return 0;
}

#define REF_GRID_CALL(name)                                                                                            \
    int nntc_ref_##name(nntc_hip_ctx* ctx, nntc_model_state* st, int policy)                                          \
    {                                                                                                                  \
        REF_CTX;                                                                                                       \
        state_to_model(st, m);                                                                                         \
        be::name(r->d, m, policy_name(policy));                                                                        \
        model_to_state(m, st);                                                                                         \
        return 0;                                                                                                      \
    }
REF_GRID_CALL(quantise_level1)
REF_GRID_CALL(freeze_level1_grid)
REF_GRID_CALL(quantise_level0)
REF_GRID_CALL(freeze_level0_grid)

int nntc_ref_snap_level0_on_grid(nntc_hip_ctx* ctx, nntc_model_state* st)
{
    REF_CTX;
    state_to_model(st, m);
    be::snap_level0_on_grid(r->d, m);
    model_to_state(m, st);
    return 0;
}

int nntc_ref_solve_decoder(nntc_hip_ctx* ctx, nntc_model_state* st, int k, const double* per_site, double* ms)
{
    REF_CTX;
    state_to_model(st, m);
    std::vector<double> ps(per_site, per_site + m.planes.size());
    const double t = be::solve_decoder(r->d, m, k, ps);
    model_to_state(m, st);
    if (ms)
        *ms = t;
    return 0;
}

void nntc_ref_decoder_ridge_tally(long long out[4])
{
    be::decoder_ridge_tally(out);
}

#define REF_BLOCK_CALL(name)                                                                                           \
    int nntc_ref_##name(nntc_hip_ctx* ctx, const nntc_model_state* st, int k, double ridge, nntc_level1_report* rep,  \
                        double* ms)                                                                                    \
    {                                                                                                                  \
        REF_CTX;                                                                                                       \
        state_to_model(st, m);                                                                                         \
        std::vector<Level1Report> lr;                                                                                  \
        const double t = be::name(r->d, m, k, ridge, lr);                                                              \
        for (size_t i = 0; i < lr.size(); i++)                                                                         \
            copy_l1(lr[i], rep[i]);                                                                                    \
        if (ms)                                                                                                        \
            *ms = t;                                                                                                   \
        return 0;                                                                                                      \
    }
#define REF_SWEEP_CALL(name)                                                                                           \
    int nntc_ref_##name(nntc_hip_ctx* ctx, const nntc_model_state* st, int k, double ridge, int sweeps,               \
                        nntc_level1_report* rep, double* ms)                                                           \
    {                                                                                                                  \
        REF_CTX;                                                                                                       \
        state_to_model(st, m);                                                                                         \
        std::vector<Level1Report> lr;                                                                                  \
        const double t = be::name(r->d, m, k, ridge, sweeps, lr);                                                      \
        for (size_t i = 0; i < lr.size(); i++)                                                                         \
            copy_l1(lr[i], rep[i]);                                                                                    \
        if (ms)                                                                                                        \
            *ms = t;                                                                                                   \
        return 0;                                                                                                      \
    }
int nntc_ref_solve_level0_all(nntc_hip_ctx* ctx, const nntc_model_state* st, int k, int sweeps,
                              nntc_level0_report* rep, double* ms)
{
    REF_CTX;
    state_to_model(st, m);
    std::vector<Level0Report> lr;
    const double t = be::solve_level0_all(r->d, m, k, sweeps, lr);
    for (size_t i = 0; i < lr.size(); i++)
    {
        rep[i] = nntc_level0_report {};
        for (int q = 0; q < 4; q++)
            rep[i].moved[q] = lr[i].moved[q];
        rep[i].moved_total = lr[i].moved_total;
        rep[i].ms = lr[i].ms;
        rep[i].joint = lr[i].joint ? 1 : 0;
        rep[i].states = lr[i].states;
    }
    if (ms)
        *ms = t;
    return 0;
}

REF_BLOCK_CALL(solve_level1_all)
REF_BLOCK_CALL(solve_level0_cont_all)
REF_SWEEP_CALL(sweep_level1_all)
REF_SWEEP_CALL(sweep_level0_cont_all)

int nntc_ref_assemble_level0_for_refine(nntc_hip_ctx* ctx, const nntc_model_state* st, int k, double* ms)
{
    REF_CTX;
    state_to_model(st, m);
    const double t = be::assemble_level0_for_refine(r->d, m, k);
    if (ms)
        *ms = t;
    return 0;
}

int nntc_ref_level1_download_indices(nntc_hip_ctx* ctx, int plane, uint8_t* k1, size_t n, size_t* count)
{
    REF_CTX;
    std::vector<uint8_t> v;
    be::level1_download_indices(r->d, m, plane, v);
    return download(v, k1, n, count);
}

int nntc_ref_level1_upload(nntc_hip_ctx* ctx, int plane, const float* v1, size_t n)
{
    REF_CTX;
    be::level1_upload(r->d, m, plane, std::vector<float>(v1, v1 + n));
    return 0;
}

int nntc_ref_level1_upload_indices(nntc_hip_ctx* ctx, int plane, const uint8_t* k1, size_t n)
{
    REF_CTX;
    be::level1_upload_indices(r->d, m, plane, std::vector<uint8_t>(k1, k1 + n));
    return 0;
}

int nntc_ref_level0_upload_indices(nntc_hip_ctx* ctx, int plane, const uint8_t* k0, size_t n)
{
    REF_CTX;
    be::level0_upload_indices(r->d, m, plane, std::vector<uint8_t>(k0, k0 + n));
    return 0;
}

int nntc_ref_upload_decoder_to_device(nntc_hip_ctx* ctx, const nntc_model_state* st)
{
    REF_CTX;
    state_to_model(st, m);
    be::upload_decoder_to_device(r->d, m);
    return 0;
}

int nntc_ref_upload_grids_to_device(nntc_hip_ctx* ctx, const nntc_model_state* st)
{
    REF_CTX;
    state_to_model(st, m);
    be::upload_grids_to_device(r->d, m);
    return 0;
}

int nntc_ref_level1_download(nntc_hip_ctx* ctx, int plane, float* v1, size_t n, size_t* count)
{
    REF_CTX;
    std::vector<float> v;
    be::level1_download(r->d, m, plane, v);
    return download(v, v1, n, count);
}

// The reference's own check (src/main.cpp:2696-2711): both planes downloaded, verify_level1_on_grid over them.
int nntc_ref_level1_off_grid(nntc_hip_ctx* ctx, const nntc_model_state* st, size_t* values, size_t* off)
{
    REF_CTX;
    state_to_model(st, m);
    std::vector<std::vector<float>> v1(m.planes.size());
    for (size_t i = 0; i < m.planes.size(); i++)
    {
        be::level1_download_indices(r->d, m, (int)i, m.k1[i]);
        be::level1_download(r->d, m, (int)i, v1[i]);
    }
    size_t nv = 0, no = 0;
    verify_level1_on_grid(m, v1, nv, no);
    if (values)
        *values = nv;
    if (off)
        *off = no;
    return 0;
}

int nntc_ref_level1_delta(nntc_hip_ctx* ctx, int plane, double* delta, size_t n, size_t* count)
{
    REF_CTX;
    std::vector<double> v;
    be::level1_delta(r->d, m, plane, v);
    return download(v, delta, n, count);
}

int nntc_ref_level1_poke(nntc_hip_ctx* ctx, int plane, size_t index, float value)
{
    REF_CTX;
    (void)m;
    be::level1_poke(r->d, plane, index, value);
    return 0;
}

int nntc_ref_level1_range(nntc_hip_ctx* ctx, int plane, float* lo, float* hi)
{
    REF_CTX;
    std::vector<float> l, h;
    be::level1_range(r->d, m, plane, l, h);
    for (size_t c = 0; c < l.size(); c++)
    {
        lo[c] = l[c];
        hi[c] = h[c];
    }
    return 0;
}

int nntc_ref_level0_download(nntc_hip_ctx* ctx, int plane, uint8_t* k0, size_t n, size_t* count)
{
    REF_CTX;
    std::vector<uint8_t> v;
    be::level0_download(r->d, m, plane, v);
    return download(v, k0, n, count);
}

int nntc_ref_level0_upload(nntc_hip_ctx* ctx, int plane, const float* v0, size_t n)
{
    REF_CTX;
    be::level0_upload(r->d, m, plane, std::vector<float>(v0, v0 + n));
    return 0;
}

int nntc_ref_level0_download_values(nntc_hip_ctx* ctx, int plane, float* v0, size_t n, size_t* count)
{
    REF_CTX;
    std::vector<float> v;
    be::level0_download_values(r->d, m, plane, v);
    return download(v, v0, n, count);
}

int nntc_ref_decode_plane(nntc_hip_ctx* ctx, int plane, uint8_t* rgb8, size_t n)
{
    REF_CTX;
    std::vector<uint8_t> v;
    be::decode_plane(r->d, m, plane, v);
    return download(v, rgb8, n, nullptr);
}

// The reference's image_psnr_texture (src/image.h:387-407) squared-error sum over decode_plane's bytes, per texture.
int nntc_ref_texture_sse(nntc_hip_ctx* ctx, int width, int height, unsigned long long* sse)
{
    REF_CTX;
    std::vector<uint8_t> rec;
    be::decode_plane(r->d, m, 0, rec);
    for (int t = 0; t < m.textures; t++)
    {
        unsigned long long s = 0;
        for (int y = 0; y < height; y++)
            for (int x = 0; x < width; x++)
                for (int c = 0; c < 3; c++)
                {
                    const float f = r->base.at(x, y)[3 * t + c];
                    const int a = (int)std::lround(std::min(1.0f, std::max(0.0f, f)) * 255.0f);
                    const int d = a - (int)rec[((size_t)y * r->base.w + x) * m.nout + 3 * t + c];
                    s += (unsigned long long)(d * d);
                }
        sse[t] = s;
    }
    return 0;
}

int nntc_ref_objective_eval(nntc_hip_ctx* ctx, int k, const double* per_site, nntc_objective* o)
{
    REF_CTX;
    std::vector<double> ps(per_site, per_site + m.planes.size());
    Objective e;
    be::objective_eval(r->d, m, k, ps, e);
    o->e = e.e;
    o->centre_mse = e.centre_mse;
    o->sampled_mse = e.sampled_mse;
    o->ms = e.ms;
    o->nplanes = (int32_t)m.planes.size();
    for (size_t p = 0; p < m.planes.size(); p++)
    {
        o->e_plane[p] = e.e_plane[p];
        o->centre_plane[p] = e.centre_plane[p];
    }
    return 0;
}

int nntc_ref_objective_check_host(nntc_hip_ctx* ctx, const nntc_model_state* st, int k, const double* per_site,
                                  nntc_objective* o)
{
    REF_CTX;
    state_to_model(st, m);
    std::vector<double> ps(per_site, per_site + m.planes.size());
    Objective e;
    be::objective_check_host(r->d, m, k, ps, e);
    o->e = e.e;
    o->centre_mse = e.centre_mse;
    o->sampled_mse = e.sampled_mse;
    o->nplanes = (int32_t)m.planes.size();
    for (size_t p = 0; p < m.planes.size(); p++)
    {
        o->e_plane[p] = e.e_plane[p];
        o->centre_plane[p] = e.centre_plane[p];
    }
    return 0;
}

int nntc_ref_solve_level1_dense_host(nntc_hip_ctx* ctx, const nntc_model_state* st, int k, int plane, double lambda,
                                     const float* c_prev, size_t n_prev, double* delta, size_t n_delta,
                                     int max_unknowns, int* solved)
{
    REF_CTX;
    state_to_model(st, m);
    std::vector<float> prev(c_prev, c_prev + n_prev);
    std::vector<double> d;
    *solved = be::solve_level1_dense_host(r->d, m, k, plane, lambda, prev, d, max_unknowns) ? 1 : 0;
    if (*solved && n_delta < d.size())
        return 1;
    for (size_t i = 0; i < d.size() && *solved; i++)
        delta[i] = d[i];
    return 0;
}

double nntc_ref_objective_total_ms(void)
{
    return be::objective_total_ms();
}

long long nntc_ref_objective_passes(void)
{
    return be::objective_passes();
}

// Model::k0 as the C ABI defines it: the caller's planes when given; otherwise, for a seeding prepare, the device's
// indices as they stand, and for the other calls the mirror already held.
static void set_k0_mirror(RefCtx* r, const uint8_t* const* k0, bool snapshot)
{
    Model& m = r->m;
    if (k0)
    {
        m.k0.assign(m.planes.size(), std::vector<uint8_t>());
        for (size_t i = 0; i < m.planes.size(); i++)
            m.k0[i].assign(k0[i], k0[i] + (size_t)m.planes[i].w0 * m.planes[i].h0 * m.c0);
    }
    else if (snapshot)
    {
        m.k0.assign(m.planes.size(), std::vector<uint8_t>());
        for (size_t i = 0; i < m.planes.size(); i++)
            be::level0_download(r->d, m, (int)i, m.k0[i]);
    }
}

int nntc_ref_bc_pack_prepare(nntc_hip_ctx* ctx, const nntc_model_state* st, int k, int seed,
                             const uint8_t* const* k0_mirror, double* pack_psnr, size_t* pack_texels, double* ms)
{
    REF_CTX;
    state_to_model(st, m);
    set_k0_mirror(r, k0_mirror, seed != 0);
    double psnr = 0.0;
    size_t texels = 0;
    const double t = be::bc_pack_prepare(r->d, m, k, seed != 0, psnr, texels);
    if (pack_psnr)
        *pack_psnr = psnr;
    if (pack_texels)
        *pack_texels = texels;
    if (ms)
        *ms = t;
    return 0;
}

int nntc_ref_bc_refine(nntc_hip_ctx* ctx, int passes, const uint8_t* const* k0_mirror, nntc_bc_refine_report* rep,
                       double* pack_psnr, size_t* pack_texels, double* ms)
{
    REF_CTX;
    set_k0_mirror(r, k0_mirror, false);
    std::vector<BcRefineReport> rr;
    double psnr = 0.0;
    size_t texels = 0;
    const double t = be::bc_refine(r->d, m, passes, rr, psnr, texels);
    for (size_t i = 0; i < rr.size(); i++)
    {
        rep[i].pass = rr[i].pass;
        rep[i].blocks = rr[i].blocks;
        rep[i].improved = rr[i].improved;
        rep[i].delta_e = rr[i].delta_e;
        rep[i].ms = rr[i].ms;
    }
    if (pack_psnr)
        *pack_psnr = psnr;
    if (pack_texels)
        *pack_texels = texels;
    if (ms)
        *ms = t;
    return 0;
}

int nntc_ref_bc_blocks_from_device(nntc_hip_ctx* ctx, const uint8_t* const* k0_mirror, double* pack_psnr,
                                   size_t* pack_texels)
{
    REF_CTX;
    set_k0_mirror(r, k0_mirror, false);
    double psnr = 0.0;
    size_t texels = 0;
    be::bc_blocks_from_device(r->d, m, psnr, texels);
    if (pack_psnr)
        *pack_psnr = psnr;
    if (pack_texels)
        *pack_texels = texels;
    return 0;
}

int nntc_ref_bc_blocks_download(nntc_hip_ctx* ctx, int plane, int file, uint8_t* blocks, size_t n, size_t* count)
{
    REF_CTX;
    if ((size_t)plane >= m.bc0_blocks.size() || (size_t)file >= m.bc0_blocks[(size_t)plane].size())
    {
        g_error = "no blocks";
        return 1;
    }
    return download(m.bc0_blocks[(size_t)plane][(size_t)file], blocks, n, count);
}

int nntc_ref_bc_state_save(nntc_hip_ctx* ctx, int plane, uint8_t* ep, size_t n_ep, uint8_t* sel, size_t n_sel)
{
    REF_CTX;
    std::vector<std::vector<uint8_t>> e, s;
    be::bc_state_save(r->d, m, e, s);
    download(e[(size_t)plane], ep, n_ep, nullptr);
    download(s[(size_t)plane], sel, n_sel, nullptr);
    return 0;
}

int nntc_ref_bc_state_restore(nntc_hip_ctx* ctx, int plane, const uint8_t* ep, size_t n_ep, const uint8_t* sel,
                              size_t n_sel)
{
    REF_CTX;
    std::vector<std::vector<uint8_t>> e, s;
    be::bc_state_save(r->d, m, e, s);
    e[(size_t)plane].assign(ep, ep + n_ep);
    s[(size_t)plane].assign(sel, sel + n_sel);
    be::bc_state_restore(r->d, m, e, s);
    return 0;
}

// The reference's own round trip (src/main.cpp:2561-2588): every plane downloaded into the context, Model::k0 beside
// them, and uploaded back on restore.
int nntc_ref_state_snapshot(nntc_hip_ctx* ctx)
{
    REF_CTX;
    RefCtx::Saved& s = r->saved;
    const size_t np = m.planes.size();
    s.v0.assign(np, {});
    s.v1.assign(np, {});
    s.k0.assign(np, {});
    s.k1.assign(np, {});
    for (size_t p = 0; p < np; p++)
    {
        be::level0_download_values(r->d, m, (int)p, s.v0[p]);
        be::level0_download(r->d, m, (int)p, s.k0[p]);
        be::level1_download(r->d, m, (int)p, s.v1[p]);
        be::level1_download_indices(r->d, m, (int)p, s.k1[p]);
    }
    be::bc_state_save(r->d, m, s.ep, s.sel);
    s.mirror = m.k0;
    s.held = true;
    return 0;
}

int nntc_ref_state_restore(nntc_hip_ctx* ctx)
{
    REF_CTX;
    RefCtx::Saved& s = r->saved;
    if (!s.held)
    {
        g_error = "state_restore: no snapshot is held";
        return 1;
    }
    for (size_t p = 0; p < m.planes.size(); p++)
    {
        be::level0_upload(r->d, m, (int)p, s.v0[p]);
        be::level0_upload_indices(r->d, m, (int)p, s.k0[p]);
        be::level1_upload(r->d, m, (int)p, s.v1[p]);
        be::level1_upload_indices(r->d, m, (int)p, s.k1[p]);
    }
    be::bc_state_restore(r->d, m, s.ep, s.sel);
    m.k0 = s.mirror;
    return 0;
}

}   // extern "C"
