// encode.cpp: the encode driver - the reference's ALS schedule over the seam's C ABI (include/nntc_encode.h).
//
// This is the caller the gfx950 arm sits under: src/main.cpp's encode() from the packed source to the shipped
// objective (src/main.cpp:1838-2743), for the --l0 bc8 path the reference ships by default, restated call for call
// so that on the same inputs the same seam calls are made in the same order with the same arguments:
//   init_level0 / init_level1 (box) and the exact-copy channel order of the box init (main.cpp:1930-2027);
//   block (a) and the residual seed channel by channel with a refit after each (main.cpp:2032-2121);
//   the rounds: (a), the grid freeze at --q1-start, (b) continuous or quantised, (c') continuous or quantised, the
//   E-rise warning and the no-move / tol stops (main.cpp:2165-2390);
//   the final sweeps and block (a) (main.cpp:2397-2434);
//   the post-fit BC pack, its refinement, the refit, --bc-refine-after, the outer repack loop with its full save /
//   restore, and the last refit (main.cpp:2474-2694);
//   the on-grid verification of level 1, the shipped objective and the base plane's decode (main.cpp:2696-2743).
// The progress lines mirror the reference's round lines so two logs can be read side by side.

#include "nntc_encode.h"
#include "nntc_host.h"

#include <dlfcn.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace nntc_host
{

thread_local std::string g_error;

namespace
{

const double E_NOISE = 1e-8;
const double E_REL = 1e-6;
const float LUMA[3] = { 0.299f, 0.587f, 0.114f };

double mse_to_db(double mse)
{
    return mse > 0.0 ? 10.0 * std::log10(1.0 / mse) : 100.0;
}

template <class F>
bool bind(void* lib, const std::string& prefix, const char* name, F& f)
{
    f = reinterpret_cast<F>(dlsym(lib, (prefix + name).c_str()));
    return f != nullptr;
}

}   // namespace

bool Seam::load(const char* lib_path, const char* prefix)
{
    if (!lib_path)
    {
#define NNTC_DIRECT(n) n = &nntc_hip_##n;
        NNTC_SEAM_FUNCTIONS(NNTC_DIRECT)
#undef NNTC_DIRECT
        name = "hip";
        return true;
    }
    void* lib = dlopen(lib_path, RTLD_NOW | RTLD_GLOBAL);
    if (!lib)
    {
        g_error = std::string("cannot load ") + lib_path + ": " + dlerror();
        return false;
    }
    const std::string p = prefix ? prefix : "nntc_hip_";
    bool ok = true;
#define NNTC_BIND(n) ok = bind(lib, p, #n, n) && ok;
    NNTC_SEAM_FUNCTIONS(NNTC_BIND)
#undef NNTC_BIND
    if (!ok)
    {
        g_error = std::string(lib_path) + " does not export the whole seam under " + p;
        return false;
    }
    name = p;
    return true;
}

void Encoder::check(int status, const char* what)
{
    if (status != 0)
        throw Failure { std::string(what) + ": " + seam.last_error() };
    if (!call_times)
        return;
    const auto now = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(now - last_check).count();
    last_check = now;
    for (auto& e : call_ms)
        if (e.first == what)
        {
            e.second.first += ms;
            e.second.second++;
            return;
        }
    call_ms.push_back({ what, { ms, 1 } });
}

void Encoder::print_call_times()
{
    if (!call_times)
        return;
    std::vector<std::pair<std::string, std::pair<double, long long>>> v = call_ms;
    std::sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.second.first > b.second.first; });
    double total = 0.0;
    for (auto& e : v)
        total += e.second.first;
    std::fprintf(stderr, "call times (wall ms, the call and the host work before it): total %.1f\n", total);
    for (auto& e : v)
        std::fprintf(stderr, "  %-32s %6lld calls %9.2f ms\n", e.first.c_str(), e.second.second, e.second.first);
}

}   // namespace nntc_host

using namespace nntc_host;

namespace
{

// What every run ends with: level 1 checked on its grid, the shipped objective, the per-texture error and the report.
struct Totals
{
    double ms_a = 0.0, ms_b = 0.0, ms_c = 0.0, ms_pack = 0.0, ms_refine = 0.0;
    double pack_psnr = 0.0, pack_psnr_seed = 0.0, obj_ms0 = 0.0;
    long long tally0[4] = {};
    long long moved0_final = 0;
    std::chrono::steady_clock::time_point t_start {};
};

void ship(Encoder& E, const Totals& tt)
{
    Seam& S = E.seam;
    const nntc_encode_options& o = E.opts;
    nntc_hip_ctx* d = E.ctx;
    nntc_encode_report& rep = E.report;
    const int k = o.k;
    // ---- level 1 must be exactly the grid value of every index about to be written
    E.verify_level1_on_grid();

    nntc_objective shipped;
    E.check(S.objective_eval(d, k, E.per_site.data(), &shipped), "objective_eval");
    // The shipped decode's per-texture error, summed on the device; the bytes themselves are fetched only when asked.
    E.check(S.texture_sse(d, E.source_w, E.source_h, E.texture_se), "texture_sse");
    E.recon.clear();
    rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tt.t_start).count();
    rep.e_shipped = shipped.e;
    rep.centre_psnr = mse_to_db(shipped.centre_mse);
    rep.sampled_psnr = k > 0 ? mse_to_db(shipped.sampled_mse) : 0.0;
    rep.pack_psnr = tt.pack_psnr;
    rep.pack_psnr_seed = tt.pack_psnr_seed;
    rep.moved0_final = tt.moved0_final;
    rep.ms_a = tt.ms_a;
    rep.ms_b = tt.ms_b;
    rep.ms_c = tt.ms_c;
    rep.ms_e = S.objective_total_ms() - tt.obj_ms0;
    rep.ms_pack = tt.ms_pack;
    rep.ms_refine = tt.ms_refine;
    long long tally[4];
    S.decoder_ridge_tally(tally);
    for (int i = 0; i < 4; i++)
        rep.ridge_tally[i] = tally[i] - tt.tally0[i];
    for (int tx = 0; tx < E.desc.textures; tx++)
        rep.psnr[tx] = E.texture_psnr(tx);
    E.print_call_times();
    if (!o.quiet)
    {
        std::printf("\nreport\n  backend      %s\n", S.name.c_str());
        for (int tx = 0; tx < E.desc.textures; tx++)
            std::printf("  psnr         texture %d  %.2f dB\n", tx, rep.psnr[tx]);
        std::printf("  centre psnr  %.2f dB\n  E shipped    %.9e\n  time         %.3f s total, (a) %.3f ms / (b) "
                    "%.3f ms / (c) %.3f ms / E %.3f ms\n",
                    rep.centre_psnr, rep.e_shipped, rep.seconds, tt.ms_a, tt.ms_b, tt.ms_c, rep.ms_e);
    }
}

// --l0 palette's end (src/main.cpp:2713-2737): at 4 bits a BC4 / BC5 level 0 is a lossy pack of the indices, so the
// plane is packed as the writer packs it and decoded back onto the device before anything is measured; at 1-3 bits
// (and uncompressed) the device's plane is already what the file decodes to.
void finish_palette(Encoder& E, Totals tt)
{
    const nntc_model_desc& dd = E.desc;
    bool lossy = false;
    for (int c = 0; c < dd.c0; c++)
        lossy = lossy || dd.bits0[c] > 3;
    tt.pack_psnr = tt.pack_psnr_seed = 100.0;
    if (dd.bc0 && lossy)
    {
        double se = 0.0;
        size_t n = 0;
        const int stored = dd.c0 == 3 ? 4 : dd.c0;
        E.packed0.assign((size_t)level0_files(dd.c0), std::vector<std::vector<uint8_t>>((size_t)dd.nplanes));
        for (int i = 0; i < dd.nplanes; i++)
        {
            const int w = dd.planes[i].w0, h = dd.planes[i].h0, bx = (w + 3) / 4;
            const size_t texels = (size_t)w * h;
            std::vector<uint8_t> k0(texels * dd.c0), bytes;
            E.check(E.seam.level0_download(E.ctx, i, k0.data(), k0.size(), nullptr), "level0_download");
            level0_index_bytes(k0.data(), texels, dd.c0, dd.bits0, bytes);
            std::vector<float> v0(texels * dd.c0);
            for (int f = 0; f < level0_files(dd.c0); f++)
            {
                const int nc = level0_file_channels(dd.c0, f), bpb = nc == 1 ? 8 : 16;
                std::vector<uint8_t> blocks =
                    bc_pack_indices(bytes.data(), w, h, stored, 2 * f, nc, dd.bits0, dd.c0, se, n);
                parallel_rows((h + 3) / 4, [&](int b0, int b1) {
                    for (int c = 0; c < nc; c++)
                        for (int by = b0; by < b1; by++)
                            for (int bxi = 0; bxi < bx; bxi++)
                            {
                                float vals[16];
                                bc4_block_values(&blocks[((size_t)by * bx + bxi) * bpb + 8 * c], vals);
                                for (int t = 0; t < 16; t++)
                                {
                                    const int x = 4 * bxi + (t & 3), y = 4 * by + (t >> 2);
                                    if (x < w && y < h)
                                        v0[((size_t)y * w + x) * dd.c0 + 2 * f + c] =
                                            (float)(-1.0 + (double)vals[t] / 255.0 * 2.0);
                                }
                            }
                });
                E.packed0[(size_t)f][(size_t)i] = std::move(blocks);
            }
            E.check(E.seam.level0_upload(E.ctx, i, v0.data(), v0.size()), "level0_upload");
        }
        tt.pack_psnr = tt.pack_psnr_seed = se > 0.0 ? 10.0 * std::log10(255.0 * 255.0 / (se / (double)n)) : 100.0;
        if (!E.opts.quiet)
            std::printf("level 0 packed: the report below is of the PACKED plane (packing psnr %.2f dB against the "
                        "exact index values), decoded back from the blocks the .dds holds\n", tt.pack_psnr);
    }
    ship(E, tt);
}

void run_encode(Encoder& E)
{
    Seam& S = E.seam;
    const nntc_encode_options& o = E.opts;
    nntc_hip_ctx* d = E.ctx;
    nntc_encode_report& rep = E.report;
    nntc_model_state& st = E.state;
    const int np = E.desc.nplanes;
    const int k = o.k;
    const double* ps = E.per_site.data();
    auto cks = [&](int status, const char* what) { E.check(status, what); };
    auto obj = [&](nntc_objective& r) { cks(S.objective_eval(d, k, ps, &r), "objective_eval"); };
    double ms_a = 0.0, ms_b = 0.0, ms_c = 0.0, t = 0.0;
    std::vector<nntc_level1_report> lr((size_t)np), lr0((size_t)np);

    const auto t_start = std::chrono::steady_clock::now();
    E.call_times = std::getenv("NNTC_CALL_TIMES") != nullptr;
    E.call_ms.clear();
    E.last_check = t_start;
    const long long passes0 = S.objective_passes();
    const double obj_ms0 = S.objective_total_ms();
    long long tally0[4];
    S.decoder_ridge_tally(tally0);

    // ---- a fresh state: the decoder at zero on the device, no decoder on the host, level 0's grid [-1, 1].
    std::memset(&st, 0, sizeof(st));
    st.has_decoder = 1;
    cks(S.upload_decoder_to_device(d, &st), "upload_decoder_to_device");
    st.has_decoder = 0;
    const bool palette = E.desc.l0_bc8 == 0;
    if (palette)
    {
        // --l0 palette: the values the shipped format decodes each index to (src/model.h:205-222)
        st.has_palette0 = 1;
        for (int q = 0; q < E.desc.c0; q++)
            for (int kk = 0; kk < (1 << E.desc.bits0[q]); kk++)
                st.palette0[q][kk] = level0_value(E.desc.bits0[q], kk, E.desc.bc0 != 0);
    }
    else
    {
        st.has_grid0 = 1;
        for (int q = 0; q < E.desc.c0; q++)
        {
            st.lo0[q] = -1.0f;
            st.hi0[q] = 1.0f;
        }
    }
    std::vector<nntc_level0_report> l0((size_t)np);
    long long moved0_search = 0;
    // --l0 palette's block (c): the per-texel palette search on every plane; returns the texels it moved.
    auto search = [&]() {
        double tc = 0.0;
        cks(S.solve_level0_all(d, &st, k, o.sweeps, l0.data(), &tc), "solve_level0_all");
        ms_c += tc;
        long long moved = 0;
        for (int p = 0; p < np; p++)
            moved += l0[(size_t)p].moved_total;
        return moved;
    };

    // ---- the latents
    cks(S.init_level0(d, &st, LUMA, E.opts.init0_luma ? 1 : 0), "init_level0");
    nntc_level1_init_report irep;
    cks(S.init_level1(d, &st, E.opts.init1_pca ? 1 : 0, &irep), "init_level1");
    if (!E.opts.init1_pca)
        E.box_reorder();   // the box init's channel order (src/main.cpp:1938-2027); --init pca takes directions instead

    nntc_objective before, after, ea, eb, ec, probe;
    obj(before);
    cks(S.solve_decoder(d, &st, k, ps, &t), "solve_decoder");
    ms_a += t;
    obj(after);

    // ---- the residual seed, one channel at a time with a refit after each (--init0 residual or texture; --init0
    // luma seeds level 0's channel 0 from the luminance at init_level0 and takes no residual)
    for (int c = 0; c < E.desc.c0 && !E.opts.init0_luma; c++)
    {
        nntc_init0_report cr;
        cks(S.init0_residual_channel(d, &st, c, ps, &cr), "init0_residual_channel");
        nntc_objective seeded;
        obj(seeded);
        cks(S.solve_decoder(d, &st, k, ps, &t), "solve_decoder");
        ms_a += t;
        obj(after);
        const double e_refit = after.e;
        if (palette)
        {
            // The snap of the seed is not a minimisation: one search and a refit follow each seeded channel.
            search();
            cks(S.solve_decoder(d, &st, k, ps, &t), "solve_decoder");
            ms_a += t;
            obj(after);
        }
        if (!o.quiet)
        {
            std::printf("level 0 init: channel %d = the residual's leading direction, %.1f %% of its variance "
                        "(eigenvalue %.3e of a trace of %.3e, divisor %.4f of a peak of %.4f)\n",
                        c, 100.0 * cr.share, cr.eigenvalue, cr.total_variance, cr.divisor, cr.peak);
            std::printf("              E %.9e -> %.9e across the refit\n", seeded.e, e_refit);
            if (palette)
                std::printf("              E %.9e -> %.9e across the search and the refit after it\n", e_refit,
                            after.e);
        }
    }

    // ---- the rounds
    bool frozen = false;
    int frozen_round = 0, tol_hits = 0;
    double e_prev = after.e;
    rep.rounds_used = 0;
    for (int r = 1; r <= o.rounds; r++)
    {
        cks(S.solve_decoder(d, &st, k, ps, &t), "solve_decoder");
        const double round_a = t;
        ms_a += t;
        obj(ea);
        if (o.q1_start > 0 && r >= o.q1_start && !frozen)
        {
            cks(S.freeze_level1_grid(d, &st, o.q1_range), "freeze_level1_grid");
            if (!palette)   // a palette level 0 has no grid to fit
                cks(S.freeze_level0_grid(d, &st, o.q1_range), "freeze_level0_grid");
            frozen = true;
            frozen_round = r;
            if (!o.quiet)
            {
                std::printf("grid frozen: round %d, level 1", r);
                for (int q = 0; q < E.desc.c1; q++)
                    std::printf("  ch%d [%.4f, %.4f]", q, (double)st.lo1[q], (double)st.hi1[q]);
                if (!palette)
                {
                    std::printf("; level 0");
                    for (int q = 0; q < E.desc.c0; q++)
                        std::printf("  ch%d [%.4f, %.4f]", q, (double)st.lo0[q], (double)st.hi0[q]);
                }
                std::printf("\n");
            }
        }
        double round_b = 0.0;
        if (frozen)
            cks(S.sweep_level1_all(d, &st, k, o.ridge, o.q1_sweeps, lr.data(), &round_b), "sweep_level1_all");
        else
            cks(S.solve_level1_all(d, &st, k, o.ridge, lr.data(), &round_b), "solve_level1_all");
        ms_b += round_b;
        long long moved1 = 0;
        double delta_norm = 0.0, plane_norm = 0.0, min_diag = 1e300;
        for (int p = 0; p < np; p++)
        {
            min_diag = std::min(min_diag, lr[(size_t)p].min_diag);
            moved1 += lr[(size_t)p].quantised ? lr[(size_t)p].moved : lr[(size_t)p].moved_values;
            delta_norm += lr[(size_t)p].delta2;
            plane_norm += lr[(size_t)p].plane2;
        }
        delta_norm = std::sqrt(delta_norm);
        plane_norm = std::sqrt(plane_norm);
        obj(eb);
        double round_c = 0.0;
        long long moved0 = 0;
        if (palette)
        {
            const double before_c = ms_c;
            moved0 = search();
            round_c = ms_c - before_c;
        }
        else
        {
            if (frozen)
                cks(S.sweep_level0_cont_all(d, &st, k, o.ridge, o.q1_sweeps, lr0.data(), &round_c),
                    "sweep_level0_cont_all");
            else
                cks(S.solve_level0_cont_all(d, &st, k, o.ridge, lr0.data(), &round_c), "solve_level0_cont_all");
            ms_c += round_c;
            for (int p = 0; p < np; p++)
                moved0 += lr0[(size_t)p].quantised ? lr0[(size_t)p].moved : lr0[(size_t)p].moved_values;
        }
        obj(ec);

        const bool rose_a = ea.e > e_prev * (1.0 + E_REL) + E_NOISE;
        const bool rose_b = eb.e > ea.e * (1.0 + E_REL) + E_NOISE && r != frozen_round;
        const bool rose_c = ec.e > eb.e * (1.0 + E_REL) + E_NOISE;
        if (rose_a || rose_b || rose_c)
            std::printf("WARNING: round %d raised E across block (%s), which that block's own minimisation forbids\n",
                        r, rose_a ? "a" : (rose_b ? "b" : "c"));
        rep.rounds_used = r;
        if (r <= 64)
        {
            rep.e_round[r - 1][0] = ea.e;
            rep.e_round[r - 1][1] = eb.e;
            rep.e_round[r - 1][2] = ec.e;
        }
        const double drop = ec.e > 0.0 ? (e_prev - ec.e) / ec.e : 0.0;
        if (!o.quiet)
        {
            std::printf("round %3d  E %.6e %.6e %.6e  psnr %.2f  sampled psnr %.2f", r, ea.e, eb.e, ec.e,
                        mse_to_db(ec.centre_mse), k > 0 ? mse_to_db(ec.sampled_mse) : 0.0);
            if (np > 1)
            {
                std::printf("  mip psnr");
                for (int p = 1; p < np; p++)
                    std::printf(" M%d %.2f", p, mse_to_db(ec.centre_plane[p]));
            }
            std::printf("  moved0 %lld  moved1 %lld  minDiagH %.3e  %.1f ms\n", moved0, moved1, min_diag,
                        round_a + round_b + round_c);
            std::fflush(stdout);
        }
        e_prev = ec.e;
        const bool level1_still = frozen ? moved1 == 0 : delta_norm <= 1e-7 * (plane_norm > 0.0 ? plane_norm : 1.0);
        if (moved0 == 0 && level1_still)
            break;
        tol_hits = drop < o.tol ? tol_hits + 1 : 0;
        if (tol_hits >= 2)
            break;
    }

    // ---- what the asset stores
    if (frozen)
    {
        cks(S.sweep_level1_all(d, &st, k, o.ridge, o.q1_sweeps, lr.data(), &t), "sweep_level1_all");
        ms_b += t;
    }
    else
        cks(S.quantise_level1(d, &st, o.q1_range), "quantise_level1");
    if (palette)
    {
        // one last search against the level 1 the asset stores, one last block (a), and the asset as it stands
        moved0_search = search();
        cks(S.solve_decoder(d, &st, k, ps, &t), "solve_decoder");
        ms_a += t;
        Totals tt;
        tt.ms_a = ms_a;
        tt.ms_b = ms_b;
        tt.ms_c = ms_c;
        tt.obj_ms0 = obj_ms0;
        tt.t_start = t_start;
        tt.moved0_final = moved0_search;
        std::memcpy(tt.tally0, tally0, sizeof(tally0));
        finish_palette(E, tt);
        return;
    }
    if (frozen)
    {
        cks(S.sweep_level0_cont_all(d, &st, k, o.ridge, o.q1_sweeps, lr0.data(), &t), "sweep_level0_cont_all");
        ms_c += t;
    }
    else
        cks(S.quantise_level0(d, &st, o.q1_range), "quantise_level0");
    cks(S.solve_decoder(d, &st, k, ps, &t), "solve_decoder");
    ms_a += t;
    // Model::k0, the index mirror the bc calls read, stays on the device: a seeding pack takes the device's indices as
    // they stand (where the reference downloads them), the other calls the mirror already held.
    auto mirror = []() -> const uint8_t* const* { return nullptr; };

    // ---- the post-fit BC pack, its refinement and the refit
    double pack_psnr = 0.0, pack_psnr_seed = 0.0, ms_pack = 0.0, ms_refine = 0.0;
    size_t pack_texels = 0;
    auto refit_level1 = [&]() {
        double tb = 0.0;
        cks(S.sweep_level1_all(d, &st, k, o.ridge, o.q1_sweeps, lr.data(), &tb), "sweep_level1_all");
        ms_b += tb;
    };
    obj(probe);
    const double e_prepack = probe.e;
    cks(S.bc_pack_prepare(d, &st, k, 1, mirror(), &pack_psnr_seed, &pack_texels, &t), "bc_pack_prepare");
    ms_pack += t;
    obj(probe);
    const double e_packed = probe.e;
    pack_psnr = pack_psnr_seed;
    std::vector<nntc_bc_refine_report> rr((size_t)std::max(1, std::max(o.bc_refine, o.bc_refine_after)));
    cks(S.bc_refine(d, o.bc_refine, mirror(), rr.data(), &pack_psnr, &pack_texels, &t), "bc_refine");
    ms_refine += t;
    obj(probe);
    const double e_refined = probe.e;
    refit_level1();
    cks(S.solve_decoder(d, &st, k, ps, &t), "solve_decoder");
    ms_a += t;
    obj(probe);
    double e_refit = probe.e, psnr_refit = mse_to_db(probe.centre_mse);
    if (!o.quiet)
        std::printf("level 0 packed: E %.9e before the pack, %.9e after it, %.9e after the refinement, %.9e after "
                    "the refit; packing psnr %.2f -> %.2f dB\n",
                    e_prepack, e_packed, e_refined, e_refit, pack_psnr_seed, pack_psnr);
    if (o.bc_refine_after > 0)
    {
        cks(S.bc_pack_prepare(d, &st, k, 0, mirror(), &pack_psnr, &pack_texels, &t), "bc_pack_prepare");
        ms_pack += t;
        cks(S.bc_refine(d, o.bc_refine_after, mirror(), rr.data(), &pack_psnr, &pack_texels, &t), "bc_refine");
        ms_refine += t;
        obj(probe);
        e_refit = probe.e;
        psnr_refit = mse_to_db(probe.centre_mse);
    }
    double e_ship = e_refit, psnr_ship = psnr_refit;

    // ---- the outer repack loop
    rep.outer_taken = rep.outer_tried = 0;
    for (int pass = 1; pass <= o.bc_outer; pass++)
    {
        rep.outer_tried++;
        // The state a rejected pass is put back to, kept on the device (the reference's host round trip,
        // src/main.cpp:2561-2588).
        cks(S.state_snapshot(d), "state_snapshot");
        const nntc_model_state keep_state = st;
        const double keep_pack_psnr = pack_psnr, keep_pack_psnr_seed = pack_psnr_seed;
        const size_t keep_pack_texels = pack_texels;

        refit_level1();
        cks(S.solve_decoder(d, &st, k, ps, &t), "solve_decoder");
        ms_a += t;
        if (frozen)
            cks(S.sweep_level0_cont_all(d, &st, k, o.ridge, o.q1_sweeps, lr0.data(), &t), "sweep_level0_cont_all");
        else
        {
            cks(S.solve_level0_cont_all(d, &st, k, o.ridge, lr0.data(), &t), "solve_level0_cont_all");
            cks(S.snap_level0_on_grid(d, &st), "snap_level0_on_grid");
        }
        ms_c += t;
        cks(S.bc_pack_prepare(d, &st, k, 1, mirror(), &pack_psnr_seed, &pack_texels, &t), "bc_pack_prepare");
        ms_pack += t;
        cks(S.bc_refine(d, o.bc_refine, mirror(), rr.data(), &pack_psnr, &pack_texels, &t), "bc_refine");
        ms_refine += t;
        obj(probe);
        if (probe.e < e_ship)
        {
            if (!o.quiet)
                std::printf("level 0 outer : pass %d ACCEPTED: shipped E %.9e psnr %.2f -> E %.9e psnr %.2f\n", pass,
                            e_ship, psnr_ship, probe.e, mse_to_db(probe.centre_mse));
            e_ship = probe.e;
            psnr_ship = mse_to_db(probe.centre_mse);
            rep.outer_taken++;
            continue;
        }
        if (!o.quiet)
            std::printf("level 0 outer : pass %d rejected: shipped E %.9e -> E %.9e, the previous pass restored\n",
                        pass, e_ship, probe.e);
        cks(S.state_restore(d), "state_restore");
        st = keep_state;
        pack_psnr = keep_pack_psnr;
        pack_psnr_seed = keep_pack_psnr_seed;
        pack_texels = keep_pack_texels;
        cks(S.upload_decoder_to_device(d, &st), "upload_decoder_to_device");
        cks(S.upload_grids_to_device(d, &st), "upload_grids_to_device");
        // The blocks the file holds are a function of the restored endpoints and selectors.
        cks(S.bc_blocks_from_device(d, mirror(), nullptr, nullptr), "bc_blocks_from_device");
        break;
    }

    // ---- one last free refit
    refit_level1();
    cks(S.solve_decoder(d, &st, k, ps, &t), "solve_decoder");
    ms_a += t;
    obj(probe);
    if (!o.quiet)
        std::printf("level 0 final : one more refit of level 1 and the decoder against the packed plane: shipped E "
                    "%.9e psnr %.2f -> E %.9e psnr %.2f\n",
                    e_ship, psnr_ship, probe.e, mse_to_db(probe.centre_mse));

    Totals tt;
    tt.ms_a = ms_a;
    tt.ms_b = ms_b;
    tt.ms_c = ms_c;
    tt.ms_pack = ms_pack;
    tt.ms_refine = ms_refine;
    tt.pack_psnr = pack_psnr;
    tt.pack_psnr_seed = pack_psnr_seed;
    tt.obj_ms0 = obj_ms0;
    tt.t_start = t_start;
    std::memcpy(tt.tally0, tally0, sizeof(tally0));
    (void)passes0;
    ship(E, tt);
}

}   // namespace

// ---------------------------------------------------------------------------------------------------------------------
// Encoder: the host half
// ---------------------------------------------------------------------------------------------------------------------

// The plane layout (src/mips.cpp:25-40), the mip weights and the model descriptor of textures x width x height.
void Encoder::describe(int textures, int width, int height)
{
    if (textures < 1 || textures > 6)
        throw Failure { "one to six textures" };
    if (opts.c0 < 1 || opts.c0 > 4 || opts.c1 < 1 || opts.c1 > 4)
        throw Failure { "--c0 and --c1 take 1..4" };
    if (width < 1 || height < 1)
        throw Failure { "an empty texture" };
    const int nout = 3 * textures;
    source_w = width;
    source_h = height;
    const int wp = (width + 3) / 4 * 4, hp = (height + 3) / 4 * 4;
    std::vector<nntc_plane_size> planes;
    plane_layout(wp, hp, opts.mips != 0, opts.mip_min, planes);
    mip_weights(planes, opts.mip_weight, opts.mip_mix, opts.k, per_site);

    std::memset(&desc, 0, sizeof(desc));
    desc.textures = textures;
    desc.nout = nout;
    desc.c0 = opts.c0;
    desc.c1 = opts.c1;
    for (int q = 0; q < 4; q++)
        desc.bits0[q] = 8;
    desc.bits1 = opts.bits1;
    desc.bc0 = 1;
    desc.l0_bc8 = 1;
    if (opts.l0_palette)
    {
        // --l0 palette: an index of 1-4 bits per level-0 channel on a fixed palette, block-compressed or not
        // (src/main.cpp:1295-1356)
        if (opts.bc0 != 0 && opts.bc0 != 1)
            throw Failure { "--bc0 must be 0 or 1" };
        if (opts.sweeps < 1)
            throw Failure { "--sweeps takes at least 1" };
        for (int q = 0; q < opts.c0; q++)
        {
            if (opts.bits0[q] < 1 || opts.bits0[q] > 4)
                throw Failure { "--bits0 " + std::to_string(opts.bits0[q]) + " is out of range (1..4)" };
            desc.bits0[q] = opts.bits0[q];
        }
        desc.bc0 = opts.bc0;
        desc.l0_bc8 = 0;
    }
    // --init0 residual or texture both seed from the residual; luma does not (src/main.cpp:1904-1906).
    desc.init0_residual = opts.init0_luma ? 0 : 1;
    desc.init0_texture = opts.init0_texture && !opts.init0_luma ? 1 : 0;
    desc.init0_chain = opts.init0_chain;
    desc.nplanes = (int)planes.size();
    for (size_t i = 0; i < planes.size(); i++)
        desc.planes[i] = planes[i];
    for (int c = 0; c < nout; c++)
    {
        const double rs = opts.rgb_weights[0] + opts.rgb_weights[1] + opts.rgb_weights[2];
        desc.cw[c] = (float)(opts.weights[c / 3] * opts.rgb_weights[c % 3] * 3.0 / rs);
    }
}

void Encoder::setup(int textures, int width, int height, const uint8_t* const* rgb)
{
    describe(textures, width, height);
    const int nout = desc.nout, wp = desc.planes[0].w0, hp = desc.planes[0].h0, np = desc.nplanes;
    // The packed source, padded by edge replication (src/image.h:298-331, src/main.cpp:1865-1884), then the box chain
    // (src/image.h:343-366), each plane's rows shared among host threads (every value is computed on its own, so the
    // split changes no bits).
    chain.assign((size_t)np, std::vector<float>());
    chain[0].resize((size_t)wp * hp * nout);
    const float inv = 1.0f / 255.0f;
    float* base = chain[0].data();
    parallel_rows(hp, [&](int y0, int y1) {
        for (int y = y0; y < y1; y++)
        {
            const int sy = std::min(y, height - 1);
            for (int x = 0; x < wp; x++)
            {
                const int sx = std::min(x, width - 1);
                for (int t = 0; t < textures; t++)
                    for (int c = 0; c < 3; c++)
                        base[((size_t)y * wp + x) * nout + 3 * t + c] =
                            (float)rgb[t][((size_t)sy * width + sx) * 3 + c] * inv;
            }
        }
    });
    for (int i = 1; i < np; i++)
    {
        const nntc_plane_size &a = desc.planes[i - 1], &b = desc.planes[i];
        chain[(size_t)i].resize((size_t)b.w0 * b.h0 * nout);
        parallel_rows(b.h0, [&](int y0, int y1) {
            box_halve_rows(chain[(size_t)i - 1], a.w0, a.h0, nout, chain[(size_t)i], b.w0, y0, y1);
        });
    }
    std::vector<const float*> ptrs((size_t)np);
    for (int i = 0; i < np; i++)
        ptrs[(size_t)i] = chain[(size_t)i].data();
    ctx = seam.device_create(&desc, ptrs.data());
    if (!ctx)
        throw Failure { std::string("device_create: ") + seam.last_error() };
    if (opts.form != NNTC_FORM_EXACT)
        check(seam.set_form(ctx, opts.form), "set_form");
    if (opts.matrix_free != -1 && seam.name == "hip")
        check(nntc_hip_set_matrix_free(ctx, opts.matrix_free), "set_matrix_free");
    if (opts.sweep_lanes != -1 && seam.name == "hip")
        check(nntc_hip_set_sweep_lanes(ctx, opts.sweep_lanes), "set_sweep_lanes");
    if (opts.graphs != -1 && seam.name == "hip")
        check(nntc_hip_set_graphs(ctx, opts.graphs), "set_graphs");
}

void Encoder::setup_device(int textures, int width, int height, const uint8_t* const* rgb_device)
{
    if (seam.name != "hip")
        throw Failure { "textures in device memory need the gfx950 arm" };
    describe(textures, width, height);
    ctx = nntc_hip_device_create_rgb8(&desc, rgb_device, width, height);
    if (!ctx)
        throw Failure { std::string("device_create_rgb8: ") + nntc_hip_last_error() };
    chain_on_device = true;
    if (opts.form != NNTC_FORM_EXACT)
        check(seam.set_form(ctx, opts.form), "set_form");
    if (opts.matrix_free != -1 && seam.name == "hip")
        check(nntc_hip_set_matrix_free(ctx, opts.matrix_free), "set_matrix_free");
    if (opts.sweep_lanes != -1 && seam.name == "hip")
        check(nntc_hip_set_sweep_lanes(ctx, opts.sweep_lanes), "set_sweep_lanes");
    if (opts.graphs != -1 && seam.name == "hip")
        check(nntc_hip_set_graphs(ctx, opts.graphs), "set_graphs");
}

// The box init takes the source channels with every exact copy of an earlier channel moved to the back; when that
// changes the SET of channels level 1 starts from, the block means are formed here and uploaded
// (src/main.cpp:1938-2027).
void Encoder::box_reorder()
{
    const int nout = desc.nout, c1 = desc.c1;
    std::vector<int> order, copies;
    uint32_t dev_copies[NNTC_MAX_NOUT] = {};
    if (chain_on_device)
        check(nntc_hip_source_copies(ctx, dev_copies), "source_copies");
    for (int s = 0; s < nout; s++)
    {
        bool copy = chain_on_device && dev_copies[s] != 0;
        for (int e = 0; e < s && !copy && !chain_on_device; e++)
        {
            bool same = true;
            for (size_t i = 0; i < chain.size() && same; i++)
            {
                const std::vector<float>& v = chain[i];
                for (size_t t = 0; t < v.size() / (size_t)nout && same; t++)
                    same = v[t * nout + s] == v[t * nout + e];
            }
            copy = same;
        }
        (copy ? copies : order).push_back(s);
    }
    order.insert(order.end(), copies.begin(), copies.end());
    const int taken = std::min(c1, nout);
    std::vector<int> chosen(order.begin(), order.begin() + taken);
    std::sort(chosen.begin(), chosen.end());
    bool plain = true;
    for (int c = 0; c < taken; c++)
        plain = plain && chosen[(size_t)c] == c;
    if (plain)
        return;
    if (chain_on_device && chain.empty())
    {
        // The reordered init reads the chain on the host: fetch it once from the device.
        chain.assign((size_t)desc.nplanes, std::vector<float>());
        for (int i = 0; i < desc.nplanes; i++)
        {
            chain[(size_t)i].resize((size_t)desc.planes[i].w0 * desc.planes[i].h0 * nout);
            check(nntc_hip_source_download(ctx, i, chain[(size_t)i].data(), chain[(size_t)i].size(), nullptr),
                  "source_download");
        }
    }
    for (int i = 0; i < desc.nplanes; i++)
    {
        const nntc_plane_size& pl = desc.planes[i];
        const int w0 = pl.w0, h0 = pl.h0, w1 = pl.w1, h1 = pl.h1;
        std::vector<float> v1((size_t)w1 * h1 * c1);
        for (int y = 0; y < h1; y++)
            for (int x = 0; x < w1; x++)
            {
                int xa = (int)((long long)x * w0 / w1), xb = (int)((long long)(x + 1) * w0 / w1);
                int ya = (int)((long long)y * h0 / h1), yb = (int)((long long)(y + 1) * h0 / h1);
                xb = std::min(w0, xb <= xa ? xa + 1 : xb);
                yb = std::min(h0, yb <= ya ? ya + 1 : yb);
                const size_t t1 = (size_t)y * w1 + x;
                for (int c = 0; c < c1; c++)
                {
                    if (c >= nout)
                    {
                        v1[t1 * c1 + c] = 0.0f;
                        continue;
                    }
                    const int sc = order[(size_t)c];
                    float acc = 0.0f;
                    int n = 0;
                    for (int sy = ya; sy < yb; sy++)
                        for (int sx = xa; sx < xb; sx++)
                        {
                            acc += chain[(size_t)i][((size_t)sy * w0 + sx) * nout + sc];
                            n++;
                        }
                    const float inv = n > 0 ? 1.0f / (float)n : 0.0f;
                    v1[t1 * c1 + c] = 2.0f * acc * inv - 1.0f;
                }
            }
        check(seam.level1_upload(ctx, i, v1.data(), v1.size()), "level1_upload");
    }
    check(seam.quantise_level1(ctx, &state, NNTC_RANGE_MINMAX), "quantise_level1");
    if (!opts.quiet)
        std::printf("level 1 init: exact copies of earlier source channels are taken last\n");
}

void Encoder::verify_level1_on_grid()
{
    size_t values = 0, off = 0;
    check(seam.level1_off_grid(ctx, &state, &values, &off), "level1_off_grid");
    if (off)
        throw Failure { std::to_string(off) + " of " + std::to_string(values) +
                        " level-1 values are not the grid value of the index beside them" };
}

void Encoder::decode_base()
{
    const nntc_plane_size& pl = desc.planes[0];
    recon.resize((size_t)pl.w0 * pl.h0 * desc.nout);
    check(seam.decode_plane(ctx, 0, recon.data(), recon.size()), "decode_plane");
}

// image_psnr_texture (src/image.h:387-407) over the original extent, both sides rounded to 8 bits, from the arm's
// integer error sum (the reference's double sum of the same integer squares has the same value).
double Encoder::texture_psnr(int t) const
{
    const size_t n = (size_t)source_w * source_h * 3;
    if (n == 0)
        return 0.0;
    const double mse = (double)texture_se[t] / (double)n;
    return mse > 0.0 ? 10.0 * std::log10(255.0 * 255.0 / mse) : 100.0;
}

Encoder::~Encoder()
{
    if (ctx)
        seam.device_destroy(ctx);
}

// ---------------------------------------------------------------------------------------------------------------------
// The C ABI
// ---------------------------------------------------------------------------------------------------------------------

extern "C"
{

void nntc_encode_default_options(nntc_encode_options* o)
{
    std::memset(o, 0, sizeof(*o));
    o->c0 = 2;
    o->c1 = 4;
    o->bits1 = 8;
    o->k = 4;
    o->rounds = 20;
    o->tol = 1e-4;
    o->q1_start = 3;
    o->q1_sweeps = 4;
    o->q1_range = NNTC_RANGE_PCT;
    o->ridge = 1e-4;
    o->mips = 1;
    o->mip_min = 8;
    o->mip_weight = 0;
    o->mip_mix = 0.5;
    o->bc_refine = 4;
    o->bc_refine_after = 0;
    o->bc_outer = 2;
    o->init0_chain = 0;
    o->form = NNTC_FORM_EXACT;
    o->quiet = 0;
    for (int t = 0; t < 6; t++)
        o->weights[t] = 1.0;
    for (int c = 0; c < 3; c++)
        o->rgb_weights[c] = 1.0;
    o->matrix_free = -1;
    o->sweep_lanes = -1;
    o->graphs = -1;
    o->init1_pca = 0;
    o->init0_texture = 0;
    o->init0_luma = 0;
    o->l0_palette = 0;
    for (int q = 0; q < 4; q++)
        o->bits0[q] = 3;
    o->bc0 = 1;
    o->sweeps = 2;
}

const char* nntc_encode_last_error(void)
{
    return g_error.c_str();
}

nntc_encoder* nntc_encoder_create(const nntc_encode_options* o, int textures, int width, int height,
                                  const uint8_t* const* rgb, const char* arm_lib, const char* arm_prefix)
{
    Encoder* e = new Encoder();
    try
    {
        e->opts = *o;
        if (!e->seam.load(arm_lib, arm_prefix))
            throw Failure { g_error };
        e->setup(textures, width, height, rgb);
        return reinterpret_cast<nntc_encoder*>(e);
    }
    catch (const Failure& f)
    {
        g_error = f.what;
    }
    delete e;
    return nullptr;
}

nntc_encoder* nntc_encoder_create_device(const nntc_encode_options* o, int textures, int width, int height,
                                         const uint8_t* const* rgb_device)
{
    Encoder* e = new Encoder();
    try
    {
        e->opts = *o;
        if (!e->seam.load(nullptr, nullptr))
            throw Failure { g_error };
        if (!rgb_device)
            throw Failure { "no textures" };
        e->setup_device(textures, width, height, rgb_device);
        return reinterpret_cast<nntc_encoder*>(e);
    }
    catch (const Failure& f)
    {
        g_error = f.what;
    }
    delete e;
    return nullptr;
}

nntc_encoder* nntc_encoder_create_files(const nntc_encode_options* o, int textures, const char* const* paths,
                                        const char* arm_lib, const char* arm_prefix)
{
    if (textures < 1 || textures > 6)
    {
        g_error = "one to six textures";
        return nullptr;
    }
    // The textures are read side by side, one host thread each (PNG decoding is sequential within a file).
    std::vector<std::vector<uint8_t>> images((size_t)textures);
    std::vector<int> ws((size_t)textures, 0), hs((size_t)textures, 0), alphas((size_t)textures, 0);
    std::vector<std::string> whys((size_t)textures);
    std::vector<char> ok((size_t)textures, 0);
    {
        std::vector<std::thread> readers;
        for (int t = 0; t < textures; t++)
            readers.emplace_back([&, t] {
                bool alpha = false;
                ok[(size_t)t] = load_png_rgb(paths[t], images[(size_t)t], ws[(size_t)t], hs[(size_t)t], alpha,
                                             whys[(size_t)t]);
                alphas[(size_t)t] = alpha;
            });
        for (std::thread& r : readers)
            r.join();
    }
    int w0 = 0, h0 = 0;
    for (int t = 0; t < textures; t++)
    {
        if (!ok[(size_t)t])
        {
            g_error = std::string("cannot read '") + paths[t] + "': " + whys[(size_t)t];
            return nullptr;
        }
        if (alphas[(size_t)t])
            std::printf("WARNING: %s has an alpha channel; it is ignored (the tool encodes 24-bit RGB)\n", paths[t]);
        if (t == 0)
        {
            w0 = ws[0];
            h0 = hs[0];
        }
        else if (ws[(size_t)t] != w0 || hs[(size_t)t] != h0)
        {
            g_error = std::string("'") + paths[t] + "' is not the size of '" + paths[0] + "'";
            return nullptr;
        }
    }
    std::vector<const uint8_t*> ptrs((size_t)textures);
    for (int t = 0; t < textures; t++)
        ptrs[(size_t)t] = images[(size_t)t].data();
    nntc_encoder* e = nntc_encoder_create(o, textures, w0, h0, ptrs.data(), arm_lib, arm_prefix);
    if (e)
        reinterpret_cast<Encoder*>(e)->inputs.assign(paths, paths + textures);
    return e;
}

int nntc_encoder_run(nntc_encoder* h)
{
    Encoder& e = *reinterpret_cast<Encoder*>(h);
    try
    {
        e.report = nntc_encode_report {};
        e.report.textures = e.desc.textures;
        e.report.nplanes = e.desc.nplanes;
        e.report.mpix = (double)e.source_w * e.source_h * e.desc.textures / 1e6;
        e.packed0.clear();
        run_encode(e);
        e.ran = true;
        return 0;
    }
    catch (const Failure& f)
    {
        g_error = f.what;
    }
    return 1;
}

void nntc_encoder_destroy(nntc_encoder* h)
{
    delete reinterpret_cast<Encoder*>(h);
}

nntc_hip_ctx* nntc_encoder_context(const nntc_encoder* h)
{
    const Encoder* e = reinterpret_cast<const Encoder*>(h);
    return e->seam.name == "hip" ? e->ctx : nullptr;
}

int nntc_encoder_report(const nntc_encoder* h, nntc_encode_report* r)
{
    *r = reinterpret_cast<const Encoder*>(h)->report;
    return 0;
}

int nntc_encoder_state(const nntc_encoder* h, nntc_model_state* s)
{
    *s = reinterpret_cast<const Encoder*>(h)->state;
    return 0;
}

int nntc_encoder_desc(const nntc_encoder* h, nntc_model_desc* d)
{
    *d = reinterpret_cast<const Encoder*>(h)->desc;
    return 0;
}

int nntc_encoder_level1_indices(const nntc_encoder* h, int plane, uint8_t* out, size_t n, size_t* count)
{
    const Encoder& e = *reinterpret_cast<const Encoder*>(h);
    return e.seam.level1_download_indices(e.ctx, plane, out, n, count);
}

int nntc_encoder_level0_blocks(const nntc_encoder* h, int plane, int file, uint8_t* out, size_t n, size_t* count)
{
    const Encoder& e = *reinterpret_cast<const Encoder*>(h);
    return e.seam.bc_blocks_download(e.ctx, plane, file, out, n, count);
}

int nntc_encoder_recon(const nntc_encoder* h, uint8_t* out, size_t n, size_t* count)
{
    Encoder& e = *const_cast<Encoder*>(reinterpret_cast<const Encoder*>(h));
    if (e.ran && e.recon.empty())
    {
        try
        {
            e.decode_base();
        }
        catch (const Failure& f)
        {
            g_error = f.what;
            return 1;
        }
    }
    if (count)
        *count = e.recon.size();
    std::memcpy(out, e.recon.data(), std::min(n, e.recon.size()));
    return 0;
}

}   // extern "C"
