// main.cpp: nntc_hip_encode, the command line of the reference's nntc_encode (src/main.cpp:256-352) for the encode
// path this arm builds, over include/nntc_encode.h.
//
//   nntc_hip_encode a.png [b.png ...] [--c0 N] [--c1 N] [--rounds R] [--tol X] [--k K] [--mips 0|1] [--mip-min S]
//                   [--mip-weight pixels|sqrt|uniform] [--mip-mix X] [--q1-start R0] [--q1-sweeps N]
//                   [--q1-range minmax|pct] [--ridge X] [--bc-refine N] [--bc-refine-after N] [--bc-outer N]
//                   [--init box|pca] [--init0 residual|texture|luma] [--init0-scope base|chain] [--weights W,..]
//                   [--rgb-weights R,G,B] [--l0 bc8|palette] [--bits0 B[,B..]] [--bc0 0|1] [--sweeps N]
//                   [--form exact|plain|fast] [--quiet] [-o PREFIX]
//
// Options the reference has and this path does not build (named mip filters, --backend cuda|cpu|check, --bc0 both,
// the recon PNGs) are refused by name rather than ignored.

#include "nntc_encode.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static bool parse_list(const char* s, std::vector<double>& out)
{
    out.clear();
    std::string t(s);
    size_t p = 0;
    while (p <= t.size())
    {
        const size_t q = t.find(',', p);
        const std::string part = t.substr(p, q == std::string::npos ? std::string::npos : q - p);
        char* end = nullptr;
        const double v = std::strtod(part.c_str(), &end);
        if (part.empty() || *end)
            return false;
        out.push_back(v);
        if (q == std::string::npos)
            break;
        p = q + 1;
    }
    return true;
}

int main(int argc, char** argv)
{
    nntc_encode_options o;
    nntc_encode_default_options(&o);
    std::vector<const char*> inputs;
    std::string prefix;
    std::vector<double> bits0;
    bool bc0_given = false, pack_flags = false;
    for (int i = 1; i < argc; i++)
    {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc)
            {
                std::fprintf(stderr, "ERROR: %s needs a value\n", a.c_str());
                std::exit(1);
            }
            return argv[++i];
        };
        auto refuse = [&](const char* why) {
            std::fprintf(stderr, "ERROR: %s %s\n", a.c_str(), why);
            std::exit(1);
        };
        if (a == "--c0")
            o.c0 = std::atoi(next());
        else if (a == "--c1")
            o.c1 = std::atoi(next());
        else if (a == "--bits1")
            o.bits1 = std::atoi(next());
        else if (a == "--rounds")
            o.rounds = std::atoi(next());
        else if (a == "--tol")
            o.tol = std::atof(next());
        else if (a == "--k")
            o.k = std::atoi(next());
        else if (a == "--mips")
            o.mips = std::atoi(next());
        else if (a == "--mip-min")
            o.mip_min = std::atoi(next());
        else if (a == "--mip-weight")
        {
            const std::string v = next();
            o.mip_weight = v == "pixels" ? 0 : v == "sqrt" ? 1 : v == "uniform" ? 2 : -1;
            if (o.mip_weight < 0)
                refuse("takes pixels, sqrt or uniform");
        }
        else if (a == "--mip-mix")
            o.mip_mix = std::atof(next());
        else if (a == "--q1-start")
            o.q1_start = std::atoi(next());
        else if (a == "--q1-sweeps")
            o.q1_sweeps = std::atoi(next());
        else if (a == "--q1-range")
        {
            const std::string v = next();
            if (v != "pct" && v != "minmax")
                refuse("takes minmax or pct");
            o.q1_range = v == "pct" ? NNTC_RANGE_PCT : NNTC_RANGE_MINMAX;
        }
        else if (a == "--ridge")
            o.ridge = std::atof(next());
        else if (a == "--bc-refine")
        {
            o.bc_refine = std::atoi(next());
            pack_flags = true;
        }
        else if (a == "--bc-refine-after")
        {
            o.bc_refine_after = std::atoi(next());
            pack_flags = true;
        }
        else if (a == "--bc-outer")
        {
            o.bc_outer = std::atoi(next());
            pack_flags = true;
        }
        else if (a == "--init0-scope")
        {
            const std::string v = next();
            if (v != "base" && v != "chain")
                refuse("takes base or chain");
            o.init0_chain = v == "chain";
        }
        else if (a == "--weights")
        {
            std::vector<double> w;
            if (!parse_list(next(), w) || w.size() > 6)
                refuse("takes one weight per texture");
            for (size_t t = 0; t < w.size(); t++)
                o.weights[t] = w[t];
        }
        else if (a == "--rgb-weights")
        {
            std::vector<double> w;
            if (!parse_list(next(), w) || w.size() != 3)
                refuse("takes R,G,B");
            for (int c = 0; c < 3; c++)
                o.rgb_weights[c] = w[(size_t)c];
        }
        else if (a == "--form")
        {
            const std::string v = next();
            if (v != "exact" && v != "fast" && v != "plain")
                refuse("takes exact, plain or fast");
            o.form = v == "exact" ? NNTC_FORM_EXACT : v == "plain" ? NNTC_FORM_PLAIN : NNTC_FORM_FAST;
        }
        else if (a == "--quiet")
            o.quiet = 1;
        else if (a == "-o")
            prefix = next();
        else if (a == "--l0")
        {
            const std::string v = next();
            if (v != "bc8" && v != "palette")
                refuse("must be bc8 or palette");
            o.l0_palette = v == "palette";
        }
        else if (a == "--bits0")
        {
            std::vector<double> b;
            if (!parse_list(next(), b) || b.empty() || b.size() > 4)
                refuse("takes B or one B per level-0 channel");
            bits0.assign(b.begin(), b.end());
        }
        else if (a == "--bc0")
        {
            const std::string v = next();
            if (v != "0" && v != "1")
                refuse("must be 0 or 1");
            o.bc0 = v == "1";
            bc0_given = true;
        }
        else if (a == "--sweeps")
            o.sweeps = std::atoi(next());
        else if (a == "--init")
        {
            const std::string v = next();
            if (v != "box" && v != "pca")
                refuse("must be box or pca");
            o.init1_pca = v == "pca";
        }
        else if (a == "--init0")
        {
            const std::string v = next();
            if (v != "residual" && v != "texture" && v != "luma")
                refuse("must be residual, texture or luma");
            o.init0_texture = v == "texture";
            o.init0_luma = v == "luma";
        }
        else if (a == "--backend")
        {
            if (std::string(next()) != "hip")
                refuse("this binary is the gfx950 arm: --backend hip");
        }
        else if (a == "--png")
        {
            if (std::atoi(next()) != 0)
                refuse("1 (the recon PNGs) is not built");
        }
        else if (!a.empty() && a[0] == '-')
            refuse("is not an option of this encoder");
        else
            inputs.push_back(argv[i]);
    }
    if (inputs.empty() || inputs.size() > 6)
    {
        std::fprintf(stderr, "usage: nntc_hip_encode a.png [b.png ... up to 6] [options]\n");
        return 1;
    }
    // --l0 bc8 fixes --bits0 and --bc0, and the BC pack's flags belong to it (src/main.cpp:1310-1356)
    if (!o.l0_palette && (!bits0.empty() || (bc0_given && o.bc0 == 0)))
    {
        std::fprintf(stderr, "ERROR: --bits0 and --bc0 0 apply under --l0 palette only\n");
        return 1;
    }
    if (o.l0_palette && pack_flags)
    {
        std::fprintf(stderr, "ERROR: --bc-refine, --bc-refine-after and --bc-outer belong to --l0 bc8\n");
        return 1;
    }
    if (bits0.size() > 1 && (int)bits0.size() != o.c0)
    {
        std::fprintf(stderr, "ERROR: --bits0 gives %zu values for %d level-0 channels\n", bits0.size(), o.c0);
        return 1;
    }
    for (int q = 0; q < 4 && !bits0.empty(); q++)
        o.bits0[q] = (int)bits0[bits0.size() == 1 ? 0 : (size_t)std::min(q, (int)bits0.size() - 1)];
    if (prefix.empty())
    {
        prefix = inputs[0];
        const size_t dot = prefix.rfind('.');
        if (dot != std::string::npos)
            prefix = prefix.substr(0, dot);
    }
    nntc_encoder* e = nntc_encoder_create_files(&o, (int)inputs.size(), inputs.data(), nullptr, nullptr);
    if (!e)
    {
        std::fprintf(stderr, "ERROR: %s\n", nntc_encode_last_error());
        return 1;
    }
    if (nntc_encoder_run(e) != 0)
    {
        std::fprintf(stderr, "ERROR: %s\n", nntc_encode_last_error());
        nntc_encoder_destroy(e);
        return 1;
    }
    const int status = nntc_encoder_write(e, prefix.c_str());
    if (status != 0)
        std::fprintf(stderr, "ERROR: %s\n", nntc_encode_last_error());
    nntc_encoder_destroy(e);
    return status == 0 ? 0 : 1;
}
