// image.cpp: the encode's host-side image work - PNG in and out, the plane layout, the box mip chain, the mip
// weights - and the grid value the on-grid check compares against.
//
// The PNG reader takes 8-bit non-interlaced PNGs of every colour type (grey, RGB, palette, grey + alpha, RGBA) through
// zlib's inflate and the five scanline filters of the PNG specification, and hands back RGB the way the reference's
// loader does (src/image.h:123-160): a grey image replicated into three channels and an alpha channel dropped, with a
// flag when it was not opaque everywhere.

#include "nntc_host.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

namespace nntc_host
{

namespace
{

uint32_t be32(const uint8_t* p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}

int paeth(int a, int b, int c)
{
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc)
        return a;
    return pb <= pc ? b : c;
}

void put32(std::vector<uint8_t>& o, uint32_t v)
{
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

void chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data)
{
    put32(o, (uint32_t)data.size());
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    const uLong crc = crc32(0L, o.data() + start, (uInt)(o.size() - start));
    put32(o, (uint32_t)crc);
}

}   // namespace

bool load_png_rgb(const char* path, std::vector<uint8_t>& rgb, int& w, int& h, bool& alpha_ignored, std::string& why)
{
    alpha_ignored = false;
    FILE* f = std::fopen(path, "rb");
    if (!f)
    {
        why = "cannot open";
        return false;
    }
    std::vector<uint8_t> file;
    uint8_t buf[65536];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0)
        file.insert(file.end(), buf, buf + got);
    std::fclose(f);
    static const uint8_t sig[8] = { 137, 80, 78, 71, 13, 10, 26, 10 };
    if (file.size() < 8 || std::memcmp(file.data(), sig, 8) != 0)
    {
        why = "not a PNG";
        return false;
    }
    int depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat, plte;
    size_t pos = 8;
    while (pos + 12 <= file.size())
    {
        const uint32_t len = be32(&file[pos]);
        const char* type = (const char*)&file[pos + 4];
        if (pos + 12 + (size_t)len > file.size())
        {
            why = "truncated chunk";
            return false;
        }
        const uint8_t* data = &file[pos + 8];
        if (!std::memcmp(type, "IHDR", 4))
        {
            w = (int)be32(data);
            h = (int)be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
        }
        else if (!std::memcmp(type, "PLTE", 4))
            plte.assign(data, data + len);
        else if (!std::memcmp(type, "IDAT", 4))
            idat.insert(idat.end(), data, data + len);
        else if (!std::memcmp(type, "IEND", 4))
            break;
        pos += 12 + (size_t)len;
    }
    if (depth != 8 || interlace != 0 || w <= 0 || h <= 0)
    {
        why = "only 8-bit, non-interlaced PNGs are read";
        return false;
    }
    const int comp = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (comp == 0 || (ctype == 3 && plte.empty()))
    {
        why = "unsupported colour type";
        return false;
    }
    const size_t stride = (size_t)w * comp;
    std::vector<uint8_t> raw((stride + 1) * (size_t)h);
    uLongf rlen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rlen, idat.data(), (uLong)idat.size()) != Z_OK || rlen != raw.size())
    {
        why = "corrupt image data";
        return false;
    }
    std::vector<uint8_t> px(stride * (size_t)h);
    for (int y = 0; y < h; y++)
    {
        const uint8_t* in = &raw[(size_t)y * (stride + 1)];
        uint8_t* out = &px[(size_t)y * stride];
        const uint8_t* up = y > 0 ? &px[(size_t)(y - 1) * stride] : nullptr;
        const int filter = in[0];
        in++;
        for (size_t i = 0; i < stride; i++)
        {
            const int a = i >= (size_t)comp ? out[i - comp] : 0;
            const int b = up ? up[i] : 0;
            const int c = (up && i >= (size_t)comp) ? up[i - comp] : 0;
            int v = in[i];
            switch (filter)
            {
            case 0:
                break;
            case 1:
                v += a;
                break;
            case 2:
                v += b;
                break;
            case 3:
                v += (a + b) / 2;
                break;
            case 4:
                v += paeth(a, b, c);
                break;
            default:
                why = "bad scanline filter";
                return false;
            }
            out[i] = (uint8_t)v;
        }
    }
    rgb.resize((size_t)w * h * 3);
    for (size_t i = 0; i < (size_t)w * h; i++)
    {
        const uint8_t* s = &px[i * comp];
        uint8_t r, g, b;
        if (ctype == 3)
        {
            const size_t e = (size_t)s[0] * 3;
            r = e + 2 < plte.size() ? plte[e] : 0;
            g = e + 2 < plte.size() ? plte[e + 1] : 0;
            b = e + 2 < plte.size() ? plte[e + 2] : 0;
        }
        else if (comp >= 3)
        {
            r = s[0];
            g = s[1];
            b = s[2];
        }
        else
            r = g = b = s[0];
        if ((comp == 2 || comp == 4) && s[comp - 1] != 255)
            alpha_ignored = true;
        rgb[i * 3] = r;
        rgb[i * 3 + 1] = g;
        rgb[i * 3 + 2] = b;
    }
    return true;
}

bool save_png_rgb(const char* path, const uint8_t* rgb, int w, int h)
{
    std::vector<uint8_t> raw(((size_t)w * 3 + 1) * (size_t)h);
    for (int y = 0; y < h; y++)
    {
        raw[(size_t)y * ((size_t)w * 3 + 1)] = 0;
        std::memcpy(&raw[(size_t)y * ((size_t)w * 3 + 1) + 1], rgb + (size_t)y * w * 3, (size_t)w * 3);
    }
    uLongf zlen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return false;
    z.resize(zlen);
    std::vector<uint8_t> o = { 137, 80, 78, 71, 13, 10, 26, 10 };
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), { 8, 2, 0, 0, 0 });
    chunk(o, "IHDR", ihdr);
    chunk(o, "IDAT", z);
    chunk(o, "IEND", {});
    FILE* f = std::fopen(path, "wb");
    if (!f)
        return false;
    const bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
    return std::fclose(f) == 0 && ok;
}

// build_plane_sizes (src/mips.cpp:25-40) with mip_count (src/model.h:230-240) and mip_dim (src/image.h:335-338).
void plane_layout(int width, int height, bool mips, int mip_min, std::vector<nntc_plane_size>& planes)
{
    int levels = 0;
    if (mips)
    {
        int w = width, h = height;
        while (levels < 12 && w / 2 >= mip_min && h / 2 >= mip_min)
        {
            w /= 2;
            h /= 2;
            levels++;
        }
    }
    auto half = [](int n) { return n / 2 > 0 ? n / 2 : 1; };
    planes.assign((size_t)levels + 1, nntc_plane_size {});
    planes[0] = { width, height, width / 4, height / 4 };
    for (int i = 1; i <= levels; i++)
        planes[(size_t)i] = { half(planes[(size_t)i - 1].w0), half(planes[(size_t)i - 1].h0),
                              half(planes[(size_t)i - 1].w1), half(planes[(size_t)i - 1].h1) };
}

// image_box (src/image.h:343-366): 0.25 ((a + b) + (c + d)) of the 2x2 block, the last row/column repeated on a
// one-wide plane.
void box_halve(const std::vector<float>& in, int w, int h, int nc, std::vector<float>& out, int ow, int oh)
{
    out.resize((size_t)ow * oh * nc);
    for (int y = 0; y < oh; y++)
        for (int x = 0; x < ow; x++)
        {
            const int xa = 2 * x < w ? 2 * x : w - 1, xb = 2 * x + 1 < w ? 2 * x + 1 : w - 1;
            const int ya = 2 * y < h ? 2 * y : h - 1, yb = 2 * y + 1 < h ? 2 * y + 1 : h - 1;
            const float* a = &in[((size_t)ya * w + xa) * nc];
            const float* b = &in[((size_t)ya * w + xb) * nc];
            const float* c = &in[((size_t)yb * w + xa) * nc];
            const float* d = &in[((size_t)yb * w + xb) * nc];
            float* o = &out[((size_t)y * ow + x) * nc];
            for (int k = 0; k < nc; k++)
                o[k] = 0.25f * ((a[k] + b[k]) + (c[k] + d[k]));
        }
}

void box_halve_rows(const std::vector<float>& in, int w, int h, int nc, std::vector<float>& out, int ow, int y0,
                    int y1)
{
    for (int y = y0; y < y1; y++)
        for (int x = 0; x < ow; x++)
        {
            const int xa = 2 * x < w ? 2 * x : w - 1, xb = 2 * x + 1 < w ? 2 * x + 1 : w - 1;
            const int ya = 2 * y < h ? 2 * y : h - 1, yb = 2 * y + 1 < h ? 2 * y + 1 : h - 1;
            const float* a = &in[((size_t)ya * w + xa) * nc];
            const float* b = &in[((size_t)ya * w + xb) * nc];
            const float* c = &in[((size_t)yb * w + xa) * nc];
            const float* d = &in[((size_t)yb * w + xb) * nc];
            float* o = &out[((size_t)y * ow + x) * nc];
            for (int k = 0; k < nc; k++)
                o[k] = 0.25f * ((a[k] + b[k]) + (c[k] + d[k]));
        }
}

void parallel_rows(int rows, const std::function<void(int, int)>& body)
{
    const int hw = (int)std::thread::hardware_concurrency();
    const int n = std::max(1, std::min({ hw > 0 ? hw : 1, 16, rows / 64 }));
    if (n <= 1)
    {
        body(0, rows);
        return;
    }
    std::vector<std::thread> pool;
    for (int i = 0; i < n; i++)
    {
        const int y0 = (int)((long long)rows * i / n), y1 = (int)((long long)rows * (i + 1) / n);
        pool.emplace_back([&body, y0, y1] { body(y0, y1); });
    }
    for (std::thread& t : pool)
        t.join();
}

// build_mip_weights (src/mips.cpp:141-169): kind 0 pixels, 1 sqrt, 2 uniform.
void mip_weights(const std::vector<nntc_plane_size>& planes, int kind, double mix, int k, std::vector<double>& per_site)
{
    const size_t n = planes.size();
    per_site.assign(n, 0.0);
    if (n == 1)
    {
        per_site[0] = 1.0;
        return;
    }
    std::vector<double> r(n, 0.0), share(n, 0.0);
    double total = 0.0;
    for (size_t m = 1; m < n; m++)
    {
        const double pixels = (double)planes[m].w0 * (double)planes[m].h0;
        r[m] = kind == 2 ? 1.0 : (kind == 1 ? std::sqrt(pixels) : pixels);
        total += r[m];
    }
    share[0] = 1.0 - mix;
    for (size_t m = 1; m < n; m++)
        share[m] = total > 0.0 ? mix * r[m] / total : 0.0;
    for (size_t m = 0; m < n; m++)
    {
        const double sites = (double)planes[m].w0 * (double)planes[m].h0 * (double)(1 + k);
        per_site[m] = sites > 0.0 ? share[m] / sites : 0.0;
    }
}

int replicate_index(int k, int bits)
{
    unsigned b8 = 0;
    for (int sh = 8 - bits; sh > -bits; sh -= bits)
        b8 |= sh >= 0 ? ((unsigned)k << sh) : ((unsigned)k >> -sh);
    return (int)(b8 & 0xFFu);
}

float level1_value(float lo, float hi, int bits, int k)
{
    const double span = (double)hi - (double)lo;
    const double frac = (double)replicate_index(k, bits) / 255.0;
    return (float)((double)lo + frac * span);
}

float level0_value(int bits, int k, bool bc)
{
    const int levels = (1 << bits) - 1;
    const double frac = bc ? (double)k / (double)levels : (double)replicate_index(k, bits) / 255.0;
    return (float)(-1.0 + 2.0 * frac);
}

}   // namespace nntc_host
