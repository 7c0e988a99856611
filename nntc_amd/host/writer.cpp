// writer.cpp: the asset - PREFIX_lat0.dds (PREFIX_lat0a/b.dds for three or four level-0 channels), PREFIX_lat1.dds
// and PREFIX_nntc.json - in the reference's format (src/export.cpp:89-147, 470-812, docs/FORMAT.md).
//
// .dds: the DX10 header (magic + 31-dword DDS_HEADER + 5-dword DDS_HEADER_DXT10 = 148 bytes, little-endian), then
// every mip level base first, tightly packed. Level 0 is BC4_UNORM (dxgi 80) or BC5_UNORM (dxgi 83) holding the blocks
// the refinement chose; level 1 is R8 / R8G8 / R8G8B8A8_UNORM (61 / 49 / 28) holding each index replicated over its
// byte, a three-channel plane padded to four with 255. The header fields are the reference's: the flags for a chain
// or a single level, the base level's linear size (compressed) or row pitch (uncompressed).
//
// JSON: the same keys and values a consumer reads (sizes, block, LOD bias, per-level files, formats, bits, mip sizes,
// the dequantisation lo/hi/levels, and the decoder's C0, C1, nin, nout, weights and bias), every float as %.9g.

#include "nntc_host.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

using namespace nntc_host;

namespace
{

bool dds_write(const std::string& name, int w, int h, int dxgi, bool compressed, uint32_t size_field,
               const std::vector<std::vector<uint8_t>>& levels)
{
    const bool chain = levels.size() > 1;
    uint32_t hdr[37] = { 0 };
    hdr[0] = 0x20534444u;                                                    // "DDS "
    hdr[1] = 124;                                                            // dwSize
    hdr[2] = 0x00001007u | (chain ? 0x00020000u : 0u) | (compressed ? 0x00080000u : 0x8u);
    hdr[3] = (uint32_t)h;
    hdr[4] = (uint32_t)w;
    hdr[5] = size_field;
    hdr[6] = 1;
    hdr[7] = (uint32_t)levels.size();
    hdr[19] = 32;
    hdr[20] = 0x4u;                                                          // DDPF_FOURCC
    hdr[21] = 0x30315844u;                                                   // "DX10"
    hdr[27] = chain ? 0x00401008u : 0x00001000u;
    hdr[32] = (uint32_t)dxgi;
    hdr[33] = 3;                                                             // TEXTURE2D
    hdr[35] = 1;                                                             // arraySize
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f)
    {
        g_error = "cannot write '" + name + "'";
        return false;
    }
    bool ok = std::fwrite(hdr, 4, 37, f) == 37;
    for (const std::vector<uint8_t>& level : levels)
        ok = ok && std::fwrite(level.data(), 1, level.size(), f) == level.size();
    if (std::fclose(f) != 0)
        ok = false;
    if (!ok)
        g_error = "writing '" + name + "' failed";
    return ok;
}

std::string num(double x)
{
    char b[32];
    std::snprintf(b, sizeof(b), "%.9g", std::isfinite(x) ? x : 0.0);
    return b;
}

std::string nums(const float* v, size_t n)
{
    std::string s = "[";
    for (size_t i = 0; i < n; i++)
        s += num(v[i]) + (i + 1 < n ? "," : "");
    return s + "]";
}

std::string base_name(const std::string& p)
{
    const size_t slash = p.find_last_of("/\\");
    return slash == std::string::npos ? p : p.substr(slash + 1);
}

// The standard BC4 palette of endpoints r0, r1 (the Direct3D functional specification): r0 > r1 gives eight values,
// r0, r1 and ((8 - i) r0 + (i - 1) r1) / 7; otherwise six, ((6 - i) r0 + (i - 1) r1) / 5 between them, then 0 and 255.
void bc4_palette(int r0, int r1, float pal[8])
{
    pal[0] = (float)r0;
    pal[1] = (float)r1;
    if (r0 > r1)
    {
        for (int i = 2; i < 8; i++)
            pal[i] = (float)((8 - i) * r0 + (i - 1) * r1) / 7.0f;
        return;
    }
    for (int i = 2; i < 6; i++)
        pal[i] = (float)((6 - i) * r0 + (i - 1) * r1) / 5.0f;
    pal[6] = 0.0f;
    pal[7] = 255.0f;
}

// One channel's 4x4 block of index bytes (each an index of `bits` bits replicated over its byte) into a BC4 block;
// returns the squared byte error against the exact index values k * 255 / (2^bits - 1) (shared/bc_pack.h).
//
// At 1-3 bits fixed endpoints make the block's palette contain the whole index grid, so the pack is a relabelling:
//   1 bit   (0, 255), six-value mode: 0 and 255 are entries 6 and 7;
//   2 bits  (85, 170), six-value mode: 0, 85, 170, 255 are entries 6, 0, 1, 7;
//   3 bits  (255, 0), eight-value mode: entry 1 is 0, entry 0 is 255, and entry 8 - k is k * 255 / 7 between.
// At 4 bits sixteen values do not fit in eight: the block's highest and lowest index values, rounded to bytes, are the
// endpoints (the high one first, which selects the eight-value mode) and each texel takes the nearest entry, the
// first of equals.
double bc4_index_block(const uint8_t v[16], int bits, uint8_t blk[8])
{
    static const int ENDPOINTS[3][2] = { { 0, 255 }, { 85, 170 }, { 255, 0 } };
    static const int ENTRY1[2] = { 6, 7 };
    static const int ENTRY2[4] = { 6, 0, 1, 7 };
    static const int ENTRY3[8] = { 1, 7, 6, 5, 4, 3, 2, 0 };
    const int levels = (1 << bits) - 1;
    int r0 = 0, r1 = 0, sel[16] = {};
    double err = 0.0;
    if (bits <= 3)
    {
        r0 = ENDPOINTS[bits - 1][0];
        r1 = ENDPOINTS[bits - 1][1];
        const int* entry = bits == 1 ? ENTRY1 : (bits == 2 ? ENTRY2 : ENTRY3);
        for (int t = 0; t < 16; t++)
            sel[t] = entry[v[t] >> (8 - bits)];
    }
    else
    {
        int lo = levels, hi = 0;
        for (int t = 0; t < 16; t++)
        {
            lo = std::min(lo, v[t] >> (8 - bits));
            hi = std::max(hi, v[t] >> (8 - bits));
        }
        r0 = (int)std::lround(hi * 255.0 / levels);
        r1 = (int)std::lround(lo * 255.0 / levels);
        if (r0 != r1)
        {
            float pal[8];
            bc4_palette(r0, r1, pal);
            for (int t = 0; t < 16; t++)
            {
                const float want = (float)(v[t] >> (8 - bits)) * 255.0f / (float)levels;
                float best = std::fabs(pal[0] - want);
                for (int i = 1; i < 8; i++)
                {
                    const float d = std::fabs(pal[i] - want);
                    if (d < best)
                    {
                        best = d;
                        sel[t] = i;
                    }
                }
                err += (double)best * (double)best;
            }
        }
    }
    blk[0] = (uint8_t)r0;
    blk[1] = (uint8_t)r1;
    uint64_t packed = 0;
    for (int t = 0; t < 16; t++)
        packed |= (uint64_t)(sel[t] & 7) << (3 * t);
    for (int i = 0; i < 6; i++)
        blk[2 + i] = (uint8_t)(packed >> (8 * i));
    return err;
}

}   // namespace

namespace nntc_host
{

void bc4_block_values(const uint8_t blk[8], float out[16])
{
    float pal[8];
    bc4_palette(blk[0], blk[1], pal);
    uint64_t packed = 0;
    for (int i = 0; i < 6; i++)
        packed |= (uint64_t)blk[2 + i] << (8 * i);
    for (int t = 0; t < 16; t++)
        out[t] = pal[(packed >> (3 * t)) & 7];
}

void level0_index_bytes(const uint8_t* k0, size_t texels, int c0, const int* bits0, std::vector<uint8_t>& bytes)
{
    const int stored = c0 == 3 ? 4 : c0;
    bytes.assign(texels * (size_t)stored, 255);
    uint8_t rep[4][16];
    for (int c = 0; c < c0; c++)
        for (int k = 0; k < (1 << bits0[c]); k++)
            rep[c][k] = (uint8_t)replicate_index(k, bits0[c]);
    const size_t rows = (texels + 4095) / 4096;
    parallel_rows((int)rows, [&](int r0, int r1) {
        const size_t end = std::min(texels, (size_t)r1 * 4096);
        for (size_t t = (size_t)r0 * 4096; t < end; t++)
            for (int c = 0; c < c0; c++)
                bytes[t * stored + c] = rep[c][k0[t * c0 + c] & ((1 << bits0[c]) - 1)];
    });
}

std::vector<uint8_t> bc_pack_indices(const uint8_t* bytes, int w, int h, int stored, int first, int nc,
                                     const int* bits, int used, double& se, size_t& n)
{
    const int bx = (w + 3) / 4, by = (h + 3) / 4, bpb = nc == 1 ? 8 : 16;
    std::vector<uint8_t> out((size_t)bx * by * bpb);
    // Block rows packed side by side on host threads; each block's error is kept and the errors are added afterwards
    // in block order, so the sums do not depend on the split.
    std::vector<double> err((size_t)bx * by * nc);
    parallel_rows(by, [&](int y0, int y1) {
        for (int y = y0; y < y1; y++)
            for (int x = 0; x < bx; x++)
                for (int c = 0; c < nc; c++)
                {
                    uint8_t v[16];
                    for (int t = 0; t < 16; t++)
                    {
                        const int px = std::min(w - 1, 4 * x + (t & 3)), py = std::min(h - 1, 4 * y + (t >> 2));
                        v[t] = bytes[((size_t)py * w + px) * stored + first + c];
                    }
                    err[((size_t)y * bx + x) * nc + c] =
                        bc4_index_block(v, bits[first + c], &out[((size_t)y * bx + x) * bpb + 8 * c]);
                }
    });
    for (int y = 0; y < by; y++)
        for (int x = 0; x < bx; x++)
        {
            // the texels of the block inside the plane (the rest repeat the last row and column)
            const int inside = (std::min(w, 4 * x + 4) - 4 * x) * (std::min(h, 4 * y + 4) - 4 * y);
            for (int c = 0; c < nc; c++)
                if (first + c < used)
                {
                    se += err[((size_t)y * bx + x) * nc + c] * inside / 16.0;
                    n += (size_t)inside;
                }
        }
    return out;
}

}   // namespace nntc_host

extern "C" int nntc_encoder_write(const nntc_encoder* h, const char* prefix_c)
{
    const Encoder& e = *reinterpret_cast<const Encoder*>(h);
    if (!e.ran)
    {
        g_error = "nothing to write: the encoder has not run";
        return 1;
    }
    const nntc_model_desc& d = e.desc;
    const nntc_model_state& st = e.state;
    const int nin = d.c1 + d.c0 + d.c0 * d.c1;
    for (int i = 0; i < d.nout * nin; i++)
        if (!std::isfinite(st.w[i]))
        {
            g_error = "the decoder carries a value that is not finite; the asset is not written";
            return 1;
        }
    const std::string prefix = prefix_c;
    const int np = d.nplanes;
    // The .dds files, written side by side on host threads once every level is on the host (the context is driven
    // from this thread only, so the downloads stay here).
    struct PendingFile
    {
        std::string name;
        int w, h, dxgi;
        bool compressed;
        uint32_t size_field;
        std::vector<std::vector<uint8_t>> levels;
    };
    std::vector<PendingFile> pending;
    const int files = d.c0 <= 2 ? 1 : 2;
    std::string js = "{\n  \"format\": \"nntc-dds-1\",\n";
    js += "  \"source\": { \"width\": " + std::to_string(e.source_w) + ", \"height\": " + std::to_string(e.source_h);
    if (e.source_w != d.planes[0].w0 || e.source_h != d.planes[0].h0)
        js += ", \"padded\": true, \"padded_width\": " + std::to_string(d.planes[0].w0) + ", \"padded_height\": " +
              std::to_string(d.planes[0].h0);
    // What each input was (informational): the driver reads every input as linear RGB8 with the box chain and edge
    // clamping, so only the name and the two weights vary. Textures handed over as arrays have no file name.
    js += ", \"inputs\": [";
    for (int t = 0; t < d.textures; t++)
    {
        const std::string file = (size_t)t < e.inputs.size() ? base_name(e.inputs[(size_t)t]) : std::string();
        js += std::string(t ? ", " : "") + "{ \"file\": \"" + file + "\", \"type\": \"\", \"filter\": \"default\", "
              "\"srgb\": false, \"edge\": \"clamp\", \"normal_map\": false, \"weight\": " + num(e.opts.weights[t]) +
              ", \"rgb_weights\": [" + num(e.opts.rgb_weights[0]) + "," + num(e.opts.rgb_weights[1]) + "," +
              num(e.opts.rgb_weights[2]) + "] }";
    }
    js += "]";
    js += " },\n";
    js += "  \"decode\": { \"width\": " + std::to_string(d.planes[0].w0) + ", \"height\": " +
          std::to_string(d.planes[0].h0) + ", \"textures_out\": " + std::to_string(d.textures) + " },\n";
    js += "  \"block\": 4,\n  \"lod_bias_level1\": 2,\n";
    js += "  \"sampling\": \"both latents are sampled bilinearly at the same coordinate, level 1 with MipLODBias = "
          "lod_bias_level1; output mip m reads mip m of both; dequantise after sampling: value = lo + sample * "
          "(hi - lo) per channel\",\n";
    js += "  \"textures\": [\n";

    std::string sizes0 = "[", sizes1 = "[";
    for (int i = 0; i < np; i++)
    {
        sizes0 += "[" + std::to_string(d.planes[i].w0) + "," + std::to_string(d.planes[i].h0) + "]" +
                  (i + 1 < np ? "," : "");
        sizes1 += "[" + std::to_string(d.planes[i].w1) + "," + std::to_string(d.planes[i].h1) + "]" +
                  (i + 1 < np ? "," : "");
    }
    sizes0 += "]";
    sizes1 += "]";

    // ---- level 0. --l0 bc8: the refined BC4 / BC5 blocks, as the device holds them. --l0 palette: the indices,
    // replicated over their bytes, packed here (block-compressed) or stored as they are (--bc0 0).
    const bool palette = d.l0_bc8 == 0;
    const bool compress = !palette || d.bc0 != 0;
    const int stored0 = d.c0 == 3 ? 4 : d.c0;
    // The blocks the encoder already packed for its report, when it did (lossy indices): the same indices give the
    // same blocks, so they are not packed a second time.
    const bool packed = palette && compress && e.packed0.size() == (size_t)files;
    std::vector<std::vector<uint8_t>> bytes0;
    if (palette && !packed)
    {
        bytes0.resize((size_t)np);
        for (int i = 0; i < np; i++)
        {
            const size_t n = (size_t)d.planes[i].w0 * d.planes[i].h0;
            std::vector<uint8_t> k(n * d.c0);
            if (e.seam.level0_download(e.ctx, i, k.data(), k.size(), nullptr) != 0)
            {
                g_error = std::string("level0_download: ") + e.seam.last_error();
                return 1;
            }
            level0_index_bytes(k.data(), n, d.c0, d.bits0, bytes0[(size_t)i]);
        }
    }
    std::string files_js;
    if (compress)
        for (int f = 0; f < files; f++)
        {
            const int nc = level0_file_channels(d.c0, f);
            const int dxgi = nc == 1 ? 80 : 83;
            const int bpb = nc == 1 ? 8 : 16;
            std::vector<std::vector<uint8_t>> levels((size_t)np);
            for (int i = 0; i < np; i++)
            {
                if (packed)
                {
                    levels[(size_t)i] = e.packed0[(size_t)f][(size_t)i];
                    continue;
                }
                if (palette)
                {
                    double se = 0.0;
                    size_t n = 0;
                    levels[(size_t)i] = bc_pack_indices(bytes0[(size_t)i].data(), d.planes[i].w0, d.planes[i].h0,
                                                        stored0, 2 * f, nc, d.bits0, d.c0, se, n);
                    continue;
                }
                const size_t nb = (size_t)((d.planes[i].w0 + 3) / 4) * ((d.planes[i].h0 + 3) / 4);
                levels[(size_t)i].resize(nb * bpb);
                if (e.seam.bc_blocks_download(e.ctx, i, f, levels[(size_t)i].data(), levels[(size_t)i].size(),
                                              nullptr) != 0)
                {
                    g_error = std::string("bc_blocks_download: ") + e.seam.last_error();
                    return 1;
                }
            }
            const std::string file = prefix + "_lat0" + (files == 2 ? (f == 0 ? "a" : "b") : "") + ".dds";
            const uint32_t linear = (uint32_t)levels[0].size();
            pending.push_back({ file, d.planes[0].w0, d.planes[0].h0, dxgi, true, linear, std::move(levels) });
            const std::string fname = nc == 1 ? "BC4_UNORM" : "BC5_UNORM";
            if (files == 2)
                files_js += std::string(f ? ", " : "") + "{ \"file\": \"" + base_name(file) + "\", \"dxgi_format\": \"" +
                            fname + "\", \"dxgi_format_id\": " + std::to_string(dxgi) + ", \"channels_stored\": " +
                            std::to_string(nc) + " }";
            else
                files_js += "\"" + base_name(file) + "\"";
        }
    int dxgi0 = 0, stored_out0 = 0;
    std::string format0;
    if (compress)
    {
        stored_out0 = level0_file_channels(d.c0, 0);
        dxgi0 = stored_out0 == 1 ? 80 : 83;
        format0 = stored_out0 == 1 ? "BC4_UNORM" : "BC5_UNORM";
    }
    else
    {
        stored_out0 = stored0;
        dxgi0 = d.c0 == 1 ? 61 : (d.c0 == 2 ? 49 : 28);
        format0 = d.c0 == 1 ? "R8_UNORM" : (d.c0 == 2 ? "R8G8_UNORM" : "R8G8B8A8_UNORM");
        const std::string file = prefix + "_lat0.dds";
        pending.push_back({ file, d.planes[0].w0, d.planes[0].h0, dxgi0, false, (uint32_t)(d.planes[0].w0 * stored0),
                            std::move(bytes0) });
        files_js = "\"" + base_name(file) + "\"";
    }
    const bool two_files = compress && files == 2;
    js += "    {\n      " + std::string(two_files ? "\"files\": [" + files_js + "]" : "\"file\": " + files_js) +
          ", \"level\": 0, \"width\": " + std::to_string(d.planes[0].w0) + ", \"height\": " +
          std::to_string(d.planes[0].h0) + ",\n";
    js += "      \"dxgi_format\": \"" + format0 + "\", \"dxgi_format_id\": " + std::to_string(dxgi0) +
          ", \"channels_used\": " + std::to_string(d.c0) + ", \"channels_stored\": " + std::to_string(stored_out0) +
          ",\n";
    if (compress)
        js += "      \"bc_palette\": \"standard\",\n";
    js += "      \"bits_per_channel\": [";
    for (int c = 0; c < d.c0; c++)
        js += std::to_string(palette ? d.bits0[c] : 8) + (c + 1 < d.c0 ? "," : "");
    js += "],\n      \"mip_count\": " + std::to_string(np) + ", \"mip_sizes\": " + sizes0 + ",\n";
    if (!palette)
    {
        js += "      \"texel\": \"an 8-bit continuous value, BC4 / BC5 encoded\",\n";
        js += "      \"dequantise\": { \"kind\": \"range\", \"lo\": " + nums(st.lo0, (size_t)d.c0) + ", \"hi\": " +
              nums(st.hi0, (size_t)d.c0) + ", \"levels\": 255 }\n    },\n";
    }
    else
    {
        // The palette each index decodes to under THIS file's format, then its ends and its number of steps.
        js += std::string("      \"texel\": \"") +
              (compress ? "BC4 / BC5 blocks of the index, lossless at 1-3 bits" : "the index replicated over its byte") +
              "\",\n      \"dequantise\": { \"kind\": \"palette\", \"palette\": [";
        std::string lo = "[", hi = "[", lv = "[";
        for (int c = 0; c < d.c0; c++)
        {
            const int levels = (1 << d.bits0[c]) - 1;
            js += nums(st.palette0[c], (size_t)levels + 1) + (c + 1 < d.c0 ? "," : "");
            lo += num(st.palette0[c][0]) + (c + 1 < d.c0 ? "," : "");
            hi += num(st.palette0[c][levels]) + (c + 1 < d.c0 ? "," : "");
            lv += std::to_string(levels) + (c + 1 < d.c0 ? "," : "");
        }
        js += "], \"lo\": " + lo + "], \"hi\": " + hi + "], \"levels\": " + lv + "] }\n    },\n";
    }

    // ---- level 1: the indices replicated over their bytes, three channels padded to four
    {
        const int ch = d.c1, stored = ch == 3 ? 4 : ch;
        const int dxgi = stored == 1 ? 61 : (stored == 2 ? 49 : 28);
        std::vector<std::vector<uint8_t>> levels((size_t)np);
        for (int i = 0; i < np; i++)
        {
            const size_t n = (size_t)d.planes[i].w1 * d.planes[i].h1;
            std::vector<uint8_t> k(n * ch);
            if (e.seam.level1_download_indices(e.ctx, i, k.data(), k.size(), nullptr) != 0)
            {
                g_error = std::string("level1_download_indices: ") + e.seam.last_error();
                return 1;
            }
            levels[(size_t)i].assign(n * stored, 255);
            for (size_t t = 0; t < n; t++)
                for (int c = 0; c < ch; c++)
                    levels[(size_t)i][t * stored + c] = (uint8_t)replicate_index(k[t * ch + c], d.bits1);
        }
        const std::string file = prefix + "_lat1.dds";
        pending.push_back({ file, d.planes[0].w1, d.planes[0].h1, dxgi, false, (uint32_t)(d.planes[0].w1 * stored),
                            std::move(levels) });
        const std::string fname = stored == 1 ? "R8_UNORM" : (stored == 2 ? "R8G8_UNORM" : "R8G8B8A8_UNORM");
        js += "    {\n      \"file\": \"" + base_name(file) + "\", \"level\": 1, \"width\": " +
              std::to_string(d.planes[0].w1) + ", \"height\": " + std::to_string(d.planes[0].h1) + ",\n";
        js += "      \"dxgi_format\": \"" + fname + "\", \"dxgi_format_id\": " + std::to_string(dxgi) +
              ", \"channels_used\": " + std::to_string(ch) + ", \"channels_stored\": " + std::to_string(stored) +
              ",\n      \"bits_per_channel\": [";
        for (int c = 0; c < ch; c++)
            js += std::to_string(d.bits1) + (c + 1 < ch ? "," : "");
        js += "],\n      \"mip_count\": " + std::to_string(np) + ", \"mip_sizes\": " + sizes1 + ",\n";
        js += "      \"texel\": \"the quantisation index in the byte's top bits, replicated\",\n";
        js += "      \"dequantise\": { \"kind\": \"range\", \"lo\": " + nums(st.lo1, (size_t)ch) + ", \"hi\": " +
              nums(st.hi1, (size_t)ch) + ", \"levels\": " + std::to_string((1 << d.bits1) - 1) + " }\n    }\n";
    }
    js += "  ],\n  \"decoder\": {\n    \"type\": \"bilinear\",\n    \"terms\": \"a b sc\", \"C0\": " +
          std::to_string(d.c0) + ", \"C1\": " + std::to_string(d.c1) + ",\n";
    js += "    \"features\": \"[c_j for j < C1][s_i for i < C0][s_i c_j, i outer, j inner]\",\n";
    js += "    \"nin\": " + std::to_string(nin) + ", \"nout\": " + std::to_string(d.nout) +
          ", \"hidden\": [], \"activation\": \"leaky_relu\", \"leak\": 0,\n    \"output\": \"identity\",\n";
    js += "    \"layers\": [\n      { \"rows\": " + std::to_string(d.nout) + ", \"cols\": " + std::to_string(nin) +
          ", \"weights\": " + nums(st.w, (size_t)d.nout * nin) + ", \"bias\": " + nums(st.b, (size_t)d.nout) +
          " }\n    ]\n  }\n}\n";
    std::vector<std::string> errors(pending.size());
    std::vector<std::thread> writers;
    for (size_t i = 0; i < pending.size(); i++)
        writers.emplace_back([&, i] {
            const PendingFile& pf = pending[i];
            if (!dds_write(pf.name, pf.w, pf.h, pf.dxgi, pf.compressed, pf.size_field, pf.levels))
                errors[i] = g_error;
        });
    for (std::thread& t : writers)
        t.join();
    for (const std::string& e : errors)
        if (!e.empty())
        {
            g_error = e;
            return 1;
        }
    const std::string jname = prefix + "_nntc.json";
    FILE* f = std::fopen(jname.c_str(), "wb");
    if (!f)
    {
        g_error = "cannot write '" + jname + "'";
        return 1;
    }
    const bool ok = std::fwrite(js.data(), 1, js.size(), f) == js.size();
    if (std::fclose(f) != 0 || !ok)
    {
        g_error = "writing '" + jname + "' failed";
        return 1;
    }
    return 0;
}
