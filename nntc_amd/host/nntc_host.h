// nntc_host.h: the encode driver's internals - the seam's function table, the encoder state, the host helpers.
#pragma once

#include "nntc_encode.h"
#include "nntc_hip.h"

#include <chrono>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace nntc_host
{

extern thread_local std::string g_error;

struct Failure
{
    std::string what;
};

// Every seam entry point the driver calls, by name (the C ABI of include/nntc_hip.h minus the prefix).
#define NNTC_SEAM_FUNCTIONS(X)                                                                                         \
    X(last_error)                                                                                                      \
    X(device_create)                                                                                                   \
    X(device_destroy)                                                                                                  \
    X(device_memory)                                                                                                   \
    X(set_form)                                                                                                        \
    X(init_level0)                                                                                                     \
    X(init_level1)                                                                                                     \
    X(init0_residual_channel)                                                                                          \
    X(quantise_level1)                                                                                                 \
    X(freeze_level1_grid)                                                                                              \
    X(quantise_level0)                                                                                                 \
    X(freeze_level0_grid)                                                                                              \
    X(snap_level0_on_grid)                                                                                             \
    X(solve_decoder)                                                                                                   \
    X(decoder_ridge_tally)                                                                                             \
    X(solve_level1_all)                                                                                                \
    X(sweep_level1_all)                                                                                                \
    X(solve_level0_cont_all)                                                                                           \
    X(sweep_level0_cont_all)                                                                                           \
    X(level1_download_indices)                                                                                         \
    X(level1_upload)                                                                                                   \
    X(level1_upload_indices)                                                                                           \
    X(level0_upload_indices)                                                                                           \
    X(upload_decoder_to_device)                                                                                        \
    X(upload_grids_to_device)                                                                                          \
    X(level1_download)                                                                                                 \
    X(level1_off_grid)                                                                                                 \
    X(level0_download)                                                                                                 \
    X(level0_upload)                                                                                                   \
    X(level0_download_values)                                                                                          \
    X(decode_plane)                                                                                                    \
    X(texture_sse)                                                                                                     \
    X(objective_eval)                                                                                                  \
    X(objective_total_ms)                                                                                              \
    X(objective_passes)                                                                                                \
    X(bc_pack_prepare)                                                                                                 \
    X(bc_refine)                                                                                                       \
    X(bc_blocks_from_device)                                                                                           \
    X(bc_blocks_download)                                                                                              \
    X(bc_state_save)                                                                                                   \
    X(bc_state_restore)                                                                                                \
    X(solve_level0_all)                                                                                                \
    X(state_snapshot)                                                                                                  \
    X(state_restore)

struct Seam
{
#define NNTC_MEMBER(n) decltype(&nntc_hip_##n) n = nullptr;
    NNTC_SEAM_FUNCTIONS(NNTC_MEMBER)
#undef NNTC_MEMBER
    std::string name;
    bool load(const char* lib_path, const char* prefix);
};

struct Encoder
{
    Seam seam;
    nntc_encode_options opts {};
    nntc_model_desc desc {};
    nntc_model_state state {};
    nntc_hip_ctx* ctx = nullptr;
    std::vector<std::vector<float>> chain;   // the source at every stored level, interleaved 3T channels
    std::vector<double> per_site;            // the mip weights' per-site weight of every plane
    std::vector<uint8_t> recon;              // the base plane decoded to bytes, fetched when first asked for
    unsigned long long texture_se[NNTC_MAX_NOUT / 3] = {};   // the shipped decode's squared byte error per texture
    std::vector<std::string> inputs;
    int source_w = 0, source_h = 0;
    nntc_encode_report report {};
    bool ran = false;
    // --l0 palette with lossy BC4 indices (a channel above 3 bits): the blocks packed for the report, [file][plane].
    // The report's plane is decoded from them and the writer stores them; a run clears them.
    std::vector<std::vector<std::vector<uint8_t>>> packed0;

    // Wall time per seam call name, when NNTC_CALL_TIMES is set in the environment: each check() charges the time
    // since the previous one (the call itself and the host work that prepared it) to the call's name.
    bool call_times = false;
    std::vector<std::pair<std::string, std::pair<double, long long>>> call_ms;
    std::chrono::steady_clock::time_point last_check {};

    void check(int status, const char* what);
    void print_call_times();
    bool chain_on_device = false;            // the chain was built on the device; `chain` is fetched if needed
    void describe(int textures, int width, int height);
    void setup(int textures, int width, int height, const uint8_t* const* rgb);
    void setup_device(int textures, int width, int height, const uint8_t* const* rgb_device);
    void box_reorder();
    void verify_level1_on_grid();
    void decode_base();
    double texture_psnr(int t) const;
    ~Encoder();
};

// image.cpp
bool load_png_rgb(const char* path, std::vector<uint8_t>& rgb, int& w, int& h, bool& alpha_ignored, std::string& why);
bool save_png_rgb(const char* path, const uint8_t* rgb, int w, int h);
void plane_layout(int width, int height, bool mips, int mip_min, std::vector<nntc_plane_size>& planes);
void box_halve(const std::vector<float>& in, int w, int h, int nc, std::vector<float>& out, int ow, int oh);
// Rows [y0, y1) of box_halve's output, which is already sized ow x oh x nc.
void box_halve_rows(const std::vector<float>& in, int w, int h, int nc, std::vector<float>& out, int ow, int y0,
                    int y1);
// body(y0, y1) over [0, rows) in contiguous pieces, one host thread each (up to the hardware's threads, 16 at most).
void parallel_rows(int rows, const std::function<void(int, int)>& body);
void mip_weights(const std::vector<nntc_plane_size>& planes, int kind, double mix, int k, std::vector<double>& per_site);

// level1_value (src/model.h:159-168), for the host's on-grid check.
int replicate_index(int k, int bits);
float level1_value(float lo, float hi, int bits, int k);
// --l0 palette's value of index k (src/model.h:205-210): -1 + 2 k / (2^bits - 1) when level 0 is block-compressed,
// -1 + 2 rep(k) / 255 when it is stored uncompressed - what the shipped format decodes the index to.
float level0_value(int bits, int k, bool bc);

// writer.cpp: --l0 palette's level 0 as the asset stores it (src/export.cpp:235-246, shared/bc_pack.h).
// One plane's indices (texels x c0) as the stored bytes: each index replicated over its byte, three channels stored
// as four with 255 in the fourth.
void level0_index_bytes(const uint8_t* k0, size_t texels, int c0, const int* bits0, std::vector<uint8_t>& bytes);
// The channels first .. first + nc - 1 of a w x h plane of `stored` bytes a texel into BC4 (nc 1) or BC5 (nc 2)
// blocks, row-major, 8 or 16 bytes a block; se / n accumulate the squared byte error against the exact index values
// over the texels inside the plane, for channels below `used`.
std::vector<uint8_t> bc_pack_indices(const uint8_t* bytes, int w, int h, int stored, int first, int nc,
                                     const int* bits, int used, double& se, size_t& n);
// The sixteen values of one BC4 block, the standard palette evaluated in float.
void bc4_block_values(const uint8_t blk[8], float out[16]);
// How level 0's channels are spread over .dds files: BC4 for one, BC5 for two, BC5 + BC4 for three, two BC5s for four.
inline int level0_files(int c0)
{
    return c0 <= 2 ? 1 : 2;
}
inline int level0_file_channels(int c0, int file)
{
    return file == 0 ? (c0 == 1 ? 1 : 2) : c0 - 2;
}

}   // namespace nntc_host
