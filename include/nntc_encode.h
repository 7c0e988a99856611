/* nntc_encode.h: the encode driver over the gfx950 arm - the reference's `nntc_encode` schedule as a C ABI.
 *
 * The reference's encoder is one function, encode() (src/main.cpp:1753-3300): load the textures, build the source
 * chain and the mip weights, create the device, initialise both latents, seed level 0 from the residual, run the
 * alternating rounds, freeze the grids, pack level 0 to BC4/BC5 and refine the pack, and report. This header exposes
 * the same run in two halves so that a caller can keep the device resident between runs:
 *
 *   nntc_encoder_create   the host half before the first seam call: the packed source, the chain, the plane layout,
 *                         the mip weights, the output weights, and device_create (the upload of the chain);
 *   nntc_encoder_run      everything from init_level0 to the shipped objective and the decode of the base plane,
 *                         restarted from scratch on every call (the device's latents, decoder and grids are reset);
 *
 * and the run's outputs through accessors. The arm is the gfx950 library by default; a different library that exports
 * the same entry points under another prefix may be named instead (the tests bind the reference CPU backend this way
 * to check the driver against the reference binary).
 */
#ifndef NNTC_ENCODE_H
#define NNTC_ENCODE_H

#include <stddef.h>
#include <stdint.h>

#include "nntc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The options of Options (src/options.h:67-191) that the encode path reads, with the reference's defaults. */
typedef struct nntc_encode_options
{
    int32_t c0, c1;              /* --c0 2, --c1 4 */
    int32_t bits1;               /* --bits1 8 */
    int32_t k;                   /* --k 4 */
    int32_t rounds;              /* --rounds 20 */
    double tol;                  /* --tol 1e-4 */
    int32_t q1_start;            /* --q1-start 3 */
    int32_t q1_sweeps;           /* --q1-sweeps 4 */
    int32_t q1_range;            /* --q1-range pct (NNTC_RANGE_PCT) */
    double ridge;                /* --ridge 1e-4 */
    int32_t mips;                /* --mips 1 */
    int32_t mip_min;             /* --mip-min 8 */
    int32_t mip_weight;          /* --mip-weight pixels: 0 pixels, 1 sqrt, 2 uniform */
    double mip_mix;              /* --mip-mix 0.5 */
    int32_t bc_refine;           /* --bc-refine 4 */
    int32_t bc_refine_after;     /* --bc-refine-after 0 */
    int32_t bc_outer;            /* --bc-outer 2 */
    int32_t init0_chain;         /* --init0-scope chain */
    int32_t form;                /* NNTC_FORM_EXACT, NNTC_FORM_PLAIN or NNTC_FORM_FAST (not a reference option) */
    int32_t quiet;               /* --quiet */
    double weights[6];           /* --weights (per texture; 1 when unset) */
    double rgb_weights[3];       /* --rgb-weights 1,1,1 */
    int32_t matrix_free;         /* the fast form's level-0 operator: -1 by size, 0 assembled, 1 matrix-free
                                    (nntc_hip_set_matrix_free; not a reference option) */
    int32_t sweep_lanes;         /* the fast sweeps at three or four channels: -1 the tuner's choice, 1 or 4 lanes
                                    a texel (nntc_hip_set_sweep_lanes; not a reference option) */
    int32_t graphs;              /* the fast conjugate gradients' probe periods as HIP graphs: -1 or 1 on, 0 off
                                    (nntc_hip_set_graphs; not a reference option) */
    int32_t init1_pca;           /* --init pca (1) or box (0, the default) */
    int32_t init0_texture;       /* --init0 texture (1) or residual (0, the default) */
    int32_t init0_luma;          /* --init0 luma (1): level 0's channel 0 from the luminance, no residual seed */
    int32_t l0_palette;          /* --l0 palette (1) or bc8 (0, the default) */
    int32_t bits0[4];            /* --bits0 under --l0 palette: bits per level-0 channel, 1..4 (default 3) */
    int32_t bc0;                 /* --bc0 under --l0 palette: 1 BC4 / BC5 (default), 0 an uncompressed level 0 */
    int32_t sweeps;              /* --sweeps 2: the palette search's coordinate sweeps above 8 bits a texel */
} nntc_encode_options;

void nntc_encode_default_options(nntc_encode_options* o);

typedef struct nntc_encoder nntc_encoder;

/* The host half and the device's creation. rgb[t] is texture t as width * height * 3 bytes, row-major RGB.
 * arm_lib / arm_prefix name the seam's implementation; NULL for both is the gfx950 arm. Returns NULL on failure
 * (nntc_encode_last_error says why). */
nntc_encoder* nntc_encoder_create(const nntc_encode_options* o, int textures, int width, int height,
                                  const uint8_t* const* rgb, const char* arm_lib, const char* arm_prefix);

/* The same from image files (PNG, 8-bit RGB / RGBA / grey), as `nntc_encode a.png b.png ...` reads them. */
nntc_encoder* nntc_encoder_create_files(const nntc_encode_options* o, int textures, const char* const* paths,
                                        const char* arm_lib, const char* arm_prefix);

/* The same from textures already in device memory: rgb_device[t] is texture t's width * height * 3 bytes (row-major
 * RGB) on the gfx950 device. The packed source and its mip chain are built there (nntc_hip_device_create_rgb8), so no
 * source crosses PCIe; the gfx950 arm only. */
nntc_encoder* nntc_encoder_create_device(const nntc_encode_options* o, int textures, int width, int height,
                                         const uint8_t* const* rgb_device);

/* One whole encode on the resident device. Returns 0, or nonzero with the reason in nntc_encode_last_error. */
int nntc_encoder_run(nntc_encoder* e);

void nntc_encoder_destroy(nntc_encoder* e);

/* The gfx950 arm's context the encoder drives (NULL for another arm), for the arm's own queries
 * (nntc_hip_device_memory, nntc_hip_graph_counts, ...). Owned by the encoder. */
nntc_hip_ctx* nntc_encoder_context(const nntc_encoder* e);

const char* nntc_encode_last_error(void);

/* ---- the last run's outputs */
typedef struct nntc_encode_report
{
    int32_t textures, nplanes, rounds_used;
    int32_t outer_taken, outer_tried;
    double psnr[6];              /* per texture, 8-bit, base plane, original extent (src/image.h:387-407) */
    double centre_psnr, sampled_psnr;
    double e_shipped;
    double e_round[64][3];       /* per round: E after blocks (a), (b), (c') (under --l0 palette: (c), the search) */
    double pack_psnr, pack_psnr_seed;  /* the refined and the seed BC pack's psnr; under --l0 palette both are the
                                          host pack's psnr against the exact index values (100 when lossless) */
    long long moved0_final;      /* --l0 palette: the last search's moved texels over every plane */
    double seconds;              /* the run's wall clock */
    double ms_a, ms_b, ms_c, ms_e, ms_pack, ms_refine;
    long long ridge_tally[4];
    double mpix;                 /* width * height * textures / 1e6 */
} nntc_encode_report;

int nntc_encoder_report(const nntc_encoder* e, nntc_encode_report* r);
int nntc_encoder_state(const nntc_encoder* e, nntc_model_state* s);
/* Plane `plane`'s level-1 indices (w1*h1*c1 bytes) and level-0 .dds blocks of file `file`, as the asset stores them
 * (--l0 bc8; under --l0 palette the writer packs the indices on the host and there are no device blocks). */
int nntc_encoder_level1_indices(const nntc_encoder* e, int plane, uint8_t* out, size_t n, size_t* count);
int nntc_encoder_level0_blocks(const nntc_encoder* e, int plane, int file, uint8_t* out, size_t n, size_t* count);
/* The base plane decoded to bytes (w0*h0*3T). */
int nntc_encoder_recon(const nntc_encoder* e, uint8_t* out, size_t n, size_t* count);
/* The model descriptor the encoder built (plane layout, channels, output weights). */
int nntc_encoder_desc(const nntc_encoder* e, nntc_model_desc* d);

/* Write the asset: PREFIX_lat0.dds (or _lat0a/_lat0b), PREFIX_lat1.dds and PREFIX_nntc.json (src/export.cpp:470). */
int nntc_encoder_write(const nntc_encoder* e, const char* prefix);

/* ---- a library of materials (no counterpart in the reference, which encodes one material per process) */
typedef struct nntc_material
{
    int32_t textures, width, height;   /* 1..6 textures of width x height */
    int32_t c0, c1;                    /* the layout (the options' c0 and c1 are replaced by these) */
    int32_t on_device;                 /* rgb[t] is in gfx950 device memory (as nntc_encoder_create_device), else host */
    const uint8_t* const* rgb;         /* texture t: width * height * 3 bytes, row-major RGB */
    const char* prefix;                /* the asset's prefix (nntc_encoder_write), or NULL to write nothing */
} nntc_material;

/* Encodes `count` materials on the gfx950 arm, `workers` at a time: each worker is a host thread that owns one
 * encoder at a time (its own context, so its own stream), takes the next material from a queue ordered largest first
 * (width * height * textures), encodes it with the options `o`, writes its asset and frees it. The launches of up to
 * `workers` materials are therefore in flight together, and one material's host decisions overlap the others'
 * kernels. Every material runs the launches an encode of it alone runs, in the same order on its own stream: its
 * outputs are that encode's bits. reports[i] (may be NULL) receives material i's report and status[i] (may be NULL)
 * 0 or nonzero. Returns how many materials failed; nntc_encode_last_error names the first. */
int nntc_encode_library(const nntc_material* materials, int count, const nntc_encode_options* o, int workers,
                        nntc_encode_report* reports, int32_t* status);

#ifdef __cplusplus
}
#endif

#endif /* NNTC_ENCODE_H */
