/* nntc_hip.h: the C ABI of the MI355X (gfx950) arm of the NNTC encoder.
 *
 * The reference encoder reaches its device work through one seam of 46 C++ functions in namespace `be`
 * (src/backend.h:86-150), each taking `be::Device*`, `Model&` and `std::vector` arguments. This header is the same seam
 * as a C ABI: an opaque context in place of `be::Device*`, a plain model descriptor and a plain mutable model state in
 * place of `Model&`, and pointer + length pairs in place of vectors. Every entry point names the `be::` function it
 * replaces. A caller in the reference's own tree binds these from a `nntc_hip::` namespace that converts the plain
 * structs to and from `Model` (INTEGRATION.md shows that binding).
 *
 * Conventions.
 *   - Every entry point that can fail returns int: 0 on success, nonzero on failure, with the reason readable from
 *     nntc_hip_last_error(). The reference prints one `ERROR:` line and exits (src/device.cuh:45-76); the C ABI
 *     reports instead, and the C++ layer above it keeps the reference's exit behaviour.
 *   - Block functions write their elapsed milliseconds to *ms (may be NULL), as the reference returns them.
 *   - `model` is the layout the encode was created with; `state` is the part of the reference's Model that the seam
 *     mutates (decoder, grids, palette) and it is read and written by the calls that read or write those fields.
 *   - All arrays are host memory. Plane buffers are interleaved [y][x][channel], exactly the reference's layout.
 *   - One context is driven from one host thread. Results never depend on scheduling: every reduction is folded in a
 *     fixed order, so repeated runs are byte-identical.
 */
#ifndef NNTC_HIP_H
#define NNTC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NNTC_MAX_PLANES 13   /* MAX_MIP_LEVELS + 1 (src/model.h:32) */
#define NNTC_MAX_CHANNELS 4  /* per-level channel cap (src/cpu/cpu_constants.h:24) */
#define NNTC_MAX_NIN 24      /* |phi| cap (src/cpu/cpu_constants.h:25) */
#define NNTC_MAX_NOUT 18     /* 3 * MAX_TEXTURES (src/model.h:39) */

/* One stored mip level: level 0's plane and level 1's plane (src/model.h:42-46). */
typedef struct nntc_plane_size
{
    int32_t w0, h0, w1, h1;
} nntc_plane_size;

/* The fixed layout of one encode: the fields of Model (src/model.h:56-114) that no seam function changes. */
typedef struct nntc_model_desc
{
    int32_t textures;                       /* T */
    int32_t nout;                           /* 3 T */
    int32_t c0, c1;                         /* latent channels */
    int32_t bits0[NNTC_MAX_CHANNELS];       /* level-0 bits per channel (8 under --l0 bc8) */
    int32_t bits1;                          /* level-1 bits */
    int32_t bc0;                            /* level 0 shipped block-compressed */
    int32_t l0_bc8;                         /* --l0 bc8 */
    int32_t init0_residual, init0_texture, init0_chain;
    int32_t nplanes;
    nntc_plane_size planes[NNTC_MAX_PLANES];
    float cw[NNTC_MAX_NOUT];                /* per-output objective weight (src/main.cpp:1856-1862) */
} nntc_model_desc;

/* The mutable part of Model: the decoder (Model::dec), both grids and the palette. */
typedef struct nntc_model_state
{
    int32_t has_decoder;                    /* Model::dec.w non-empty */
    float w[NNTC_MAX_NOUT * NNTC_MAX_NIN];  /* nout x nin, row-major */
    float b[NNTC_MAX_NOUT];
    int32_t has_grid0, has_grid1;           /* Model::lo0 / lo1 non-empty */
    float lo0[NNTC_MAX_CHANNELS], hi0[NNTC_MAX_CHANNELS];
    float lo1[NNTC_MAX_CHANNELS], hi1[NNTC_MAX_CHANNELS];
    int32_t has_palette0;                   /* palette mode only */
    float palette0[NNTC_MAX_CHANNELS][16];
} nntc_model_state;

/* DeviceInfo (src/model.h:255-261). */
typedef struct nntc_device_info
{
    char name[256];
    int32_t major, minor;
    char error[256];
} nntc_device_info;

/* Level1InitReport (src/model.h:281-289). */
typedef struct nntc_level1_init_report
{
    int32_t pca, components;
    double eigenvalue[4], peak[4], total_variance;
} nntc_level1_init_report;

/* Init0ChannelReport (src/model.h:314-326). */
typedef struct nntc_init0_report
{
    double eigenvalue, total_variance, share, peak, divisor;
    double dir[NNTC_MAX_NOUT];
    int32_t nout, texture;
} nntc_init0_report;

/* Level1Report (src/model.h:374-398); one per plane. */
typedef struct nntc_level1_report
{
    int32_t iterations;
    double residual, lambda, min_diag, max_diag;
    double ms_assemble, ms_precond, ms_cg;
    int64_t moved_values;
    double delta2, plane2;
    float lo[NNTC_MAX_CHANNELS], hi[NNTC_MAX_CHANNELS];
    int32_t quantised, sweeps;
    int64_t moved;
    double ms_sweeps;
} nntc_level1_report;

/* Level0Report (src/model.h:440-447); one per plane. */
typedef struct nntc_level0_report
{
    int64_t moved[4], moved_total;
    double ms;
    int32_t joint;
    int64_t states;
} nntc_level0_report;

/* Objective (src/model.h:467-478), with its per-plane vectors as fixed arrays. */
typedef struct nntc_objective
{
    double e, centre_mse, sampled_mse, ms;
    int32_t nplanes;
    double e_plane[NNTC_MAX_PLANES];
    double centre_plane[NNTC_MAX_PLANES];
} nntc_objective;

/* BcRefineReport (src/model.h:512-519); one per pass. */
typedef struct nntc_bc_refine_report
{
    int32_t pass;
    int64_t blocks, improved;
    double delta_e, ms;
} nntc_bc_refine_report;

/* Range policies of the grid fits (src/model.h:334-345): "minmax" or "pct". */
#define NNTC_RANGE_MINMAX 0
#define NNTC_RANGE_PCT 1

/* Execution forms. Not in the reference's seam: the reference has one form.
 *   EXACT  replays the reference CPU backend's reduction orders (fixed 4096-element chunks folded in chunk order, the
 *          256-site decoder tiles, the covariance grid) so that every result is the reference's bits; conjugate
 *          gradients in fp64 to the reference's 1e-10 relative residual (src/solve_level1.cu:1257).
 *   PLAIN  one kernel per operation with wavefront-tree reductions, the assembled fp32 stencil, fp32 conjugate gradients
 *          to a 1e-6 relative residual (at most 200 iterations), and the decoder Gram in fp64 on the vector ALUs
 *          (the reference's 512 site blocks x 256-site tiles, src/solve_decoder.cu:87-161).
 *   FAST   the chosen form: PLAIN with the decoder Gram as bf16 three-part split products on the matrix cores (fp32
 *          accumulation within a wavefront's sites, fp64 beyond), the decoder's normal equations solved on the device
 *          (fp64 Cholesky, one refinement step), the level-0 conjugate gradients on the matrix-free decode adjoint for
 *          planes of 4096^2 and above, each probe group of conjugate-gradient iterations replayed from a captured HIP
 *          graph, fp32 arithmetic in the BC refinement, and the sweeps' launch form at three or four channels taken
 *          from the launch-shape tuner. */
#define NNTC_FORM_EXACT 0
#define NNTC_FORM_FAST 1
#define NNTC_FORM_PLAIN 2

typedef struct nntc_hip_ctx nntc_hip_ctx;

/* The reason the last failing call on this thread gave; empty when none has failed. */
const char* nntc_hip_last_error(void);

/* Replaces be::device_select (src/backend.h:86). Returns 1 when device `index` is usable, 0 otherwise. */
int nntc_hip_device_select(int index, nntc_device_info* info);

/* Replaces be::device_create (src/backend.h:87): allocates every plane buffer and uploads the source chain.
 * source_chain[p] holds planes[p].w0 * planes[p].h0 * nout floats. Returns NULL on failure. */
nntc_hip_ctx* nntc_hip_device_create(const nntc_model_desc* model, const float* const* source_chain);

/* device_create for textures already in device memory: no counterpart in the reference's seam, whose caller builds the
 * source chain on the host (src/main.cpp:1792-1895) and hands it to be::device_create. Here the packed source plane
 * (every texture's bytes times 1/255, edge-padded to the base plane's size: src/image.h:298-331) and its box mip chain
 * (src/image.h:343-366) are built on the device from rgb8_device[t], texture t's width * height * 3 bytes (row-major
 * RGB) in device memory, with the same float arithmetic, so the chain holds the bits device_create would be handed.
 * model->planes[0] must be (width, height) rounded up to multiples of 4. Returns NULL on failure. */
nntc_hip_ctx* nntc_hip_device_create_rgb8(const nntc_model_desc* model, const uint8_t* const* rgb8_device, int width,
                                          int height);

/* The exact-copy test of the box init's channel order (src/main.cpp:1938-1960) on the device's source chain: copies[s]
 * (nout entries) has bit e set when channel e < s holds the same value as channel s at every texel of every plane.
 * No counterpart in the reference's seam (its caller compares the host chain). */
int nntc_hip_source_copies(nntc_hip_ctx* ctx, uint32_t* copies);

/* One plane of the source chain as the device holds it (w0 * h0 * nout floats); n and count as the other downloads.
 * No counterpart in the reference's seam. */
int nntc_hip_source_download(nntc_hip_ctx* ctx, int plane, float* out, size_t n, size_t* count);

/* Replaces be::device_destroy (src/backend.h:88); NULL is ignored. */
void nntc_hip_device_destroy(nntc_hip_ctx* ctx);

/* No counterpart in the reference (which allocates and frees per run): device_destroy gives a drained context's device
 * buffers and pinned host areas to a cache of this process (at most 96 GiB), and the contexts created after it take
 * their memory from there before they ask the runtime; this returns everything cached to the runtime now and says
 * how many bytes that was. An allocation the runtime refuses for lack of memory empties the cache and is tried
 * again. */
size_t nntc_hip_release_cached_memory(void);

/* Replaces be::device_memory (src/backend.h:89): bytes of device memory the context holds. */
size_t nntc_hip_device_memory(const nntc_hip_ctx* ctx);

/* Selects the execution form (NNTC_FORM_EXACT by default). Not in the reference's seam: the reference has one form. */
int nntc_hip_set_form(nntc_hip_ctx* ctx, int form);

/* Replaces be::init_level0 (src/backend.h:91). k0 mirrors are read with nntc_hip_level0_download. */
int nntc_hip_init_level0(nntc_hip_ctx* ctx, nntc_model_state* state, const float luma_weights[3], int luma_channel0);

/* Replaces be::init_level1 (src/backend.h:92). Fits and writes state->lo1/hi1. */
int nntc_hip_init_level1(nntc_hip_ctx* ctx, nntc_model_state* state, int pca, nntc_level1_init_report* rep);

/* Replaces be::init0_residual_channel (src/backend.h:93). plane_weight has nplanes entries. */
int nntc_hip_init0_residual_channel(nntc_hip_ctx* ctx, nntc_model_state* state, int channel,
                                    const double* plane_weight, nntc_init0_report* rep);

/* Replace be::quantise_level1, freeze_level1_grid, quantise_level0, freeze_level0_grid, snap_level0_on_grid
 * (src/backend.h:96-100). policy is NNTC_RANGE_MINMAX or NNTC_RANGE_PCT. */
int nntc_hip_quantise_level1(nntc_hip_ctx* ctx, nntc_model_state* state, int policy);
int nntc_hip_freeze_level1_grid(nntc_hip_ctx* ctx, nntc_model_state* state, int policy);
int nntc_hip_quantise_level0(nntc_hip_ctx* ctx, nntc_model_state* state, int policy);
int nntc_hip_freeze_level0_grid(nntc_hip_ctx* ctx, nntc_model_state* state, int policy);
int nntc_hip_snap_level0_on_grid(nntc_hip_ctx* ctx, nntc_model_state* state);

/* The fast form's level-0 operator (no counterpart in the reference: it has one). -1 (default): the conjugate
 * gradients of block (c') apply the matrix-free operator - the decode linearised at every site and taken back through
 * its adjoint - when the base plane has 4096^2 texels or more, the assembled fp32 stencil below that; 0: always the
 * assembled stencil; 1: always the matrix-free operator. The quantised sweeps and the BC refinement always read the
 * assembled stencil. */
int nntc_hip_set_matrix_free(nntc_hip_ctx* ctx, int mode);

/* The fast form's quantised sweeps at three or four channels (no counterpart in the reference): -1 (default) takes
 * the launch-shape tuner's choice (below), 1 gives each texel one lane, 4 gives each texel four lanes (one channel row
 * each). Both forms take the same decisions from the same sums and write the same bits. */
int nntc_hip_set_sweep_lanes(nntc_hip_ctx* ctx, int lanes);

/* The fast form's conjugate gradients as HIP graphs (no counterpart in the reference, whose iterations are launched
 * one kernel at a time): -1 (default) or 1, each probe period of CG_PROBE = 4 iterations is captured once per set of
 * launch arguments and replayed; 0, every kernel launched directly. The launches and their order are the same either
 * way. The counts say how many periods were captured and how many replayed on this context. */
int nntc_hip_set_graphs(nntc_hip_ctx* ctx, int mode);
int nntc_hip_graph_counts(const nntc_hip_ctx* ctx, long long* captures, long long* launches);

/* The launch-shape tuner's table (no counterpart in the reference). Where a kernel family has launch forms that write
 * the same bits (today: the fast sweeps' one or four lanes a texel), the first call of a shape - latent channels and
 * texel count to a power of two - on a device times each form once on that call's own state, restoring what the
 * trials moved, and every later call of the shape in the process takes the fastest. Entry i (0 up): the family (0 the
 * sweep lanes), the channels, the bucket (texels <= 2^bucket), the form chosen, and each form tried with its time in
 * ms; nonzero past the last entry. The reset empties the table, so the next calls tune again. */
int nntc_hip_tuning_entry(int i, int* family, int* channels, int* bucket, int* choice, int* forms, double* ms,
                          int* nforms);
void nntc_hip_tuning_reset(void);

/* Block (c')'s operator applied once, for checks (no counterpart in the reference): level 0's system is assembled for
 * the state, as solve_level0_cont_all assembles it, and out = (H + lambda I) p on every plane, from the form's
 * assembled stencil (mode 0) or the fast form's matrix-free operator (mode 1). p and out hold every plane's level-0
 * values (the chain's n0 * c0, interleaved like the planes). */
int nntc_hip_level0_apply(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, double ridge, int mode,
                          const double* p, double* out);

/* Replaces be::stencil_monomials (src/backend.h:102): 1 + c + c (c + 1) / 2. */
int nntc_hip_stencil_monomials(int c0);

/* Replaces be::solve_decoder (src/backend.h:104): block (a). per_site has nplanes entries. Writes state's decoder. */
int nntc_hip_solve_decoder(nntc_hip_ctx* ctx, nntc_model_state* state, int k, const double* per_site, double* ms);

/* Replaces be::decoder_ridge_tally (src/backend.h:105): the rungs the block (a) calls made from this host thread
 * ended on (the reference counts per process; per thread here, so encodes driven from other threads do not mix in). */
void nntc_hip_decoder_ridge_tally(long long out[4]);

/* No counterpart in the reference (its one arithmetic has no such case): how many of this host thread's fast-form
 * block (a) calls the split-bf16 Gram could not settle - a feature whose Schur complement is below 1e-4 of its
 * diagonal entry, or every rung refused - and were redone from the fp64 Gram of the plain form. The state that makes
 * them is a level-0 channel holding one nonzero value everywhere (a palette seed snapped to one entry), whose features
 * are collinear with the bias and level 1's. */
long long nntc_hip_decoder_fast_fallbacks(void);

/* Replace be::solve_level1_all, sweep_level1_all, solve_level0_cont_all, sweep_level0_cont_all
 * (src/backend.h:107-111): blocks (b) and (c'). rep has nplanes entries. */
int nntc_hip_solve_level1_all(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, double ridge,
                              nntc_level1_report* rep, double* ms);
int nntc_hip_sweep_level1_all(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, double ridge, int sweeps,
                              nntc_level1_report* rep, double* ms);
int nntc_hip_solve_level0_cont_all(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, double ridge,
                                   nntc_level1_report* rep, double* ms);
int nntc_hip_sweep_level0_cont_all(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, double ridge, int sweeps,
                                   nntc_level1_report* rep, double* ms);

/* Replaces be::solve_level0_all (src/backend.h:127): --l0 palette's block (c), the exact per-texel search over the
 * palette in four colour passes on every plane. rep has nplanes entries (the block's time on the base plane's). */
int nntc_hip_solve_level0_all(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, int sweeps,
                              nntc_level0_report* rep, double* ms);

/* Replaces be::assemble_level0_for_refine (src/backend.h:112). */
int nntc_hip_assemble_level0_for_refine(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, double* ms);

/* The transfers (src/backend.h:114-123, 128-130). `n` is the element count the caller's buffer holds; a download
 * writes min(n, plane size) elements and returns the plane size through *count when count is not NULL. An upload of
 * a size other than the plane's is ignored, as the reference ignores it (src/cpu/cpu_model.cpp:191-197). */
int nntc_hip_level1_download_indices(nntc_hip_ctx* ctx, int plane, uint8_t* k1, size_t n, size_t* count);
int nntc_hip_level1_upload(nntc_hip_ctx* ctx, int plane, const float* v1, size_t n);
int nntc_hip_level1_upload_indices(nntc_hip_ctx* ctx, int plane, const uint8_t* k1, size_t n);
int nntc_hip_level0_upload_indices(nntc_hip_ctx* ctx, int plane, const uint8_t* k0, size_t n);
int nntc_hip_upload_decoder_to_device(nntc_hip_ctx* ctx, const nntc_model_state* state);
int nntc_hip_upload_grids_to_device(nntc_hip_ctx* ctx, const nntc_model_state* state);
int nntc_hip_level1_download(nntc_hip_ctx* ctx, int plane, float* v1, size_t n, size_t* count);
/* The check of src/main.cpp:2696-2711 (verify_level1_on_grid, src/export.cpp:70-84): of the *values level-1 values of
 * every plane as the device holds them, *off are not the grid value, under the state's level-1 grid, of the index
 * beside them. Replaces the download of both and the host loop. */
int nntc_hip_level1_off_grid(nntc_hip_ctx* ctx, const nntc_model_state* state, size_t* values, size_t* off);
int nntc_hip_level1_delta(nntc_hip_ctx* ctx, int plane, double* delta, size_t n, size_t* count);
int nntc_hip_level1_poke(nntc_hip_ctx* ctx, int plane, size_t index, float value);
int nntc_hip_level1_range(nntc_hip_ctx* ctx, int plane, float* lo, float* hi);
int nntc_hip_level0_download(nntc_hip_ctx* ctx, int plane, uint8_t* k0, size_t n, size_t* count);
int nntc_hip_level0_upload(nntc_hip_ctx* ctx, int plane, const float* v0, size_t n);
int nntc_hip_level0_download_values(nntc_hip_ctx* ctx, int plane, float* v0, size_t n, size_t* count);

/* Replaces be::decode_plane (src/backend.h:132): out = saturate(W phi + b) rounded to bytes, w0*h0*nout bytes. */
int nntc_hip_decode_plane(nntc_hip_ctx* ctx, int plane, uint8_t* rgb8, size_t n);

/* Replaces the host loop image_psnr_texture (src/image.h:387-407) that the reference runs over decode_plane's bytes:
 * the base plane is decoded on the device and each texture's squared byte error against the source over the top-left
 * width x height is summed there. sse has textures (nout / 3) entries; the sums are integers, the reference's double
 * sums' exact values. */
int nntc_hip_texture_sse(nntc_hip_ctx* ctx, int width, int height, unsigned long long* sse);

/* No counterpart in the reference: nntc_hip_texture_sse for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). It allocates nothing and does not wait: the decode and the
 * error sums run in one pass with no decoded plane stored, the per-texture counters live in a fixed 48 KiB device array
 * of the library (one per GPU, held until the process ends), and the sums are written to sse_device, `textures`
 * unsigned 64-bit values in device memory, by work enqueued on `stream` (NULL: the context's stream). The inputs are
 * the context's latents and decoder as its stream leaves them. Limits: at most 6 textures (the counters' capacity) and
 * an extent within the base plane; beyond them it fails with a message. Calls that overlap on one GPU race on the
 * shared counters: a caller issues them one after another (on one stream, or with its own ordering). */
int nntc_hip_texture_sse_noalloc(nntc_hip_ctx* ctx, int width, int height, unsigned long long* sse_device,
                                 void* stream);

/* ---- The decode and its adjoint on one plane (no counterpart in the reference's seam; its arms never differentiate the
 * decode). The sites of a w0 x h0 level-0 plane are numbered s = (py w0 + px)(1 + k) + j for sub-texel site j of pixel
 * (px, py) (src/sample.cuh:87-149); at each the level-0 sample s0 (the pixel's own texel at j = 0, else the bilinear
 * blend of its clamped taps) and the level-1 sample s1 of the w1 x h1 plane give phi = [s1][s0][s0 s1] and
 * out = W phi + b (src/sample.cuh:74-85, src/objective.cu:53-75 before saturation). Every array is device memory,
 * double when fp64 is nonzero, else float: v0 (w0 h0 c0), v1 (w1 h1 c1), W (nout x nin, nin = c1 + c0 + c0 c1), b
 * (nout), out and g (sites x nout). The work runs on `stream` (NULL: the null stream) and the call returns when it is
 * done. */
typedef struct nntc_sites_desc
{
    int32_t w0, h0, w1, h1;
    int32_t c0, c1, nout, k;
} nntc_sites_desc;

int nntc_hip_decode_sites(const nntc_sites_desc* d, int fp64, const void* v0, const void* v1, const void* w,
                          const void* b, void* out, void* stream);

/* The adjoint of nntc_hip_decode_sites at (v0, v1, W, b) for the cotangent g: dv0, dv1, dW and db are the gradients of
 * sum_s,r g_sr out_sr (the transposed bilinear sampling into both latent planes, and the decoder's gradient), each
 * written whole, every sum in a fixed order. */
int nntc_hip_decode_adjoint(const nntc_sites_desc* d, int fp64, const void* v0, const void* v1, const void* w,
                            const void* g, void* dv0, void* dv1, void* dw, void* db, void* stream);

/* Replaces be::objective_eval (src/backend.h:133). per_site has nplanes entries. */
int nntc_hip_objective_eval(nntc_hip_ctx* ctx, int k, const double* per_site, nntc_objective* r);

/* Replaces be::objective_check_host (src/backend.h:136; src/objective.cu:317-380): the --check twin of
 * objective_eval, E by brute force on the host in fp64 from the device's current planes and the state's decoder - every
 * site sampled by the bilinear clamp rule, decoded and weighed - sharing nothing with the device's passes but the
 * model's definition. A checker for small planes; the encode never calls it. */
int nntc_hip_objective_check_host(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, const double* per_site,
                                  nntc_objective* r);

/* Replaces be::solve_level1_dense_host (src/backend.h:124; src/solve_level1.cu:1502-1688): the --check twin of block
 * (b)'s solve on plane `plane`. Level 1's Gauss-Newton system at c_prev (the plane's w1 * h1 * c1 level-1 values) is
 * scattered into a dense matrix on the host in fp64, site by site, and (H + lambda I) delta = -grad is solved by
 * Cholesky into delta (n_delta >= w1 * h1 * c1). *solved is 0, and delta untouched, when the plane has more than
 * max_unknowns unknowns or the factorisation meets a pivot that is not positive (lambda 0 on a singular H), as the
 * reference returns false. */
int nntc_hip_solve_level1_dense_host(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, int plane, double lambda,
                                     const float* c_prev, size_t n_prev, double* delta, size_t n_delta,
                                     int max_unknowns, int* solved);

/* Replace be::objective_total_ms and be::objective_passes (src/backend.h:134-135), counted per host thread. */
double nntc_hip_objective_total_ms(void);
long long nntc_hip_objective_passes(void);

/* Replace be::bc_pack_prepare, bc_refine, bc_blocks_from_device (src/backend.h:138-141). The refined blocks are read
 * with nntc_hip_bc_blocks_download; rep has `passes` entries.
 *
 * The reference reads the host's index mirror Model::k0 in two places: the seed pack packs it
 * (src/export.cpp:294-331) and the packing error is measured against it (src/export.cpp:283-292). Model::k0 is the
 * level-0 index plane as the encode last downloaded it, so here it is an argument: `k0_mirror[plane]` points at the
 * caller's copy of each plane's indices, or k0_mirror is NULL, which means "the device's level-0 indices as they stand"
 * for a seeding bc_pack_prepare (the reference downloads Model::k0 right before each) and "the mirror the previous
 * bc_* call used" otherwise. */
int nntc_hip_bc_pack_prepare(nntc_hip_ctx* ctx, const nntc_model_state* state, int k, int seed,
                             const uint8_t* const* k0_mirror, double* pack_psnr, size_t* pack_texels, double* ms);
int nntc_hip_bc_refine(nntc_hip_ctx* ctx, int passes, const uint8_t* const* k0_mirror, nntc_bc_refine_report* rep,
                       double* pack_psnr, size_t* pack_texels, double* ms);
int nntc_hip_bc_blocks_from_device(nntc_hip_ctx* ctx, const uint8_t* const* k0_mirror, double* pack_psnr,
                                   size_t* pack_texels);

/* The BC4/BC5 blocks of one plane's level-0 file `file` (0, or 1 when c0 > 2), as the .dds stores them: the model's
 * bc0_blocks[plane][file] after a bc_* call (src/model.h:108). */
int nntc_hip_bc_blocks_download(nntc_hip_ctx* ctx, int plane, int file, uint8_t* blocks, size_t n, size_t* count);

/* Replace be::bc_state_save / bc_state_restore (src/backend.h:142-145): one plane's endpoints (blocks*c0*2 bytes) and
 * selectors (blocks*c0*16 bytes). */
int nntc_hip_bc_state_save(nntc_hip_ctx* ctx, int plane, uint8_t* ep, size_t n_ep, uint8_t* sel, size_t n_sel);
int nntc_hip_bc_state_restore(nntc_hip_ctx* ctx, int plane, const uint8_t* ep, size_t n_ep, const uint8_t* sel,
                              size_t n_sel);

/* Replace the outer repack loop's host round trip (src/main.cpp:2561-2588: every plane's level-0 values and indices,
 * level-1 values and indices and BC endpoints and selectors downloaded before a pass and uploaded again when it is
 * rejected, with Model::k0 kept beside them) by a copy on the device: state_snapshot keeps all of it, the index mirror
 * included, and state_restore puts it back. One snapshot is held; restoring without one is an error. */
int nntc_hip_state_snapshot(nntc_hip_ctx* ctx);
int nntc_hip_state_restore(nntc_hip_ctx* ctx);

/* Per-kernel timing (not in the reference's seam; the benchmark's instrument). While enabled, every launch of the
 * arm's named kernels is bracketed by HIP events on the arm's stream and its algorithmic bytes and flops are added
 * up. nntc_hip_kernel_stats reads one kernel's totals (index-based enumeration: name_out receives the name); it
 * returns 0, or nonzero past the last kernel. nntc_hip_profile_only limits the timing to the kernel of that name
 * (NULL: every kernel), so that a timed run carries one kernel's events only. */
void nntc_hip_profile_enable(int on);
void nntc_hip_profile_only(const char* kernel);
void nntc_hip_profile_reset(void);
int nntc_hip_kernel_stats(int index, char* name_out, size_t name_len, long long* launches, double* ms, double* bytes,
                          double* flops);

#ifdef __cplusplus
}
#endif

#endif /* NNTC_HIP_H */
