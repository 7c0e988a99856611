// arm.h: the gfx950 arm's device model and the launchers each kernel file provides.
//
// The device model is the MI355X counterpart of the reference's DeviceModel / CpuModel (src/device.cuh:193-375,
// src/cpu/cpu_model.h:76-174), laid out for batched launches: every per-plane buffer of the chain is ONE allocation
// with the planes concatenated (the base first), and every kernel covers every plane of the chain in one launch,
// mapping its global index to (plane, local index) through the plane table. The deep mip planes are a few hundred
// texels each, so one launch per plane (the reference's stream-per-plane fan-out) would be launch-bound.

#pragma once

#include "nntc_common.h"
#include "nntc_hip.h"

#include <string>
#include <mutex>
#include <vector>

namespace nntc
{

#define HIP_OK(expr)                                                                                                   \
    do                                                                                                                 \
    {                                                                                                                  \
        hipError_t e_ = (expr);                                                                                        \
        if (e_ != hipSuccess)                                                                                          \
            throw ::nntc::HipFailure(#expr, e_, __FILE__, __LINE__);                                                   \
    } while (0)

struct HipFailure
{
    std::string what;
    HipFailure(const char* expr, hipError_t e, const char* file, int line);
};

struct Refused
{
    std::string what;
};

// A typed device allocation that frees itself.
template <class T>
struct DevBuf
{
    T* p = nullptr;
    size_t n = 0;
    size_t cap = 0;        // the bytes the block holds (at least n * sizeof(T); more when it came from the cache)
    int device = -1;
    void alloc(size_t count);
    void release();
    ~DevBuf() { release(); }
    size_t bytes() const { return n * sizeof(T); }
};

// Memory a destroyed context gives back, kept for the contexts this process creates after it (device buffers by device
// and size, pinned host areas by size). A context gives its memory back only once its stream has drained, so a block
// handed out again is never still in use; cache_take returns null when nothing cached fits.
void* cache_take(size_t bytes, int device, bool host, size_t& got);
bool cache_put(void* p, size_t bytes, int device, bool host);
size_t cache_flush();
void* host_alloc(size_t bytes, size_t& got);   // pinned host memory, from the cache when it can
void host_release(void* p, size_t bytes);
// Set while a drained context releases its memory: releases then go to the cache instead of the runtime.
extern thread_local bool g_recycling;

// Scalar row per plane (the reference's PlaneScalar slots, src/cpu/cpu_constants.h:82-101).
enum Scalar
{
    SC_DIAG_SUM = 0,
    SC_DIAG_MIN,
    SC_DIAG_MAX,
    SC_LAMBDA,
    SC_RHS2,
    SC_RZ,
    SC_RZ_NEW,
    SC_RZ_PREV,   // the fast form at level 1: r.z before the update, which the fused p = z + beta p divides by
    SC_PAP,
    SC_RR,
    SC_RESID,
    SC_ALPHA,
    SC_BETA,
    SC_MOVED,
    SC_DELTA2,
    SC_PLANE2,
    SC_LO,
    SC_HI = SC_LO + MAX_CH,
    SC_ACTIVE = SC_HI + MAX_CH,   // 1 while the plane's conjugate gradients still iterate
    SC_COUNT = SC_ACTIVE + 1
};

// One latent's solve workspace: level 1 (block (b)) or level 0 under --l0 bc8 (block (c')).
struct Latent
{
    bool level0 = false;
    int c = 0;              // channels
    int bits = 8;
    DevBuf<double> stencil; // texels * 5 * c * c
    DevBuf<float> stencil32;   // the fast form's fp32 copy, written by its gathers, read by its sweeps and mat-vec
    DevBuf<double> grad;    // texels * c
    DevBuf<float> precond;  // texels * c * c
    DevBuf<double> x, r, z, p, ap;
    DevBuf<float> prev;
    DevBuf<double> mono;    // nmono * c * c
    int nmono = 0;
};

// Per-kernel timing for the benchmark: when on, every timed launch is bracketed by a pair of events on the arm's
// stream, and its algorithmic bytes and flops (computed from the shapes, not measured) are added up by kernel name.
struct KernelStat
{
    long long launches = 0;
    double bytes = 0.0, flops = 0.0, ms = 0.0;
    std::vector<hipEvent_t> pending;   // begin/end pairs not yet read
};

struct Profiler
{
    bool on = false;
    std::string only;                   // when set, the one kernel timed
    std::vector<std::pair<std::string, KernelStat>> stats;
    std::vector<hipEvent_t> pool;
    std::recursive_mutex lock;          // contexts driven from several host threads time into one table
    KernelStat& get(const char* name);
    hipEvent_t event();
    void settle();                      // read every pending pair (synchronizes on the last)
    void reset();
    ~Profiler();
};

Profiler& profiler();
bool profiled(const char* name);   // would a KernelTimer of this name record events now

// RAII bracket around one launch: `KernelTimer t(c.stream, "k_matvec", bytes, flops); hipLaunchKernelGGL(...);`
struct KernelTimer
{
    KernelStat* st = nullptr;
    hipStream_t stream;
    hipEvent_t b = nullptr;
    KernelTimer(hipStream_t s, const char* name, double bytes, double flops);
    ~KernelTimer();
};

struct Ctx
{
    nntc_model_desc md {};
    int nin = 0, nout = 0, c0 = 0, c1 = 0;
    int form = NNTC_FORM_EXACT;
    int device = -1;                       // the HIP device the context was created on
    int matrix_free = -1;                  // the fast level-0 operator: -1 by size (4096^2 and up), 0 assembled, 1 matrix-free
    int assembled_k = -1;                  // the K of the last assembly
    int sweep_lanes = -1;                  // the fast sweeps at C >= 3: -1 the tuner's choice, 1 or 4 lanes a texel
    int graphs = -1;                       // the fast CG's probe periods as HIP graphs: -1 / 1 on, 0 off
    // The fast CG's instantiated probe periods, by their launch arguments (kept for the context's lifetime: an
    // encode's calls of one latent repeat the same launches on the same buffers).
    struct CgGraph
    {
        std::vector<unsigned char> key;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<CgGraph> cg_graphs;
    long long graph_launches = 0, graph_captures = 0;
    Planes planes {};
    long long n0_total = 0, n1_total = 0, nb_total = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;

    DevBuf<float> src, v0, v1;
    DevBuf<uint8_t> k0, k1;
    DevBuf<float> weights, bias, cw;       // nout * nin, nout, nout
    DevBuf<float> grid;                    // lo0[4] hi0[4] lo1[4] hi1[4] palette[4*16]
    Latent lat1, lat0;
    DevBuf<double> scalars;                // nplanes * SC_COUNT
    DevBuf<double> partial;                // chunk partials of the fixed-order reductions
    DevBuf<double> ls_partial;             // nplanes * entries * SITE_BLOCKS
    DevBuf<double> ls_sum;                 // entries
    DevBuf<double> dec_io;                 // the device decoder solve's incoming decoder, result and status
    DevBuf<double> obj;                    // nplanes * 3
    DevBuf<float> resid;                   // base texels * nout
    DevBuf<double> cov_partial;            // SITE_BLOCKS * entries
    DevBuf<unsigned int> counters;         // per plane, MOVED_SLOTS counters MOVED_STRIDE apart
    DevBuf<unsigned int> pal_counters;     // the palette search's: the same per colour, then the folded counts
    DevBuf<double> cg_partial;             // the fast form's CG: three partials per chunk of 256 texels
    DevBuf<double> gather_tab;             // the fast level-0 gather's interior site lists (4 x 45 entries of 7)
    int gather_tab_k = -1;                 // the K the list was built for
    DevBuf<uint8_t> bc_ep, bc_sel;         // blocks * c0 * 2, blocks * c0 * 16
    DevBuf<uint8_t> bc_blocks;             // blocks * files * 16 (file f of block b at (b * files + f) * 16)
    DevBuf<uint8_t> k0_mirror;             // the host's Model::k0 as the bc calls read it
    struct Snapshot                         // the outer repack loop's saved state (state_snapshot / state_restore)
    {
        DevBuf<float> v0, v1;
        DevBuf<uint8_t> k0, k1, ep, sel, mirror;
        bool held = false;
    } snap;
    DevBuf<double> bc_taken;               // per block: the accepted decrease
    DevBuf<double> bc_se;                  // per block: the packing error
    DevBuf<double> bc_fold;                // the packing error's total, then each refinement pass's figures per plane
    std::vector<hipEvent_t> pass_ev;       // brackets of the refinement passes
    DevBuf<float> work_f;                  // staging for grid fits (values of one channel)
    DevBuf<unsigned int> work_u;           // radix-select histograms
    DevBuf<uint8_t> rgb8;                  // one plane's decode
    DevBuf<float> site_f;                  // per site of the chain: the other latent's sample (assembly pass 1)
    DevBuf<double> site_d;                 // per site: the residual pulled back through the decoder
    DevBuf<float> site_target;             // the fast form's per-site source sample (nout floats), built once per K
    int site_target_k = -1;
    DevBuf<unsigned long long> tex_sse;    // per texture, SSE_SLOTS counters: the base decode's squared byte error

    // Pinned host staging for the small per-call readbacks.
    double* host_scalars = nullptr;
    size_t host_scalars_bytes = 0;
    // Pinned host staging: a copy between device memory and pageable host memory is staged by the runtime through
    // buffers of its own, in pieces, each a copy kernel and a wait. Readbacks use one area (each waits for its copy);
    // uploads take successive pieces of a second area, so a new upload never waits for an earlier one still in flight,
    // and the area is reused from its start once the stream has drained.
    void* host_stage = nullptr;
    size_t host_stage_bytes = 0;
    char* host_up = nullptr;
    size_t host_up_bytes = 0, up_head = 0;
    void* stage(size_t bytes);          // the readback area, at least `bytes` long
    void* upload_slot(size_t bytes);    // a piece of the upload area no queued copy still reads

    size_t bytes = 0;

    // Host-side view of the decoder and grids as last uploaded.
    std::vector<float> h_w, h_b;
    float h_lo0[4] = {}, h_hi0[4] = {}, h_lo1[4] = {}, h_hi1[4] = {};
    float h_grid[16 + 64] = {};                 // the grid buffer as last uploaded
    bool decoder_on_device = false, grid_on_device = false;

    Latent& latent(bool level0) { return level0 ? lat0 : lat1; }
    const PlaneGeom& g(int p) const { return planes.p[p]; }
    float ms_since(hipEvent_t a);
};

// n values read back from the device through the pinned readback area; the pointer holds until the next readback.
template <class T>
const T* read_back(Ctx& c, const T* dev, size_t n)
{
    T* h = static_cast<T*>(c.stage(n * sizeof(T)));
    if (n)
        HIP_OK(hipMemcpyAsync(h, dev, n * sizeof(T), hipMemcpyDeviceToHost, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
    c.up_head = 0;   // the stream has drained: no upload reads the upload area any more
    return h;
}

// n values uploaded through the pinned upload area, asynchronously: the launches that read them follow on the stream.
template <class T>
void upload_staged(Ctx& c, T* dev, const T* host, size_t n)
{
    if (!n)
        return;
    T* h = static_cast<T*>(c.upload_slot(n * sizeof(T)));
    for (size_t i = 0; i < n; i++)
        h[i] = host[i];
    HIP_OK(hipMemcpyAsync(dev, h, n * sizeof(T), hipMemcpyHostToDevice, c.stream));
}

// The stencil layout of one latent's chain (see StencilMap).
inline StencilMap stencil_map(const Ctx& c, bool level0)
{
    StencilMap m {};
    long long acc = 0;
    for (int p = 0; p < c.planes.n; p++)
    {
        const PlaneGeom& g = c.planes.p[p];
        const int w = level0 ? g.w0 : g.w1, h = level0 ? g.h0 : g.h1;
        for (int col = 0; col < 4; col++)
        {
            m.cpos[p][col] = acc;
            acc += colour_count(w, h, col);
        }
    }
    m.sn = acc;
    return m;
}

// The quantised sweeps' moved-index counters: MOVED_SLOTS per plane, MOVED_STRIDE unsigned ints (one 128-byte line)
// apart, summed on the host after the sweeps.
constexpr int MOVED_SLOTS = 64, MOVED_STRIDE = 32;
// The PSNR sums' counters: SSE_SLOTS per texture, SSE_STRIDE 64-bit counters (one 128-byte line) apart.
constexpr int SSE_SLOTS = 64, SSE_STRIDE = 16;

// The fast-form objective's tile of pixels (one workgroup each) and the number of tiles of a plane.
constexpr int OBJ_TW = 32, OBJ_TH = 8;
inline long long objective_tiles(const PlaneGeom& g)
{
    return (long long)((g.w0 + OBJ_TW - 1) / OBJ_TW) * ((g.h0 + OBJ_TH - 1) / OBJ_TH);
}

// ---- launchers (kernel files)
// objective.hip
void objective_pass(Ctx& c, int k, double* sums_out /* nplanes * 3 */);
void decode_pass(Ctx& c, int plane, uint8_t* out);
void residual_pass(Ctx& c, int plane, float* resid);
void texture_sse_pass(Ctx& c, int width, int height, unsigned long long* sse_out);
// The allocation-free variant's scratch address on `device` (the current device), taken once per device.
void noalloc_setup(int device);
// texture_sse_pass without its plane buffer, with counters in the file-scope scratch; enqueued on `stream`, the sums
// written to sse_device (device memory, one per texture).
void texture_sse_noalloc_pass(Ctx& c, int width, int height, unsigned long long* sse_device, hipStream_t stream);
// decoder.hip
// Block (a)'s normal equations into c.ls_sum (and, when to_host, into acc).
void ls_accumulate(Ctx& c, int k, const double* omega, std::vector<double>& acc, bool to_host = true);
// The fast form's solve of the normal equations in c.ls_sum on the device (fp64 Cholesky, one refinement step, the
// ridge ladder and acceptance rule): the rung taken (3: refused), whether x holds a decoder, and x (nout x (nin + 1)).
void decoder_solve_device(Ctx& c, const nntc_model_state* st, double inv_norm, bool norm_ok, const double ladder[3],
                          double accept_rise, int& taken, bool& wrote, std::vector<float>& x);
// latent.hip
void build_block_matrices(Ctx& c, bool level0);
void assemble(Ctx& c, bool level0, int k, double ridge);
void cg_solve(Ctx& c, bool level0, std::vector<int>& iterations);
// (H + lambda I) p for level 0's last assembled system, every plane: mode 0 the form's assembled stencil, mode 1 the
// fast form's matrix-free operator. p and out hold n0_total * c0 values.
void level0_apply(Ctx& c, int mode, const double* p_host, double* out_host);
void correction_and_probe(Ctx& c, bool level0);
void quant_sweeps(Ctx& c, bool level0, int sweeps);
void copy_prev_zero_delta(Ctx& c, bool level0);
// The fast form's per-site source samples (site s's nout values at [s nout]), built when K changes or an encode starts.
const float* site_targets(Ctx& c, int k);
// The source chain built on the device from the textures' RGB8 bytes in device memory (init.hip).
void build_source_chain(Ctx& c, const uint8_t* const* rgb8_device, int width, int height);
// copies[s] bit e (e < s): source channel e is an exact copy of channel s over the whole chain.
void channel_copies(Ctx& c, unsigned int* copies);
// init.cu counterparts (init.hip)
void init_level0_kernel(Ctx& c, const float* luma, int luma_channel0);
void init_level1_box(Ctx& c);
void block_means(Ctx& c, int plane, float* out /* device, n1 * nout */);   // the pca init's per-texel source means
void quantise_values(Ctx& c, bool level0, const float* in_device, int plane);
void fit_range(Ctx& c, bool level0, int policy, float* lo, float* hi);
void cov_accumulate(Ctx& c, const float* data, long long texels, double weight, bool first);
void cov_read(Ctx& c, std::vector<double>& raw);
double proj_peak_chain(Ctx& c, const float* resid, const double* e, const double* mu);
void seed_channel(Ctx& c, int plane, const float* resid, const double* e, const double* mu, double inv, int channel);
// The palette mode's seed divisor: the 99th percentile of |projection| over the chain, from a 4096-bin histogram on
// [0, peak], taken at its bin's upper edge (the peak itself when every texel falls in the first bin).
double proj_percentile_chain(Ctx& c, const float* resid, const double* e, const double* mu, double peak);
// palette.hip: --l0 palette
void init_level0_palette(Ctx& c, const float* luma, int luma_channel0);
void solve_level0_all(Ctx& c, int k, int sweeps, nntc_level0_report* rep);
unsigned long long level1_off_grid(Ctx& c, const float* lo, const float* hi);
// twins.hip: the --check twins (host, fp64, from the device's current planes)
void objective_check_host(Ctx& c, const nntc_model_state* st, int k, double* acc /* nplanes * 3 */);
bool solve_level1_dense_host(Ctx& c, const nntc_model_state* st, int k, int plane, double lambda,
                             const float* c_prev, std::vector<double>& delta, int max_unknowns);
// refine.hip
void bc_decode(Ctx& c);
void bc_seed(Ctx& c);
void bc_refine_pass(Ctx& c, double* fold /* device, nplanes * 2 */);
void bc_pack_blocks(Ctx& c, double* se_total /* device, one value */);

// ---- shared launch helpers
inline unsigned int blocks_for(long long n, int threads)
{
    return (unsigned int)((n + threads - 1) / threads);
}

inline int ilog2_ceil(long long n)
{
    int b = 0;
    while ((1ll << b) < n)
        b++;
    return b;
}

// ---- the launch-shape tuner (arm.hip): for a kernel family with several launch forms that write the same bits, the
// form each shape takes on each device, picked by timing the forms on the first call of that shape.
enum TuneFamily
{
    TUNE_SWEEP_LANES = 0,   // the fast quantised sweeps at C >= 3: one lane or four lanes a texel
    TUNE_FAMILIES
};
struct TuneKey
{
    int device, family, channels, bucket;   // bucket: the texel count's power of two, rounded up
};
int tuned_choice(const TuneKey& k);          // 0 when the shape has not been tuned
void record_choice(const TuneKey& k, int choice, const int* forms, const double* ms, int nforms);

}   // namespace nntc
