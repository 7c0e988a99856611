// arm.hip: the gfx950 arm's lifecycle, transfers and host-side numerics, and the C ABI of include/nntc_hip.h.
//
// Host-side numerics kept here, each the reference's arithmetic on the reference's operands:
//   - block (a)'s solve: mean diagonal, the incoming decoder as x, Q's normaliser, qform and its rounding bound, the
//     ridge ladder {1e-9, 1e-12, 0} x mean diag with Gauss-Jordan on [A + ridge I | rhs^T], the float round trip before
//     judging, the acceptance bar, the tally and the constant-fit fallback (src/solve_decoder.cu:184-458,
//     src/cpu/cpu_solve_decoder.cpp:199-361);
//   - the residual seed's covariance, cyclic Jacobi eigensolver, sign rule and divisor (src/init.cu:813-1400,
//     src/cpu/cpu_init.cpp:372-419, 729-910);
//   - the grid fits' fallback to [-1, 1] on an empty range (src/cpu/cpu_init.cpp:428-507).
// The host compiler runs with -ffp-contract=off, as the reference's host build does.

#include "arm.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

namespace nntc
{

HipFailure::HipFailure(const char* expr, hipError_t e, const char* file, int line)
{
    char buf[512];
    std::snprintf(buf, sizeof(buf), "HIP error %d (%s) at %s:%d in %s", (int)e, hipGetErrorString(e), file, line, expr);
    what = buf;
}

template <class T>
void DevBuf<T>::alloc(size_t count)
{
    release();
    if (count == 0)
        count = 1;
    const size_t want = count * sizeof(T);
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    size_t got = 0;
    void* ptr = cache_take(want, dev, false, got);
    if (!ptr)
    {
        hipError_t e = hipMalloc(&ptr, want);
        if (e == hipErrorOutOfMemory && cache_flush() > 0)
        {
            (void)hipGetLastError();
            e = hipMalloc(&ptr, want);
        }
        HIP_OK(e);
        got = want;
    }
    p = static_cast<T*>(ptr);
    n = count;
    cap = got;
    device = dev;
}

template <class T>
void DevBuf<T>::release()
{
    if (p && !(g_recycling && cache_put(p, cap, device, false)))
        (void)hipFree(p);
    p = nullptr;
    n = 0;
    cap = 0;
}

namespace
{

struct Cached
{
    void* p;
    size_t bytes;
    int device;
    bool host;
};

std::mutex g_cache_lock;
std::vector<Cached> g_cache;
size_t g_cache_bytes = 0;
// At most this much is held for later contexts; beyond it a released block goes back to the runtime.
constexpr size_t CACHE_CAP = (size_t)96 << 30;

}   // namespace

thread_local bool g_recycling = false;

void* cache_take(size_t bytes, int device, bool host, size_t& got)
{
    std::lock_guard<std::mutex> g(g_cache_lock);
    // The smallest cached block that holds the request and wastes at most a quarter of itself.
    int best = -1;
    for (int i = 0; i < (int)g_cache.size(); i++)
    {
        const Cached& c = g_cache[(size_t)i];
        if (c.host == host && c.device == device && c.bytes >= bytes && c.bytes - bytes <= c.bytes / 4 &&
            (best < 0 || c.bytes < g_cache[(size_t)best].bytes))
            best = i;
    }
    if (best < 0)
        return nullptr;
    const Cached c = g_cache[(size_t)best];
    g_cache.erase(g_cache.begin() + best);
    g_cache_bytes -= c.bytes;
    got = c.bytes;
    return c.p;
}

bool cache_put(void* p, size_t bytes, int device, bool host)
{
    std::lock_guard<std::mutex> g(g_cache_lock);
    if (g_cache_bytes + bytes > CACHE_CAP)
        return false;
    g_cache.push_back(Cached { p, bytes, device, host });
    g_cache_bytes += bytes;
    return true;
}

size_t cache_flush()
{
    std::vector<Cached> all;
    {
        std::lock_guard<std::mutex> g(g_cache_lock);
        all.swap(g_cache);
        g_cache_bytes = 0;
    }
    size_t freed = 0;
    for (const Cached& c : all)
    {
        if (c.host)
            (void)hipHostFree(c.p);
        else
            (void)hipFree(c.p);
        freed += c.bytes;
    }
    return freed;
}

void* host_alloc(size_t bytes, size_t& got)
{
    void* p = cache_take(bytes, -1, true, got);
    if (p)
        return p;
    HIP_OK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    got = bytes;
    return p;
}

void host_release(void* p, size_t bytes)
{
    if (p && !(g_recycling && cache_put(p, bytes, -1, true)))
        (void)hipHostFree(p);
}

template struct DevBuf<float>;
template struct DevBuf<double>;
template struct DevBuf<uint8_t>;
template struct DevBuf<unsigned int>;
template struct DevBuf<unsigned long long>;

Profiler& profiler()
{
    static Profiler p;
    return p;
}

KernelStat& Profiler::get(const char* name)
{
    for (auto& kv : stats)
        if (kv.first == name)
            return kv.second;
    stats.emplace_back(name, KernelStat());
    return stats.back().second;
}

hipEvent_t Profiler::event()
{
    if (!pool.empty())
    {
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    HIP_OK(hipEventCreate(&e));
    return e;
}

void Profiler::settle()
{
    std::lock_guard<std::recursive_mutex> g(lock);
    for (auto& kv : stats)
    {
        KernelStat& st = kv.second;
        for (size_t i = 0; i + 1 < st.pending.size(); i += 2)
        {
            HIP_OK(hipEventSynchronize(st.pending[i + 1]));
            float ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&ms, st.pending[i], st.pending[i + 1]));
            st.ms += ms;
            pool.push_back(st.pending[i]);
            pool.push_back(st.pending[i + 1]);
        }
        st.pending.clear();
    }
}

void Profiler::reset()
{
    std::lock_guard<std::recursive_mutex> g(lock);
    settle();
    stats.clear();
}

Profiler::~Profiler()
{
    for (hipEvent_t e : pool)
        (void)hipEventDestroy(e);
}

bool profiled(const char* name)
{
    Profiler& p = profiler();
    if (!p.on)
        return false;
    std::lock_guard<std::recursive_mutex> g(p.lock);
    return p.only.empty() || p.only == name;
}

KernelTimer::KernelTimer(hipStream_t s, const char* name, double bytes, double flops) : stream(s)
{
    Profiler& p = profiler();
    if (!p.on)
        return;
    std::lock_guard<std::recursive_mutex> g(p.lock);
    if (!p.only.empty() && p.only != name)
        return;
    st = &p.get(name);
    st->launches++;
    st->bytes += bytes;
    st->flops += flops;
    b = p.event();
    HIP_OK(hipEventRecord(b, stream));
}

KernelTimer::~KernelTimer()
{
    if (!st)
        return;
    Profiler& p = profiler();
    std::lock_guard<std::recursive_mutex> g(p.lock);
    hipEvent_t e = p.event();
    if (hipEventRecord(e, stream) == hipSuccess)
    {
        st->pending.push_back(b);
        st->pending.push_back(e);
    }
}

void* Ctx::stage(size_t bytes)
{
    if (bytes > host_stage_bytes)
    {
        host_release(host_stage, host_stage_bytes);
        host_stage = nullptr;
        host_stage_bytes = 0;
        const size_t want = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
        host_stage = host_alloc(want, host_stage_bytes);
    }
    return host_stage;
}

void* Ctx::upload_slot(size_t bytes)
{
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (up_head + need > host_up_bytes)
    {
        if (up_head)
        {
            HIP_OK(hipStreamSynchronize(stream));
            up_head = 0;
        }
        if (need > host_up_bytes)
        {
            host_release(host_up, host_up_bytes);
            host_up = nullptr;
            host_up_bytes = 0;
            const size_t want = need < ((size_t)1 << 20) ? ((size_t)1 << 20) : need;
            host_up = static_cast<char*>(host_alloc(want, host_up_bytes));
        }
    }
    void* slot = host_up + up_head;
    up_head += need;
    return slot;
}

float Ctx::ms_since(hipEvent_t a)
{
    HIP_OK(hipEventRecord(ev1, stream));
    HIP_OK(hipEventSynchronize(ev1));
    float ms = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms, a, ev1));
    return ms;
}

namespace
{

thread_local std::string g_error;
// The reference's process-global counters (src/backend.h:105, 134-135), kept per host thread: the driver reads them
// before and after an encode, and each encode runs on one thread, so encodes on other threads do not mix in.
thread_local double g_objective_ms = 0.0;
thread_local long long g_objective_calls = 0;
thread_local long long g_ridge_tally[4] = { 0, 0, 0, 0 };
// The fast form's decoder solves that every rung refused on the split-bf16 Gram and were redone from the fp64 one.
thread_local long long g_fast_fallbacks = 0;

const double RIDGE_LADDER[] = { 1e-9, 1e-12, 0.0 };
const double LS_ACCEPT_RISE = 1e-12;
const size_t RIDGE_REFUSED = 3;

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Run a body, turning HIP failures and refusals into a status and a message.
template <class F>
int guarded(F&& f)
{
    try
    {
        f();
        return 0;
    }
    catch (const HipFailure& e)
    {
        g_error = e.what;
    }
    catch (const Refused& e)
    {
        g_error = e.what;
    }
    catch (const std::bad_alloc&)
    {
        g_error = "out of host memory";
    }
    return 1;
}

}   // namespace

// guarded() for the other source files of the arm.
int guarded_call(const std::function<void()>& f)
{
    return guarded(f);
}

namespace
{

inline int tri_index(int i, int j, int mm)
{
    return i * mm - i * (i - 1) / 2 + (j - i);
}

bool gauss_jordan(std::vector<double>& g, int m, int nrhs)
{
    const int cols = m + nrhs;
    for (int k = 0; k < m; k++)
    {
        int pivot = k;
        double best = std::fabs(g[(size_t)k * cols + k]);
        for (int i = k + 1; i < m; i++)
        {
            const double a = std::fabs(g[(size_t)i * cols + k]);
            if (a > best)
            {
                best = a;
                pivot = i;
            }
        }
        if (best == 0.0)
            return false;
        if (pivot != k)
            for (int j = 0; j < cols; j++)
                std::swap(g[(size_t)k * cols + j], g[(size_t)pivot * cols + j]);
        const double inv = 1.0 / g[(size_t)k * cols + k];
        for (int j = 0; j < cols; j++)
            g[(size_t)k * cols + j] *= inv;
        for (int i = 0; i < m; i++)
        {
            if (i == k)
                continue;
            const double f = g[(size_t)i * cols + k];
            if (f == 0.0)
                continue;
            for (int j = 0; j < cols; j++)
                g[(size_t)i * cols + j] -= f * g[(size_t)k * cols + j];
        }
    }
    return true;
}

// The cyclic Jacobi eigendecomposition of a small symmetric matrix; vec's COLUMNS are the eigenvectors.
void jacobi_eigen(std::vector<double>& a, int n, std::vector<double>& vec, std::vector<double>& val)
{
    vec.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++)
        vec[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; sweep++)
    {
        double off = 0.0;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++)
                off += a[(size_t)p * n + q] * a[(size_t)p * n + q];
        if (off <= 1e-30)
            break;
        for (int p = 0; p < n; p++)
            for (int q = p + 1; q < n; q++)
            {
                const double apq = a[(size_t)p * n + q];
                if (std::fabs(apq) <= 1e-300)
                    continue;
                const double theta = (a[(size_t)q * n + q] - a[(size_t)p * n + p]) / (2.0 * apq);
                const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++)
                {
                    const double akp = a[(size_t)k * n + p], akq = a[(size_t)k * n + q];
                    a[(size_t)k * n + p] = c * akp - s * akq;
                    a[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; k++)
                {
                    const double apk = a[(size_t)p * n + k], aqk = a[(size_t)q * n + k];
                    a[(size_t)p * n + k] = c * apk - s * aqk;
                    a[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; k++)
                {
                    const double vkp = vec[(size_t)k * n + p], vkq = vec[(size_t)k * n + q];
                    vec[(size_t)k * n + p] = c * vkp - s * vkq;
                    vec[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    val.assign((size_t)n, 0.0);
    for (int i = 0; i < n; i++)
        val[(size_t)i] = a[(size_t)i * n + i];
}

void upload_decoder(Ctx& c, const nntc_model_state* st)
{
    const size_t nw = (size_t)c.nout * c.nin;
    if (st && st->has_decoder)
    {
        // Nothing but this upload writes the device's decoder, so when the values are those of the last upload (bit for
        // bit) the device already holds them.
        if (c.decoder_on_device && c.h_w.size() == nw && c.h_b.size() == (size_t)c.nout &&
            std::memcmp(c.h_w.data(), st->w, nw * sizeof(float)) == 0 &&
            std::memcmp(c.h_b.data(), st->b, (size_t)c.nout * sizeof(float)) == 0)
            return;
        c.h_w.assign(st->w, st->w + nw);
        c.h_b.assign(st->b, st->b + c.nout);
        upload_staged(c, c.weights.p, st->w, nw);
        upload_staged(c, c.bias.p, st->b, (size_t)c.nout);
        c.decoder_on_device = true;
    }
}

// Both grids written to the device, zero-padded (src/cpu/cpu_model.cpp:385-413).
void upload_grids(Ctx& c, const nntc_model_state* st)
{
    float g[16 + 64] = {};
    if (st)
    {
        if (st->has_grid0)
            for (int q = 0; q < c.c0; q++)
            {
                g[q] = st->lo0[q];
                g[4 + q] = st->hi0[q];
            }
        if (st->has_grid1)
            for (int q = 0; q < c.c1; q++)
            {
                g[8 + q] = st->lo1[q];
                g[12 + q] = st->hi1[q];
            }
        if (st->has_palette0)
            for (int q = 0; q < 4; q++)
                for (int k = 0; k < 16; k++)
                    g[16 + q * 16 + k] = st->palette0[q][k];
    }
    std::memcpy(c.h_lo0, g, sizeof(c.h_lo0));
    std::memcpy(c.h_hi0, g + 4, sizeof(c.h_hi0));
    std::memcpy(c.h_lo1, g + 8, sizeof(c.h_lo1));
    std::memcpy(c.h_hi1, g + 12, sizeof(c.h_hi1));
    // As with the decoder, only this upload writes the device's grids.
    if (c.grid_on_device && std::memcmp(c.h_grid, g, sizeof(g)) == 0)
        return;
    std::memcpy(c.h_grid, g, sizeof(g));
    upload_staged(c, c.grid.p, g, sizeof(g) / sizeof(g[0]));
    c.grid_on_device = true;
}

// A plane-sized download (or, with n smaller, its head).
template <class T>
void download(Ctx& c, const T* src, size_t count, T* out, size_t n, size_t* total)
{
    if (total)
        *total = count;
    const size_t m = count < n ? count : n;
    if (out && m)
    {
        HIP_OK(hipMemcpyAsync(out, src, m * sizeof(T), hipMemcpyDeviceToHost, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
    }
}

template <class T>
void upload(Ctx& c, T* dst, size_t count, const T* in, size_t n)
{
    if (n != count)
        return;   // the reference ignores an upload of the wrong size (src/cpu/cpu_model.cpp:191-197)
    HIP_OK(hipMemcpyAsync(dst, in, n * sizeof(T), hipMemcpyHostToDevice, c.stream));
    HIP_OK(hipStreamSynchronize(c.stream));
}

void check_plane(const Ctx& c, int plane)
{
    if (plane < 0 || plane >= c.planes.n)
        throw Refused { "plane index out of range" };
}

// Fit one latent's grid over every plane, write it into the state and the device, snap every plane onto it.
void fit_and_snap(Ctx& c, nntc_model_state* st, bool level0, int policy)
{
    float lo[4], hi[4];
    fit_range(c, level0, policy, lo, hi);
    const int ch = level0 ? c.c0 : c.c1;
    for (int q = 0; q < ch; q++)
        if (!(hi[q] > lo[q]))
        {
            lo[q] = -1.0f;
            hi[q] = 1.0f;
        }
    for (int q = 0; q < ch; q++)
    {
        (level0 ? st->lo0 : st->lo1)[q] = lo[q];
        (level0 ? st->hi0 : st->hi1)[q] = hi[q];
    }
    (level0 ? st->has_grid0 : st->has_grid1) = 1;
    upload_grids(c, st);
    quantise_values(c, level0, level0 ? c.v0.p : c.v1.p, -1);
    HIP_OK(hipStreamSynchronize(c.stream));
}

// The scalar rows copied into host_scalars; they are there once the stream has next been synchronized.
void fetch_scalars(Ctx& c)
{
    HIP_OK(hipMemcpyAsync(c.host_scalars, c.scalars.p, sizeof(double) * SC_COUNT * c.planes.n,
                          hipMemcpyDeviceToHost, c.stream));
}

// The report fields of the scalar rows fetched last.
void parse_scalars(Ctx& c, int channels, nntc_level1_report* rep)
{
    for (int p = 0; p < c.planes.n; p++)
    {
        const double* s = c.host_scalars + (size_t)p * SC_COUNT;
        rep[p].lambda = s[SC_LAMBDA];
        rep[p].min_diag = s[SC_DIAG_MIN];
        rep[p].max_diag = s[SC_DIAG_MAX];
        rep[p].residual = s[SC_RESID];
        rep[p].moved_values = (long long)s[SC_MOVED];
        rep[p].delta2 = s[SC_DELTA2];
        rep[p].plane2 = s[SC_PLANE2];
        for (int q = 0; q < channels; q++)
        {
            rep[p].lo[q] = (float)s[SC_LO + q];
            rep[p].hi[q] = (float)s[SC_HI + q];
        }
    }
}

// Blocks (b) and (c'), continuous: assemble, solve, correct, probe (src/cpu/cpu_solve_level1.cpp:969-986).
double solve_block(Ctx& c, const nntc_model_state* st, int k, double ridge, bool level0, nntc_level1_report* rep)
{
    const double t0 = now_ms();
    upload_decoder(c, st);
    for (int p = 0; p < c.planes.n; p++)
        rep[p] = nntc_level1_report {};
    build_block_matrices(c, level0);
    HIP_OK(hipEventRecord(c.ev0, c.stream));
    assemble(c, level0, k, ridge);
    HIP_OK(hipEventRecord(c.ev1, c.stream));
    std::vector<int> iterations;
    cg_solve(c, level0, iterations);
    correction_and_probe(c, level0);
    HIP_OK(hipEventRecord(c.ev2, c.stream));
    fetch_scalars(c);
    HIP_OK(hipStreamSynchronize(c.stream));
    c.up_head = 0;
    const int ch = level0 ? c.c0 : c.c1;
    parse_scalars(c, ch, rep);
    float ms_asm = 0.0f, ms_cg = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms_asm, c.ev0, c.ev1));
    HIP_OK(hipEventElapsedTime(&ms_cg, c.ev1, c.ev2));
    // Every plane of the chain runs in the same launches, so each plane's report carries the block's kernel time.
    for (int p = 0; p < c.planes.n; p++)
    {
        rep[p].iterations = iterations[(size_t)p];
        rep[p].quantised = 0;
        rep[p].ms_assemble = ms_asm;
        rep[p].ms_cg = ms_cg;
    }
    return now_ms() - t0;
}

// The same block on a frozen grid: assemble, then `sweeps` four-colour Gauss-Seidel passes
// (src/cpu/cpu_solve_level1.cpp:991-1054).
double sweep_block(Ctx& c, const nntc_model_state* st, int k, double ridge, int sweeps, bool level0,
                   nntc_level1_report* rep)
{
    const double t0 = now_ms();
    upload_decoder(c, st);
    for (int p = 0; p < c.planes.n; p++)
        rep[p] = nntc_level1_report {};
    build_block_matrices(c, level0);
    HIP_OK(hipEventRecord(c.ev0, c.stream));
    assemble(c, level0, k, ridge);
    HIP_OK(hipEventRecord(c.ev1, c.stream));
    copy_prev_zero_delta(c, level0);
    quant_sweeps(c, level0, sweeps);
    HIP_OK(hipEventRecord(c.ev2, c.stream));
    fetch_scalars(c);
    const unsigned int* slots = read_back(c, (const unsigned int*)c.counters.p,
                                          (size_t)c.planes.n * MOVED_SLOTS * MOVED_STRIDE);
    const int ch = level0 ? c.c0 : c.c1;
    parse_scalars(c, ch, rep);
    float ms_asm = 0.0f, ms_sw = 0.0f;
    HIP_OK(hipEventElapsedTime(&ms_asm, c.ev0, c.ev1));
    HIP_OK(hipEventElapsedTime(&ms_sw, c.ev1, c.ev2));
    std::vector<unsigned long long> moved((size_t)c.planes.n, 0);
    for (int p = 0; p < c.planes.n; p++)
        for (int s = 0; s < MOVED_SLOTS; s++)
            moved[(size_t)p] += slots[((size_t)p * MOVED_SLOTS + s) * MOVED_STRIDE];
    for (int p = 0; p < c.planes.n; p++)
    {
        rep[p].quantised = 1;
        rep[p].sweeps = sweeps;
        rep[p].ms_assemble = ms_asm;
        rep[p].ms_sweeps = ms_sw;
        // The scalar row persists between blocks, so the continuous-only fields are cleared.
        rep[p].iterations = 0;
        rep[p].residual = 0.0;
        rep[p].moved_values = 0;
        rep[p].delta2 = 0.0;
        rep[p].plane2 = 0.0;
        rep[p].moved = (long long)moved[(size_t)p];
    }
    return now_ms() - t0;
}

// Block (c')'s assembly alone at zero ridge, prev = v0, delta = 0 (src/cpu/cpu_solve_level1.cpp:1082-1097).
double assemble_for_refine(Ctx& c, const nntc_model_state* st, int k)
{
    const double t0 = now_ms();
    upload_decoder(c, st);
    build_block_matrices(c, true);
    assemble(c, true, k, 0.0);
    copy_prev_zero_delta(c, true);
    HIP_OK(hipStreamSynchronize(c.stream));
    return now_ms() - t0;
}

void set_k0_mirror(Ctx& c, const uint8_t* const* k0, bool snapshot)
{
    if (k0)
    {
        for (int p = 0; p < c.planes.n; p++)
            HIP_OK(hipMemcpyAsync(c.k0_mirror.p + c.g(p).off0 * c.c0, k0[p], (size_t)c.g(p).n0 * c.c0,
                                  hipMemcpyHostToDevice, c.stream));
    }
    else if (snapshot)
        HIP_OK(hipMemcpyAsync(c.k0_mirror.p, c.k0.p, (size_t)c.n0_total * c.c0, hipMemcpyDeviceToDevice, c.stream));
}

// The packing error and texel count of the pack on the device, from se, its total squared error.
void pack_figures(const Ctx& c, double se, double* pack_psnr, size_t* pack_texels)
{
    size_t n = 0;
    for (int p = 0; p < c.planes.n; p++)
        n += (size_t)c.g(p).n0 * (size_t)c.c0;
    if (pack_texels)
        *pack_texels = n;
    if (pack_psnr)
        *pack_psnr = n > 0 && se > 0.0 ? 10.0 * std::log10(255.0 * 255.0 / (se / (double)n)) : 100.0;
}

// --init pca (src/init.cu:861-987): the base plane's block means of every source channel, their covariance (the
// device's sums in the reference's k_cov_partial order, as the residual seed takes them), the host's Jacobi
// eigensolver, the directions in order of their eigenvalues; every plane's block means projected onto the first C1
// directions in fp64, each channel divided by the largest magnitude it reaches over the chain (clamped to [-1, 1]),
// and written as the plane's level 1 for the grid fit that follows. The projection runs on the host, over the means
// read back, as the reference's does, so the exact form keeps its bits.
void init_level1_pca(Ctx& c, nntc_level1_init_report* rep)
{
    const int nout = c.nout, c1 = c.c1;
    const int entries = nout + nout * nout;
    std::vector<double> raw((size_t)entries, 0.0);
    block_means(c, 0, c.resid.p);
    cov_accumulate(c, c.resid.p, c.g(0).n1, 1.0, true);
    cov_read(c, raw);
    const double count = (double)c.g(0).n1;
    std::vector<double> mu((size_t)nout, 0.0), cov((size_t)nout * nout, 0.0);
    for (int r = 0; r < nout; r++)
        mu[(size_t)r] = count > 0.0 ? raw[(size_t)r] / count : 0.0;
    for (int r = 0; r < nout; r++)
        for (int s = 0; s < nout; s++)
            cov[(size_t)r * nout + s] =
                (count > 0.0 ? raw[(size_t)(nout + r * nout + s)] / count : 0.0) - mu[(size_t)r] * mu[(size_t)s];
    double total = 0.0;
    for (int r = 0; r < nout; r++)
        total += cov[(size_t)r * nout + r];
    std::vector<double> work = cov, vec, val;
    jacobi_eigen(work, nout, vec, val);
    std::vector<int> order((size_t)nout);
    for (int i = 0; i < nout; i++)
        order[(size_t)i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return val[(size_t)a] > val[(size_t)b]; });
    const int components = c1 < nout ? c1 : nout;
    if (rep)
    {
        rep->total_variance = total;
        rep->components = components;
        for (int j = 0; j < components && j < MAX_CH; j++)
            rep->eigenvalue[j] = val[(size_t)order[(size_t)j]];
    }
    std::vector<std::vector<float>> cont((size_t)c.planes.n);
    std::vector<double> peak((size_t)c1, 0.0);
    std::vector<float> means;
    for (int p = 0; p < c.planes.n; p++)
    {
        const long long texels = c.g(p).n1;
        block_means(c, p, c.resid.p);
        means.resize((size_t)texels * nout);
        const float* h = read_back(c, (const float*)c.resid.p, means.size());
        std::copy(h, h + means.size(), means.begin());
        cont[(size_t)p].assign((size_t)texels * c1, 0.0f);
        for (long long t = 0; t < texels; t++)
            for (int j = 0; j < components; j++)
            {
                const int col = order[(size_t)j];
                double acc = 0.0;
                for (int r = 0; r < nout; r++)
                    acc += vec[(size_t)r * nout + col] * ((double)means[(size_t)t * nout + r] - mu[(size_t)r]);
                cont[(size_t)p][(size_t)t * c1 + j] = (float)acc;
                peak[(size_t)j] = std::max(peak[(size_t)j], std::fabs(acc));
            }
    }
    for (int j = 0; j < c1; j++)
    {
        if (rep && j < MAX_CH)
            rep->peak[j] = peak[(size_t)j];
        if (!(peak[(size_t)j] > 0.0))
            continue;
        const double inv = 1.0 / peak[(size_t)j];
        for (std::vector<float>& v : cont)
            for (size_t t = (size_t)j; t < v.size(); t += (size_t)c1)
            {
                const double x = (double)v[t] * inv;
                v[t] = (float)(x < -1.0 ? -1.0 : (x > 1.0 ? 1.0 : x));
            }
    }
    for (int p = 0; p < c.planes.n; p++)
        upload_staged(c, c.v1.p + c.g(p).off1 * c1, cont[(size_t)p].data(), cont[(size_t)p].size());
    HIP_OK(hipStreamSynchronize(c.stream));
}

void blocks_from_device(Ctx& c, double* pack_psnr, size_t* pack_texels)
{
    bc_pack_blocks(c, c.bc_fold.p);
    pack_figures(c, *read_back(c, (const double*)c.bc_fold.p, 1), pack_psnr, pack_texels);
}

}   // namespace

}   // namespace nntc

// ---- the launch-shape tuner's table: one entry per (device, family, channels, bucket), for the process's lifetime.
namespace nntc
{
namespace
{
struct TuneEntry
{
    TuneKey key;
    int choice;
    int forms[4];
    double ms[4];
    int nforms;
};
std::mutex g_tune_lock;
std::vector<TuneEntry> g_tune;
}   // namespace

int tuned_choice(const TuneKey& k)
{
    std::lock_guard<std::mutex> g(g_tune_lock);
    for (const TuneEntry& e : g_tune)
        if (e.key.device == k.device && e.key.family == k.family && e.key.channels == k.channels &&
            e.key.bucket == k.bucket)
            return e.choice;
    return 0;
}

void record_choice(const TuneKey& k, int choice, const int* forms, const double* ms, int nforms)
{
    std::lock_guard<std::mutex> g(g_tune_lock);
    // Two contexts may tune one shape at once (a library of materials): the first to finish keeps the entry.
    for (const TuneEntry& t : g_tune)
        if (t.key.device == k.device && t.key.family == k.family && t.key.channels == k.channels &&
            t.key.bucket == k.bucket)
            return;
    TuneEntry e {};
    e.key = k;
    e.choice = choice;
    e.nforms = nforms < 4 ? nforms : 4;
    for (int i = 0; i < e.nforms; i++)
    {
        e.forms[i] = forms[i];
        e.ms[i] = ms[i];
    }
    g_tune.push_back(e);
}
}   // namespace nntc

using namespace nntc;

struct nntc_hip_ctx
{
    Ctx c;
};

extern "C"
{

const char* nntc_hip_last_error(void)
{
    return g_error.c_str();
}

int nntc_hip_device_select(int index, nntc_device_info* info)
{
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || index < 0 || index >= count)
    {
        if (info)
            std::snprintf(info->error, sizeof(info->error), "%s",
                          e != hipSuccess ? hipGetErrorString(e) : "no such device");
        return 0;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, index) != hipSuccess || hipSetDevice(index) != hipSuccess)
        return 0;
    if (info)
    {
        std::snprintf(info->name, sizeof(info->name), "%s (%s)", prop.name, prop.gcnArchName);
        info->major = prop.major;
        info->minor = prop.minor;
        info->error[0] = 0;
    }
    return 1;
}

namespace
{

// Everything device_create does but the source chain's upload: the plane table, the stream, every buffer, the
// zeroed state and the starting grids.
void create_model(Ctx& c, const nntc_model_desc* md)
{
    {
        if (!md || md->nplanes < 1 || md->nplanes > MAX_PLANES || md->c0 < 1 || md->c0 > MAX_CH || md->c1 < 1 ||
            md->c1 > MAX_CH || md->nout < 3 || md->nout > MAX_NOUT)
            throw Refused { "model descriptor out of range" };
        if (!md->l0_bc8)
            for (int q = 0; q < md->c0; q++)
                if (md->bits0[q] < 1 || md->bits0[q] > 4)
                    throw Refused { "--l0 palette takes 1 to 4 bits per level-0 channel" };
        c.md = *md;
        c.nout = md->nout;
        c.c0 = md->c0;
        c.c1 = md->c1;
        c.nin = c.c1 + c.c0 + c.c0 * c.c1;
        c.planes.n = md->nplanes;
        long long o0 = 0, o1 = 0, ob = 0;
        for (int p = 0; p < md->nplanes; p++)
        {
            PlaneGeom& g = c.planes.p[p];
            g.w0 = md->planes[p].w0;
            g.h0 = md->planes[p].h0;
            g.w1 = md->planes[p].w1;
            g.h1 = md->planes[p].h1;
            if (g.w0 < 1 || g.h0 < 1 || g.w1 < 1 || g.h1 < 1)
                throw Refused { "empty plane" };
            g.n0 = (long long)g.w0 * g.h0;
            g.n1 = (long long)g.w1 * g.h1;
            // The kernels index within a plane in 32 bits: every plane's values (at most MAX_NOUT per texel), and its
            // sites (at most five per pixel), stay below 2^31.
            if (g.n0 * std::max(MAX_NOUT, 5) >= (1ll << 31) || g.n1 * MAX_CH >= (1ll << 31))
                throw Refused { "plane too large: a plane's values must number fewer than 2^31" };
            g.nb = (long long)((g.w0 + 3) / 4) * ((g.h0 + 3) / 4);
            g.off0 = o0;
            g.off1 = o1;
            g.offb = ob;
            o0 += g.n0;
            o1 += g.n1;
            ob += g.nb;
        }
        c.n0_total = o0;
        c.n1_total = o1;
        c.nb_total = ob;
        HIP_OK(hipGetDevice(&c.device));
        noalloc_setup(c.device);
        HIP_OK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        HIP_OK(hipEventCreate(&c.ev0));
        HIP_OK(hipEventCreate(&c.ev1));
        HIP_OK(hipEventCreate(&c.ev2));

        const int nout = c.nout, c0 = c.c0, c1 = c.c1;
        c.src.alloc((size_t)o0 * nout);
        c.v0.alloc((size_t)o0 * c0);
        c.v1.alloc((size_t)o1 * c1);
        c.k0.alloc((size_t)o0 * c0);
        c.k1.alloc((size_t)o1 * c1);
        c.weights.alloc((size_t)nout * c.nin);
        c.bias.alloc((size_t)nout);
        c.cw.alloc((size_t)nout);
        c.grid.alloc(16 + 64);
        for (int lvl = 0; lvl < 2; lvl++)
        {
            const bool level0 = lvl == 0;
            Latent& L = c.latent(level0);
            const long long t = level0 ? o0 : o1;
            L.level0 = level0;
            L.c = level0 ? c0 : c1;
            L.bits = level0 ? 8 : md->bits1;
            L.nmono = stencil_monomials(level0 ? c1 : c0);
            L.stencil.alloc((size_t)t * STENCIL_BLOCKS * L.c * L.c);
            L.stencil32.alloc((size_t)t * STENCIL_BLOCKS * L.c * L.c);
            L.grad.alloc((size_t)t * L.c);
            L.precond.alloc((size_t)t * L.c * L.c);
            L.x.alloc((size_t)t * L.c);
            L.r.alloc((size_t)t * L.c);
            L.z.alloc((size_t)t * L.c);
            L.p.alloc((size_t)t * L.c);
            L.ap.alloc((size_t)t * L.c);
            L.prev.alloc((size_t)t * L.c);
            L.mono.alloc((size_t)L.nmono * L.c * L.c);
        }
        c.scalars.alloc((size_t)SC_COUNT * md->nplanes);
        long long chunks_obj = 0;
        for (int p = 0; p < md->nplanes; p++)
            chunks_obj += std::max((c.planes.p[p].n0 * 5 + CHUNK - 1) / CHUNK + 1, objective_tiles(c.planes.p[p]));
        c.partial.alloc((size_t)std::max<long long>(chunks_obj * 3, (o0 / CHUNK + md->nplanes + 1) * 8));
        const int mm = c.nin + 1;
        const int entries = mm * (mm + 1) / 2 + nout * mm;
        long long tiles = 0;   // the fast form's Gram keeps one partial per pixel tile
        for (int p = 0; p < md->nplanes; p++)
            tiles += objective_tiles(c.planes.p[p]);
        c.ls_partial.alloc((size_t)entries * (size_t)std::max<long long>(md->nplanes * SITE_BLOCKS, tiles));
        c.ls_sum.alloc((size_t)entries + MAX_PLANES);
        c.dec_io.alloc(1024);
        c.obj.alloc((size_t)md->nplanes * 3);
        c.resid.alloc((size_t)o0 * nout);
        c.cov_partial.alloc((size_t)SITE_BLOCKS * (nout + nout * nout) + (nout + nout * nout));
        c.counters.alloc((size_t)md->nplanes * MOVED_SLOTS * MOVED_STRIDE);
        c.cg_partial.alloc(3 * ((size_t)std::max(o0, o1) / 256 + MAX_PLANES + 1));
        c.gather_tab.alloc(4 * 9 * 5 * 7 + 4);
        c.bc_ep.alloc((size_t)ob * c0 * 2);
        c.bc_sel.alloc((size_t)ob * c0 * 16);
        c.bc_blocks.alloc((size_t)ob * 32);
        c.k0_mirror.alloc((size_t)o0 * c0);
        c.bc_taken.alloc((size_t)ob);
        c.bc_se.alloc((size_t)ob);
        c.bc_fold.alloc(1 + 2 * (size_t)md->nplanes * 16);
        c.work_f.alloc(256 * 2 * MAX_CH);
        c.work_u.alloc(MAX_CH * 2 * 256 + 64);
        c.rgb8.alloc((size_t)c.planes.p[0].n0 * nout);
        c.tex_sse.alloc((size_t)(NNTC_MAX_NOUT / 3) * SSE_SLOTS * SSE_STRIDE);
        // The assembly's site buffer: every site of the chain at the largest K the encoder takes (4).
        c.site_f.alloc((size_t)o0 * 5 * std::max({ c0, c1, c0 * (c0 + 1) / 2 }));
        c.site_d.alloc((size_t)o0 * 5 * std::max(c0, c1));
        {
            size_t got = 0;
            c.host_scalars = static_cast<double*>(host_alloc(sizeof(double) * SC_COUNT * MAX_PLANES, got));
            c.host_scalars_bytes = got;
        }

        c.bytes = c.src.bytes() + c.v0.bytes() + c.v1.bytes() + c.k0.bytes() + c.k1.bytes() + c.weights.bytes() +
                  c.bias.bytes() + c.cw.bytes() + c.grid.bytes() + c.scalars.bytes() + c.partial.bytes() +
                  c.ls_partial.bytes() + c.ls_sum.bytes() + c.obj.bytes() + c.resid.bytes() + c.cov_partial.bytes() +
                  c.counters.bytes() + c.bc_ep.bytes() + c.bc_sel.bytes() + c.bc_blocks.bytes() +
                  c.k0_mirror.bytes() + c.bc_taken.bytes() + c.bc_se.bytes() + c.work_f.bytes() + c.work_u.bytes() +
                  c.rgb8.bytes() + c.site_f.bytes() + c.site_d.bytes();
        for (int lvl = 0; lvl < 2; lvl++)
        {
            Latent& L = c.latent(lvl == 0);
            c.bytes += L.stencil.bytes() + L.grad.bytes() + L.precond.bytes() + L.x.bytes() + L.r.bytes() +
                       L.z.bytes() + L.p.bytes() + L.ap.bytes() + L.prev.bytes() + L.mono.bytes();
        }

        // The decoder starts at zero, the grids at the state device_create is handed (src/cpu/cpu_model.cpp:208-237).
        HIP_OK(hipMemcpyAsync(c.cw.p, md->cw, nout * sizeof(float), hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipMemsetAsync(c.weights.p, 0, c.weights.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.bias.p, 0, c.bias.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.v0.p, 0, c.v0.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.v1.p, 0, c.v1.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.k0.p, 0, c.k0.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.k1.p, 0, c.k1.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.scalars.p, 0, c.scalars.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.bc_ep.p, 0, c.bc_ep.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.bc_sel.p, 0, c.bc_sel.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.lat0.x.p, 0, c.lat0.x.bytes(), c.stream));
        HIP_OK(hipMemsetAsync(c.lat1.x.p, 0, c.lat1.x.bytes(), c.stream));
        c.h_w.assign((size_t)nout * c.nin, 0.0f);
        c.h_b.assign((size_t)nout, 0.0f);
        nntc_model_state s {};
        s.has_grid0 = 1;
        for (int q = 0; q < c0; q++)
        {
            s.lo0[q] = -1.0f;
            s.hi0[q] = 1.0f;
        }
        upload_grids(c, &s);
    }
}

}   // namespace

nntc_hip_ctx* nntc_hip_device_create(const nntc_model_desc* md, const float* const* source_chain)
{
    nntc_hip_ctx* h = new nntc_hip_ctx();
    Ctx& c = h->c;
    const int st = guarded([&] {
        create_model(c, md);
        // The source chain, the one upload before any block.
        for (int p = 0; p < md->nplanes; p++)
            HIP_OK(hipMemcpyAsync(c.src.p + c.g(p).off0 * c.nout, source_chain[p],
                                  (size_t)c.g(p).n0 * c.nout * sizeof(float), hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
    });
    if (st)
    {
        nntc_hip_device_destroy(h);
        return nullptr;
    }
    return h;
}

nntc_hip_ctx* nntc_hip_device_create_rgb8(const nntc_model_desc* md, const uint8_t* const* rgb8_device, int width,
                                          int height)
{
    nntc_hip_ctx* h = new nntc_hip_ctx();
    Ctx& c = h->c;
    const int st = guarded([&] {
        if (!md || !rgb8_device || width < 1 || height < 1 || md->nout < 3 || md->nout > MAX_NOUT)
            throw Refused { "device_create_rgb8: no textures" };
        const int wp = (width + 3) / 4 * 4, hp = (height + 3) / 4 * 4;
        if (md->nplanes < 1 || md->planes[0].w0 != wp || md->planes[0].h0 != hp)
            throw Refused { "device_create_rgb8: the base plane is not the textures padded to whole 4x4 blocks" };
        for (int t = 0; t < md->nout / 3; t++)
            if (!rgb8_device[t])
                throw Refused { "device_create_rgb8: a texture pointer is null" };
        create_model(c, md);
        build_source_chain(c, rgb8_device, width, height);
        HIP_OK(hipStreamSynchronize(c.stream));
    });
    if (st)
    {
        nntc_hip_device_destroy(h);
        return nullptr;
    }
    return h;
}

int nntc_hip_source_copies(nntc_hip_ctx* h, uint32_t* copies)
{
    return guarded([&] {
        unsigned int m[NNTC_MAX_NOUT] = {};
        channel_copies(h->c, m);
        for (int s = 0; s < h->c.nout; s++)
            copies[s] = m[s];
    });
}

int nntc_hip_source_download(nntc_hip_ctx* h, int plane, float* out, size_t n, size_t* count)
{
    return guarded([&] {
        Ctx& c = h->c;
        if (plane < 0 || plane >= c.planes.n)
            throw Refused { "source_download: no such plane" };
        const size_t total = (size_t)c.g(plane).n0 * c.nout;
        if (count)
            *count = total;
        const size_t m = std::min(n, total);
        if (m)
            HIP_OK(hipMemcpy(out, c.src.p + c.g(plane).off0 * c.nout, m * sizeof(float), hipMemcpyDeviceToHost));
    });
}

void nntc_hip_device_destroy(nntc_hip_ctx* h)
{
    if (!h)
        return;
    Ctx& c = h->c;
    // Once the stream has drained nothing of the context is in use: its memory goes to the process's cache, for the
    // next context to take (a stream that failed to drain returns its memory to the runtime instead).
    const bool drained = !c.stream || hipStreamSynchronize(c.stream) == hipSuccess;
    g_recycling = drained;
    host_release(c.host_scalars, c.host_scalars_bytes);
    host_release(c.host_stage, c.host_stage_bytes);
    host_release(c.host_up, c.host_up_bytes);
    if (c.ev0)
        (void)hipEventDestroy(c.ev0);
    if (c.ev1)
        (void)hipEventDestroy(c.ev1);
    if (c.ev2)
        (void)hipEventDestroy(c.ev2);
    for (hipEvent_t e : c.pass_ev)
        (void)hipEventDestroy(e);
    for (Ctx::CgGraph& g : c.cg_graphs)
        if (g.exec)
            (void)hipGraphExecDestroy(g.exec);
    hipStream_t s = c.stream;
    delete h;
    g_recycling = false;
    if (s)
        (void)hipStreamDestroy(s);
}

size_t nntc_hip_release_cached_memory(void)
{
    return cache_flush();
}

size_t nntc_hip_device_memory(const nntc_hip_ctx* h)
{
    return h ? h->c.bytes : 0;
}

int nntc_hip_set_form(nntc_hip_ctx* h, int form)
{
    return guarded([&] {
        if (form != NNTC_FORM_EXACT && form != NNTC_FORM_FAST && form != NNTC_FORM_PLAIN)
            throw Refused { "unknown form" };
        h->c.form = form;
    });
}

int nntc_hip_set_matrix_free(nntc_hip_ctx* h, int mode)
{
    return guarded([&] {
        if (mode < -1 || mode > 1)
            throw Refused { "set_matrix_free takes -1, 0 or 1" };
        h->c.matrix_free = mode;
    });
}

int nntc_hip_tuning_entry(int i, int* family, int* channels, int* bucket, int* choice, int* forms, double* ms,
                          int* nforms)
{
    std::lock_guard<std::mutex> g(nntc::g_tune_lock);
    if (i < 0 || (size_t)i >= nntc::g_tune.size())
        return 1;
    const nntc::TuneEntry& e = nntc::g_tune[(size_t)i];
    *family = e.key.family;
    *channels = e.key.channels;
    *bucket = e.key.bucket;
    *choice = e.choice;
    *nforms = e.nforms;
    for (int f = 0; f < e.nforms; f++)
    {
        forms[f] = e.forms[f];
        ms[f] = e.ms[f];
    }
    return 0;
}

void nntc_hip_tuning_reset(void)
{
    std::lock_guard<std::mutex> g(nntc::g_tune_lock);
    nntc::g_tune.clear();
}

int nntc_hip_set_sweep_lanes(nntc_hip_ctx* h, int lanes)
{
    return guarded([&] {
        if (lanes != -1 && lanes != 1 && lanes != 4)
            throw Refused { "set_sweep_lanes takes -1, 1 or 4" };
        h->c.sweep_lanes = lanes;
    });
}

int nntc_hip_set_graphs(nntc_hip_ctx* h, int mode)
{
    return guarded([&] {
        if (mode < -1 || mode > 1)
            throw Refused { "set_graphs takes -1, 0 or 1" };
        h->c.graphs = mode;
    });
}

int nntc_hip_graph_counts(const nntc_hip_ctx* h, long long* captures, long long* launches)
{
    return guarded([&] {
        *captures = h->c.graph_captures;
        *launches = h->c.graph_launches;
    });
}

int nntc_hip_level0_apply(nntc_hip_ctx* h, const nntc_model_state* st, int k, double ridge, int mode, const double* p,
                          double* out)
{
    return guarded([&] {
        Ctx& c = h->c;
        if (mode != 0 && mode != 1)
            throw Refused { "level0_apply: mode 0 (assembled) or 1 (matrix-free)" };
        upload_decoder(c, st);
        build_block_matrices(c, true);
        assemble(c, true, k, ridge);
        level0_apply(c, mode, p, out);
    });
}

int nntc_hip_stencil_monomials(int c0)
{
    return stencil_monomials(c0);
}

int nntc_hip_init_level0(nntc_hip_ctx* h, nntc_model_state* st, const float luma[3], int luma_channel0)
{
    return guarded([&] {
        Ctx& c = h->c;
        // An encode starts here: the tables the fast form derives while it runs (the per-site source samples, the
        // interior site list) are built again inside it, not carried over from an earlier encode.
        c.site_target_k = -1;
        c.gather_tab_k = -1;
        upload_grids(c, st);
        if (c.md.l0_bc8)
            init_level0_kernel(c, luma, luma_channel0);
        else
        {
            if (!st->has_palette0)
                throw Refused { "init_level0: the palette mode needs the state's palette" };
            init_level0_palette(c, luma, luma_channel0);
        }
        HIP_OK(hipStreamSynchronize(c.stream));
    });
}

int nntc_hip_init_level1(nntc_hip_ctx* h, nntc_model_state* st, int pca, nntc_level1_init_report* rep)
{
    return guarded([&] {
        Ctx& c = h->c;
        if (rep)
        {
            *rep = nntc_level1_init_report {};
            rep->pca = pca ? 1 : 0;
        }
        if (pca)
            init_level1_pca(c, rep);
        else
            init_level1_box(c);
        fit_and_snap(c, st, false, NNTC_RANGE_MINMAX);
    });
}

int nntc_hip_init0_residual_channel(nntc_hip_ctx* h, nntc_model_state* st, int channel, const double* plane_weight,
                                    nntc_init0_report* rep)
{
    return guarded([&] {
        Ctx& c = h->c;
        const int nout = c.nout;
        if (channel < 0 || channel >= c.c0)
            throw Refused { "channel out of range" };
        upload_decoder(c, st);
        nntc_init0_report r {};
        r.nout = nout;
        r.texture = -1;
        // (1) the residual of every plane at once: each plane's depends on that plane alone, and the seed below
        // writes a plane only after its residual is read.
        for (int p = 0; p < c.planes.n; p++)
            residual_pass(c, p, c.resid.p + c.g(p).off0 * nout);
        // (2) the covariance over the base plane, or every plane weighted under --init0-scope chain.
        const int entries = nout + nout * nout;
        std::vector<double> raw((size_t)entries, 0.0);
        double count = 0.0;
        const int nplanes = c.md.init0_chain ? c.planes.n : 1;
        bool first = true;
        for (int p = 0; p < nplanes; p++)
        {
            const double w = c.md.init0_chain ? plane_weight[p] : 1.0;
            if (!(w > 0.0))
                continue;
            cov_accumulate(c, c.resid.p + c.g(p).off0 * nout, c.g(p).n0, w, first);
            first = false;
            count += w * (double)c.g(p).n0;
        }
        if (!first)
            cov_read(c, raw);
        std::vector<double> mu((size_t)nout, 0.0), cov((size_t)nout * nout, 0.0);
        for (int rr = 0; rr < nout; rr++)
            mu[(size_t)rr] = count > 0.0 ? raw[(size_t)rr] / count : 0.0;
        for (int rr = 0; rr < nout; rr++)
            for (int s = 0; s < nout; s++)
                cov[(size_t)rr * nout + s] = (count > 0.0 ? raw[(size_t)(nout + rr * nout + s)] / count : 0.0) -
                                             mu[(size_t)rr] * mu[(size_t)s];
        for (int rr = 0; rr < nout; rr++)
            r.total_variance += cov[(size_t)rr * nout + rr];
        // (3) the direction, its largest component made positive: the leading eigenvector of the whole covariance
        // (--init0 residual), or (--init0 texture, src/init.cu:1219-1260) that of the one texture's 3 x 3 block whose
        // leading eigenvalue is the largest, every other output at zero.
        std::vector<double> e((size_t)nout, 0.0);
        if (c.md.init0_texture)
        {
            int best = 0;
            double best_value = -1.0, best_vec[3] = { 0.0, 0.0, 0.0 };
            for (int t = 0; t < c.md.textures; t++)
            {
                std::vector<double> block(9), vec, val;
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        block[(size_t)(a * 3 + b)] = cov[(size_t)(3 * t + a) * nout + (3 * t + b)];
                jacobi_eigen(block, 3, vec, val);
                int top = 0;
                for (int i = 1; i < 3; i++)
                    if (val[(size_t)i] > val[(size_t)top])
                        top = i;
                if (val[(size_t)top] > best_value)
                {
                    best_value = val[(size_t)top];
                    best = t;
                    for (int a = 0; a < 3; a++)
                        best_vec[a] = vec[(size_t)(a * 3 + top)];
                }
            }
            r.texture = best;
            r.eigenvalue = best_value > 0.0 ? best_value : 0.0;
            for (int a = 0; a < 3; a++)
                e[(size_t)(3 * best + a)] = best_vec[a];
        }
        else
        {
            std::vector<double> work = cov, vec, val;
            jacobi_eigen(work, nout, vec, val);
            int top = 0;
            for (int i = 1; i < nout; i++)
                if (val[(size_t)i] > val[(size_t)top])
                    top = i;
            r.eigenvalue = val[(size_t)top];
            for (int rr = 0; rr < nout; rr++)
                e[(size_t)rr] = vec[(size_t)rr * nout + top];
        }
        r.share = r.total_variance > 0.0 ? r.eigenvalue / r.total_variance : 0.0;
        int biggest = 0;
        for (int rr = 1; rr < nout; rr++)
            if (std::fabs(e[(size_t)rr]) > std::fabs(e[(size_t)biggest]))
                biggest = rr;
        if (e[(size_t)biggest] < 0.0)
            for (int rr = 0; rr < nout; rr++)
                e[(size_t)rr] = -e[(size_t)rr];
        for (int rr = 0; rr < nout && rr < MAX_NOUT; rr++)
            r.dir[rr] = e[(size_t)rr];
        // (4a) the divisor: the largest |projection| over the whole chain.
        const double peak = proj_peak_chain(c, c.resid.p, e.data(), mu.data());
        r.peak = peak;
        // Under --l0 palette the divisor is the 99th percentile of |projection| instead (src/init.cu:1327-1377): the
        // palette is fixed on [-1, 1], so the seed must use its range.
        double divisor = peak;
        if (!c.md.l0_bc8 && peak > 0.0)
            divisor = proj_percentile_chain(c, c.resid.p, e.data(), mu.data(), peak);
        r.divisor = divisor;
        // (4b) the channel on every plane: continuous under --l0 bc8, snapped onto the palette otherwise.
        const double inv = divisor > 0.0 ? 1.0 / divisor : 0.0;
        for (int p = 0; p < c.planes.n; p++)
            seed_channel(c, p, c.resid.p + c.g(p).off0 * nout, e.data(), mu.data(), inv, channel);
        HIP_OK(hipStreamSynchronize(c.stream));
        if (rep)
            *rep = r;
    });
}

int nntc_hip_solve_level0_all(nntc_hip_ctx* h, const nntc_model_state* st, int k, int sweeps,
                              nntc_level0_report* rep, double* ms)
{
    return guarded([&] {
        Ctx& c = h->c;
        if (c.md.l0_bc8)
            throw Refused { "solve_level0_all is the palette mode's block (c); --l0 bc8 runs solve_level0_cont_all" };
        if (!st->has_decoder || !st->has_palette0)
            throw Refused { "solve_level0_all: the state needs its decoder and its palette" };
        const double t0 = now_ms();
        upload_decoder(c, st);
        upload_grids(c, st);
        solve_level0_all(c, k, sweeps, rep);
        if (ms)
            *ms = now_ms() - t0;
    });
}

int nntc_hip_quantise_level1(nntc_hip_ctx* h, nntc_model_state* st, int policy)
{
    return guarded([&] { fit_and_snap(h->c, st, false, policy); });
}

int nntc_hip_freeze_level1_grid(nntc_hip_ctx* h, nntc_model_state* st, int policy)
{
    return guarded([&] { fit_and_snap(h->c, st, false, policy); });
}

int nntc_hip_quantise_level0(nntc_hip_ctx* h, nntc_model_state* st, int policy)
{
    return guarded([&] { fit_and_snap(h->c, st, true, policy); });
}

int nntc_hip_freeze_level0_grid(nntc_hip_ctx* h, nntc_model_state* st, int policy)
{
    return guarded([&] { fit_and_snap(h->c, st, true, policy); });
}

int nntc_hip_snap_level0_on_grid(nntc_hip_ctx* h, nntc_model_state* st)
{
    return guarded([&] {
        Ctx& c = h->c;
        upload_grids(c, st);
        quantise_values(c, true, c.v0.p, -1);
        HIP_OK(hipStreamSynchronize(c.stream));
    });
}

int nntc_hip_solve_decoder(nntc_hip_ctx* h, nntc_model_state* st, int k, const double* per_site, double* ms_out)
{
    return guarded([&] {
        Ctx& c = h->c;
        const int nin = c.nin, nout = c.nout, mm = nin + 1;
        const int tri = mm * (mm + 1) / 2;
        const int planes = c.planes.n;
        HIP_OK(hipEventRecord(c.ev0, c.stream));
        std::vector<double> acc;
        const bool on_device = c.form == NNTC_FORM_FAST;
        ls_accumulate(c, k, per_site, acc, !on_device);
        double ms = on_device ? 0.0 : c.ms_since(c.ev0);
        const bool had_previous = st->has_decoder != 0;
        std::vector<double> x_prev((size_t)nout * mm, 0.0);
        if (had_previous)
            for (int cc = 0; cc < nout; cc++)
            {
                for (int i = 0; i < nin; i++)
                    x_prev[(size_t)cc * mm + i] = (double)st->w[(size_t)cc * nin + i];
                x_prev[(size_t)cc * mm + nin] = (double)st->b[cc];
            }
        double sum_cw = 0.0;
        for (int cc = 0; cc < nout; cc++)
            sum_cw += (double)c.md.cw[cc];
        double den = 0.0;
        for (int i = 0; i < planes; i++)
            den += per_site[i] * (double)c.g(i).w0 * (double)c.g(i).h0 * (double)(1 + k);
        const bool norm_ok = sum_cw * den > 0.0;
        const double inv_norm = norm_ok ? 1.0 / (sum_cw * den) : 0.0;
        if (on_device)
        {
            // The fast form: the ladder solved on the device from the fp64 Gram there; one readback of the result.
            int rung = 3;
            bool wrote = false;
            std::vector<float> x;
            decoder_solve_device(c, st, inv_norm, norm_ok, RIDGE_LADDER, LS_ACCEPT_RISE, rung, wrote, x);
            if (rung < 3)
            {
                g_ridge_tally[rung]++;
                st->has_decoder = 1;
                for (int cc = 0; cc < nout; cc++)
                {
                    for (int i = 0; i < nin; i++)
                        st->w[(size_t)cc * nin + i] = x[(size_t)cc * mm + i];
                    st->b[cc] = x[(size_t)cc * mm + nin];
                }
                upload_decoder(c, st);
                if (ms_out)
                    *ms_out = c.ms_since(c.ev0);
                return;
            }
            // The split-bf16 Gram cannot settle this system (status 4: a feature nearly a combination of the ones
            // before it, where the 1e-9 ridge needs the Gram exact to far below its rounding - a level-0 channel
            // that holds one nonzero palette value everywhere makes s_i a multiple of the bias and s_i c_j of c_j),
            // or every rung was refused (3). The same call is redone from the fp64 Gram of the plain form and solved
            // on the host as that form solves it: the answer of the exact arithmetic, at a cost paid only in such a
            // state.
            g_fast_fallbacks++;
            c.form = NNTC_FORM_PLAIN;
            try
            {
                ls_accumulate(c, k, per_site, acc, true);
            }
            catch (...)
            {
                c.form = NNTC_FORM_FAST;
                throw;
            }
            c.form = NNTC_FORM_FAST;
            ms = c.ms_since(c.ev0);
        }
        double trace = 0.0;
        for (int i = 0; i < mm; i++)
            trace += acc[(size_t)tri_index(i, i, mm)];
        const double mean_diag = trace > 0.0 ? trace / mm : 1.0;
        auto qform = [&](const std::vector<double>& x, double& err) {
            double q = 0.0, e = 0.0;
            for (int cc = 0; cc < nout; cc++)
            {
                double qc = 0.0, ec = 0.0;
                for (int i = 0; i < mm; i++)
                {
                    for (int j = 0; j < mm; j++)
                    {
                        const double a = i <= j ? acc[(size_t)tri_index(i, j, mm)] : acc[(size_t)tri_index(j, i, mm)];
                        qc += x[(size_t)cc * mm + i] * a * x[(size_t)cc * mm + j];
                        ec += std::fabs(x[(size_t)cc * mm + i]) * std::fabs(a) * std::fabs(x[(size_t)cc * mm + j]);
                    }
                    qc -= 2.0 * x[(size_t)cc * mm + i] * acc[(size_t)tri + (size_t)cc * mm + i];
                    ec += 2.0 * std::fabs(x[(size_t)cc * mm + i]) * std::fabs(acc[(size_t)tri + (size_t)cc * mm + i]);
                }
                q += (double)c.md.cw[cc] * qc;
                e += (double)c.md.cw[cc] * ec;
            }
            err = DBL_EPSILON * (double)mm * e * inv_norm;
            return q;
        };
        double err_prev = 0.0;
        const double q_prev = had_previous ? qform(x_prev, err_prev) : 0.0;
        const int cols = mm + nout;
        std::vector<double> g((size_t)mm * cols, 0.0);
        std::vector<double> x_new((size_t)nout * mm, 0.0);
        bool accepted = false;
        size_t taken = 0;
        for (size_t step = 0; step < RIDGE_REFUSED && !accepted; step++)
        {
            const double ridge = RIDGE_LADDER[step] * mean_diag;
            std::fill(g.begin(), g.end(), 0.0);
            for (int i = 0; i < mm; i++)
            {
                for (int j = i; j < mm; j++)
                {
                    const double a = acc[(size_t)tri_index(i, j, mm)];
                    g[(size_t)i * cols + j] = a;
                    g[(size_t)j * cols + i] = a;
                }
                g[(size_t)i * cols + i] += ridge;
                for (int cc = 0; cc < nout; cc++)
                    g[(size_t)i * cols + mm + cc] = acc[(size_t)tri + (size_t)cc * mm + i];
            }
            if (!gauss_jordan(g, mm, nout))
                continue;
            bool finite = true;
            for (int cc = 0; cc < nout; cc++)
                for (int i = 0; i < mm; i++)
                {
                    const double v = (double)(float)g[(size_t)i * cols + mm + cc];
                    x_new[(size_t)cc * mm + i] = v;
                    finite = finite && std::isfinite(v);
                }
            if (!finite)
                continue;
            double err_cand = 0.0;
            const double q_cand = qform(x_new, err_cand);
            if (!had_previous || (norm_ok && (q_cand - q_prev) * inv_norm + err_prev + err_cand <= LS_ACCEPT_RISE))
            {
                accepted = true;
                taken = step;
            }
        }
        g_ridge_tally[accepted ? taken : RIDGE_REFUSED]++;
        if (accepted)
        {
            st->has_decoder = 1;
            for (int cc = 0; cc < nout; cc++)
            {
                for (int i = 0; i < nin; i++)
                    st->w[(size_t)cc * nin + i] = (float)x_new[(size_t)cc * mm + i];
                st->b[cc] = (float)x_new[(size_t)cc * mm + nin];
            }
        }
        else if (!had_previous)
        {
            st->has_decoder = 1;
            const double count = acc[(size_t)tri_index(nin, nin, mm)];
            for (int cc = 0; cc < nout; cc++)
            {
                for (int i = 0; i < nin; i++)
                    st->w[(size_t)cc * nin + i] = 0.0f;
                st->b[cc] = count > 0.0 ? (float)(acc[(size_t)tri + (size_t)cc * mm + nin] / count) : 0.0f;
            }
        }
        upload_decoder(c, st);
        if (ms_out)
            *ms_out = ms;
    });
}

void nntc_hip_decoder_ridge_tally(long long out[4])
{
    for (int i = 0; i < 4; i++)
        out[i] = g_ridge_tally[i];
}

long long nntc_hip_decoder_fast_fallbacks(void)
{
    return g_fast_fallbacks;
}

int nntc_hip_solve_level1_all(nntc_hip_ctx* h, const nntc_model_state* st, int k, double ridge,
                              nntc_level1_report* rep, double* ms)
{
    return guarded([&] {
        const double t = solve_block(h->c, st, k, ridge, false, rep);
        if (ms)
            *ms = t;
    });
}

int nntc_hip_sweep_level1_all(nntc_hip_ctx* h, const nntc_model_state* st, int k, double ridge, int sweeps,
                              nntc_level1_report* rep, double* ms)
{
    return guarded([&] {
        const double t = sweep_block(h->c, st, k, ridge, sweeps, false, rep);
        if (ms)
            *ms = t;
    });
}

int nntc_hip_solve_level0_cont_all(nntc_hip_ctx* h, const nntc_model_state* st, int k, double ridge,
                                   nntc_level1_report* rep, double* ms)
{
    return guarded([&] {
        const double t = solve_block(h->c, st, k, ridge, true, rep);
        if (ms)
            *ms = t;
    });
}

int nntc_hip_sweep_level0_cont_all(nntc_hip_ctx* h, const nntc_model_state* st, int k, double ridge, int sweeps,
                                   nntc_level1_report* rep, double* ms)
{
    return guarded([&] {
        const double t = sweep_block(h->c, st, k, ridge, sweeps, true, rep);
        if (ms)
            *ms = t;
    });
}

int nntc_hip_assemble_level0_for_refine(nntc_hip_ctx* h, const nntc_model_state* st, int k, double* ms)
{
    return guarded([&] {
        const double t = assemble_for_refine(h->c, st, k);
        if (ms)
            *ms = t;
    });
}

int nntc_hip_level1_download_indices(nntc_hip_ctx* h, int plane, uint8_t* k1, size_t n, size_t* count)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        download(c, c.k1.p + c.g(plane).off1 * c.c1, (size_t)c.g(plane).n1 * c.c1, k1, n, count);
    });
}

int nntc_hip_level1_off_grid(nntc_hip_ctx* h, const nntc_model_state* st, size_t* values, size_t* off)
{
    return guarded([&] {
        Ctx& c = h->c;
        const unsigned long long n = level1_off_grid(c, st->lo1, st->hi1);
        if (values)
            *values = (size_t)(c.n1_total * c.c1);
        if (off)
            *off = (size_t)n;
    });
}

int nntc_hip_level1_upload(nntc_hip_ctx* h, int plane, const float* v1, size_t n)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        upload(c, c.v1.p + c.g(plane).off1 * c.c1, (size_t)c.g(plane).n1 * c.c1, v1, n);
    });
}

int nntc_hip_level1_upload_indices(nntc_hip_ctx* h, int plane, const uint8_t* k1, size_t n)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        upload(c, c.k1.p + c.g(plane).off1 * c.c1, (size_t)c.g(plane).n1 * c.c1, k1, n);
    });
}

int nntc_hip_level0_upload_indices(nntc_hip_ctx* h, int plane, const uint8_t* k0, size_t n)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        upload(c, c.k0.p + c.g(plane).off0 * c.c0, (size_t)c.g(plane).n0 * c.c0, k0, n);
    });
}

int nntc_hip_upload_decoder_to_device(nntc_hip_ctx* h, const nntc_model_state* st)
{
    return guarded([&] { upload_decoder(h->c, st); });
}

int nntc_hip_upload_grids_to_device(nntc_hip_ctx* h, const nntc_model_state* st)
{
    return guarded([&] { upload_grids(h->c, st); });
}

int nntc_hip_level1_download(nntc_hip_ctx* h, int plane, float* v1, size_t n, size_t* count)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        download(c, c.v1.p + c.g(plane).off1 * c.c1, (size_t)c.g(plane).n1 * c.c1, v1, n, count);
    });
}

int nntc_hip_level1_delta(nntc_hip_ctx* h, int plane, double* delta, size_t n, size_t* count)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        download(c, c.lat1.x.p + c.g(plane).off1 * c.c1, (size_t)c.g(plane).n1 * c.c1, delta, n, count);
    });
}

int nntc_hip_level1_poke(nntc_hip_ctx* h, int plane, size_t index, float value)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        if (index >= (size_t)c.g(plane).n1 * c.c1)
            throw Refused { "level1_poke past the end of the plane" };
        HIP_OK(hipMemcpyAsync(c.v1.p + c.g(plane).off1 * c.c1 + index, &value, sizeof(float), hipMemcpyHostToDevice,
                              c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
    });
}

int nntc_hip_level1_range(nntc_hip_ctx* h, int plane, float* lo, float* hi)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        std::vector<float> v((size_t)c.g(plane).n1 * c.c1);
        download(c, c.v1.p + c.g(plane).off1 * c.c1, v.size(), v.data(), v.size(), nullptr);
        for (int q = 0; q < c.c1; q++)
        {
            lo[q] = 1e30f;
            hi[q] = -1e30f;
        }
        for (size_t i = 0; i < v.size(); i++)
        {
            const int q = (int)(i % (size_t)c.c1);
            lo[q] = v[i] < lo[q] ? v[i] : lo[q];
            hi[q] = v[i] > hi[q] ? v[i] : hi[q];
        }
    });
}

int nntc_hip_level0_download(nntc_hip_ctx* h, int plane, uint8_t* k0, size_t n, size_t* count)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        download(c, c.k0.p + c.g(plane).off0 * c.c0, (size_t)c.g(plane).n0 * c.c0, k0, n, count);
    });
}

int nntc_hip_level0_upload(nntc_hip_ctx* h, int plane, const float* v0, size_t n)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        upload(c, c.v0.p + c.g(plane).off0 * c.c0, (size_t)c.g(plane).n0 * c.c0, v0, n);
    });
}

int nntc_hip_level0_download_values(nntc_hip_ctx* h, int plane, float* v0, size_t n, size_t* count)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        download(c, c.v0.p + c.g(plane).off0 * c.c0, (size_t)c.g(plane).n0 * c.c0, v0, n, count);
    });
}

int nntc_hip_decode_plane(nntc_hip_ctx* h, int plane, uint8_t* rgb8, size_t n)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        decode_pass(c, plane, c.rgb8.p);
        download(c, (const uint8_t*)c.rgb8.p, (size_t)c.g(plane).n0 * c.nout, rgb8, n, nullptr);
    });
}

int nntc_hip_texture_sse(nntc_hip_ctx* h, int width, int height, unsigned long long* sse)
{
    return guarded([&] {
        Ctx& c = h->c;
        if (width < 0 || height < 0 || width > c.g(0).w0 || height > c.g(0).h0)
            throw Refused { "texture_sse: the extent exceeds the base plane" };
        texture_sse_pass(c, width, height, sse);
    });
}

int nntc_hip_texture_sse_noalloc(nntc_hip_ctx* h, int width, int height, unsigned long long* sse_device,
                                 void* stream)
{
    return guarded([&] {
        Ctx& c = h->c;
        texture_sse_noalloc_pass(c, width, height, sse_device, stream ? (hipStream_t)stream : c.stream);
    });
}

}   // extern "C"

namespace
{

// E and its report from the per-plane sums [weighted SE, centre SE, fractional SE] (src/objective.cu:233-279).
void objective_report(const Ctx& c, int k, const double* per_site, const std::vector<double>& acc,
                      nntc_objective* r)
{
    double sum_cw = 0.0;
    for (int q = 0; q < c.nout; q++)
        sum_cw += (double)c.md.cw[q];
    double num = 0.0, den = 0.0;
    *r = nntc_objective {};
    r->nplanes = c.planes.n;
    for (int p = 0; p < c.planes.n; p++)
    {
        const double pixels = (double)c.g(p).w0 * (double)c.g(p).h0;
        const double sites = pixels * (double)(1 + k);
        r->e_plane[p] = acc[(size_t)3 * p] / (sum_cw * sites);
        r->centre_plane[p] = acc[(size_t)3 * p + 1] / (pixels * (double)c.nout);
        num += per_site[p] * acc[(size_t)3 * p];
        den += per_site[p] * sites;
    }
    r->e = den > 0.0 ? num / (sum_cw * den) : 0.0;
    const double base = (double)c.g(0).w0 * (double)c.g(0).h0 * (double)c.nout;
    r->centre_mse = acc[1] / base;
    r->sampled_mse = k > 0 ? acc[2] / (base * (double)k) : 0.0;
}

}   // namespace

extern "C"
{

int nntc_hip_objective_eval(nntc_hip_ctx* h, int k, const double* per_site, nntc_objective* r)
{
    return guarded([&] {
        Ctx& c = h->c;
        const double t0 = now_ms();
        std::vector<double> acc((size_t)c.planes.n * 3);
        objective_pass(c, k, acc.data());
        const double ms = now_ms() - t0;
        objective_report(c, k, per_site, acc, r);
        r->ms = ms;
        g_objective_ms += ms;
        g_objective_calls++;
    });
}

int nntc_hip_objective_check_host(nntc_hip_ctx* h, const nntc_model_state* st, int k, const double* per_site,
                                  nntc_objective* r)
{
    return guarded([&] {
        Ctx& c = h->c;
        if (!st->has_decoder)
            throw Refused { "objective_check_host: the state has no decoder" };
        std::vector<double> acc((size_t)c.planes.n * 3);
        objective_check_host(c, st, k, acc.data());
        objective_report(c, k, per_site, acc, r);
        r->ms = 0.0;
    });
}

int nntc_hip_solve_level1_dense_host(nntc_hip_ctx* h, const nntc_model_state* st, int k, int plane, double lambda,
                                     const float* c_prev, size_t n_prev, double* delta, size_t n_delta,
                                     int max_unknowns, int* solved)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        if (!st->has_decoder)
            throw Refused { "solve_level1_dense_host: the state has no decoder" };
        const size_t nunk = (size_t)c.g(plane).n1 * c.c1;
        if (n_prev != nunk)
            throw Refused { "solve_level1_dense_host: c_prev must hold the plane's level-1 values" };
        std::vector<double> d;
        *solved = solve_level1_dense_host(c, st, k, plane, lambda, c_prev, d, max_unknowns) ? 1 : 0;
        if (*solved && n_delta < d.size())
            throw Refused { "solve_level1_dense_host: delta is shorter than the plane's unknowns" };
        for (size_t i = 0; i < d.size() && *solved; i++)
            delta[i] = d[i];
    });
}

double nntc_hip_objective_total_ms(void)
{
    return g_objective_ms;
}

long long nntc_hip_objective_passes(void)
{
    return g_objective_calls;
}

int nntc_hip_bc_pack_prepare(nntc_hip_ctx* h, const nntc_model_state* st, int k, int seed,
                             const uint8_t* const* k0_mirror, double* pack_psnr, size_t* pack_texels, double* ms_out)
{
    return guarded([&] {
        Ctx& c = h->c;
        double ms = assemble_for_refine(c, st, k);
        const double t0 = now_ms();
        set_k0_mirror(c, k0_mirror, seed != 0);
        if (seed)
            bc_seed(c);
        bc_decode(c);
        HIP_OK(hipStreamSynchronize(c.stream));
        ms += now_ms() - t0;
        blocks_from_device(c, pack_psnr, pack_texels);
        if (ms_out)
            *ms_out = ms;
    });
}

int nntc_hip_bc_refine(nntc_hip_ctx* h, int passes, const uint8_t* const* k0_mirror, nntc_bc_refine_report* rep,
                       double* pack_psnr, size_t* pack_texels, double* ms_out)
{
    return guarded([&] {
        Ctx& c = h->c;
        double ms = 0.0;
        set_k0_mirror(c, k0_mirror, false);
        if (passes <= 0)
        {
            if (ms_out)
                *ms_out = 0.0;
            return;
        }
        // Every pass's per-plane figures land in bc_fold after its first slot (the packing error's), so the passes,
        // the pack and its error run back to back and are read once at the end.
        const size_t np = (size_t)c.planes.n, need = 1 + 2 * np * (size_t)passes;
        if (c.bc_fold.n < need)
            c.bc_fold.alloc(need);
        while (c.pass_ev.size() < (size_t)passes + 1)
        {
            hipEvent_t e = nullptr;
            HIP_OK(hipEventCreate(&e));
            c.pass_ev.push_back(e);
        }
        const double t0 = now_ms();
        HIP_OK(hipEventRecord(c.pass_ev[0], c.stream));
        for (int pass = 0; pass < passes; pass++)
        {
            bc_refine_pass(c, c.bc_fold.p + 1 + 2 * np * (size_t)pass);
            HIP_OK(hipEventRecord(c.pass_ev[(size_t)pass + 1], c.stream));
        }
        bc_pack_blocks(c, c.bc_fold.p);
        const double* f = read_back(c, (const double*)c.bc_fold.p, need);
        ms = now_ms() - t0;
        for (int pass = 0; pass < passes; pass++)
        {
            float pass_ms = 0.0f;
            HIP_OK(hipEventElapsedTime(&pass_ms, c.pass_ev[(size_t)pass], c.pass_ev[(size_t)pass + 1]));
            nntc_bc_refine_report r {};
            r.pass = pass + 1;
            r.ms = pass_ms;
            const double* fp = f + 1 + 2 * np * (size_t)pass;
            for (size_t p = 0; p < np; p++)
            {
                r.delta_e += fp[2 * p];
                r.improved += (long long)fp[2 * p + 1];
                r.blocks += c.g((int)p).nb;
            }
            if (rep)
                rep[pass] = r;
        }
        pack_figures(c, f[0], pack_psnr, pack_texels);
        if (ms_out)
            *ms_out = ms;
    });
}

int nntc_hip_bc_blocks_from_device(nntc_hip_ctx* h, const uint8_t* const* k0_mirror, double* pack_psnr,
                                   size_t* pack_texels)
{
    return guarded([&] {
        Ctx& c = h->c;
        set_k0_mirror(c, k0_mirror, false);
        blocks_from_device(c, pack_psnr, pack_texels);
    });
}

int nntc_hip_bc_blocks_download(nntc_hip_ctx* h, int plane, int file, uint8_t* blocks, size_t n, size_t* count)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        const int files = c.c0 <= 2 ? 1 : 2;
        if (file < 0 || file >= files)
            throw Refused { "no such level-0 file" };
        const int nc = file == 0 ? (c.c0 == 1 ? 1 : 2) : c.c0 - 2;
        const int bpb = nc == 1 ? 8 : 16;
        const PlaneGeom& g = c.g(plane);
        download(c, (const uint8_t*)(c.bc_blocks.p + g.offb * 32 + (long long)file * g.nb * 16), (size_t)g.nb * bpb,
                 blocks, n, count);
    });
}

int nntc_hip_bc_state_save(nntc_hip_ctx* h, int plane, uint8_t* ep, size_t n_ep, uint8_t* sel, size_t n_sel)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        const PlaneGeom& g = c.g(plane);
        download(c, (const uint8_t*)(c.bc_ep.p + g.offb * c.c0 * 2), (size_t)g.nb * c.c0 * 2, ep, n_ep, nullptr);
        download(c, (const uint8_t*)(c.bc_sel.p + g.offb * c.c0 * 16), (size_t)g.nb * c.c0 * 16, sel, n_sel, nullptr);
    });
}

void nntc_hip_profile_enable(int on)
{
    profiler().on = on != 0;
}

void nntc_hip_profile_only(const char* kernel)
{
    Profiler& p = profiler();
    std::lock_guard<std::recursive_mutex> g(p.lock);
    p.only = kernel ? kernel : "";
}

void nntc_hip_profile_reset(void)
{
    (void)guarded([] { profiler().reset(); });
}

int nntc_hip_kernel_stats(int index, char* name_out, size_t name_len, long long* launches, double* ms, double* bytes,
                          double* flops)
{
    Profiler& p = profiler();
    std::lock_guard<std::recursive_mutex> g(p.lock);
    if (guarded([&] { p.settle(); }) != 0)
        return 1;
    if (index < 0 || (size_t)index >= p.stats.size())
        return 1;
    const auto& kv = p.stats[(size_t)index];
    if (name_out && name_len)
        std::snprintf(name_out, name_len, "%s", kv.first.c_str());
    if (launches)
        *launches = kv.second.launches;
    if (ms)
        *ms = kv.second.ms;
    if (bytes)
        *bytes = kv.second.bytes;
    if (flops)
        *flops = kv.second.flops;
    return 0;
}

int nntc_hip_bc_state_restore(nntc_hip_ctx* h, int plane, const uint8_t* ep, size_t n_ep, const uint8_t* sel,
                              size_t n_sel)
{
    return guarded([&] {
        Ctx& c = h->c;
        check_plane(c, plane);
        const PlaneGeom& g = c.g(plane);
        if (n_ep > (size_t)g.nb * c.c0 * 2 || n_sel > (size_t)g.nb * c.c0 * 16)
            throw Refused { "BC state restore was handed more blocks than the plane holds" };
        HIP_OK(hipMemcpyAsync(c.bc_ep.p + g.offb * c.c0 * 2, ep, n_ep, hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipMemcpyAsync(c.bc_sel.p + g.offb * c.c0 * 16, sel, n_sel, hipMemcpyHostToDevice, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
    });
}

int nntc_hip_state_snapshot(nntc_hip_ctx* h)
{
    return guarded([&] {
        Ctx& c = h->c;
        Ctx::Snapshot& s = c.snap;
        if (!s.v0.p)
        {
            s.v0.alloc(c.v0.n);
            s.v1.alloc(c.v1.n);
            s.k0.alloc(c.k0.n);
            s.k1.alloc(c.k1.n);
            s.ep.alloc(c.bc_ep.n);
            s.sel.alloc(c.bc_sel.n);
            s.mirror.alloc(c.k0_mirror.n);
            c.bytes += s.v0.bytes() + s.v1.bytes() + s.k0.bytes() + s.k1.bytes() + s.ep.bytes() + s.sel.bytes() +
                       s.mirror.bytes();
        }
        const auto d2d = hipMemcpyDeviceToDevice;
        HIP_OK(hipMemcpyAsync(s.v0.p, c.v0.p, c.v0.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(s.v1.p, c.v1.p, c.v1.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(s.k0.p, c.k0.p, c.k0.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(s.k1.p, c.k1.p, c.k1.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(s.ep.p, c.bc_ep.p, c.bc_ep.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(s.sel.p, c.bc_sel.p, c.bc_sel.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(s.mirror.p, c.k0_mirror.p, c.k0_mirror.bytes(), d2d, c.stream));
        s.held = true;
    });
}

int nntc_hip_state_restore(nntc_hip_ctx* h)
{
    return guarded([&] {
        Ctx& c = h->c;
        Ctx::Snapshot& s = c.snap;
        if (!s.held)
            throw Refused { "state_restore: no snapshot is held" };
        const auto d2d = hipMemcpyDeviceToDevice;
        HIP_OK(hipMemcpyAsync(c.v0.p, s.v0.p, c.v0.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(c.v1.p, s.v1.p, c.v1.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(c.k0.p, s.k0.p, c.k0.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(c.k1.p, s.k1.p, c.k1.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(c.bc_ep.p, s.ep.p, c.bc_ep.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(c.bc_sel.p, s.sel.p, c.bc_sel.bytes(), d2d, c.stream));
        HIP_OK(hipMemcpyAsync(c.k0_mirror.p, s.mirror.p, c.k0_mirror.bytes(), d2d, c.stream));
        HIP_OK(hipStreamSynchronize(c.stream));
    });
}

}   // extern "C"
