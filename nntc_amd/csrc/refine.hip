// refine.hip: level 0's BC4/BC5 pack and its refinement by analysis by synthesis, on gfx950 (src/refine_bc.cu,
// src/cpu/cpu_refine_bc.cpp, shared/bc_pack.h, src/export.cpp:294-363).
//
//   seed    every 4x4 block of every channel packed stb_dxt-style from the 8-bit index plane: the block's largest and
//           smallest byte as the endpoints (eight-value mode), each texel's nearest palette entry as its selector, one
//           thread per block-channel (integer, bit-exact);
//   decode  every texel's plane value, index and correction from its block's bytes;
//   refine  one lane per 4x4 block of one colour of the block grid (blocks two apart share no site): the block's
//           exact quadratic taken from block (c')'s stencil with every texel outside it held, each channel's
//           endpoints by a least-squares step and nine integer candidates, then two Gauss-Seidel passes of exact
//           8-way selector argmins, every step taken only on a strict decrease; every sum that feeds a decision is
//           folded over the sixteen texels in increasing order, as the reference CPU backend folds it;
//   pack    the endpoints and selectors written as the .dds's 8-byte BC4 blocks, and the pack's squared error in byte
//           units against the index plane the host mirrors (a report figure).

#include "arm.h"

#include <cmath>

namespace nntc
{

namespace
{

struct BcArgs
{
    Planes planes;
    int c0;
    const uint8_t* k0m;     // the index mirror (seed, error)
    uint8_t* ep;
    uint8_t* sel;
    float* v0;
    uint8_t* k0;
    double* delta;
    const float* prev;
    const double* stencil;  // level 0's, entry e of position s at stencil[e * sm.sn + s]
    const float* stencil32; // its fp32 copy, the one the fast form's gathers build
    StencilMap sm;
    const double* grad;
    const float* lo0;
    const float* hi0;
    double* taken;          // per block
    uint8_t* blocks;
    double* se;             // per block: packing error
};

// Block b (global) of plane p and channel c, seeded: bc_pack_level at 8 bits (shared/bc_pack.h:101-133, 196-224).
__global__ __launch_bounds__(256) void k_bc_seed(BcArgs a, long long total)
{
    const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total * a.c0)
        return;
    const long long gb = gi / a.c0;
    const int c = (int)(gi % a.c0);
    int p = 0;
    while (p + 1 < a.planes.n && gb >= a.planes.p[p + 1].offb)
        p++;
    const PlaneGeom& g = a.planes.p[p];
    const long long b = gb - g.offb;
    const int bx = (g.w0 + 3) / 4;
    const int x = (int)(b % bx), y = (int)(b / bx);
    const uint8_t* k0 = a.k0m + g.off0 * a.c0;
    int v[16];
    int kmin = 255, kmax = 0;
    for (int t = 0; t < 16; t++)
    {
        const int px = min(g.w0 - 1, x * 4 + t % 4), py = min(g.h0 - 1, y * 4 + t / 4);
        v[t] = k0[((size_t)py * g.w0 + px) * a.c0 + c];
        kmin = min(kmin, v[t]);
        kmax = max(kmax, v[t]);
    }
    const long long slot = gb * a.c0 + c;
    const int r0 = kmax, r1 = kmin;   // lround(k * 255.0 / 255) is k itself at 8 bits
    a.ep[slot * 2] = (uint8_t)r0;
    a.ep[slot * 2 + 1] = (uint8_t)r1;
    if (r0 == r1)
    {
        for (int t = 0; t < 16; t++)
            a.sel[slot * 16 + t] = 0;
        return;
    }
    float pal[8];
    for (int i = 0; i < 8; i++)
        pal[i] = bc4_entry(r0, r1, i);
    for (int t = 0; t < 16; t++)
    {
        const float target = (float)v[t] * 255.0f / 255.0f;
        int best = 0;
        float best_d = 1e30f;
        for (int i = 0; i < 8; i++)
        {
            const float d = fabsf(pal[i] - target);
            if (d < best_d)
            {
                best_d = d;
                best = i;
            }
        }
        a.sel[slot * 16 + t] = (uint8_t)best;
    }
}

// k_bc_decode: texel ti's plane value, index and correction from its block's bytes.
__global__ __launch_bounds__(256) void k_bc_decode(BcArgs a, long long total)
{
    const long long gt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt >= total)
        return;
    int p = 0;
    while (p + 1 < a.planes.n && gt >= a.planes.p[p + 1].off0)
        p++;
    const PlaneGeom& g = a.planes.p[p];
    const long long ti = gt - g.off0;
    const int px = (int)(ti % g.w0), py = (int)(ti / g.w0);
    const int bx = (g.w0 + 3) / 4;
    const long long gb = g.offb + (long long)(py / 4) * bx + (px / 4);
    const int t = (py % 4) * 4 + (px % 4);
    for (int c = 0; c < a.c0; c++)
    {
        const long long b = gb * a.c0 + c;
        const float byte = bc4_entry8(a.ep[b * 2], a.ep[b * 2 + 1], a.sel[b * 16 + t]);
        const float value = bc0_value(a.lo0[c], a.hi0[c], byte);
        const long long idx = gt * a.c0 + c;
        a.v0[idx] = value;
        a.k0[idx] = bc0_index(byte);
        a.delta[idx] = (double)value - (double)a.prev[idx];
    }
}

// ---------------------------------------------------------------------------------------------------------------------
// The refinement, one sixteen-lane DPP row per 4x4 block (src/cpu/cpu_refine_bc.cpp:138-384)
// ---------------------------------------------------------------------------------------------------------------------
//
// The reference refines a block with a chain of small serial decisions (a least-squares endpoint step, nine candidates,
// then 2 x 16 x c0 selector argmins, each one seeing the moves before it). Here a row of sixteen lanes carries one
// block and a wavefront four: lane u holds texel u's rows of the block's two n-vectors (a1: the linear term from the
// texels around the block; uv: the quadratic applied to the block's own corrections, kept current as the block moves)
// in registers, and reads its rows of the block's quadratic a2 straight from block (c')'s stencil: entry (texel u,
// channel i; texel v, channel j) of a2 is one entry of a stencil block of u or of v (the transpose of v's when v owns
// the pair), zero unless u and v are neighbours in the plane. Every decision is taken by all sixteen lanes from the same
// row-broadcast operands, so the row agrees on it without a barrier or a vote.
//
// Every sum folds its terms in the reference's order: a sum over the block's texels is gathered lane by lane with row
// broadcasts and added in increasing texel order, as the reference CPU backend folds it. The reference's row sums run
// over all sixteen texels of the block; the terms skipped here are products with a zero entry of a2, which add a zero
// to a sum that starts at +0.0 - a sum such a zero cannot change.

// Texel (lx, ly) of block (gx, gy), lx and ly in -1..4, in the stencil's runs: base[col] + (ly >> 1) cw[col] + (lx >> 1).
struct BlockPos
{
    // The colour's base as b0 + ox d1 + oy d2 + ox oy d3 and its row length as cw0 - ox (cw0 - cw1): arithmetic, not a
    // table or a select between fields, which the compiler turns into an indexed load from scratch when the colour is
    // not a constant.
    long long b0, d1, d2, d3;
    int cw0, cw1;
    __device__ __forceinline__ void init(const long long* cpos, int w, int gx, int gy)
    {
        cw0 = (w + 1) >> 1;
        cw1 = w >> 1;
        const long long c0 = cpos[0] + (long long)(2 * gy) * cw0 + 2 * gx;
        const long long c1 = cpos[1] + (long long)(2 * gy) * cw1 + 2 * gx;
        const long long c2 = cpos[2] + (long long)(2 * gy) * cw0 + 2 * gx;
        const long long c3 = cpos[3] + (long long)(2 * gy) * cw1 + 2 * gx;
        b0 = c0;
        d1 = c1 - c0;
        d2 = c2 - c0;
        d3 = c3 - c2 - c1 + c0;
    }
    __device__ __forceinline__ long long at(int lx, int ly) const
    {
        const long long ox = lx & 1, oy = ly & 1;
        const long long b = b0 + ox * d1 + oy * d2 + (ox & oy) * d3;
        const long long w = cw0 - ox * (cw0 - cw1);
        return b + (long long)(ly >> 1) * w + (lx >> 1);
    }
};

// Texel u of a block, as its lane sees the block; ST is the stencil's element type (double, or the fast form's float).
template <int C0, class ST>
struct BlockLane
{
    const ST* S;
    long long sn;
    BlockPos bp;
    int u, ux, uy;
    long long pu;           // texel u's stencil position
    unsigned int inside;    // bit v: texel v of the block is in the plane
    unsigned int nbmask;    // bit k = (dy + 1) * 3 + dx + 1: texel u + (dx, dy) is in the block, and it and u in the plane

    __device__ __forceinline__ bool in(int v) const { return (inside >> v) & 1u; }

    __device__ __forceinline__ void init_neighbours()
    {
        nbmask = 0;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++)
            {
                const bool ok = ux + dx >= 0 && ux + dx < 4 && uy + dy >= 0 && uy + dy < 4 && in(u) &&
                                in((u + 4 * dy + dx) & 15);
                nbmask |= (ok ? 1u : 0u) << ((dy + 1) * 3 + dx + 1);
            }
    }

    // The lane's geometry made opaque to the optimiser, so that addresses and offsets derived from it are computed where
    // they are used rather than hoisted out of a loop and held in registers across it.
    __device__ __forceinline__ void opaque()
    {
        asm volatile("" : "+v"(pu), "+v"(bp.b0), "+v"(bp.d1), "+v"(bp.d2), "+v"(bp.d3));
        asm volatile("" : "+v"(ux), "+v"(uy), "+v"(u), "+v"(nbmask), "+v"(bp.cw0), "+v"(bp.cw1));
    }

    // 1.0 for a neighbour k whose entries a2 holds, 0.0 otherwise; opaque, so that an entry weighted by it is loaded
    // whatever its value rather than behind a branch.
    __device__ __forceinline__ double keep(int k) const
    {
        double w = ((nbmask >> k) & 1u) ? 1.0 : 0.0;
        asm volatile("" : "+v"(w));
        return w;
    }

    // a2's entry (texel u, channel i; texel u + (dx, dy), channel j), weighted by keep: dx, dy, i and j are constants
    // after unrolling, so the entry's offset is uniform. The forward slots (1, 0), (-1, 1), (0, 1), (1, 1) are u's
    // blocks, the other four the neighbour's, transposed. A neighbour a2 does not hold reads the lane's own run instead
    // (a valid address, a finite value) and is weighted by 0.0: the entry is then a zero, as a2's is.
    __device__ __forceinline__ double a2(int i, int dx, int dy, int j) const
    {
        const int k = (dy + 1) * 3 + dx + 1;
        const ST* e;
        if (dx == 0 && dy == 0)
            e = S + (long long)(i * C0 + j) * sn + pu;
        else if (dy == 1 || (dy == 0 && dx == 1))
            e = S + (long long)((dy == 0 ? 1 : dx + 3) * C0 * C0 + i * C0 + j) * sn + pu;
        else
        {
            const long long pv = ((nbmask >> k) & 1u) ? bp.at(ux + dx, uy + dy) : pu;
            e = S + (long long)((dy == 0 ? 1 : 3 - dx) * C0 * C0 + j * C0 + i) * sn + pv;
        }
        return (double)*e * keep(k);
    }

};


// Lane K of each sixteen-lane row, to the whole row (DPP row_newbcast).
template <int K>
__device__ __forceinline__ double row_bcast(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), 0x150 + K, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), 0x150 + K, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Lane i + OFF of lane i's row (DPP row_shl / row_shr); a lane whose source falls outside the row reads zero.
template <int OFF>
__device__ __forceinline__ double row_from(double x)
{
    if constexpr (OFF == 0)
        return x;
    else
    {
        constexpr int ctrl = OFF > 0 ? 0x100 + OFF : 0x110 - OFF;
        const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), ctrl, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), ctrl, 0xF, 0xF, false);
        return __hiloint2double(hi, lo);
    }
}

// ((0.0 + t_0) + t_1) + ... + t_15 over the row's lanes in lane order: a sum over the block's texels in increasing
// texel order, the same on every lane of the row.
template <int K = 0>
__device__ __forceinline__ double row_fold(double t, double acc = 0.0)
{
    acc += row_bcast<K>(t);
    if constexpr (K < 15)
        return row_fold<K + 1>(t, acc);
    else
        return acc;
}

// v(u + (DX, DY)) for each of the nine neighbours, read from their lanes (a neighbour outside the block reads another
// lane's or a zero: a2 weighs it by zero).
#define NNTC_FROM_NEIGHBOUR(DX, DY) out[((DY) + 1) * 3 + (DX) + 1] = row_from<4 * (DY) + (DX)>(v);
__device__ __forceinline__ void from_neighbours(double v, double (&out)[9])
{
    NNTC_FROM_NEIGHBOUR(-1, -1)
    NNTC_FROM_NEIGHBOUR(0, -1)
    NNTC_FROM_NEIGHBOUR(1, -1)
    NNTC_FROM_NEIGHBOUR(-1, 0)
    NNTC_FROM_NEIGHBOUR(0, 0)
    NNTC_FROM_NEIGHBOUR(1, 0)
    NNTC_FROM_NEIGHBOUR(-1, 1)
    NNTC_FROM_NEIGHBOUR(0, 1)
    NNTC_FROM_NEIGHBOUR(1, 1)
}
#undef NNTC_FROM_NEIGHBOUR

__device__ __forceinline__ int sel_get(unsigned long long bits, int t)
{
    return (int)((bits >> (3 * t)) & 7ull);
}

__device__ __forceinline__ unsigned long long sel_set(unsigned long long bits, int t, int v)
{
    return (bits & ~(7ull << (3 * t))) | ((unsigned long long)v << (3 * t));
}

// The plane value of selector s of a block with endpoints r0 > r1 (eight-value mode), computed rather than looked up
// (a table indexed by a lane's own selector would live in scratch).
__device__ __forceinline__ float level0_value(float lo, float hi, int r0, int r1, int s)
{
    return bc0_value(lo, hi, bc4_entry8(r0, r1, s));
}

// The block's texels (x + dx, y + dy) that are texel (x, y) itself or its 8-neighbours, in increasing index order
// (dy, then dx); dx and dy are constants in BODY.
#define FOR_NEIGHBOURS(x, y, dx, dy, BODY)                                                                             \
    _Pragma("unroll") for (int dy = -1; dy <= 1; dy++)                                                                 \
        _Pragma("unroll") for (int dx = -1; dx <= 1; dx++)                                                             \
        {                                                                                                              \
            if ((x) + dx >= 0 && (x) + dx < 4 && (y) + dy >= 0 && (y) + dy < 4)                                        \
            {                                                                                                          \
                BODY                                                                                                   \
            }                                                                                                          \
        }

// The nine texels (x + DX, y + DY) around texel (x, y), itself included, in increasing index order (DY, then DX), with
// DX and DY literal constants (for the DPP shifts).
#define NEIGHBOURS9(M) M(-1, -1) M(0, -1) M(1, -1) M(-1, 0) M(0, 0) M(1, 0) M(-1, 1) M(0, 1) M(1, 1)

// Selector steps T..15 of one channel of one Gauss-Seidel pass: texel T's exact 8-way argmin, taken from texel T's
// linear term and diagonal broadcast over the row, and on a strict decrease the move added into the rows of T's
// neighbours. Lane u's entry (u, ii; T, c) sits in its LDS column at e[k * C0 * 64 * C0 + ii] for T = u + offset k: an
// offset that differs from lane to lane, which in registers would take a select per slot.
template <int C0, int T, class BL, class ST>
__device__ __forceinline__ void selector_steps(const BL& L, int c, const double (&a1)[C0], double (&uv)[C0],
                                               const ST* e, double diag, const float (&pal)[8], float lo, float hi,
                                               int r0, int r1, unsigned long long& sc, double& taken)
{
    constexpr int tx = T & 3, ty = T >> 2;
    const double b = row_bcast<T>(a1[c] + uv[c]);
    const double aa = row_bcast<T>(diag);
    const int s = sel_get(sc, T);
    // The texel's value now, its selector's palette entry (picked by selects: an array indexed by a lane's value would
    // live in scratch).
    float pnow = pal[0];
#pragma unroll
    for (int kk = 1; kk < 8; kk++)
        pnow = s == kk ? pal[kk] : pnow;
    const double now = (double)pnow;
    double best = 0.0, move = 0.0;
    int which = s;
#pragma unroll
    for (int kk = 0; kk < 8; kk++)
    {
        const double d = (double)pal[kk] - now;
        const double de = 2.0 * b * d + aa * d * d;
        if (de < best)
        {
            best = de;
            move = d;
            which = kk;
        }
    }
    if (best < 0.0)
    {
        taken += best;
        const int dx = tx - L.ux, dy = ty - L.uy;
        if (dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1)
        {
            const ST* ek = e + ((dy + 1) * 3 + dx + 1) * C0 * 64 * C0;
#pragma unroll
            for (int ii = 0; ii < C0; ii++)
                uv[ii] += (double)ek[ii] * move;
        }
        // Opaque, so the shift is not folded into the argmin's selects as eight shifted constants (which VOP3 cannot
        // encode, so each would hold a register).
        asm volatile("" : "+v"(which));
        sc = sel_set(sc, T, which);
    }
    if constexpr (T < 15)
        selector_steps<C0, T + 1>(L, c, a1, uv, e, diag, pal, lo, hi, r0, r1, sc, taken);
}

// The lane's entries a2(i, neighbour K, j) of the initial uv = a2 dv, e[j][i]. The geometry is made opaque first, so
// the loads cannot be issued before this point.
template <int C0, int K, class BL>
__device__ __forceinline__ void uv_entries(BL& L, double (&e)[C0][C0])
{
    constexpr int dx = K % 3 - 1, dy = K / 3 - 1;
    L.opaque();
#pragma unroll
    for (int j = 0; j < C0; j++)
#pragma unroll
        for (int i = 0; i < C0; i++)
            e[j][i] = L.a2(i, dx, dy, j);
}

// Neighbour K's terms of the lane's rows of uv, channel by channel, with the neighbour's correction read from its lane;
// uv is made opaque after them, so they are added before the loads that follow are issued. The entries go to the
// lane's LDS column, col[((K C0 + j) 64) C0 + i], where every later step reads them.
template <int C0, int K, class ST>
__device__ __forceinline__ void uv_add(const double (&e)[C0][C0], const double (&dv)[C0], double (&uv)[C0], ST* col)
{
#pragma unroll
    for (int j = 0; j < C0; j++)
#pragma unroll
        for (int i = 0; i < C0; i++)
            col[((K * C0 + j) * 64) * C0 + i] = (ST)e[j][i];
    constexpr int dx = K % 3 - 1, dy = K / 3 - 1;
    double dn[C0];
#pragma unroll
    for (int j = 0; j < C0; j++)
        dn[j] = row_from<4 * dy + dx>(dv[j]);
#pragma unroll
    for (int j = 0; j < C0; j++)
#pragma unroll
        for (int i = 0; i < C0; i++)
            uv[i] += e[j][i] * dn[j];
#pragma unroll
    for (int i = 0; i < C0; i++)
        asm volatile("" : "+v"(uv[i]));
}

// uv = a2 dv over the nine neighbours in increasing order (the order of the reference's row sums), neighbour K + 1's
// entries loaded while neighbour K's terms are added: two neighbours' entries in flight, not all nine.
template <int C0, int K, class BL, class ST>
__device__ __forceinline__ void uv_pipeline(BL& L, const double (&dv)[C0], double (&uv)[C0],
                                            const double (&cur)[C0][C0], ST* col)
{
    if constexpr (K < 8)
    {
        double nxt[C0][C0];
        uv_entries<C0, K + 1>(L, nxt);
        uv_add<C0, K>(cur, dv, uv, col);
        uv_pipeline<C0, K + 1>(L, dv, uv, nxt, col);
    }
    else
        uv_add<C0, K>(cur, dv, uv, col);
}

template <int C0, class ST>
__global__ __launch_bounds__(64) void k_bc_refine(BcArgs a, Ranges colour, int colour_index)
{
    // Four blocks per wavefront, one per sixteen-lane row; a row past the colour's last block leaves as a whole (the
    // rows of a wavefront never read each other's lanes).
    const long long gi = (long long)xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 4);
    if (gi >= colour.start[colour.n])
        return;
    // The lane's entries of a2, loaded from the stencil once (by the start of uv) and read from here by every later
    // step: one column per lane, entry (u, i; neighbour k, j) at [((k C0 + j) 64 + lane) C0 + i]. A lane reads only
    // its own column, so no barrier is needed. Reading the stencil again in each step instead would miss the L2 (the
    // rows a CU's wavefronts hold exceed it) and fetch the entries from HBM once per step.
    __shared__ ST s_E[9 * C0 * 64 * C0];
    ST* const col = s_E + threadIdx.x * C0;

    const int p = range_plane(colour, gi);
    const PlaneGeom& g = a.planes.p[p];
    const long long ci = gi - colour.start[p];
    const int bxn = (g.w0 + 3) / 4;
    const int cx = colour_index & 1, cy = colour_index >> 1;
    const int cbx = (bxn - cx + 1) / 2;
    const int gx = cx + 2 * (int)(ci % cbx);
    const int gy = cy + 2 * (int)(ci / cbx);
    const int w0 = g.w0, h0 = g.h0;
    const long long base = g.off0;
    const int px0 = gx * 4, py0 = gy * 4;
    const ST* S;
    if constexpr (sizeof(ST) == sizeof(float))
        S = a.stencil32;
    else
        S = a.stencil;
    const long long sn = a.sm.sn;

    BlockLane<C0, ST> L;   // not const: opaque() launders its fields
    L.S = S;
    L.sn = sn;
    L.bp.init(a.sm.cpos[p], w0, gx, gy);
    const int u = threadIdx.x & 15, ux = u & 3, uy = u >> 2;
    L.u = u;
    L.ux = ux;
    L.uy = uy;
    {
        const int nx = min(4, w0 - px0), ny = min(4, h0 - py0);
        const unsigned int row = (1u << nx) - 1u;
        L.inside = (row | row << 4 | row << 8 | row << 12) & ((1u << (4 * ny)) - 1u);
    }
    const bool in_u = L.in(u);
    // A texel outside the plane takes the block's first texel's position: its entries are read (weighted by zero), so
    // the address must be one the stencil holds.
    L.pu = in_u ? L.bp.at(ux, uy) : L.bp.at(0, 0);
    L.init_neighbours();
    auto inside_xy = [=](int lx, int ly) {
        return px0 + lx >= 0 && px0 + lx < w0 && py0 + ly >= 0 && py0 + ly < h0;
    };
    // The chain index of texel (lx, ly) of the block (or of a neighbour just outside it).
    auto tix = [=](int lx, int ly) { return base + (long long)(py0 + ly) * w0 + (px0 + lx); };
    const long long tu = tix(ux, uy);

    // ---- a1: the texel's gradient plus the shares of its neighbours outside the block, slot by slot, the forward
    // neighbour's then the backward one's (slot 0 is the texel itself, inside the block).
    double a1[C0], uv[C0];
#pragma unroll
    for (int i = 0; i < C0; i++)
        a1[i] = 0.0;
    if (in_u)
    {
        double acc[C0];
#pragma unroll
        for (int i = 0; i < C0; i++)
            acc[i] = a.grad[tu * C0 + i];
#pragma unroll
        for (int slot = 1; slot < STENCIL_BLOCKS; slot++)
        {
            int dx, dy;
            stencil_offset(slot, dx, dy);
            // A neighbour that adds nothing (inside the block, or outside the plane) reads the texel's own values (valid
            // addresses, finite values) with an opaque weight 0.0: its terms are zeros. A zero leaves the sum unchanged
            // but for turning a -0.0 into +0.0, and a1 enters every use as a1 + uv, with uv a sum from +0.0 that never
            // holds -0.0: either zero gives the same a1 + uv.
            const int fx = ux + dx, fy = uy + dy;
            const bool fok = !(fx >= 0 && fx < 4 && fy >= 0 && fy < 4) && inside_xy(fx, fy);
            double fkeep = fok ? 1.0 : 0.0;
            asm volatile("" : "+v"(fkeep));
            {
                const double* xn = a.delta + (fok ? tix(fx, fy) : tu) * C0;
                double xv[C0];
#pragma unroll
                for (int j = 0; j < C0; j++)
                    xv[j] = xn[j];
#pragma unroll
                for (int i = 0; i < C0; i++)
#pragma unroll
                    for (int j = 0; j < C0; j++)
                        acc[i] += ((double)S[(long long)(slot * C0 * C0 + i * C0 + j) * sn + L.pu] * xv[j]) * fkeep;
            }
            const int bx = ux - dx, by = uy - dy;
            const bool bok = !(bx >= 0 && bx < 4 && by >= 0 && by < 4) && inside_xy(bx, by);
            double bkeep = bok ? 1.0 : 0.0;
            asm volatile("" : "+v"(bkeep));
            {
                const double* xn = a.delta + (bok ? tix(bx, by) : tu) * C0;
                const long long pb = bok ? L.bp.at(bx, by) : L.pu;
                double xv[C0];
#pragma unroll
                for (int j = 0; j < C0; j++)
                    xv[j] = xn[j];
#pragma unroll
                for (int i = 0; i < C0; i++)
#pragma unroll
                    for (int j = 0; j < C0; j++)
                        acc[i] += ((double)S[(long long)(slot * C0 * C0 + j * C0 + i) * sn + pb] * xv[j]) * bkeep;
            }
        }
#pragma unroll
        for (int i = 0; i < C0; i++)
            a1[i] = acc[i];
    }

    // ---- uv = a2 dv: the lane's rows, their terms added in increasing column order (neighbour, then channel), the
    // order of the reference's row sums; each neighbour's correction is read from its lane.
    double dv[C0];
#pragma unroll
    for (int j = 0; j < C0; j++)
    {
        dv[j] = in_u ? (double)a.v0[tu * C0 + j] - (double)a.prev[tu * C0 + j] : 0.0;
        uv[j] = 0.0;
    }
    {
        double e0[C0][C0];
        uv_entries<C0, 0>(L, e0);
        uv_pipeline<C0, 0>(L, dv, uv, e0, col);
    }

    const long long gb = g.offb + (long long)gy * bxn + gx;
    int ep0[C0], ep1[C0];
    unsigned long long sel[C0];
#pragma unroll
    for (int c = 0; c < C0; c++)
    {
        ep0[c] = a.ep[(gb * C0 + c) * 2];
        ep1[c] = a.ep[(gb * C0 + c) * 2 + 1];
        const uint8_t* s = a.sel + (gb * C0 + c) * 16;
        unsigned long long bits = 0;
#pragma unroll
        for (int t = 0; t < 16; t++)
            bits |= (unsigned long long)(s[t] & 7) << (3 * t);
        sel[c] = bits;
    }
    double taken = 0.0;

    // ---- (i) the endpoints of each channel: each lane's terms of the five sums, folded over the row in texel order
#pragma unroll
    for (int c = 0; c < C0; c++)
    {
        const float lo = a.lo0[c], hi = a.hi0[c];
        const double step = ((double)hi - (double)lo) / 255.0;
        const unsigned long long sc = sel[c];
        // The lane's rows of a2's column c (zero where a2 holds nothing): the endpoint step, its check and its update
        // read these alone, so a term for a neighbour outside the block or the plane is a zero, which leaves the sum
        // (from +0.0, never -0.0) unchanged, as leaving it out would.
        double E[9][C0];
#pragma unroll
        for (int k = 0; k < 9; k++)
#pragma unroll
            for (int i = 0; i < C0; i++)
                E[k][i] = col[((k * C0 + c) * 64) * C0 + i];
        double bp_, bq, pap, paq, qaq;
        {
            // The texel's weight on the high endpoint, and its neighbours' from their lanes.
            const double aw = bc4_weight8(sel_get(sc, u));
            double aj[9];
            from_neighbours(aw, aj);
            const double pp = step * aw;
            const double qq = step * (1.0 - aw);
            const double b = a1[c] + uv[c];
            double ap = 0.0, aq = 0.0;
#pragma unroll
            for (int k = 0; k < 9; k++)
            {
                ap += E[k][c] * step * aj[k];
                aq += E[k][c] * step * (1.0 - aj[k]);
            }
            bp_ = row_fold(b * pp);
            bq = row_fold(b * qq);
            pap = row_fold(pp * ap);
            paq = row_fold(pp * aq);
            qaq = row_fold(qq * aq);
        }
        const int r0 = ep0[c], r1 = ep1[c];
        const double det = pap * qaq - paq * paq;
        double d0 = 0.0, d1 = 0.0;
        if (det > 1e-12 * pap * qaq)
        {
            d0 = (paq * bq - qaq * bp_) / det;
            d1 = (paq * bp_ - pap * bq) / det;
        }
        const double v0r = (double)r0 + d0, v1r = (double)r1 + d1;
        const int c0r = (int)floor(v0r + 0.5), c1r = (int)floor(v1r + 0.5);
        double best = 0.0;
        int best0 = r0, best1 = r1;
#pragma unroll
        for (int e0 = -1; e0 <= 1; e0++)
#pragma unroll
            for (int e1 = -1; e1 <= 1; e1++)
            {
                const int x0 = c0r + e0 < 0 ? 0 : (c0r + e0 > 255 ? 255 : c0r + e0);
                const int x1 = c1r + e1 < 0 ? 0 : (c1r + e1 > 255 ? 255 : c1r + e1);
                if (x0 <= x1)
                    continue;
                const double f0 = (double)(x0 - r0), f1 = (double)(x1 - r1);
                const double de = 2.0 * (bp_ * f0 + bq * f1) + pap * f0 * f0 + 2.0 * paq * f0 * f1 + qaq * f1 * f1;
                if (de < best)
                {
                    best = de;
                    best0 = x0;
                    best1 = x1;
                }
            }
        if (best0 == r0 && best1 == r1)
            continue;
        // The candidate weighed again in the arithmetic the plane is written in: texel v moves by
        // next(v) - now(v), the two palettes' values at its selector; the lane works out its own texel's move and reads
        // its neighbours' from their lanes.
        const int su = sel_get(sc, u);
        const double dlt = (double)level0_value(lo, hi, best0, best1, su) - (double)level0_value(lo, hi, r0, r1, su);
        double dn[9];
        from_neighbours(dlt, dn);
        double quad = 0.0;
#pragma unroll
        for (int k = 0; k < 9; k++)
            quad += E[k][c] * dn[k];
        const double exact = row_fold(dlt * (2.0 * (a1[c] + uv[c]) + quad));
        if (!(exact < 0.0))
            continue;
        taken += exact;
#pragma unroll
        for (int i = 0; i < C0; i++)
        {
            double add = 0.0;
#pragma unroll
            for (int k = 0; k < 9; k++)
                add += E[k][i] * dn[k];
            uv[i] += add;
        }
        ep0[c] = best0;
        ep1[c] = best1;
    }

    // ---- (ii) the selectors, two Gauss-Seidel passes of exact 8-way argmins
    float pal[C0][8];
#pragma unroll
    for (int c = 0; c < C0; c++)
#pragma unroll
        for (int k = 0; k < 8; k++)
            pal[c][k] = bc0_value(a.lo0[c], a.hi0[c], bc4_entry8(ep0[c], ep1[c], k));
#pragma unroll 1
    for (int pass = 0; pass < 2; pass++)
#pragma unroll
        for (int c = 0; c < C0; c++)
        {
            // Addresses are formed in the channel that uses them: hoisted out of the pass loop, every channel's and
            // every step's would be live at once.
            L.opaque();
            int lane_off = (int)threadIdx.x * C0 + c * 64 * C0;
            asm volatile("" : "+v"(lane_off));
            const ST* const ec = s_E + lane_off;   // the lane's column of a2's column channel c
            const double diag = ec[4 * C0 * 64 * C0 + c];
            selector_steps<C0, 0>(L, c, a1, uv, ec, diag, pal[c], a.lo0[c], a.hi0[c], ep0[c], ep1[c], sel[c], taken);
        }

    // ---- what the block leaves behind: lane u its texel's selectors and plane values, lane 0 the block's endpoints
#pragma unroll
    for (int c = 0; c < C0; c++)
    {
        if (u == 0)
        {
            a.ep[(gb * C0 + c) * 2] = (uint8_t)ep0[c];
            a.ep[(gb * C0 + c) * 2 + 1] = (uint8_t)ep1[c];
        }
        a.sel[(gb * C0 + c) * 16 + u] = (uint8_t)sel_get(sel[c], u);
    }
    if (in_u)
    {
#pragma unroll
        for (int c = 0; c < C0; c++)
        {
            const float byte = bc4_entry8(ep0[c], ep1[c], sel_get(sel[c], u));
            const float value = bc0_value(a.lo0[c], a.hi0[c], byte);
            const long long idx = tu * C0 + c;
            a.v0[idx] = value;
            a.k0[idx] = bc0_index(byte);
            a.delta[idx] = (double)value - (double)a.prev[idx];
        }
    }
    if (u == 0)
        a.taken[gb] = taken;
}

// ---------------------------------------------------------------------------------------------------------------------
// The fast form's refinement: the same lanes, stencil reads and decisions, in fp32
// ---------------------------------------------------------------------------------------------------------------------
//
// k_bc_refine_fast takes the reference's decisions (src/cpu/cpu_refine_bc.cpp:138-384: the least-squares endpoint step,
// nine candidates weighed again exactly, two Gauss-Seidel passes of 8-way selector argmins, a move only on a strict
// decrease) with three changes of arithmetic and schedule that only the exact form's bit-identity forbids:
//   - fp32 throughout (the stencil is the fast form's fp32 copy already), the 2 x 2 endpoint solve in fp64;
//   - a sum over the block is a butterfly across the row's sixteen lanes (quad xor 1, xor 2, half-row and row
//     mirrors), which leaves the same value on every lane, so the row still agrees on every decision;
//   - a selector step weighs its eight candidates one per lane (lane u and u + 8 weigh selector u & 7) and finds the
//     winner by a three-level max over a packed key: the decrease's fp32 bits with its low three bits replaced by
//     7 - selector, so the largest decrease wins and a tie goes to the lower selector, as the sequential scan takes it.
// The LDS column holds each lane's five forward blocks only (center, right, down-left, down, down-right), half the
// nine a full column takes: a2(u, i; v, j) for v behind u is v's forward entry (j, i) toward u, read from v's column.
// A texel outside the block or the plane has its entries zeroed in its own column, so every read that reaches past the
// block lands on a zero.

template <int CTRL>
__device__ __forceinline__ float dpp_f(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}

template <int CTRL>
__device__ __forceinline__ unsigned int dpp_u(unsigned int x)
{
    return (unsigned int)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}

// Lane K of each sixteen-lane row, to the whole row (DPP row_newbcast).
template <int K>
__device__ __forceinline__ float row_bcast_f(float x)
{
    return dpp_f<0x150 + K>(x);
}

// Lane i + OFF of lane i's row; a lane whose source falls outside the row reads zero.
template <int OFF>
__device__ __forceinline__ float row_from_f(float x)
{
    if constexpr (OFF == 0)
        return x;
    else
        return dpp_f<(OFF > 0 ? 0x100 + OFF : 0x110 - OFF)>(x);
}

// The sum of x over the row's sixteen lanes, the same value on every lane.
__device__ __forceinline__ float row_sum_f(float x)
{
    x += dpp_f<0xB1>(x);    // quad_perm [1, 0, 3, 2]
    x += dpp_f<0x4E>(x);    // quad_perm [2, 3, 0, 1]
    x += dpp_f<0x141>(x);   // row_half_mirror
    x += dpp_f<0x140>(x);   // row_mirror
    return x;
}

// The largest key over each eight-lane half of the row, on all eight lanes.
__device__ __forceinline__ unsigned int oct_max(unsigned int k)
{
    k = max(k, dpp_u<0xB1>(k));
    k = max(k, dpp_u<0x4E>(k));
    k = max(k, dpp_u<0x141>(k));
    return k;
}

// Direction k = (dy + 1) 3 + dx + 1 of the nine neighbours, as a lane offset in the row, and the forward slot of the
// opposite direction (the slot of the neighbour's block that holds the pair) for k < 4.
constexpr int nb_off(int k)
{
    return 4 * (k / 3 - 1) + (k % 3 - 1);
}

// x of the lane at direction k from this one (k a constant once the loop over the neighbours is unrolled).
__device__ __forceinline__ float from_dir(float x, int k)
{
    switch (k)
    {
    case 0: return row_from_f<nb_off(0)>(x);
    case 1: return row_from_f<nb_off(1)>(x);
    case 2: return row_from_f<nb_off(2)>(x);
    case 3: return row_from_f<nb_off(3)>(x);
    case 5: return row_from_f<nb_off(5)>(x);
    case 6: return row_from_f<nb_off(6)>(x);
    case 7: return row_from_f<nb_off(7)>(x);
    case 8: return row_from_f<nb_off(8)>(x);
    default: return x;
    }
}

__device__ __forceinline__ float bc4_weight8f(int sel)
{
    return sel == 0 ? 1.0f : (sel == 1 ? 0.0f : (float)(8 - sel) / 7.0f);
}

// Selector step T of channel c, then the steps after it. nbm bit T: texel T is u or one of its eight neighbours.
template <int C0, int T>
__device__ __forceinline__ void fast_selector_steps(const float* s_F, int lane, int rowbase, int u, unsigned int nbm,
                                                    int c, float a1c, float (&uv)[C0], float diag, float palk,
                                                    int& selu, float& cur, float& taken)
{
    // The lane's entries a2(u, ii; T, c): its own forward block toward T (T at or ahead of u in the row), or T's
    // forward block toward u read from T's column. The forward slot of lane offset o = |T - u| is o - (o >= 3); a lane
    // that is not T's neighbour reads a valid entry and weighs it by zero. The geometry is made opaque first, so that
    // the step's addresses are formed here, not all sixteen at once.
    asm volatile("" : "+v"(lane), "+v"(rowbase), "+v"(u), "+v"(nbm));
    const int off = T - u;
    const bool fwd = off >= 0;
    const int ao = fwd ? off : -off;
    const int kf = min(4, ao - (ao >= 3 ? 1 : 0));
    const int at = kf * (C0 * 64 * C0) + (fwd ? (lane * C0 + c * 64 * C0) : ((rowbase + T) * C0 + c));
    const int stride = fwd ? 1 : 64 * C0;
    float e[C0];
#pragma unroll
    for (int ii = 0; ii < C0; ii++)
        e[ii] = s_F[at + ii * stride];
    const bool nb = (nbm >> T) & 1u;

    const float b = row_bcast_f<T>(a1c + uv[c]);
    const float aa = row_bcast_f<T>(diag);
    const float now = row_bcast_f<T>(cur);
    const float d = palk - now;
    const float de = d * fmaf(aa, d, 2.0f * b);
    const unsigned int kk = (unsigned int)lane & 7u;
    unsigned int key = de < 0.0f ? ((__float_as_uint(de) & ~7u) | (7u - kk)) : 0u;
    key = oct_max(key);
    const bool win = key != 0u;
    const int wk = 7 - (int)(key & 7u);
    // The winner's palette value from the row's lane wk (lane j of the row holds selector j & 7's value).
    const float pw = __int_as_float(__builtin_amdgcn_ds_bpermute((rowbase + wk) << 2, __float_as_int(palk)));
    const float move = win ? pw - now : 0.0f;
    const float mv = nb ? move : 0.0f;
    taken += win ? __uint_as_float(key & ~7u) : 0.0f;
#pragma unroll
    for (int ii = 0; ii < C0; ii++)
        uv[ii] = fmaf(e[ii], mv, uv[ii]);
    const bool mine = win && u == T;
    selu = mine ? wk : selu;
    cur = mine ? pw : cur;
    // Each step's updates made here, not deferred to the channel's end with every step's key held until then.
    asm volatile("" : "+v"(selu), "+v"(taken), "+v"(cur));
    if constexpr (T < 15)
        fast_selector_steps<C0, T + 1>(s_F, lane, rowbase, u, nbm, c, a1c, uv, diag, palk, selu, cur, taken);
}

// Channels C..C0-1 of one Gauss-Seidel pass (a template over the channel: the channel indexes the lane's arrays, which a
// loop the compiler leaves rolled would put in scratch).
template <int C0, int C>
__device__ __forceinline__ void fast_selector_pass(const float* s_F, int lane, int rowbase, int u, unsigned int nbm,
                                                   const float (&a1)[C0], float (&uv)[C0], const float (&diag)[C0],
                                                   const float (&palk)[C0], int (&selu)[C0], float (&cur)[C0],
                                                   float& taken)
{
    fast_selector_steps<C0, 0>(s_F, lane, rowbase, u, nbm, C, a1[C], uv, diag[C], palk[C], selu[C], cur[C], taken);
    if constexpr (C + 1 < C0)
        fast_selector_pass<C0, C + 1>(s_F, lane, rowbase, u, nbm, a1, uv, diag, palk, selu, cur, taken);
}

template <int C0>
__global__ __launch_bounds__(64) void k_bc_refine_fast(BcArgs a, Ranges colour, int colour_index)
{
    const long long gi = (long long)xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 4);
    if (gi >= colour.start[colour.n])
        return;
    // Lane l's forward blocks: entry (kf, j, i) = a2(u, i; u + d(kf), j) at [((kf C0 + j) 64 + l) C0 + i].
    __shared__ float s_F[STENCIL_BLOCKS * C0 * 64 * C0];
    const int lane = (int)threadIdx.x, rowbase = lane & ~15;

    const int p = range_plane(colour, gi);
    const PlaneGeom& g = a.planes.p[p];
    const long long ci = gi - colour.start[p];
    const int bxn = (g.w0 + 3) / 4;
    const int cx = colour_index & 1, cy = colour_index >> 1;
    const int cbx = (bxn - cx + 1) / 2;
    const int gx = cx + 2 * (int)(ci % cbx);
    const int gy = cy + 2 * (int)(ci / cbx);
    const int w0 = g.w0, h0 = g.h0;
    const long long base = g.off0;
    const int px0 = gx * 4, py0 = gy * 4;
    const float* S = a.stencil32;
    const long long sn = a.sm.sn;

    BlockLane<C0, float> L;
    L.S = S;
    L.sn = sn;
    L.bp.init(a.sm.cpos[p], w0, gx, gy);
    const int u = lane & 15, ux = u & 3, uy = u >> 2;
    L.u = u;
    L.ux = ux;
    L.uy = uy;
    {
        const int nx = min(4, w0 - px0), ny = min(4, h0 - py0);
        const unsigned int row = (1u << nx) - 1u;
        L.inside = (row | row << 4 | row << 8 | row << 12) & ((1u << (4 * ny)) - 1u);
    }
    const bool in_u = L.in(u);
    L.pu = in_u ? L.bp.at(ux, uy) : L.bp.at(0, 0);
    L.init_neighbours();
    auto inside_xy = [=](int lx, int ly) {
        return px0 + lx >= 0 && px0 + lx < w0 && py0 + ly >= 0 && py0 + ly < h0;
    };
    auto tix = [=](int lx, int ly) { return base + (long long)(py0 + ly) * w0 + (px0 + lx); };
    const long long tu = tix(ux, uy);

    // ---- slot by slot: the lane's forward block (raw, its entries toward texels outside the block feed the linear
    // term a1), the backward term of the neighbour behind it, and the block masked to the lane's column (zero where
    // the pair is not inside the block and the plane)
    // A neighbour that adds nothing reads a texel the plane holds (the lane's own, or the block's first for a lane
    // outside the plane) and weighs it by zero.
    const long long tsafe = in_u ? tu : tix(0, 0);
    float a1[C0];
#pragma unroll
    for (int i = 0; i < C0; i++)
        a1[i] = in_u ? (float)a.grad[tu * C0 + i] : 0.0f;
#pragma unroll
    for (int slot = 0; slot < STENCIL_BLOCKS; slot++)
    {
        // The geometry made opaque per slot: the slot's loads are issued here, not all five slots' at once.
        L.opaque();
        float Fs[C0][C0];
#pragma unroll
        for (int i = 0; i < C0; i++)
#pragma unroll
            for (int j = 0; j < C0; j++)
                Fs[i][j] = S[(long long)(slot * C0 * C0 + i * C0 + j) * sn + L.pu];
        if (slot > 0)
        {
            int dx, dy;
            stencil_offset(slot, dx, dy);
            const int fx = ux + dx, fy = uy + dy;
            const bool fok = in_u && !(fx >= 0 && fx < 4 && fy >= 0 && fy < 4) && inside_xy(fx, fy);
            {
                const double* xn = a.delta + (fok ? tix(fx, fy) : tsafe) * C0;
                const float keep = fok ? 1.0f : 0.0f;
#pragma unroll
                for (int j = 0; j < C0; j++)
                {
                    const float xv = (float)xn[j] * keep;
#pragma unroll
                    for (int i = 0; i < C0; i++)
                        a1[i] = fmaf(Fs[i][j], xv, a1[i]);
                }
            }
            const int bx = ux - dx, by = uy - dy;
            const bool bok = in_u && !(bx >= 0 && bx < 4 && by >= 0 && by < 4) && inside_xy(bx, by);
            {
                const double* xn = a.delta + (bok ? tix(bx, by) : tsafe) * C0;
                const long long pb = bok ? L.bp.at(bx, by) : L.pu;
                const float keep = bok ? 1.0f : 0.0f;
#pragma unroll
                for (int j = 0; j < C0; j++)
                {
                    const float xv = (float)xn[j] * keep;
#pragma unroll
                    for (int i = 0; i < C0; i++)
                        a1[i] = fmaf(S[(long long)(slot * C0 * C0 + j * C0 + i) * sn + pb], xv, a1[i]);
                }
            }
        }
        const bool held = (L.nbmask >> (slot + 4)) & 1u;
#pragma unroll
        for (int i = 0; i < C0; i++)
#pragma unroll
            for (int j = 0; j < C0; j++)
                s_F[((slot * C0 + j) * 64 + lane) * C0 + i] = held ? Fs[i][j] : 0.0f;
    }
    // The columns are read across lanes from here on.
    __syncthreads();

    // ---- uv = a2 dv: the forward blocks' terms with the neighbours' corrections pulled from their lanes, and the
    // backward ones as the neighbours' own forward products pulled from theirs
    float dv[C0], uv[C0];
#pragma unroll
    for (int j = 0; j < C0; j++)
    {
        dv[j] = in_u ? a.v0[tu * C0 + j] - a.prev[tu * C0 + j] : 0.0f;
        uv[j] = 0.0f;
    }
    const float* own = s_F + lane * C0;   // entry (kf, j, i) of the lane's column at own[(kf C0 + j) 64 C0 + i]
#pragma unroll
    for (int k = 4; k < 9; k++)
#pragma unroll
        for (int j = 0; j < C0; j++)
        {
            const float x = from_dir(dv[j], k);
#pragma unroll
            for (int i = 0; i < C0; i++)
                uv[i] = fmaf(own[((k - 4) * C0 + j) * 64 * C0 + i], x, uv[i]);
        }
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int i = 0; i < C0; i++)
        {
            float w = 0.0f;
#pragma unroll
            for (int j = 0; j < C0; j++)
                w = fmaf(own[((4 - k) * C0 + i) * 64 * C0 + j], dv[j], w);
            uv[i] += from_dir(w, k);
        }

    const long long gb = g.offb + (long long)gy * bxn + gx;
    int ep0[C0], ep1[C0], selu[C0];
#pragma unroll
    for (int c = 0; c < C0; c++)
    {
        ep0[c] = a.ep[(gb * C0 + c) * 2];
        ep1[c] = a.ep[(gb * C0 + c) * 2 + 1];
        selu[c] = a.sel[(gb * C0 + c) * 16 + u] & 7;
    }
    float taken = 0.0f;

    // ---- (i) the endpoints of each channel
#pragma unroll
    for (int c = 0; c < C0; c++)
    {
        const float lo = a.lo0[c], hi = a.hi0[c];
        const float stp = (hi - lo) / 255.0f;
        // a2(u, i; neighbour k, c) for the nine neighbours: the lane's own column ahead of it, the neighbour's behind
        float E[9][C0];
#pragma unroll
        for (int k = 0; k < 9; k++)
        {
            const int v = rowbase + ((u + nb_off(k)) & 15);
#pragma unroll
            for (int i = 0; i < C0; i++)
                E[k][i] = k >= 4 ? s_F[(((k - 4) * C0 + c) * 64 + lane) * C0 + i]
                                 : s_F[(((4 - k) * C0 + i) * 64 + v) * C0 + c];
        }
        const float aw = bc4_weight8f(selu[c]);
        float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
        for (int k = 0; k < 9; k++)
        {
            s1 += E[k][c];
            s2 = fmaf(E[k][c], from_dir(aw, k), s2);
        }
        const float pp = stp * aw, qq = stp * (1.0f - aw);
        const float ap = stp * s2, aq = stp * (s1 - s2);
        const float b = a1[c] + uv[c];
        const float bp_ = row_sum_f(b * pp), bq = row_sum_f(b * qq);
        const float pap = row_sum_f(pp * ap), paq = row_sum_f(pp * aq), qaq = row_sum_f(qq * aq);
        const int r0 = ep0[c], r1 = ep1[c];
        const double det = (double)pap * qaq - (double)paq * paq;
        double d0 = 0.0, d1 = 0.0;
        if (det > 1e-6 * (double)pap * qaq)
        {
            d0 = ((double)paq * bq - (double)qaq * bp_) / det;
            d1 = ((double)paq * bp_ - (double)pap * bq) / det;
        }
        const int c0r = (int)floor((double)r0 + d0 + 0.5), c1r = (int)floor((double)r1 + d1 + 0.5);
        float best = 0.0f;
        int best0 = r0, best1 = r1;
#pragma unroll
        for (int e0 = -1; e0 <= 1; e0++)
#pragma unroll
            for (int e1 = -1; e1 <= 1; e1++)
            {
                const int x0 = min(255, max(0, c0r + e0)), x1 = min(255, max(0, c1r + e1));
                if (x0 <= x1)
                    continue;
                const float f0 = (float)(x0 - r0), f1 = (float)(x1 - r1);
                const float de = 2.0f * (bp_ * f0 + bq * f1) + pap * f0 * f0 + 2.0f * paq * f0 * f1 + qaq * f1 * f1;
                if (de < best)
                {
                    best = de;
                    best0 = x0;
                    best1 = x1;
                }
            }
        if (best0 == r0 && best1 == r1)
            continue;
        const float dlt = level0_value(lo, hi, best0, best1, selu[c]) - level0_value(lo, hi, r0, r1, selu[c]);
        float dn[9];
#pragma unroll
        for (int k = 0; k < 9; k++)
            dn[k] = from_dir(dlt, k);
        float quad = 0.0f;
#pragma unroll
        for (int k = 0; k < 9; k++)
            quad = fmaf(E[k][c], dn[k], quad);
        const float exact = row_sum_f(dlt * (2.0f * (a1[c] + uv[c]) + quad));
        if (!(exact < 0.0f))
            continue;
        taken += exact;
#pragma unroll
        for (int i = 0; i < C0; i++)
        {
            float add = 0.0f;
#pragma unroll
            for (int k = 0; k < 9; k++)
                add = fmaf(E[k][i], dn[k], add);
            uv[i] += add;
        }
        ep0[c] = best0;
        ep1[c] = best1;
    }

    // ---- (ii) the selectors, two Gauss-Seidel passes of 8-way argmins
    // nbm bit t: texel t of the block is the lane's texel or one of its eight neighbours.
    unsigned int nbm = 0;
#pragma unroll
    for (int t = 0; t < 16; t++)
    {
        const int ddx = (t & 3) - ux, ddy = (t >> 2) - uy;
        nbm |= (ddx >= -1 && ddx <= 1 && ddy >= -1 && ddy <= 1 ? 1u : 0u) << t;
    }
    float pal[C0][8], cur[C0], diag[C0], palk[C0];
#pragma unroll
    for (int c = 0; c < C0; c++)
    {
#pragma unroll
        for (int k = 0; k < 8; k++)
            pal[c][k] = level0_value(a.lo0[c], a.hi0[c], ep0[c], ep1[c], k);
        cur[c] = pal[c][0];
        palk[c] = pal[c][0];
#pragma unroll
        for (int k = 1; k < 8; k++)
        {
            cur[c] = selu[c] == k ? pal[c][k] : cur[c];
            palk[c] = (lane & 7) == k ? pal[c][k] : palk[c];
        }
        diag[c] = s_F[(c * 64 + lane) * C0 + c];
    }
#pragma unroll 1
    for (int pass = 0; pass < 2; pass++)
        fast_selector_pass<C0, 0>(s_F, lane, rowbase, u, nbm, a1, uv, diag, palk, selu, cur, taken);

    // ---- what the block leaves behind
#pragma unroll
    for (int c = 0; c < C0; c++)
    {
        if (u == 0)
        {
            a.ep[(gb * C0 + c) * 2] = (uint8_t)ep0[c];
            a.ep[(gb * C0 + c) * 2 + 1] = (uint8_t)ep1[c];
        }
        a.sel[(gb * C0 + c) * 16 + u] = (uint8_t)selu[c];
    }
    if (in_u)
    {
#pragma unroll
        for (int c = 0; c < C0; c++)
        {
            const float byte = bc4_entry8(ep0[c], ep1[c], selu[c]);
            const long long idx = tu * C0 + c;
            a.v0[idx] = cur[c];
            a.k0[idx] = bc0_index(byte);
            a.delta[idx] = (double)cur[c] - (double)a.prev[idx];
        }
    }
    if (u == 0)
        a.taken[gb] = (double)taken;
}

// The .dds blocks of every file of every plane (bc4_pack, shared/bc_pack.h:51-60), and each block's squared error in
// byte units against the mirror's exact index values over the texels inside the plane
// (src/export.cpp:283-292).
__global__ __launch_bounds__(256) void k_bc_pack(BcArgs a, long long total)
{
    const long long gb = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gb >= total)
        return;
    int p = 0;
    while (p + 1 < a.planes.n && gb >= a.planes.p[p + 1].offb)
        p++;
    const PlaneGeom& g = a.planes.p[p];
    const long long b = gb - g.offb;
    const int bx = (g.w0 + 3) / 4;
    const int x = (int)(b % bx), y = (int)(b / bx);
    const int files = a.c0 <= 2 ? 1 : 2;
    double se = 0.0;
    for (int f = 0; f < files; f++)
    {
        const int nc = f == 0 ? (a.c0 == 1 ? 1 : 2) : a.c0 - 2;
        const int bpb = nc == 1 ? 8 : 16;
        uint8_t* file = a.blocks + g.offb * 32 + (long long)f * g.nb * 16;
        for (int cc = 0; cc < nc; cc++)
        {
            const int channel = 2 * f + cc;
            const long long slot = gb * a.c0 + channel;
            const int r0 = a.ep[slot * 2], r1 = a.ep[slot * 2 + 1];
            unsigned long long bits = 0;
            for (int t = 0; t < 16; t++)
                bits |= (unsigned long long)(a.sel[slot * 16 + t] & 7) << (3 * t);
            uint8_t* out = file + b * bpb + cc * 8;
            out[0] = (uint8_t)r0;
            out[1] = (uint8_t)r1;
            for (int i = 0; i < 6; i++)
                out[2 + i] = (uint8_t)(bits >> (8 * i));
            for (int t = 0; t < 16; t++)
            {
                const int px = x * 4 + t % 4, py = y * 4 + t / 4;
                if (px >= g.w0 || py >= g.h0)
                    continue;
                const float dec = bc4_entry(r0, r1, a.sel[slot * 16 + t] & 7);
                const int k = a.k0m[(g.off0 + (long long)py * g.w0 + px) * a.c0 + channel];
                const double e = (double)dec - (double)k * 255.0 / 255.0;
                se += e * e;
            }
        }
    }
    a.se[gb] = se;
}

// One pass's figures per plane, folded in the reference's order (src/cpu/cpu_refine_bc.cpp:398-447): per colour, the
// colour's blocks in chunks of 64, each chunk's accepted decreases (the negative ones) and their count summed in block
// order, the chunk sums added in chunk order; then the four colours added in order. One workgroup per plane, a group of
// 256 lanes per colour: lane t sums chunk t of each 256, the group's lane 0 adds those in order. Skipping a block that
// took nothing and adding +0.0 for it give the same sums, since no partial sum is ever -0.0.
__global__ __launch_bounds__(1024) void k_bc_fold(Planes planes, const double* taken, double* out)
{
    __shared__ double s_t[4][256], s_m[4][256], s_col[4][2];
    const int p = blockIdx.x, col = threadIdx.x >> 8, t = threadIdx.x & 255;
    const PlaneGeom& g = planes.p[p];
    const int bx = (g.w0 + 3) / 4, by = (g.h0 + 3) / 4;
    const int cx = col & 1, cy = col >> 1;
    const int cbx = (bx - cx + 1) / 2;
    const int count = cbx * ((by - cy + 1) / 2);
    const int chunks = (count + 63) / 64;
    const int chunks_max = (((bx + 1) / 2) * ((by + 1) / 2) + 63) / 64;   // colour 0 has the most blocks
    const double* const tp = taken + g.offb;
    double ct = 0.0, cm = 0.0;
    for (int base = 0; base < chunks_max; base += 256)
    {
        const int ch = base + t;
        if (ch < chunks)
        {
            double pt = 0.0, pm = 0.0;
            const int b0 = ch * 64, b1 = min(b0 + 64, count);
            int bxi = b0 % cbx, byi = b0 / cbx;
            // Sixteen blocks' values loaded ahead of their (ordered) additions.
            for (int b = b0; b < b1; b += 16)
            {
                double v[16];
#pragma unroll
                for (int u = 0; u < 16; u++)
                {
                    v[u] = b + u < b1 ? tp[(long long)(cy + 2 * byi) * bx + cx + 2 * bxi] : 0.0;
                    if (++bxi == cbx)
                    {
                        bxi = 0;
                        byi++;
                    }
                }
#pragma unroll
                for (int u = 0; u < 16; u++)
                {
                    const bool neg = v[u] < 0.0;
                    pt += neg ? v[u] : 0.0;
                    pm += neg ? 1.0 : 0.0;
                }
            }
            s_t[col][t] = pt;
            s_m[col][t] = pm;
        }
        __syncthreads();
        if (t == 0)
        {
            const int n = min(chunks - base, 256);
            for (int j = 0; j < n; j++)
            {
                ct += s_t[col][j];
                cm += s_m[col][j];
            }
        }
        __syncthreads();
    }
    if (t == 0)
    {
        s_col[col][0] = ct;
        s_col[col][1] = cm;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double tk = 0.0, im = 0.0;
        for (int q = 0; q < 4; q++)
        {
            tk += s_col[q][0];
            im += s_col[q][1];
        }
        out[2 * p] = tk;
        out[2 * p + 1] = im;
    }
}

// The fast form's fold of the same figures: every block of the plane (the four colours together) read coalesced, lane t
// taking blocks t, t + 1024, ... into four accumulators, then a tree over the workgroup - a fixed order, not the
// reference's.
__global__ __launch_bounds__(1024) void k_bc_fold_fast(Planes planes, const double* taken, double* out)
{
    __shared__ double s_t[1024], s_m[1024];
    const int p = blockIdx.x, t = threadIdx.x;
    const PlaneGeom& g = planes.p[p];
    const int count = ((g.w0 + 3) / 4) * ((g.h0 + 3) / 4);
    const double* const tp = taken + g.offb;
    double pt[4] = { 0.0, 0.0, 0.0, 0.0 }, pm[4] = { 0.0, 0.0, 0.0, 0.0 };
    int b = t;
    // Sixteen loads in flight; accumulator u % 4 takes the same values in the same order as the loop below.
    for (; b + 15 * 1024 < count; b += 16 * 1024)
    {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; u++)
            v[u] = tp[b + u * 1024];
#pragma unroll
        for (int u = 0; u < 16; u++)
        {
            const bool neg = v[u] < 0.0;
            pt[u & 3] += neg ? v[u] : 0.0;
            pm[u & 3] += neg ? 1.0 : 0.0;
        }
    }
    for (; b < count; b += 4 * 1024)
    {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
            v[u] = b + u * 1024 < count ? tp[b + u * 1024] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            const bool neg = v[u] < 0.0;
            pt[u] += neg ? v[u] : 0.0;
            pm[u] += neg ? 1.0 : 0.0;
        }
    }
    s_t[t] = (pt[0] + pt[1]) + (pt[2] + pt[3]);
    s_m[t] = (pm[0] + pm[1]) + (pm[2] + pm[3]);
    __syncthreads();
    for (int h = 512; h > 0; h >>= 1)
    {
        if (t < h)
        {
            s_t[t] += s_t[t + h];
            s_m[t] += s_m[t + h];
        }
        __syncthreads();
    }
    if (t == 0)
    {
        out[2 * p] = s_t[0];
        out[2 * p + 1] = s_m[0];
    }
}

// The packing error's total over every block of the chain, in this kernel's fixed order (a report figure only: the
// reference adds the same squared errors texel by texel, so the last bits may differ).
__global__ __launch_bounds__(1024) void k_bc_se_total(const double* se, long long n, double* out)
{
    __shared__ double s[1024];
    // Lane t adds elements t, t + 1024, ... in that order, eight loads in flight.
    double acc = 0.0;
    long long i = threadIdx.x;
    for (; i + 7 * 1024 < n; i += 8 * 1024)
    {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            v[u] = se[i + u * 1024];
#pragma unroll
        for (int u = 0; u < 8; u++)
            acc += v[u];
    }
    for (; i < n; i += 1024)
        acc += se[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1)
    {
        if ((int)threadIdx.x < w)
            s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        *out = s[0];
}

// The exact form reads the fp64 stencil in fp64; the plain form the fp32 copy its gathers build (the only one they
// build) in fp64 arithmetic and the exact form's order; the fast form the fp32 copy in fp32 (k_bc_refine_fast).
template <int C0>
void launch_refine(int form, dim3 grid, hipStream_t s, const BcArgs& a, const Ranges& colour, int col)
{
    if (form == NNTC_FORM_FAST)
        hipLaunchKernelGGL((k_bc_refine_fast<C0>), grid, dim3(64), 0, s, a, colour, col);
    else if (form == NNTC_FORM_PLAIN)
        hipLaunchKernelGGL((k_bc_refine<C0, float>), grid, dim3(64), 0, s, a, colour, col);
    else
        hipLaunchKernelGGL((k_bc_refine<C0, double>), grid, dim3(64), 0, s, a, colour, col);
}

BcArgs bc_args(Ctx& c)
{
    BcArgs a;
    a.planes = c.planes;
    a.c0 = c.c0;
    a.k0m = c.k0_mirror.p;
    a.ep = c.bc_ep.p;
    a.sel = c.bc_sel.p;
    a.v0 = c.v0.p;
    a.k0 = c.k0.p;
    a.delta = c.lat0.x.p;
    a.prev = c.lat0.prev.p;
    a.stencil = c.lat0.stencil.p;
    a.stencil32 = c.lat0.stencil32.p;
    a.sm = stencil_map(c, true);
    a.grad = c.lat0.grad.p;
    a.lo0 = c.grid.p;
    a.hi0 = c.grid.p + 4;
    a.taken = c.bc_taken.p;
    a.blocks = c.bc_blocks.p;
    a.se = c.bc_se.p;
    return a;
}

}   // namespace

void bc_seed(Ctx& c)
{
    BcArgs a = bc_args(c);
    hipLaunchKernelGGL(k_bc_seed, dim3(blocks_for(c.nb_total * c.c0, 256)), dim3(256), 0, c.stream, a, c.nb_total);
    HIP_OK(hipGetLastError());
}

void bc_decode(Ctx& c)
{
    BcArgs a = bc_args(c);
    hipLaunchKernelGGL(k_bc_decode, dim3(blocks_for(c.n0_total, 256)), dim3(256), 0, c.stream, a, c.n0_total);
    HIP_OK(hipGetLastError());
}

// One pass: the four colours over every plane's block grid, then the pass figures folded per plane on the device into
// fold[2p] (the decrease taken) and fold[2p + 1] (the blocks improved).
void bc_refine_pass(Ctx& c, double* fold)
{
    BcArgs a = bc_args(c);
    Ranges colour[4];
    for (int col = 0; col < 4; col++)
    {
        const int cx = col & 1, cy = col >> 1;
        colour[col].n = c.planes.n;
        long long acc = 0;
        for (int p = 0; p < c.planes.n; p++)
        {
            colour[col].start[p] = acc;
            const int bx = (c.g(p).w0 + 3) / 4, by = (c.g(p).h0 + 3) / 4;
            acc += (long long)((bx - cx + 1) / 2) * ((by - cy + 1) / 2);
        }
        colour[col].start[c.planes.n] = acc;
    }
    const bool fast = c.form != NNTC_FORM_EXACT;
    for (int col = 0; col < 4; col++)
    {
        const long long n = colour[col].start[colour[col].n];
        if (n == 0)
            continue;
        // Each block reads its sixteen texels' five stencil blocks (8-byte entries, 4-byte in the fast form),
        // gradients and corrections (8 bytes a channel), and planes and previous values (4 bytes a channel), and
        // writes its planes, corrections and indices (4 + 8 + 1 bytes a channel).
        const double c0 = c.c0, sb = fast ? 4.0 : 8.0;
        KernelTimer kt(c.stream, "k_bc_refine", (double)n * 16.0 * (5.0 * sb * c0 * c0 + 8.0 * c0 + 8.0 * c0 + 9.0 * c0),
                       0.0);
        const dim3 grid(blocks_for(n, 4));   // four blocks per 64-lane workgroup
        switch (c.c0)
        {
        case 1: launch_refine<1>(c.form, grid, c.stream, a, colour[col], col); break;
        case 2: launch_refine<2>(c.form, grid, c.stream, a, colour[col], col); break;
        case 3: launch_refine<3>(c.form, grid, c.stream, a, colour[col], col); break;
        default: launch_refine<4>(c.form, grid, c.stream, a, colour[col], col); break;
        }
        HIP_OK(hipGetLastError());
    }
    if (fast)
        hipLaunchKernelGGL(k_bc_fold_fast, dim3((unsigned)c.planes.n), dim3(1024), 0, c.stream, c.planes,
                           (const double*)c.bc_taken.p, fold);
    else
        hipLaunchKernelGGL(k_bc_fold, dim3((unsigned)c.planes.n), dim3(1024), 0, c.stream, c.planes,
                           (const double*)c.bc_taken.p, fold);
    HIP_OK(hipGetLastError());
}

void bc_pack_blocks(Ctx& c, double* se_total)
{
    BcArgs a = bc_args(c);
    hipLaunchKernelGGL(k_bc_pack, dim3(blocks_for(c.nb_total, 256)), dim3(256), 0, c.stream, a, c.nb_total);
    HIP_OK(hipGetLastError());
    hipLaunchKernelGGL(k_bc_se_total, dim3(1), dim3(1024), 0, c.stream, (const double*)c.bc_se.p, c.nb_total, se_total);
    HIP_OK(hipGetLastError());
}

}   // namespace nntc
