"""The oracle's restatement of the reference encoder's arithmetic, in numpy.

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use what is under oracle/,
and only as the checker. The compiled reference itself (oracle/_ref, built by oracle/Makefile from the reference's own
sources) is the primary oracle; this file restates the pieces of its arithmetic that the tests pin independently of
it - against the reference's own known-answer vectors, closed forms, brute force and textbook linear algebra - so that
a mistake in either would show:

  replicate_index, level1_value, level1_index      src/model.h:141-199
  bc0_value, bc4_weight8                           src/refine_bc.cu:102-131, src/cpu/cpu_refine_bc.cpp:62-100
  bc4_palette, bc4_encode_block, bc4_decode_block  shared/bc_pack.h:62-151
  tap / sample (the hardware bilinear rule)        src/sample.cuh:43-72, src/cpu/cpu_sample.h:36-66
  subtexel_offset                                  src/sample.cuh:87-128
  objective_brute (E, the brute-force twin)        src/objective.cu:317-380, src/cpu/cpu_objective.cpp:298-360
  normal_equations (block (a)'s A and rhs)         src/solve_decoder.cu:87-180 (a plain sum over every site)
  dense_level1_system (block (b)'s H and grad)     src/solve_level1.cu:1502-1688 (the dense twin's scatter)

Every check of this file is in tests/test_oracle.py; nothing here is pinned by retyping it in a test.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32
f64 = np.float64


# ---------------------------------------------------------------------------------------------------------------------
# The grid (src/model.h:141-199)
# ---------------------------------------------------------------------------------------------------------------------

def replicate_index(k: int, bits: int) -> int:
    b8 = 0
    sh = 8 - bits
    while sh > -bits:
        b8 |= (k << sh) if sh >= 0 else (k >> -sh)
        sh -= bits
    return b8 & 0xFF


def level1_value(lo: float, hi: float, bits: int, k: int) -> np.float32:
    """lo + rep(k)/255 (hi - lo), in double with the product and the sum rounded apart, rounded to float once."""
    span = f64(f32(hi)) - f64(f32(lo))
    frac = f64(replicate_index(k, bits)) / f64(255.0)
    return f32(f64(f32(lo)) + frac * span)


def level1_index(lo: float, hi: float, bits: int, value: float) -> int:
    """The index of the nearest grid value; a tie keeps the lower index."""
    lo, hi, value = f32(lo), f32(hi), f32(value)
    span = f64(hi) - f64(lo)
    levels = (1 << bits) - 1
    if not span > 0.0:
        return 0
    x = (f64(value) - f64(lo)) / span * f64(levels)
    x = min(max(x, 0.0), float(levels))
    k = int(x + 0.5)
    k = min(max(k, 0), levels)
    best, best_d = k, abs(f64(value) - f64(level1_value(lo, hi, bits, k)))
    for n in (k - 1, k + 1):
        if 0 <= n <= levels:
            d = abs(f64(value) - f64(level1_value(lo, hi, bits, n)))
            if d < best_d or (d == best_d and n < best):
                best, best_d = n, d
    return best


# ---------------------------------------------------------------------------------------------------------------------
# BC4 (shared/bc_pack.h, src/refine_bc.cu:102-131)
# ---------------------------------------------------------------------------------------------------------------------

def bc0_value(lo: float, hi: float, byte: float) -> np.float32:
    frac = f64(f32(byte)) / f64(255.0)
    span = f64(f32(hi)) - f64(f32(lo))
    return f32(f64(f32(lo)) + frac * span)


def bc4_weight8(sel: int) -> float:
    return 1.0 if sel == 0 else (0.0 if sel == 1 else (8 - sel) / 7.0)


def bc4_palette(r0: int, r1: int) -> np.ndarray:
    p = np.zeros(8, dtype=np.float32)
    p[0], p[1] = r0, r1
    if r0 > r1:
        for i in range(2, 8):
            p[i] = f32(f32((8 - i) * r0 + (i - 1) * r1) / f32(7.0))
    else:
        for i in range(2, 6):
            p[i] = f32(f32((6 - i) * r0 + (i - 1) * r1) / f32(5.0))
        p[6], p[7] = 0.0, 255.0
    return p


def bc4_encode_block(v: np.ndarray, bits: int) -> tuple[int, int, list[int], float]:
    """(r0, r1, selectors, squared error) of the stb_dxt-style pack of sixteen bytes holding `bits`-bit indices."""
    v = [int(x) for x in v]
    if bits <= 3:
        if bits == 3:
            sel = [0 if (x >> 5) == 7 else (1 if (x >> 5) == 0 else 8 - (x >> 5)) for x in v]
            return 255, 0, sel, 0.0
        if bits == 2:
            m = {0: 6, 1: 0, 2: 1, 3: 7}
            return 85, 170, [m[x >> 6] for x in v], 0.0
        return 0, 255, [7 if (x >> 7) else 6 for x in v], 0.0
    levels = (1 << bits) - 1
    ks = [x >> (8 - bits) for x in v]
    r0 = int(np.floor(max(ks) * 255.0 / levels + 0.5))
    r1 = int(np.floor(min(ks) * 255.0 / levels + 0.5))
    if r0 == r1:
        return r0, r1, [0] * 16, 0.0
    p = bc4_palette(r0, r1)
    sel, err = [], 0.0
    for k in ks:
        target = f32(f32(f32(k) * f32(255.0)) / f32(levels))
        d = np.abs(p - target)
        best = int(np.argmin(d))   # the first minimum, as the reference's strict < keeps it
        sel.append(best)
        err += float(d[best]) ** 2
    return r0, r1, sel, err


def bc4_pack(r0: int, r1: int, sel: list[int]) -> bytes:
    bits = 0
    for t, s in enumerate(sel):
        bits |= (s & 7) << (3 * t)
    return bytes([r0, r1] + [(bits >> (8 * i)) & 0xFF for i in range(6)])


def bc4_decode_block(blk: bytes) -> np.ndarray:
    p = bc4_palette(blk[0], blk[1])
    bits = int.from_bytes(bytes(blk[2:8]), "little")
    return np.array([p[(bits >> (3 * t)) & 7] for t in range(16)], dtype=np.float32)


# ---------------------------------------------------------------------------------------------------------------------
# Sampling (src/sample.cuh:43-128)
# ---------------------------------------------------------------------------------------------------------------------

def subtexel_offset(k_count: int, j: int, tx: int, ty: int) -> tuple[float, float]:
    if k_count == 2:
        return (-0.25, 0.25) if j == 0 else (0.25, -0.25)
    ox, oy = 0.25, 0.25
    if k_count == 4:
        ox = [-0.375, -0.125, 0.125, 0.375][j]
        oy = [-0.125, 0.375, -0.375, 0.125][j]
    for _ in range((ty & 1) * 2 + (tx & 1)):
        ox, oy = -oy, ox
    return ox, oy


def twin_sample(plane: np.ndarray, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """The objective's double bilinear rule (src/objective.cu:104-118) at normalised (u, v); plane is (h, w, c)."""
    h, w = plane.shape[:2]
    x = u * w - 0.5
    y = v * h - 0.5
    xf, yf = np.floor(x), np.floor(y)
    fx, fy = x - xf, y - yf
    x0 = np.clip(xf.astype(np.int64), 0, w - 1)
    x1 = np.clip(xf.astype(np.int64) + 1, 0, w - 1)
    y0 = np.clip(yf.astype(np.int64), 0, h - 1)
    y1 = np.clip(yf.astype(np.int64) + 1, 0, h - 1)
    p = plane.astype(np.float64)
    return ((1 - fx) * (1 - fy))[:, None] * p[y0, x0] + (fx * (1 - fy))[:, None] * p[y0, x1] + \
        ((1 - fx) * fy)[:, None] * p[y1, x0] + (fx * fy)[:, None] * p[y1, x1]


def sites(w0: int, h0: int, k: int):
    """Every site of a w0 x h0 plane: pixel coordinates, site index j (0 = centre) and normalised (u, v)."""
    py, px = np.mgrid[0:h0, 0:w0]
    px, py = px.reshape(-1), py.reshape(-1)
    out = []
    for j in range(k + 1):
        if j == 0:
            dx = np.zeros(px.size)
            dy = np.zeros(px.size)
        else:
            off = np.array([subtexel_offset(k, j - 1, int(a), int(b)) for a, b in zip(px, py)])
            dx, dy = off[:, 0], off[:, 1]
        dx32 = dx.astype(np.float32).astype(np.float64)
        dy32 = dy.astype(np.float32).astype(np.float64)
        u = (px + 0.5 + dx32) / w0
        v = (py + 0.5 + dy32) / h0
        out.append((px, py, j, u, v))
    return out


def features(s: np.ndarray, c: np.ndarray) -> np.ndarray:
    """phi = [c_j][s_i][s_i c_j] (src/sample.cuh:74-85), rows of sites."""
    c0, c1 = s.shape[1], c.shape[1]
    prod = (s[:, :, None] * c[:, None, :]).reshape(s.shape[0], c0 * c1)
    return np.concatenate([c, s, prod], axis=1)


def objective_brute(planes, v0s, v1s, srcs, w, b, cw, k, per_site):
    """E and the per-plane sums by the brute-force twin: every site of every plane in double."""
    acc = []
    for (w0, h0, w1, h1), v0, v1, src in zip(planes, v0s, v1s, srcs):
        sums = np.zeros(3)
        for px, py, j, u, v in sites(w0, h0, k):
            c = twin_sample(v1.reshape(h1, w1, -1), u, v)
            if j == 0:
                idx = py * w0 + px
                s = v0.reshape(-1, v0.size // (w0 * h0))[idx].astype(np.float64)
                t = src.reshape(-1, src.size // (w0 * h0))[idx].astype(np.float64)
            else:
                s = twin_sample(v0.reshape(h0, w0, -1), u, v)
                t = twin_sample(src.reshape(h0, w0, -1), u, v)
            y = features(s, c) @ np.asarray(w, np.float64).T + np.asarray(b, np.float64)
            e = y - t
            sums[0] += float(np.sum(np.asarray(cw, np.float64) * e * e))
            sums[1 if j == 0 else 2] += float(np.sum(e * e))
        acc.append(sums)
    sum_cw = float(np.sum(np.asarray(cw, np.float64)))
    num = sum(ps * a[0] for ps, a in zip(per_site, acc))
    den = sum(ps * pl[0] * pl[1] * (1 + k) for ps, pl in zip(per_site, planes))
    return num / (sum_cw * den), acc


def tap(px, py, dx, dy, w0, h0, w, h):
    """The fp32 bilinear addressing from pixel coordinates (src/cpu/cpu_sample.h:36-51), vectorised."""
    px, py = px.astype(np.float32), py.astype(np.float32)
    dx, dy = dx.astype(np.float32), dy.astype(np.float32)
    if w == w0:
        x = px + dx
    else:
        x = ((px + f32(0.5) + dx) * f32(w)) / f32(w0) - f32(0.5)
    if h == h0:
        y = py + dy
    else:
        y = ((py + f32(0.5) + dy) * f32(h)) / f32(h0) - f32(0.5)
    x0 = np.floor(x).astype(np.int64)
    y0 = np.floor(y).astype(np.int64)
    fx = (x - x0.astype(np.float32)).astype(np.float32)
    fy = (y - y0.astype(np.float32)).astype(np.float32)
    return (np.clip(x0, 0, w - 1), np.clip(x0 + 1, 0, w - 1), np.clip(y0, 0, h - 1), np.clip(y0 + 1, 0, h - 1),
            fx, fy)


def sample32(plane, w, t):
    """The fp32 four-tap blend w00 a + w10 b + w01 e + w11 f (src/cpu/cpu_sample.h:54-66); plane is (texels, c)."""
    x0, x1, y0, y1, fx, fy = t
    p = plane
    one = f32(1.0)
    w00 = ((one - fx) * (one - fy))[:, None]
    w10 = (fx * (one - fy))[:, None]
    w01 = ((one - fx) * fy)[:, None]
    w11 = (fx * fy)[:, None]
    a, b_, e, f = p[y0 * w + x0], p[y0 * w + x1], p[y1 * w + x0], p[y1 * w + x1]
    return (((w00 * a + w10 * b_) + w01 * e) + w11 * f).astype(np.float32)


def site_rows(plane, v0, v1, src, c0, c1, nout, k):
    """Every site's fp32 features and target (src/cpu/cpu_sample.h:115-139)."""
    w0, h0, w1, h1 = plane
    v0 = v0.reshape(-1, c0)
    v1 = v1.reshape(-1, c1)
    src = src.reshape(-1, nout)
    phis, tgts = [], []
    py, px = np.mgrid[0:h0, 0:w0]
    px, py = px.reshape(-1), py.reshape(-1)
    for j in range(k + 1):
        if j == 0:
            dx = np.zeros(px.size, np.float32)
            dy = np.zeros(px.size, np.float32)
        else:
            off = np.array([subtexel_offset(k, j - 1, int(a), int(b)) for a, b in zip(px, py)], np.float32)
            dx, dy = off[:, 0], off[:, 1]
        t1 = tap(px, py, dx, dy, w0, h0, w1, h1)
        c = sample32(v1, w1, t1)
        if j == 0:
            idx = py * w0 + px
            s, t = v0[idx], src[idx]
        else:
            t0 = tap(px, py, dx, dy, w0, h0, w0, h0)
            s, t = sample32(v0, w0, t0), sample32(src, w0, t0)
        phis.append(features(s.astype(np.float64), c.astype(np.float64)).astype(np.float32))
        tgts.append(t)
    return np.concatenate(phis), np.concatenate(tgts)


def normal_equations(planes, v0s, v1s, srcs, c0, c1, nout, k, per_site):
    """A = sum omega [phi;1][phi;1]^T and rhs = sum omega t [phi;1]^T over every site of every plane, in double."""
    mm = c1 + c0 + c0 * c1 + 1
    a = np.zeros((mm, mm))
    rhs = np.zeros((nout, mm))
    for plane, v0, v1, src, om in zip(planes, v0s, v1s, srcs, per_site):
        phi, t = site_rows(plane, v0, v1, src, c0, c1, nout, k)
        v = np.concatenate([phi.astype(np.float64), np.ones((phi.shape[0], 1))], axis=1)
        a += om * (v.T @ v)
        rhs += om * (t.astype(np.float64).T @ v)
    return a, rhs


def dense_level1_system(plane, v0, v1, src, w, b, cw, c0, c1, nout, k):
    """Block (b)'s exact quadratic in the level-1 plane, by scatter: H (n x n) and grad (n), n = w1 h1 c1
    (src/solve_level1.cu:1502-1688: every site in double, level 1's four taps merged by texel)."""
    w0, h0, w1, h1 = plane
    n = w1 * h1 * c1
    hm = np.zeros((n, n))
    gv = np.zeros(n)
    w = np.asarray(w, np.float64)
    b = np.asarray(b, np.float64)
    cw = np.asarray(cw, np.float64)
    v0r = v0.reshape(-1, c0).astype(np.float64)
    v1r = v1.reshape(-1, c1).astype(np.float64)
    srcr = src.reshape(-1, nout).astype(np.float64)
    for px, py, j, u, v in sites(w0, h0, k):
        for i in range(px.size):
            gx, gy = u[i] * w1 - 0.5, v[i] * h1 - 0.5
            ix, iy = int(np.floor(gx)), int(np.floor(gy))
            ax, ay = gx - ix, gy - iy
            xs = [min(max(ix, 0), w1 - 1), min(max(ix + 1, 0), w1 - 1)]
            ys = [min(max(iy, 0), h1 - 1), min(max(iy + 1, 0), h1 - 1)]
            taps = {}
            for bb in range(2):
                for aa in range(2):
                    tid = ys[bb] * w1 + xs[aa]
                    taps[tid] = taps.get(tid, 0.0) + [1 - ax, ax][aa] * [1 - ay, ay][bb]
            cbar = sum(wt * v1r[t] for t, wt in taps.items())
            if j == 0:
                pi = py[i] * w0 + px[i]
                sbar, tgt = v0r[pi], srcr[pi]
            else:
                sbar = twin_sample(v0.reshape(h0, w0, c0), np.array([u[i]]), np.array([v[i]]))[0]
                tgt = twin_sample(src.reshape(h0, w0, nout), np.array([u[i]]), np.array([v[i]]))[0]
            g = w[:, :c1] + sum(sbar[q] * w[:, c1 + c0 + q * c1: c1 + c0 + (q + 1) * c1] for q in range(c0))
            res = b + w[:, c1:c1 + c0] @ sbar + g @ cbar - tgt
            ga = (cw[:, None] * g).T @ res
            gg = (cw[:, None] * g).T @ g
            for te, we in taps.items():
                gv[te * c1:(te + 1) * c1] += we * ga
                for tf, wf in taps.items():
                    hm[te * c1:(te + 1) * c1, tf * c1:(tf + 1) * c1] += we * wf * gg
    return hm, gv
