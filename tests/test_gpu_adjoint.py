"""The decode's hand-written adjoint and the matrix-free level-0 operator built on it.

nntc_hip_decode_sites is the decode at every sample site of one plane (out = W phi(s0, s1) + b) and
nntc_hip_decode_adjoint its adjoint: the transposed bilinear sampling of a cotangent into both latent planes and the
decoder's gradient. They are checked
  - against the arm's own decode: at the centre sites the fp32 forward rounds to decode_plane's bytes;
  - against finite differences, through a test-only torch.autograd.Function over the two entry points, by
    torch.autograd.gradcheck in fp64 (step 1e-6) on a 64x64 crop with two channels per latent (random projections)
    and on an 8x8 crop (the whole Jacobian);
  - against the assembled stencil: J^T C J p built from the fp64 adjoint (J p = decode(p) - decode(0) at every
    site, C = diag(cw)) equals the exact form's fp64 level-0 stencil applied to p (src/solve_level1.cu:440-581,
    781-828) to fp32's resolution (the assembly forms each site's rows in fp32, as the reference does);
    in fp32, the fast form's matrix-free operator equals its assembled fp32 stencil's (H + lambda I) p to rounding;
  - and in use: block (c') solved by conjugate gradients on the matrix-free operator lands on the assembled solve's
    correction.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from nntc_amd import abi, chain, synth
from schedule import Options, steps

pytestmark = pytest.mark.gpu


class SitesDesc(C.Structure):
    _fields_ = [("w0", C.c_int32), ("h0", C.c_int32), ("w1", C.c_int32), ("h1", C.c_int32),
                ("c0", C.c_int32), ("c1", C.c_int32), ("nout", C.c_int32), ("k", C.c_int32)]


def lib():
    L = C.CDLL(abi.HIP_LIB, mode=C.RTLD_GLOBAL)
    L.nntc_hip_decode_sites.argtypes = [C.POINTER(SitesDesc), C.c_int] + [C.c_void_p] * 6
    L.nntc_hip_decode_adjoint.argtypes = [C.POINTER(SitesDesc), C.c_int] + [C.c_void_p] * 9
    L.nntc_hip_last_error.restype = C.c_char_p
    return L


def decode_sites(d, v0, v1, w, b, fp64=True):
    """out (sites x nout) through the C ABI; every array uploaded to the device and back."""
    L = lib()
    dt = np.float64 if fp64 else np.float32
    ins = [abi.DeviceBytes(np.ascontiguousarray(x, dtype=dt)) for x in (v0, v1, w, b)]
    sites = d.w0 * d.h0 * (1 + d.k)
    out = abi.DeviceBytes(np.zeros((sites, d.nout), dt))
    st = L.nntc_hip_decode_sites(C.byref(d), 1 if fp64 else 0, *[x.address for x in ins], out.address, None)
    assert st == 0, L.nntc_hip_last_error()
    return out.read()


def decode_adjoint(d, v0, v1, w, g, fp64=True):
    L = lib()
    dt = np.float64 if fp64 else np.float32
    ins = [abi.DeviceBytes(np.ascontiguousarray(x, dtype=dt)) for x in (v0, v1, w, g)]
    nin = d.c1 + d.c0 + d.c0 * d.c1
    outs = [abi.DeviceBytes(np.zeros(n, dt)) for n in (d.w0 * d.h0 * d.c0, d.w1 * d.h1 * d.c1, d.nout * nin, d.nout)]
    st = L.nntc_hip_decode_adjoint(C.byref(d), 1 if fp64 else 0, *[x.address for x in ins], *[o.address for o in outs],
                                   None)
    assert st == 0, L.nntc_hip_last_error()
    return [o.read() for o in outs]


def random_model(rng, w0, h0, c0, c1, nout, k=4):
    d = SitesDesc(w0, h0, max(1, w0 // 4), max(1, h0 // 4), c0, c1, nout, k)
    nin = c1 + c0 + c0 * c1
    v0 = rng.uniform(-1, 1, (w0 * h0 * c0,))
    v1 = rng.uniform(-1, 1, (d.w1 * d.h1 * c1,))
    w = rng.uniform(-0.5, 0.5, (nout, nin))
    b = rng.uniform(0, 1, (nout,))
    return d, v0, v1, w, b


def test_decode_sites_centre_matches_decode_plane(hip_arm):
    """At the centre sites the fp32 forward is the arm's decode before saturation: rounded like decode_plane, the bytes
    are decode_plane's."""
    name = "small_2x100x76"
    spec = synth.WORKLOADS[name]
    tex = synth.workload(name)
    planes, ch = chain.source_chain(tex)
    _, per_site = chain.mip_weights(planes)
    desc = abi.make_desc(len(tex), spec["c0"], spec["c1"], planes, chain.output_weights(len(tex)))
    s = abi.Session(hip_arm, desc, [c.reshape(-1) for c in ch])
    print("decode_sites vs decode_plane small_2x100x76", flush=True)
    for i, (n, _) in enumerate(steps(s, Options(), per_site)):
        if n == "solve_level1_all":
            break
    pl = s.plane(0)
    w, b = s.decoder()
    d = SitesDesc(pl.w0, pl.h0, pl.w1, pl.h1, desc.c0, desc.c1, desc.nout, 4)
    out = decode_sites(d, s.v0(0), s.v1(0), w, b, fp64=False)
    centre = out.reshape(pl.h0 * pl.w0, 5, desc.nout)[:, 0, :]
    sat = np.clip(centre, np.float32(0), np.float32(1)) * np.float32(255)
    want = np.floor(sat.astype(np.float64) + 0.5).astype(np.uint8)   # roundf: half away from zero, sat >= 0
    got = s.decode_plane(0).reshape(-1, desc.nout)
    np.testing.assert_array_equal(want, got)


def test_adjoint_is_the_transpose():
    """<g, J_i delta_i> = <adjoint(g)_i, delta_i> for a random direction in each input i (v0, v1, W, b), with J_i delta_i
    from a central difference in that input alone: the decode is affine in each of its inputs when the others are held,
    so the difference is exact up to rounding."""
    rng = np.random.default_rng(7)
    d, v0, v1, w, b = random_model(rng, 24, 20, 3, 2, 6)
    sites = d.w0 * d.h0 * 5
    g = rng.normal(size=(sites, d.nout))
    grads = decode_adjoint(d, v0, v1, w, g)
    inputs = [v0, v1, w, b]
    eps = 1e-3
    for i, name in enumerate(("v0", "v1", "W", "b")):
        e = rng.normal(size=inputs[i].shape)
        plus = [x + eps * e if k == i else x for k, x in enumerate(inputs)]
        minus = [x - eps * e if k == i else x for k, x in enumerate(inputs)]
        lhs = float((g * (decode_sites(d, *plus) - decode_sites(d, *minus)) / (2 * eps)).sum())
        rhs = float((grads[i].reshape(-1) * e.reshape(-1)).sum())
        print(f"  {name}: <g, J delta> {lhs:.15e}  <adjoint g, delta> {rhs:.15e}", flush=True)
        assert abs(lhs - rhs) <= 1e-10 * abs(rhs), name


def autograd_decode():
    torch = pytest.importorskip("torch")

    class Decode(torch.autograd.Function):
        """The decode at every site of one plane, forward and backward through the C ABI (fp64)."""

        @staticmethod
        def forward(ctx, v0, v1, w, b, d):
            ctx.d = d
            ctx.save_for_backward(v0, v1, w)
            out = decode_sites(d, v0.detach().numpy(), v1.detach().numpy(), w.detach().numpy(), b.detach().numpy())
            return torch.from_numpy(out)

        @staticmethod
        def backward(ctx, g):
            v0, v1, w = ctx.saved_tensors
            dv0, dv1, dw, db = decode_adjoint(ctx.d, v0.numpy(), v1.numpy(), w.numpy(), g.contiguous().numpy())
            return (torch.from_numpy(dv0), torch.from_numpy(dv1), torch.from_numpy(dw.reshape(w.shape)),
                    torch.from_numpy(db), None)

    return torch, Decode


@pytest.mark.parametrize("size,fast", [((64, 64), True), ((8, 8), False)], ids=["64x64-projected", "8x8-full"])
def test_gradcheck(size, fast):
    torch, Decode = autograd_decode()
    rng = np.random.default_rng(11)
    d, v0, v1, w, b = random_model(rng, size[0], size[1], 2, 2, 3)
    args = [torch.tensor(x, dtype=torch.float64, requires_grad=True) for x in (v0, v1, w, b)]
    print(f"gradcheck {size} fast_mode={fast}", flush=True)
    assert torch.autograd.gradcheck(lambda a, bb, c, e: Decode.apply(a, bb, c, e, d), args, eps=1e-6, atol=1e-6,
                                    rtol=1e-5, fast_mode=fast)


def set_matrix_free(s, mode):
    f = s.arm.lib.nntc_hip_set_matrix_free
    f.argtypes = [C.c_void_p, C.c_int]
    assert f(s.ctx, mode) == 0


def level0_session(hip_arm, name, form):
    if name in synth.WORKLOADS:
        spec = synth.WORKLOADS[name]
        tex, c0, c1 = synth.workload(name), spec["c0"], spec["c1"]
    else:
        tex, c0, c1 = synth.make_material(2, (64, 64), 32, 3, 1.0, config=9), 2, 2
    planes, ch = chain.source_chain(tex)
    _, per_site = chain.mip_weights(planes)
    desc = abi.make_desc(len(tex), c0, c1, planes, chain.output_weights(len(tex)))
    s = abi.Session(hip_arm, desc, [c.reshape(-1) for c in ch])
    s.call("set_form", form)
    set_matrix_free(s, 0)
    gen = steps(s, Options(), per_site)
    for n, _ in gen:
        if n == "solve_level1_all":
            break
    return s, desc, per_site


def apply(s, mode, p, ridge):
    L = s.arm.lib
    L.nntc_hip_level0_apply.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_void_p, C.c_void_p]
    out = np.zeros_like(p)
    st = L.nntc_hip_level0_apply(s.ctx, C.addressof(s.state), 4, ridge, mode, p.ctypes.data, out.ctypes.data)
    assert st == 0, s.arm.fn["last_error"]().decode()
    return out


def test_adjoint_operator_matches_assembled_stencil_fp64(hip_arm):
    """J^T C J p from the fp64 decode and its adjoint equals the exact form's assembled level-0 stencil times p on the
    base plane (ridge 0, so lambda = 0)."""
    s, desc, _ = level0_session(hip_arm, "crop_2x64x64", abi.FORM_EXACT)
    print("J^T C J vs the assembled stencil (fp64), 64x64, c0 2 c1 2", flush=True)
    rng = np.random.default_rng(5)
    n0 = sum(s.plane(p).w0 * s.plane(p).h0 for p in range(s.nplanes)) * desc.c0
    p = rng.normal(size=n0)
    hp = apply(s, 0, p, 0.0)
    pl = s.plane(0)
    base = pl.w0 * pl.h0 * desc.c0
    w, b = s.decoder()
    d = SitesDesc(pl.w0, pl.h0, pl.w1, pl.h1, desc.c0, desc.c1, desc.nout, 4)
    v1 = s.v1(0).astype(np.float64)
    jp = decode_sites(d, p[:base], v1, w, b) - decode_sites(d, np.zeros(base), v1, w, b)
    cw = np.array([desc.cw[r] for r in range(desc.nout)], dtype=np.float64)
    dv0, _, _, _ = decode_adjoint(d, np.zeros(base), v1, w, jp * cw[None, :])
    err = np.abs(dv0 - hp[:base]).max() / np.abs(hp[:base]).max()
    print(f"  max |J^T C J p - H p| / max |H p| = {err:.3e}", flush=True)
    # The assembly forms each site's rows d out / d s0 and the level-1 sample in fp32 (the reference's arithmetic), the
    # adjoint in fp64 from the same inputs: they meet to fp32's resolution, not fp64's.
    assert err <= 1e-5


@pytest.mark.parametrize("name", ["crop_2x64x64", "plain_cpu_4x256"])
def test_matrix_free_operator_matches_assembled_fp32(name, hip_arm):
    """The fast form's matrix-free operator against its assembled fp32 stencil on every plane: (H + lambda I) p to
    fp32 rounding."""
    s, desc, _ = level0_session(hip_arm, name, abi.FORM_FAST)
    print(f"matrix-free vs assembled fp32 operator {name}", flush=True)
    rng = np.random.default_rng(3)
    n0 = sum(s.plane(p).w0 * s.plane(p).h0 for p in range(s.nplanes)) * desc.c0
    p = rng.normal(size=n0)
    a = apply(s, 0, p, 1e-4)
    m = apply(s, 1, p, 1e-4)
    off = 0
    for q in range(s.nplanes):
        n = s.plane(q).w0 * s.plane(q).h0 * desc.c0
        err = np.abs(m[off:off + n] - a[off:off + n]).max() / np.abs(a[off:off + n]).max()
        assert err <= 2e-5, f"plane {q}: {err:.3e}"
        off += n
    print(f"  worst plane {err:.3e} (last)", flush=True)


@pytest.mark.parametrize("name", ["crop_2x64x64", "plain_cpu_4x256"])
def test_matrix_free_solve_matches_assembled(name, hip_arm):
    """Block (c') by the fast conjugate gradients on the matrix-free operator and on the assembled stencil, from one
    state: the corrections agree to the solves' own tolerance and the iteration counts to a probe."""
    sa, desc, per_site = level0_session(hip_arm, name, abi.FORM_FAST)
    sm, _, _ = level0_session(hip_arm, name, abi.FORM_FAST)
    set_matrix_free(sm, 1)
    print(f"matrix-free vs assembled block (c') {name}", flush=True)
    before = [sa.v0(p).astype(np.float64) for p in range(sa.nplanes)]
    ra, _ = sa.block("solve_level0_cont_all", 4, 1e-4)
    rm, _ = sm.block("solve_level0_cont_all", 4, 1e-4)
    for p in range(sa.nplanes):
        da = sa.v0(p) - before[p]
        dm = sm.v0(p) - before[p]
        scale = np.abs(da).max()
        err = np.abs(dm - da).max() / scale if scale > 0 else 0.0
        assert err <= 1e-3, f"plane {p}: correction deviation {err:.3e} of its size"
        assert abs(ra[p].iterations - rm[p].iterations) <= 8, (p, ra[p].iterations, rm[p].iterations)
    print(f"  iterations {[r.iterations for r in ra]} vs {[r.iterations for r in rm]}", flush=True)
