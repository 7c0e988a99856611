"""Per-call check of the plain and the fast (chosen) form against the exact form, the reference's bits.

The reference's `--backend check` (src/backend_check.cpp) mirrors the CUDA arm's state into the CPU arm before every
seam call, makes the call on both and holds the two results to measured bars (src/backend_check.cpp:91-151). This is
the same harness for the forms of the gfx950 arm: the exact form (bit-identical to the reference CPU backend, as
tests/test_gpu_lockstep.py shows call by call) is driven through the encode's whole schedule (tests/schedule.py); before
each call its state - both latents' values and indices on every plane, the decoder, the grids, the BC endpoints and
selectors - is copied into a session of the other form, the same call is made on both, and the other form's outputs
are compared with the exact ones. The schedule follows the exact form's results, so every call starts from the same
state on both sides and a deviation belongs to that call alone.

The bars, per function (the reference's where the form's arithmetic allows them):
  objective_eval           E within 1e-6 relative. The reference's BAR_SUM (1e-9) holds for fp64 decodes; the
                           plain and the fast form decode each site in fp32, whose rounding of the 18 outputs' error
                           terms moves E by a few 1e-8.
  solve_decoder            the coefficient vector within 1e-5 of its largest coefficient (normwise; ten times inside
                           "within 1e-4 relative"), and every coefficient within 1e-3 of its own magnitude, floored at
                           1e-3 of the largest (a coefficient a thousand times below the largest is held to 1e-6 of
                           the largest). The plain form's fp64 Gram meets the exact bits; the fast form's bf16
                           three-part Gram, accumulated in fp32 over 64 sites at a time, sits near 2e-6 normwise and
                           a few 1e-4 per coefficient (the worst of each call is printed).
  solve_*_all (continuous) every value within BAR_ENTRY = 1e-5 of its plane's range after the correction
                           (src/backend_check.cpp:108): the conjugate gradients stop at 1e-6 in these forms, so the
                           corrections of two forms land about that far apart relative to the correction's own size,
                           which is a small part of the range.
  sweep_*_all              at most FLIP_SHARE = 1 % of the indices differ (src/backend_check.cpp:145-151).
  bc_refine                at most 1 % of the blocks' endpoint and selector bytes differ.
  init0_residual_channel   every seeded value within BAR_ENTRY_DIRECT = 1e-6 of its plane's range
                           (src/backend_check.cpp:120): the fast covariance adds the same products in another order.
  the other inits, grids   every output identical (the forms share those kernels).
  and the seed pack
The worst figure of every function is printed, so a run shows how far inside its bar each one sits.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from nntc_amd import abi, chain, synth
from schedule import Options, steps
from test_gpu_lockstep import material

pytestmark = pytest.mark.gpu

BAR_E = 1e-6
BAR_DECODER_NORMWISE = 1e-5
BAR_DECODER = 1e-3
DECODER_FLOOR = 1e-3
BAR_ENTRY = 1e-5
FLIP_SHARE = 0.01


class Twin:
    """The exact session as the schedule sees it; every compared call is also made on the other form's session from
    the exact state, and its deviation recorded."""

    def __init__(self, ex: abi.Session, other: abi.Session):
        self.ex, self.o = ex, other
        self.nplanes, self.desc = ex.nplanes, ex.desc
        self.worst: dict[str, float] = {}
        self.calls: dict[str, int] = {}
        self.log: list[str] = []

    # ---- what the schedule reads from its session
    @property
    def state(self):
        return self.ex.state

    def plane(self, p):
        return self.ex.plane(p)

    def v0(self, p):
        return self.ex.v0(p)

    def k0(self, p):
        return self.ex.k0(p)

    def v1(self, p):
        return self.ex.v1(p)

    def k1(self, p):
        return self.ex.k1(p)

    def call(self, name, *args):
        self.ex.call(name, *args)

    def upload_v0(self, p, v):
        self.ex.upload_v0(p, v)

    def upload_v1(self, p, v):
        self.ex.upload_v1(p, v)

    def _mirror(self, k0):
        return self.ex._mirror(k0)

    # ---- mirroring and recording
    def sync(self):
        """The exact session's state copied into the other session."""
        ex, o = self.ex, self.o
        C.memmove(C.addressof(o.state), C.addressof(ex.state), C.sizeof(abi.ModelState))
        o.call("upload_decoder_to_device", C.byref(o.state))
        o.call("upload_grids_to_device", C.byref(o.state))
        for p in range(self.nplanes):
            o.upload_v0(p, ex.v0(p))
            o.upload_v1(p, ex.v1(p))
            a = ex.k0(p)
            o.call("level0_upload_indices", p, abi._ptr(a, C.c_uint8), a.size)
            a = ex.k1(p)
            o.call("level1_upload_indices", p, abi._ptr(a, C.c_uint8), a.size)
            pl = ex.plane(p)
            nb = ((pl.w0 + 3) // 4) * ((pl.h0 + 3) // 4)
            ep = np.zeros(nb * self.desc.c0 * 2, np.uint8)
            sel = np.zeros(nb * self.desc.c0 * 16, np.uint8)
            ex.call("bc_state_save", p, abi._ptr(ep, C.c_uint8), ep.size, abi._ptr(sel, C.c_uint8), sel.size)
            o.call("bc_state_restore", p, abi._ptr(ep, C.c_uint8), ep.size, abi._ptr(sel, C.c_uint8), sel.size)

    def record(self, fn, dev):
        self.worst[fn] = max(self.worst.get(fn, 0.0), float(dev))
        self.calls[fn] = self.calls.get(fn, 0) + 1

    def latent_dev(self, level0):
        """Worst |value difference| over a plane's range, over the planes of one latent."""
        worst = 0.0
        for p in range(self.nplanes):
            a = self.ex.v0(p) if level0 else self.ex.v1(p)
            b = self.o.v0(p) if level0 else self.o.v1(p)
            rng = float(a.max() - a.min()) or 1.0
            worst = max(worst, float(np.abs(a.astype(np.float64) - b).max()) / rng)
        return worst

    def flips(self, level0):
        n = bad = 0
        for p in range(self.nplanes):
            a = self.ex.k0(p) if level0 else self.ex.k1(p)
            b = self.o.k0(p) if level0 else self.o.k1(p)
            n += a.size
            bad += int((a != b).sum())
        return bad / max(1, n)

    def bc_flips(self):
        n = bad = 0
        for p in range(self.nplanes):
            for f in range(1 if self.desc.c0 <= 2 else 2):
                a, b = self.ex.bc_blocks(p, f), self.o.bc_blocks(p, f)
                n += a.size
                bad += int((a != b).sum())
        return bad / max(1, n)

    # ---- the compared calls
    def objective(self, k, per_site):
        self.sync()
        a = self.ex.objective(k, per_site)
        b = self.o.objective(k, per_site)
        self.record("objective_eval", abs(b.e - a.e) / a.e)
        return a

    def tally(self):
        out = (C.c_longlong * 4)()
        self.ex.arm.fn["decoder_ridge_tally"](out)
        return list(out)

    def solve_decoder(self, k, per_site):
        self.sync()
        t0 = self.tally()
        ms = self.ex.solve_decoder(k, per_site)
        t1 = self.tally()
        self.o.solve_decoder(k, per_site)
        t2 = self.tally()
        # The rung of the ridge ladder each form's solve ended on (3: refused).
        rung_ex = [b - a for a, b in zip(t0, t1)].index(1)
        rung_o = [b - a for a, b in zip(t1, t2)].index(1)
        (wa, ba), (wb, bb) = self.ex.decoder(), self.o.decoder()
        x = np.concatenate([wa.ravel(), ba]).astype(np.float64)
        y = np.concatenate([wb.ravel(), bb]).astype(np.float64)
        floor = DECODER_FLOOR * np.abs(x).max()
        dev = (np.abs(y - x) / np.maximum(np.abs(x), floor)).max()
        normwise = np.abs(y - x).max() / np.abs(x).max()
        self.log.append(f"solve_decoder #{self.calls.get('solve_decoder', 0)}: rungs {rung_ex}/{rung_o}, "
                        f"worst {dev:.3e}, normwise {normwise:.3e}")
        self.record("solve_decoder", dev)
        self.record("solve_decoder normwise", normwise)
        return ms

    def block(self, name, k, ridge, sweeps=None):
        self.sync()
        r = self.ex.block(name, k, ridge, sweeps)
        self.o.block(name, k, ridge, sweeps)
        level0 = "level0" in name
        if sweeps is None:
            self.record(name, self.latent_dev(level0))
        else:
            self.record(name, self.flips(level0))
        return r

    def grid(self, name, policy):
        self.sync()
        self.ex.grid(name, policy)
        self.o.grid(name, policy)
        same = all(np.array_equal(np.array(getattr(self.ex.state, f)), np.array(getattr(self.o.state, f)))
                   for f in ("lo0", "hi0", "lo1", "hi1"))
        self.record(name, 0.0 if same else 1.0)

    def init_level0(self, **kw):
        self.ex.init_level0(**kw)
        self.o.init_level0(**kw)
        self.record("init_level0", self.latent_dev(True))

    def init_level1(self, **kw):
        r = self.ex.init_level1(**kw)
        self.o.init_level1(**kw)
        self.record("init_level1", self.latent_dev(False))
        return r

    def init0_residual_channel(self, channel, plane_weight):
        self.sync()
        r = self.ex.init0_residual_channel(channel, plane_weight)
        self.o.init0_residual_channel(channel, plane_weight)
        self.record("init0_residual_channel", self.latent_dev(True))
        return r

    def bc_pack_prepare(self, k, seed, k0_mirror=None):
        self.sync()
        r = self.ex.bc_pack_prepare(k, seed, k0_mirror)
        self.o.bc_pack_prepare(k, seed, k0_mirror)
        self.record("bc_pack_prepare", self.bc_flips())
        return r

    def bc_refine(self, passes, k0_mirror=None):
        self.sync()
        # The other form's refinement reads its own level-0 stencil, which its last assembly built from the state it
        # was given then; assemble it again from the exact state so both refine the same quadratic's form.
        self.o.call("assemble_level0_for_refine", C.byref(self.o.state), 4, None)
        r = self.ex.bc_refine(passes, k0_mirror)
        self.o.bc_refine(passes, k0_mirror)
        self.record("bc_refine", self.bc_flips())
        return r


BARS = {
    "objective_eval": BAR_E,
    "init0_residual_channel": 1e-6,   # BAR_ENTRY_DIRECT (src/backend_check.cpp:120): the covariance's sums in a new order
    "solve_decoder": BAR_DECODER,
    "solve_decoder normwise": BAR_DECODER_NORMWISE,
    "solve_level1_all": BAR_ENTRY,
    "solve_level0_cont_all": BAR_ENTRY,
    "sweep_level1_all": FLIP_SHARE,
    "sweep_level0_cont_all": FLIP_SHARE,
    "bc_refine": FLIP_SHARE,
}


@pytest.mark.parametrize("form", [abi.FORM_FAST, abi.FORM_PLAIN], ids=["fast", "plain"])
@pytest.mark.parametrize("name", ["small_6x100x76_c4c4", "plain_cpu_4x256"])
def test_per_call_against_exact(name, form, hip_arm):
    tex, c0, c1 = material(name)
    planes, ch = chain.source_chain(tex)
    _, per_site = chain.mip_weights(planes)
    desc = abi.make_desc(len(tex), c0, c1, planes, chain.output_weights(len(tex)))
    flat = [c.reshape(-1) for c in ch]
    ex = abi.Session(hip_arm, desc, flat)
    other = abi.Session(hip_arm, desc, flat)
    other.call("set_form", form)
    print(f"per-call {name} form {form}", flush=True)
    t = Twin(ex, other)
    n = sum(1 for _ in steps(t, Options(), per_site))
    assert n > 40
    lines = [f"  {fn:24s} {t.calls[fn]:3d} calls  worst {t.worst[fn]:.3e}  bar {BARS.get(fn, 0.0):.0e}"
             for fn in sorted(t.worst)]
    print("\n".join(t.log + lines), flush=True)
    failed = [fn for fn, w in t.worst.items() if w > BARS.get(fn, 0.0)]
    assert not failed, "over the bar: " + ", ".join(f"{fn} {t.worst[fn]:.3e}" for fn in failed)
