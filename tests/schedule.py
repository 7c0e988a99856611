"""The encode's call sequence over the seam, as a generator, for lockstep parity tests.

`steps(session, opts)` issues the reference's schedule for the default --l0 bc8 / --init0 residual / --init box path
(src/main.cpp:1930-2694) and yields (name, info) after every seam call, so a test can advance two arms together and
compare their state after each call. It restates the schedule only; every number comes from the arm.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from nntc_amd import abi

E_REL = 1e-6
E_NOISE = 1e-8


@dataclass
class Options:
    k: int = 4
    rounds: int = 8
    tol: float = 0.0
    q1_start: int = 3
    q1_sweeps: int = 4
    q1_range: int = abi.RANGE_PCT
    ridge: float = 1e-4
    bc_refine: int = 4
    bc_outer: int = 2
    pca: bool = False           # --init pca (the box init otherwise)
    luma: bool = False          # --init0 luma (no residual seed)


def steps(s: abi.Session, o: Options, per_site):
    nplanes = s.nplanes
    s.init_level0(luma_channel0=o.luma)
    yield "init_level0", {}
    irep = s.init_level1(pca=o.pca)
    yield "init_level1", {"eigenvalue": list(irep.eigenvalue), "peak": list(irep.peak),
                          "total_variance": irep.total_variance, "components": irep.components}
    e = s.objective(o.k, per_site)
    yield "objective", {"e": e.e}
    s.solve_decoder(o.k, per_site)
    yield "solve_decoder", {}
    after = s.objective(o.k, per_site)
    yield "objective", {"e": after.e}
    for c in range(0 if o.luma else s.desc.c0):
        rep = s.init0_residual_channel(c, per_site)
        yield "init0_residual_channel", {"eigenvalue": rep.eigenvalue, "peak": rep.peak, "texture": rep.texture}
        seeded = s.objective(o.k, per_site)
        yield "objective", {"e": seeded.e}
        s.solve_decoder(o.k, per_site)
        yield "solve_decoder", {}
        after = s.objective(o.k, per_site)
        yield "objective", {"e": after.e}

    frozen = False
    e_prev = after.e
    tol_hits = 0
    for r in range(1, o.rounds + 1):
        s.solve_decoder(o.k, per_site)
        yield "solve_decoder", {"round": r}
        ea = s.objective(o.k, per_site)
        yield "objective", {"e": ea.e}
        if o.q1_start > 0 and r >= o.q1_start and not frozen:
            s.grid("freeze_level1_grid", o.q1_range)
            yield "freeze_level1_grid", {}
            s.grid("freeze_level0_grid", o.q1_range)
            yield "freeze_level0_grid", {}
            frozen = True
        if frozen:
            lr, _ = s.block("sweep_level1_all", o.k, o.ridge, o.q1_sweeps)
            yield "sweep_level1_all", {"moved": [x.moved for x in lr]}
        else:
            lr, _ = s.block("solve_level1_all", o.k, o.ridge)
            yield "solve_level1_all", {"iterations": [x.iterations for x in lr], "lambda": [x.lambda_ for x in lr]}
        eb = s.objective(o.k, per_site)
        yield "objective", {"e": eb.e}
        if frozen:
            lr0, _ = s.block("sweep_level0_cont_all", o.k, o.ridge, o.q1_sweeps)
            yield "sweep_level0_cont_all", {"moved": [x.moved for x in lr0]}
        else:
            lr0, _ = s.block("solve_level0_cont_all", o.k, o.ridge)
            yield "solve_level0_cont_all", {"iterations": [x.iterations for x in lr0]}
        ec = s.objective(o.k, per_site)
        yield "objective", {"e": ec.e, "round_end": r}
        moved1 = sum(x.moved if x.quantised else x.moved_values for x in lr)
        moved0 = sum(x.moved if x.quantised else x.moved_values for x in lr0)
        delta = np.sqrt(sum(x.delta2 for x in lr))
        plane = np.sqrt(sum(x.plane2 for x in lr))
        drop = (e_prev - ec.e) / ec.e if ec.e > 0.0 else 0.0
        e_prev = ec.e
        still = moved1 == 0 if frozen else delta <= 1e-7 * (plane if plane > 0.0 else 1.0)
        if moved0 == 0 and still:
            break
        tol_hits = tol_hits + 1 if drop < o.tol else 0
        if tol_hits >= 2:
            break

    assert frozen, "the lockstep schedule covers the --q1-start path only"
    s.block("sweep_level1_all", o.k, o.ridge, o.q1_sweeps)
    yield "sweep_level1_all", {}
    s.block("sweep_level0_cont_all", o.k, o.ridge, o.q1_sweeps)
    yield "sweep_level0_cont_all", {}
    s.solve_decoder(o.k, per_site)
    yield "solve_decoder", {}
    k0 = [s.k0(p) for p in range(nplanes)]

    probe = s.objective(o.k, per_site)
    yield "objective", {"e": probe.e}
    psnr_seed, _ = s.bc_pack_prepare(o.k, True, k0)
    yield "bc_pack_prepare", {"pack_psnr": psnr_seed}
    probe = s.objective(o.k, per_site)
    yield "objective", {"e": probe.e}
    rep, psnr, _ = s.bc_refine(o.bc_refine, k0)
    yield "bc_refine", {"improved": [x.improved for x in rep], "delta_e": [x.delta_e for x in rep]}
    probe = s.objective(o.k, per_site)
    yield "objective", {"e": probe.e}
    s.block("sweep_level1_all", o.k, o.ridge, o.q1_sweeps)
    yield "sweep_level1_all", {}
    s.solve_decoder(o.k, per_site)
    yield "solve_decoder", {}
    probe = s.objective(o.k, per_site)
    yield "objective", {"e": probe.e}
    e_ship = probe.e
    for pss in range(1, o.bc_outer + 1):
        keep = {
            "v0": [s.v0(p) for p in range(nplanes)],
            "k0dev": [s.k0(p) for p in range(nplanes)],
            "v1": [s.v1(p) for p in range(nplanes)],
            "k1dev": [s.k1(p) for p in range(nplanes)],
            "k0": [a.copy() for a in k0],
            "state": abi.ModelState.from_buffer_copy(s.state),
        }
        keep_bc = []
        for p in range(nplanes):
            pl = s.plane(p)
            nb = ((pl.w0 + 3) // 4) * ((pl.h0 + 3) // 4)
            ep = np.zeros(nb * s.desc.c0 * 2, np.uint8)
            sel = np.zeros(nb * s.desc.c0 * 16, np.uint8)
            s.call("bc_state_save", p, abi._ptr(ep, abi.C.c_uint8), ep.size, abi._ptr(sel, abi.C.c_uint8), sel.size)
            keep_bc.append((ep, sel))
        s.block("sweep_level1_all", o.k, o.ridge, o.q1_sweeps)
        yield "sweep_level1_all", {"outer": pss}
        s.solve_decoder(o.k, per_site)
        yield "solve_decoder", {}
        s.block("sweep_level0_cont_all", o.k, o.ridge, o.q1_sweeps)
        yield "sweep_level0_cont_all", {}
        k0 = [s.k0(p) for p in range(nplanes)]
        s.bc_pack_prepare(o.k, True, k0)
        yield "bc_pack_prepare", {}
        s.bc_refine(o.bc_refine, k0)
        yield "bc_refine", {}
        probe = s.objective(o.k, per_site)
        accepted = probe.e < e_ship
        yield "objective", {"e": probe.e, "accepted": accepted}
        if accepted:
            e_ship = probe.e
            continue
        for p in range(nplanes):
            s.upload_v0(p, keep["v0"][p])
            a = keep["k0dev"][p]
            s.call("level0_upload_indices", p, abi._ptr(a, abi.C.c_uint8), a.size)
            s.upload_v1(p, keep["v1"][p])
            a = keep["k1dev"][p]
            s.call("level1_upload_indices", p, abi._ptr(a, abi.C.c_uint8), a.size)
            ep, sel = keep_bc[p]
            s.call("bc_state_restore", p, abi._ptr(ep, abi.C.c_uint8), ep.size, abi._ptr(sel, abi.C.c_uint8),
                   sel.size)
        k0 = keep["k0"]
        abi.C.memmove(abi.C.addressof(s.state), abi.C.addressof(keep["state"]), abi.C.sizeof(abi.ModelState))
        s.call("upload_decoder_to_device", abi.C.byref(s.state))
        s.call("upload_grids_to_device", abi.C.byref(s.state))
        s.call("bc_blocks_from_device", s._mirror(k0)[0], None, None)
        yield "restore", {}
        break
    s.block("sweep_level1_all", o.k, o.ridge, o.q1_sweeps)
    yield "sweep_level1_all", {"final": True}
    s.solve_decoder(o.k, per_site)
    yield "solve_decoder", {}
    shipped = s.objective(o.k, per_site)
    yield "shipped", {"e": shipped.e, "centre_mse": shipped.centre_mse}


def steps_palette(s: abi.Session, o: Options, per_site, sweeps: int = 2):
    """The --l0 palette schedule (src/main.cpp:1930-2440 under !l0_bc8): block (c) is the per-texel palette search,
    each seeded channel is followed by one search and a refit, level 0 has no grid to freeze, and the run ends with a
    last search and a last block (a)."""
    nplanes = s.nplanes
    s.init_level0(luma_channel0=o.luma)
    yield "init_level0", {}
    irep = s.init_level1(pca=o.pca)
    yield "init_level1", {"eigenvalue": list(irep.eigenvalue), "peak": list(irep.peak)}
    s.objective(o.k, per_site)
    s.solve_decoder(o.k, per_site)
    yield "solve_decoder", {}
    after = s.objective(o.k, per_site)
    yield "objective", {"e": after.e}
    for c in range(0 if o.luma else s.desc.c0):
        rep = s.init0_residual_channel(c, per_site)
        yield "init0_residual_channel", {"eigenvalue": rep.eigenvalue, "peak": rep.peak, "divisor": rep.divisor,
                                         "texture": rep.texture}
        seeded = s.objective(o.k, per_site)
        yield "objective", {"e": seeded.e}
        s.solve_decoder(o.k, per_site)
        yield "solve_decoder", {}
        s.objective(o.k, per_site)
        l0, _ = s.solve_level0_all(o.k, sweeps)
        yield "solve_level0_all", {"moved": [list(x.moved) for x in l0]}
        s.solve_decoder(o.k, per_site)
        yield "solve_decoder", {}
        after = s.objective(o.k, per_site)
        yield "objective", {"e": after.e}

    frozen = False
    e_prev = after.e
    tol_hits = 0
    for r in range(1, o.rounds + 1):
        s.solve_decoder(o.k, per_site)
        yield "solve_decoder", {"round": r}
        s.objective(o.k, per_site)
        if o.q1_start > 0 and r >= o.q1_start and not frozen:
            s.grid("freeze_level1_grid", o.q1_range)
            yield "freeze_level1_grid", {}
            frozen = True
        if frozen:
            lr, _ = s.block("sweep_level1_all", o.k, o.ridge, o.q1_sweeps)
            yield "sweep_level1_all", {"moved": [x.moved for x in lr]}
        else:
            lr, _ = s.block("solve_level1_all", o.k, o.ridge)
            yield "solve_level1_all", {"iterations": [x.iterations for x in lr]}
        s.objective(o.k, per_site)
        l0, _ = s.solve_level0_all(o.k, sweeps)
        yield "solve_level0_all", {"moved": [list(x.moved) for x in l0], "joint": [x.joint for x in l0]}
        ec = s.objective(o.k, per_site)
        yield "objective", {"e": ec.e, "round_end": r}
        moved1 = sum(x.moved if x.quantised else x.moved_values for x in lr)
        moved0 = sum(x.moved_total for x in l0)
        delta = np.sqrt(sum(x.delta2 for x in lr))
        plane = np.sqrt(sum(x.plane2 for x in lr))
        drop = (e_prev - ec.e) / ec.e if ec.e > 0.0 else 0.0
        e_prev = ec.e
        still = moved1 == 0 if frozen else delta <= 1e-7 * (plane if plane > 0.0 else 1.0)
        if moved0 == 0 and still:
            break
        tol_hits = tol_hits + 1 if drop < o.tol else 0
        if tol_hits >= 2:
            break

    if frozen:
        s.block("sweep_level1_all", o.k, o.ridge, o.q1_sweeps)
        yield "sweep_level1_all", {}
    else:
        s.grid("quantise_level1", o.q1_range)
        yield "quantise_level1", {}
    l0, _ = s.solve_level0_all(o.k, sweeps)
    yield "solve_level0_all", {"moved": [list(x.moved) for x in l0]}
    s.solve_decoder(o.k, per_site)
    yield "solve_decoder", {}
    shipped = s.objective(o.k, per_site)
    yield "objective", {"e": shipped.e, "shipped": True}
    del nplanes
