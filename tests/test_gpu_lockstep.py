"""Parity of the gfx950 arm with the reference CPU backend, call by call, over the whole encode schedule.

Both arms are created from the same seeded material and driven through the same schedule (tests/schedule.py); after
every seam call the two arms' state is compared: the objective, the decoder, both grids, both latents' values and
indices on every plane, and the BC blocks. The exact form is held to the reference's bits.
"""

import numpy as np
import pytest

from nntc_amd import abi, chain, synth
from schedule import Options, steps, steps_palette

pytestmark = pytest.mark.gpu


def material(name):
    if name in synth.WORKLOADS:
        spec = synth.WORKLOADS[name]
        return synth.workload(name), spec["c0"], spec["c1"]
    textures, w, h, c0, c1, smooth = {
        "small_2x100x76_c3c3": (2, 100, 76, 3, 3, 3),
        "small_1x64x40_c1c2": (1, 64, 40, 1, 2, 3),
        # more level-1 channels than the source has directions (the pca init's zero channel)
        "small_1x64x40_c2c4": (1, 64, 40, 2, 4, 3),
        # six textures (nout 18) and four level-0 channels (two BC5 files), as the 2048^2 precision workload
        "small_6x100x76_c4c4": (6, 100, 76, 4, 4, 5),
    }[name]
    tex = synth.make_material(textures, (w, h), 32, smooth, 1.0, config=9)
    return tex, c0, c1


def sessions(name, hip_arm, ref_arm, init0_texture=False, init0_luma=False, bits0=None, bc0=1):
    tex, c0, c1 = material(name)
    planes, ch = chain.source_chain(tex)
    _, per_site = chain.mip_weights(planes)
    desc = abi.make_desc(len(tex), c0, c1, planes, chain.output_weights(len(tex)), bits0=bits0, bc0=bc0)
    desc.init0_texture = 1 if init0_texture else 0
    desc.init0_residual = 0 if init0_luma else 1
    flat = [c.reshape(-1) for c in ch]
    return abi.Session(hip_arm, desc, flat), abi.Session(ref_arm, desc, flat), per_site


def compare(a, b, step, i):
    where = f"step {i} ({step})"
    assert a.state.has_decoder == b.state.has_decoder, where
    wa, ba = a.decoder()
    wb, bb = b.decoder()
    np.testing.assert_array_equal(wa, wb, err_msg=f"decoder W after {where}")
    np.testing.assert_array_equal(ba, bb, err_msg=f"decoder b after {where}")
    for f in ("lo0", "hi0", "lo1", "hi1"):
        np.testing.assert_array_equal(np.array(getattr(a.state, f)), np.array(getattr(b.state, f)),
                                      err_msg=f"{f} after {where}")
    for p in range(a.nplanes):
        for fn in ("v0", "k0", "v1", "k1"):
            x, y = getattr(a, fn)(p), getattr(b, fn)(p)
            if not np.array_equal(x, y):
                bad = np.flatnonzero(x != y)
                raise AssertionError(f"{fn} plane {p} after {where}: {bad.size} of {x.size} differ, first at "
                                     f"{bad[0]}: {x[bad[0]]!r} vs {y[bad[0]]!r}")


@pytest.mark.parametrize("name", ["small_1x64x40_c1c2", "small_2x100x76_c3c3", "small_6x100x76_c4c4",
                                  "plain_cpu_4x256"])
def test_lockstep_exact(name, hip_arm, ref_arm):
    print(f"lockstep {name}", flush=True)
    h, r, per_site = sessions(name, hip_arm, ref_arm)
    run_lockstep(h, r, per_site, Options())


@pytest.mark.parametrize("name", ["small_2x100x76_c3c3", "small_1x64x40_c2c4"])
def test_lockstep_exact_pca_and_texture_inits(name, hip_arm, ref_arm):
    """--init pca (level 1 from the principal directions of the base plane's block means, src/init.cu:861-987) and
    --init0 texture (each level-0 channel seeded along the leading direction of one texture's block of the residual's
    covariance, src/init.cu:1219-1260), call by call against the reference; the second material has more level-1
    channels than its source has directions, so one pca channel is all zeros."""
    print(f"lockstep {name}, pca + init0 texture", flush=True)
    h, r, per_site = sessions(name, hip_arm, ref_arm, init0_texture=True)
    run_lockstep(h, r, per_site, Options(pca=True))


def test_lockstep_exact_luma_init(hip_arm, ref_arm):
    """--init0 luma: level 0's channel 0 from the source luminance, no residual seed (src/init.cu:389-482)."""
    print("lockstep small_2x100x76_c3c3, init0 luma", flush=True)
    h, r, per_site = sessions("small_2x100x76_c3c3", hip_arm, ref_arm, init0_luma=True)
    run_lockstep(h, r, per_site, Options(luma=True))


@pytest.mark.parametrize("name,bits0,bc0", [("small_2x100x76_c3c3", [3, 2, 3], 1), ("small_2x100x76_c3c3", [3, 3, 3], 1),
                                             ("small_1x64x40_c2c4", [4, 4], 0), ("small_1x64x40_c1c2", [1], 1),
                                             ("small_6x100x76_c4c4", [4, 4, 4, 4], 0)],
                         ids=["8bits-joint-bc", "9bits-sweeps-bc", "8bits-joint-uncompressed", "1bit-joint-bc",
                              "16bits-sweeps-uncompressed"])
def test_lockstep_exact_palette(name, bits0, bc0, hip_arm, ref_arm):
    """--l0 palette (src/solve_level0.cu, src/init.cu:389-428, 1327-1377): the per-texel palette search (the joint
    enumeration at 8 bits or fewer, coordinate sweeps above), the palette init and the seed's percentile divisor and
    snap, call by call against the reference, on both palettes (block-compressed k / levels, uncompressed replicated
    bytes)."""
    print(f"lockstep {name}, --l0 palette bits {bits0} bc0 {bc0}", flush=True)
    h, r, per_site = sessions(name, hip_arm, ref_arm, bits0=bits0, bc0=bc0)
    run_lockstep(h, r, per_site, Options(rounds=5), palette=True)


def run_lockstep(h, r, per_site, o, palette=False):
    step = steps_palette if palette else steps
    gh, gr = step(h, o, per_site), step(r, o, per_site)
    least = 20 if palette else 40
    i = 0
    for (nh, ih), (nr, ir) in zip(gh, gr):
        assert nh == nr
        for key in ih:
            if key == "pack_psnr":
                # A report figure: the packing error is summed per block on the device, not in the reference's
                # row-major order, so only its last bits may differ. No decision reads it.
                assert abs(ih[key] - ir[key]) <= 1e-12 * abs(ir[key]), f"{key} at step {i}"
                continue
            assert ih[key] == ir[key], f"{key} at step {i} ({nh}): {ih[key]!r} vs {ir[key]!r}"
        compare(h, r, nh, i)
        i += 1
    assert i > least
    for p in range(h.nplanes):
        # The palette mode packs level 0 on the host when it writes the asset: no blocks live on either arm.
        for f in range(0 if palette else 1 if h.desc.c0 <= 2 else 2):
            np.testing.assert_array_equal(h.bc_blocks(p, f), r.download_blocks(p, f, h.bc_blocks(p, f).size * 2),
                                          err_msg=f"blocks plane {p} file {f}")
        np.testing.assert_array_equal(h.decode_plane(p), r.decode_plane(p), err_msg=f"decode plane {p}")
    # The per-texture error sums the PSNR report is made of, over the whole base and over a crop.
    base = h.plane(0)
    for wh in ((base.w0, base.h0), (max(1, base.w0 - 3), max(1, base.h0 // 2))):
        assert h.texture_sse(*wh) == r.texture_sse(*wh), f"texture sse over {wh}"
    # The on-grid check made before the asset is written: no level-1 value off its index's grid value, then one value
    # moved by an ulp on both arms, counted by both.
    values, off = h.level1_off_grid()
    assert (values, off) == r.level1_off_grid() and off == 0
    v = np.float32(h.v1(0)[0])
    for s in (h, r):
        s.call("level1_poke", 0, 0, float(np.nextafter(v, np.float32(np.inf))))
    assert h.level1_off_grid() == r.level1_off_grid() == (values, 1)
