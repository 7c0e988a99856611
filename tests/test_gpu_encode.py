"""The whole encode on the gfx950 arm against the reference encoder's own outputs.

The driver (include/nntc_encode.h) runs the reference's schedule over the HIP arm on a seeded workload and writes the
asset; the fixtures in tests/golden/<workload>/report.json are the reference CPU encoder's outputs on the same
workload (tests/golden/make_golden.py). The exact form must reproduce them: the same .dds bytes, the same decoder
coefficients and grids, the same per-texture psnr, the same E in every round.
"""

import ctypes as C
import hashlib
import json
import os

import numpy as np
import pytest

from nntc_amd import abi, asset, chain, encoder, synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden(name):
    p = os.path.join(GOLDEN, name, "report.json")
    if not os.path.exists(p):
        pytest.skip(f"no golden fixture for {name}")
    return json.load(open(p))


def texture_sse(e, tex):
    """Each texture's sum of squared byte errors of the encoder's base decode against the source (the quantity the
    reference's psnr line is made of, src/image.h:387-407)."""
    h, w = tex[0].shape[:2]
    pl = e.desc.planes[0]
    rec = e.recon().reshape(pl.h0, pl.w0, -1)[:h, :w].astype(np.int64)
    out = []
    for t, src in enumerate(tex):
        d = rec[:, :, 3 * t: 3 * t + 3] - src.astype(np.int64)
        out.append(int((d * d).sum()))
    return out


def reference_psnr(g):
    """The reference's per-texture PSNR to full precision, from its integer error sums (tests/golden/make_golden.py
    --sse), or None for a fixture made without them."""
    if "sse" not in g:
        return None
    n = g["sse_samples"]
    return [10.0 * np.log10(255.0 * 255.0 / (s / n)) for s in g["sse"]]


def encode(name, tmp_path, **kw):
    spec = synth.WORKLOADS[name]
    tex = synth.workload(name)
    e = encoder.Encoder(tex, c0=spec["c0"], c1=spec["c1"], rounds=8, tol=0.0, quiet=1, **kw)
    rep = e.run()
    prefix = str(tmp_path / "m")
    e.write(prefix)
    return e, rep, prefix


@pytest.mark.parametrize("name", ["plain_cpu_4x256", "plain_gpu_4x1024", "precision_6x2048", "differentiation_4x4096"])
def test_encode_exact_matches_reference(name, tmp_path):
    """The exact form reproduces the reference CPU encoder: every round's E, the shipped E, the decoder's bits, the
    per-texture error sums and the .dds bytes. The 2048^2 workload covers six textures (nout 18) and four level-0
    channels (two BC5 files); the 4096^2 one three level-0 channels (a BC5 and a BC4 file) on ten planes."""
    g = golden(name)
    print(f"encode {name}", flush=True)
    e, rep, prefix = encode(name, tmp_path)
    print(f"  {rep.seconds:.3f} s, {rep.mpix / rep.seconds:.3f} Mpix/s", flush=True)
    if "sse" in g:
        assert texture_sse(e, synth.workload(name)) == g["sse"]
    assert rep.rounds_used == g["rounds_used"]
    for r, (ea, eb, ec) in enumerate(g["rounds"]):
        # The reference prints E to 7 significant digits; the exact form's E is the reference's double.
        for x, y in zip((rep.e_round[r][0], rep.e_round[r][1], rep.e_round[r][2]), (ea, eb, ec)):
            assert float(f"{x:.6e}") == y, f"round {r + 1}: {x!r} vs {y!r}"
    assert float(f"{rep.e_shipped:.9e}") == g["e_shipped"]
    for t, p in enumerate(g["psnr"]):
        assert round(rep.psnr[t], 2) == p
    w = np.array(e.state().w[: len(g["decoder"]["weights"])], dtype=np.float32)
    np.testing.assert_array_equal(w, np.array(g["decoder"]["weights"], dtype=np.float32))
    for suffix, digest in g["dds"].items():
        data = open(prefix + suffix, "rb").read()
        assert hashlib.sha256(data).hexdigest() == digest, f".dds {suffix} differs from the reference's"


def test_encode_repeatable(tmp_path):
    """Two runs on one resident device give the same bytes (the driver resets every piece of state)."""
    spec = synth.WORKLOADS["plain_cpu_4x256"]
    e = encoder.Encoder(synth.workload("plain_cpu_4x256"), c0=spec["c0"], c1=spec["c1"], rounds=3, tol=0.0, quiet=1)
    a = e.run()
    ka = [e.level1_indices(p).copy() for p in range(e.desc.nplanes)]
    ba = [e.level0_blocks(p, 0).copy() for p in range(e.desc.nplanes)]
    b = e.run()
    assert a.e_shipped == b.e_shipped
    for p in range(e.desc.nplanes):
        np.testing.assert_array_equal(ka[p], e.level1_indices(p))
        np.testing.assert_array_equal(ba[p], e.level0_blocks(p, 0))


@pytest.mark.parametrize("name", ["plain_cpu_4x256", "small_2x100x76"])
def test_fast_form_within_psnr_bar(name, tmp_path):
    """The fast form (tree reductions, tiled fp32 objective) against the exact form on the same workload: decoded
    PSNR within 0.05 dB on every texture (the bar the chosen form is measured at), the shipped E within 1 %, and the
    same bytes on a second run (no order depends on scheduling)."""
    spec = synth.WORKLOADS[name]
    tex = synth.workload(name)
    print(f"fast vs exact {name}", flush=True)
    ex = encoder.Encoder(tex, c0=spec["c0"], c1=spec["c1"], rounds=8, tol=0.0, quiet=1)
    rx = ex.run()
    fa = encoder.Encoder(tex, c0=spec["c0"], c1=spec["c1"], rounds=8, tol=0.0, quiet=1, form=1)
    r1 = fa.run()
    k1 = [fa.level0_blocks(p, 0).copy() for p in range(fa.desc.nplanes)]
    r2 = fa.run()
    k2 = [fa.level0_blocks(p, 0).copy() for p in range(fa.desc.nplanes)]
    for t in range(len(tex)):
        assert abs(r1.psnr[t] - rx.psnr[t]) <= 0.05, f"texture {t}: {r1.psnr[t]:.4f} vs {rx.psnr[t]:.4f} dB"
    assert abs(r1.e_shipped - rx.e_shipped) <= 0.01 * rx.e_shipped
    assert r1.e_shipped == r2.e_shipped
    for a, b in zip(k1, k2):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("form", [abi.FORM_FAST, abi.FORM_PLAIN], ids=["fast", "plain"])
@pytest.mark.parametrize("name", ["plain_gpu_4x1024", "precision_6x2048", "differentiation_4x4096"])
def test_form_psnr_against_reference(name, form, tmp_path):
    """The chosen (fast) and the plain form against the reference CPU encoder on the benchmark workloads: decoded PSNR
    within 0.05 dB of the reference's on every texture (its integer error sums, so the bar is read to full precision),
    and the written asset's level-1 values on the grid. These workloads have more than 256 chunks of 256 texels per
    level-0 plane, so the fast conjugate gradients fold their partials in their own launches (the path a small plane
    does not take); at 4096^2 the fast form's level-0 conjugate gradients run on the matrix-free operator. A second
    run must repeat the first's bits."""
    g = golden(name)
    ref = reference_psnr(g)
    if ref is None:
        pytest.skip(f"{name}: the fixture has no error sums")
    print(f"form {form} {name}", flush=True)
    e, rep, prefix = encode(name, tmp_path, form=form)
    tex = synth.workload(name)
    sse = texture_sse(e, tex)
    n = tex[0].size
    ours = [10.0 * np.log10(255.0 * 255.0 / (x / n)) for x in sse]
    gap = [ours[t] - ref[t] for t in range(len(tex))]
    print(f"  {rep.seconds:.3f} s; psnr gap per texture (dB): {', '.join(f'{x:+.5f}' for x in gap)}", flush=True)
    for t in range(len(tex)):
        assert abs(rep.psnr[t] - ours[t]) < 1e-9
        assert abs(gap[t]) <= 0.05, f"texture {t}: {ours[t]:.5f} vs the reference's {ref[t]:.5f} dB"
    assert rep.rounds_used == g["rounds_used"]
    if tex[0].shape[0] <= 2048:
        # The written asset read back by the reader that shares no code with the writer (nntc_amd/asset.py): its
        # float64 decode within one step of the encoder's, its PSNR the report's to 0.01 dB.
        rgb = asset.decode(asset.load(prefix))
        pl = e.desc.planes[0]
        rec = e.recon().reshape(pl.h0, pl.w0, -1)
        assert np.abs(rgb.astype(np.int64) - rec.astype(np.int64)).max() <= 1
        for t, p in enumerate(asset.psnr(rgb, tex)):
            assert abs(p - rep.psnr[t]) <= 0.01, f"texture {t}: reader {p:.4f} vs report {rep.psnr[t]:.4f} dB"
    # A second run on the resident device gives the same bits: no reduction order depends on scheduling.
    blocks = [e.level0_blocks(p, 0).copy() for p in range(e.desc.nplanes)]
    again = e.run()
    assert again.e_shipped == rep.e_shipped
    for p in range(e.desc.nplanes):
        np.testing.assert_array_equal(blocks[p], e.level0_blocks(p, 0), err_msg=f"plane {p} on the second run")


def _source_planes(lib, ctx, desc):
    f = lib.nntc_hip_source_download
    f.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_size_t)]
    out = []
    for p in range(desc.nplanes):
        pl = desc.planes[p]
        a = np.zeros(pl.w0 * pl.h0 * desc.nout, np.float32)
        assert f(ctx, p, a.ctypes.data_as(C.POINTER(C.c_float)), a.size, None) == 0
        out.append(a)
    return out


@pytest.mark.parametrize("name", ["small_2x100x76", "plain_cpu_4x256"])
def test_device_resident_source_chain(name, hip_arm, tmp_path):
    """nntc_hip_device_create_rgb8 builds, from the textures' bytes in device memory, the source chain the host builds
    (the packed, edge-padded base and every box-filtered plane, bit for bit), and an exact-form encode from
    device-resident textures writes the reference's .dds bytes."""
    spec = synth.WORKLOADS[name]
    tex = synth.workload(name)
    print(f"device-resident source {name}", flush=True)
    h, w = tex[0].shape[:2]
    planes, ch = chain.source_chain(tex)
    desc = abi.make_desc(len(tex), spec["c0"], spec["c1"], planes, chain.output_weights(len(tex)))
    host = abi.Session(hip_arm, desc, [c.reshape(-1) for c in ch])
    dev = [abi.DeviceBytes(t) for t in tex]
    lib = hip_arm.lib
    lib.nntc_hip_device_create_rgb8.restype = C.c_void_p
    lib.nntc_hip_device_create_rgb8.argtypes = [C.POINTER(abi.ModelDesc), C.POINTER(C.c_void_p), C.c_int, C.c_int]
    ptrs = (C.c_void_p * len(dev))(*[C.c_void_p(d.address) for d in dev])
    ctx = lib.nntc_hip_device_create_rgb8(C.byref(desc), ptrs, w, h)
    assert ctx, lib.nntc_hip_last_error()
    try:
        for p, (a, b) in enumerate(zip(_source_planes(lib, host.ctx, desc), _source_planes(lib, ctx, desc))):
            np.testing.assert_array_equal(a, b, err_msg=f"source plane {p}")
            np.testing.assert_array_equal(a, ch[p].reshape(-1), err_msg=f"source plane {p} against the host chain")
    finally:
        hip_arm.fn["device_destroy"](ctx)
    host.close()
    e = encoder.Encoder(device=(w, h, [d.address for d in dev]), c0=spec["c0"], c1=spec["c1"], rounds=8, tol=0.0,
                        quiet=1)
    rep = e.run()
    prefix = str(tmp_path / "m")
    e.write(prefix)
    g = golden(name)
    for suffix, digest in g["dds"].items():
        assert hashlib.sha256(open(prefix + suffix, "rb").read()).hexdigest() == digest
    assert float(f"{rep.e_shipped:.9e}") == g["e_shipped"]
    e.close()


def test_device_resident_channel_copies(hip_arm, tmp_path):
    """The box init takes exact copies of earlier source channels last (src/main.cpp:1938-2027). With texture 1 a copy
    of texture 0, the device's copy test finds channels 3-5 copies of 0-2, and the encode from device-resident textures
    (which fetches the chain for the reordered init) equals the encode from host textures."""
    base = synth.make_material(2, (64, 40), 32, 3, 1.0, config=9)
    tex = [base[0], base[0].copy(), base[1]]
    h, w = tex[0].shape[:2]
    dev = [abi.DeviceBytes(t) for t in tex]
    kw = dict(c0=2, c1=4, rounds=4, tol=0.0, quiet=1)
    a = encoder.Encoder(tex, **kw)
    ra = a.run()
    b = encoder.Encoder(device=(w, h, [d.address for d in dev]), **kw)
    rb = b.run()
    assert ra.e_shipped == rb.e_shipped
    for p in range(a.desc.nplanes):
        np.testing.assert_array_equal(a.level1_indices(p), b.level1_indices(p))
        np.testing.assert_array_equal(a.level0_blocks(p, 0), b.level0_blocks(p, 0))
    lib = hip_arm.lib
    planes, ch = chain.source_chain(tex)
    desc = abi.make_desc(len(tex), 2, 4, planes, chain.output_weights(len(tex)))
    s = abi.Session(hip_arm, desc, [c.reshape(-1) for c in ch])
    m = (C.c_uint32 * 18)()
    lib.nntc_hip_source_copies.argtypes = [C.c_void_p, C.POINTER(C.c_uint32)]
    assert lib.nntc_hip_source_copies(s.ctx, m) == 0
    assert [m[i] for i in range(9)] == [0, 0, 0, 1, 2, 4, 0, 0, 0]


@pytest.mark.parametrize("name,bits0,bc0", [("small_2x100x76", [4], 1), ("plain_cpu_4x256", [3], 1),
                                             ("plain_cpu_4x256", [4, 2], 0), ("plain_gpu_4x1024", [4], 1)],
                         ids=["small-4bit-bc", "256-3bit-bc", "256-4+2bit-uncompressed", "1024-4bit-bc"])
def test_encode_palette_matches_reference(name, bits0, bc0, ref_arm, tmp_path):
    """--l0 palette through the driver on the gfx950 arm (exact form) against the same driver over the reference CPU
    backend: every round's E, the shipped E (of the packed plane when the pack is lossy), the per-texture psnr and
    the .dds bytes."""
    spec = synth.WORKLOADS[name]
    tex = synth.workload(name)
    print(f"palette encode {name} bits {bits0} bc0 {bc0}", flush=True)
    kw = dict(c0=spec["c0"], c1=spec["c1"], rounds=6, tol=0.0, quiet=1, l0_palette=1, bits0=bits0, bc0=bc0)
    h = encoder.Encoder(tex, **kw)
    rh = h.run()
    h.write(str(tmp_path / "h"))
    r = encoder.Encoder(tex, arm_lib=ref_arm.path, arm_prefix="nntc_ref_", **kw)
    rr = r.run()
    r.write(str(tmp_path / "r"))
    print(f"  {rh.seconds:.3f} s on the arm, {rr.seconds:.3f} s on the reference backend", flush=True)
    assert rh.rounds_used == rr.rounds_used
    for i in range(rh.rounds_used):
        assert list(rh.e_round[i]) == list(rr.e_round[i]), f"round {i + 1}"
    assert rh.e_shipped == rr.e_shipped
    assert rh.pack_psnr == rr.pack_psnr
    assert [rh.psnr[t] for t in range(len(tex))] == [rr.psnr[t] for t in range(len(tex))]
    files = sorted(f for f in os.listdir(tmp_path) if f.startswith("r_"))
    assert any(f.endswith(".dds") for f in files)
    for f in files:
        a = open(tmp_path / f, "rb").read()
        b = open(tmp_path / ("h_" + f[2:]), "rb").read()
        if f.endswith(".json"):
            a = a.replace(b'"r_lat', b'"h_lat')   # the descriptor names its own prefix's files
        assert a == b, f"{f[1:]} differs between the arm and the reference backend"


@pytest.mark.parametrize("bits0", [[3], [4]], ids=["3bit", "4bit"])
def test_encode_palette_fast_form_within_psnr_bar(bits0, hip_arm, tmp_path):
    """--l0 palette under the fast form against the exact form: decoded psnr within 0.05 dB on every texture and the
    shipped E within 0.5 %. The palette seed leaves each level-0 channel at one nonzero value until its turn, which
    makes the decoder's features collinear; the fast form's split-bf16 Gram cannot settle that system and hands those
    solves to the fp64 Gram (counted by nntc_hip_decoder_fast_fallbacks), which is what keeps the run on the exact
    form's track."""
    spec = synth.WORKLOADS["plain_gpu_4x1024"]
    tex = synth.workload("plain_gpu_4x1024")
    kw = dict(c0=spec["c0"], c1=spec["c1"], rounds=6, tol=0.0, quiet=1, l0_palette=1, bits0=bits0)
    counter = hip_arm.lib.nntc_hip_decoder_fast_fallbacks
    counter.restype = C.c_longlong
    rx = encoder.Encoder(tex, **kw).run()
    before = counter()
    rf = encoder.Encoder(tex, form=1, **kw).run()
    assert counter() > before
    for t in range(len(tex)):
        assert abs(rf.psnr[t] - rx.psnr[t]) <= 0.05, f"texture {t}: {rf.psnr[t]:.4f} vs {rx.psnr[t]:.4f} dB"
    assert abs(rf.e_shipped - rx.e_shipped) <= 0.005 * rx.e_shipped


def test_fast_decoder_solve_keeps_the_split_gram_on_a_full_rank_state(hip_arm):
    """The referral is for degenerate systems only: a --l0 bc8 encode under the fast form settles every block (a) on
    the split-bf16 Gram (no call redone in fp64)."""
    spec = synth.WORKLOADS["plain_gpu_4x1024"]
    counter = hip_arm.lib.nntc_hip_decoder_fast_fallbacks
    counter.restype = C.c_longlong
    before = counter()
    encoder.Encoder(synth.workload("plain_gpu_4x1024"), c0=spec["c0"], c1=spec["c1"], rounds=6, tol=0.0, quiet=1,
                    form=1).run()
    assert counter() == before


def test_recycled_memory_gives_the_same_bits(hip_arm):
    """A destroyed context's device buffers and pinned areas go to the process's cache and the next context takes them
    (holding the first encode's data): the second encode's outputs are the first's bits, and releasing the cache
    returns what the destroy gave back."""
    rel = hip_arm.lib.nntc_hip_release_cached_memory
    rel.restype = C.c_size_t
    rel()
    spec = synth.WORKLOADS["small_2x100x76"]
    tex = synth.workload("small_2x100x76")
    out = []
    for run in range(2):
        e = encoder.Encoder(tex, c0=spec["c0"], c1=spec["c1"], rounds=4, tol=0.0, quiet=1, form=abi.FORM_FAST)
        r = e.run()
        blocks = [e.level0_blocks(p, f).tobytes() for p in range(e.desc.nplanes) for f in range(2)]
        out.append((r.e_shipped, hashlib.sha256(b"".join(blocks)).hexdigest()))
        e.close()
    assert out[0] == out[1]
    assert rel() > 0
    assert rel() == 0
