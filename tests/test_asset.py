"""The written asset read back by a reader that shares no code with the writer (nntc_amd/asset.py).

The reference's bars for its own reader (tests/run_checks.py:637-685 over tools/dds_decode.py): the reader's decode is
within one step of the encoder's own decode, and its per-texture PSNR equals the encoder's report to 0.01 dB. They are
checked here on the reference encoder's own asset (the fixture under tests/golden/plain_cpu_4x256/), on assets the
driver writes over the reference CPU backend (one BC4 file, and a BC5 + BC4 pair at three level-0 channels), and the
descriptor the driver writes is compared key by key with the reference's.
"""

import ctypes
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from nntc_amd import asset, encoder, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libnntc_ref.so")
# Prose the format marks informational (docs/FORMAT.md section 4): free text, not compared.
PROSE = {"sampling", "features", "texel", "note"}


def bc4_block(r0, r1, sel):
    bits = 0
    for t, s in enumerate(sel):
        bits |= int(s) << (3 * t)
    return np.array([r0, r1] + [(bits >> (8 * i)) & 255 for i in range(6)], dtype=np.uint8)


def test_bc4_palettes_closed_form():
    """Both BC4 modes against their closed forms, texel order row-major with texel 0 in the lowest bits."""
    sel = list(range(8)) * 2
    blk = bc4_block(200, 10, sel)
    v = asset.decode_bc(blk.tobytes(), 4, 4, 8)[:, :, 0].reshape(16)
    want = [200, 10] + [((7 - i) * 200 + i * 10) / 7 for i in range(1, 7)]
    np.testing.assert_allclose(v, want * 2)
    blk = bc4_block(10, 200, sel)
    v = asset.decode_bc(blk.tobytes(), 4, 4, 8)[:, :, 0].reshape(16)
    want = [10, 200] + [((5 - i) * 10 + i * 200) / 5 for i in range(1, 5)] + [0, 255]
    np.testing.assert_allclose(v, want * 2)
    # At three bits the endpoints 255 and 0 make the palette the grid k * 255 / 7 (docs/FORMAT.md section 7).
    blk = bc4_block(255, 0, list(range(8)) + [0] * 8)
    v = asset.decode_bc(blk.tobytes(), 4, 4, 8)[:, :, 0].reshape(16)
    np.testing.assert_allclose(v[:8], [255, 0] + [k * 255 / 7 for k in range(6, 0, -1)])


def test_bc5_is_two_bc4_blocks_and_crops_to_the_extent():
    a = bc4_block(255, 0, [1] * 16)
    b = bc4_block(0, 255, [0] * 16)
    data = np.concatenate([a, b]).tobytes() * 2          # two blocks side by side: an extent of 6 x 3
    v = asset.decode_bc(data, 6, 3, 16)
    assert v.shape == (3, 6, 2)
    assert (v[:, :, 0] == 0).all() and (v[:, :, 1] == 0).all()


def test_bilinear_taps_clamp_and_centre():
    """At a plane's own extent every centre lands on its texel; at a quarter, taps clamp at the edges."""
    plane = np.arange(12, dtype=np.float64).reshape(3, 4, 1)
    np.testing.assert_array_equal(asset.sample(plane, 4, 3), plane)
    x0, x1, f = asset._taps(8, 2)
    assert x0[0] == 0 and x1[0] == 0 and x0[-1] == 1 and x1[-1] == 1   # both ends clamp
    # x = (p + 0.5) * 2 / 8 - 0.5: the interior pixels fall between texels 0 and 1 at eighths.
    np.testing.assert_allclose(f[2:6], [0.125, 0.375, 0.625, 0.875])
    assert (x0[2:6] == 0).all() and (x1[2:6] == 1).all()


def reference_asset(tmp_path):
    js = json.load(open(os.path.join(GOLDEN, "plain_cpu_4x256", "ref_nntc.json")))
    shutil.copy(os.path.join(GOLDEN, "plain_cpu_4x256", "ref_lat0.dds"), tmp_path / js["textures"][0]["file"])
    shutil.copy(os.path.join(GOLDEN, "plain_cpu_4x256", "ref_lat1.dds"), tmp_path / js["textures"][1]["file"])
    shutil.copy(os.path.join(GOLDEN, "plain_cpu_4x256", "ref_nntc.json"), tmp_path / "m_nntc.json")
    return str(tmp_path / "m_nntc.json")


def test_reference_asset_decodes_to_its_reported_psnr(tmp_path):
    """The reference encoder's own asset (4 x 256^2, BC5 level 0 with six mips): the reader's per-texture PSNR is
    the reference's report to 0.01 dB, and every mip decodes at its stored size."""
    a = asset.load(reference_asset(tmp_path))
    rep = json.load(open(os.path.join(GOLDEN, "plain_cpu_4x256", "report.json")))
    tex = synth.workload("plain_cpu_4x256")
    rgb = asset.decode(a)
    ps = asset.psnr(rgb, tex)
    for t, p in enumerate(rep["psnr"]):
        assert abs(ps[t] - p) <= 0.01, f"texture {t}: {ps[t]:.4f} vs {p}"
    # The integer error sums differ from the reference's only by the rounding steps of its float32 decode.
    for t, s in enumerate(rep["sse"]):
        assert abs(asset.psnr(rgb, tex)[t] - 10 * np.log10(255.0 ** 2 / (s / rep["sse_samples"]))) < 0.002
    for m, (w, h) in enumerate(a.meta["textures"][0]["mip_sizes"]):
        assert asset.decode(a, m).shape == (h, w, 12)


def test_reader_refuses_what_the_format_does_not_define(tmp_path):
    path = reference_asset(tmp_path)
    js = json.load(open(path))
    for mutate in (lambda j: j["decoder"].update(terms="b a sc"),
                   lambda j: j["decoder"].update(hidden=[8]),
                   lambda j: j.update(format="other"),
                   lambda j: j["textures"][0].update(dxgi_format_id=80),
                   lambda j: j["decoder"].update(nin=18)):
        bad = json.loads(json.dumps(js))
        mutate(bad)
        p = str(tmp_path / "bad_nntc.json")
        json.dump(bad, open(p, "w"))
        with pytest.raises(asset.AssetError):
            asset.load(p)


def ref_encode(name, rounds, files=None, **opts):
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref is not built")
    ctypes.CDLL(REF_LIB, mode=ctypes.RTLD_GLOBAL).nntc_ref_set_threads(min(8, os.cpu_count() or 1))
    spec = synth.WORKLOADS[name]
    kw = dict(arm_lib=REF_LIB, arm_prefix="nntc_ref_", c0=spec["c0"], c1=spec["c1"], rounds=rounds, tol=0.0, quiet=1,
              **opts)
    e = encoder.Encoder(files=files, **kw) if files else encoder.Encoder(synth.workload(name), **kw)
    return e, e.run()


@pytest.mark.parametrize("name,rounds,opts", [
    ("smoke_1x64x40", 4, {}), ("small_2x100x76", 4, {}),
    ("small_2x100x76", 4, dict(l0_palette=1, bits0=[3])), ("small_2x100x76", 4, dict(l0_palette=1, bits0=[4])),
    ("smoke_1x64x40", 4, dict(l0_palette=1, bits0=[2], bc0=0))],
    ids=["bc8-bc4", "bc8-bc5+bc4", "palette-3bit", "palette-4bit-lossy", "palette-2bit-uncompressed"])
def test_written_asset_reads_back(name, rounds, opts, tmp_path):
    """The driver's asset over the reference CPU backend (--l0 bc8: one BC4 file at one level-0 channel, a BC5 + BC4
    pair at three; --l0 palette: the indices block-compressed at 3 and 4 bits and stored as bytes at 2): the reader's
    decode is within one step of the encoder's own, its PSNR the report's to 0.01 dB."""
    e, rep = ref_encode(name, rounds, **opts)
    e.write(str(tmp_path / "m"))
    a = asset.load(str(tmp_path / "m"))
    rgb = asset.decode(a)
    pl = e.desc.planes[0]
    rec = e.recon().reshape(pl.h0, pl.w0, -1)
    assert np.abs(rgb.astype(np.int64) - rec.astype(np.int64)).max() <= 1
    ps = asset.psnr(rgb, synth.workload(name))
    for t in range(len(ps)):
        assert abs(ps[t] - rep.psnr[t]) <= 0.01


def strip_prose(x):
    if isinstance(x, dict):
        return {k: strip_prose(v) for k, v in x.items() if k not in PROSE}
    if isinstance(x, list):
        return [strip_prose(v) for v in x]
    return x


def test_descriptor_matches_the_reference(tmp_path):
    """The driver's _nntc.json over the reference CPU backend, from the same PNG files the reference encoder read,
    equals the reference's own descriptor key by key (prose aside, and the .dds names, which follow the prefix)."""
    paths = synth.write_pngs(synth.workload("plain_cpu_4x256"), str(tmp_path / "m"))
    e, rep = ref_encode("plain_cpu_4x256", 8, files=paths)
    e.write(str(tmp_path / "out"))
    ours = strip_prose(json.load(open(str(tmp_path / "out") + "_nntc.json")))
    ref = strip_prose(json.load(open(os.path.join(GOLDEN, "plain_cpu_4x256", "ref_nntc.json"))))
    for lvl in range(2):
        assert ours["textures"][lvl].pop("file") == f"out_lat{lvl}.dds"
        assert ref["textures"][lvl].pop("file") == f"m_t0_lat{lvl}.dds"
    assert ours == ref


@pytest.mark.parametrize("bits0,bc0", [("4", 1), ("2", 0)], ids=["4bit-bc", "2bit-uncompressed"])
def test_palette_descriptor_matches_the_reference(bits0, bc0, tmp_path):
    """--l0 palette's _nntc.json from the driver over the reference CPU backend equals the reference encoder's own,
    run here on the same PNG files, key by key (prose aside, and the .dds names, which follow the prefix)."""
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "nntc_encode")
    if not (os.path.exists(REF_LIB) and os.path.exists(ref_bin)):
        pytest.skip("oracle/_ref is not built")
    paths = synth.write_pngs(synth.workload("small_2x100x76"), str(tmp_path / "m"))
    out = str(tmp_path / "ref") + os.sep
    os.makedirs(out)
    spec = synth.WORKLOADS["small_2x100x76"]
    r = subprocess.run([ref_bin, *paths, "--c0", str(spec["c0"]), "--c1", str(spec["c1"]), "--rounds", "4", "--tol",
                        "0", "--png", "0", "--backend", "cpu", "-j", "8", "--quiet", "--l0", "palette", "--bits0",
                        bits0, "--bc0", str(bc0), "-o", out], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    e, _ = ref_encode("small_2x100x76", 4, files=paths, l0_palette=1, bits0=[int(bits0)], bc0=bc0)
    e.write(str(tmp_path / "out"))
    ours = strip_prose(json.load(open(str(tmp_path / "out") + "_nntc.json")))
    ref = strip_prose(json.load(open(os.path.join(out, "m_t0_nntc.json"))))
    for lvl in range(2):
        for key in ("file", "files"):
            a, b = ours["textures"][lvl].pop(key, None), ref["textures"][lvl].pop(key, None)
            assert json.dumps(a).replace("out_lat", "m_t0_lat") == json.dumps(b)
    assert ours == ref
    # The reader takes the asset, and refuses it once a palette entry is not what its byte decodes to.
    asset.load(str(tmp_path / "out"))
    bad = json.load(open(str(tmp_path / "out") + "_nntc.json"))
    bad["textures"][0]["dequantise"]["palette"][0][1] += 0.01
    json.dump(bad, open(str(tmp_path / "out") + "_nntc.json", "w"))
    with pytest.raises(asset.AssetError):
        asset.load(str(tmp_path / "out"))
