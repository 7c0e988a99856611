"""CPU-side checks of the boundary and the host path (no GPU calls).

- the built libraries export every entry point include/*.h declares;
- the encode driver (include/nntc_encode.h) run over the reference CPU backend reproduces the reference binary's
  assets byte for byte (the driver and the writer are thereby pinned independently of the gfx950 arm);
- the PNG path and the array path give the same encode;
- the synthetic generator is deterministic;
- bench.py refuses a multi-GPU run with a message.
"""

import ctypes
import glob
import hashlib
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from nntc_amd import abi, encoder, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libnntc_ref.so")


def header_symbols(name):
    return abi.declared_symbols(os.path.join(ROOT, "include", name))


@pytest.mark.parametrize("header,lib", [("nntc_hip.h", abi.HIP_LIB), ("nntc_encode.h", encoder.HOST_LIB)])
def test_library_exports_every_declared_symbol(header, lib):
    names = header_symbols(header)
    assert len(names) > 10
    so = ctypes.CDLL(lib, mode=ctypes.RTLD_GLOBAL)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, f"{os.path.basename(lib)} lacks {missing}"


def test_c_abi_struct_layouts_match_header():
    """The ctypes mirrors have the sizes the C compiler gives the header's structs."""
    src = os.path.join(ROOT, "_build", "layout_probe.c")
    os.makedirs(os.path.dirname(src), exist_ok=True)
    structs = ["nntc_model_desc", "nntc_model_state", "nntc_level1_report", "nntc_objective", "nntc_init0_report",
               "nntc_level1_init_report", "nntc_bc_refine_report", "nntc_device_info", "nntc_encode_options",
               "nntc_encode_report"]
    with open(src, "w") as f:
        f.write('#include <stdio.h>\n#include "nntc_encode.h"\nint main(void){\n')
        for s in structs:
            f.write(f'printf("%zu\\n", sizeof({s}));\n')
        f.write("return 0;}\n")
    exe = src[:-2]
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    sizes = [int(x) for x in subprocess.run([exe], stdout=subprocess.PIPE, text=True).stdout.split()]
    mirrors = [abi.ModelDesc, abi.ModelState, abi.Level1Report, abi.Objective, abi.Init0Report,
               abi.Level1InitReport, abi.BcRefineReport, abi.DeviceInfo, encoder.Options, encoder.Report]
    for s, m, n in zip(structs, mirrors, sizes):
        assert ctypes.sizeof(m) == n, f"{s}: C {n} bytes, ctypes {ctypes.sizeof(m)}"


def ref_encode(name, tmp_path, rounds):
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref is not built")
    ctypes.CDLL(REF_LIB, mode=ctypes.RTLD_GLOBAL).nntc_ref_set_threads(min(8, os.cpu_count() or 1))
    spec = synth.WORKLOADS[name]
    e = encoder.Encoder(synth.workload(name), arm_lib=REF_LIB, arm_prefix="nntc_ref_", c0=spec["c0"],
                        c1=spec["c1"], rounds=rounds, tol=0.0, quiet=1)
    rep = e.run()
    e.write(str(tmp_path / "m"))
    return e, rep


@pytest.mark.parametrize("name,rounds", [("smoke_1x64x40", 4), ("small_2x100x76", 8)])
def test_driver_over_reference_reproduces_reference_binary(name, rounds, tmp_path):
    g = json.load(open(os.path.join(GOLDEN, name, "report.json")))
    e, rep = ref_encode(name, tmp_path, rounds)
    assert f"{rep.e_shipped:.9e}" == f"{g['e_shipped']:.9e}"
    assert [round(rep.psnr[t], 2) for t in range(len(g["psnr"]))] == g["psnr"]
    assert rep.rounds_used == g["rounds_used"]
    for suffix, digest in g["dds"].items():
        data = open(str(tmp_path / "m") + suffix, "rb").read()
        assert hashlib.sha256(data).hexdigest() == digest, suffix
    js = json.load(open(str(tmp_path / "m") + "_nntc.json"))
    np.testing.assert_array_equal(np.array(js["decoder"]["layers"][0]["weights"], np.float32),
                                  np.array(g["decoder"]["weights"], np.float32))
    for lvl in range(2):
        assert js["textures"][lvl]["dequantise"]["lo"] == g["grids"][lvl]["lo"]
        assert js["textures"][lvl]["dequantise"]["hi"] == g["grids"][lvl]["hi"]


def test_png_path_matches_array_path(tmp_path):
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref is not built")
    ctypes.CDLL(REF_LIB, mode=ctypes.RTLD_GLOBAL).nntc_ref_set_threads(4)
    tex = synth.workload("smoke_1x64x40")
    paths = synth.write_pngs(tex, str(tmp_path / "in"))
    kw = dict(arm_lib=REF_LIB, arm_prefix="nntc_ref_", c0=1, c1=2, rounds=2, tol=0.0, quiet=1)
    a = encoder.Encoder(tex, **kw).run()
    b = encoder.Encoder(files=paths, **kw).run()
    assert a.e_shipped == b.e_shipped


def test_synthetic_generator_is_deterministic():
    a = synth.workload("smoke_1x64x40")
    b = synth.workload("smoke_1x64x40")
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    h = hashlib.sha256(b"".join(x.tobytes() for x in synth.workload("plain_cpu_4x256"))).hexdigest()
    g = json.load(open(os.path.join(GOLDEN, "synth_digests.json")))
    assert h == g["plain_cpu_4x256"]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True)
    assert r.returncode != 0
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("inits", [("pca", "texture"), ("box", "luma")], ids=["pca-texture", "box-luma"])
@pytest.mark.parametrize("name", ["smoke_1x64x40", "small_2x100x76"])
def test_driver_inits_match_reference_binary(name, inits, tmp_path):
    """--init pca with --init0 texture, and --init0 luma: the driver over the reference CPU backend writes the bytes
    the reference encoder itself writes for the same PNG files and options (both run here, live)."""
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "nntc_encode")
    if not (os.path.exists(REF_LIB) and os.path.exists(ref_bin)):
        pytest.skip("oracle/_ref is not built")
    spec = synth.WORKLOADS[name]
    paths = synth.write_pngs(synth.workload(name), str(tmp_path / "m"))
    out = str(tmp_path / "ref") + os.sep
    os.makedirs(out)
    cmd = [ref_bin, *paths, "--c0", str(spec["c0"]), "--c1", str(spec["c1"]), "--rounds", "4", "--tol", "0",
           "--png", "0", "--backend", "cpu", "-j", "8", "--quiet", "--init", inits[0], "--init0", inits[1], "-o", out]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    ctypes.CDLL(REF_LIB, mode=ctypes.RTLD_GLOBAL).nntc_ref_set_threads(8)
    e = encoder.Encoder(files=paths, arm_lib=REF_LIB, arm_prefix="nntc_ref_", c0=spec["c0"], c1=spec["c1"], rounds=4,
                        tol=0.0, quiet=1, init1_pca=int(inits[0] == "pca"), init0_texture=int(inits[1] == "texture"),
                        init0_luma=int(inits[1] == "luma"))
    e.run()
    e.write(str(tmp_path / "drv"))
    ref_files = sorted(f for f in os.listdir(out) if f.endswith(".dds"))
    assert ref_files
    for f in ref_files:
        suffix = f[len("m_t0"):]
        a = hashlib.sha256(open(os.path.join(out, f), "rb").read()).hexdigest()
        b = hashlib.sha256(open(str(tmp_path / "drv") + suffix, "rb").read()).hexdigest()
        assert a == b, f"{suffix} differs from the reference's"


@pytest.mark.parametrize("bits0,bc0", [("3", 1), ("4", 1), ("2", 0)], ids=["3bit-bc", "4bit-bc-lossy", "2bit-uncompressed"])
@pytest.mark.parametrize("name", ["smoke_1x64x40", "small_2x100x76"])
def test_driver_palette_matches_reference_binary(name, bits0, bc0, tmp_path):
    """--l0 palette: the driver over the reference CPU backend, with the writer's own host pack of the indices (lossless
    BC4 / BC5 at 1-3 bits, the nearest-entry pack at 4, the replicated bytes under --bc0 0), writes the .dds bytes the
    reference encoder writes, and the same palette in the descriptor."""
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "nntc_encode")
    if not (os.path.exists(REF_LIB) and os.path.exists(ref_bin)):
        pytest.skip("oracle/_ref is not built")
    spec = synth.WORKLOADS[name]
    paths = synth.write_pngs(synth.workload(name), str(tmp_path / "m"))
    out = str(tmp_path / "ref") + os.sep
    os.makedirs(out)
    cmd = [ref_bin, *paths, "--c0", str(spec["c0"]), "--c1", str(spec["c1"]), "--rounds", "4", "--tol", "0",
           "--png", "0", "--backend", "cpu", "-j", "8", "--l0", "palette", "--bits0", bits0, "--bc0",
           str(bc0), "-o", out]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-2000:]
    ctypes.CDLL(REF_LIB, mode=ctypes.RTLD_GLOBAL).nntc_ref_set_threads(8)
    e = encoder.Encoder(files=paths, arm_lib=REF_LIB, arm_prefix="nntc_ref_", c0=spec["c0"], c1=spec["c1"], rounds=4,
                        tol=0.0, quiet=1, l0_palette=1, bits0=[int(bits0)], bc0=bc0)
    rep = e.run()
    e.write(str(tmp_path / "drv"))
    ref_files = sorted(f for f in os.listdir(out) if f.endswith(".dds"))
    assert ref_files
    for f in ref_files:
        suffix = f[len("m_t0"):]
        a = hashlib.sha256(open(os.path.join(out, f), "rb").read()).hexdigest()
        b = hashlib.sha256(open(str(tmp_path / "drv") + suffix, "rb").read()).hexdigest()
        assert a == b, f"{suffix} differs from the reference's"
    jref = json.load(open(os.path.join(out, [f for f in os.listdir(out) if f.endswith("_nntc.json")][0])))
    jdrv = json.load(open(str(tmp_path / "drv") + "_nntc.json"))
    for key in ("palette", "lo", "hi", "levels", "kind"):
        assert jdrv["textures"][0]["dequantise"][key] == jref["textures"][0]["dequantise"][key], key
    for key in ("dxgi_format_id", "channels_stored", "bits_per_channel", "mip_sizes"):
        assert jdrv["textures"][0][key] == jref["textures"][0][key], key
    assert (rep.pack_psnr == 100.0) == (int(bits0) <= 3 or bc0 == 0)
    # The report: at 4 bits block-compressed it is of the packed plane, decoded back from the blocks.
    shipped = [ln for ln in r.stdout.splitlines() if ln.strip().startswith("E shipped")]
    assert shipped and shipped[0].split()[2] == f"{rep.e_shipped:.9e}"
    psnr = [ln.split()[3] for ln in r.stdout.splitlines() if ln.strip().startswith("psnr         texture")]
    assert psnr == [f"{rep.psnr[t]:.2f}" for t in range(len(psnr))] and psnr


@pytest.mark.parametrize("args,message", [
    (["--l0", "palette", "--bc-refine", "2"], "belong to --l0 bc8"),
    (["--bits0", "3"], "apply under --l0 palette only"),
    (["--bc0", "0"], "apply under --l0 palette only"),
    (["--l0", "palette", "--c0", "3", "--bits0", "3,3"], "gives 2 values for 3 level-0 channels"),
    (["--l0", "rgb"], "must be bc8 or palette"),
    (["--mip-filter", "mitchell"], "is not an option of this encoder"),
])
def test_cli_refuses_what_it_does_not_build(args, message, tmp_path):
    """The command line refuses, by name and before touching a device, the options that do not apply (--bits0 and
    --bc0 0 under --l0 bc8, the BC pack flags under --l0 palette, src/main.cpp:1310-1356) and the reference options
    this encoder does not build."""
    exe = os.path.join(ROOT, "nntc_amd", "bin", "nntc_hip_encode")
    if not os.path.exists(exe):
        pytest.skip("the CLI is not built")
    png = synth.write_pngs(synth.workload("smoke_1x64x40"), str(tmp_path / "m"))[0]
    r = subprocess.run([exe, png, *args], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode != 0
    assert message in r.stderr, r.stderr
