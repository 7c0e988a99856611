"""Golden fixtures: the reference encoder's outputs on the seeded synthetic workloads.

Runs oracle/_ref/nntc_encode (the reference CPU encoder, built from /root/reference by oracle/Makefile) on a workload
from nntc_amd.synth and records, under tests/golden/<workload>/:
    report.json   per-texture psnr, centre psnr, E shipped, every round's three E values, the decoder (W, b) and both
                  grids from the asset's JSON, the sha256 of each .dds, and the reference's own wall clock and threads;
                  with --sse, each texture's integer sum of squared byte errors at the base level (from the
                  reference's own reconstruction PNGs, written by its --png 1), so PSNR is known to full precision
    *.dds         the level-0 and level-1 files themselves (small workloads only)

    python tests/golden/make_golden.py plain_cpu_4x256 [--rounds 8] [--threads 8] [--sse] [--keep-dir DIR]
"""

from __future__ import annotations

import argparse
import hashlib
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nntc_amd import synth  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "nntc_encode")


def parse_log(log: str) -> dict:
    out = {"psnr": [], "rounds": []}
    for m in re.finditer(r"psnr\s+texture (\d+)\s+([\d.]+) dB", log):
        out["psnr"].append(float(m.group(2)))
    m = re.search(r"centre psnr\s+([\d.]+) dB", log)
    out["centre_psnr"] = float(m.group(1))
    m = re.search(r"E shipped\s+([\d.e+-]+)", log)
    out["e_shipped"] = float(m.group(1))
    for m in re.finditer(r"^round\s+(\d+)\s+E ([\d.e+-]+) ([\d.e+-]+) ([\d.e+-]+)", log, flags=re.M):
        out["rounds"].append([float(m.group(2)), float(m.group(3)), float(m.group(4))])
    m = re.search(r"time\s+([\d.]+) s total", log)
    out["seconds"] = float(m.group(1))
    m = re.search(r"rounds\s+(\d+) \(stopped on (\w+)\)", log)
    out["rounds_used"] = int(m.group(1))
    out["stopped_on"] = m.group(2)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload")
    ap.add_argument("--rounds", type=int, default=8)
    ap.add_argument("--threads", type=int, default=8)
    ap.add_argument("--keep-dds", action="store_true", help="store the .dds files themselves")
    ap.add_argument("--keep-json", action="store_true", help="store the asset's _nntc.json itself")
    ap.add_argument("--sse", action="store_true",
                    help="run with --png 1 and record each texture's integer squared-error sum at the base level")
    ap.add_argument("--keep-dir", help="copy the whole reference asset (dds + json) into this directory")
    a = ap.parse_args()
    spec = synth.WORKLOADS[a.workload]
    tex = synth.workload(a.workload)
    out_dir = os.path.join(os.path.dirname(os.path.abspath(__file__)), a.workload)
    os.makedirs(out_dir, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        paths = synth.write_pngs(tex, os.path.join(tmp, "m"))
        prefix = os.path.join(tmp, "out") + os.sep   # a directory: the asset is named after the first input
        cmd = [REF_BIN, *paths, "--c0", str(spec["c0"]), "--c1", str(spec["c1"]), "--rounds", str(a.rounds),
               "--tol", "0", "--png", "1" if a.sse else "0", "--backend", "cpu", "-j", str(a.threads), "-o", prefix]
        t0 = time.time()
        r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
        wall = time.time() - t0
        if r.returncode != 0:
            raise SystemExit(r.stdout)
        rep = parse_log(r.stdout)
        rep["wall_seconds"] = wall
        rep["threads"] = a.threads
        rep["command"] = " ".join(os.path.basename(c) if c.startswith(tmp) else c for c in cmd)
        od = os.path.join(tmp, "out")
        files = sorted(f for f in os.listdir(od) if f.startswith("m_t0") and f.endswith(".dds"))
        rep["dds"] = {}
        for f in files:
            data = open(os.path.join(od, f), "rb").read()
            key = f[len("m_t0"):]
            rep["dds"][key] = hashlib.sha256(data).hexdigest()
            if a.keep_dds:
                shutil.copy(os.path.join(od, f), os.path.join(out_dir, "ref" + key))
        js = json.load(open(os.path.join(od, "m_t0_nntc.json")))
        if a.keep_json:
            shutil.copy(os.path.join(od, "m_t0_nntc.json"), os.path.join(out_dir, "ref_nntc.json"))
        if a.keep_dir:
            os.makedirs(a.keep_dir, exist_ok=True)
            for f in os.listdir(od):
                if f.endswith(".dds") or f.endswith(".json"):
                    shutil.copy(os.path.join(od, f), os.path.join(a.keep_dir, f))
        if a.sse:
            from PIL import Image

            rep["sse"] = []
            for t, src in enumerate(tex):
                name = f"m_t0_recon_t{t}_M0.png" if len(tex) > 1 else "m_t0_recon_M0.png"
                rec = np.asarray(Image.open(os.path.join(od, name)).convert("RGB"), dtype=np.int64)
                d = rec - src.astype(np.int64)
                rep["sse"].append(int((d * d).sum()))
            rep["sse_samples"] = int(tex[0].size)
        rep["decoder"] = {"weights": js["decoder"]["layers"][0]["weights"], "bias": js["decoder"]["layers"][0]["bias"]}
        rep["grids"] = [t["dequantise"] for t in js["textures"]]
    json.dump(rep, open(os.path.join(out_dir, "report.json"), "w"), indent=1)
    print(json.dumps({k: rep[k] for k in ("psnr", "e_shipped", "seconds", "rounds_used")}))


if __name__ == "__main__":
    main()
