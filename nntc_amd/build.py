"""In-tree build of the gfx950 arm and its host driver.

    python -m nntc_amd.build          # everything
    python -m nntc_amd.build hip      # libnntc_hip.so only

Outputs (git-ignored, but part of the tree the GPU runs use):
    nntc_amd/libnntc_hip.so    the HIP arm behind include/nntc_hip.h
    nntc_amd/libnntc_host.so   the encode driver behind include/nntc_encode.h (links the arm)
    nntc_amd/bin/nntc_hip_encode   the command-line encoder
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
OBJ = os.path.join(ROOT, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIP_SOURCES = ["arm", "objective", "decoder", "latent", "init", "refine", "adjoint", "twins", "palette"]
HOST_SOURCES = ["encode", "image", "writer", "library"]

# -ffp-contract=off: no fused multiply-add anywhere, as the reference builds (CMakeLists.txt:183-187, 206-208).
# The packed-fp32 feature is switched off for the device: the compiler's own packed FP32 forms are not used.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + INCLUDE]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + INCLUDE]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}")


def hipcc() -> str:
    for p in ("/opt/rocm/bin/hipcc",):
        if os.path.exists(p):
            return p
    return "hipcc"


def build_hip(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    cc = hipcc()

    def one(name: str) -> str:
        out = os.path.join(OBJ, name + ".o")
        _run([cc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, name + ".hip"), "-o", out])
        return out

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    lib = os.path.join(HERE, "libnntc_hip.so")
    _run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", lib])
    return lib


def build_host(jobs: int = 8) -> str:
    """The encode driver: host C++ over the arm's C ABI (no device code of its own)."""
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(os.path.join(HERE, "bin"), exist_ok=True)
    cxx = "g++"
    present = [s for s in HOST_SOURCES if os.path.exists(os.path.join(HOST, s + ".cpp"))]

    def one(name: str) -> str:
        out = os.path.join(OBJ, "host_" + name + ".o")
        _run([cxx, *CXX_FLAGS, "-c", os.path.join(HOST, name + ".cpp"), "-o", out])
        return out

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, present))
    lib = os.path.join(HERE, "libnntc_host.so")
    _run([cxx, "-shared", "-fPIC", *objs, "-L" + HERE, "-lnntc_hip", "-Wl,-rpath,$ORIGIN", "-lz", "-ldl",
          "-lpthread", "-o", lib])
    if os.path.exists(os.path.join(HOST, "main.cpp")):
        main_o = os.path.join(OBJ, "host_main.o")
        _run([cxx, *CXX_FLAGS, "-c", os.path.join(HOST, "main.cpp"), "-o", main_o])
        _run([cxx, main_o, "-L" + HERE, "-lnntc_host", "-lnntc_hip", "-Wl,-rpath,$ORIGIN/..", "-o",
              os.path.join(HERE, "bin", "nntc_hip_encode")])
    return lib


def build_oracle() -> bool:
    """The reference CPU encoder from /root/reference (this container only; the GPU box uses the built files)."""
    if not os.path.isdir("/root/reference/src"):
        return False
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "-j8"])
    return True


def build_all() -> None:
    build_hip()
    build_host()
    build_oracle()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "host": build_host, "oracle": build_oracle}[w]()
    print("built:", ", ".join(what))
