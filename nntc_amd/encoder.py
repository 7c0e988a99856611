"""ctypes view of include/nntc_encode.h: the encode driver over the gfx950 arm.

    enc = Encoder(textures_rgb8, c0=2, c1=4, rounds=8, tol=0.0)   # host half + device_create
    rep = enc.run()                                             # one whole encode on the resident device
    enc.write("out/material")                                   # the .dds files and the JSON

`arm_lib` / `arm_prefix` select the seam's implementation (default: the gfx950 arm).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi

HOST_LIB = os.path.join(abi.HERE, "libnntc_host.so")


class Options(C.Structure):
    _fields_ = [
        ("c0", C.c_int32), ("c1", C.c_int32), ("bits1", C.c_int32), ("k", C.c_int32), ("rounds", C.c_int32),
        ("tol", C.c_double), ("q1_start", C.c_int32), ("q1_sweeps", C.c_int32), ("q1_range", C.c_int32),
        ("ridge", C.c_double), ("mips", C.c_int32), ("mip_min", C.c_int32), ("mip_weight", C.c_int32),
        ("mip_mix", C.c_double), ("bc_refine", C.c_int32), ("bc_refine_after", C.c_int32),
        ("bc_outer", C.c_int32), ("init0_chain", C.c_int32), ("form", C.c_int32), ("quiet", C.c_int32),
        ("weights", C.c_double * 6), ("rgb_weights", C.c_double * 3), ("matrix_free", C.c_int32),
        ("sweep_lanes", C.c_int32),
        ("graphs", C.c_int32),
        ("init1_pca", C.c_int32), ("init0_texture", C.c_int32), ("init0_luma", C.c_int32),
        ("l0_palette", C.c_int32), ("bits0", C.c_int32 * 4), ("bc0", C.c_int32), ("sweeps", C.c_int32),
    ]


class Report(C.Structure):
    _fields_ = [
        ("textures", C.c_int32), ("nplanes", C.c_int32), ("rounds_used", C.c_int32),
        ("outer_taken", C.c_int32), ("outer_tried", C.c_int32),
        ("psnr", C.c_double * 6), ("centre_psnr", C.c_double), ("sampled_psnr", C.c_double),
        ("e_shipped", C.c_double), ("e_round", (C.c_double * 3) * 64),
        ("pack_psnr", C.c_double), ("pack_psnr_seed", C.c_double), ("moved0_final", C.c_longlong),
        ("seconds", C.c_double),
        ("ms_a", C.c_double), ("ms_b", C.c_double), ("ms_c", C.c_double), ("ms_e", C.c_double),
        ("ms_pack", C.c_double), ("ms_refine", C.c_double), ("ridge_tally", C.c_longlong * 4),
        ("mpix", C.c_double),
    ]


class Material(C.Structure):
    _fields_ = [
        ("textures", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("c0", C.c_int32), ("c1", C.c_int32),
        ("on_device", C.c_int32), ("rgb", C.c_void_p), ("prefix", C.c_char_p),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(HOST_LIB):
            raise FileNotFoundError(f"{HOST_LIB} is not built; run __graft_entry__.build()")
        _lib = C.CDLL(HOST_LIB, mode=C.RTLD_GLOBAL)
        L = _lib
        L.nntc_encode_default_options.argtypes = [C.POINTER(Options)]
        L.nntc_encoder_create.restype = C.c_void_p
        L.nntc_encoder_create.argtypes = [C.POINTER(Options), C.c_int, C.c_int, C.c_int,
                                          C.POINTER(C.POINTER(C.c_uint8)), C.c_char_p, C.c_char_p]
        L.nntc_encoder_create_files.restype = C.c_void_p
        L.nntc_encoder_create_files.argtypes = [C.POINTER(Options), C.c_int, C.POINTER(C.c_char_p), C.c_char_p,
                                                C.c_char_p]
        L.nntc_encoder_create_device.restype = C.c_void_p
        L.nntc_encoder_create_device.argtypes = [C.POINTER(Options), C.c_int, C.c_int, C.c_int,
                                                 C.POINTER(C.c_void_p)]
        L.nntc_encoder_run.argtypes = [C.c_void_p]
        L.nntc_encoder_destroy.argtypes = [C.c_void_p]
        L.nntc_encode_last_error.restype = C.c_char_p
        L.nntc_encoder_report.argtypes = [C.c_void_p, C.POINTER(Report)]
        L.nntc_encoder_state.argtypes = [C.c_void_p, C.POINTER(abi.ModelState)]
        L.nntc_encoder_desc.argtypes = [C.c_void_p, C.POINTER(abi.ModelDesc)]
        L.nntc_encoder_level1_indices.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_uint8), C.c_size_t,
                                                  C.POINTER(C.c_size_t)]
        L.nntc_encoder_level0_blocks.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_size_t,
                                                 C.POINTER(C.c_size_t)]
        L.nntc_encoder_recon.argtypes = [C.c_void_p, C.POINTER(C.c_uint8), C.c_size_t, C.POINTER(C.c_size_t)]
        L.nntc_encoder_write.argtypes = [C.c_void_p, C.c_char_p]
        L.nntc_encode_library.argtypes = [C.POINTER(Material), C.c_int, C.POINTER(Options), C.c_int,
                                          C.POINTER(Report), C.POINTER(C.c_int32)]
    return _lib


def options(**kw) -> Options:
    o = Options()
    lib().nntc_encode_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "bits0":
            v = list(v) + [v[-1]] * (4 - len(v)) if len(v) else [3] * 4
            v = (C.c_int32 * 4)(*v)
        setattr(o, k, v)
    return o


class Encoder:
    def __init__(self, textures: list[np.ndarray] | None = None, files: list[str] | None = None,
                 arm_lib: str | None = None, arm_prefix: str | None = None, device=None, **opts):
        """textures: host RGB8 arrays (H, W, 3); files: PNG paths; device: (width, height, [device pointers of each
        texture's W*H*3 bytes]) for textures already in GPU memory (the caller keeps them alive)."""
        L = lib()
        self.opts = options(**opts)
        a = arm_lib.encode() if arm_lib else None
        p = arm_prefix.encode() if arm_prefix else None
        if device is not None:
            w, h, ptrs_dev = device
            arr = (C.c_void_p * len(ptrs_dev))(*[C.c_void_p(int(x)) for x in ptrs_dev])
            self.h = L.nntc_encoder_create_device(C.byref(self.opts), len(ptrs_dev), w, h, arr)
        elif files is not None:
            arr = (C.c_char_p * len(files))(*[f.encode() for f in files])
            self.h = L.nntc_encoder_create_files(C.byref(self.opts), len(files), arr, a, p)
        else:
            self._keep = [np.ascontiguousarray(t, dtype=np.uint8) for t in textures]
            h, w = self._keep[0].shape[:2]
            ptrs = (C.POINTER(C.c_uint8) * len(self._keep))(
                *[t.ctypes.data_as(C.POINTER(C.c_uint8)) for t in self._keep])
            self.h = L.nntc_encoder_create(C.byref(self.opts), len(self._keep), w, h, ptrs, a, p)
        if not self.h:
            raise RuntimeError(L.nntc_encode_last_error().decode())
        self.desc = abi.ModelDesc()
        L.nntc_encoder_desc(self.h, C.byref(self.desc))

    def run(self) -> Report:
        L = lib()
        if L.nntc_encoder_run(self.h) != 0:
            raise RuntimeError(L.nntc_encode_last_error().decode())
        return self.report()

    def report(self) -> Report:
        r = Report()
        lib().nntc_encoder_report(self.h, C.byref(r))
        return r

    def state(self) -> abi.ModelState:
        s = abi.ModelState()
        lib().nntc_encoder_state(self.h, C.byref(s))
        return s

    def level1_indices(self, p: int) -> np.ndarray:
        pl = self.desc.planes[p]
        out = np.zeros(pl.w1 * pl.h1 * self.desc.c1, np.uint8)
        lib().nntc_encoder_level1_indices(self.h, p, out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size, None)
        return out

    def level0_blocks(self, p: int, f: int) -> np.ndarray:
        pl = self.desc.planes[p]
        nb = ((pl.w0 + 3) // 4) * ((pl.h0 + 3) // 4)
        out = np.zeros(nb * 16, np.uint8)
        n = C.c_size_t(0)
        lib().nntc_encoder_level0_blocks(self.h, p, f, out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size,
                                         C.byref(n))
        return out[: n.value]

    def recon(self) -> np.ndarray:
        pl = self.desc.planes[0]
        out = np.zeros(pl.w0 * pl.h0 * self.desc.nout, np.uint8)
        lib().nntc_encoder_recon(self.h, out.ctypes.data_as(C.POINTER(C.c_uint8)), out.size, None)
        return out

    def write(self, prefix: str) -> None:
        L = lib()
        if L.nntc_encoder_write(self.h, prefix.encode()) != 0:
            raise RuntimeError(L.nntc_encode_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            lib().nntc_encoder_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def encode_library(materials: list[dict], workers: int = 4, **opts) -> list[Report]:
    """Encode a library of materials side by side (include/nntc_encode.h, nntc_encode_library). Each material is a
    dict with c0, c1 and either `textures` (host RGB8 arrays) or `device` = (width, height, [device pointers]), and
    optionally `prefix` (where to write its asset). Returns the reports in the materials' order; raises if any
    material failed."""
    L = lib()
    o = options(**opts)
    arr = (Material * len(materials))()
    keep = []
    for i, m in enumerate(materials):
        if "device" in m:
            w, h, ptrs = m["device"]
            rgb = (C.c_void_p * len(ptrs))(*[C.c_void_p(int(x)) for x in ptrs])
            on_device = 1
        else:
            tex = [np.ascontiguousarray(t, dtype=np.uint8) for t in m["textures"]]
            h, w = tex[0].shape[:2]
            rgb = (C.c_void_p * len(tex))(*[t.ctypes.data for t in tex])
            keep.append(tex)
            on_device = 0
        keep.append(rgb)
        prefix = m.get("prefix")
        arr[i] = Material(len(rgb), w, h, m["c0"], m["c1"], on_device, C.cast(rgb, C.c_void_p),
                          prefix.encode() if prefix else None)
    reports = (Report * len(materials))()
    status = (C.c_int32 * len(materials))()
    failed = L.nntc_encode_library(arr, len(materials), C.byref(o), workers, reports, status)
    if failed:
        raise RuntimeError(f"{failed} of {len(materials)} materials failed: {L.nntc_encode_last_error().decode()}")
    return list(reports)
