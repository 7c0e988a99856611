"""ctypes view of include/nntc_hip.h, for tests and the benchmark.

`Arm(lib_path, prefix)` binds every entry point of the C ABI from one shared library: the gfx950 arm
(`libnntc_hip.so`, prefix "nntc_hip_") or any library exporting the same functions under another prefix. The plain
structs below mirror the header field for field.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

MAX_PLANES = 13
MAX_CH = 4
MAX_NIN = 24
MAX_NOUT = 18
RANGE_MINMAX = 0
RANGE_PCT = 1
FORM_EXACT = 0
FORM_FAST = 1
FORM_PLAIN = 2

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(HERE, "libnntc_hip.so")
HOST_LIB = os.path.join(HERE, "libnntc_host.so")


class PlaneSize(C.Structure):
    _fields_ = [("w0", C.c_int32), ("h0", C.c_int32), ("w1", C.c_int32), ("h1", C.c_int32)]


class ModelDesc(C.Structure):
    _fields_ = [
        ("textures", C.c_int32),
        ("nout", C.c_int32),
        ("c0", C.c_int32),
        ("c1", C.c_int32),
        ("bits0", C.c_int32 * MAX_CH),
        ("bits1", C.c_int32),
        ("bc0", C.c_int32),
        ("l0_bc8", C.c_int32),
        ("init0_residual", C.c_int32),
        ("init0_texture", C.c_int32),
        ("init0_chain", C.c_int32),
        ("nplanes", C.c_int32),
        ("planes", PlaneSize * MAX_PLANES),
        ("cw", C.c_float * MAX_NOUT),
    ]


class ModelState(C.Structure):
    _fields_ = [
        ("has_decoder", C.c_int32),
        ("w", C.c_float * (MAX_NOUT * MAX_NIN)),
        ("b", C.c_float * MAX_NOUT),
        ("has_grid0", C.c_int32),
        ("has_grid1", C.c_int32),
        ("lo0", C.c_float * MAX_CH),
        ("hi0", C.c_float * MAX_CH),
        ("lo1", C.c_float * MAX_CH),
        ("hi1", C.c_float * MAX_CH),
        ("has_palette0", C.c_int32),
        ("palette0", (C.c_float * 16) * MAX_CH),
    ]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 256), ("major", C.c_int32), ("minor", C.c_int32), ("error", C.c_char * 256)]


class Level1InitReport(C.Structure):
    _fields_ = [("pca", C.c_int32), ("components", C.c_int32), ("eigenvalue", C.c_double * 4),
                ("peak", C.c_double * 4), ("total_variance", C.c_double)]


class Init0Report(C.Structure):
    _fields_ = [("eigenvalue", C.c_double), ("total_variance", C.c_double), ("share", C.c_double),
                ("peak", C.c_double), ("divisor", C.c_double), ("dir", C.c_double * MAX_NOUT),
                ("nout", C.c_int32), ("texture", C.c_int32)]


class Level0Report(C.Structure):
    _fields_ = [("moved", C.c_int64 * 4), ("moved_total", C.c_int64), ("ms", C.c_double), ("joint", C.c_int32),
                ("states", C.c_int64)]


class Level1Report(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("residual", C.c_double),
        ("lambda_", C.c_double),
        ("min_diag", C.c_double),
        ("max_diag", C.c_double),
        ("ms_assemble", C.c_double),
        ("ms_precond", C.c_double),
        ("ms_cg", C.c_double),
        ("moved_values", C.c_int64),
        ("delta2", C.c_double),
        ("plane2", C.c_double),
        ("lo", C.c_float * MAX_CH),
        ("hi", C.c_float * MAX_CH),
        ("quantised", C.c_int32),
        ("sweeps", C.c_int32),
        ("moved", C.c_int64),
        ("ms_sweeps", C.c_double),
    ]


class Objective(C.Structure):
    _fields_ = [("e", C.c_double), ("centre_mse", C.c_double), ("sampled_mse", C.c_double), ("ms", C.c_double),
                ("nplanes", C.c_int32), ("e_plane", C.c_double * MAX_PLANES),
                ("centre_plane", C.c_double * MAX_PLANES)]


class BcRefineReport(C.Structure):
    _fields_ = [("pass_", C.c_int32), ("blocks", C.c_int64), ("improved", C.c_int64), ("delta_e", C.c_double),
                ("ms", C.c_double)]


_P = C.POINTER
_u8p = _P(C.c_uint8)
_f32p = _P(C.c_float)
_f64p = _P(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "last_error": (C.c_char_p, []),
    "device_select": (C.c_int, [C.c_int, _P(DeviceInfo)]),
    "device_create": (C.c_void_p, [_P(ModelDesc), _P(_f32p)]),
    "device_destroy": (None, [C.c_void_p]),
    "device_memory": (C.c_size_t, [C.c_void_p]),
    "set_form": (C.c_int, [C.c_void_p, C.c_int]),
    "init_level0": (C.c_int, [C.c_void_p, _P(ModelState), _f32p, C.c_int]),
    "init_level1": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, _P(Level1InitReport)]),
    "init0_residual_channel": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, _f64p, _P(Init0Report)]),
    "quantise_level1": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int]),
    "freeze_level1_grid": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int]),
    "quantise_level0": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int]),
    "freeze_level0_grid": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int]),
    "snap_level0_on_grid": (C.c_int, [C.c_void_p, _P(ModelState)]),
    "stencil_monomials": (C.c_int, [C.c_int]),
    "solve_decoder": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, _f64p, _f64p]),
    "decoder_ridge_tally": (None, [_P(C.c_longlong)]),
    "solve_level1_all": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, C.c_double, _P(Level1Report), _f64p]),
    "sweep_level1_all": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, C.c_double, C.c_int, _P(Level1Report),
                                   _f64p]),
    "solve_level0_cont_all": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, C.c_double, _P(Level1Report), _f64p]),
    "sweep_level0_cont_all": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, C.c_double, C.c_int,
                                        _P(Level1Report), _f64p]),
    "assemble_level0_for_refine": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, _f64p]),
    "level1_download_indices": (C.c_int, [C.c_void_p, C.c_int, _u8p, C.c_size_t, _P(C.c_size_t)]),
    "level1_upload": (C.c_int, [C.c_void_p, C.c_int, _f32p, C.c_size_t]),
    "level1_upload_indices": (C.c_int, [C.c_void_p, C.c_int, _u8p, C.c_size_t]),
    "level0_upload_indices": (C.c_int, [C.c_void_p, C.c_int, _u8p, C.c_size_t]),
    "upload_decoder_to_device": (C.c_int, [C.c_void_p, _P(ModelState)]),
    "upload_grids_to_device": (C.c_int, [C.c_void_p, _P(ModelState)]),
    "level1_download": (C.c_int, [C.c_void_p, C.c_int, _f32p, C.c_size_t, _P(C.c_size_t)]),
    "level1_delta": (C.c_int, [C.c_void_p, C.c_int, _f64p, C.c_size_t, _P(C.c_size_t)]),
    "level1_off_grid": (C.c_int, [C.c_void_p, _P(ModelState), _P(C.c_size_t), _P(C.c_size_t)]),
    "level1_poke": (C.c_int, [C.c_void_p, C.c_int, C.c_size_t, C.c_float]),
    "level1_range": (C.c_int, [C.c_void_p, C.c_int, _f32p, _f32p]),
    "level0_download": (C.c_int, [C.c_void_p, C.c_int, _u8p, C.c_size_t, _P(C.c_size_t)]),
    "level0_upload": (C.c_int, [C.c_void_p, C.c_int, _f32p, C.c_size_t]),
    "level0_download_values": (C.c_int, [C.c_void_p, C.c_int, _f32p, C.c_size_t, _P(C.c_size_t)]),
    "decode_plane": (C.c_int, [C.c_void_p, C.c_int, _u8p, C.c_size_t]),
    "texture_sse": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_ulonglong)]),
    "objective_eval": (C.c_int, [C.c_void_p, C.c_int, _f64p, _P(Objective)]),
    "objective_check_host": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, _f64p, _P(Objective)]),
    "solve_level0_all": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, C.c_int, _P(Level0Report), _f64p]),
    "solve_level1_dense_host": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, C.c_int, C.c_double, _f32p,
                                          C.c_size_t, _f64p, C.c_size_t, C.c_int, _P(C.c_int)]),
    "objective_total_ms": (C.c_double, []),
    "objective_passes": (C.c_longlong, []),
    "bc_pack_prepare": (C.c_int, [C.c_void_p, _P(ModelState), C.c_int, C.c_int, _P(_u8p), _f64p, _P(C.c_size_t),
                                  _f64p]),
    "bc_refine": (C.c_int, [C.c_void_p, C.c_int, _P(_u8p), _P(BcRefineReport), _f64p, _P(C.c_size_t), _f64p]),
    "bc_blocks_from_device": (C.c_int, [C.c_void_p, _P(_u8p), _f64p, _P(C.c_size_t)]),
    "bc_blocks_download": (C.c_int, [C.c_void_p, C.c_int, C.c_int, _u8p, C.c_size_t, _P(C.c_size_t)]),
    "bc_state_save": (C.c_int, [C.c_void_p, C.c_int, _u8p, C.c_size_t, _u8p, C.c_size_t]),
    "bc_state_restore": (C.c_int, [C.c_void_p, C.c_int, _u8p, C.c_size_t, _u8p, C.c_size_t]),
    "state_snapshot": (C.c_int, [C.c_void_p]),
    "state_restore": (C.c_int, [C.c_void_p]),
}


def declared_symbols(header: str) -> list[str]:
    """Every function name a header declares (`<type> name(` at line start, after the type)."""
    import re

    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(nntc_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(_P(ct))


class ArmError(RuntimeError):
    pass


class Arm:
    """One loaded implementation of the seam."""

    def __init__(self, lib_path: str = HIP_LIB, prefix: str = "nntc_hip_"):
        if not os.path.exists(lib_path):
            raise FileNotFoundError(f"{lib_path} is not built; run __graft_entry__.build()")
        self.lib = C.CDLL(lib_path, mode=C.RTLD_GLOBAL)
        self.path = lib_path
        self.prefix = prefix
        self.fn = {}
        for name, (res, args) in SIGNATURES.items():
            f = getattr(self.lib, prefix + name)
            f.restype = res
            f.argtypes = args
            self.fn[name] = f

    def check(self, status: int, what: str):
        if status != 0:
            raise ArmError(f"{self.prefix}{what}: {self.fn['last_error']().decode()}")


class Session:
    """A model on one arm: the descriptor, the mutable state, and typed wrappers over the seam."""

    def __init__(self, arm: Arm, desc: ModelDesc, chain: list[np.ndarray]):
        self.arm = arm
        self.desc = desc
        self.state = ModelState()
        if desc.l0_bc8:
            self.state.has_grid0 = 1
            for q in range(desc.c0):
                self.state.lo0[q] = -1.0
                self.state.hi0[q] = 1.0
        else:
            self.state.has_palette0 = 1
            for q in range(desc.c0):
                for k, v in enumerate(level0_palette(desc.bits0[q], bool(desc.bc0))):
                    self.state.palette0[q][k] = v
        self.chain = [np.ascontiguousarray(c, dtype=np.float32) for c in chain]
        ptrs = (_f32p * len(self.chain))(*[_ptr(c, C.c_float) for c in self.chain])
        self.ctx = arm.fn["device_create"](C.byref(desc), ptrs)
        if not self.ctx:
            raise ArmError(f"{arm.prefix}device_create: {arm.fn['last_error']().decode()}")
        self.nplanes = desc.nplanes
        self.nin = desc.c1 + desc.c0 + desc.c0 * desc.c1
        self.nout = desc.nout

    def close(self):
        if self.ctx:
            self.arm.fn["device_destroy"](self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        self.arm.check(self.arm.fn[name](self.ctx, *args), name)

    # ---- typed helpers
    def plane(self, p):
        return self.desc.planes[p]

    def decoder(self):
        w = np.ctypeslib.as_array(self.state.w)[: self.nout * self.nin].reshape(self.nout, self.nin).copy()
        b = np.ctypeslib.as_array(self.state.b)[: self.nout].copy()
        return w, b

    def set_decoder(self, w, b):
        self.state.has_decoder = 1
        flat = np.asarray(w, dtype=np.float32).reshape(-1)
        for i, v in enumerate(flat):
            self.state.w[i] = float(v)
        for i, v in enumerate(np.asarray(b, dtype=np.float32)):
            self.state.b[i] = float(v)
        self.call("upload_decoder_to_device", C.byref(self.state))

    def download(self, name, plane, dtype, count):
        out = np.zeros(count, dtype=dtype)
        ct = {np.uint8: C.c_uint8, np.float32: C.c_float, np.float64: C.c_double}[dtype]
        n = C.c_size_t(0)
        self.call(name, plane, _ptr(out, ct), count, C.byref(n))
        return out[: n.value] if n.value < count else out

    def v0(self, p):
        pl = self.plane(p)
        return self.download("level0_download_values", p, np.float32, pl.w0 * pl.h0 * self.desc.c0)

    def k0(self, p):
        pl = self.plane(p)
        return self.download("level0_download", p, np.uint8, pl.w0 * pl.h0 * self.desc.c0)

    def v1(self, p):
        pl = self.plane(p)
        return self.download("level1_download", p, np.float32, pl.w1 * pl.h1 * self.desc.c1)

    def k1(self, p):
        pl = self.plane(p)
        return self.download("level1_download_indices", p, np.uint8, pl.w1 * pl.h1 * self.desc.c1)

    def level1_off_grid(self):
        """(values, off): the level-1 values, and those not the grid value of their index under the state's grid."""
        values, off = C.c_size_t(0), C.c_size_t(0)
        self.call("level1_off_grid", C.byref(self.state), C.byref(values), C.byref(off))
        return values.value, off.value

    def delta1(self, p):
        pl = self.plane(p)
        return self.download("level1_delta", p, np.float64, pl.w1 * pl.h1 * self.desc.c1)

    def upload_v0(self, p, v):
        v = np.ascontiguousarray(v, dtype=np.float32)
        self.call("level0_upload", p, _ptr(v, C.c_float), v.size)

    def upload_v1(self, p, v):
        v = np.ascontiguousarray(v, dtype=np.float32)
        self.call("level1_upload", p, _ptr(v, C.c_float), v.size)

    def objective(self, k, per_site):
        o = Objective()
        ps = np.ascontiguousarray(per_site, dtype=np.float64)
        self.call("objective_eval", k, _ptr(ps, C.c_double), C.byref(o))
        return o

    def objective_check(self, k, per_site):
        """The --check twin of objective(): E by brute force on the host."""
        o = Objective()
        ps = np.ascontiguousarray(per_site, dtype=np.float64)
        self.call("objective_check_host", C.byref(self.state), k, _ptr(ps, C.c_double), C.byref(o))
        return o

    def level1_dense(self, k, plane, lam, c_prev, max_unknowns=1024):
        """The --check twin of block (b) on one plane: the dense solve at c_prev, or None when it declines."""
        prev = np.ascontiguousarray(c_prev, dtype=np.float32)
        out = np.zeros(prev.size, dtype=np.float64)
        solved = C.c_int(0)
        self.call("solve_level1_dense_host", C.byref(self.state), k, plane, lam, _ptr(prev, C.c_float), prev.size,
                  _ptr(out, C.c_double), out.size, max_unknowns, C.byref(solved))
        return out if solved.value else None

    def solve_decoder(self, k, per_site):
        ps = np.ascontiguousarray(per_site, dtype=np.float64)
        ms = C.c_double(0)
        self.call("solve_decoder", C.byref(self.state), k, _ptr(ps, C.c_double), C.byref(ms))
        return ms.value

    def block(self, name, k, ridge, sweeps=None):
        rep = (Level1Report * self.nplanes)()
        ms = C.c_double(0)
        if sweeps is None:
            self.call(name, C.byref(self.state), k, ridge, rep, C.byref(ms))
        else:
            self.call(name, C.byref(self.state), k, ridge, sweeps, rep, C.byref(ms))
        return list(rep), ms.value

    def solve_level0_all(self, k, sweeps):
        """--l0 palette's block (c): the per-texel palette search on every plane."""
        rep = (Level0Report * self.nplanes)()
        ms = C.c_double(0)
        self.call("solve_level0_all", C.byref(self.state), k, sweeps, rep, C.byref(ms))
        return list(rep), ms.value

    def decode_plane(self, p):
        pl = self.plane(p)
        out = np.zeros(pl.w0 * pl.h0 * self.nout, dtype=np.uint8)
        self.call("decode_plane", p, _ptr(out, C.c_uint8), out.size)
        return out

    def texture_sse(self, width, height):
        out = (C.c_ulonglong * (self.nout // 3))()
        self.call("texture_sse", width, height, out)
        return list(out)

    def init_level0(self, luma=(0.299, 0.587, 0.114), luma_channel0=False):
        l = np.array(luma, dtype=np.float32)
        self.call("init_level0", C.byref(self.state), _ptr(l, C.c_float), 1 if luma_channel0 else 0)

    def init_level1(self, pca=False):
        rep = Level1InitReport()
        self.call("init_level1", C.byref(self.state), 1 if pca else 0, C.byref(rep))
        return rep

    def init0_residual_channel(self, channel, plane_weight):
        rep = Init0Report()
        w = np.ascontiguousarray(plane_weight, dtype=np.float64)
        self.call("init0_residual_channel", C.byref(self.state), channel, _ptr(w, C.c_double), C.byref(rep))
        return rep

    def grid(self, name, policy):
        self.call(name, C.byref(self.state), policy)

    def bc_pack_prepare(self, k, seed, k0_mirror=None):
        psnr = C.c_double(0)
        texels = C.c_size_t(0)
        ms = C.c_double(0)
        mir, keep = self._mirror(k0_mirror)
        self.call("bc_pack_prepare", C.byref(self.state), k, 1 if seed else 0, mir, C.byref(psnr), C.byref(texels),
                  C.byref(ms))
        return psnr.value, texels.value

    def bc_refine(self, passes, k0_mirror=None):
        rep = (BcRefineReport * max(1, passes))()
        psnr = C.c_double(0)
        texels = C.c_size_t(0)
        ms = C.c_double(0)
        mir, keep = self._mirror(k0_mirror)
        self.call("bc_refine", passes, mir, rep, C.byref(psnr), C.byref(texels), C.byref(ms))
        return list(rep)[:passes], psnr.value, texels.value

    def _mirror(self, k0_mirror):
        if k0_mirror is None:
            return None, None
        arrs = [np.ascontiguousarray(a, dtype=np.uint8) for a in k0_mirror]
        ptrs = (_u8p * len(arrs))(*[_ptr(a, C.c_uint8) for a in arrs])
        return ptrs, arrs

    def bc_blocks(self, p, f):
        pl = self.plane(p)
        nb = ((pl.w0 + 3) // 4) * ((pl.h0 + 3) // 4)
        return self.download_blocks(p, f, nb * 16)

    def download_blocks(self, p, f, cap):
        out = np.zeros(cap, dtype=np.uint8)
        n = C.c_size_t(0)
        self.call("bc_blocks_download", p, f, _ptr(out, C.c_uint8), cap, C.byref(n))
        return out[: n.value]


def plane_sizes(width: int, height: int, mips: bool = True, mip_min: int = 8) -> list[tuple[int, int, int, int]]:
    """build_plane_sizes (src/mips.cpp:25-40): the base, then floor halvings while both sides stay >= mip_min."""
    levels = 0
    w, h = width, height
    if mips:
        while levels < 12 and w // 2 >= mip_min and h // 2 >= mip_min:
            w //= 2
            h //= 2
            levels += 1
    half = lambda n: n // 2 if n // 2 > 0 else 1
    planes = [(width, height, width // 4, height // 4)]
    for _ in range(levels):
        w0, h0, w1, h1 = planes[-1]
        planes.append((half(w0), half(h0), half(w1), half(h1)))
    return planes


def replicate_byte(k: int, bits: int) -> int:
    """The index k of `bits` bits in a byte's top bits, replicated downward (the uncompressed texel rule)."""
    byte, sh = 0, 8 - bits
    while sh > -bits:
        byte |= (k << sh) if sh >= 0 else (k >> -sh)
        sh -= bits
    return byte & 255


def level0_palette(bits: int, bc: bool) -> list:
    """--l0 palette's values of one channel, what the shipped format decodes index k to: -1 + 2 k / (2^bits - 1)
    block-compressed, -1 + 2 rep(k) / 255 uncompressed, rounded to float (src/model.h:205-222)."""
    levels = (1 << bits) - 1
    out = []
    for k in range(levels + 1):
        frac = k / levels if bc else replicate_byte(k, bits) / 255.0
        out.append(float(np.float32(-1.0 + 2.0 * frac)))
    return out


def make_desc(textures: int, c0: int, c1: int, planes, cw=None, bits1: int = 8, bits0=None, bc0: int = 1) -> ModelDesc:
    """The model descriptor; bits0 (one entry per level-0 channel, 1 to 4) selects --l0 palette."""
    d = ModelDesc()
    d.textures = textures
    d.nout = 3 * textures
    d.c0 = c0
    d.c1 = c1
    for q in range(MAX_CH):
        d.bits0[q] = 8 if bits0 is None or q >= len(bits0) else bits0[q]
    d.bits1 = bits1
    d.bc0 = bc0
    d.l0_bc8 = 1 if bits0 is None else 0
    d.init0_residual = 1
    d.nplanes = len(planes)
    for i, (w0, h0, w1, h1) in enumerate(planes):
        d.planes[i] = PlaneSize(w0, h0, w1, h1)
    cw = cw if cw is not None else [1.0] * (3 * textures)
    for i, v in enumerate(cw):
        d.cw[i] = v
    return d


class DeviceBytes:
    """A host array copied into device memory through the HIP runtime the arm runs on (the runtime libnntc_hip.so
    loaded, found by its soname), for entry points that take device pointers. Freed by close()."""

    def __init__(self, data: np.ndarray):
        C.CDLL(HIP_LIB, mode=C.RTLD_GLOBAL)   # the arm first, so the soname below resolves to its runtime
        self.rt = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.host = np.ascontiguousarray(data)
        self.ptr = C.c_void_p(0)
        if self.rt.hipMalloc(C.byref(self.ptr), self.host.nbytes) != 0:
            raise ArmError("hipMalloc failed")
        if self.rt.hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) != 0:   # hipMemcpyHostToDevice
            raise ArmError("hipMemcpy failed")

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def read(self) -> np.ndarray:
        """The device bytes back in an array shaped and typed like the one uploaded, after every queued device
        operation has finished."""
        self.rt.hipDeviceSynchronize()
        out = np.empty_like(self.host)
        if self.rt.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) != 0:   # hipMemcpyDeviceToHost
            raise ArmError("hipMemcpy failed")
        return out

    def close(self):
        if self.ptr.value:
            self.rt.hipFree(self.ptr)
            self.ptr = C.c_void_p(0)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
