"""Seeded synthetic materials: the encoder's benchmark and test inputs.

A material is T RGB8 textures built from F fields over one W x H grid:

* F - 1 smooth fields, each 4-octave value noise: a lattice of U(0, 1) values per octave with smoothstep
  interpolation, octave o at period P / 2**o and amplitude 0.5**o, normalised by the sum of the amplitudes;
* 1 detail field, per-pixel U(0, 1), scaled by `detail` (1.0, or 0.3 for the 4096^2 workload).

Texture t is clamp(round(255 * (A_t f + b_t + N(0, 0.02)))) with A_t a random 3 x F matrix of U(-1, 1) entries, each
output row rescaled (the affine b_t) so that the noiseless combination spans [0.1, 0.9]. The fields come from a PCG64
generator seeded with 0x4E4E5443 ^ (config << 16) ^ material; texture t draws its matrix and noise from its own
stream (the same seed with t + 1 appended).
"""

from __future__ import annotations

import numpy as np

SEED_BASE = 0x4E4E5443

# name -> (textures, size, base period, smooth fields, detail amplitude, c0, c1)
WORKLOADS = {
    "plain_cpu_4x256": dict(textures=4, size=(256, 256), period=64, smooth=3, detail=1.0, c0=2, c1=4, config=1),
    "plain_gpu_4x1024": dict(textures=4, size=(1024, 1024), period=128, smooth=3, detail=1.0, c0=2, c1=4, config=2),
    "precision_6x2048": dict(textures=6, size=(2048, 2048), period=128, smooth=5, detail=1.0, c0=4, c1=4, config=3),
    "differentiation_4x4096": dict(textures=4, size=(4096, 4096), period=256, smooth=3, detail=0.3, c0=3, c1=4,
                                   config=4),
    # small cases for the smoke run and the parity tests
    "smoke_1x64x40": dict(textures=1, size=(64, 40), period=32, smooth=3, detail=1.0, c0=1, c1=2, config=9),
    "small_2x100x76": dict(textures=2, size=(100, 76), period=32, smooth=3, detail=1.0, c0=3, c1=3, config=9),
}


def material_seed(config: int, material: int = 0) -> int:
    return SEED_BASE ^ (config << 16) ^ material


def value_noise(rng: np.random.Generator, w: int, h: int, period: float, octaves: int = 4) -> np.ndarray:
    """4-octave value noise on [0, 1], smoothstep-interpolated, amplitude-normalised."""
    out = np.zeros((h, w), dtype=np.float64)
    total = 0.0
    for o in range(octaves):
        p = period / (2 ** o)
        amp = 0.5 ** o
        gw = int(np.ceil(w / p)) + 2
        gh = int(np.ceil(h / p)) + 2
        lattice = rng.random((gh, gw))
        u = np.arange(w, dtype=np.float64) / p
        v = np.arange(h, dtype=np.float64) / p
        iu = np.floor(u).astype(np.int64)
        iv = np.floor(v).astype(np.int64)
        fu = u - iu
        fv = v - iv
        su = fu * fu * (3.0 - 2.0 * fu)
        sv = fv * fv * (3.0 - 2.0 * fv)
        a = lattice[iv][:, iu]
        b = lattice[iv][:, iu + 1]
        c = lattice[iv + 1][:, iu]
        d = lattice[iv + 1][:, iu + 1]
        top = a + (b - a) * su[None, :]
        bot = c + (d - c) * su[None, :]
        out += amp * (top + (bot - top) * sv[:, None])
        total += amp
    return out / total


def make_material(textures: int, size: tuple[int, int], period: float, smooth: int, detail: float, config: int,
                  material: int = 0) -> list[np.ndarray]:
    """The material's textures as uint8 arrays of shape (H, W, 3)."""
    w, h = size
    seed = material_seed(config, material)
    rng = np.random.Generator(np.random.PCG64(seed))
    fields = [value_noise(rng, w, h, period) for _ in range(smooth)]
    fields.append(detail * rng.random((h, w)))
    f = np.stack(fields, axis=0)                       # F x H x W
    out = []
    for t in range(textures):
        trng = np.random.Generator(np.random.PCG64([seed, t + 1]))
        a = trng.uniform(-1.0, 1.0, size=(3, f.shape[0]))
        y = np.tensordot(a, f, axes=(1, 0))             # 3 x H x W
        lo = y.reshape(3, -1).min(axis=1)
        hi = y.reshape(3, -1).max(axis=1)
        span = np.where(hi > lo, hi - lo, 1.0)
        y = 0.1 + 0.8 * (y - lo[:, None, None]) / span[:, None, None]
        y = y + trng.normal(0.0, 0.02, size=y.shape)
        rgb = np.clip(np.rint(255.0 * y), 0, 255).astype(np.uint8)
        out.append(np.ascontiguousarray(np.transpose(rgb, (1, 2, 0))))
    return out


def workload(name: str, material: int = 0) -> list[np.ndarray]:
    spec = WORKLOADS[name]
    return make_material(spec["textures"], spec["size"], spec["period"], spec["smooth"], spec["detail"],
                         spec["config"], material)


def write_pngs(textures: list[np.ndarray], prefix: str) -> list[str]:
    """Write each texture as PREFIX_t<i>.png; returns the paths."""
    from PIL import Image  # imported here: only the PNG-writing path needs it

    paths = []
    for i, t in enumerate(textures):
        p = f"{prefix}_t{i}.png"
        Image.fromarray(t, mode="RGB").save(p)
        paths.append(p)
    return paths


# The batching workload's layouts (c0, c1).
LIBRARY_LAYOUTS = [(2, 4), (3, 4), (4, 4), (1, 2)]


def library_spec(count: int = 64, config: int = 5) -> list[dict]:
    """A library of `count` seeded materials of uneven shapes: per material, width and height uniform over the
    multiples of 4 in [128, 2048] (so not square in general), 1 to 6 textures, and a layout drawn from
    LIBRARY_LAYOUTS; the textures themselves are make_material's at base period 64 with 3 smooth fields and a full
    detail field, seeded by (config, material)."""
    rng = np.random.Generator(np.random.PCG64(material_seed(config, 0xFFFF)))
    out = []
    for m in range(count):
        w = 4 * int(rng.integers(32, 513))
        h = 4 * int(rng.integers(32, 513))
        t = int(rng.integers(1, 7))
        c0, c1 = LIBRARY_LAYOUTS[int(rng.integers(0, len(LIBRARY_LAYOUTS)))]
        out.append(dict(textures=t, size=(w, h), period=64, smooth=3, detail=1.0, c0=c0, c1=c1, config=config,
                        material=m))
    return out


def library(specs: list[dict], threads: int = 8) -> list[list[np.ndarray]]:
    """The textures of every material of `specs` (library_spec), made on `threads` host threads."""
    import concurrent.futures as cf

    def one(s):
        return make_material(s["textures"], s["size"], s["period"], s["smooth"], s["detail"], s["config"],
                             s["material"])

    with cf.ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, specs))
