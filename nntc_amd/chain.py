"""The encode's host-side inputs: the packed source, its box mip chain, the plane layout and the mip weights.

Each function restates the reference's host code in float32 numpy with the reference's operation order:
  - RGB8 -> float by x * (1/255) (src/image.h:146-155), textures interleaved into 3T channels (src/image.h:321-331),
    edge-replication padding to a multiple of 4 (src/image.h:298-319, src/main.cpp:1865-1880);
  - the 2x2 box chain 0.25 ((a + b) + (c + d)) with floor halving (src/image.h:335-366);
  - build_plane_sizes and build_mip_weights (src/mips.cpp:25-40, 141-169);
  - the objective's per-output weight cw (src/main.cpp:1856-1862).
"""

from __future__ import annotations

import numpy as np

from . import abi


def pack_textures(textures: list[np.ndarray]) -> np.ndarray:
    """T uint8 (H, W, 3) arrays -> one float32 (H, W, 3T) image."""
    inv = np.float32(1.0) / np.float32(255.0)
    return np.concatenate([t.astype(np.float32) * inv for t in textures], axis=2)


def pad_to_block(img: np.ndarray, block: int = 4) -> np.ndarray:
    h, w = img.shape[:2]
    hp = (h + block - 1) // block * block
    wp = (w + block - 1) // block * block
    if hp == h and wp == w:
        return img
    ys = np.minimum(np.arange(hp), h - 1)
    xs = np.minimum(np.arange(wp), w - 1)
    return img[ys][:, xs]


def mip_dim(n: int) -> int:
    return n // 2 if n // 2 > 0 else 1


def box(img: np.ndarray) -> np.ndarray:
    h, w = img.shape[:2]
    oh, ow = mip_dim(h), mip_dim(w)
    xa = np.minimum(2 * np.arange(ow), w - 1)
    xb = np.minimum(2 * np.arange(ow) + 1, w - 1)
    ya = np.minimum(2 * np.arange(oh), h - 1)
    yb = np.minimum(2 * np.arange(oh) + 1, h - 1)
    a = img[ya][:, xa]
    b = img[ya][:, xb]
    c = img[yb][:, xa]
    d = img[yb][:, xb]
    return (np.float32(0.25) * ((a + b) + (c + d))).astype(np.float32)


def source_chain(textures: list[np.ndarray], mips: bool = True, mip_min: int = 8):
    """(planes, chain): the plane layout and the float32 source at every stored level."""
    base = pad_to_block(pack_textures(textures))
    h, w = base.shape[:2]
    planes = abi.plane_sizes(w, h, mips, mip_min)
    chain = [np.ascontiguousarray(base)]
    for _ in range(1, len(planes)):
        chain.append(box(chain[-1]))
    return planes, chain


def mip_weights(planes, kind: str = "pixels", mix: float = 0.5, k: int = 4):
    n = len(planes)
    if n == 1:
        return [1.0], [1.0]
    r = [0.0] * n
    total = 0.0
    for m in range(1, n):
        pixels = float(planes[m][0]) * float(planes[m][1])
        r[m] = 1.0 if kind == "uniform" else (np.sqrt(pixels) if kind == "sqrt" else pixels)
        total += r[m]
    share = [0.0] * n
    share[0] = 1.0 - mix
    for m in range(1, n):
        share[m] = mix * r[m] / total if total > 0.0 else 0.0
    per_site = []
    for m in range(n):
        sites = float(planes[m][0]) * float(planes[m][1]) * float(1 + k)
        per_site.append(share[m] / sites if sites > 0.0 else 0.0)
    return share, per_site


def output_weights(textures: int, weights=None, rgb=(1.0, 1.0, 1.0)) -> list[float]:
    cw = []
    for c in range(3 * textures):
        wt = 1.0 if weights is None else weights[c // 3]
        s = rgb[0] + rgb[1] + rgb[2]
        cw.append(float(np.float32(wt * rgb[c % 3] * 3.0 / s)))
    return cw
