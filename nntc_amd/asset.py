"""A reader of the NNTC asset that shares no code with the writer (nntc_amd/host/writer.cpp).

It does what a consumer does with the files nntc_encode writes (docs/FORMAT.md of the reference; its own reader is
tools/dds_decode.py): parse the DX10 .dds headers, decode the BC4 / BC5 blocks of level 0 and the bytes of level 1,
sample both latents at the centres of an output mip's pixels with the bilinear clamp rule, dequantise AFTER the blend
(value = lo + sample * (hi - lo)), build phi in the descriptor's term order [a b sc] and apply the one affine layer,
then clamp to [0, 1] and round to bytes. Everything is float64, as the reference's reader computes; the encoder decodes
in float32 (its decode_plane), so the two agree within one step of 255.

    asset = load("out/m_nntc.json")     # or the prefix "out/m"
    rgb = decode(asset, mip=0)          # (height, width, 3T) uint8
    psnr(rgb, textures)                 # per texture, the reference's 8-bit PSNR over the source extent
"""

from __future__ import annotations

import json
import os
import struct
from dataclasses import dataclass, field

import numpy as np

DDS_MAGIC = 0x20534444
FOURCC_DX10 = 0x30315844
# dxgi id -> (block-compressed, bytes per block or per texel, channels stored)
FORMATS = {61: (False, 1, 1), 49: (False, 2, 2), 28: (False, 4, 4), 80: (True, 8, 1), 83: (True, 16, 2)}


class AssetError(ValueError):
    """The files do not describe an asset this format defines (the reader refuses rather than guesses)."""


@dataclass
class Dds:
    width: int
    height: int
    dxgi: int
    mips: list = field(default_factory=list)    # per level: float64 (h, w, channels) of sampled bytes, 0..255


def _bc4_blocks(raw: np.ndarray) -> np.ndarray:
    """(n, 8) BC4 blocks -> (n, 16) texel values as the standard palette gives them (floats: an interior entry such as
    (6 r0 + r1) / 7 is not a whole byte, and a sampler returns it unrounded)."""
    r0 = raw[:, 0].astype(np.float64)
    r1 = raw[:, 1].astype(np.float64)
    bits = np.zeros(raw.shape[0], dtype=np.uint64)
    for i in range(6):
        bits |= raw[:, 2 + i].astype(np.uint64) << np.uint64(8 * i)
    sel = np.stack([(bits >> np.uint64(3 * t)) & np.uint64(7) for t in range(16)], axis=1).astype(np.int64)
    pal = np.empty((raw.shape[0], 8), dtype=np.float64)
    pal[:, 0] = r0
    pal[:, 1] = r1
    eight = r0 > r1
    for i in range(1, 7):
        # r0 > r1: (7 - i) r0 + i r1 over 7 in entries 2..7
        pal[:, 1 + i] = np.where(eight, ((7 - i) * r0 + i * r1) / 7.0, 0.0)
    for i in range(1, 5):
        # r0 <= r1: (5 - i) r0 + i r1 over 5 in entries 2..5, then 0 and 255
        pal[:, 1 + i] = np.where(eight, pal[:, 1 + i], ((5 - i) * r0 + i * r1) / 5.0)
    pal[:, 6] = np.where(eight, pal[:, 6], 0.0)
    pal[:, 7] = np.where(eight, pal[:, 7], 255.0)
    return np.take_along_axis(pal, sel, axis=1)


def decode_bc(data: bytes, w: int, h: int, block_bytes: int) -> np.ndarray:
    """One level of BC4 (8-byte blocks) or BC5 (16: red's BC4 block, then green's) -> (h, w, channels) floats."""
    bx, by = (w + 3) // 4, (h + 3) // 4
    ch = block_bytes // 8
    raw = np.frombuffer(data, dtype=np.uint8).reshape(by * bx, ch, 8)
    out = np.empty((by * 4, bx * 4, ch), dtype=np.float64)
    for c in range(ch):
        v = _bc4_blocks(raw[:, c, :]).reshape(by, bx, 4, 4)
        out[:, :, c] = v.transpose(0, 2, 1, 3).reshape(by * 4, bx * 4)
    return out[:h, :w]


def read_dds(path: str) -> Dds:
    """The 4-byte magic, the 124-byte header, the 20-byte DX10 header, then every level base first, packed."""
    blob = open(path, "rb").read()
    if len(blob) < 148:
        raise AssetError(f"{path}: shorter than a DX10 header")
    d = struct.unpack_from("<32I", blob, 0)
    magic, size, _flags, height, width = d[0], d[1], d[2], d[3], d[4]
    mip_count = d[7]
    if magic != DDS_MAGIC or size != 124 or d[21] != FOURCC_DX10:
        raise AssetError(f"{path}: not a DX10 .dds")
    dxgi, dim, _misc, array = struct.unpack_from("<4I", blob, 128)
    if dxgi not in FORMATS or dim != 3 or array != 1:
        raise AssetError(f"{path}: format {dxgi}, dimension {dim}, array {array} is not an NNTC plane")
    compressed, unit, ch = FORMATS[dxgi]
    out = Dds(width, height, dxgi)
    off = 148
    w, h = width, height
    for _ in range(max(1, mip_count)):
        if compressed:
            n = ((w + 3) // 4) * ((h + 3) // 4) * unit
            out.mips.append(decode_bc(blob[off: off + n], w, h, unit))
        else:
            n = w * h * unit
            out.mips.append(np.frombuffer(blob[off: off + n], dtype=np.uint8).reshape(h, w, ch).astype(np.float64))
        if off + n > len(blob):
            raise AssetError(f"{path}: truncated at level {len(out.mips) - 1}")
        off += n
        w, h = max(1, w // 2), max(1, h // 2)
    if off != len(blob):
        raise AssetError(f"{path}: {len(blob) - off} bytes past the last level")
    return out


@dataclass
class Asset:
    meta: dict
    c0: int
    c1: int
    nout: int
    weights: np.ndarray            # (nout, nin)
    bias: np.ndarray               # (nout,)
    lo: list                       # per level: (C,) float64
    hi: list
    planes: list                   # per level: per mip, (h, w, C) sampled bytes 0..255
    width: int
    height: int
    source: tuple


def _file_entries(tex: dict) -> list:
    if "files" in tex:
        ents = tex["files"]
        if len(ents) != 2:
            raise AssetError("a two-file level must name two files")
        return [e if isinstance(e, dict) else {"file": e} for e in ents]
    if "file" not in tex:
        raise AssetError(f"level {tex.get('level')} names no file")
    return [{"file": tex["file"], "dxgi_format_id": tex.get("dxgi_format_id"),
             "channels_stored": tex.get("channels_stored")}]


def _replicate(k: int, bits: int) -> int:
    """Index k of `bits` bits replicated over a byte: its bits from the top down, repeated."""
    byte, sh = 0, 8 - bits
    while sh > -bits:
        byte |= (k << sh) if sh >= 0 else (k >> -sh)
        sh -= bits
    return byte & 255


def _check_palette(lvl: int, dq: dict, bits: list, lo: np.ndarray, hi: np.ndarray, compressed: bool):
    """A --l0 palette level publishes palette[c][k], the value index k decodes to; the format promises it equals the
    range rule on the byte a sampler returns, k / (2^bits - 1) of full scale when block-compressed and the replicated
    index over 255 when not, so a palette that disagrees (beyond float rounding) is refused."""
    pal = dq.get("palette")
    if pal is None or len(pal) != len(bits):
        raise AssetError(f"level {lvl}: a palette per channel is required")
    for c, (p, b) in enumerate(zip(pal, bits)):
        levels = (1 << b) - 1
        if len(p) != levels + 1:
            raise AssetError(f"level {lvl}: channel {c}'s palette has {len(p)} entries for {b} bits")
        for k, v in enumerate(p):
            frac = k / levels if compressed else _replicate(k, b) / 255.0
            if abs(v - (lo[c] + frac * (hi[c] - lo[c]))) > 1e-6:
                raise AssetError(f"level {lvl}: palette entry {k} of channel {c} is not what its byte decodes to")


def load(path: str) -> Asset:
    """The descriptor (PREFIX_nntc.json, or a prefix) and the .dds files it names, checked as the format requires."""
    if not path.endswith(".json"):
        path = path + "_nntc.json"
    meta = json.load(open(path))
    directory = os.path.dirname(os.path.abspath(path))
    if meta.get("format") not in ("nntc-dds-1", "ntc-dds-1"):
        raise AssetError(f"format {meta.get('format')!r}")
    texs = meta["textures"]
    if len(texs) != 2 or any(t["level"] != i for i, t in enumerate(texs)):
        raise AssetError("the textures must be level 0 then level 1")
    dec = meta["decoder"]
    if dec["type"] != "bilinear" or dec.get("hidden", []) != [] or dec.get("output", "identity") != "identity":
        raise AssetError("the decoder is not the one affine bilinear layer")
    if dec["terms"].split() != ["a", "b", "sc"]:
        raise AssetError(f"terms {dec['terms']!r}")
    c0, c1 = int(dec["C0"]), int(dec["C1"])
    nin, nout = c1 + c0 + c0 * c1, int(dec["nout"])
    if int(dec["nin"]) != nin or len(dec["layers"]) != 1:
        raise AssetError("nin does not match the terms, or the decoder has more than one layer")
    layer = dec["layers"][0]
    if layer["rows"] != nout or layer["cols"] != nin:
        raise AssetError("the layer is not nout x nin")
    wts = np.asarray(layer["weights"], dtype=np.float64).reshape(nout, nin)
    bias = np.asarray(layer["bias"], dtype=np.float64)
    lo, hi, planes = [], [], []
    for lvl, tex in enumerate(texs):
        used = int(tex["channels_used"])
        if used != (c0 if lvl == 0 else c1):
            raise AssetError(f"level {lvl}: channels_used {used} does not match the decoder")
        bits = tex.get("bits_per_channel", [8] * used)
        if any(not (1 <= b <= 8) for b in bits) or (lvl == 1 and any(b < 4 for b in bits)):
            raise AssetError(f"level {lvl}: bits_per_channel {bits}")
        files = []
        for ent in _file_entries(tex):
            name = ent["file"]
            dds = read_dds(name if os.path.isabs(name) else os.path.join(directory, name))
            stored = FORMATS[dds.dxgi][2]
            if ent.get("dxgi_format_id") is not None and ent["dxgi_format_id"] != dds.dxgi:
                raise AssetError(f"{name}: header format {dds.dxgi}, descriptor {ent['dxgi_format_id']}")
            if ent.get("channels_stored") is not None and ent["channels_stored"] != stored:
                raise AssetError(f"{name}: stores {stored} channels, descriptor says {ent['channels_stored']}")
            files.append(dds)
        if len(files) == 2 and [len(f.mips) for f in files] != [len(files[0].mips)] * 2:
            raise AssetError(f"level {lvl}: the two files hold different level counts")
        mips = [np.concatenate([f.mips[m] for f in files], axis=2)[:, :, :used] for m in range(len(files[0].mips))]
        if mips[0].shape[2] != used:
            raise AssetError(f"level {lvl}: the files store {mips[0].shape[2]} channels, {used} used")
        sizes = tex.get("mip_sizes")
        if sizes is not None and [[m.shape[1], m.shape[0]] for m in mips] != sizes:
            raise AssetError(f"level {lvl}: mip sizes differ from the descriptor's")
        dq = tex["dequantise"]
        lo.append(np.asarray(dq["lo"], dtype=np.float64)[:used])
        hi.append(np.asarray(dq["hi"], dtype=np.float64)[:used])
        if dq.get("kind") == "palette":
            _check_palette(lvl, dq, bits, lo[-1], hi[-1], "bc_palette" in tex)
        planes.append(mips)
    src = meta.get("source", {})
    return Asset(meta, c0, c1, nout, wts, bias, lo, hi, planes, int(meta["decode"]["width"]),
                 int(meta["decode"]["height"]), (int(src.get("width", 0)), int(src.get("height", 0))))


def _taps(n_out: int, n_in: int):
    """Bilinear taps at the centres of n_out pixels over a plane of n_in texels: x = (p + 0.5) n_in / n_out - 0.5,
    the two indices floor(x) and floor(x) + 1 clamped into [0, n_in - 1], weights 1 - f and f."""
    x = (np.arange(n_out, dtype=np.float64) + 0.5) * n_in / n_out - 0.5
    i0 = np.floor(x)
    f = x - i0
    i0 = i0.astype(np.int64)
    return np.clip(i0, 0, n_in - 1), np.clip(i0 + 1, 0, n_in - 1), f


def sample(plane: np.ndarray, w: int, h: int, rows: slice = slice(None)) -> np.ndarray:
    """Plane (H, W, C) sampled bilinearly at the centres of a w x h grid's pixels (rows `rows` of it)."""
    x0, x1, fx = _taps(w, plane.shape[1])
    y0, y1, fy = _taps(h, plane.shape[0])
    y0, y1, fy = y0[rows], y1[rows], fy[rows]
    fx = fx[None, :, None]
    fy = fy[:, None, None]
    top = plane[y0][:, x0] * (1 - fx) + plane[y0][:, x1] * fx
    bot = plane[y1][:, x0] * (1 - fx) + plane[y1][:, x1] * fx
    return top * (1 - fy) + bot * fy


def decode(asset: Asset, mip: int = 0, band: int = 256) -> np.ndarray:
    """Output mip `mip` from mip `mip` of both latents (the 1:1 rule), at its pixel centres, as 8-bit RGB per texture:
    (h, w, 3T) uint8. Rows are decoded `band` at a time to bound the float64 working set."""
    p0 = asset.planes[0][mip]
    p1 = asset.planes[1][mip]
    h, w = p0.shape[:2]
    out = np.empty((h, w, asset.nout), dtype=np.uint8)
    for r in range(0, h, band):
        rows = slice(r, min(h, r + band))
        s0 = sample(p0, w, h, rows) / 255.0
        s1 = sample(p1, w, h, rows) / 255.0
        z0 = asset.lo[0] + s0 * (asset.hi[0] - asset.lo[0])
        z1 = asset.lo[1] + s1 * (asset.hi[1] - asset.lo[1])
        sc = (z0[:, :, :, None] * z1[:, :, None, :]).reshape(z0.shape[0], w, asset.c0 * asset.c1)
        phi = np.concatenate([z1, z0, sc], axis=2)
        y = phi @ asset.weights.T + asset.bias
        # floor(x + 0.5) after the clamp: half away from zero, as the encoder's lroundf rounds a non-negative value.
        out[rows] = np.floor(np.clip(y, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)
    return out


def psnr(rgb: np.ndarray, textures: list) -> list:
    """Per texture, 10 log10(255^2 / MSE) over the source extent and its three channels (src/image.h:387-407)."""
    out = []
    for t, src in enumerate(textures):
        h, w = src.shape[:2]
        d = rgb[:h, :w, 3 * t: 3 * t + 3].astype(np.int64) - src.astype(np.int64)
        mse = float((d * d).sum()) / src.size
        out.append(10.0 * np.log10(255.0 * 255.0 / mse) if mse > 0 else float("inf"))
    return out
