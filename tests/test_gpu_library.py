"""A library of materials encoded side by side (nntc_encode_library) against each material encoded alone.

The library runs its materials on several host threads at once, each with its own context and stream, so their
kernels overlap on the GPU; each material must still come out exactly as an encode of it alone: the same report
figures and the same asset bytes. The library here is the batching workload's generator (uneven sizes, non-square,
one to six textures, the four layouts) at small sizes.
"""

import hashlib
import os

import pytest

from nntc_amd import abi, encoder, synth

pytestmark = pytest.mark.gpu


def small_library(count):
    specs = synth.library_spec(count, config=15)
    for i, s in enumerate(specs):
        # Sizes a twentieth of the batching workload's, so the test runs in seconds; one non-multiple-of-4 extent;
        # every layout of the workload among the materials.
        s["c0"], s["c1"] = synth.LIBRARY_LAYOUTS[i % len(synth.LIBRARY_LAYOUTS)]
        w, h = s["size"]
        s["size"] = (max(16, (w // 20) // 4 * 4), max(16, (h // 20) // 4 * 4 + (2 if i == 1 else 0)))
    return specs, synth.library(specs)


def asset_digest(prefix):
    """The asset's files by suffix; the descriptor names its .dds files by the prefix, so that name is taken out."""
    out = {}
    for suffix in ("_lat0.dds", "_lat0a.dds", "_lat0b.dds", "_lat1.dds", "_nntc.json"):
        if os.path.exists(prefix + suffix):
            data = open(prefix + suffix, "rb").read()
            if suffix.endswith(".json"):
                data = data.replace(os.path.basename(prefix).encode(), b"PREFIX")
            out[suffix] = hashlib.sha256(data).hexdigest()
    return out


@pytest.mark.parametrize("form", [abi.FORM_FAST, abi.FORM_EXACT], ids=["fast", "exact"])
def test_library_matches_materials_alone(form, tmp_path, hip_arm):
    specs, textures = small_library(8)
    assert {(s["c0"], s["c1"]) for s in specs} == set(synth.LIBRARY_LAYOUTS)
    alone = []
    for i, (s, tex) in enumerate(zip(specs, textures)):
        e = encoder.Encoder(tex, c0=s["c0"], c1=s["c1"], rounds=6, tol=0.0, quiet=1, form=form)
        rep = e.run()
        e.write(str(tmp_path / f"alone{i}"))
        alone.append((rep.e_shipped, [rep.psnr[t] for t in range(len(tex))], rep.rounds_used))
        e.close()
    mats = [dict(textures=tex, c0=s["c0"], c1=s["c1"], prefix=str(tmp_path / f"lib{i}"))
            for i, (s, tex) in enumerate(zip(specs, textures))]
    print(f"library of {len(mats)} on 4 workers, form {form}", flush=True)
    reps = encoder.encode_library(mats, workers=4, rounds=6, tol=0.0, quiet=1, form=form)
    for i, rep in enumerate(reps):
        got = (rep.e_shipped, [rep.psnr[t] for t in range(specs[i]["textures"])], rep.rounds_used)
        assert got == alone[i], f"material {i}: {got} vs alone {alone[i]}"
        assert asset_digest(str(tmp_path / f"lib{i}")) == asset_digest(str(tmp_path / f"alone{i}")), f"material {i}"
        assert len(asset_digest(str(tmp_path / f"lib{i}"))) >= 3


def test_library_from_device_memory(tmp_path, hip_arm):
    """The same library with its textures resident in device memory (the benchmark's inputs)."""
    specs, textures = small_library(4)
    dev = [[abi.DeviceBytes(t) for t in tex] for tex in textures]
    try:
        mats = [dict(device=(tex[0].shape[1], tex[0].shape[0], [d.address for d in ds]), c0=s["c0"], c1=s["c1"])
                for s, tex, ds in zip(specs, textures, dev)]
        a = encoder.encode_library(mats, workers=3, rounds=4, tol=0.0, quiet=1, form=abi.FORM_FAST)
        hosts = [dict(textures=tex, c0=s["c0"], c1=s["c1"]) for s, tex in zip(specs, textures)]
        b = encoder.encode_library(hosts, workers=1, rounds=4, tol=0.0, quiet=1, form=abi.FORM_FAST)
        assert [r.e_shipped for r in a] == [r.e_shipped for r in b]
    finally:
        for ds in dev:
            for d in ds:
                d.close()
