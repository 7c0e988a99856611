"""The launch-shape tuner: the forms it chooses between write the same bits, and it records a choice per shape.

The fast sweeps at three or four channels have two launch forms, one lane a texel and four lanes a texel (one channel
row each); the tuner times both on the first call of a shape and keeps the faster. The choice may differ between runs
and machines, so the forms must not differ in anything but time: a whole fast-form encode with either form forced, and
with the tuner's choice, must give the same .dds blocks, level-1 indices and objective.
"""

import ctypes as C
import hashlib

import pytest

from nntc_amd import abi, encoder, synth

pytestmark = pytest.mark.gpu


def digest(e, c0):
    parts = [e.level0_blocks(p, f).tobytes() for p in range(e.desc.nplanes) for f in range(1 if c0 <= 2 else 2)]
    parts += [e.level1_indices(p).tobytes() for p in range(e.desc.nplanes)]
    return hashlib.sha256(b"".join(parts)).hexdigest()


def tuning_table():
    lib = C.CDLL(abi.HIP_LIB, mode=C.RTLD_GLOBAL)
    f = lib.nntc_hip_tuning_entry
    ip = C.POINTER(C.c_int)
    f.argtypes = [C.c_int, ip, ip, ip, ip, ip, C.POINTER(C.c_double), ip]
    out = []
    i = 0
    while True:
        fam, ch, bk, choice, n = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
        forms = (C.c_int * 4)()
        ms = (C.c_double * 4)()
        if f(i, C.byref(fam), C.byref(ch), C.byref(bk), C.byref(choice), forms, ms, C.byref(n)) != 0:
            return out
        out.append({"family": fam.value, "channels": ch.value, "bucket": bk.value, "choice": choice.value,
                    "forms": list(forms[: n.value]), "ms": list(ms[: n.value])})
        i += 1


@pytest.mark.parametrize("name,spec", [
    ("small_2x100x76", None),
    ("small_6x100x76_c4c4", dict(textures=6, size=(100, 76), period=32, smooth=5, detail=1.0, c0=4, c1=4, config=9)),
])
def test_sweep_forms_write_the_same_bits(name, spec, hip_arm):
    lib = C.CDLL(abi.HIP_LIB, mode=C.RTLD_GLOBAL)
    lib.nntc_hip_tuning_reset()
    if spec is None:
        spec = synth.WORKLOADS[name]
        tex = synth.workload(name)
    else:
        tex = synth.make_material(**{k: v for k, v in spec.items() if k not in ("c0", "c1")})
    seen = {}
    for lanes in (1, 4, -1):
        print(f"{name}: sweep lanes {lanes}", flush=True)
        e = encoder.Encoder(tex, c0=spec["c0"], c1=spec["c1"], rounds=6, tol=0.0, quiet=1, form=abi.FORM_FAST,
                            sweep_lanes=lanes)
        rep = e.run()
        seen[lanes] = (rep.e_shipped, digest(e, spec["c0"]))
        e.close()
    assert seen[1] == seen[4], "one and four lanes a texel gave different bits"
    assert seen[-1] == seen[1]
    table = tuning_table()
    # The tuned run met both latents' shapes (C1 >= 3 and C0 >= 3) and chose one of the two forms for each.
    assert {t["channels"] for t in table} == {spec["c0"], spec["c1"]}
    for t in table:
        assert t["family"] == 0 and t["forms"] == [1, 4] and t["choice"] in (1, 4)
        assert t["choice"] == t["forms"][min(range(2), key=lambda i: t["ms"][i])]


def graph_counts(e):
    lib = C.CDLL(abi.HIP_LIB, mode=C.RTLD_GLOBAL)
    ctx_of = encoder.lib().nntc_encoder_context
    ctx_of.restype = C.c_void_p
    ctx_of.argtypes = [C.c_void_p]
    f = lib.nntc_hip_graph_counts
    f.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    caps, runs = C.c_longlong(), C.c_longlong()
    assert f(ctx_of(e.h), C.byref(caps), C.byref(runs)) == 0
    return caps.value, runs.value


@pytest.mark.parametrize("name", ["plain_cpu_4x256", "small_2x100x76"])
def test_graph_replayed_cg_writes_the_same_bits(name, hip_arm):
    """The fast conjugate gradients with each probe period replayed from a captured HIP graph, against the same
    launches made one by one: identical bits over a whole encode (every replay's output feeds the next), on two runs in
    a row of one context (the graphs captured by the first are replayed by the second, with the driver's own eager
    work and readbacks between the replays)."""
    spec = synth.WORKLOADS[name]
    tex = synth.workload(name)
    out = {}
    for graphs in (0, 1):
        print(f"{name}: graphs {graphs}", flush=True)
        e = encoder.Encoder(tex, c0=spec["c0"], c1=spec["c1"], rounds=8, tol=0.0, quiet=1, form=abi.FORM_FAST,
                            graphs=graphs)
        a = e.run()
        da = digest(e, spec["c0"])
        caps1, runs1 = graph_counts(e)
        b = e.run()
        caps2, runs2 = graph_counts(e)
        assert digest(e, spec["c0"]) == da and b.e_shipped == a.e_shipped
        out[graphs] = (a.e_shipped, da)
        if graphs:
            assert caps1 > 0 and runs1 > caps1, (caps1, runs1)
            assert caps2 == caps1 and runs2 > runs1, "the second run should replay the first run's graphs"
        else:
            assert (caps2, runs2) == (0, 0)
        e.close()
    assert out[0] == out[1]
