"""The oracle pinned: oracle/restate.py against the reference's own known-answer vectors, closed forms, brute force and
textbook linear algebra, and cross-checked against the compiled reference (oracle/_ref) where it is built.

CPU only. Each check is chosen so that a plausible slip in the restatement (a fused multiply-add, a dropped term, a
swapped operand, a wrong tie rule) fails it.
"""

import json
import os
import struct

import numpy as np
import pytest

import oracle.restate as R
from nntc_amd import abi, chain, synth

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits_to_float(u):
    return np.frombuffer(struct.pack("<I", u), dtype=np.float32)[0]


def float_bits(x):
    return struct.unpack("<I", np.float32(x).tobytes())[0]


W = json.load(open(os.path.join(GOLDEN, "fp_contract_witnesses.json")))


@pytest.mark.parametrize("w", W["level1_value"])
def test_level1_value_witness(w):
    """The reference's 20 witnesses where a fused multiply-add changes the last bit (tests/fp_contract_check.cpp)."""
    got = R.level1_value(bits_to_float(w["lo"]), bits_to_float(w["hi"]), w["bits"], w["k"])
    assert float_bits(got) == w["separate"]
    assert float_bits(got) != w["fused"]


@pytest.mark.parametrize("w", W["bc0_value"])
def test_bc0_value_witness(w):
    got = R.bc0_value(bits_to_float(w["lo"]), bits_to_float(w["hi"]), float(w["byte"]))
    assert float_bits(got) == w["separate"]
    assert float_bits(got) != w["fused"]


def test_replicate_index_closed_form():
    # 8 bits is the index itself, 4 bits k * 17, 1 bit 0 / 255, 2 bits k * 85.
    for k in range(256):
        assert R.replicate_index(k, 8) == k
    for k in range(16):
        assert R.replicate_index(k, 4) == k * 17
    for k in range(4):
        assert R.replicate_index(k, 2) == k * 85
    assert [R.replicate_index(k, 1) for k in (0, 1)] == [0, 255]
    # 3 bits: k k k, i.e. (k << 5) | (k << 2) | (k >> 1)
    for k in range(8):
        assert R.replicate_index(k, 3) == ((k << 5) | (k << 2) | (k >> 1))


@pytest.mark.parametrize("bits", [4, 5, 8])
def test_level1_index_is_the_nearest_value(bits):
    """level1_index against brute force over every index, and every index re-quantising to itself."""
    rng = np.random.default_rng(bits)
    for _ in range(20):
        lo, hi = sorted(rng.uniform(-3, 3, 2))
        lo, hi = np.float32(lo), np.float32(hi)
        levels = (1 << bits) - 1
        grid = np.array([R.level1_value(lo, hi, bits, k) for k in range(levels + 1)], np.float64)
        for k in range(levels + 1):
            assert R.level1_index(lo, hi, bits, grid[k]) == k
        for v in rng.uniform(lo - 0.5, hi + 0.5, 40).astype(np.float32):
            d = np.abs(float(v) - grid)
            assert R.level1_index(lo, hi, bits, v) == int(np.flatnonzero(d == d.min())[0])


def test_bc4_palette_closed_form():
    for r0, r1 in [(255, 0), (200, 13), (7, 6)]:
        p = R.bc4_palette(r0, r1)
        want = [r0, r1] + [((8 - i) * r0 + (i - 1) * r1) / 7.0 for i in range(2, 8)]
        np.testing.assert_allclose(p, want, rtol=0, atol=1e-4)
    p = R.bc4_palette(10, 200)
    assert p[6] == 0.0 and p[7] == 255.0


@pytest.mark.parametrize("bits", [1, 2, 3])
def test_bc4_lossless_at_three_bits_and_below(bits):
    """At 1-3 bits the pack is exact: every index decodes to k * 255 / (2^bits - 1) (shared/bc_pack.h:1-29)."""
    levels = (1 << bits) - 1
    rng = np.random.default_rng(bits)
    for _ in range(50):
        ks = rng.integers(0, levels + 1, 16)
        v = np.array([R.replicate_index(int(k), bits) for k in ks], np.uint8)
        r0, r1, sel, err = R.bc4_encode_block(v, bits)
        dec = R.bc4_decode_block(R.bc4_pack(r0, r1, sel))
        np.testing.assert_allclose(dec, ks * 255.0 / levels, atol=1e-3)
        assert err == 0.0


def test_bc4_eight_bit_pack_is_nearest_entry():
    """At 8 bits the endpoints are the block's extremes and every selector the nearest palette entry."""
    rng = np.random.default_rng(8)
    for _ in range(200):
        v = rng.integers(0, 256, 16).astype(np.uint8)
        r0, r1, sel, err = R.bc4_encode_block(v, 8)
        assert r0 == v.max() and r1 == v.min()
        dec = R.bc4_decode_block(R.bc4_pack(r0, r1, sel))
        if r0 == r1:
            assert (dec == r0).all()
            continue
        p = R.bc4_palette(r0, r1)
        best = np.min(np.abs(p[None, :] - v[:, None].astype(np.float32)), axis=1)
        np.testing.assert_array_equal(np.abs(dec - v.astype(np.float32)), best)
        assert abs(err - float(np.sum(best.astype(np.float64) ** 2))) < 1e-6


def small_model(seed=0, textures=1, w0=16, h0=12, c0=2, c1=2, k=4):
    planes = abi.plane_sizes(w0, h0, True, 4)
    rng = np.random.default_rng(seed)
    nout = 3 * textures
    nin = c1 + c0 + c0 * c1
    v0s = [rng.uniform(-1, 1, (p[0] * p[1] * c0)).astype(np.float32) for p in planes]
    v1s = [rng.uniform(-1, 1, (p[2] * p[3] * c1)).astype(np.float32) for p in planes]
    tex = synth.make_material(textures, (w0, h0), 8, 3, 1.0, config=11, material=seed)
    _, srcs = chain.source_chain(tex, True, 4)
    w = rng.normal(0, 0.3, (nout, nin)).astype(np.float32)
    b = rng.normal(0.5, 0.1, nout).astype(np.float32)
    cw = chain.output_weights(textures)
    _, per_site = chain.mip_weights(planes, k=k)
    return dict(planes=planes, v0s=v0s, v1s=v1s, srcs=[s.reshape(-1) for s in srcs], w=w, b=b, cw=cw, k=k,
                per_site=per_site, c0=c0, c1=c1, nout=nout)


def test_objective_sites_at_k0_is_centre_mse():
    """At K = 0 and one plane E is the cw-weighted centre MSE computed directly (src/objective.cu; run_checks.py)."""
    m = small_model(k=0)
    p0 = [m["planes"][0]]
    e, acc = R.objective_brute(p0, m["v0s"][:1], m["v1s"][:1], m["srcs"][:1], m["w"], m["b"], m["cw"], 0, [1.0])
    w0, h0, w1, h1 = p0[0]
    c = R.twin_sample(m["v1s"][0].reshape(h1, w1, -1), *[(np.mgrid[0:h0, 0:w0][i].reshape(-1) + 0.5) / n
                                                          for i, n in ((1, w0), (0, h0))])
    s = m["v0s"][0].reshape(-1, m["c0"]).astype(np.float64)
    y = R.features(s, c) @ m["w"].astype(np.float64).T + m["b"].astype(np.float64)
    t = m["srcs"][0].reshape(-1, m["nout"]).astype(np.float64)
    direct = np.mean((y - t) ** 2 * np.asarray(m["cw"])[None, :]) / np.mean(m["cw"])
    assert abs(e - direct) <= 1e-12 * direct


def test_normal_equations_solve_is_least_squares():
    """The decoder from A and rhs equals numpy's weighted least squares on the stacked site rows."""
    m = small_model(seed=3)
    a, rhs = R.normal_equations(m["planes"], m["v0s"], m["v1s"], m["srcs"], m["c0"], m["c1"], m["nout"], m["k"],
                                m["per_site"])
    x = np.linalg.solve(a, rhs.T)
    rows, tg, wt = [], [], []
    for plane, v0, v1, src, om in zip(m["planes"], m["v0s"], m["v1s"], m["srcs"], m["per_site"]):
        phi, t = R.site_rows(plane, v0, v1, src, m["c0"], m["c1"], m["nout"], m["k"])
        rows.append(np.concatenate([phi.astype(np.float64), np.ones((phi.shape[0], 1))], axis=1))
        tg.append(t.astype(np.float64))
        wt.append(np.full(phi.shape[0], np.sqrt(om)))
    A = np.concatenate(rows) * np.concatenate(wt)[:, None]
    T = np.concatenate(tg) * np.concatenate(wt)[:, None]
    lsq = np.linalg.lstsq(A, T, rcond=None)[0]
    np.testing.assert_allclose(x, lsq, rtol=1e-7, atol=1e-9)


def test_dense_level1_gradient_matches_finite_differences():
    """The dense quadratic's gradient is E's gradient in the level-1 plane (brute-force E, central differences)."""
    m = small_model(seed=5, w0=8, h0=8)
    p0 = m["planes"][:1]
    h, g = R.dense_level1_system(p0[0], m["v0s"][0], m["v1s"][0], m["srcs"][0], m["w"], m["b"], m["cw"], m["c0"],
                                 m["c1"], m["nout"], m["k"])
    assert np.allclose(h, h.T)

    def esum(v1):
        _, acc = R.objective_brute(p0, m["v0s"][:1], [v1], m["srcs"][:1], m["w"], m["b"], m["cw"], m["k"], [1.0])
        return acc[0][0]

    for i in (0, 3, 7):
        step = 1e-3
        vp = m["v1s"][0].copy()
        vm = m["v1s"][0].copy()
        vp[i] += step
        vm[i] -= step
        fd = (esum(vp) - esum(vm)) / (2 * (np.float64(vp[i]) - np.float64(m["v1s"][0][i])) * 2) * 2
        # E = sum cw e^2, so dE/dv = 2 grad; the quadratic is exact, so the central difference is exact up to rounding
        assert abs(fd - 2 * g[i]) <= 1e-5 * max(1.0, abs(2 * g[i]))


# ---------------------------------------------------------------------------------------------------------------------
# Against the compiled reference
# ---------------------------------------------------------------------------------------------------------------------

def ref_session(ref_arm, m):
    desc = abi.make_desc(m["nout"] // 3, m["c0"], m["c1"], m["planes"], m["cw"])
    s = abi.Session(ref_arm, desc, m["srcs"])
    for p in range(len(m["planes"])):
        s.upload_v0(p, m["v0s"][p])
        s.upload_v1(p, m["v1s"][p])
    s.set_decoder(m["w"], m["b"])
    return s


def test_objective_matches_reference(ref_arm):
    m = small_model(seed=1, textures=2, w0=20, h0=16)
    s = ref_session(ref_arm, m)
    o = s.objective(m["k"], m["per_site"])
    e, _ = R.objective_brute(m["planes"], m["v0s"], m["v1s"], m["srcs"], m["w"], m["b"], m["cw"], m["k"],
                             m["per_site"])
    assert abs(o.e - e) <= 1e-12 * e


def test_decoder_matches_reference(ref_arm):
    m = small_model(seed=2, textures=2, w0=20, h0=16)
    s = ref_session(ref_arm, m)
    s.state.has_decoder = 0
    s.solve_decoder(m["k"], m["per_site"])
    w, b = s.decoder()
    a, rhs = R.normal_equations(m["planes"], m["v0s"], m["v1s"], m["srcs"], m["c0"], m["c1"], m["nout"], m["k"],
                                m["per_site"])
    ridge = 1e-9 * np.trace(a) / a.shape[0]
    x = np.linalg.solve(a + ridge * np.eye(a.shape[0]), rhs.T).T
    np.testing.assert_allclose(w, x[:, :-1], rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(b, x[:, -1], rtol=1e-4, atol=1e-5)


def test_level1_solve_matches_dense_reference(ref_arm):
    """Block (b)'s PCG correction against the dense direct solve of the restated quadratic (DENSE_BAR 1e-5 of range,
    tests/cpu_check.cpp:839)."""
    m = small_model(seed=4, textures=1, w0=16, h0=16)
    m["planes"] = m["planes"][:1]
    for key in ("v0s", "v1s", "srcs", "per_site"):
        m[key] = m[key][:1]
    s = ref_session(ref_arm, m)
    ridge = 1e-4
    rep, _ = s.block("solve_level1_all", m["k"], ridge)
    delta = s.delta1(0)
    h, g = R.dense_level1_system(m["planes"][0], m["v0s"][0], m["v1s"][0], m["srcs"][0], m["w"], m["b"], m["cw"],
                                 m["c0"], m["c1"], m["nout"], m["k"])
    lam = ridge * np.trace(h) / h.shape[0]
    assert abs(rep[0].lambda_ - lam) <= 1e-9 * lam
    d = np.linalg.solve(h + lam * np.eye(h.shape[0]), -g)
    assert np.max(np.abs(delta - d)) <= 1e-5 * (d.max() - d.min())


def test_reference_schedule_matches_golden(ref_arm):
    """The lockstep schedule on the compiled reference reproduces the reference binary's own run."""
    from schedule import Options, steps

    g = json.load(open(os.path.join(GOLDEN, "smoke_1x64x40", "report.json")))
    spec = synth.WORKLOADS["smoke_1x64x40"]
    tex = synth.workload("smoke_1x64x40")
    planes, ch = chain.source_chain(tex)
    _, per_site = chain.mip_weights(planes)
    desc = abi.make_desc(1, spec["c0"], spec["c1"], planes, chain.output_weights(1))
    s = abi.Session(ref_arm, desc, [c.reshape(-1) for c in ch])
    last = None
    for name, info in steps(s, Options(rounds=4), per_site):
        last = (name, info)
    assert last[0] == "shipped"
    assert f"{last[1]['e']:.9e}" == f"{g['e_shipped']:.9e}"
