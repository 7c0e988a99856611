"""The allocation-free entry point, nntc_hip_texture_sse_noalloc, against nntc_hip_texture_sse.

Both sum each texture's squared byte error of the decoded base plane against the source over a top-left extent; the
allocation-free one decodes and sums in one pass with its counters in the library's fixed device array and writes to
device memory. On the same state the two must give the same integers, on two calls in a row (the second over a
smaller extent, so the counters' contents from the first call must not leak into the second), and the limits must be
refused with a message.
"""

import ctypes as C

import numpy as np
import pytest

from nntc_amd import abi, chain, synth
from schedule import Options, steps

pytestmark = pytest.mark.gpu


def test_texture_sse_noalloc_matches_allocating_form(hip_arm):
    name = "plain_cpu_4x256"
    spec = synth.WORKLOADS[name]
    tex = synth.workload(name)
    planes, ch = chain.source_chain(tex)
    _, per_site = chain.mip_weights(planes)
    desc = abi.make_desc(len(tex), spec["c0"], spec["c1"], planes, chain.output_weights(len(tex)))
    s = abi.Session(hip_arm, desc, [c.reshape(-1) for c in ch])
    print("texture_sse_noalloc plain_cpu_4x256", flush=True)
    # A state with a fitted decoder and seeded latents: the schedule up to the first round's block (b).
    for i, (n, _) in enumerate(steps(s, Options(), per_site)):
        if n == "solve_level1_all":
            break
    f = hip_arm.lib.nntc_hip_texture_sse_noalloc
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    base = s.plane(0)
    out = abi.DeviceBytes(np.full(desc.nout // 3, 12345, np.uint64))
    for wh in ((base.w0, base.h0), (base.w0 - 37, base.h0 // 3)):
        want = s.texture_sse(*wh)
        assert f(s.ctx, wh[0], wh[1], out.address, None) == 0, hip_arm.fn["last_error"]().decode()
        got = [int(x) for x in out.read()]
        print(f"  extent {wh}: {got}", flush=True)
        assert got == want, (wh, got, want)
        assert all(x > 0 for x in got)
    assert f(s.ctx, base.w0 + 1, base.h0, out.address, None) != 0
    assert b"exceeds" in hip_arm.fn["last_error"]()
    out.close()
    s.close()
