"""The --check twins of the gfx950 arm (objective_check_host, solve_level1_dense_host) against the device's answers
and against the reference CPU backend's own twins, on a state the encode's schedule reaches.

The reference's check mode holds the twins to its bars (src/backend_check.cpp:91-151): E's brute-force twin within
BAR_SUM 1e-9 relative of the objective, and the dense level-1 solve within 1e-5 of the correction's range of the
device's conjugate gradients. Here both arms are driven in lockstep (exact form) to the first block (b) of round 1;
the twins of the two arms, which sum in different orders (and solve by Cholesky here, Gauss-Jordan there), agree far
inside those bars.
"""

import numpy as np
import pytest

from schedule import Options, steps
from test_gpu_lockstep import sessions

pytestmark = pytest.mark.gpu

K = 4


@pytest.mark.parametrize("name", ["small_2x100x76_c3c3", "small_1x64x40_c1c2"])
def test_check_twins(name, hip_arm, ref_arm):
    a, b, per_site = sessions(name, hip_arm, ref_arm)
    o = Options()
    ga, gb = steps(a, o, per_site), steps(b, o, per_site)
    # Up to the state block (b) of round 1 starts from: the decoder solved for that round.
    while True:
        sa, ia = next(ga)
        sb, ib = next(gb)
        assert sa == sb
        if sa == "solve_decoder" and ia.get("round") == 1:
            break
    e_dev = a.objective(K, per_site)
    e_twin = a.objective_check(K, per_site)
    e_ref = b.objective_check(K, per_site)
    print(f"{name}: E {e_dev.e!r} device, {e_twin.e!r} twin, {e_ref.e!r} reference twin", flush=True)
    assert abs(e_twin.e - e_dev.e) <= 1e-9 * e_dev.e
    assert abs(e_twin.e - e_ref.e) <= 1e-12 * e_ref.e
    for p in range(a.nplanes):
        assert abs(e_twin.e_plane[p] - e_ref.e_plane[p]) <= 1e-12 * e_ref.e_plane[p]
        assert abs(e_twin.centre_plane[p] - e_ref.centre_plane[p]) <= 1e-12 * e_ref.centre_plane[p]
    prev = [a.v1(p) for p in range(a.nplanes)]
    rep_a, _ = a.block("solve_level1_all", K, o.ridge)
    b.block("solve_level1_all", K, o.ridge)
    checked = 0
    for p in range(a.nplanes):
        if prev[p].size > 1024:
            assert a.level1_dense(K, p, rep_a[p].lambda_, prev[p]) is None
            continue
        dense = a.level1_dense(K, p, rep_a[p].lambda_, prev[p])
        dense_ref = b.level1_dense(K, p, rep_a[p].lambda_, prev[p])
        cg = a.delta1(p)
        span = max(np.ptp(dense_ref), 1e-30)
        print(f"  plane {p}: {prev[p].size} unknowns, |dense - reference dense| {np.abs(dense - dense_ref).max():.3e}, "
              f"|dense - CG| {np.abs(dense - cg).max():.3e}, range {span:.3e}", flush=True)
        assert np.abs(dense - dense_ref).max() <= 1e-9 * span
        assert np.abs(dense - cg).max() <= 1e-5 * span
        checked += 1
    assert checked >= 2
