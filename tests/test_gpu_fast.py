"""The fast form's own kernels against the exact form on the gfx950 arm.

The exact form is the reference CPU backend bit for bit (test_gpu_lockstep.py); the fast form replaces its serial
folds with trees and two kernels with new ones: the tiled objective (fp32 sampling and decode, LDS-staged tiles) and
the decoder Gram on the matrix cores (bf16 three-part split products, fp32 accumulation over 64 sites of a wavefront,
fp64 across rounds, wavefronts and tiles). Each is checked here against the exact form on the same state: decoder coefficients within
1e-5 of the coefficients' scale, E within 1e-4 relative. small_6x100x76_c4c4 has D = mm + nout = 25 + 18 = 43 Gram
features, so its Gram runs the two-block (64-feature) form of the matrix-core kernel. The per-call check over a whole
encode schedule is tests/test_gpu_percall.py.
"""

import numpy as np
import pytest

from nntc_amd import abi, chain
from test_gpu_lockstep import material

pytestmark = pytest.mark.gpu


def session(hip_arm, name, form):
    tex, c0, c1 = material(name)
    planes, ch = chain.source_chain(tex)
    _, per_site = chain.mip_weights(planes)
    desc = abi.make_desc(len(tex), c0, c1, planes, chain.output_weights(len(tex)))
    s = abi.Session(hip_arm, desc, [c.reshape(-1) for c in ch])
    s.call("set_form", form)
    return s, per_site


@pytest.mark.parametrize("name", ["plain_cpu_4x256", "small_2x100x76", "small_6x100x76_c4c4"])
def test_fast_gram_and_objective_match_exact(name, hip_arm):
    print(f"fast gram/objective {name}", flush=True)
    ex, per_site = session(hip_arm, name, abi.FORM_EXACT)
    fa, _ = session(hip_arm, name, abi.FORM_FAST)
    for s in (ex, fa):
        s.init_level0(luma_channel0=False)
        s.init_level1(pca=False)
    # Identical latents (the inits share their kernels): the two objectives of the same state.
    e_ex = ex.objective(4, per_site).e
    e_fa = fa.objective(4, per_site).e
    assert abs(e_fa - e_ex) <= 1e-4 * e_ex, (e_fa, e_ex)
    # Block (a) on the same state: the Gram's two forms through the same host solve.
    ex.solve_decoder(4, per_site)
    fa.solve_decoder(4, per_site)
    (we, be), (wf, bf) = ex.decoder(), fa.decoder()
    scale = max(np.abs(we).max(), np.abs(be).max())
    dev = max(np.abs(wf - we).max(), np.abs(bf - be).max()) / scale
    print(f"  decoder deviation {dev:.3e} of scale {scale:.3e}; E {e_fa:.9e} vs {e_ex:.9e}", flush=True)
    assert dev <= 1e-5
    # And the objective of the two decoders on their own (identical) latents.
    a, b = ex.objective(4, per_site).e, fa.objective(4, per_site).e
    assert abs(a - b) <= 1e-4 * a, (a, b)
