"""The no-allocation gate (vib_nis_gate_static) against the allocating one (vib_nis_gate).

Two calls in a row, the second on a smaller window, so the second reads scratch the first one
wrote: decisions, statistics and the rejections applied must equal the allocating gate's bit
for bit (the same kernels and summation order). Runs alone in a fresh process and after the
other GPU tests.
"""
import numpy as np
import pytest

from parity_helpers import say

pytestmark = pytest.mark.gpu


def _stats(s):
    return (s.factors, s.components, s.squared_norm, s.over_gate, s.max_nis, s.worst_factor, s.robust_cost,
            s.median_nis, s.p95_nis)


def test_static_gate_equals_allocating_gate(abi):
    from vibpy import synth
    ctx = abi.Context(0)
    try:
        big = synth.generate(K=10, L=2500, n_cams=4, obs_range=(2, 8), seed=95, outlier_fraction=0.05)
        small = synth.generate(K=6, L=400, n_cams=3, obs_range=(2, 6), seed=96, outlier_fraction=0.05)
        for name, w in (("big", big), ("small", small)):
            say(f"[gpu] static gate {name}: F={w.F}")
            ctx.upload(0, w)
            ctx.upload(1, w)
            st0, rep0, dec0 = ctx.nis_gate(0, apply=True)
            st1, rep1, dec1 = ctx.nis_gate_static(1, apply=True)
            assert st0 == 0 and st1 == 0, (st0, st1, ctx.error())
            assert rep0.rejected > 0 and rep1.rejected == rep0.rejected
            np.testing.assert_array_equal(dec1, dec0)
            assert _stats(rep1.evaluated) == _stats(rep0.evaluated)
            assert _stats(rep1.admitted) == _stats(rep0.admitted)
            assert (rep1.skipped, rep1.invalid) == (rep0.skipped, rep0.invalid)
            # the rejections reached the device window alike: a second gate skips the same factors
            st0, rep0, dec0 = ctx.nis_gate(0)
            st1, rep1, dec1 = ctx.nis_gate_static(1)
            np.testing.assert_array_equal(dec1, dec0)
            assert rep1.skipped == rep0.skipped > 0
            w0, w1 = ctx.read_window(0), ctx.read_window(1)
            np.testing.assert_array_equal(w1.obs_sqrt_info, w0.obs_sqrt_info)
    finally:
        ctx.close()
