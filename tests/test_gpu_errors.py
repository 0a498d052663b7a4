"""Failure classes at the C-ABI boundary and window limits, on the device.

Each test forces its condition through the inputs (never by faulting a kernel) and checks the
status code the reference's fallback reason maps to (BackendOptimizer.cpp:32-308), against the
oracle where the oracle has the same class:
  invalid residual         an inverse range <= 0 (VisualSphereBatchMath.hpp:88: rho > 0 required)
  not positive definite    a prior row with a 1e200 speed/bias coefficient: A^T A overflows, the damped
                           reduced system has an infinite pivot (potrf info != 0)
  unsupported dimensions   21 frames or 5 cameras (the library's window limits)
  unsupported loss         a negative Huber scale
  out of memory            an upload beyond the context's HBM budget
A failed upload leaves its slot empty. A landmark may have any number of factors (tracks of up to
23 observations solve like the oracle).
"""
import numpy as np
import pytest

from parity_helpers import rel, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(abi):
    c = abi.Context(0)
    yield c
    c.close()


def test_nonpositive_inverse_range_is_an_invalid_residual(ctx, oracle, abi):
    from vibpy import synth
    w = synth.generate(K=6, L=200, seed=82)
    w.lm_rho[5] = -0.5
    ost, _ = oracle.solve_step(w, 1e-4)
    assert ost == abi.VIB_ERR_INVALID_RESIDUAL
    say("[gpu] errors: rho <= 0")
    ctx.upload(0, w)
    st, rep = ctx.solve(0, abi.options(max_iterations=3))
    assert st == abi.VIB_ERR_INVALID_RESIDUAL and rep.status == abi.VIB_ERR_INVALID_RESIDUAL
    ctx.check(ctx.linearize(0), "linearize")
    st, _ = ctx.solve_step(0, 1e-4)
    assert st == abi.VIB_ERR_INVALID_RESIDUAL


def test_overflowing_prior_is_not_positive_definite(ctx, oracle, abi):
    from vibpy import synth
    w = synth.generate(K=6, L=200, seed=83)
    # a speed/bias column: its chart x - x_ref is exactly zero at the reference, so the cost stays
    # finite while A^T A overflows
    w.prior_A[6, 6] = 1e200
    ost, _ = oracle.solve_step(w, 1e-4)
    assert ost == abi.VIB_ERR_NOT_POSITIVE_DEFINITE
    assert np.isfinite(oracle.cost(w)[1])
    say("[gpu] errors: overflowing prior")
    ctx.upload(0, w)
    ctx.check(ctx.linearize(0), "linearize")
    st, _ = ctx.solve_step(0, 1e-4)
    assert st == abi.VIB_ERR_NOT_POSITIVE_DEFINITE
    ctx.upload(0, w)
    st, rep = ctx.solve(0, abi.options(max_iterations=3))
    assert st == abi.VIB_ERR_NOT_POSITIVE_DEFINITE


def test_window_limits_and_loss(ctx, abi):
    from vibpy import synth
    say("[gpu] errors: window limits")
    w = synth.generate(K=21, L=100, seed=5)
    assert ctx.try_upload(1, w) == abi.VIB_ERR_UNSUPPORTED_DIMENSIONS
    w = synth.generate(K=5, L=100, seed=6)
    w5 = w.copy()
    w5.cam_rotation = np.ascontiguousarray(np.concatenate([w.cam_rotation, w.cam_rotation[:1]]))
    w5.cam_translation = np.ascontiguousarray(np.concatenate([w.cam_translation, w.cam_translation[:1]]))
    assert ctx.try_upload(1, w5) == abi.VIB_ERR_UNSUPPORTED_DIMENSIONS
    wl = w.copy()
    wl.huber_delta = -1.0
    assert ctx.try_upload(1, wl) == abi.VIB_ERR_UNSUPPORTED_LOSS
    # a failed upload leaves the slot empty, also when it held a window before
    ctx.upload(1, w)
    assert ctx.try_upload(1, wl) == abi.VIB_ERR_UNSUPPORTED_LOSS
    st, _ = ctx.solve(1, abi.options(max_iterations=1))
    assert st == abi.VIB_ERR_NOT_FOUND


def test_hbm_budget_admission(ctx, abi):
    from vibpy import synth
    say("[gpu] errors: HBM budget")
    for s in (20, 21, 22):
        ctx.release_if_present(s)
    a = synth.generate(K=6, L=300, seed=90)
    b = synth.generate(K=8, L=600, seed=91)
    fa, fb = abi.window_footprint(a), abi.window_footprint(b)
    base = ctx.resident_bytes()
    ctx.set_hbm_budget(base + max(fa, fb) + min(fa, fb) // 2)  # either window fits, not both
    try:
        ctx.upload(20, a)
        assert ctx.resident_bytes() == base + fa
        assert ctx.try_upload(21, b) == abi.VIB_ERR_OUT_OF_MEMORY
        st, _ = ctx.solve(21, abi.options(max_iterations=1))
        assert st == abi.VIB_ERR_NOT_FOUND
        # re-uploading a slot does not count its own previous window against the budget
        ctx.upload(20, a)
        ctx.release(20)
        ctx.upload(21, b)
        st, _ = ctx.solve(21, abi.options(max_iterations=2))
        assert st == 0, ctx.error()
    finally:
        ctx.set_hbm_budget(0)
        for s in (20, 21):
            ctx.release_if_present(s)


def test_long_tracks_match_oracle(ctx, oracle, abi):
    """Tracks of 18-23 observations (no per-landmark factor cap): one step and a solve equal the
    oracle's, at the small-window parity tolerances."""
    from vibpy import synth
    w = synth.generate(K=20, L=300, n_cams=4, obs_range=(30, 60), seed=84, speed=0.3, rig_free=True, td_free=True)
    counts = np.bincount(w.obs_landmark)
    assert counts.max() > 16
    lay = w.layout()
    say(f"[gpu] long tracks: max {counts.max()} factors per landmark, F={w.F}")
    ctx.upload(0, w)
    ctx.check(ctx.linearize(0), "linearize")
    st, rep = ctx.solve_step(0, 1e-4)
    assert st == 0, ctx.error()
    delta = ctx.read_step(0, lay["nr"] + lay["ne"])
    ost, ref = oracle.solve_step(w, 1e-4)
    assert ost == 0
    assert rel(delta, ref["delta"]) < 1e-6
    assert abs(rep.cost - ref["cost"]) <= 1e-10 * ref["cost"]
    assert abs(rep.trial_cost - ref["trial_cost"]) <= 1e-6 * ref["trial_cost"]
    opts = abi.options(max_iterations=4)
    ctx.upload(0, w)
    st, srep = ctx.solve(0, opts)
    assert st == 0, ctx.error()
    got = w.copy()
    ctx.read_state(0, got)
    exp = w.copy()
    ost, orep = oracle.solve(exp, opts)
    assert ost == 0
    assert srep.iterations == orep.iterations and srep.accepted_steps == orep.accepted_steps
    assert abs(srep.final_cost - orep.final_cost) <= 1e-8 * orep.final_cost
    assert np.abs(got.pose - exp.pose).max() < 1e-6


def test_host_retained_terms_equal_the_device_prior(ctx, oracle, abi):
    """Host-evaluated factors at the boundary (solveOnline's retainedH / retainedG): the window
    uploaded without its prior, with the prior's normal equations (H, g at the current state,
    from the oracle) passed as host terms, takes the same step as the window with its prior
    (step 1e-6, model decrease 1e-6); its reported cost lacks exactly the prior's cost. The
    device-resident LM loop refuses such a window."""
    from vibpy import synth
    w = synth.generate(K=7, L=400, n_cams=4, seed=93, rig_free=True, td_free=True)
    w.pose[0, :3] += 0.01  # move frame 0 off the prior's reference: the prior has a gradient
    bare = w.copy()
    bare.prior_A = np.zeros((0, w.prior_A.shape[1]))
    bare.prior_r0 = np.zeros(0)
    full = oracle.linearize(w)
    nop = oracle.linearize(bare)
    nr = full["nr"]
    Hh = full["H"] - nop["H"]
    gh = full["g"][:nr] - nop["g"][:nr]
    prior_cost = full["cost"] - nop["cost"]
    assert prior_cost > 0
    lay = w.layout()
    say("[gpu] host retained terms")
    ctx.upload(0, bare)
    ctx.set_host_retained(0, Hh, gh)
    ctx.check(ctx.linearize(0), "linearize")
    st, rep = ctx.solve_step(0, 1e-4)
    assert st == 0, ctx.error()
    delta = ctx.read_step(0, lay["nr"] + lay["ne"])
    ost, ref = oracle.solve_step(w, 1e-4)
    assert ost == 0
    assert rel(delta, ref["delta"]) < 1e-6
    assert abs(rep.predicted_reduction - ref["predicted"]) <= 1e-6 * abs(ref["predicted"])
    assert abs(rep.cost + prior_cost - ref["cost"]) <= 1e-9 * ref["cost"]
    st, _ = ctx.solve(0, abi.options(max_iterations=2))
    assert st == abi.VIB_ERR_UNSUPPORTED
    ctx.set_host_retained(0, None, None)
    st, _ = ctx.solve(0, abi.options(max_iterations=2))
    assert st == 0, ctx.error()
