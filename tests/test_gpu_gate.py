"""Statistical gate and outlier refits on the device (vib_nis_gate, vib_solve_refit) against the
oracle's whitened residuals and against restatements of BackendOptimizer::statisticalGate
(src/backend/BackendOptimizer.cpp:394-472) and of its refit loop (:680-690, 720-737, 1036-1079).
Decisions are exact away from the threshold; statistics agree to rtol 1e-10 (sums in another
order); a rejected factor behaves exactly like a factor uploaded with zero square-root
information (bitwise equal solves)."""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(abi):
    c = abi.Context(0)
    yield c
    c.close()


def _window(seed, L=600):
    from vibpy import synth
    return synth.generate(K=8, L=L, n_cams=4, obs_range=(2, 8), seed=seed, outlier_fraction=0.05)


def _energies(oracle, w):
    """NIS of every factor in landmark order (the device order: stable sort by landmark)."""
    perm = np.argsort(w.obs_landmark, kind="stable")
    e = np.empty(len(perm))
    for i, f in enumerate(perm):
        ok, r, _, _ = oracle.visual_factor(w, int(f))
        assert ok
        e[i] = r @ r
    return perm, e


def _huber(delta, s):
    return np.where(s > delta * delta, delta * (np.sqrt(s) - 0.5 * delta), 0.5 * s)


def _check_stats(st, e, delta):
    n = len(e)
    assert st.factors == n and st.components == 2 * n
    if n == 0:
        assert st.worst_factor == -1
        return
    np.testing.assert_allclose(st.squared_norm, e.sum(), rtol=1e-10)
    np.testing.assert_allclose(st.robust_cost, _huber(delta, e).sum(), rtol=1e-10)
    s = np.sort(e)
    np.testing.assert_allclose(st.median_nis, s[n // 2], rtol=1e-10)
    np.testing.assert_allclose(st.p95_nis, s[math.ceil(0.95 * n) - 1], rtol=1e-10)
    np.testing.assert_allclose(st.max_nis, e.max(), rtol=1e-10)


def test_gate_matches_oracle(ctx, abi, oracle):
    w = _window(130)
    ctx.upload(0, w)
    thr = abi.VISUAL_NIS_THRESHOLD
    say(f"[gpu] gate: F={w.F}")
    st, rep, dec = ctx.nis_gate(0, thr, apply=False)
    assert st == abi.VIB_OK and rep.status == abi.VIB_OK
    perm, e = _energies(oracle, w)
    near = np.abs(e - thr) <= 1e-9 * thr
    want = (e > thr).astype(np.uint8)
    assert (dec[~near] == want[~near]).all()
    assert rep.rejected == 0 and rep.invalid == 0 and rep.skipped == 0
    over = int((dec == 1).sum())
    assert rep.evaluated.over_gate == over and over > 0
    _check_stats(rep.evaluated, e, w.huber_delta)
    _check_stats(rep.admitted, e[dec == 0], w.huber_delta)
    assert rep.admitted.over_gate == 0
    assert e[rep.evaluated.worst_factor] == pytest.approx(e.max(), rel=1e-10)
    say(f"[gpu] gate: {over} over the gate, median {rep.evaluated.median_nis:.3f}, p95 {rep.evaluated.p95_nis:.3f}")


def test_gate_apply_rejects_for_good(ctx, abi):
    w = _window(131)
    ctx.upload(0, w)
    st, rep, dec = ctx.nis_gate(0, apply=True)
    assert st == abi.VIB_OK
    over = int((dec == 1).sum())
    assert rep.rejected == over > 0
    st2, rep2, dec2 = ctx.nis_gate(0, apply=False)  # same state: the rest stay admitted
    assert st2 == abi.VIB_OK
    assert rep2.skipped == over and rep2.evaluated.factors == rep.evaluated.factors - over
    assert rep2.evaluated.over_gate == 0
    assert (dec2[dec == 1] == 3).all() and (dec2[dec != 1] == 0).all()
    assert rep2.evaluated.squared_norm == pytest.approx(rep.admitted.squared_norm, rel=1e-12)
    back = ctx.read_window(0)  # landmark order
    perm = np.argsort(w.obs_landmark, kind="stable")
    assert not back.obs_sqrt_info[dec == 1].any()
    np.testing.assert_array_equal(back.obs_sqrt_info[dec != 1], w.obs_sqrt_info[perm][dec != 1])


def test_rejection_equals_zero_information_upload(ctx, abi):
    w = _window(132)
    ctx.upload(0, w)
    st, rep, dec = ctx.nis_gate(0, apply=True)
    assert st == abi.VIB_OK and rep.rejected > 0
    perm = np.argsort(w.obs_landmark, kind="stable")
    w2 = w.copy()
    w2.obs_sqrt_info[perm[dec == 1]] = 0.0
    ctx.upload(1, w2)
    opts = abi.options(max_iterations=6)
    s0, r0 = ctx.solve(0, opts)
    s1, r1 = ctx.solve(1, opts)
    assert s0 == s1 == abi.VIB_OK
    assert r0.final_cost == r1.final_cost and r0.iterations == r1.iterations
    a, b = w.copy(), w.copy()
    ctx.read_state(0, a)
    ctx.read_state(1, b)
    np.testing.assert_array_equal(a.pose, b.pose)
    np.testing.assert_array_equal(a.lm_rho, b.lm_rho)
    ctx.release(1)


def _refit_restated(ctx, abi, slot, opts, thr, max_refits):
    """BackendOptimizer.cpp:680-690, 720-737, 1036-1079 with the C-ABI pieces."""
    st, pre, _ = ctx.nis_gate(slot, thr, apply=True)
    assert st == abi.VIB_OK
    remaining, used, passes, post_total = max(1, opts.max_iterations), 0, 0, 0
    term = None
    for p in range(max_refits + 1):
        if remaining <= 0:
            term = abi.TERM_ITERATION_BUDGET
            break
        minimum_remaining = max(0, opts.min_iterations - used)
        reserve = min(max_refits - p, max(0, remaining - max(1, minimum_remaining)))
        po = abi.options(max_iterations=remaining - max(0, reserve), min_iterations=opts.min_iterations)
        s, rep = ctx.solve(slot, po)
        assert s == abi.VIB_OK
        passes += 1
        used += rep.iterations
        remaining -= max(1, rep.iterations)
        st, post, _ = ctx.nis_gate(slot, thr, apply=True)
        assert st == abi.VIB_OK
        post_total += post.rejected
        if not post.rejected or remaining <= 0 or p == max_refits:
            term = (abi.TERM_ITERATION_BUDGET if remaining <= 0 else
                    abi.TERM_REFIT_BUDGET if post.rejected and p == max_refits else abi.TERM_SOLVER_FINISHED)
            break
    return dict(pre=pre.rejected, passes=passes, used=used, post=post_total, term=term, final=rep.final_cost)


@pytest.mark.parametrize("max_iterations,max_refits", [(12, 3), (3, 3), (8, 0)])
def test_refit_loop_matches_restatement(ctx, abi, max_iterations, max_refits):
    w = _window(133, L=800)
    thr = abi.VISUAL_NIS_THRESHOLD
    opts = abi.options(max_iterations=max_iterations)
    ctx.upload(0, w)
    st, rep = ctx.solve_refit(0, opts, thr, max_refits)
    assert st == abi.VIB_OK
    ctx.upload(1, w)
    ref = _refit_restated(ctx, abi, 1, opts, thr, max_refits)
    say(f"[gpu] refit {max_iterations}/{max_refits}: passes {rep.passes}, pre {rep.pre_rejected}, "
        f"post {rep.post_rejected}, termination {rep.termination}")
    assert (rep.pre_rejected, rep.passes, rep.solve.iterations, rep.post_rejected, rep.termination) == \
        (ref["pre"], ref["passes"], ref["used"], ref["post"], ref["term"])
    assert rep.solve.final_cost == ref["final"]
    assert rep.solve.iterations <= max_iterations
    a, b = w.copy(), w.copy()
    ctx.read_state(0, a)
    ctx.read_state(1, b)
    np.testing.assert_array_equal(a.pose, b.pose)
    if rep.termination == abi.TERM_SOLVER_FINISHED:
        _, fin, _ = ctx.nis_gate(0, thr, apply=False)
        assert fin.evaluated.over_gate == 0
    ctx.release(1)


def _without(w, drop):
    """The window without the observations `drop` (window order) and without the landmarks left
    with no observation."""
    keep = np.ones(w.F, bool)
    keep[drop] = False
    had = np.bincount(w.obs_landmark, minlength=w.L) > 0
    left = np.bincount(w.obs_landmark[keep], minlength=w.L) > 0
    keep_lm = ~had | left
    new = np.cumsum(keep_lm) - 1
    o = w.copy()
    for name in ("lm_anchor_frame", "lm_anchor_cam", "lm_bearing", "lm_velocity", "lm_tref", "lm_rho", "lm_free"):
        v = getattr(o, name)
        if v is not None:
            setattr(o, name, np.ascontiguousarray(v[keep_lm]))
    for name in ("obs_landmark", "obs_frame", "obs_cam", "obs_bearing", "obs_velocity", "obs_tref", "obs_sqrt_info"):
        setattr(o, name, np.ascontiguousarray(getattr(o, name)[keep]))
    o.obs_landmark = new[o.obs_landmark].astype(np.int32)
    return o


def test_slide_after_rejection_drops_rejected_factors(ctx, abi):
    w = _window(134)
    ctx.upload(0, w)
    opts = abi.options(max_iterations=4)
    st, rep = ctx.solve_refit(0, opts, abi.VISUAL_NIS_THRESHOLD, 1)
    assert st == abi.VIB_OK and rep.pre_rejected + rep.post_rejected > 0
    cur = ctx.read_window(0)  # landmark order, rejected factors with zero information
    rejected = np.flatnonzero(~cur.obs_sqrt_info.any(axis=1))
    assert len(rejected) == rep.pre_rejected + rep.post_rejected
    ctx.upload(1, _without(cur, rejected))
    s0, m0 = ctx.marginalize(0, abi.VIB_MARGIN_OLD)
    s1, m1 = ctx.marginalize(1, abi.VIB_MARGIN_OLD)
    assert s0 == s1 == abi.VIB_OK
    p0, p1 = ctx.read_prior(0), ctx.read_prior(1)
    assert np.isfinite(p0["A"]).all() and np.isfinite(p0["r0"]).all()
    # the square-root factor is free up to a rotation: compare H = A^T A and b = A^T r0
    assert list(p0["kind"]) == list(p1["kind"]) and list(p0["id"]) == list(p1["id"])
    H0, H1 = p0["A"].T @ p0["A"], p1["A"].T @ p1["A"]
    b0, b1 = p0["A"].T @ p0["r0"], p1["A"].T @ p1["r0"]
    np.testing.assert_allclose(H0, H1, rtol=1e-8, atol=1e-8 * np.abs(H1).max())
    np.testing.assert_allclose(b0, b1, rtol=1e-8, atol=1e-8 * np.abs(b1).max())
    a, b = ctx.read_window(0), ctx.read_window(1)
    assert a.F == b.F and a.L == b.L
    assert a.obs_sqrt_info.any(axis=1).all()
    ctx.release(1)
