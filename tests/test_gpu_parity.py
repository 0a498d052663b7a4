"""HIP path vs the FP64 oracle on identical seeded windows (parity tests proper).

Every call goes through the C-ABI of libvib.so. Tolerances (FP64 on both sides; the device
forms tangent Jacobians directly and sums in a different order, and the prior chart derivative
is a central difference with h = 1e-7 evaluated with the device's own sin/cos/atan2):
  normal equations   ||dH|| / ||H|| <= 1e-8, ||dg|| / ||g|| <= 1e-8, cost rel 1e-10
  reduced system     ||dS|| / ||S|| <= 1e-8
  LM step            ||dx_gpu - dx_cpu|| / ||dx_cpu|| <= 1e-6
  full solve         final cost rel 1e-8, states within 1e-6 (reference onlineSchurMatchesJoint-
                     CpuSolve criteria: cost 1e-7, state 1e-4)
"""
import sys

import numpy as np
import pytest

from parity_helpers import assert_imu_records_close

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(abi):
    c = abi.Context(0)
    yield c
    c.close()


def small_windows():
    from vibpy import synth
    return [
        ("k5_l150", dict(K=5, L=150, n_cams=4, obs_range=(2, 6), seed=11)),
        ("k8_l400_calib", dict(K=8, L=400, n_cams=4, obs_range=(2, 8), seed=12, rig_free=True, td_free=True)),
        ("k10_l600_c2", dict(K=10, L=600, n_cams=2, obs_range=(2, 10), seed=13, outlier_fraction=0.05)),
    ]


def rel(a, b):
    d = np.linalg.norm(a - b)
    n = max(np.linalg.norm(b), 1e-300)
    return d / n


@pytest.mark.parametrize("name,cfg", small_windows())
def test_linearize_matches_oracle(ctx, oracle, abi, name, cfg):
    from vibpy import synth
    w = synth.generate(**cfg)
    say(f"[gpu] linearize {name}: K={w.K} L={w.L} F={w.F}")
    ctx.upload(0, w)
    ctx.check(ctx.linearize(0), "linearize")
    st, rep = ctx.solve_step(0, 1e-4)
    assert st == 0, ctx.error()
    nr, nv = ctx.dims(0)
    H, g, cost = ctx.read_normal(0, nr)
    ref = oracle.linearize(w)
    assert ref["nr"] == nr
    assert rel(H, ref["H"]) < 1e-8
    assert rel(g, ref["g"][:nr]) < 1e-8
    assert abs(cost - ref["cost"]) <= 1e-10 * abs(ref["cost"])


@pytest.mark.parametrize("name,cfg", small_windows())
def test_schur_step_matches_oracle(ctx, oracle, abi, name, cfg):
    from vibpy import synth
    w = synth.generate(**cfg)
    lay = w.layout()
    for lam in (1e-4, 1e-1):
        say(f"[gpu] step {name} lambda={lam}")
        ctx.upload(0, w)
        ctx.check(ctx.linearize(0), "linearize")
        st, rep = ctx.solve_step(0, lam)
        assert st == 0, ctx.error()
        nr = lay["nr"]
        S, rhs = ctx.read_reduced(0, nr)
        delta = ctx.read_step(0, nr + lay["ne"])
        ost, ref = oracle.solve_step(w, lam)
        assert ost == 0
        assert rel(S, ref["S"]) < 1e-8
        assert rel(rhs, ref["rhs"]) < 1e-8
        assert rel(delta, ref["delta"]) < 1e-6
        assert abs(rep.cost - ref["cost"]) <= 1e-10 * ref["cost"]
        assert abs(rep.predicted_reduction - ref["predicted"]) <= 1e-6 * abs(ref["predicted"])
        assert rep.trial_valid == ref["trial_valid"]
        assert abs(rep.trial_cost - ref["trial_cost"]) <= 1e-6 * abs(ref["trial_cost"]) + 1e-9
        assert abs(rep.max_increment - ref["max_increment"]) <= 1e-6 * ref["max_increment"]


@pytest.mark.parametrize("name,cfg", small_windows())
def test_full_solve_matches_oracle(ctx, oracle, abi, name, cfg):
    from vibpy import synth
    w = synth.generate(**cfg)
    opts = abi.options(max_iterations=6)
    say(f"[gpu] solve {name}")
    ctx.upload(0, w)
    st, rep = ctx.solve(0, opts)
    assert st == 0, ctx.error()
    got = w.copy()
    ctx.read_state(0, got)
    exp = w.copy()
    ost, orep = oracle.solve(exp, opts)
    assert ost == 0
    assert rep.iterations == orep.iterations
    assert rep.accepted_steps == orep.accepted_steps
    assert abs(rep.final_cost - orep.final_cost) <= 1e-8 * orep.final_cost
    assert abs(rep.initial_cost - orep.initial_cost) <= 1e-10 * orep.initial_cost
    assert np.abs(got.pose - exp.pose).max() < 1e-6
    assert np.abs(got.speed_bias - exp.speed_bias).max() < 1e-6
    assert np.abs(got.lm_rho - exp.lm_rho).max() < 1e-6 * max(1.0, np.abs(exp.lm_rho).max())


def test_merged_chunks_match_oracle(ctx, oracle, abi, monkeypatch):
    """Landmark chunks and visual factor ranges merged three at a time per workgroup (the
    large-batch launch shape, here forced on a small window with ragged last groups) give the
    same LM step as the oracle."""
    from vibpy import synth
    monkeypatch.setenv("VIB_CHUNK_MERGE", "3")
    monkeypatch.setenv("VIB_LIN_MERGE", "3")
    w = synth.generate(K=8, L=900, n_cams=4, obs_range=(2, 8), seed=31, rig_free=True, td_free=True)
    lay = w.layout()
    say("[gpu] step with merged chunks")
    ctx.upload(0, w)
    ctx.check(ctx.linearize(0), "linearize")
    st, rep = ctx.solve_step(0, 1e-4)
    assert st == 0, ctx.error()
    nr = lay["nr"]
    S, rhs = ctx.read_reduced(0, nr)
    delta = ctx.read_step(0, nr + lay["ne"])
    ost, ref = oracle.solve_step(w, 1e-4)
    assert ost == 0
    assert rel(S, ref["S"]) < 1e-8
    assert rel(rhs, ref["rhs"]) < 1e-8
    assert rel(delta, ref["delta"]) < 1e-6
    assert abs(rep.predicted_reduction - ref["predicted"]) <= 1e-6 * abs(ref["predicted"])
    assert abs(rep.trial_cost - ref["trial_cost"]) <= 1e-6 * abs(ref["trial_cost"]) + 1e-9


def test_batch_alone_equals_mixed(ctx, abi):
    """Each window's result is the same solved alone and inside a batch of uneven windows, within
    1e-10 relative cost and 1e-7 absolute state: the per-workgroup LDS accumulation adds in
    hardware order, so sums differ at the last bit and LM iterations carry that forward."""
    from vibpy import synth
    cfgs = [dict(K=6, L=300, seed=21), dict(K=9, L=700, n_cams=3, seed=22, obs_range=(2, 9)),
            dict(K=5, L=120, n_cams=2, seed=23), dict(K=7, L=500, seed=24, rig_free=True, td_free=True)]
    wins = [synth.generate(**c) for c in cfgs]
    opts = abi.options(max_iterations=4)
    alone = []
    for i, w in enumerate(wins):
        say(f"[gpu] alone {i}")
        ctx.upload(10 + i, w)
        st, rep = ctx.solve(10 + i, opts)
        assert st == 0, ctx.error()
        out = w.copy()
        ctx.read_state(10 + i, out)
        alone.append((rep.final_cost, out, rep))
    for i, w in enumerate(wins):
        ctx.upload(10 + i, w)
    say("[gpu] batch of 4")
    st, reps = ctx.solve_batch([10, 11, 12, 13], opts)
    assert st == 0, ctx.error()
    for i, w in enumerate(wins):
        out = w.copy()
        ctx.read_state(10 + i, out)
        c0, o0, _ = alone[i]
        assert reps[i].iterations == alone[i][2].iterations
        assert reps[i].accepted_steps == alone[i][2].accepted_steps
        assert abs(reps[i].final_cost - c0) <= 1e-10 * c0
        assert np.abs(out.pose - o0.pose).max() < 1e-7
        assert np.abs(out.lm_rho - o0.lm_rho).max() < 1e-7 * max(1.0, np.abs(o0.lm_rho).max())


def test_latest_two_matches_oracle(ctx, oracle, abi):
    """Two-frame lightweight solve on the device vs the oracle restatement of optimizeLatestTwo:
    same iteration/accept counts, poses and velocities of the two newest frames within 1e-6,
    IMU NIS before/after within 1e-6 relative, nothing else changed."""
    from vibpy import synth
    w = synth.generate(K=8, L=400, n_cams=4, obs_range=(2, 8), seed=31)
    st, _ = oracle.solve(w, abi.options(max_iterations=12))
    assert st == 0
    pert = w.copy()
    pert.pose[-2:, :3] += 0.03
    pert.speed_bias[-2:, :3] += 0.05
    opts = abi.options(max_iterations=5)
    say("[gpu] latest two")
    ctx.upload(0, pert)
    st, rep = ctx.solve_latest_two(0, opts)
    assert st == 0, ctx.error()
    got = pert.copy()
    ctx.read_state(0, got)
    exp = pert.copy()
    ost, orep, nis, support = oracle.solve_latest_two(exp, opts)
    assert ost == 0
    assert rep.solve.iterations == orep.iterations
    assert rep.solve.accepted_steps == orep.accepted_steps
    assert rep.solve.retained_dimension == 18
    assert rep.support_landmarks == support
    assert np.abs(got.pose - exp.pose).max() < 1e-6
    assert np.abs(got.speed_bias - exp.speed_bias).max() < 1e-6
    assert np.array_equal(got.lm_rho, pert.lm_rho)
    dev = [rep.old_imu_nis_before, rep.old_imu_nis_after, rep.new_imu_nis_before, rep.new_imu_nis_after]
    for a, b in zip(dev, nis):
        assert abs(a - b) <= 1e-6 * max(1.0, abs(b)), (dev, nis)


def test_latest_two_rejects_missing_bridge(ctx, abi):
    from vibpy import synth
    w = synth.generate(K=6, L=300, seed=32)
    keep = w.imu_to != w.K - 2
    w.imu_from, w.imu_to, w.imu_data = w.imu_from[keep], w.imu_to[keep], w.imu_data[keep]
    say("[gpu] latest two without bridge")
    ctx.upload(0, w)
    st, _ = ctx.solve_latest_two(0, abi.options(max_iterations=3))
    assert st == abi.VIB_ERR_IMU_BRIDGE


def _prior_columns(kind, ids):
    widths = {0: 6, 1: 9, 2: 6, 3: 1}
    out, c = {}, 0
    for k, i in zip(kind, ids):
        key = (int(k), int(i) if k in (0, 1) else 0)
        out[key] = (c, widths[int(k)])
        c += widths[int(k)]
    return out, c


def test_marginalize_old_matches_oracle(ctx, oracle, abi):
    """OLD marginalization on the device (landmark pseudo-inverse Schur, semidefinite MFMA tile
    Cholesky of the reordered system) vs the oracle's eigen-based prior: A^T A and A^T r0 agree
    within 1e-6 relative on the oracle's kept subspace (the metric's prior bound); the window
    slides by one frame and solves afterwards."""
    from vibpy import synth
    w = synth.generate(K=6, L=200, n_cams=4, obs_range=(2, 6), seed=41, rig_free=True, td_free=True)
    st, pr = oracle.marginalize_old(w)
    assert st == 0
    say("[gpu] marginalize OLD")
    ctx.upload(0, w)
    st, rep = ctx.marginalize(0, abi.VIB_MARGIN_OLD)
    assert st == 0, ctx.error()
    dev = ctx.read_prior(0)
    co, no = _prior_columns(pr["kind"], pr["id"])
    cd, nd = _prior_columns(dev["kind"], dev["id"])
    assert set(co) == set(cd) and no == nd
    perm = np.concatenate([np.arange(cd[k][0], cd[k][0] + cd[k][1]) for k in co])
    Ad = dev["A"][:, perm]
    Gd, hd = Ad.T @ Ad, Ad.T @ dev["r0"]
    Ao = pr["A"]
    Go, ho = Ao.T @ Ao, Ao.T @ pr["r0"]
    P = np.linalg.pinv(Ao) @ Ao
    assert np.linalg.norm(P @ Gd @ P - Go) <= 1e-6 * np.linalg.norm(Go)
    assert np.linalg.norm(P @ hd - ho) <= 1e-6 * np.linalg.norm(ho)
    assert rep.eliminated_landmarks == pr["landmarks"]
    K, L, F, E, m, n = ctx.shape(0)
    assert K == w.K - 1 and E == w.E - 1 and L == w.L - pr["landmarks"] and n == no
    say("[gpu] solve after marginalization")
    st, srep = ctx.solve(0, abi.options(max_iterations=3))
    assert st == 0, ctx.error()


def _second_new_case(name):
    from vibpy import synth
    if name == "retire":
        w = synth.generate(K=6, L=300, n_cams=4, obs_range=(2, 7), seed=54, rig_free=True, td_free=True)
        w, _ = synth.reanchor_at(w, w.K - 2)
        return w, True
    if name == "reset":
        w = synth.generate(K=7, L=400, n_cams=4, obs_range=(2, 8), seed=53, static_bearings=True)
        w, _ = synth.reanchor_at(w, w.K - 2)
        v = w.K - 2
        seen = {}
        for f in range(w.F):
            if w.obs_frame[f] != v:
                seen[w.obs_landmark[f]] = seen.get(w.obs_landmark[f], 0) + 1
        # tracks anchored in the frame with no surviving view would be retired into the prior
        drop = [l for l in range(w.L) if w.lm_anchor_frame[l] == v and seen.get(l, 0) == 0]
        return synth.drop_landmarks(w, drop), True
    if name == "reuse":
        return synth.generate(K=8, L=600, n_cams=4, obs_range=(2, 8), seed=41, static_bearings=True), True
    w = synth.generate(K=6, L=250, n_cams=3, obs_range=(2, 6), seed=55)
    return w, False


def _canonical_obs(w):
    order = np.lexsort((w.obs_cam, w.obs_frame, w.obs_landmark))
    return {k: getattr(w, k)[order] for k in ("obs_landmark", "obs_frame", "obs_cam", "obs_bearing", "obs_velocity",
                                               "obs_tref", "obs_sqrt_info")}


@pytest.mark.parametrize("name", ["retire", "reset", "reuse", "bridge"])
def test_second_new_matches_oracle(ctx, oracle, abi, name):
    """SECOND_NEW on the device vs the oracle: the same IMU merge, track re-anchoring, demotion
    and retirement; the window after the slide agrees exactly in topology, within 1e-12 in the
    re-anchored inverse ranges and 1e-9 in the re-integrated IMU record (whitening compared as
    W^T W); the prior is the same
    (reused bit for bit, or A^T A and A^T r0 within 1e-6 on the oracle's kept subspace); both
    windows then solve to the same state."""
    w, with_samples = _second_new_case(name)
    m = w.meta
    ft = m["frame_t"]
    args = (m["samples"], float(ft[w.K - 3]), float(ft[w.K - 1]), m["imu_params"]) if with_samples else \
        (None, 0.0, 0.0, None)
    st, orep, nxt = oracle.marginalize_second_new(w, *args)
    assert st == 0
    say(f"[gpu] SECOND_NEW {name}: K={w.K} L={w.L} F={w.F}")
    ctx.upload(0, w)
    st, rep = ctx.marginalize(0, abi.VIB_MARGIN_SECOND_NEW, *args[:3], params=args[3]) if with_samples else \
        ctx.marginalize(0, abi.VIB_MARGIN_SECOND_NEW)
    assert st == 0, ctx.error()
    assert rep.imu_mode == orep["imu_mode"] and bool(rep.prior_reused) == orep["prior_reused"]
    assert rep.reset_landmarks == orep["reset"] and rep.demoted_landmarks == orep["demoted"]
    assert rep.eliminated_landmarks == orep["retired"] and rep.dropped_factors == orep["dropped"]
    dev = ctx.read_window(0)
    assert (dev.K, dev.L, dev.F, dev.E) == (nxt.K, nxt.L, nxt.F, nxt.E)
    for k in ("lm_anchor_frame", "lm_anchor_cam", "lm_bearing", "lm_velocity", "imu_from", "imu_to", "frame_id"):
        np.testing.assert_array_equal(getattr(dev, k), getattr(nxt, k), err_msg=k)
    np.testing.assert_allclose(dev.lm_rho, nxt.lm_rho, rtol=1e-12)
    assert_imu_records_close(dev.imu_data, nxt.imu_data)
    cd, co = _canonical_obs(dev), _canonical_obs(nxt)
    for k in cd:
        np.testing.assert_array_equal(cd[k], co[k], err_msg=k)
    if orep["prior_reused"]:
        np.testing.assert_array_equal(dev.prior_A, nxt.prior_A)
        np.testing.assert_array_equal(dev.prior_r0, nxt.prior_r0)
    else:
        co_, no = _prior_columns(nxt.prior_key_kind, nxt.prior_key_id)
        cd_, nd = _prior_columns(dev.prior_key_kind, dev.prior_key_id)
        assert set(co_) == set(cd_) and no == nd
        perm = np.concatenate([np.arange(cd_[k][0], cd_[k][0] + cd_[k][1]) for k in co_])
        Ad = dev.prior_A[:, perm]
        Gd, hd = Ad.T @ Ad, Ad.T @ dev.prior_r0
        Ao = nxt.prior_A
        Go, ho = Ao.T @ Ao, Ao.T @ nxt.prior_r0
        P = np.linalg.pinv(Ao) @ Ao
        assert np.linalg.norm(P @ Gd @ P - Go) <= 1e-6 * np.linalg.norm(Go)
        assert np.linalg.norm(P @ hd - ho) <= 1e-6 * np.linalg.norm(ho)
    say(f"[gpu] solve after SECOND_NEW {name}")
    opts = abi.options(max_iterations=3)
    st, srep = ctx.solve(0, opts)
    assert st == 0, ctx.error()
    got = ctx.read_window(0)
    exp = nxt.copy()
    ost, orep2 = oracle.solve(exp, opts)
    assert ost == 0
    assert abs(srep.final_cost - orep2.final_cost) <= 1e-8 * orep2.final_cost
    assert np.abs(got.pose - exp.pose).max() < 1e-6
    assert np.abs(got.lm_rho - exp.lm_rho).max() < 1e-6 * max(1.0, np.abs(exp.lm_rho).max())


def test_step_is_bitwise_deterministic(ctx, abi):
    """The same step from the same state, solved five times: costs and the full tangent step are
    bit-identical (every sum on the device, including the visual normal block accumulated in LDS
    by the workgroups' waves, runs in a fixed order)."""
    from vibpy import synth
    w = synth.generate(K=12, L=4000, n_cams=4, obs_range=(2, 8), seed=66, rig_free=True, td_free=True)
    lay = w.layout()
    n = lay["nr"] + lay["ne"]
    say(f"[gpu] determinism K={w.K} L={w.L} F={w.F}")
    runs = []
    for _ in range(5):
        ctx.upload(0, w)
        ctx.check(ctx.linearize(0, abi.VIB_FP64), "linearize")
        st, rep = ctx.solve_step(0, 1e-4, abi.VIB_FP64, 1)
        assert st == 0, ctx.error()
        runs.append((rep.cost, rep.trial_cost, ctx.read_step(0, n)))
        ctx.release(0)
    for cost, trial, dx in runs[1:]:
        assert cost == runs[0][0] and trial == runs[0][1]
        assert np.array_equal(dx, runs[0][2])


@pytest.mark.parametrize("name,cfg,solve_tol", [
    ("k10_l1500", dict(K=10, L=1500, n_cams=4, obs_range=(2, 8), seed=61, rig_free=True, td_free=True), (1e-9, 1e-7)),
    # depths 0.5-200 m, 30% low-parallax tracks, a near-static stretch. Its second accepted step
    # (iteration 4) moves weak inverse ranges so far that two FP64 solvers part there too: the
    # FP64 device solve and the FP64 oracle differ by 7.7e-6 in one inverse range and 2e-11 in
    # the cost, mixed and FP64 by 1.4e-5 and 3e-9, with every step agreeing to 1e-12 relative.
    ("k12_l3000_ill", dict(K=12, L=3000, n_cams=4, obs_range=(2, 8), seed=62, rig_free=True, td_free=True,
                           ill_conditioned=True), (1e-8, 1e-4)),
])
def test_mixed_step_within_1e6_of_fp64(ctx, abi, name, cfg, solve_tol):
    """Mixed precision (coupling streams as FP32 hi/lo pairs, landmark Schur products from the hi
    parts on FP32 MFMA, two refinement passes against the exact FP64 operator) vs plain FP64 at
    identical lambda from the same linearization point: ||dx_mixed - dx_fp64|| / ||dx_fp64|| <=
    1e-6 over the full tangent step (the metric's bound), the linearization-point cost equal
    (FP64 in both) and the trial cost within 1e-9. Without refinement the step keeps the FP32
    products' error, and the reduced system (FP32 products) stays within 1e-6 of FP64. Then a
    4-iteration solve with refinement reaches the FP64 solve: same counts, poses 1e-8, cost and
    inverse ranges within the case's tolerances."""
    from vibpy import synth
    w = synth.generate(**cfg)
    lay = w.layout()
    nr, n = lay["nr"], lay["nr"] + lay["ne"]
    res = {}
    for prec, refine in ((abi.VIB_FP64, 0), (abi.VIB_MIXED, 2), (abi.VIB_MIXED, 0)):
        say(f"[gpu] step {name} precision={prec} refinement={refine}")
        ctx.upload(0, w)
        ctx.check(ctx.linearize(0, prec), "linearize")
        st, rep = ctx.solve_step(0, 1e-4, prec, refine)
        assert st == 0, ctx.error()
        S, rhs = ctx.read_reduced(0, nr)
        res[(prec, refine)] = (ctx.read_step(0, n), S, rhs, rep)
    d64, S64, b64, r64 = res[(abi.VIB_FP64, 0)]
    d32, S32, b32, r32 = res[(abi.VIB_MIXED, 2)]
    d0 = res[(abi.VIB_MIXED, 0)][0]
    say(f"[gpu] {name}: dx rel {rel(d32, d64):.3e} (unrefined {rel(d0, d64):.3e}), S rel {rel(S32, S64):.3e}, "
        f"rhs rel {rel(b32, b64):.3e}, trial cost rel {abs(r32.trial_cost - r64.trial_cost) / r64.trial_cost:.3e}")
    assert rel(d32, d64) <= 1e-6
    assert rel(S32, S64) <= 1e-6
    assert rel(b32, b64) <= 1e-6
    assert r32.cost == r64.cost
    assert abs(r32.trial_cost - r64.trial_cost) <= 1e-9 * r64.trial_cost
    assert abs(r32.predicted_reduction - r64.predicted_reduction) <= 1e-6 * abs(r64.predicted_reduction)
    # a linearization in one precision cannot be stepped in the other
    st, _ = ctx.solve_step(0, 1e-4, abi.VIB_FP64, 0)
    assert st == abi.VIB_ERR_INVALID_ARGUMENT
    out = {}
    for prec, refine in ((abi.VIB_FP64, 0), (abi.VIB_MIXED, 2)):
        say(f"[gpu] solve {name} precision={prec}")
        ctx.upload(0, w)
        st, srep = ctx.solve(0, abi.options(max_iterations=4, precision=prec, refinement_steps=refine))
        assert st == 0, ctx.error()
        got = w.copy()
        ctx.read_state(0, got)
        out[prec] = (srep, got)
    (s64, g64), (s32, g32) = out[abi.VIB_FP64], out[abi.VIB_MIXED]
    say(f"[gpu] {name}: solve final cost rel {abs(s32.final_cost - s64.final_cost) / s64.final_cost:.3e}, "
        f"pose max {np.abs(g32.pose - g64.pose).max():.3e}, rho max {np.abs(g32.lm_rho - g64.lm_rho).max():.3e}")
    assert s32.iterations == s64.iterations and s32.accepted_steps == s64.accepted_steps
    assert abs(s32.final_cost - s64.final_cost) <= solve_tol[0] * s64.final_cost
    assert np.abs(g32.pose - g64.pose).max() < 1e-8
    assert np.abs(g32.lm_rho - g64.lm_rho).max() < solve_tol[1] * max(1.0, np.abs(g64.lm_rho).max())
