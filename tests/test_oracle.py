"""Pins of the FP64 CPU oracle (no GPU). Each test checks the oracle against something it does not
restate: central finite differences, closed forms, the reference's numeric factor path, textbook
routines (numpy/scipy), brute-force dense solves and invariants. Sizes are small so the CPU
suite runs in seconds."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def win():
    from vibpy import synth
    return synth.generate(K=5, L=80, n_cams=4, obs_range=(2, 6), seed=101, rig_free=True, td_free=True)


def tangent_perturb(win, block, idx, h, oracle):
    """Copy of win with a left-SE3 / additive perturbation of one tangent coordinate."""
    w = win.copy()
    kind, k = block
    d = np.zeros(6)
    if kind == "pose":
        d[idx] = h
        w.pose[k] = oracle.pose_plus(w.pose[k], d)
    elif kind == "rig":
        d[idx] = h
        w.rig[:] = oracle.pose_plus(w.rig, d)
    elif kind == "td":
        w.td[0] += h
    elif kind == "rho":
        w.lm_rho[k] += h
    elif kind == "sb":
        w.speed_bias[k, idx] += h
    return w


def test_visual_jacobian_central_differences(win, oracle):
    """Analytic tangent Jacobian (ambient J x PlusJacobian) vs central differences on the left-SE3
    tangent (ManagedWindowRegression.cpp:119-167)."""
    checked = 0
    for f in range(0, win.F, max(1, win.F // 25)):
        l = win.obs_landmark[f]
        a, b = win.lm_anchor_frame[l], win.obs_frame[f]
        ok, r, J, _ = oracle.visual_factor(win, f)
        assert ok
        cols = []
        for i in range(6):
            cols.append((("pose", a), i, J[:, i], 1e-6))
            if a != b:
                cols.append((("pose", b), i, J[:, 6 + i], 1e-6))
            cols.append((("rig", 0), i, J[:, 12 + i], 1e-6))
        cols.append((("td", 0), 0, J[:, 18], 1e-6))
        rho = win.lm_rho[l]
        cols.append((("rho", l), 0, J[:, 19], 1e-6 * rho))
        for block, i, col, h in cols:
            wp = tangent_perturb(win, block, i, h, oracle)
            wm = tangent_perturb(win, block, i, -h, oracle)
            okp, rp, _, _ = oracle.visual_factor(wp, f)
            okm, rm, _, _ = oracle.visual_factor(wm, f)
            assert okp and okm
            fd = (rp - rm) / (2 * h)
            scale = max(1.0, np.abs(col).max())
            assert np.abs(fd - col).max() < 2e-5 * scale, (f, block, i, fd, col)
        checked += 1
    assert checked > 10


def test_visual_rho_closed_form(win, oracle):
    """d r / d rho = reduce R (-b_a / rho^2): scaling rho by (1+e) moves r by J_rho rho e."""
    for f in range(0, win.F, 7):
        l = win.obs_landmark[f]
        ok, r, J, _ = oracle.visual_factor(win, f)
        w2 = win.copy()
        e = 1e-7
        w2.lm_rho[l] *= 1 + e
        _, r2, _, _ = oracle.visual_factor(w2, f)
        pred = J[:, 19] * win.lm_rho[l] * e
        assert np.allclose(r2 - r, pred, rtol=1e-4, atol=1e-10 * max(1, np.abs(r).max()))


def test_same_epoch_pose_and_rig_jacobians_are_zero(win, oracle):
    win = win.copy()
    for f in range(0, win.F, 4):  # turn some factors into same-epoch (other camera, same frame)
        l = win.obs_landmark[f]
        win.obs_frame[f] = win.lm_anchor_frame[l]
        win.obs_cam[f] = (win.lm_anchor_cam[l] + 1) % 4
    found = 0
    for f in range(win.F):
        l = win.obs_landmark[f]
        if win.lm_anchor_frame[l] != win.obs_frame[f]:
            continue
        ok, r, J, amb = oracle.visual_factor(win, f)
        assert ok
        assert np.all(J[:, :18] == 0)
        found += 1
    assert found > 0


def test_td_jacobian_zero_for_static_bearings(oracle):
    from vibpy import synth
    w = synth.generate(K=4, L=40, seed=102, td_free=True)
    w.lm_velocity[:] = 0
    w.obs_velocity[:] = 0
    for f in range(0, w.F, 5):
        ok, r, J, _ = oracle.visual_factor(w, f)
        assert ok and J[0, 18] == 0 and J[1, 18] == 0


def test_analytic_residual_matches_numeric_reference_path(win, oracle):
    """Analytic S^2 residual vs the FromTwoVectors / plain-log residual of the reference's numeric
    factor (WindowVisualFactor.cpp:411-430); reference criterion 1e-8 (1 + |r|)."""
    for f in range(win.F):
        ok, r, _, _ = oracle.visual_factor(win, f)
        ok2, r2 = oracle.visual_residual_reference(win, f)
        assert ok and ok2
        assert np.abs(r - r2).max() < 1e-8 * (1 + np.abs(r).max())


def test_visual_residual_zero_at_noiseless_truth(oracle):
    from vibpy import synth
    w = synth.generate(K=4, L=60, seed=103, pixel_noise=0.0, init_pos=0.0, init_rot_deg=0.0, init_rho=0.0)
    for f in range(w.F):
        ok, r, _, _ = oracle.visual_factor(w, f)
        assert ok
        assert np.abs(r).max() < 1e-6


def test_whitening_closed_form(abi, oracle):
    """sqrtInfo^T sqrtInfo = Sigma^-1 with Sigma = sigma^2 (B D)(B D)^T, and the native whitening
    equals the oracle's."""
    rng = np.random.default_rng(5)
    px = np.stack([rng.uniform(0, 720, 50), rng.uniform(0, 540, 50)], 1)
    b, s = abi.pinhole_observations(350, 350, 360, 270, 1.5, px)
    for i in range(len(px)):
        st, bo, so = oracle.pinhole_observation(350, 350, 360, 270, px[i, 0], px[i, 1], 1.5)
        assert st == 0
        assert np.allclose(b[i], bo, atol=1e-15) and np.allclose(s[i], so, rtol=1e-12)
        # closed form with an independent basis/Jacobian construction
        ray = np.array([(px[i, 0] - 360) / 350, (px[i, 1] - 270) / 350, 1.0])
        bb = ray / np.linalg.norm(ray)
        J = np.array([[1 / 350, 0], [0, 1 / 350], [0, 0]])
        D = (np.eye(3) - np.outer(bb, bb)) @ J / np.linalg.norm(ray)
        axis = np.array([1.0, 0, 0]) if abs(bb[0]) < 0.8 else np.array([0, 1.0, 0])
        c0 = axis - bb * bb.dot(axis)
        c0 /= np.linalg.norm(c0)
        B = np.stack([c0, np.cross(bb, c0)])
        Sig = 1.5 ** 2 * (B @ D) @ (B @ D).T
        L = np.array([[s[i, 0], 0], [s[i, 1], s[i, 2]]])
        assert np.allclose(L.T @ L, np.linalg.inv(Sig), rtol=1e-9)


def test_huber_closed_form_cost_and_weight(oracle, abi):
    """Cost 0.5 rho(s) with Huber delta = 1 and sqrt(rho') row scaling, checked through the
    normal equations of a one-factor window against a direct evaluation."""
    from vibpy import synth
    w = synth.generate(K=3, L=10, seed=104, outlier_fraction=0.5)
    st, c = oracle.cost(w)
    assert st == 0
    tot = 0.0
    for f in range(w.F):
        _, r, _, _ = oracle.visual_factor(w, f)
        s2 = r.dot(r)
        tot += 0.5 * s2 if s2 <= 1 else np.sqrt(s2) - 0.5
    # retained part: the oracle's cost minus the visual part must equal IMU + prior 0.5|r|^2
    imu = 0.0
    for e in range(w.E):
        _, r, _ = oracle.imu_factor(w, e)
        imu += 0.5 * r.dot(r)
    prior = 0.5 * np.sum((w.prior_r0) ** 2)  # state at the reference: chart 0
    assert abs(c - (tot + imu + prior)) <= 1e-10 * c


def test_imu_residual_small_on_analytic_motion(oracle, abi):
    """Noise-free IMU, true biases and the true state give a near-zero unwhitened error
    (NativeSlamRegression.cpp:678-688 tolerances: 1e-5 position, 1e-6 rotation, 2e-5 velocity;
    relaxed by the midpoint discretization of our faster trajectory)."""
    from vibpy import synth
    w = synth.generate(K=4, L=20, seed=105, imu_noise=False, init_pos=0, init_rot_deg=0, init_rho=0)
    w.pose[:] = w.meta["true_pose"]
    w.speed_bias[:] = w.meta["true_sb"]
    # re-preintegrate at the true bias so the bias correction terms vanish
    for e in range(w.E):
        t = w.meta["frame_t"]
        st, rec = oracle.preintegrate(w.meta["samples"], t[e], t[e + 1], w.speed_bias[e], w.meta["imu_params"],
                                      w.imu_data[e, 9:12], w.imu_data[e, 12:15])
        assert st == 0
        w.imu_data[e] = rec
    for e in range(w.E):
        ok, r, _ = oracle.imu_factor(w, e)
        W = w.imu_data[e, 76:].reshape(15, 15)
        err = np.linalg.solve(W, r)
        assert np.abs(err[0:3]).max() < 2e-4
        assert np.abs(err[3:6]).max() < 1e-5
        assert np.abs(err[6:9]).max() < 2e-3


def test_imu_jacobian_central_differences(oracle):
    from vibpy import synth
    w = synth.generate(K=3, L=10, seed=106)
    w.speed_bias[:, 3:] += 0.01  # nonzero bias correction
    for e in range(w.E):
        ok, r, J = oracle.imu_factor(w, e)
        s0, s1 = w.imu_from[e], w.imu_to[e]
        blocks = [("pose", s0, 6), ("sb", s0, 9), ("pose", s1, 6), ("sb", s1, 9)]
        for bi, (kind, k, width) in enumerate(blocks):
            for i in range(width):
                h = 1e-6
                wp = tangent_perturb(w, (kind, k), i, h, oracle)
                wm = tangent_perturb(w, (kind, k), i, -h, oracle)
                _, rp, _ = oracle.imu_factor(wp, e)
                _, rm, _ = oracle.imu_factor(wm, e)
                fd = (rp - rm) / (2 * h)
                scale = max(1.0, np.abs(J[bi, :, i]).max())
                assert np.abs(fd - J[bi, :, i]).max() < 1e-4 * scale, (kind, i)


def test_preintegration_native_matches_oracle(oracle, abi):
    """Native host preintegration (vib_preintegrate) equals the oracle restatement; whitening
    compared through W^T W (eigenvector signs are free)."""
    from vibpy import synth
    w = synth.generate(K=4, L=20, seed=107)
    t = w.meta["frame_t"]
    for e in range(w.E):
        st, ref = oracle.preintegrate(w.meta["samples"], t[e], t[e + 1], w.speed_bias[e], w.meta["imu_params"],
                                      w.imu_data[e, 9:12], w.imu_data[e, 12:15])
        assert st == 0
        got = w.imu_data[e]
        assert np.allclose(got[:76], ref[:76], rtol=1e-12, atol=1e-14)
        Wg = got[76:].reshape(15, 15)
        Wr = ref[76:].reshape(15, 15)
        assert np.allclose(Wg.T @ Wg, Wr.T @ Wr, rtol=1e-8)


def test_preintegration_covariance_scaling(oracle, abi):
    """Doubling every noise density scales the covariance by 4 (W^T W by 1/4)."""
    from vibpy import synth
    w = synth.generate(K=3, L=10, seed=108)
    t = w.meta["frame_t"]
    p = w.meta["imu_params"]
    p2 = abi.imu_params(sigma_g_c=2 * p.sigma_g_c, sigma_a_c=2 * p.sigma_a_c, sigma_gw_c=2 * p.sigma_gw_c,
                        sigma_aw_c=2 * p.sigma_aw_c)
    args = (w.meta["samples"], t[0], t[1], w.speed_bias[0])
    _, r1 = oracle.preintegrate(*args, p, w.imu_data[0, 9:12], w.imu_data[0, 12:15])
    _, r2 = oracle.preintegrate(*args, p2, w.imu_data[0, 9:12], w.imu_data[0, 12:15])
    W1 = r1[76:].reshape(15, 15)
    W2 = r2[76:].reshape(15, 15)
    assert np.allclose(W2.T @ W2 * 4, W1.T @ W1, rtol=1e-8)



# ---------------------------------------------------------------------------------------------
# Preintegration pins (ImuError::redoPreintegration, ImuError.cpp:254-462): the bias Jacobians
# against central differences of the deltas re-integrated at b +- h, and the 15x15 covariance
# against closed-form sums of the error-state recursion for motions where it solves by hand.
# ---------------------------------------------------------------------------------------------
def _so3_log(R):
    c = np.clip((np.trace(R) - 1) / 2, -1, 1)
    a = np.arccos(c)
    v = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]]) / 2
    return v if a < 1e-12 else v * a / np.sin(a)


def _skew(a):
    return np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])


def _preint_pair(oracle, abi, samples, t0, t1, sb, params, g0, g1):
    """(oracle record, native vib_preintegrate record) of one interval."""
    st, ref = oracle.preintegrate(samples, t0, t1, sb, params, g0, g1)
    assert st == 0
    return ref, abi.preintegrate(samples, t0, t1, sb, params, g0, g1)


@pytest.mark.parametrize("impl", ["oracle", "native"])
def test_preintegration_bias_jacobians_central_differences(oracle, abi, impl):
    """Record entries 31-75 (d alpha/d b_g, dv/d b_g, dp/d b_g, the C integral = -dv/d b_a and the
    C double integral = -dp/d b_a) equal central differences of the preintegrated deltas with the
    interval re-integrated at the bias +- h. Rotation: Dq(b_g + d) = Exp(-A d) Dq(b_g) (left
    perturbation in the interval's start frame)."""
    from vibpy import synth
    w = synth.generate(K=4, L=20, seed=109)
    t = w.meta["frame_t"]
    smp, prm = w.meta["samples"], w.meta["imu_params"]
    sb0 = w.speed_bias[1].copy()
    sb0[3:6] += [0.02, -0.015, 0.01]   # bias far from zero so the Jacobians are exercised
    sb0[6:9] += [0.05, -0.03, 0.04]
    g0, g1 = w.imu_data[1, 9:12], w.imu_data[1, 12:15]

    def rec(sb):
        pair = _preint_pair(oracle, abi, smp, t[1], t[2], sb, prm, g0, g1)
        return pair[0] if impl == "oracle" else pair[1]

    r0 = rec(sb0)
    R0 = synth.matrix_from_quat(r0[15:19])
    A, Vg, Pg = r0[31:40].reshape(3, 3), r0[40:49].reshape(3, 3), r0[49:58].reshape(3, 3)
    Cint, Cdint = r0[58:67].reshape(3, 3), r0[67:76].reshape(3, 3)
    for i in range(3):
        h = 1e-5
        sp, sm = sb0.copy(), sb0.copy()
        sp[3 + i] += h
        sm[3 + i] -= h
        rp, rm = rec(sp), rec(sm)
        dth = (_so3_log(synth.matrix_from_quat(rp[15:19]) @ R0.T) - _so3_log(synth.matrix_from_quat(rm[15:19]) @ R0.T)) / (2 * h)
        np.testing.assert_allclose(dth, -A[:, i], rtol=1e-6, atol=1e-9 * np.abs(A).max())
        np.testing.assert_allclose((rp[19:22] - rm[19:22]) / (2 * h), Vg[:, i], rtol=1e-6, atol=1e-8 * np.abs(Vg).max())
        np.testing.assert_allclose((rp[22:25] - rm[22:25]) / (2 * h), Pg[:, i], rtol=1e-6, atol=1e-8 * np.abs(Pg).max())
        # accelerometer bias: the deltas are linear in b_a (exact differences)
        h = 1e-3
        sp, sm = sb0.copy(), sb0.copy()
        sp[6 + i] += h
        sm[6 + i] -= h
        rp, rm = rec(sp), rec(sm)
        np.testing.assert_allclose((rp[19:22] - rm[19:22]) / (2 * h), -Cint[:, i], rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose((rp[22:25] - rm[22:25]) / (2 * h), -Cdint[:, i], rtol=1e-9, atol=1e-12)
        np.testing.assert_array_equal(rp[15:19], r0[15:19])  # the rotation does not depend on b_a


def _still_samples(N, T, accel):
    """N + 1 samples over [0, T]: zero angular rate and a constant specific force."""
    smp = np.zeros((N + 1, 7))
    smp[:, 0] = np.arange(N + 1) * (T / N)
    smp[-1, 0] = T
    smp[:, 4:7] = accel
    return smp


def _cov(oracle, abi, smp, T, sig, accel=(0.0, 0.0, 0.0)):
    p = abi.imu_params(sigma_g_c=sig[0], sigma_a_c=sig[1], sigma_gw_c=sig[2], sigma_aw_c=sig[3])
    st, rec, P = oracle.preintegrate_cov(smp, 0.0, T, np.zeros(9), p, np.zeros(3), np.zeros(3))
    assert st == 0
    return rec, P


def _blocks(P):
    return {(a, b): P[3 * a:3 * a + 3, 3 * b:3 * b + 3] for a in range(5) for b in range(5)}


def test_preintegration_covariance_closed_forms_zero_motion(oracle, abi):
    """Zero rate, zero specific force, zero bias, N steps of dt (T = N dt), one noise density at a
    time. The error-state recursion (state [p, alpha, v, b_g, b_a], ImuError.cpp:370-420) reduces
    per axis to scalar sums with these closed forms (S2 = sum_{k<N} k^2 = (N-1)N(2N-1)/6):
      gyro noise:        P_aa = s^2 T
      accel noise:       P_vv = s^2 T, P_pv = s^2 dt^2 N(N-1)/2, P_pp = s^2 dt^3 (S2 + N/2)
                         (the last term is the reference's 0.5 dt^3 position noise)
      gyro bias walk:    P_gg = s^2 T, P_ag = -s^2 dt^2 N(N-1)/2, P_aa = s^2 dt^3 S2
      accel bias walk:   P_bb = s^2 T, P_vb = -s^2 dt^2 N(N-1)/2, P_vv = s^2 dt^3 S2, and the
                         position rows from the explicit solution p_N = sum_j a_j w_j with
                         a_j = -dt^2 [m(m-1)/2 + N(N-1)/2 - j(j+1)/2 - m/2], m = N-1-j
                         (F_pb = -dt C_int + dt^2/2 I double-counts the earlier steps' bias)
    every other block is zero. Also the deltas: Dq = 1, dv = dp = 0, A = T I, C_int = T I,
    C_dint = T^2/2 I."""
    N, T = 20, 0.1
    dt = T / N
    I3 = np.eye(3)
    smp = _still_samples(N, T, (0.0, 0.0, 0.0))
    S2 = (N - 1) * N * (2 * N - 1) / 6
    tri = N * (N - 1) / 2
    s = 0.3
    rec, _ = _cov(oracle, abi, smp, T, (s, 0, 0, 0))
    np.testing.assert_allclose(rec[15:19], [0, 0, 0, 1], atol=1e-15)
    np.testing.assert_allclose(rec[19:25], 0, atol=1e-15)
    np.testing.assert_allclose(rec[31:40].reshape(3, 3), T * I3, rtol=1e-13)
    np.testing.assert_allclose(rec[58:67].reshape(3, 3), T * I3, rtol=1e-13)
    np.testing.assert_allclose(rec[67:76].reshape(3, 3), T * T / 2 * I3, rtol=1e-13)
    np.testing.assert_allclose(rec[40:58], 0, atol=1e-15)
    cases = []
    exp = np.zeros((15, 15))
    exp[3:6, 3:6] = s * s * T * I3
    cases.append(((s, 0, 0, 0), exp))
    exp = np.zeros((15, 15))
    exp[6:9, 6:9] = s * s * T * I3
    exp[0:3, 6:9] = exp[6:9, 0:3] = s * s * dt * dt * tri * I3
    exp[0:3, 0:3] = s * s * dt ** 3 * (S2 + N / 2) * I3
    cases.append(((0, s, 0, 0), exp))
    exp = np.zeros((15, 15))
    exp[9:12, 9:12] = s * s * T * I3
    exp[3:6, 9:12] = exp[9:12, 3:6] = -s * s * dt * dt * tri * I3
    exp[3:6, 3:6] = s * s * dt ** 3 * S2 * I3
    cases.append(((0, 0, s, 0), exp))
    exp = np.zeros((15, 15))
    j = np.arange(N)
    m = N - 1 - j
    a = -dt * dt * (m * (m - 1) / 2 + N * (N - 1) / 2 - j * (j + 1) / 2 - m / 2)
    b = -dt * m
    coef = np.stack([a, b, np.ones(N)])       # rows p, v, b_a
    G = s * s * dt * coef @ coef.T
    idx = [slice(0, 3), slice(6, 9), slice(12, 15)]
    for r in range(3):
        for c in range(3):
            exp[idx[r], idx[c]] = G[r, c] * I3
    np.testing.assert_allclose(G[2, 2], s * s * T, rtol=1e-13)
    np.testing.assert_allclose(G[1, 2], -s * s * dt * dt * tri, rtol=1e-13)
    np.testing.assert_allclose(G[1, 1], s * s * dt ** 3 * S2, rtol=1e-13)
    cases.append(((0, 0, 0, s), exp))
    for sig, exp in cases:
        _, P = _cov(oracle, abi, smp, T, sig)
        np.testing.assert_allclose(P, exp, rtol=1e-12, atol=1e-12 * np.abs(exp).max(), err_msg=str(sig))


def test_preintegration_covariance_closed_forms_constant_force(oracle, abi):
    """Zero rate and a constant specific force a: the rotation error moves the velocity and
    position through -[a]x (F_va = -dt [a]x, F_pa = -dt^2 (k + 1/2) [a]x at step k), and the gyro
    bias through dv/db_g = dt^2 k^2 / 2 [a]x (F_vg = dt^2 (2k+1)/2 [a]x, F_pg = dt^3 (k^2/2 + k/2 +
    1/4) [a]x). Gyro noise and gyro bias walk, one at a time, against the explicit solutions
    x_N = sum_j G_j w_j of the recursion (coefficients summed per step below); the deltas are
    dv = T a, dp = T^2/2 a, dv/db_g = T^2/2 [a]x."""
    N, T = 20, 0.1
    dt = T / N
    acc = np.array([0.7, -1.3, 9.81])
    ax = _skew(acc)
    I3 = np.eye(3)
    smp = _still_samples(N, T, acc)
    s = 0.3
    rec, P = _cov(oracle, abi, smp, T, (s, 0, 0, 0))
    np.testing.assert_allclose(rec[19:22], T * acc, rtol=1e-13)
    np.testing.assert_allclose(rec[22:25], T * T / 2 * acc, rtol=1e-13)
    np.testing.assert_allclose(rec[40:49].reshape(3, 3), T * T / 2 * ax, rtol=1e-12, atol=1e-15)
    # gyro white noise w_j enters alpha after step j
    G = []
    for jj in range(N):
        ca = 1.0
        cv = -dt * sum(1.0 for k in range(jj + 1, N))                      # v_N coefficient of [a]x
        cp = sum(-dt * dt * (k - 1 - jj) for k in range(jj + 2, N)) - sum(dt * dt * (k + 0.5) for k in range(jj + 1, N))
        G.append(np.vstack([cp * ax, ca * I3, cv * ax]))                   # rows p, alpha, v
    G = np.array(G)
    exp = s * s * dt * np.einsum("jab,jcb->ac", G, G)
    np.testing.assert_allclose(P[0:9, 0:9], exp, rtol=1e-11, atol=1e-12 * np.abs(exp).max())
    np.testing.assert_allclose(P[9:, :], 0, atol=1e-18)
    # gyro bias random walk w_j enters b_g after step j
    rec, P = _cov(oracle, abi, smp, T, (0, 0, s, 0))
    G = []
    for jj in range(N):
        def bg(k):
            return 1.0 if k > jj else 0.0

        def al(k):  # alpha_k coefficient
            return -dt * max(k - 1 - jj, 0)

        def vk(k):  # v_k coefficient (of [a]x)
            return sum(dt * dt * max(i - 1 - jj, 0) + 0.5 * dt * dt * (2 * i + 1) * bg(i) for i in range(k))

        cv = vk(N)
        cp = sum(dt * vk(k) + dt * dt * (k + 0.5) * dt * max(k - 1 - jj, 0) + dt ** 3 * (0.5 * k * k + 0.5 * k + 0.25) * bg(k)
                 for k in range(N))
        G.append(np.vstack([cp * ax, al(N) * I3, cv * ax, bg(N) * I3]))      # rows p, alpha, v, b_g
    G = np.array(G)
    exp = s * s * dt * np.einsum("jab,jcb->ac", G, G)
    np.testing.assert_allclose(P[0:12, 0:12], exp, rtol=1e-11, atol=1e-12 * np.abs(exp).max())


def test_preintegration_whitening_inverts_the_covariance(oracle, abi):
    """The record's whitening W (entries 76-300) satisfies W^T W = P^-1 for a full-rank P, in
    both the oracle and the native vib_preintegrate (IMU residual r = W e, ImuError.cpp:436-462)."""
    from vibpy import synth
    w = synth.generate(K=3, L=10, seed=110)
    t = w.meta["frame_t"]
    smp, prm = w.meta["samples"], w.meta["imu_params"]
    st, rec, P = oracle.preintegrate_cov(smp, t[0], t[1], w.speed_bias[0], prm, w.imu_data[0, 9:12], w.imu_data[0, 12:15])
    assert st == 0
    native = abi.preintegrate(smp, t[0], t[1], w.speed_bias[0], prm, w.imu_data[0, 9:12], w.imu_data[0, 12:15])
    Pinv = np.linalg.inv(P)
    for r in (rec, native):
        W = r[76:].reshape(15, 15)
        np.testing.assert_allclose(W.T @ W, Pinv, rtol=1e-7, atol=1e-9 * np.abs(Pinv).max())


def dense_system(win, oracle, lam):
    """Brute force: full dense normal equations over [states, rho] and the LM-damped solve."""
    lin = oracle.linearize(win)
    nr, ne = lin["nr"], lin["ne"]
    n = nr + ne
    H = np.zeros((n, n))
    H[:nr, :nr] = lin["H"]
    H[nr:, :nr] = lin["cross"]
    H[:nr, nr:] = lin["cross"].T
    H[nr:, nr:] = np.diag(lin["diag"])
    g = lin["g"]
    Hd = H + np.diag(lam * np.maximum(np.diag(H), 1e-6))
    return H, g, Hd, np.linalg.solve(Hd, -g)


def test_schur_step_equals_dense_solve(oracle):
    """The Schur-eliminated step equals a dense solve of the full damped system."""
    from vibpy import synth
    w = synth.generate(K=4, L=40, seed=109, rig_free=True, td_free=True)
    for lam in (1e-4, 1e-1, 10.0):
        H, g, Hd, dx = dense_system(w, oracle, lam)
        st, step = oracle.solve_step(w, lam)
        assert st == 0
        assert np.linalg.norm(step["delta"] - dx) <= 1e-8 * np.linalg.norm(dx)
        # model decrease -dx^T (g + 0.5 H dx) with the undamped H
        pred = -dx.dot(g + 0.5 * H.dot(dx))
        assert abs(step["predicted"] - pred) <= 1e-8 * abs(pred)


def test_normal_equations_equal_stacked_jacobian_products(oracle):
    """H = J^T J and g = J^T r from the oracle's per-factor rows (visual with Huber weights, IMU
    and prior), stacked into one dense Jacobian."""
    from vibpy import synth
    w = synth.generate(K=3, L=15, seed=110, outlier_fraction=0.3)
    lay = w.layout()
    n = lay["nr"] + lay["ne"]
    rows, res = [], []
    for f in range(w.F):
        ok, r, J, _ = oracle.visual_factor(w, f)
        l = w.obs_landmark[f]
        a, b = w.lm_anchor_frame[l], w.obs_frame[f]
        s = r.dot(r)
        wt = 1.0 if s <= 1 else np.sqrt(1 / np.sqrt(s))
        row = np.zeros((2, n))
        row[:, lay["pose"][a]:lay["pose"][a] + 6] += J[:, :6]
        if a != b:
            row[:, lay["pose"][b]:lay["pose"][b] + 6] += J[:, 6:12]
        row[:, lay["lm"][l]] += J[:, 19]
        rows.append(wt * row)
        res.append(wt * r)
    for e in range(w.E):
        ok, r, J = oracle.imu_factor(w, e)
        s0, s1 = w.imu_from[e], w.imu_to[e]
        row = np.zeros((15, n))
        row[:, lay["pose"][s0]:lay["pose"][s0] + 6] = J[0, :, :6]
        row[:, lay["sb"][s0]:lay["sb"][s0] + 9] = J[1]
        row[:, lay["pose"][s1]:lay["pose"][s1] + 6] = J[2, :, :6]
        row[:, lay["sb"][s1]:lay["sb"][s1] + 9] = J[3]
        rows.append(row)
        res.append(r)
    # prior at its reference: chart 0, chart derivative = identity (up to the FD error)
    A = w.prior_A
    row = np.zeros((A.shape[0], n))
    row[:, lay["pose"][0]:lay["pose"][0] + 6] = A[:, :6]
    row[:, lay["sb"][0]:lay["sb"][0] + 9] = A[:, 6:15]
    rows.append(row)
    res.append(w.prior_r0)
    Jd = np.concatenate(rows)
    r = np.concatenate(res)
    lin = oracle.linearize(w)
    nr = lin["nr"]
    assert np.allclose(lin["H"], (Jd.T @ Jd)[:nr, :nr], rtol=1e-6, atol=1e-6 * np.abs(lin["H"]).max())
    assert np.allclose(lin["g"], Jd.T @ r, rtol=1e-6, atol=1e-6 * np.abs(lin["g"]).max())
    assert np.allclose(lin["diag"], np.diag(Jd.T @ Jd)[nr:], rtol=1e-9)


def test_jacobi_eigensolver_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    for n in (5, 15, 40):
        X = rng.normal(size=(n, n))
        A = X @ X.T
        st, ev, V = oracle.eigen(A)
        assert st == 0
        assert np.allclose(ev, np.linalg.eigvalsh(A), rtol=1e-10, atol=1e-10 * ev.max())
        assert np.allclose(V @ np.diag(ev) @ V.T, A, atol=1e-9 * np.abs(A).max())


def test_pose_plus_minus_roundtrip(oracle):
    rng = np.random.default_rng(4)
    for _ in range(20):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        x = np.concatenate([rng.normal(size=3), q])
        d = rng.normal(size=6) * 0.3
        y = oracle.pose_plus(x, d)
        back = oracle.pose_minus(y, x)
        assert np.allclose(back, d, atol=1e-12)


def test_lm_reduces_cost_and_recovers_truth(oracle, abi):
    from vibpy import synth
    w = synth.generate(K=6, L=300, seed=111)
    out = w.copy()
    st, rep = oracle.solve(out, abi.options(max_iterations=10))
    assert st == 0
    assert rep.final_cost < 1e-2 * rep.initial_cost
    assert np.abs(out.pose[:, :3] - w.meta["true_pose"][:, :3]).max() < 0.02


def _latest_two_case(oracle, abi, seed=31):
    """A window solved jointly (its optimum), then the two newest poses/velocities perturbed."""
    from vibpy import synth
    w = synth.generate(K=8, L=400, n_cams=4, obs_range=(2, 8), seed=seed)
    st, _ = oracle.solve(w, abi.options(max_iterations=12))
    assert st == 0
    pert = w.copy()
    pert.pose[-2:, :3] += 0.03
    pert.speed_bias[-2:, :3] += 0.05
    return w, pert


def test_latest_two_returns_to_joint_optimum(oracle, abi):
    """Restricted to the newest two poses and velocities, the solve must return them to the joint
    optimum (every other block is held at that optimum); nothing else may change."""
    opt, pert = _latest_two_case(oracle, abi)
    out = pert.copy()
    st, rep, nis, support = oracle.solve_latest_two(out, abi.options(max_iterations=5))
    assert st == 0 and rep.iterations <= 3 and rep.retained_dimension == 18
    assert support >= 6
    assert np.abs(out.pose[-2:, :3] - opt.pose[-2:, :3]).max() < 1e-4
    assert np.abs(out.speed_bias[-2:, :3] - opt.speed_bias[-2:, :3]).max() < 1e-3
    assert np.array_equal(out.pose[:-2], pert.pose[:-2])
    assert np.array_equal(out.speed_bias[:, 3:], pert.speed_bias[:, 3:])
    assert np.array_equal(out.lm_rho, pert.lm_rho)
    assert nis[1] < nis[0] and nis[3] < nis[2]


def test_latest_two_rejections(oracle, abi):
    """Reference rejection reasons: no IMU edge into K-2, too few landmarks seen in K-1, and the
    NIS gate of the new edge."""
    from vibpy import synth
    opt, pert = _latest_two_case(oracle, abi)
    w = pert.copy()
    keep = w.imu_to != w.K - 2
    w.imu_from, w.imu_to, w.imu_data = w.imu_from[keep], w.imu_to[keep], w.imu_data[keep]
    st, _, _, _ = oracle.solve_latest_two(w, abi.options(max_iterations=3))
    assert st == abi.VIB_ERR_IMU_BRIDGE
    w = pert.copy()
    latest = w.obs_frame == w.K - 1
    w.obs_frame = np.where(latest, 0, w.obs_frame).astype(w.obs_frame.dtype)
    st, _, _, support = oracle.solve_latest_two(w, abi.options(max_iterations=3))
    assert st == abi.VIB_ERR_LIGHTWEIGHT_SUPPORT and support == 0
    w = pert.copy()
    w.speed_bias[-1, :3] += 50.0  # a velocity far off makes the new edge's NIS explode
    st, _, nis, _ = oracle.solve_latest_two(w, abi.options(max_iterations=1), gate=True)
    assert st in (abi.VIB_ERR_IMU_GATE, 0)
    if st == abi.VIB_ERR_IMU_GATE:
        assert nis[3] > 60


def test_marginalize_old_is_the_schur_complement(oracle, abi):
    """OLD marginalization: the prior's quadratic (A^T A, A^T r0) equals the Schur complement of
    the eliminated block (victim pose/speed-bias and its landmarks) in the whole system of the
    consumed factors, computed here directly with numpy; the eigen cut drops nothing here."""
    from vibpy import synth
    w = synth.generate(K=6, L=200, n_cams=4, obs_range=(2, 6), seed=41, rig_free=True, td_free=True)
    st, sysm = oracle.marginal_system(w)
    assert st == 0
    st, pr = oracle.marginalize_old(w)
    assert st == 0
    H, b, e = sysm["H"], sysm["b"], sysm["mask"]
    r = ~e
    Hmm, Hmr, Hrr = H[np.ix_(e, e)], H[np.ix_(e, r)], H[np.ix_(r, r)]
    X = np.linalg.solve(Hmm, np.column_stack([Hmr, b[e]]))
    S = Hrr - Hmr.T @ X[:, :-1]
    g = b[r] - Hmr.T @ X[:, -1]
    A, r0 = pr["A"], pr["r0"]
    # eigen cut max(1e-12, 1e-9 lambda_max) (Marginalizer.cpp:441-444): the rank and the kept
    # subspace follow numpy's eigendecomposition of the Schur complement
    lam, U = np.linalg.eigh(0.5 * (S + S.T))
    cut = max(1e-12, 1e-9 * lam[-1])
    assert abs(pr["cutoff"] - cut) <= 1e-6 * cut
    keep = lam > cut
    assert A.shape == (keep.sum(), S.shape[0])
    P = U[:, keep] @ U[:, keep].T
    assert np.linalg.norm(A.T @ A - P @ S @ P) <= 1e-8 * np.linalg.norm(S)
    assert np.linalg.norm(A.T @ r0 - P @ g) <= 1e-8 * np.linalg.norm(g)
    # linear-quadratic agreement: min over the eliminated block of the whole quadratic
    rng = np.random.default_rng(3)
    for _ in range(3):
        dr = rng.normal(size=r.sum()) * 1e-2
        dm = -np.linalg.solve(Hmm, Hmr @ dr + b[e])
        d = np.zeros(len(b)); d[e] = dm; d[r] = dr
        full = 0.5 * d @ H @ d + b @ d
        dr = P @ dr  # kept subspace
        dm = -np.linalg.solve(Hmm, Hmr @ dr + b[e])
        d = np.zeros(len(b)); d[e] = dm; d[r] = dr
        full = 0.5 * d @ H @ d + b @ d
        base = -0.5 * b[e] @ np.linalg.solve(Hmm, b[e])
        prior = 0.5 * np.sum((r0 + A @ dr) ** 2) - 0.5 * np.sum(r0 ** 2) + base
        assert abs(full - prior) <= 1e-7 * max(1.0, abs(full))


# ---------------------------------------------------------------------------------------------
# SECOND_NEW (SlidingWindowManager.cpp:263-482, 587-680; BodyImuFactor.cpp:306-388)
# ---------------------------------------------------------------------------------------------
def _slide_args(w):
    m = w.meta
    ft = m["frame_t"]
    return m["samples"], float(ft[w.K - 3]), float(ft[w.K - 1]), m["imu_params"]


def _qmul(a, b):  # Hamilton product, xyzw storage
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def _edge(w, a, b):
    idx = [e for e in range(w.E) if w.imu_from[e] == a and w.imu_to[e] == b]
    assert len(idx) == 1
    return w.imu_data[idx[0]]


def test_second_new_merged_imu_edge_composes_the_two_edges(oracle, abi):
    """With the bias estimate at the edges' reference the merged edge keeps it (incremental
    mode) and, the middle frame time being a sample time, midpoint integration over the joined
    samples equals the composition of the two edges: dq = dq_a dq_b, dv = dv_a + R_a dv_b,
    dp = dp_a + dv_a dt_b + R_a dp_b."""
    from vibpy import synth
    w = synth.generate(K=6, L=150, n_cams=2, obs_range=(2, 5), seed=51)
    samples, t0, t1, params = _slide_args(w)
    tv = float(w.meta["frame_t"][w.K - 2])
    assert np.min(np.abs(samples[:, 0] - tv)) < 1e-12
    st, rep, nxt = oracle.marginalize_second_new(w, samples, t0, t1, params)
    assert st == 0 and rep["imu_mode"] == abi.VIB_IMU_MERGE_INCREMENTAL
    K = w.K
    da, db = _edge(w, K - 3, K - 2), _edge(w, K - 2, K - 1)
    dm = _edge(nxt, K - 3, K - 2)
    assert nxt.E == w.E - 1
    assert abs(dm[0] - (da[0] + db[0])) < 1e-12
    np.testing.assert_allclose(dm[25:31], da[25:31], atol=0)
    Ra = synth.matrix_from_quat(da[15:19])
    q = _qmul(da[15:19], db[15:19])
    q = q if np.dot(q, dm[15:19]) > 0 else -q
    np.testing.assert_allclose(dm[15:19], q, atol=1e-12)
    np.testing.assert_allclose(dm[19:22], da[19:22] + Ra @ db[19:22], atol=1e-11)
    np.testing.assert_allclose(dm[22:25], da[22:25] + da[19:22] * db[0] + Ra @ db[22:25], atol=1e-11)
    # the other edges move down with their frames, unchanged
    for k in range(K - 3):
        np.testing.assert_array_equal(_edge(nxt, k, k + 1), _edge(w, k, k + 1))


def test_second_new_reintegrates_at_a_moved_bias(oracle, abi):
    """A gyro-bias estimate 1e-3 rad/s away from the reference (> 3e-4) re-integrates the joined
    interval at the current estimate, with the gyroscope at both frame times interpolated from
    the samples (gyroAt)."""
    from vibpy import synth
    w = synth.generate(K=6, L=150, n_cams=2, obs_range=(2, 5), seed=52)
    w.speed_bias[w.K - 3, 3:6] += 1e-3
    samples, t0, t1, params = _slide_args(w)
    st, rep, nxt = oracle.marginalize_second_new(w, samples, t0, t1, params)
    assert st == 0 and rep["imu_mode"] == abi.VIB_IMU_MERGE_REINTEGRATED
    dm = _edge(nxt, w.K - 3, w.K - 2)
    np.testing.assert_allclose(dm[25:31], w.speed_bias[w.K - 3, 3:9], atol=0)
    st, direct = oracle.preintegrate(samples, t0, t1, w.speed_bias[w.K - 3], params,
                                     synth.gyro_at(samples, t0), synth.gyro_at(samples, t1))
    assert st == 0
    np.testing.assert_allclose(dm, direct, rtol=1e-13, atol=1e-15)
    # no samples: the two edges go into the Schur prior with the frame
    st, rep, nxt = oracle.marginalize_second_new(w, None, 0.0, 0.0, None)
    assert st == 0 and rep["imu_mode"] == abi.VIB_IMU_MERGE_BRIDGE and not rep["prior_reused"]
    assert nxt.E == w.E - 2


def test_second_new_reanchors_demotes_and_reuses_the_prior(oracle, abi):
    """Static tracks anchored in the second-newest frame move to their newest surviving view
    (lowest camera first) with rho = 1 / (ray . p_C) of the current point; a track with one
    surviving ray leaves the window; nothing is left on the frame, so the prior is reused as it
    is; the frame's other measurements are discarded."""
    from vibpy import synth
    w0 = synth.generate(K=7, L=400, n_cams=4, obs_range=(2, 8), seed=53, static_bearings=True)
    v = w0.K - 2
    w, moved = synth.reanchor_at(w0, v)
    assert len(moved) > 20
    obs_of = [[] for _ in range(w.L)]
    for f in range(w.F):
        obs_of[w.obs_landmark[f]].append(f)
    anchored = [l for l in range(w.L) if w.lm_anchor_frame[l] == v]
    surv = {l: [f for f in obs_of[l] if w.obs_frame[f] != v] for l in anchored}
    choice = {}
    for l in anchored:
        if len(surv[l]) < 2:
            continue
        Ra, ta = synth.camera_world(w, v, w.lm_anchor_cam[l])
        b = w.lm_bearing[l] / np.linalg.norm(w.lm_bearing[l])
        world = Ra @ (b / w.lm_rho[l]) + ta
        for f in sorted(surv[l], key=lambda f: (-w.obs_frame[f], w.obs_cam[f])):
            Rc, tc = synth.camera_world(w, w.obs_frame[f], w.obs_cam[f])
            d = Rc.T @ (world - tc)
            ray = w.obs_bearing[f]
            r = float(d @ ray)
            if r > 0 and np.arctan2(np.linalg.norm(np.cross(d, ray)), r) <= 0.03:
                choice[l] = (f, r)
                break
    # tracks that would be Schur-retired (no surviving view, or no acceptable one) are removed
    # first, so that nothing is left on the frame
    retired = [l for l in anchored if len(surv[l]) != 1 and l not in choice]
    assert len(choice) > 20
    w2 = synth.drop_landmarks(w, retired)
    st, rep, nxt = oracle.marginalize_second_new(w2, *_slide_args(w2))
    assert st == 0 and rep["prior_reused"]
    np.testing.assert_array_equal(nxt.prior_A, w2.prior_A)
    np.testing.assert_array_equal(nxt.prior_r0, w2.prior_r0)
    np.testing.assert_array_equal(nxt.prior_key_id, w2.prior_key_id)
    demoted = {l for l in anchored if len(surv[l]) == 1}
    keep = [l for l in range(w.L) if l not in demoted and l not in retired]
    removed_f = sum(1 for f in range(w2.F) if w2.obs_frame[f] == v) + sum(len(surv[l]) for l in demoted)
    for l, (f, r) in choice.items():
        i = keep.index(l)
        fr = w.obs_frame[f] - (1 if w.obs_frame[f] > v else 0)
        assert nxt.lm_anchor_frame[i] == fr and nxt.lm_anchor_cam[i] == w.obs_cam[f]
        np.testing.assert_array_equal(nxt.lm_bearing[i], w.obs_bearing[f])
        assert abs(nxt.lm_rho[i] - 1.0 / r) <= 1e-12 / r
        removed_f += 1
    assert rep["reset"] == len(choice) and rep["demoted"] == len(demoted) and rep["retired"] == 0
    assert nxt.K == w.K - 1 and nxt.L == len(keep) and nxt.F == w2.F - removed_f
    assert not np.any(nxt.obs_frame >= nxt.K)


def _check_prior_is_schur(H, b, e, A, r0):
    r = ~e
    Hmm, Hmr, Hrr = H[np.ix_(e, e)], H[np.ix_(e, r)], H[np.ix_(r, r)]
    X = np.linalg.solve(Hmm, np.column_stack([Hmr, b[e]]))
    S = Hrr - Hmr.T @ X[:, :-1]
    g = b[r] - Hmr.T @ X[:, -1]
    lam, U = np.linalg.eigh(0.5 * (S + S.T))
    keep = lam > max(1e-12, 1e-9 * lam[-1])
    assert A.shape == (keep.sum(), S.shape[0])
    P = U[:, keep] @ U[:, keep].T
    assert np.linalg.norm(A.T @ A - P @ S @ P) <= 1e-8 * np.linalg.norm(S)
    assert np.linalg.norm(A.T @ r0 - P @ g) <= 1e-8 * np.linalg.norm(g)


def test_second_new_retires_moving_tracks_into_the_schur_prior(oracle, abi):
    """Tracks anchored in the frame whose bearings move cannot be re-anchored: they are
    eliminated with the frame, and the new prior's quadratic is the Schur complement of that
    block in the system of the consumed factors and the previous prior."""
    from vibpy import synth
    w = synth.generate(K=6, L=300, n_cams=4, obs_range=(2, 7), seed=54, rig_free=True, td_free=True)
    v = w.K - 2
    w, _ = synth.reanchor_at(w, v)
    args = _slide_args(w)
    st, rep, nxt = oracle.marginalize_second_new(w, *args)
    anchored = [l for l in range(w.L) if w.lm_anchor_frame[l] == v]
    assert st == 0 and not rep["prior_reused"] and rep["reset"] == 0
    assert rep["retired"] == len(anchored) > 0
    st, sysm = oracle.second_new_system(w, *args)
    assert st == 0
    _check_prior_is_schur(sysm["H"], sysm["b"], sysm["mask"], nxt.prior_A, nxt.prior_r0)
    assert nxt.L == w.L - len(anchored) and nxt.E == w.E - 1


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_block_step_equals_dense_step(oracle, threads):
    """The timed CPU baseline (block-sparse per-landmark Schur on host threads) takes the dense
    oracle step: same state step within 1e-10 relative, same costs and model decrease within
    1e-11 (sums in another order)."""
    from vibpy import synth
    w = synth.generate(K=8, L=600, n_cams=4, obs_range=(2, 8), seed=111, rig_free=True, td_free=True,
                       outlier_fraction=0.05)
    for lam in (1e-4, 1e-1):
        st, ref = oracle.solve_step(w, lam)
        assert st == 0
        st, delta, m = oracle.solve_step_threaded(w, lam, threads)
        assert st == 0
        assert np.linalg.norm(delta - ref["delta"]) <= 1e-10 * np.linalg.norm(ref["delta"])
        assert abs(m["cost"] - ref["cost"]) <= 1e-12 * ref["cost"]
        assert abs(m["trial_cost"] - ref["trial_cost"]) <= 1e-11 * ref["trial_cost"]
        assert abs(m["predicted"] - ref["predicted"]) <= 1e-9 * abs(ref["predicted"])
        assert abs(m["max_increment"] - ref["max_increment"]) <= 1e-10 * ref["max_increment"]
        assert m["trial_valid"] == ref["trial_valid"]
