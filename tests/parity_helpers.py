"""Comparison helpers shared by the GPU parity tests (not collected by pytest)."""
import sys

import numpy as np


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    d = np.linalg.norm(a - b)
    n = max(np.linalg.norm(b), 1e-300)
    return d / n


def prior_columns(kind, ids):
    """Tangent column range of every prior block, keyed by (kind, frame id)."""
    widths = {0: 6, 1: 9, 2: 6, 3: 1}
    out, c = {}, 0
    for k, i in zip(kind, ids):
        key = (int(k), int(i) if k in (0, 1) else 0)
        out[key] = (c, widths[int(k)])
        c += widths[int(k)]
    return out, c


def prior_errors(A_dev, r0_dev, kind_dev, id_dev, A_ref, r0_ref, kind_ref, id_ref):
    """Relative errors of A^T A and A^T r0 of the device prior against the reference prior, on the
    reference's kept subspace (eigenvector signs and rotations inside the cut are not comparable).
    Also checks that both priors carry the same blocks."""
    co, no = prior_columns(kind_ref, id_ref)
    cd, nd = prior_columns(kind_dev, id_dev)
    assert set(co) == set(cd) and no == nd, (sorted(co), sorted(cd))
    perm = np.concatenate([np.arange(cd[k][0], cd[k][0] + cd[k][1]) for k in co])
    Ad = A_dev[:, perm]
    Gd, hd = Ad.T @ Ad, Ad.T @ r0_dev
    Go, ho = A_ref.T @ A_ref, A_ref.T @ r0_ref
    P = np.linalg.pinv(A_ref) @ A_ref
    eg = np.linalg.norm(P @ Gd @ P - Go) / np.linalg.norm(Go)
    eh = np.linalg.norm(P @ hd - ho) / max(np.linalg.norm(ho), 1e-300)
    return eg, eh


def canonical_obs(w):
    order = np.lexsort((w.obs_cam, w.obs_frame, w.obs_landmark))
    return {k: getattr(w, k)[order] for k in ("obs_landmark", "obs_frame", "obs_cam", "obs_bearing", "obs_velocity",
                                               "obs_tref", "obs_sqrt_info")}


def assert_imu_records_close(got, exp, rtol=1e-9):
    """IMU records (VIB_IMU_STRIDE rows) agree: the preintegrated deltas, reference bias and bias
    Jacobians (entries 0..75) elementwise, the whitening W (entries 76..300) through W^T W. The
    eigenvectors behind W carry free signs (the reference's eigensolver fixes none), so only
    W^T W, the inverse covariance the factor uses, is comparable."""
    import numpy as np
    got = np.atleast_2d(got)
    exp = np.atleast_2d(exp)
    np.testing.assert_allclose(got[:, :76], exp[:, :76], rtol=rtol, atol=1e-12)
    for a, b in zip(got, exp):
        Wa, Wb = a[76:].reshape(15, 15), b[76:].reshape(15, 15)
        Ga, Gb = Wa.T @ Wa, Wb.T @ Wb
        assert np.linalg.norm(Ga - Gb) <= 1e-8 * np.linalg.norm(Gb)


def same_layout(ctx, a, b):
    """The window of slot a (edited on the device) uploaded to slot b through the host path steps
    bit for bit alike: normal equations, reduced system, step and costs. Slot b is released."""
    import numpy as np
    w = ctx.read_window(a)
    ctx.upload(b, w)
    lay = w.layout()
    nr, ne = lay["nr"], lay["ne"]
    for s in (a, b):
        ctx.check(ctx.linearize(s), "linearize")
    sa, ra = ctx.solve_step(a, 1e-4)
    sb, rb = ctx.solve_step(b, 1e-4)
    assert sa == 0 and sb == 0, ctx.error()
    Ha, ga, ca = ctx.read_normal(a, nr)
    Hb, gb, cb = ctx.read_normal(b, nr)
    assert np.all(np.isfinite(Ha)) and np.abs(Ha).max() > 0
    assert np.array_equal(Ha, Hb) and np.array_equal(ga, gb) and ca == cb
    Sa, rha = ctx.read_reduced(a, nr)
    Sb, rhb = ctx.read_reduced(b, nr)
    assert np.array_equal(Sa, Sb) and np.array_equal(rha, rhb)
    assert np.array_equal(ctx.read_step(a, nr + ne), ctx.read_step(b, nr + ne))
    assert ra.cost == rb.cost and ra.trial_cost == rb.trial_cost and ra.predicted_reduction == rb.predicted_reduction
    ctx.release(b)
    return w
