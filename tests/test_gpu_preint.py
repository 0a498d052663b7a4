"""Device IMU preintegration (vib_preintegrate_batch, one wavefront per interval) against the
host routine vib_preintegrate and the oracle restatement of ImuError::redoPreintegration
(src/factor/ImuError.cpp:254-462). Records 0..75 (deltas and bias Jacobians) are compared
entrywise; the whitening W through W^T W, since the signs of the eigenvectors are free.
Tolerance: rtol 1e-11 / atol 1e-13 on entries 0..75 (the device contracts a*b+c into FMA, the
host rounds twice), rtol 1e-8 on W^T W (the eigensolver's stopping rule)."""
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


def _range(ts, t0, t1):
    b = int(np.searchsorted(ts, t0, "right")) - 1
    e = int(np.searchsorted(ts, t1, "left")) + 1
    assert ts[b] <= t0 and ts[e - 1] >= t1
    return b, e


def _intervals(w, rng, n_random):
    from vibpy import synth
    s = w.meta["samples"]
    ts = s[:, 0]
    t = w.meta["frame_t"]
    iv = []
    for e in range(w.E):  # every keyframe edge
        iv.append((t[e], t[e + 1], w.speed_bias[e]))
    for e in range(w.E - 1):  # two edges merged into one (the redundant-frame slide)
        iv.append((t[e], t[e + 2], w.speed_bias[e]))
    for _ in range(n_random):  # random sub-intervals at perturbed bias linearization points
        a, b = np.sort(rng.uniform(ts[1], ts[-2], 2))
        if b - a < 1e-3:
            b = a + 1e-3
        sb = w.speed_bias[0] + rng.normal(0, [0.1] * 3 + [1e-3] * 3 + [1e-2] * 3)
        iv.append((a, b, sb))
    t0 = np.array([x[0] for x in iv])
    t1 = np.array([x[1] for x in iv])
    sb = np.array([x[2] for x in iv])
    rg = np.array([_range(ts, a, b) for a, b in zip(t0, t1)], np.int32)
    g0 = np.array([synth.gyro_at(s, a) for a in t0])
    g1 = np.array([synth.gyro_at(s, b) for b in t1])
    return t0, t1, sb, rg[:, 0].copy(), rg[:, 1].copy(), g0, g1


def _check(got, ref):
    np.testing.assert_allclose(got[:76], ref[:76], rtol=1e-11, atol=1e-13)
    Wg = got[76:].reshape(15, 15)
    Wr = ref[76:].reshape(15, 15)
    np.testing.assert_allclose(Wg.T @ Wg, Wr.T @ Wr, rtol=1e-8, atol=1e-8 * np.abs(Wr.T @ Wr).max())


def test_device_preintegration_matches_host(ctx, abi, oracle):
    from vibpy import synth
    rng = np.random.default_rng(121)
    w = synth.generate(K=12, L=40, seed=120)
    s = w.meta["samples"]
    p = w.meta["imu_params"]
    t0, t1, sb, b, e, g0, g1 = _intervals(w, rng, 40)
    say(f"[gpu] preintegrate_batch: {len(t0)} intervals over {len(s)} samples")
    out, st = ctx.preintegrate_batch(s, b, e, t0, t1, sb, p, g0, g1)
    assert (st == abi.VIB_OK).all()
    worst = 0.0
    for i in range(len(t0)):
        ref = abi.preintegrate(s[b[i]:e[i]], t0[i], t1[i], sb[i], p, g0[i], g1[i])
        _check(out[i], ref)
        worst = max(worst, float(np.max(np.abs(out[i, :76] - ref[:76]) / (np.abs(ref[:76]) + 1e-12))))
    say(f"[gpu] worst relative deviation of entries 0..75: {worst:.3e}")
    for i in (0, w.E, len(t0) - 1):  # and the oracle restatement directly
        ost, ref = oracle.preintegrate(s, t0[i], t1[i], sb[i], p, g0[i], g1[i])
        assert ost == 0
        _check(out[i], ref)


def test_device_preintegration_whole_array_ranges(ctx, abi):
    """An interval handed the whole sample array gives the record of its tight range (the
    samples before t0 are skipped, those after t1 never reached)."""
    from vibpy import synth
    w = synth.generate(K=6, L=20, seed=122)
    s = w.meta["samples"]
    p = w.meta["imu_params"]
    t0, t1, sb, b, e, g0, g1 = _intervals(w, np.random.default_rng(123), 4)
    tight, st1 = ctx.preintegrate_batch(s, b, e, t0, t1, sb, p, g0, g1)
    n = len(t0)
    whole, st2 = ctx.preintegrate_batch(s, np.zeros(n, np.int32), np.full(n, len(s), np.int32), t0, t1, sb, p, g0, g1)
    assert (st1 == 0).all() and (st2 == 0).all()
    np.testing.assert_array_equal(tight, whole)


def test_device_preintegration_saturation_and_invalid(ctx, abi):
    """Saturated readings inflate the noise (the factor 100 of the reference); invalid
    intervals are reported per interval with a zero record and do not disturb the others."""
    from vibpy import synth
    w = synth.generate(K=5, L=20, seed=124)
    s = w.meta["samples"].copy()
    p = w.meta["imu_params"]
    t = w.meta["frame_t"]
    ts = s[:, 0]
    j = int(np.searchsorted(ts, 0.5 * (t[1] + t[2])))
    s[j, 4] = 2.0 * p.a_max  # one saturated accelerometer sample inside edge 1
    s[j + 3, 1] = -2.0 * p.g_max  # and one saturated gyro sample
    t0 = np.array([t[0], t[1], t[2], t[1], t[0]])
    t1 = np.array([t[1], t[2], t[2], t[3], ts[-1] + 1.0])  # [2]: empty interval, [4]: beyond the samples
    sb = np.array([w.speed_bias[0]] * 5)
    be = [_range(ts, t0[i], t1[i]) for i in (0, 1, 3)]
    b = np.array([be[0][0], be[1][0], 0, be[2][0], 0], np.int32)
    e = np.array([be[0][1], be[1][1], len(s), be[2][1], len(s)], np.int32)
    g0 = np.array([synth.gyro_at(s, a) for a in t0])
    g1 = np.array([synth.gyro_at(s, a) for a in t1])
    out, st = ctx.preintegrate_batch(s, b, e, t0, t1, sb, p, g0, g1)
    assert list(st) == [0, 0, abi.VIB_ERR_INVALID_ARGUMENT, 0, abi.VIB_ERR_INVALID_ARGUMENT]
    assert not out[2].any() and not out[4].any()
    for i in (0, 1, 3):
        ref = abi.preintegrate(s[b[i]:e[i]], t0[i], t1[i], sb[i], p, g0[i], g1[i])
        _check(out[i], ref)
    # the saturated edge's covariance is larger than that of the clean edge integrated alike
    clean = abi.preintegrate(w.meta["samples"][b[1]:e[1]], t0[1], t1[1], sb[1], p, g0[1], g1[1])
    Ws, Wc = out[1, 76:].reshape(15, 15), clean[76:].reshape(15, 15)
    assert np.trace(np.linalg.inv(Ws.T @ Ws)) > np.trace(np.linalg.inv(Wc.T @ Wc))
