"""Keyframe append on the device (vib_window_append) and the request cycle of the serving loop.

An appended keyframe must give exactly the window a host-side append gives: its state and IMU
edge after the window's, its landmarks after the resident ones (ids increasing), each new
observation after its landmark's resident observations in keyframe order, and observations of
landmarks no longer resident skipped (counted). The resident layout must equal the host upload of
that window (one LM step bit for bit alike). The request cycle (append, LM iterations, OLD
slide) keeps the window at its size with the prior resident, cycle after cycle.
"""
import numpy as np
import pytest

from parity_helpers import same_layout, say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(abi):
    c = abi.Context(0)
    yield c
    c.close()


def host_append(abi, w, kf):
    """The window after appending keyframe kf (the restatement the device edit is checked against);
    returns (window, skipped observations)."""
    K, L = w.K, w.L
    ids = np.concatenate([w.lm_id, kf.lm_id])
    index = {int(u): i for i, u in enumerate(ids)}
    known = np.array([int(u) in index for u in kf.obs_landmark_id], bool)
    o = np.nonzero(known)[0]
    o = o[np.argsort(kf.obs_landmark_id[o], kind="stable")]
    lm = np.concatenate([w.obs_landmark, [index[int(u)] for u in kf.obs_landmark_id[o]]]).astype(np.int32)
    order = np.argsort(lm, kind="stable")
    cat = lambda a, b: np.concatenate([a, b])[order]
    nw = abi.Window(
        frame_id=np.append(w.frame_id, kf.frame_id), pose=np.vstack([w.pose, kf.pose]),
        speed_bias=np.vstack([w.speed_bias, kf.speed_bias]), rig=w.rig, td=w.td, cam_rotation=w.cam_rotation,
        cam_translation=w.cam_translation,
        lm_anchor_frame=np.concatenate([w.lm_anchor_frame, np.full(kf.lm_id.size, K)]),
        lm_anchor_cam=np.concatenate([w.lm_anchor_cam, kf.lm_anchor_cam]),
        lm_bearing=np.vstack([w.lm_bearing, kf.lm_bearing]), lm_velocity=np.vstack([w.lm_velocity, kf.lm_velocity]),
        lm_tref=np.concatenate([w.lm_tref, kf.lm_tref]), lm_rho=np.concatenate([w.lm_rho, kf.lm_rho]),
        lm_free=np.concatenate([w.lm_free if w.lm_free is not None else np.ones(L, np.uint8),
                                np.ones(kf.lm_id.size, np.uint8)]),
        obs_landmark=lm[order], obs_frame=cat(w.obs_frame, np.full(o.size, K)), obs_cam=cat(w.obs_cam, kf.obs_cam[o]),
        obs_bearing=cat(w.obs_bearing, kf.obs_bearing[o]), obs_velocity=cat(w.obs_velocity, kf.obs_velocity[o]),
        obs_tref=cat(w.obs_tref, kf.obs_tref[o]), obs_sqrt_info=cat(w.obs_sqrt_info, kf.obs_sqrt_info[o]),
        imu_from=np.append(w.imu_from, K - 1), imu_to=np.append(w.imu_to, K),
        imu_data=np.vstack([w.imu_data, kf.imu_data[None]]), flags=w.flags, huber_delta=w.huber_delta,
        min_inverse_depth=w.min_inverse_depth, td_lower=w.td_lower, td_upper=w.td_upper,
        pose_free=np.append(w.pose_free if w.pose_free is not None else np.ones(K, np.uint8), 1),
        sb_width=np.append(w.sb_width if w.sb_width is not None else np.full(K, 9, np.uint8), 9),
        prior_key_kind=w.prior_key_kind, prior_key_id=w.prior_key_id, prior_ref=w.prior_ref, prior_A=w.prior_A,
        prior_r0=w.prior_r0, lm_id=ids)
    return nw, int((~known).sum())


FIELDS = ("frame_id", "pose", "speed_bias", "lm_id", "lm_anchor_frame", "lm_anchor_cam", "lm_bearing", "lm_velocity",
          "lm_tref", "lm_rho", "lm_free", "obs_landmark", "obs_frame", "obs_cam", "obs_bearing", "obs_velocity",
          "obs_tref", "obs_sqrt_info", "imu_from", "imu_to", "imu_data", "pose_free", "sb_width", "prior_A", "prior_r0")


def assert_same_window(got, exp):
    for k in FIELDS:
        np.testing.assert_array_equal(getattr(got, k), getattr(exp, k), err_msg=k)


def test_append_equals_the_host_append(ctx, abi):
    from vibpy import synth
    w, kfs = synth.stream(8, 12, 200, n_cams=3, seed=31, rig_free=True, td_free=True)
    say(f"[gpu] append: K={w.K} L={w.L} F={w.F}")
    ctx.upload(0, w)
    cur = ctx.read_window(0)
    for kf in kfs[:3]:
        miss = ctx.append(0, kf)
        exp, skipped = host_append(abi, cur, kf)
        assert miss == skipped == 0
        cur = ctx.read_window(0)
        assert_same_window(cur, exp)
    same_layout(ctx, 0, 1)


def test_request_cycle_append_solve_slide(ctx, abi):
    """Six requests of one session: append the next keyframe, 2 LM iterations, slide OLD. The
    window keeps its 8 frames; observations of eliminated landmarks are skipped as the host
    append predicts; the layout stays the host's."""
    from vibpy import synth
    w, kfs = synth.stream(8, 14, 250, n_cams=4, seed=32, span=12)
    ctx.upload(0, w)
    cur = ctx.read_window(0)
    total = 0
    for i, kf in enumerate(kfs):
        say(f"[gpu] request {i}")
        exp, skipped = host_append(abi, cur, kf)
        miss = ctx.append(0, kf)
        assert miss == skipped
        total += skipped
        assert_same_window(ctx.read_window(0), exp)
        st, rep = ctx.solve(0, abi.options(max_iterations=2))
        assert st == 0, ctx.error()
        st, mrep = ctx.marginalize(0, abi.VIB_MARGIN_OLD)
        assert st == 0, ctx.error()
        cur = ctx.read_window(0)
        assert cur.K == w.K and np.all(np.diff(cur.lm_id) > 0)
        assert cur.prior_rows > 0
    assert total > 0  # tracks reach past the frames their landmarks were eliminated with
    same_layout(ctx, 0, 1)


def test_append_errors(ctx, abi):
    from vibpy import synth
    w, kfs = synth.stream(6, 8, 100, n_cams=2, seed=33)
    ctx.upload(0, w)
    kf = kfs[0]
    bad = abi.Keyframe(**{k: getattr(kf, k) for k in kf.__dataclass_fields__})
    bad.lm_id = kf.lm_id - 10**6  # ids below the window's
    assert ctx.lib.vib_window_append(ctx.h, 0, bad.cstruct(), None) == abi.VIB_ERR_INVALID_ARGUMENT
    bad = abi.Keyframe(**{k: getattr(kf, k) for k in kf.__dataclass_fields__})
    bad.obs_cam = np.full_like(kf.obs_cam, 7)
    assert ctx.lib.vib_window_append(ctx.h, 0, bad.cstruct(), None) == abi.VIB_ERR_INVALID_ARGUMENT
    # the failed appends left the window as it was
    assert ctx.shape(0)[:3] == (w.K, w.L, w.F)
    big, kb = synth.stream(20, 21, 30, n_cams=2, seed=34)
    ctx.upload(1, big)
    assert ctx.lib.vib_window_append(ctx.h, 1, kb[0].cstruct(), None) == abi.VIB_ERR_UNSUPPORTED_DIMENSIONS
    ctx.release(1)
