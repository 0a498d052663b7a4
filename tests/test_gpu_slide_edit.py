"""Window edits on the device (slides, stripping of gate rejections, the two-frame window)
against the host layout.

A slide builds the next window on the device: stream compaction of the landmark and observation
streams and a stable (anchor frame, observer frame) sort of the pair-order planes. The same
window read back (vib_read_window) and uploaded through the host path (vib_window_upload, which
sorts on the host) must be the identical device layout, so the two slots take one LM step bit for
bit alike (normal equations, reduced system, step, costs). Cases: OLD after a step on the 2000-landmark window,
SECOND_NEW with the IMU merge on the 20000-landmark rig+td window, re-anchored and retired
tracks, gate rejections stripped before a slide, and three slides in a row on one slot (the
spare arena changes places each time). The topology and prior themselves are compared with the
oracle in test_gpu_parity and test_gpu_fullsize.
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


def _second_new_args(w):
    m = w.meta
    ft = m["frame_t"]
    return m["samples"], float(ft[w.K - 3]), float(ft[w.K - 1]), m["imu_params"]


def test_old_after_a_step_keeps_the_host_layout(ctx, abi):
    from vibpy import synth
    w = synth.generate(**synth.CONFIGS["c1"])
    say(f"[gpu] edit: OLD K={w.K} L={w.L} F={w.F}")
    ctx.upload(0, w)
    st, _ = ctx.solve(0, abi.options(max_iterations=1))
    assert st == 0, ctx.error()
    st, rep = ctx.marginalize(0, abi.VIB_MARGIN_OLD)
    assert st == 0, ctx.error()
    nxt = same_layout(ctx, 0, 1)
    assert nxt.K == w.K - 1 and rep.eliminated_landmarks > 0 and nxt.L == w.L - rep.eliminated_landmarks


def test_second_new_with_merge_keeps_the_host_layout(ctx, abi):
    from vibpy import synth
    w = synth.generate(**synth.CONFIGS["c2"])
    say(f"[gpu] edit: SECOND_NEW K={w.K} L={w.L} F={w.F}")
    ctx.upload(0, w)
    st, _ = ctx.solve(0, abi.options(max_iterations=2))
    assert st == 0, ctx.error()
    smp, t0, t1, prm = _second_new_args(w)
    st, rep = ctx.marginalize(0, abi.VIB_MARGIN_SECOND_NEW, smp, t0, t1, params=prm)
    assert st == 0, ctx.error()
    assert rep.imu_mode in (abi.VIB_IMU_MERGE_INCREMENTAL, abi.VIB_IMU_MERGE_REINTEGRATED)
    nxt = same_layout(ctx, 0, 1)
    assert nxt.K == w.K - 1 and nxt.E == w.E - 1


@pytest.mark.parametrize("seed", [53, 54])
def test_reanchored_tracks_keep_the_host_layout(ctx, abi, seed):
    from vibpy import synth
    w = synth.generate(K=7, L=400, n_cams=4, obs_range=(2, 8), seed=seed, static_bearings=True, rig_free=seed == 54,
                       td_free=seed == 54)
    w, _ = synth.reanchor_at(w, w.K - 2)
    say(f"[gpu] edit: re-anchor seed {seed}")
    ctx.upload(0, w)
    smp, t0, t1, prm = _second_new_args(w)
    st, rep = ctx.marginalize(0, abi.VIB_MARGIN_SECOND_NEW, smp, t0, t1, params=prm)
    assert st == 0, ctx.error()
    assert rep.reset_landmarks + rep.eliminated_landmarks + rep.demoted_landmarks > 0
    same_layout(ctx, 0, 1)


def test_gate_rejections_are_stripped_on_the_device(ctx, abi):
    from vibpy import synth
    w = synth.generate(K=10, L=600, n_cams=2, obs_range=(2, 8), seed=71, outlier_fraction=0.05)
    say("[gpu] edit: strip rejected")
    ctx.upload(0, w)
    st, _ = ctx.solve(0, abi.options(max_iterations=3))
    assert st == 0, ctx.error()
    st, grep, dec = ctx.nis_gate(0, apply=True)
    assert st == 0 and grep.rejected > 0
    K, L, F = ctx.shape(0)[:3]
    st, rep = ctx.marginalize(0, abi.VIB_MARGIN_OLD)
    assert st == 0, ctx.error()
    nxt = same_layout(ctx, 0, 1)
    # no rejected observation survives the slide
    s = nxt.obs_sqrt_info.reshape(-1, 3)
    assert not np.any(np.all(s == 0, axis=1))
    assert nxt.F < F - grep.rejected + 1


def test_three_slides_in_a_row(ctx, abi):
    from vibpy import synth
    w = synth.generate(K=12, L=1500, n_cams=4, obs_range=(2, 8), seed=77, rig_free=True, td_free=True)
    say("[gpu] edit: three slides")
    ctx.upload(0, w)
    for i, pol in enumerate((abi.VIB_MARGIN_OLD, abi.VIB_MARGIN_SECOND_NEW, abi.VIB_MARGIN_OLD)):
        st, _ = ctx.solve(0, abi.options(max_iterations=1))
        assert st == 0, ctx.error()
        st, rep = ctx.marginalize(0, pol)
        assert st == 0, ctx.error()
        nxt = same_layout(ctx, 0, 1)
        assert nxt.K == w.K - 1 - i


def test_batched_slides_equal_single_slides(ctx, abi):
    """vib_marginalize_batch of five uneven windows (one marginal-system batch, one eigen launch
    per stage) leaves each window exactly as vib_marginalize alone does: report, topology, prior
    and state bit for bit."""
    from vibpy import synth
    wins = [synth.generate(K=6 + 2 * i, L=300 + 200 * i, n_cams=2 + i % 3, seed=60 + i, rig_free=i % 2 == 1,
                           td_free=i % 2 == 1) for i in range(5)]
    say("[gpu] edit: batched slides")
    for i, w in enumerate(wins):
        ctx.upload(10 + i, w)
        ctx.upload(20 + i, w)
        ctx.check(ctx.solve(10 + i, abi.options(max_iterations=1))[0], "solve")
        ctx.check(ctx.solve(20 + i, abi.options(max_iterations=1))[0], "solve")
    st, reps = ctx.marginalize_batch([10 + i for i in range(5)], abi.VIB_MARGIN_OLD)
    assert st == 0, ctx.error()
    for i, w in enumerate(wins):
        st1, r1 = ctx.marginalize(20 + i, abi.VIB_MARGIN_OLD)
        assert st1 == 0 and reps[i].status == 0
        for f in ("eliminated_landmarks", "prior_rank", "retained_dimension", "consumed_factors", "eigen_cutoff"):
            assert getattr(reps[i], f) == getattr(r1, f), f
        a, b = ctx.read_window(10 + i), ctx.read_window(20 + i)
        for f in ("lm_id", "lm_anchor_frame", "obs_landmark", "obs_frame", "prior_A", "prior_r0", "prior_ref",
                  "pose", "speed_bias", "lm_rho", "imu_data"):
            np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f)
        ctx.release(10 + i)
        ctx.release(20 + i)
