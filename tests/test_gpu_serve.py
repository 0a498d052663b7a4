"""The request server (vib_server_*): continuous batching of sliding-window sessions.

Six sessions of uneven size stream keyframes; requests (append the session's next keyframe, 1-3
LM iterations, slide OLD back to the window size) arrive as Poisson while others run and join
the running launches of their size bucket. With several workers the sessions are spread over
contexts whose turns run concurrently on threads and streams of their own. An HBM budget below the sessions' footprint forces
evictions (window read back to host memory) and re-uploads. Every request's result (status,
final cost, iterations) and the session's window right after it (poses, inverse ranges, shape)
equal the same request sequence of that session served alone through the C-ABI (append,
request begin / iterate, marginalize): cost within 1e-10 relative, states within 1e-7.
"""
import numpy as np
import pytest

from parity_helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(abi):
    c = abi.Context(0)
    yield c
    c.close()


def _streams(synth, reqs, n):
    """Session i: a stream with a keyframe for each of its requests."""
    out = []
    for i in range(n):
        need = sum(1 for r in reqs if r["session"] == i)
        K = 6 + i
        out.append(synth.stream(K, K + need + 1, 80 + 40 * i, n_cams=2 + i % 3, seed=90 + i, obs_range=(2, 7)))
    return out


@pytest.mark.parametrize("precision,workers", [("fp64", 1), ("mixed", 1), ("fp64", 3)])
def test_served_requests_equal_alone(ctx, abi, precision, workers):
    from vibpy import serve, synth
    prec = abi.VIB_FP64 if precision == "fp64" else abi.VIB_MIXED
    refine = 2 if prec == abi.VIB_MIXED else 0
    n = 6
    reqs = serve.poisson_requests(n, 300.0, 0.12, np.random.default_rng(6))
    assert len(reqs) > 20
    streams = _streams(synth, reqs, n)
    wins = [w for w, _ in streams]
    kfs = [k for _, k in streams]
    budget = sum(sorted(abi.window_footprint(w, prec) for w in wins)[-4:])
    srv = serve.Server(ctx, wins, budget, buckets=(64, 128, 320), precision=prec, refinement=refine, workers=workers)
    snaps = {}

    def on_done(res):
        view, slot = srv.session_view(res["session"])
        cams = wins[res["session"]].cam_rotation.shape[0]
        snaps[res["id"]] = (res, view.read_window(slot, cams) if slot >= 0 else None)

    say(f"[gpu] serving {len(reqs)} requests of {n} sessions ({precision}, {workers} worker(s))")
    try:
        done = srv.run([dict(r) for r in reqs], keyframes=kfs, on_done=on_done)
        st = srv.stats()
    finally:
        srv.close()
    assert len(done) == len(reqs) and st["failed"] == 0, [r["status"] for r in done if r["status"]]
    assert st["slides"] == len(reqs) and st["appends"] == len(reqs)
    assert st["evictions"] > 0 or workers > 1  # one worker: the budget holds four of the six windows
    say(f"[gpu] served: {st['uploads']} uploads, {st['evictions']} evictions, {st['launches']} launches; replaying")
    for i in range(n):
        ctx.upload(0, wins[i])
        mine = [r for r in reqs if r["session"] == i]
        for j, r in enumerate(mine):
            ctx.append(0, kfs[i][j])
            ctx.request_begin(0, abi.options(max_iterations=r["iterations"], precision=prec, refinement_steps=refine))
            while True:
                done_, reps = ctx.request_iterate([0], prec, refine)
                if done_[0]:
                    break
            st_, _ = ctx.marginalize(0, abi.VIB_MARGIN_OLD)
            assert st_ == 0, ctx.error()
            res, win = snaps[r["id"]]
            assert res["status"] == reps[0].status == 0
            assert res["iterations_run"] == reps[0].iterations
            assert abs(res["final_cost"] - reps[0].final_cost) <= 1e-10 * abs(reps[0].final_cost)
            alone = ctx.read_window(0)
            assert (alone.K, alone.L, alone.F) == (win.K, win.L, win.F)
            np.testing.assert_array_equal(alone.lm_id, win.lm_id)
            assert np.abs(alone.pose - win.pose).max() < 1e-7
            assert np.abs(alone.lm_rho - win.lm_rho).max() < 1e-7 * max(1.0, np.abs(win.lm_rho).max())


def test_window_over_the_budget_fails_instead_of_waiting(ctx, abi):
    """A request whose window alone exceeds the server's HBM budget finishes with
    VIB_ERR_OUT_OF_MEMORY at its first turn; windows that other slots of the context hold do not
    count against the server's budget."""
    from vibpy import serve, synth
    small, kf_small = synth.stream(6, 8, 80, n_cams=2, seed=301, obs_range=(2, 6))
    large, kf_large = synth.stream(9, 11, 400, n_cams=3, seed=302, obs_range=(2, 6))
    fs = abi.window_footprint(small, abi.VIB_FP64)
    fl = abi.window_footprint(large, abi.VIB_FP64)
    assert fl > 2 * fs
    ctx.upload(0, large)  # a slot outside the server, larger than the budget
    srv = serve.Server(ctx, [small, large], hbm_budget_bytes=fl - 1, workers=2)
    try:
        reqs = [dict(id=1, session=0, iterations=1, arrival=0.0), dict(id=2, session=1, iterations=1, arrival=0.0)]
        done = {r["id"]: r for r in srv.run(reqs, keyframes=[kf_small, kf_large])}
    finally:
        srv.close()
    assert done[1]["status"] == abi.VIB_OK
    assert done[2]["status"] == abi.VIB_ERR_OUT_OF_MEMORY
