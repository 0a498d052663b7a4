"""The work-list autotuner (vib_autotune): 16 shapes measured, the windows' states restored bit
for bit afterwards, and a batched solve on the tuned context agreeing with the same solve on an
untuned one (merged work lists reorder some sums: cost within 1e-10 relative, states 1e-8)."""
import numpy as np
import pytest

from parity_helpers import say

pytestmark = pytest.mark.gpu


def test_autotune_restores_state_and_keeps_results(abi):
    from vibpy import synth
    wins = [synth.generate(K=6 + i, L=400 + 300 * i, n_cams=2 + i % 3, seed=120 + i) for i in range(6)]
    slots = list(range(30, 36))
    opts = abi.options(max_iterations=3)
    a, b = abi.Context(0), abi.Context(0)
    try:
        for s, w in zip(slots, wins):
            a.upload(s, w)
            b.upload(s, w)
        before = []
        for s, w in zip(slots, wins):
            x = w.copy()
            a.read_state(s, x)
            before.append(x)
        say("[gpu] autotune: 6 windows")
        rep = a.autotune(slots, abi.options(max_iterations=1, flags=abi.VIB_SOLVE_CONTINUOUS), 2)
        assert len(rep["candidates"]) == 16 and all(c["ms"] > 0 for c in rep["candidates"])
        assert rep["best"]["ms"] == min(c["ms"] for c in rep["candidates"])
        for s, w, x in zip(slots, wins, before):
            y = w.copy()
            a.read_state(s, y)
            for f in ("pose", "speed_bias", "rig", "td", "lm_rho"):
                np.testing.assert_array_equal(getattr(y, f), getattr(x, f), err_msg=f)
        sa, ra = a.solve_batch(slots, opts)
        sb, rb = b.solve_batch(slots, opts)
        assert sa == 0 and sb == 0, a.error()
        for s, w, p, q in zip(slots, wins, ra, rb):
            assert p.status == q.status == 0
            assert abs(p.final_cost - q.final_cost) <= 1e-10 * q.final_cost
            x, y = w.copy(), w.copy()
            a.read_state(s, x)
            b.read_state(s, y)
            assert np.abs(x.pose - y.pose).max() < 1e-8
    finally:
        a.close()
        b.close()
