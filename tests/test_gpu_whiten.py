"""Device observation whitening (vib_whiten_observations) against the host routine
vib_pinhole_observations and the oracle restatement of FrameStore::observation
(src/backend/FrameStore.cpp:102-124) for undistorted pinhole cameras. Tolerance rtol 1e-12 /
atol 1e-15 (the device contracts into FMA); failure classes exact."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INTR = np.array([[458.6, 457.3, 367.2, 248.4, 752, 480],
                 [700.0, 700.0, 640.0, 360.0, 1280, 720],
                 [320.5, 318.0, 315.0, 242.0, 640, 480],
                 [1000.0, 990.0, 800.0, 600.0, 1600, 1200]])


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx(abi):
    c = abi.Context(0)
    yield c
    c.close()


def _pixels(rng, n, cam):
    w, h = INTR[cam, 4], INTR[cam, 5]
    return np.stack([rng.uniform(0, w), rng.uniform(0, h)], 1)


def test_whitening_matches_host_and_oracle(ctx, abi, oracle):
    rng = np.random.default_rng(140)
    n = 20000
    cam = rng.integers(0, 4, n).astype(np.int32)
    px = _pixels(rng, n, cam)
    sig = rng.uniform(0.5, 2.0, n)
    say(f"[gpu] whiten {n} observations, 4 cameras")
    b, s, st = ctx.whiten(px, INTR, cam, sig)
    assert (st == 0).all()
    for c in range(4):
        for sv in (0.5, 1.0):  # the host routine takes one camera and one sigma per call
            m = cam == c
            hb, hs = abi.pinhole_observations(*INTR[c, :4], sv, px[m])
            db, ds, dst = ctx.whiten(px[m], INTR[c:c + 1], None, sv)
            assert (dst == 0).all()
            np.testing.assert_allclose(db, hb, rtol=1e-12, atol=1e-15)
            np.testing.assert_allclose(ds, hs, rtol=1e-12, atol=1e-15)
    for i in rng.choice(n, 25, replace=False):
        ost, ob, os_ = oracle.pinhole_observation(*INTR[cam[i], :4], px[i, 0], px[i, 1], sig[i])
        assert ost == 0
        np.testing.assert_allclose(b[i], ob, rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(s[i], os_, rtol=1e-12, atol=1e-15)
    # unit bearings, and the whitened covariance is the identity in the sphere basis
    np.testing.assert_allclose(np.linalg.norm(b, axis=1), 1.0, rtol=1e-14)


def test_whitening_failure_classes(ctx, abi):
    rng = np.random.default_rng(141)
    n = 12
    cam = np.zeros(n, np.int32)
    px = _pixels(rng, n, cam)
    sig = np.ones(n)
    px[1] = [-0.5, 10.0]          # left of the image
    px[2] = [INTR[0, 4], 10.0]    # x == width
    px[3] = [10.0, INTR[0, 5]]    # y == height
    px[4] = [np.nan, 10.0]
    sig[5] = 0.0                  # invalid sigma
    sig[6] = np.inf
    cam[7] = 4                    # no such camera
    cam[8] = -1
    b, s, st = ctx.whiten(px, INTR, cam, sig)
    assert list(st) == [0, 1, 1, 1, 1, 4, 4, 4, 4, 0, 0, 0]
    bad = st != 0
    assert not b[bad].any() and not s[bad].any()
    hb, hs = abi.pinhole_observations(*INTR[0, :4], 1.0, px[~bad])
    np.testing.assert_allclose(b[~bad], hb, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(s[~bad], hs, rtol=1e-12, atol=1e-15)
