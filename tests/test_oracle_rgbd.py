"""The RGBD odometry oracle (oracle/rgbd.cpp) pinned without the reference
(whose path is not functional): the twist exponential against scipy's
matrix exponential, the identity correspondence and zero right-hand side for
a frame paired with itself, symmetric positive semi-definite normal
equations, and recovery of the rendered motion between orbit frames."""
import numpy as np
import pytest
from scipy.linalg import expm

import rgbd_inputs as ri


@pytest.mark.parametrize("scale", [1e-7, 1e-5, 1e-3, 0.3, 1.5])
def test_twist_exp_matches_matrix_exponential(orc, scale):
    rng = np.random.default_rng(int(scale * 1e7) % 1000)
    for _ in range(20):
        ksi = rng.normal(size=6) * np.array([scale] * 3 + [1.0] * 3)
        T = np.zeros((4, 4))
        T[:3, :3] = [[0, -ksi[2], ksi[1]], [ksi[2], 0, -ksi[0]], [-ksi[1], ksi[0], 0]]
        T[:3, 3] = ksi[3:]
        G = expm(T)
        R, t = orc.rgbd_twist_exp(ksi)
        assert np.max(np.abs(R - G[:3, :3])) < 1e-13
        assert np.max(np.abs(t - G[:3, 3])) < 1e-12


@pytest.fixture(scope="module")
def small(kfx):
    intr = kfx.kinect_intrinsics(320, 240, 262.5)
    frames, poses = ri.rgbd_frames(kfx, intr, (10, 12))
    return intr, frames, poses


def test_frame_paired_with_itself(orc, small):
    intr, frames, _ = small
    lv = orc.rgbd_prepare(frames[0][0].view(np.uint8).reshape(intr.rows, intr.cols, 4), frames[0][1], 3)
    I, D, dx, dy = lv[0]
    corr, n = orc.rgbd_corr(lv[0], lv[0], intr, np.eye(3), np.zeros(3), 10.0, 0.07)
    textured = (D > 0) & (dx * dx + dy * dy >= 100.0)
    assert n == int(textured.sum()) > 1000
    idx = np.arange(D.size).reshape(D.shape)
    assert np.array_equal(corr[textured], idx[textured]) and np.all(corr[~textured] == -1)
    s, sigma, cnt = orc.rgbd_sums(lv[0], lv[0], intr, corr, np.eye(3), np.zeros(3))
    assert cnt == n and sigma == 0.0 and np.all(s[21:] == 0.0)
    A = np.zeros((6, 6))
    k = 0
    for r in range(6):
        for c in range(r, 6):
            A[r, c] = A[c, r] = s[k]
            k += 1
    ev = np.linalg.eigvalsh(A)
    assert ev.min() > -1e-9 * ev.max() and ev.max() > 0


def test_oracle_recovers_orbit_motion(orc, small):
    intr, frames, poses = small
    (c0, d0), (c1, d1) = frames
    ok, R, t, solves = orc.rgbd_estimate(c0.view(np.uint8), d0, c1.view(np.uint8), d1, intr, 3, (7, 7, 10), 10.0,
                                          0.07, np.eye(3), np.zeros(3))
    Rt, tt = ri.true_motion(poses, 10, 12)
    assert ok == 1 and solves > 10
    # The motion is about 6 cm and 2 degrees; the estimate must remove most
    # of it (starting from the identity).
    assert np.linalg.norm(tt) > 0.03
    assert np.linalg.norm(t - tt) < 0.1 * np.linalg.norm(tt)
    assert ri.rot_angle(R @ Rt.T) < 0.1 * ri.rot_angle(Rt)
