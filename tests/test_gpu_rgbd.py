"""GPU parity of the RGBD photometric odometry (kfx_rgbd.hip) against the
oracle (oracle/rgbd.cpp): the prepared intensity, depth and Sobel pyramids
bit for bit; the first iteration's correspondence count and normal
equations (1e-9 relative); the whole coarse-to-fine estimate against the
oracle's and against the rendered motion; argument checks."""
import ctypes as C

import numpy as np
import pytest

import rgbd_inputs as ri

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pair(gpu):
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    frames, poses = ri.rgbd_frames(kfx, intr, (10, 12), noise=True)
    return intr, frames, poses


def _handle(kfx, intr, frames, levels=4):
    h = C.c_void_p()
    kfx.call("kfx_rgbd_create", intr.cols, intr.rows, levels, C.byref(h))
    bufs = []
    for slot, (c, d) in enumerate(frames):
        bc, bd = kfx.Buffer.from_array(c), kfx.Buffer.from_array(d)
        kfx.call("kfx_rgbd_set_frame", h, slot, bc.ptr, bd.ptr)
        bufs += [bc, bd]
    return h, bufs


def test_prepared_levels_bit_exact(gpu, orc, pair):
    kfx = gpu
    intr, frames, _ = pair
    h, bufs = _handle(kfx, intr, frames)
    for slot, (c, d) in enumerate(frames):
        want = orc.rgbd_prepare(c.view(np.uint8), d, 4)
        for l in range(4):
            got = [np.zeros_like(want[l][0]) for _ in range(4)]
            kfx.call("kfx_rgbd_download_level", h, slot, l, *[g.ctypes.data_as(C.c_void_p) for g in got])
            for g, w in zip(got, want[l]):
                assert np.array_equal(g.view(np.uint32), w.view(np.uint32)), (slot, l)
    kfx.call("kfx_rgbd_destroy", h)


def test_first_iteration_normal_equations(gpu, orc, pair):
    kfx = gpu
    intr, frames, _ = pair
    h, bufs = _handle(kfx, intr, frames)
    prm = kfx.rgbd_params(levels=1, iterations=(1,))
    init = kfx.pose_from(np.eye(3), np.zeros(3))
    res = kfx.RgbdResult()
    kfx.call("kfx_rgbd_estimate", h, C.byref(intr), C.byref(prm), C.byref(init), C.byref(res))
    lp = orc.rgbd_prepare(frames[0][0].view(np.uint8), frames[0][1], 1)[0]
    lc = orc.rgbd_prepare(frames[1][0].view(np.uint8), frames[1][1], 1)[0]
    corr, n = orc.rgbd_corr(lp, lc, intr, np.eye(3), np.zeros(3), 10.0, 0.07)
    s, sigma, cnt = orc.rgbd_sums(lp, lc, intr, corr, np.eye(3), np.zeros(3))
    assert res.correspondences == n > 5000 and res.solves == 1 and res.ok == 1
    assert abs(res.sigma - sigma) <= 1e-12 * sigma
    A = np.array(res.A[:]).reshape(6, 6)
    k = 0
    for r in range(6):
        for c in range(r, 6):
            assert abs(A[r, c] - s[k]) <= 1e-9 * np.abs(s[:21]).max(), (r, c)
            k += 1
    np.testing.assert_allclose(np.array(res.b[:]), s[21:], rtol=1e-9, atol=1e-9 * np.abs(s[21:]).max())
    kfx.call("kfx_rgbd_destroy", h)


def test_estimate_matches_oracle_and_motion(gpu, orc, pair):
    kfx = gpu
    intr, frames, poses = pair
    h, bufs = _handle(kfx, intr, frames)
    prm = kfx.rgbd_params()
    init = kfx.pose_from(np.eye(3), np.zeros(3))
    res = kfx.RgbdResult()
    kfx.call("kfx_rgbd_estimate", h, C.byref(intr), C.byref(prm), C.byref(init), C.byref(res))
    R, t = np.array(res.pose.R[:]).reshape(3, 3), np.array(res.pose.t[:])
    (c0, d0), (c1, d1) = frames
    ok, Ro, to, solves = orc.rgbd_estimate(c0.view(np.uint8), d0, c1.view(np.uint8), d1, intr, 4, (7, 7, 7, 10),
                                            10.0, 0.07, np.eye(3), np.zeros(3))
    assert res.ok == ok == 1 and res.solves == solves == 31
    assert np.linalg.norm(t - to) < 1e-5 and ri.rot_angle(R @ Ro.T) < 1e-5
    Rt, tt = ri.true_motion(poses, 10, 12)
    assert np.linalg.norm(t - tt) < 0.01 and ri.rot_angle(R @ Rt.T) < 0.005
    kfx.call("kfx_rgbd_destroy", h)


def test_rgbd_rejects_bad_arguments(gpu, pair):
    kfx = gpu
    intr, frames, _ = pair
    h = C.c_void_p()
    assert kfx.lib.kfx_rgbd_create(640, 480, 5, C.byref(h)) != 0
    assert kfx.lib.kfx_rgbd_create(60, 40, 4, C.byref(h)) != 0
    kfx.call("kfx_rgbd_create", 640, 480, 2, C.byref(h))
    prm = kfx.rgbd_params(levels=2, iterations=(3, 3))
    init = kfx.pose_from(np.eye(3), np.zeros(3))
    res = kfx.RgbdResult()
    assert kfx.lib.kfx_rgbd_estimate(h, C.byref(intr), C.byref(prm), C.byref(init), C.byref(res)) != 0  # no frames
    c, d = (kfx.Buffer.from_array(a) for a in frames[0])
    assert kfx.lib.kfx_rgbd_set_frame(h, 2, c.ptr, d.ptr) != 0
    kfx.call("kfx_rgbd_set_frame", h, 0, c.ptr, d.ptr)
    kfx.call("kfx_rgbd_swap", h)
    kfx.call("kfx_rgbd_set_frame", h, 0, c.ptr, d.ptr)
    bad = kfx.rgbd_params(levels=3, iterations=(3, 3, 3))
    assert kfx.lib.kfx_rgbd_estimate(h, C.byref(intr), C.byref(bad), C.byref(init), C.byref(res)) != 0
    kfx.call("kfx_rgbd_estimate", h, C.byref(intr), C.byref(prm), C.byref(init), C.byref(res))
    # The same frame in both slots: the identity, found at once.
    assert res.ok == 1 and np.abs(np.array(res.pose.t[:])).max() < 1e-6
    kfx.call("kfx_rgbd_destroy", h)
