"""GPU parity of the ICP paths against the oracle: the reference-style
Estimate_Combined reduction bit for bit; the fused path's normal equations
(compensated FP32, plain FP32 and FP64 workgroup sums), solve and pose update against double precision; a
coarse-to-fine run within 1e-5 rad / 1e-5 m of the oracle's FP64 ICP; the
singular-system and early-exit behaviour."""
import ctypes as C

import numpy as np
import pytest

import icp_inputs as ii

pytestmark = pytest.mark.gpu


def _upload(kfx, maps):
    return [kfx.Buffer.from_array(np.ascontiguousarray(m, np.float32)) for m in maps]


@pytest.fixture(scope="module")
def pair(gpu, orc):
    intr = gpu.kinect_intrinsics()
    maps, prev, cur = ii.icp_pair(orc, intr, noise=True)
    return intr, maps, prev, cur


def test_estimate_combined_bit_exact(gpu, orc, pair):
    kfx = gpu
    intr, maps, (Rp, tp), (Rc, tc) = pair
    bufs = _upload(kfx, maps)
    stride = ((intr.cols + 31) // 32) * ((intr.rows + 7) // 8)
    gbuf = kfx.Buffer(27 * stride * 4)
    mbuf = kfx.Buffer(27 * 4)
    A = (C.c_double * 36)()
    b = (C.c_double * 6)()
    kfx.call("kfx_clean_buff", gbuf.ptr, stride)
    kfx.call("kfx_estimate_combined", C.byref(kfx.matrix3f(Rp)), C.byref(kfx.point3f(tp)), bufs[0].ptr, bufs[1].ptr,
             C.byref(kfx.matrix3f(Rp.T)), C.byref(kfx.point3f(tp)), C.byref(intr), bufs[2].ptr, bufs[3].ptr,
             ii.DIST_THR, ii.ANGLE_THR, gbuf.ptr, stride, mbuf.ptr, A, b)
    got = mbuf.download(np.float32, 27)
    want = orc.estimate_combined_f32(*maps, Rp, tp, Rp.T, tp, intr, ii.DIST_THR, ii.ANGLE_THR)
    assert np.array_equal(got, want)
    Aw, bw = orc.unpack_sums(want.astype(np.float64))
    assert np.array_equal(np.array(A[:]).reshape(6, 6), Aw) and np.array_equal(np.array(b[:]), bw)


def _icp_handle(kfx, npix):
    h = C.c_void_p()
    kfx.call("kfx_icp_create", npix, C.byref(h))
    return h


def _pose(kfx, R, t):
    p = kfx.Pose()
    p.R[:] = [float(v) for v in np.asarray(R).reshape(9)]
    p.t[:] = [float(v) for v in np.asarray(t).reshape(3)]
    return p


@pytest.mark.parametrize("level", [0, 2])
@pytest.mark.parametrize("mode,rtol", [(0, 1e-9), (2, 1e-9), (1, 1e-4)])
def test_fused_iteration_matches_double(gpu, orc, pair, mode, rtol, level):
    # mode 0: compensated FP32 workgroup sums (default), 2: FP64 sums, 1: plain
    # FP32 sums.  All three are compared with the FP32 products summed in
    # double precision; the compensated and FP64 sums agree to 1e-9, plain
    # FP32 sums over ~1e5 products only to ~1e-5.  Level 0 (640x480) takes two
    # pixels per thread, level 2 (160x120) one.
    kfx = gpu
    intr0, maps0, (Rp, tp), (Rc, tc) = pair
    intr, maps = intr0.level(level), ii.pyramid_maps(orc, maps0)[level]
    bufs = _upload(kfx, maps)
    h = _icp_handle(kfx, intr.cols * intr.rows)
    kfx.call("kfx_icp_set_sum_mode", h, mode)
    kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
    kfx.call("kfx_icp_iterate", h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, C.byref(kfx.matrix3f(Rp.T)),
             C.byref(kfx.point3f(tp)), C.byref(intr), ii.DIST_THR, ii.ANGLE_THR, -1, 0.0, 0.0)
    res = kfx.IcpResult()
    kfx.call("kfx_icp_read", h, C.byref(res))
    s, n = orc.icp_sums(*maps, Rp, tp, Rp.T, tp, intr, ii.DIST_THR, ii.ANGLE_THR, precision=0)
    A, b = orc.unpack_sums(s)
    np.testing.assert_allclose(np.array(res.A[:]).reshape(6, 6), A, rtol=rtol, atol=rtol * np.abs(A).max())
    np.testing.assert_allclose(np.array(res.b[:]), b, rtol=rtol, atol=rtol * np.abs(b).max())
    if mode != 1:
        rc, R, t, x = orc.icp_solve_update(s, Rp, tp)
        assert res.status == 0 and rc == 0
        np.testing.assert_allclose(np.array(res.x[:]), x, atol=1e-9)
        np.testing.assert_allclose(np.array(res.pose.R[:]).reshape(3, 3), R, atol=1e-9)
        np.testing.assert_allclose(np.array(res.pose.t[:]), t, atol=1e-9)
    kfx.call("kfx_icp_destroy", h)


def test_icp_sum_mode_rejects_unknown(gpu):
    kfx = gpu
    h = _icp_handle(kfx, 1024)
    assert kfx.lib.kfx_icp_set_sum_mode(h, 7) == 1
    kfx.call("kfx_icp_destroy", h)


def test_fused_coarse_to_fine_matches_oracle_pose(gpu, orc, pair):
    kfx = gpu
    intr, maps, (Rp, tp), (Rc, tc) = pair
    pyr = ii.pyramid_maps(orc, maps)
    dev = [_upload(kfx, m) for m in pyr]
    h = _icp_handle(kfx, intr.cols * intr.rows)
    kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
    for level in (2, 1, 0):
        li = intr.level(level)
        b = dev[level]
        for _ in range((10, 5, 4)[level]):
            kfx.call("kfx_icp_iterate", h, b[0].ptr, b[1].ptr, b[2].ptr, b[3].ptr, C.byref(kfx.matrix3f(Rp.T)),
                     C.byref(kfx.point3f(tp)), C.byref(li), ii.DIST_THR, ii.ANGLE_THR, -1, 0.0, 0.0)
    res = kfx.IcpResult()
    kfx.call("kfx_icp_read", h, C.byref(res))
    assert res.iterations == 19 and res.singular_count == 0
    R, t = ii.oracle_icp(orc, pyr, intr, Rp, tp, Rp, tp, precision=1)
    Rg = np.array(res.pose.R[:]).reshape(3, 3)
    tg = np.array(res.pose.t[:])
    from scipy.spatial.transform import Rotation
    assert np.linalg.norm(Rotation.from_matrix(Rg @ R.T).as_rotvec()) < 1e-5
    assert np.linalg.norm(tg - t) < 1e-5
    # and it tracked the true motion
    assert np.linalg.norm(tg - tc) < 2e-3
    kfx.call("kfx_icp_destroy", h)


def test_fused_singular_keeps_pose_and_early_exit(gpu, orc, pair):
    kfx = gpu
    intr, maps, (Rp, tp), (Rc, tc) = pair
    bufs = _upload(kfx, maps)
    nan = kfx.Buffer.from_array(np.full(maps[0].shape, np.nan, np.float32))
    h = _icp_handle(kfx, intr.cols * intr.rows)
    kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
    for _ in range(3):  # an all-invalid frame: no correspondences
        kfx.call("kfx_icp_iterate", h, nan.ptr, nan.ptr, bufs[2].ptr, bufs[3].ptr, C.byref(kfx.matrix3f(Rp.T)),
                 C.byref(kfx.point3f(tp)), C.byref(intr), ii.DIST_THR, ii.ANGLE_THR, -1, 0.0, 0.0)
    res = kfx.IcpResult()
    kfx.call("kfx_icp_read", h, C.byref(res))
    assert res.status == 5 and res.singular_count == 3
    assert np.array_equal(np.array(res.pose.t[:]), tp) and np.array_equal(np.array(res.pose.R[:]), Rp.reshape(9))
    # Early exit: a loose convergence test stops the level after one solve.
    kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
    for _ in range(6):
        kfx.call("kfx_icp_iterate", h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr,
                 C.byref(kfx.matrix3f(Rp.T)), C.byref(kfx.point3f(tp)), C.byref(intr), ii.DIST_THR, ii.ANGLE_THR,
                 7, 1.0, 1.0)
    kfx.call("kfx_icp_read", h, C.byref(res))
    assert res.iterations == 1 and res.converged_tag == 7
    kfx.call("kfx_icp_destroy", h)


@pytest.mark.parametrize("size", [(1280, 720), (1600, 1200)])
@pytest.mark.parametrize("mode", [0, 2])
def test_fused_iteration_large_images(gpu, orc, size, mode):
    # 900 and 1875 workgroups of partial sums: the one-workgroup solve then
    # loads two or three workgroups' partials per round before the remainder.
    kfx = gpu
    W, H = size
    intr = kfx.kinect_intrinsics(W, H, 525.0 * W / 640)
    maps, (Rp, tp), _ = ii.icp_pair(orc, intr, noise=True)
    bufs = _upload(kfx, maps)
    h = _icp_handle(kfx, W * H)
    kfx.call("kfx_icp_set_sum_mode", h, mode)
    kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
    kfx.call("kfx_icp_iterate", h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, C.byref(kfx.matrix3f(Rp.T)),
             C.byref(kfx.point3f(tp)), C.byref(intr), ii.DIST_THR, ii.ANGLE_THR, -1, 0.0, 0.0)
    res = kfx.IcpResult()
    kfx.call("kfx_icp_read", h, C.byref(res))
    s, n = orc.icp_sums(*maps, Rp, tp, Rp.T, tp, intr, ii.DIST_THR, ii.ANGLE_THR, precision=0)
    assert n > 10000
    A, b = orc.unpack_sums(s)
    np.testing.assert_allclose(np.array(res.A[:]).reshape(6, 6), A, rtol=1e-9, atol=1e-9 * np.abs(A).max())
    np.testing.assert_allclose(np.array(res.b[:]), b, rtol=1e-9, atol=1e-9 * np.abs(b).max())
    kfx.call("kfx_icp_destroy", h)


@pytest.mark.parametrize("conv", [0.0, 1.0])
def test_published_state_equals_read(gpu, orc, pair, conv):
    # The last solve's copy in pinned host memory equals a read of the device
    # state, also when that iteration is skipped by the early exit (conv 1.0
    # stops the level after one solve).
    kfx = gpu
    intr, maps, (Rp, tp), (Rc, tc) = pair
    bufs = _upload(kfx, maps)
    h = _icp_handle(kfx, intr.cols * intr.rows)
    kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
    kfx.call("kfx_icp_begin_frame", h)
    for it in range(4):
        if it == 3:
            kfx.call("kfx_icp_publish_next", h)
        kfx.call("kfx_icp_iterate", h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, None, None,
                 C.byref(intr), ii.DIST_THR, ii.ANGLE_THR, 0, conv, conv)
    a, b = kfx.IcpResult(), kfx.IcpResult()
    kfx.call("kfx_icp_read_published", h, C.byref(a))
    kfx.call("kfx_icp_read", h, C.byref(b))
    assert bytes(a) == bytes(b)
    assert a.iterations == (1 if conv > 0.0 else 4)
    kfx.call("kfx_icp_destroy", h)


@pytest.mark.parametrize("width", [1200, 1201, 37])
def test_clean_buff_zeroes_every_entry(gpu, width):
    # Clean_Buff (clean_buff.cl:4-11): the 27 rows of `width` partials become
    # +0.0; the words after the last row keep their contents.  1200 is the
    # reference's stride, the others leave a partial last workgroup.
    kfx = gpu
    n = 27 * width
    buf = kfx.Buffer((n + 64) * 4)
    buf.fill(0xA5)
    kfx.call("kfx_clean_buff", buf.ptr, width)
    kfx.synchronize()
    got = buf.download(np.uint32, n + 64)
    assert not got[:n].any()
    assert np.all(got[n:] == 0xA5A5A5A5)


def test_clean_buff_rejects_bad_width(gpu):
    kfx = gpu
    buf = kfx.Buffer(27 * 4)
    assert kfx.lib.kfx_clean_buff(buf.ptr, 0) != 0
    assert kfx.lib.kfx_clean_buff(None, 4) != 0


def test_read_published_tracks_what_was_enqueued(gpu, orc, pair):
    # kfx_icp_read_published fails unless a publishing iteration was enqueued
    # since the last reset or translate; a publish request is consumed by the
    # next kfx_icp_iterate call even when that call fails its checks.
    kfx = gpu
    intr, maps, (Rp, tp), _ = pair
    bufs = _upload(kfx, maps)
    h = _icp_handle(kfx, intr.cols * intr.rows)
    res = kfx.IcpResult()
    kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
    assert kfx.lib.kfx_icp_read_published(h, C.byref(res)) != 0
    assert b"publishing" in kfx.lib.kfx_last_error()

    def iterate(vc):
        return kfx.lib.kfx_icp_iterate(h, vc, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, C.byref(kfx.matrix3f(Rp.T)),
                                       C.byref(kfx.point3f(tp)), C.byref(intr), ii.DIST_THR, ii.ANGLE_THR, -1, 0.0,
                                       0.0)

    kfx.call("kfx_icp_publish_next", h)
    assert iterate(None) != 0                      # fails its checks, consumes the request
    assert iterate(bufs[0].ptr) == 0               # no request: publishes nothing
    assert kfx.lib.kfx_icp_read_published(h, C.byref(res)) != 0
    kfx.call("kfx_icp_publish_next", h)
    assert iterate(bufs[0].ptr) == 0
    kfx.call("kfx_icp_read_published", h, C.byref(res))
    full = kfx.IcpResult()
    kfx.call("kfx_icp_read", h, C.byref(full))
    assert res.iterations == full.iterations == 2
    assert np.array_equal(np.array(res.pose.t[:]), np.array(full.pose.t[:]))
    assert np.array_equal(np.array(res.pose.R[:]), np.array(full.pose.R[:]))
    kfx.call("kfx_icp_translate", h, 0.01, 0.0, 0.0)
    assert kfx.lib.kfx_icp_read_published(h, C.byref(res)) != 0
    kfx.call("kfx_icp_destroy", h)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("level", [0, 1, 2])
def test_one_launch_iteration_equals_two(gpu, orc, pair, mode, level):
    # Five iterations from the same estimate with the solve in the last
    # workgroup of the sums launch and with the solve as its own launch:
    # the states (pose, normal equations, increment) are bit-identical in
    # every sum mode, at 640x480 (two pixels per thread, 600 workgroups) and
    # on the coarse levels.
    kfx = gpu
    intr0, maps0, (Rp, tp), _ = pair
    intr, maps = intr0.level(level), ii.pyramid_maps(orc, maps0)[level]
    bufs = _upload(kfx, maps)
    out = []
    for launch in (kfx.ICP_LAUNCH_ONE, kfx.ICP_LAUNCH_TWO):
        h = _icp_handle(kfx, intr0.cols * intr0.rows)
        kfx.call("kfx_icp_set_sum_mode", h, mode)
        kfx.call("kfx_icp_set_launch_mode", h, launch)
        kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
        for _ in range(5):
            kfx.call("kfx_icp_iterate", h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr,
                     C.byref(kfx.matrix3f(Rp.T)), C.byref(kfx.point3f(tp)), C.byref(intr), ii.DIST_THR,
                     ii.ANGLE_THR, -1, 0.0, 0.0)
        res = kfx.IcpResult()
        kfx.call("kfx_icp_read", h, C.byref(res))
        out.append((np.array(res.pose.R[:]), np.array(res.pose.t[:]), np.array(res.A[:]), np.array(res.x[:]),
                    res.iterations, res.status))
        kfx.call("kfx_icp_destroy", h)
    a, b = out
    assert a[4] == b[4] == 5 and a[5] == b[5] == 0
    for u, v in zip(a[:4], b[:4]):
        assert np.array_equal(u, v)


def test_one_launch_ticket_is_reset(gpu, orc, pair):
    # Iterations on images of different sizes through one handle: each
    # launch finds the ticket at zero (the last workgroup resets it), so the
    # states still equal the two-launch form's.
    kfx = gpu
    intr0, maps0, (Rp, tp), _ = pair
    pyr = ii.pyramid_maps(orc, maps0)
    results = []
    for launch in (kfx.ICP_LAUNCH_ONE, kfx.ICP_LAUNCH_TWO):
        h = _icp_handle(kfx, intr0.cols * intr0.rows)
        kfx.call("kfx_icp_set_launch_mode", h, launch)
        kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
        for level in (2, 0, 1, 2, 0):
            bufs = _upload(kfx, pyr[level])
            kfx.call("kfx_icp_iterate", h, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr,
                     C.byref(kfx.matrix3f(Rp.T)), C.byref(kfx.point3f(tp)), C.byref(intr0.level(level)),
                     ii.DIST_THR, ii.ANGLE_THR, -1, 0.0, 0.0)
        res = kfx.IcpResult()
        kfx.call("kfx_icp_read", h, C.byref(res))
        results.append((np.array(res.pose.R[:]), np.array(res.pose.t[:])))
        kfx.call("kfx_icp_destroy", h)
    assert np.array_equal(results[0][0], results[1][0]) and np.array_equal(results[0][1], results[1][1])


def _frame_levels(kfx, intr0, pyr_bufs, iters):
    lv = (kfx.IcpLevel * 3)()
    for l in range(3):
        b = pyr_bufs[l]
        lv[l] = kfx.IcpLevel(b[0].ptr.value, b[1].ptr.value, b[2].ptr.value, b[3].ptr.value, intr0.level(l), iters[l])
    return lv


def _state(kfx, h, published=False):
    res = kfx.IcpResult()
    kfx.call("kfx_icp_read_published" if published else "kfx_icp_read", h, C.byref(res))
    return (np.array(res.pose.R[:]), np.array(res.pose.t[:]), np.array(res.A[:]), np.array(res.b[:]),
            np.array(res.x[:]), res.iterations, res.status, res.singular_count, res.converged_tag)


def _same_state(a, b):
    for u, v in zip(a[:5], b[:5]):
        assert np.array_equal(u, v)
    assert a[5:] == b[5:]


@pytest.fixture(scope="module")
def pyr_bufs(gpu, orc, pair):
    _, maps0, _, _ = pair
    return [_upload(gpu, m) for m in ii.pyramid_maps(orc, maps0)]


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("iters,tagged", [((10, 5, 4), 0), ((10, 5, 4), 1), ((3, 0, 2), 0)])
@pytest.mark.parametrize("begin", [0, 1])
def test_frame_one_launch_equals_iterations(gpu, pair, pyr_bufs, mode, iters, tagged, begin):
    # A frame's ICP as one launch (256 resident workgroups, a grid barrier
    # per iteration) and as begin-frame plus two-launch iterations: the
    # states are bit-identical in every sum mode, with and without the early
    # exit, with and without the begin-frame update, and the published copy
    # equals the state.
    kfx = gpu
    intr0, _, (Rp, tp), _ = pair
    lv = _frame_levels(kfx, intr0, pyr_bufs, iters)
    conv = 1e-5 if tagged else 0.0
    out = []
    for launch in (kfx.ICP_LAUNCH_FRAME, kfx.ICP_LAUNCH_TWO):
        h = _icp_handle(kfx, intr0.cols * intr0.rows)
        resident = C.c_int()
        kfx.call("kfx_icp_frame_resident", h, C.byref(resident))
        if launch == kfx.ICP_LAUNCH_FRAME:
            print(f"one-launch frame resident: {resident.value}")
            assert resident.value == 1
        kfx.call("kfx_icp_set_sum_mode", h, mode)
        kfx.call("kfx_icp_set_launch_mode", h, launch)
        kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
        kfx.call("kfx_icp_translate", h, 0.004, -0.003, 0.002)
        kfx.call("kfx_icp_publish_next", h)
        kfx.call("kfx_icp_frame", h, lv, begin, ii.DIST_THR, ii.ANGLE_THR, tagged, conv, conv)
        st = _state(kfx, h)
        _same_state(st, _state(kfx, h, published=True))
        out.append(st)
        kfx.call("kfx_icp_destroy", h)
    a, b = out
    _same_state(a, b)
    assert a[6] == 0
    if tagged:
        assert a[5] < sum(iters)
    else:
        assert a[5] == sum(iters)


def test_frame_one_launch_consecutive_frames(gpu, pair, pyr_bufs):
    # Several frames through one handle, with different iteration counts:
    # every launch finds the barrier counters at zero (the last workgroup
    # out resets them), so the states still equal the separate launches'.
    kfx = gpu
    intr0, _, (Rp, tp), _ = pair
    plans = [(10, 5, 4), (1, 1, 1), (0, 0, 3), (2, 0, 0), (10, 5, 4)]
    out = []
    for launch in (kfx.ICP_LAUNCH_FRAME, kfx.ICP_LAUNCH_TWO):
        h = _icp_handle(kfx, intr0.cols * intr0.rows)
        kfx.call("kfx_icp_set_launch_mode", h, launch)
        kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, Rp, tp)))
        states = []
        for iters in plans:
            kfx.call("kfx_icp_translate", h, 0.003, 0.0, -0.002)
            kfx.call("kfx_icp_frame", h, _frame_levels(kfx, intr0, pyr_bufs, iters), 0, ii.DIST_THR, ii.ANGLE_THR,
                     0, 0.0, 0.0)
            states.append(_state(kfx, h))
        out.append(states)
        kfx.call("kfx_icp_destroy", h)
    for a, b in zip(*out):
        _same_state(a, b)


def test_frame_argument_checks(gpu, pair, pyr_bufs):
    kfx = gpu
    intr0, _, _, _ = pair
    h = _icp_handle(kfx, (intr0.cols * intr0.rows) // 4)
    lv = _frame_levels(kfx, intr0, pyr_bufs, (10, 5, 4))
    assert kfx.lib.kfx_icp_frame(h, lv, 1, ii.DIST_THR, ii.ANGLE_THR, 0, 0.0, 0.0) != 0
    assert b"capacity" in kfx.lib.kfx_last_error()
    lv = _frame_levels(kfx, intr0, pyr_bufs, (0, -1, 4))
    assert kfx.lib.kfx_icp_frame(h, lv, 1, ii.DIST_THR, ii.ANGLE_THR, 0, 0.0, 0.0) != 0
    lv = _frame_levels(kfx, intr0, pyr_bufs, (0, 0, 4))
    lv[2].vmap_current = None
    assert kfx.lib.kfx_icp_frame(h, lv, 1, ii.DIST_THR, ii.ANGLE_THR, 0, 0.0, 0.0) != 0
    kfx.call("kfx_icp_destroy", h)


def _pack_sums(A, b):
    """[A | b] -> the 27 sums (row k of its upper triangle; opencl_wrappers.cpp:426-437)."""
    return np.array([A[i, j] if j < 6 else b[i] for i in range(6) for j in range(i, 7)], np.float64)


@pytest.mark.parametrize("angles", [(1e-3, -2e-3, 5e-4),      # the series
                                    (0.125, -0.1, 0.05),       # the series' boundary
                                    (0.3, -1.2, 2.9),          # halving and double-angle steps
                                    (3.5, -7.5, 100.0),        # reduction by 2 pi first
                                    (-0.2, 0.01, -1000.0)])
def test_solve_sums_matches_oracle(gpu, orc, angles):
    # The GPU's solve (Gauss-Jordan over one wave) and pose update, against
    # the oracle's Cholesky solve with libm sin/cos, on a positive definite
    # system whose solution holds the given angles; every increment angle
    # range of the update is reached (series, halving, reduction by 2 pi).
    from scipy.spatial.transform import Rotation
    kfx = gpu
    rng = np.random.default_rng(abs(hash(angles)) % 2**32)
    Q, _ = np.linalg.qr(rng.normal(size=(6, 6)))
    A = (Q * np.geomspace(1.0, 1e3, 6)) @ Q.T * 1e4
    A = (A + A.T) / 2
    x_true = np.array(list(angles) + list(rng.normal(scale=0.05, size=3)))
    sums = _pack_sums(A, A @ x_true)
    R0 = Rotation.from_rotvec(rng.normal(scale=0.5, size=3)).as_matrix()
    t0 = rng.normal(size=3)
    h = _icp_handle(kfx, 1024)
    kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, R0, t0)))
    kfx.call("kfx_icp_solve_sums", h, sums.ctypes.data_as(C.c_void_p))
    res = kfx.IcpResult()
    kfx.call("kfx_icp_read", h, C.byref(res))
    rc, R, t, x = orc.icp_solve_update(sums, R0, t0)
    assert rc == 0 and res.status == 0 and res.iterations == 1
    # Both solves are backward stable: they differ by about
    # cond(A) eps max|x| (cond(A) = 1e3), the poses by as much.
    tol = 1e3 * 1e-15 * max(1.0, np.abs(x_true).max())
    np.testing.assert_allclose(np.array(res.x[:]), x, rtol=0, atol=tol)
    np.testing.assert_allclose(x, x_true, rtol=0, atol=tol)
    np.testing.assert_allclose(np.array(res.pose.R[:]).reshape(3, 3), R, atol=tol)
    np.testing.assert_allclose(np.array(res.pose.t[:]), t, atol=tol)
    Rg = np.array(res.pose.R[:]).reshape(3, 3)
    np.testing.assert_allclose(Rg @ Rg.T, np.eye(3), atol=1e-13)
    np.testing.assert_array_equal(np.array(res.A[:]).reshape(6, 6), orc.unpack_sums(sums)[0])
    kfx.call("kfx_icp_destroy", h)


@pytest.mark.parametrize("bad", ["zero", "nan", "indefinite"])
def test_solve_sums_singular_keeps_pose(gpu, orc, bad):
    kfx = gpu
    A = np.diag([4.0, 3.0, 2.0, 5.0, 6.0, 7.0])
    if bad == "zero":
        A[:] = 0.0
    elif bad == "nan":
        A[2, 2] = np.nan
    else:
        A[4, 4] = -1.0
    sums = _pack_sums(A, np.ones(6))
    R0, t0 = np.eye(3), np.array([0.1, 0.2, 0.3])
    h = _icp_handle(kfx, 1024)
    kfx.call("kfx_icp_reset", h, C.byref(_pose(kfx, R0, t0)))
    kfx.call("kfx_icp_solve_sums", h, sums.ctypes.data_as(C.c_void_p))
    res = kfx.IcpResult()
    kfx.call("kfx_icp_read", h, C.byref(res))
    rc, *_ = orc.icp_solve_update(sums, R0, t0)
    assert rc == 1 and res.status == 5 and res.singular_count == 1 and res.iterations == 1
    assert np.array_equal(np.array(res.pose.R[:]), R0.reshape(9)) and np.array_equal(np.array(res.pose.t[:]), t0)
    assert np.array_equal(np.array(res.x[:]), np.zeros(6))
    kfx.call("kfx_icp_destroy", h)
