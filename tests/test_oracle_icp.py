"""Pins the oracle's ICP pieces: the 6x6 solve and pose increment against
numpy/scipy, the normal equations' structure, the reference reduction
against double sums, and recovery of a known camera motion.  No GPU needed."""
import numpy as np
from scipy.spatial.transform import Rotation

import icp_inputs as ii


def test_solve_matches_numpy_and_euler_convention(orc):
    rng = np.random.default_rng(0)
    for _ in range(20):
        M = rng.normal(size=(6, 6))
        A = M @ M.T + 6 * np.eye(6)
        b = rng.normal(size=6) * 0.01
        sums = [A[i, j] if j < 6 else b[i] for i in range(6) for j in range(i, 7)]
        R0 = Rotation.random(random_state=rng.integers(1 << 30)).as_matrix()
        t0 = rng.normal(size=3)
        rc, R, t, x = orc.icp_solve_update(sums, R0, t0)
        assert rc == 0
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-10, atol=1e-14)
        # R_inc = Rz(x2) Ry(x1) Rx(x0): scipy's extrinsic 'xyz' sequence.
        Ri = Rotation.from_euler("xyz", x[:3]).as_matrix()
        np.testing.assert_allclose(R, Ri @ R0, atol=1e-12)
        np.testing.assert_allclose(t, Ri @ t0 + x[3:], atol=1e-12)


def test_solve_rejects_singular_system(orc):
    R0, t0 = np.eye(3), np.array([1.0, 2.0, 3.0])
    rc, R, t, x = orc.icp_solve_update(np.zeros(27), R0, t0)
    assert rc == 1 and np.array_equal(R, R0) and np.array_equal(t, t0) and not x.any()
    # rank-deficient: only translation constraints along z
    s = np.zeros(27)
    A = np.zeros((6, 6))
    A[5, 5] = 1.0
    rc, *_ = orc.icp_solve_update([A[i, j] if j < 6 else 0.0 for i in range(6) for j in range(i, 7)], R0, t0)
    assert rc == 1


def test_normal_equations_structure(kfx, orc):
    intr = kfx.kinect_intrinsics(160, 120, 131.25)
    (vc, nc, vp, np_), (Rp, tp), (Rc, tc) = ii.icp_pair(orc, intr)
    s64, n = orc.icp_sums(vc, nc, vp, np_, Rp, tp, Rp.T, tp, intr, ii.DIST_THR, ii.ANGLE_THR, precision=1)
    assert n > 0.3 * intr.cols * intr.rows
    A, b = orc.unpack_sums(s64)
    assert np.allclose(A, A.T)
    assert np.linalg.eigvalsh(A).min() > 0
    # Entries 3..5 of the diagonal are sums of squared normal components:
    # they add up to the number of correspondences.
    assert abs(A[3, 3] + A[4, 4] + A[5, 5] - n) < 1e-3 * n
    s32, n32 = orc.icp_sums(vc, nc, vp, np_, Rp, tp, Rp.T, tp, intr, ii.DIST_THR, ii.ANGLE_THR, precision=0)
    assert abs(n32 - n) <= 2e-3 * n
    # The reference's single-precision reduction tracks the double sums.
    m = orc.estimate_combined_f32(vc, nc, vp, np_, Rp, tp, Rp.T, tp, intr, ii.DIST_THR, ii.ANGLE_THR)
    big = np.abs(s32) > 1e-3 * np.abs(s32).max()
    np.testing.assert_allclose(m[big], s32[big], rtol=1e-4)


def test_icp_recovers_known_motion(kfx, orc):
    # Noise-free frames 10 -> 11 of the C1 orbit (3 cm, 1.1 degree apart):
    # coarse-to-fine ICP from the previous pose lands on the true pose.
    intr = kfx.kinect_intrinsics(320, 240, 262.5)
    maps, (Rp, tp), (Rc, tc) = ii.icp_pair(orc, intr)
    pyr = ii.pyramid_maps(orc, maps)
    for precision in (0, 1):
        R, t = ii.oracle_icp(orc, pyr, intr, Rp, tp, Rp, tp, precision=precision)
        ang = np.linalg.norm(Rotation.from_matrix(R @ Rc.T).as_rotvec())
        assert ang < 3e-4, ang
        assert np.linalg.norm(t - tc) < 5e-4, np.linalg.norm(t - tc)  # from 30 mm off
