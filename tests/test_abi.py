"""The C ABI library: loads, exports every declared entry point, and its
host-side pieces (address formula, synthetic input) behave as specified.
No GPU needed."""
import ctypes as C

import numpy as np


def test_library_exports_every_declared_symbol(kfx):
    names = kfx.header_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(kfx.lib, n)]
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_every_declared_symbol_has_a_prototype(kfx):
    unbound = [n for n in kfx.header_functions() if n not in kfx._PROTOS]
    assert not unbound, unbound


def test_struct_layouts(kfx):
    assert C.sizeof(kfx.Intrinsics) == 24
    assert C.sizeof(kfx.RoundBuffer) == 72
    assert kfx.RoundBuffer.origin_GRID.offset == 12
    assert kfx.RoundBuffer.voxels_size.offset == 60
    assert C.sizeof(kfx.Matrix3f) == 36
    assert kfx.TSDF_DTYPE.itemsize == 4 and kfx.XYZ_DTYPE.itemsize == 12


def test_error_path_without_init_is_a_status_not_an_abort(kfx):
    # Before kfx_init (or without a GPU) entry points return a status code.
    p = C.c_void_p()
    rc = kfx.lib.kfx_malloc(C.byref(p), 16)
    if rc != 0:
        assert kfx.lib.kfx_last_error()


def test_host_cyclic_address_closed_form(kfx):
    # raycast.cl:29-47 on an 8x6x5 grid for every origin: with start_index 0
    # the address is a bijection onto [0, 240) given by the closed form.
    n = (8, 6, 5)
    for ox in range(n[0]):
        for oy in range(0, n[1], 2):
            for oz in range(n[2]):
                rb = kfx.round_buffer(n, 1.0, origin=(ox, oy, oz))
                seen = set()
                for z in range(n[2]):
                    for y in range(n[1]):
                        for x in range(n[0]):
                            a = kfx.lib.kfx_cyclic_address(x, y, z, C.byref(rb))
                            want = ((x + ox) % 8) + 8 * ((y + oy) % 6) + 48 * ((z + oz) % 5)
                            assert a == want
                            seen.add(a)
                assert seen == set(range(240))


def test_host_cyclic_address_start_index_fold(kfx):
    # start_index shifts the address; the fold applies only when strictly
    # greater than buffer_length (raycast.cl:44-45).
    rb = kfx.round_buffer((4, 4, 4), 1.0)
    rb.start_index = 10
    assert kfx.lib.kfx_cyclic_address(0, 0, 0, C.byref(rb)) == 10
    assert kfx.lib.kfx_cyclic_address(1, 3, 3, C.byref(rb)) == (1 + 12 + 48 + 10) % 64
    # 53 + 10 = 63 -> no fold; 54 + 10 = 64 == buffer_length -> no fold either.
    assert kfx.lib.kfx_cyclic_address(1, 1, 3, C.byref(rb)) == 1 + 4 + 48 + 10
    assert kfx.lib.kfx_cyclic_address(2, 1, 3, C.byref(rb)) == 64


def test_synth_render_deterministic_and_geometric(kfx):
    intr = kfx.kinect_intrinsics(160, 120, 131.25)
    scene = kfx.make_scene(3, 2, 2, (-1, 1, 2, 4), 1.0)
    pose = kfx.orbit((0.0, 0.0, 3.0), 3.0, 0.0, 0.0, 0.0, 1)[0]
    d1 = kfx.render(scene, pose, intr, 7, 0, noise=True, threads=1)
    d2 = kfx.render(scene, pose, intr, 7, 0, noise=True, threads=4)
    assert np.array_equal(d1, d2)
    d3 = kfx.render(scene, pose, intr, 7, 1, noise=True)
    assert not np.array_equal(d1, d3)
    # Noise-free ground plane: the camera sits at y=0 looking along +z, so the
    # rows below the principal point see the plane y=1 at z-depth
    # 1 / ((v - cy) / fy).
    empty = kfx.make_scene(3, 0, 0, (0, 0, 0, 0), 1.0)
    d, clean = kfx.render(empty, pose, intr, 7, 0, noise=False, want_clean=True)
    v = np.arange(intr.rows)[:, None] * np.ones((1, intr.cols))
    with np.errstate(divide="ignore"):
        z = np.where(v > intr.cy, 1.0 / ((v - intr.cy) / intr.fy), 0.0)
    np.testing.assert_allclose(clean, z, rtol=1e-6)
    zr = np.where((z >= 0.4) & (z <= 6.0), np.rint(z * 1000), 0)
    assert np.array_equal(d, zr.astype(np.uint16))


def test_synth_noise_statistics(kfx):
    # sigma(z) = 1.2 + 1.9 (z - 0.4)^2 mm and 2 % dropout on a wall at 2 m.
    intr = kfx.kinect_intrinsics(320, 240, 262.5)
    sc = kfx.make_scene(1, 1, 0, (0, 0, 0, 0), 100.0)
    sc.objects[0].center[:] = (0.0, 0.0, 102.0)
    sc.objects[0].half[:] = (100.0, 100.0, 100.0)
    pose = kfx.orbit((0.0, 0.0, 10.0), 10.0, 0.0, 0.0, 0.0, 1)[0]
    d = kfx.render(sc, pose, intr, 11, 0, noise=True).astype(np.float64)
    drop = np.mean(d == 0)
    assert 0.015 < drop < 0.025
    valid = d[d > 0]
    sigma = 1.2 + 1.9 * (2.0 - 0.4) ** 2
    assert abs(valid.mean() - 2000.0) < 0.2
    assert abs(valid.std() - np.sqrt(sigma ** 2 + 1 / 12.0)) < 0.1


def test_trajectory_file_round_trip_and_errors(kfx, tmp_path):
    gt = kfx.orbit((1.5, 2.0, 1.5), 1.5, -0.6, -0.3, 0.02, 20)
    path = str(tmp_path / "traj.txt").encode()
    kfx.call("kfx_trajectory_write", path, (kfx.Pose * 20)(*gt), 20)
    text = open(path.decode()).read().splitlines()
    assert text[0] == "0" and text[1].count("x") == 2  # reference layout: index, 3 rows, t
    back = (kfx.Pose * 30)()
    n = C.c_int()
    kfx.call("kfx_trajectory_read", path, back, 30, C.byref(n))
    assert n.value == 20
    for a, b in zip(back[:20], gt):
        assert np.allclose(a.R[:], b.R[:], atol=1e-5) and np.allclose(a.t[:], b.t[:], atol=1e-5)
    # errors: identical -> 0; a constant 1 cm offset after frame 10 -> ATE, RPE closed forms
    ate0, rt0, rr0 = kfx.trajectory_errors(list(gt), list(gt))
    assert ate0 == 0.0 and rt0 < 1e-15 and rr0 < 1e-7  # acos near 1 resolves ~1e-8 rad
    est = (kfx.Pose * 20)(*gt)
    for i in range(10, 20):
        est[i].t[0] += 0.01
    ate, rt, rr = kfx.trajectory_errors(list(est), list(gt))
    assert abs(ate - 0.01 * np.sqrt(10 / 20)) < 1e-12
    assert abs(rt - 0.01 * np.sqrt(1 / 19)) < 1e-9 and rr < 1e-7
