"""GPU parity of the volume kernels against the oracle: cyclic addressing
(bit-exact), TSDF integration (bit-exact; the bar is 1 LSB on >= 99.99 % of
voxels) and raycast (vertex within 0.1 mm, NaN masks identical)."""
import ctypes as C

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


def _integrate_gpu(kfx, dvol, depth, intr, R_inv, t, trunc, rb):
    dd = kfx.Buffer.from_array(depth)
    ds = kfx.Buffer(depth.size * 4)
    cs = kfx.cell_size(rb)
    kfx.call("kfx_integrate_tsdf", dd.ptr, ds.ptr, dvol.ptr, C.byref(intr), C.byref(kfx.matrix3f(R_inv)),
             C.byref(kfx.point3f(t)), float(trunc), C.byref(cs), C.byref(rb))
    kfx.synchronize()
    return ds


@pytest.mark.parametrize("origin", [(0, 0, 0), (3, 200, 77)])
def test_cyclic_address_batch_bit_exact(gpu, origin):
    kfx = gpu
    rb = kfx.round_buffer(256, 3.0, origin=origin)
    rng = np.random.default_rng(1)
    xyz = rng.integers(0, 256, size=(100000, 3)).astype(np.int32)
    dx = kfx.Buffer.from_array(xyz)
    do = kfx.Buffer(4 * len(xyz))
    kfx.call("kfx_cyclic_address_batch", dx.ptr, do.ptr, len(xyz), C.byref(rb))
    got = do.download(np.uint32, len(xyz))
    o = np.array(origin)
    w = (xyz + o) % 256
    want = (w[:, 0] + 256 * w[:, 1] + 65536 * w[:, 2]).astype(np.uint32)
    assert np.array_equal(got, want)


def test_init_tsdf_zeroes_volume(gpu):
    kfx = gpu
    rb = kfx.round_buffer(64, 1.0)
    dv = kfx.Buffer(rb.buffer_length * 4)
    dv.fill(0x5A)
    kfx.call("kfx_init_tsdf", dv.ptr, C.byref(rb))
    assert not dv.download(np.uint8, rb.buffer_length * 4).any()


@pytest.mark.parametrize("origin", [(0, 0, 0), (17, 101, 250)])
def test_integrate_matches_oracle_c1(gpu, orc, origin):
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    rb = kfx.round_buffer(256, 3.0, origin=origin)
    scene, poses = scenes.c1_scene(), scenes.c1_poses(30)
    ref = scenes.empty_volume(rb)
    dv = kfx.Buffer(rb.buffer_length * 4)
    kfx.call("kfx_init_tsdf", dv.ptr, C.byref(rb))
    total = 0
    for f in (0, 7, 15):
        R, t = poses[f].matrices()
        depth = kfx.render(scene, poses[f], intr, scenes.C1_SEED, f, noise=True)
        total += orc.integrate_f32(ref, depth, intr, R.T, t, scenes.TRUNC, rb)
        ds = _integrate_gpu(kfx, dv, depth, intr, R.T, t, scenes.TRUNC, rb)
        # scale_depth is bit-exact as well
        assert np.array_equal(ds.download(np.float32, depth.shape), orc.scale_depth_f32(depth, intr))
    got = dv.download(kfx.TSDF_DTYPE, rb.buffer_length)
    print(f"N_upd over 3 frames: {total}")
    dval = np.abs(got["value"].astype(np.int32) - ref["value"])
    assert np.mean(dval <= 1) >= 0.9999
    assert np.array_equal(got["weight"], ref["weight"])
    assert np.array_equal(got["value"], ref["value"]), f"{np.count_nonzero(dval)} voxels differ"


def test_integrate_rejects_bad_round_buffer(gpu):
    kfx = gpu
    rb = kfx.round_buffer(32, 1.0)
    rb.origin_GRID = kfx.Point3i(40, 0, 0)
    intr = kfx.kinect_intrinsics(64, 48, 52.5)
    dv = kfx.Buffer(rb.buffer_length * 4)
    dd = kfx.Buffer(64 * 48 * 2)
    ds = kfx.Buffer(64 * 48 * 4)
    cs = kfx.cell_size(rb)
    rc = kfx.lib.kfx_integrate_tsdf(dd.ptr, ds.ptr, dv.ptr, C.byref(intr), C.byref(kfx.matrix3f(np.eye(3))),
                                    C.byref(kfx.point3f((0, 0, 0))), 0.03, C.byref(cs), C.byref(rb))
    assert rc == 1 and b"origin_GRID" in kfx.lib.kfx_last_error()


@pytest.mark.parametrize("origin", [(0, 0, 0), (17, 101, 250)])
def test_raycast_matches_oracle_c1(gpu, orc, origin):
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    rb = kfx.round_buffer(256, 3.0, origin=origin)
    scene, poses = scenes.c1_scene(), scenes.c1_poses(30)
    ref = scenes.empty_volume(rb)
    for f in range(0, 30, 3):
        R, t = poses[f].matrices()
        depth = kfx.render(scene, poses[f], intr, scenes.C1_SEED, f, noise=True)
        orc.integrate_f32(ref, depth, intr, R.T, t, scenes.TRUNC, rb)
    dv = kfx.Buffer.from_array(ref)
    dvm = kfx.Buffer(intr.cols * intr.rows * 12)
    dnm = kfx.Buffer(intr.cols * intr.rows * 12)
    cs = kfx.cell_size(rb)
    for f in (4, 20):
        R, t = poses[f].matrices()
        kfx.call("kfx_raycast", dv.ptr, dvm.ptr, dnm.ptr, C.byref(rb), C.byref(kfx.point3f(t)),
                 C.byref(kfx.matrix3f(R)), scenes.TRUNC, C.byref(intr), C.byref(cs))
        vm = dvm.download(np.float32, (intr.rows, intr.cols, 3))
        nm = dnm.download(np.float32, (intr.rows, intr.cols, 3))
        rv, rn = orc.raycast_f32(ref, rb, t, R, scenes.TRUNC, intr)
        assert np.array_equal(np.isnan(vm), np.isnan(rv))
        assert np.array_equal(np.isnan(nm), np.isnan(rn))
        hit = ~np.isnan(rv[..., 0])
        assert hit.mean() > 0.5
        assert np.max(np.abs(vm[hit] - rv[hit])) <= 1e-4
        nh = ~np.isnan(rn[..., 0])
        assert np.max(np.abs(nm[nh] - rn[nh])) <= 1e-4
        assert np.array_equal(vm[hit], rv[hit]), "raycast vertices not bit-identical"
        # The counting variant: same maps, and the oracle's count of the
        # algorithm's voxel reads (the raycast's algorithmic bytes 4 S + 24 W H).
        ns = C.c_uint64()
        kfx.call("kfx_raycast_count_samples", dv.ptr, dvm.ptr, dnm.ptr, C.byref(rb), C.byref(kfx.point3f(t)),
                 C.byref(kfx.matrix3f(R)), scenes.TRUNC, C.byref(intr), C.byref(cs), C.byref(ns))
        _, _, s_ref = orc.raycast_f32(ref, rb, t, R, scenes.TRUNC, intr, want_samples=True)
        assert np.array_equal(dvm.download(np.float32, (intr.rows, intr.cols, 3)), vm, equal_nan=True)
        assert ns.value == s_ref and s_ref > 10 * intr.cols * intr.rows


@pytest.mark.parametrize("origin", [(0, 0, 0), (17, 101, 250)])
def test_raycast_bricks_equal_full_march(gpu, orc, origin):
    # The skipping march loads only samples in bricks flagged as holding a
    # negative value; with exact flags, and with every brick flagged, its
    # maps equal the full march's bit for bit.  The flags kernel equals a
    # CPU evaluation of the same rule.
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    rb = kfx.round_buffer(256, 3.0, origin=origin)
    scene, poses = scenes.c1_scene(), scenes.c1_poses(30)
    ref = scenes.empty_volume(rb)
    for f in range(0, 30, 3):
        R, t = poses[f].matrices()
        depth = kfx.render(scene, poses[f], intr, scenes.C1_SEED, f, noise=True)
        orc.integrate_f32(ref, depth, intr, R.T, t, scenes.TRUNC, rb)
    nbytes = C.c_size_t()
    kfx.call("kfx_brick_flags_size", C.byref(rb), C.byref(nbytes))
    assert nbytes.value == (256 // 16) ** 3
    dv = kfx.Buffer.from_array(ref)
    db = kfx.Buffer(nbytes.value)
    kfx.call("kfx_brick_flags_build", dv.ptr, C.byref(rb), db.ptr)
    kfx.synchronize()
    flags = db.download(np.uint8, nbytes.value)
    want = scenes.brick_flags_cpu(ref, rb)
    assert np.array_equal(flags != 0, want != 0)
    assert 0.01 < want.mean() < 0.5, want.mean()
    ones = kfx.Buffer.from_array(np.ones(nbytes.value, np.uint8))
    n = intr.cols * intr.rows
    vm, nm, vb, nb = (kfx.Buffer(n * 12) for _ in range(4))
    cs = kfx.cell_size(rb)
    for f in (4, 11, 20, 29):
        R, t = poses[f].matrices()
        args = (C.byref(rb), C.byref(kfx.point3f(t)), C.byref(kfx.matrix3f(R)), scenes.TRUNC, C.byref(intr),
                C.byref(cs))
        kfx.call("kfx_raycast", dv.ptr, vm.ptr, nm.ptr, *args)
        full_v = vm.download(np.uint32, n * 3)
        full_n = nm.download(np.uint32, n * 3)
        assert (~np.isnan(full_v.view(np.float32))).mean() > 0.3
        for bricks in (db, ones):
            kfx.call("kfx_raycast_bricks", dv.ptr, vb.ptr, nb.ptr, *args, bricks.ptr)
            assert np.array_equal(vb.download(np.uint32, n * 3), full_v)
            assert np.array_equal(nb.download(np.uint32, n * 3), full_n)


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


@pytest.mark.parametrize("seed", range(6))
def test_integrate_frustum_clip_random_poses(gpu, orc, seed):
    # Arbitrary camera orientations and positions (inside and outside the
    # cube) exercise every sign of the frustum-plane clip; the result must
    # equal the oracle's full sweep bit for bit.
    kfx = gpu
    rng = np.random.default_rng(100 + seed)
    intr = kfx.kinect_intrinsics(160, 120, 131.25)
    rb = kfx.round_buffer(96, 1.92, origin=tuple(int(v) for v in rng.integers(0, 96, 3)))
    R = _random_rotation(rng)
    t = rng.uniform(-0.5, 2.4, size=3)
    depth = rng.integers(300, 3000, size=(120, 160)).astype(np.uint16)
    depth[rng.random(depth.shape) < 0.05] = 0
    ref = scenes.empty_volume(rb)
    n_ref = orc.integrate_f32(ref, depth, intr, R.T, t, scenes.TRUNC, rb)
    dv = kfx.Buffer(rb.buffer_length * 4)
    kfx.call("kfx_init_tsdf", dv.ptr, C.byref(rb))
    ds = _integrate_gpu(kfx, dv, depth, intr, R.T, t, scenes.TRUNC, rb)
    got = dv.download(kfx.TSDF_DTYPE, rb.buffer_length)
    assert np.array_equal(got, ref)
    dd = kfx.Buffer.from_array(depth)
    n = C.c_uint64()
    cs = kfx.cell_size(rb)
    kfx.call("kfx_integrate_count_updates", dd.ptr, ds.ptr, C.byref(intr), C.byref(kfx.matrix3f(R.T)),
             C.byref(kfx.point3f(t)), scenes.TRUNC, C.byref(cs), C.byref(rb), C.byref(n))
    assert n.value == n_ref


def test_device_reciprocal_is_correctly_rounded(gpu):
    # The projections' division-free 1/z equals the IEEE quotient for every
    # float in [2^-126, 2^126) on this device (the range the kernels keep z
    # in); the top binade (subnormal results) is where it stops.
    kfx = gpu
    bad = C.c_uint64()
    kfx.call("kfx_check_reciprocal", -126, 125, C.byref(bad))
    assert bad.value == 0
    kfx.call("kfx_check_reciprocal", 126, 126, C.byref(bad))
    assert bad.value > 0


def test_integrate_clamps_out_of_range_weights(gpu, orc):
    # A volume not written by the update rule (random values, weights from
    # -300 to 32767): the kernel treats each weight outside [0, 128] as the
    # nearest bound, as the oracle does, so the result is still bit-exact.
    kfx = gpu
    rng = np.random.default_rng(7)
    intr = kfx.kinect_intrinsics(160, 120, 131.25)
    rb = kfx.round_buffer(96, 1.92, origin=(5, 17, 60))
    ref = scenes.empty_volume(rb)
    ref["value"] = rng.integers(-32767, 32768, rb.buffer_length).astype(np.int16)
    ref["weight"] = rng.integers(-300, 32768, rb.buffer_length).astype(np.int16)
    R = np.eye(3)
    t = np.array([0.96, 0.96, -0.2])
    depth = rng.integers(600, 2500, size=(120, 160)).astype(np.uint16)
    dv = kfx.Buffer.from_array(ref)
    n = orc.integrate_f32(ref, depth, intr, R.T, t, scenes.TRUNC, rb)
    _integrate_gpu(kfx, dv, depth, intr, R.T, t, scenes.TRUNC, rb)
    got = dv.download(kfx.TSDF_DTYPE, rb.buffer_length)
    assert n > 100000
    assert np.array_equal(got, ref)
