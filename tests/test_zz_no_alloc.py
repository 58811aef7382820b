"""kfx_integrate_tsdf_noalloc against kfx_integrate_tsdf through the C ABI:
two calls in a row on the same volume pair, the second with a smaller image,
bit-identical volumes and scaled depth; an image beyond the fixed scratch
size is rejected."""
import ctypes as C

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


def _call(kfx, name, dd, ds, dv, intr, R, t, rb):
    cs = kfx.cell_size(rb)
    kfx.call(name, dd.ptr, ds.ptr, dv.ptr, C.byref(intr), C.byref(kfx.matrix3f(R.T)), C.byref(kfx.point3f(t)),
             scenes.TRUNC, C.byref(cs), C.byref(rb))


def test_noalloc_matches_integrate_twice(gpu):
    kfx = gpu
    rb = kfx.round_buffer(256, 3.0, origin=(11, 200, 37))
    scene, poses = scenes.c1_scene(), scenes.c1_poses(30)
    va, vb = kfx.Buffer(rb.buffer_length * 4), kfx.Buffer(rb.buffer_length * 4)
    kfx.call("kfx_init_tsdf", va.ptr, C.byref(rb))
    kfx.call("kfx_init_tsdf", vb.ptr, C.byref(rb))
    # 640x480 first, then 320x240 (a smaller pyramid in the same scratch).
    for f, intr in ((3, kfx.kinect_intrinsics()), (9, kfx.kinect_intrinsics(320, 240, 262.5))):
        R, t = poses[f].matrices()
        depth = kfx.render(scene, poses[f], intr, scenes.C1_SEED, f, noise=True)
        dd = kfx.Buffer.from_array(depth)
        sa, sb = kfx.Buffer(depth.size * 4), kfx.Buffer(depth.size * 4)
        _call(kfx, "kfx_integrate_tsdf", dd, sa, va, intr, R, t, rb)
        _call(kfx, "kfx_integrate_tsdf_noalloc", dd, sb, vb, intr, R, t, rb)
        kfx.synchronize()
        assert np.array_equal(sa.download(np.float32, depth.shape), sb.download(np.float32, depth.shape))
        a = va.download(np.uint32, rb.buffer_length)
        b = vb.download(np.uint32, rb.buffer_length)
        assert np.count_nonzero(a) > 100000
        assert np.array_equal(a, b), f"frame {f}: {np.count_nonzero(a != b)} voxels differ"


def test_noalloc_rejects_images_beyond_its_scratch(gpu):
    kfx = gpu
    rb = kfx.round_buffer(32, 1.0)
    intr = kfx.kinect_intrinsics(4096, 3072, 3000.0)
    dv = kfx.Buffer(rb.buffer_length * 4)
    dd, ds = kfx.Buffer(16), kfx.Buffer(16)
    cs = kfx.cell_size(rb)
    rc = kfx.lib.kfx_integrate_tsdf_noalloc(dd.ptr, ds.ptr, dv.ptr, C.byref(intr), C.byref(kfx.matrix3f(np.eye(3))),
                                            C.byref(kfx.point3f((0, 0, 0))), 0.03, C.byref(cs), C.byref(rb))
    assert rc != 0 and b"scratch" in kfx.lib.kfx_last_error()
