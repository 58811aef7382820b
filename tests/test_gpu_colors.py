"""GPU parity of the colour volume (Integrate_Colors,
common/src/opencl_wrappers.cpp:181-214) against the oracle: bit-exact colours
and weights on the C1 scene with its raycast vertex maps, for the reference's
default weight cap (1) and the 8-bit maximum, with a cyclic origin, through
the host-pose and the device-pose entry points; the colour pipeline end to
end (colours of the reconstructed surface, colour planes of the extracted
slices)."""
import ctypes as C

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


def _model(kfx, orc, rb, frames):
    """TSDF of the C1 scene integrated (oracle) from `frames`."""
    intr = kfx.kinect_intrinsics()
    scene, poses = scenes.c1_scene(), scenes.c1_poses(30)
    vol = scenes.empty_volume(rb)
    for f in frames:
        R, t = poses[f].matrices()
        orc.integrate_f32(vol, kfx.render(scene, poses[f], intr, scenes.C1_SEED, f, noise=True), intr, R.T, t,
                          scenes.TRUNC, rb)
    return intr, scene, poses, vol


def _raycast(kfx, dv, rb, intr, R, t):
    dvm = kfx.Buffer(intr.cols * intr.rows * 12)
    dnm = kfx.Buffer(intr.cols * intr.rows * 12)
    kfx.call("kfx_raycast", dv.ptr, dvm.ptr, dnm.ptr, C.byref(rb), C.byref(kfx.point3f(t)),
             C.byref(kfx.matrix3f(R)), scenes.TRUNC, C.byref(intr), C.byref(kfx.cell_size(rb)))
    return dvm


def _random_colors(orc, rb, max_weight, seed):
    rng = np.random.default_rng(seed)
    c = np.zeros(rb.buffer_length, orc.COLOR_DTYPE)
    for ch in "rgb":
        c[ch] = rng.integers(0, 256, c.size)
    c["w"] = rng.integers(0, max_weight + 1, c.size)
    return c


@pytest.mark.parametrize("origin,max_weight", [((0, 0, 0), 1), ((17, 101, 250), 255), ((200, 3, 90), 7)])
def test_integrate_colors_matches_oracle_c1(gpu, orc, origin, max_weight):
    kfx = gpu
    rb = kfx.round_buffer(256, 3.0, origin=origin)
    intr, scene, poses, vol = _model(kfx, orc, rb, range(0, 30, 3))
    dv = kfx.Buffer.from_array(vol)
    ref = _random_colors(orc, rb, max_weight, seed=origin[0] + max_weight)
    dc = kfx.Buffer.from_array(ref)
    total = 0
    for f in (4, 13, 22):
        R, t = poses[f].matrices()
        dvm = _raycast(kfx, dv, rb, intr, R, t)
        vm = dvm.download(np.float32, (intr.rows, intr.cols, 3))
        rgb = kfx.render_rgb(scene, poses[f], intr)
        drgb = kfx.Buffer.from_array(rgb)
        kfx.call("kfx_integrate_colors", dc.ptr, drgb.ptr, dvm.ptr, C.byref(intr), C.byref(kfx.matrix3f(R.T)),
                 C.byref(kfx.point3f(t)), scenes.TRUNC, max_weight, C.byref(kfx.cell_size(rb)), C.byref(rb))
        total += orc.integrate_colors(ref, rgb, vm, intr, R.T, t, scenes.TRUNC, max_weight, rb)
    got = dc.download(orc.COLOR_DTYPE, rb.buffer_length)
    print(f"voxels touched over 3 frames: {total}")
    assert total > 100000
    diff = np.count_nonzero(got.view(np.uint32) != ref.view(np.uint32))
    assert diff == 0, f"{diff} voxels differ"


def test_integrate_colors_posed_and_skip(gpu, orc):
    kfx = gpu
    rb = kfx.round_buffer(256, 3.0, origin=(40, 0, 130))
    intr, scene, poses, vol = _model(kfx, orc, rb, range(0, 30, 5))
    dv = kfx.Buffer.from_array(vol)
    R, t = poses[11].matrices()
    dvm = _raycast(kfx, dv, rb, intr, R, t)
    drgb = kfx.Buffer.from_array(kfx.render_rgb(scene, poses[11], intr))
    start = _random_colors(orc, rb, 3, seed=9)
    host, posed = kfx.Buffer.from_array(start), kfx.Buffer.from_array(start)
    cs = kfx.cell_size(rb)
    kfx.call("kfx_integrate_colors", host.ptr, drgb.ptr, dvm.ptr, C.byref(intr), C.byref(kfx.matrix3f(R.T)),
             C.byref(kfx.point3f(t)), scenes.TRUNC, 3, C.byref(cs), C.byref(rb))
    dpose = kfx.Buffer.from_array(np.concatenate([R.reshape(9), t]).astype(np.float32))
    skip = kfx.Buffer.from_array(np.array([1], np.int32))
    kfx.call("kfx_integrate_colors_posed", posed.ptr, drgb.ptr, dvm.ptr, C.byref(intr), dpose.ptr, scenes.TRUNC, 3,
             C.byref(cs), C.byref(rb), skip.ptr)
    assert np.array_equal(posed.download(orc.COLOR_DTYPE, rb.buffer_length), start)
    kfx.call("kfx_integrate_colors_posed", posed.ptr, drgb.ptr, dvm.ptr, C.byref(intr), dpose.ptr, scenes.TRUNC, 3,
             C.byref(cs), C.byref(rb), None)
    a = host.download(orc.COLOR_DTYPE, rb.buffer_length)
    b = posed.download(orc.COLOR_DTYPE, rb.buffer_length)
    assert np.count_nonzero(a.view(np.uint32) != start.view(np.uint32)) > 10000
    # The host path takes R^-1 and t in float32 from double; the device path
    # transposes the float32 R itself: the same values.
    assert np.array_equal(a, b)


def test_integrate_colors_rejects_bad_weight(gpu, orc):
    kfx = gpu
    rb = kfx.round_buffer(32, 1.0)
    intr = kfx.kinect_intrinsics(64, 48, 52.5)
    dc, drgb, dvm = kfx.Buffer(rb.buffer_length * 4), kfx.Buffer(64 * 48 * 4), kfx.Buffer(64 * 48 * 12)
    for w in (0, 256):
        rc = kfx.lib.kfx_integrate_colors(dc.ptr, drgb.ptr, dvm.ptr, C.byref(intr), C.byref(kfx.matrix3f(np.eye(3))),
                                          C.byref(kfx.point3f((0, 0, 0))), 0.03, w, C.byref(kfx.cell_size(rb)),
                                          C.byref(rb))
        assert rc != 0 and b"max_weight" in kfx.lib.kfx_last_error()


def _palette(kfx, scene, intr, poses):
    cols = set()
    for f in range(0, len(poses), 4):
        rgb = kfx.render_rgb(scene, poses[f], intr)
        cols |= set(map(tuple, np.stack([rgb["r"], rgb["g"], rgb["b"]], -1).reshape(-1, 3)))
    return cols


@pytest.mark.parametrize("mode", ["plain", "fused"])
def test_colour_pipeline_paints_the_surface(gpu, orc, mode):
    # C1 orbit with colour frames, weight cap 1 (the reference's default):
    # the voxels the colour volume set carry the scene's albedos (with the
    # true poses the oracle pipeline sets about 75k voxels over these 20
    # frames, 96.6 % of them exactly one of the 12 albedos; the rest mix two
    # along object edges), and the slice extraction returns the colour
    # planes with the points.
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    scene, poses = scenes.c1_scene(), scenes.c1_poses(30)
    cfg = kfx.default_config(voxels=kfx.Point3i(256, 256, 256), volume_size=kfx.Point3f(3.0, 3.0, 3.0),
                             mode=kfx.FUSED if mode == "fused" else kfx.PLAIN, use_graph=1 if mode == "fused" else 0,
                             colors=1, shift_threshold=100.0)
    pipe = kfx.Pipeline(cfg, poses[0], (0.0, 0.0, 0.0))
    dd = kfx.Buffer(intr.cols * intr.rows * 2)
    dc = kfx.Buffer(intr.cols * intr.rows * 4)
    for f in range(20):
        depth = kfx.render(scene, poses[f], intr, scenes.C1_SEED, f, noise=True)
        kfx.call("kfx_memcpy_htod", dd.ptr, depth.ctypes.data_as(C.c_void_p), depth.nbytes)
        rgb = kfx.render_rgb(scene, poses[f], intr)
        kfx.call("kfx_memcpy_htod", dc.ptr, rgb.ctypes.data_as(C.c_void_p), rgb.nbytes)
        info = pipe.process_rgbd(dd.ptr, dc.ptr)
    v = pipe.view()
    rb = v.round_buffer
    cvol = kfx.download(v.color_volume, orc.COLOR_DTYPE, rb.buffer_length)
    set_ = cvol["w"] > 0
    assert set_.sum() > 50000
    assert np.all(cvol["w"][set_] == 1)
    pal = _palette(kfx, scene, intr, poses[:20])
    rgbv = np.stack([cvol["r"], cvol["g"], cvol["b"]], -1)[set_]
    in_pal = np.array([tuple(c) in pal for c in rgbv[:: max(1, len(rgbv) // 20000)]])
    assert in_pal.mean() > 0.93, in_pal.mean()
    # Extraction with colours from the pipeline's volumes.
    slicer = C.c_void_p()
    kfx.call("kfx_slicer_create", C.byref(rb), C.byref(slicer))
    pts, cols = kfx.Buffer(16 * (1 << 22)), kfx.Buffer(4 * (1 << 22))
    n = C.c_uint64()
    lo, hi = kfx.point3i((100, 1, 1)), kfx.point3i((163, 0, 0))  # x in [100, 163]; y, z ranges empty
    kfx.call("kfx_extract_slice", slicer, C.byref(lo), C.byref(hi), C.byref(rb), v.volume, v.color_volume, pts.ptr,
             cols.ptr, 1 << 22, C.byref(n))
    kfx.call("kfx_slicer_destroy", slicer)
    got_c = cols.download(orc.COLOR_DTYPE, n.value)
    assert n.value > 1000 and np.count_nonzero(got_c["w"]) > 1000
    pipe.close()
    assert info.icp_iterations == 19
