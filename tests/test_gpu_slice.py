"""GPU parity of the cyclic-shift kernels against the oracle: black-zone
clean and slice extraction (points, colours and order) bit for bit, with and
without a colour volume, including the chunked z order; plus a 512^3
full-size property check (extracted points are exactly the zero crossings
inside the zone: count equals the oracle's)."""
import ctypes as C

import numpy as np
import pytest

import scenes
import slice_inputs as si

pytestmark = pytest.mark.gpu

CASES = [
    ((64, 48, 80), (0, 0, 0), (0, 1000, 1000), (9, -1, -1)),
    ((64, 48, 80), (50, 7, 33), (60, 1000, 1000), (70, -1, -1)),
    ((64, 48, 80), (3, 40, 70), (1000, 47, 100), (-1, 52, 110)),
    ((32, 32, 32), (5, 6, 7), (0, 0, 0), (100, 100, 100)),
]


def _gpu_extract(kfx, dv, dc, rb, mn, mx):
    s = C.c_void_p()
    kfx.call("kfx_slicer_create", C.byref(rb), C.byref(s))
    n = C.c_uint64()
    rc = kfx.lib.kfx_extract_slice(s, C.byref(kfx.point3i(mn)), C.byref(kfx.point3i(mx)), C.byref(rb), dv.ptr,
                                   dc.ptr if dc is not None else None, None, None, 0, C.byref(n))
    assert rc in (0, 3)
    cap = max(1, n.value)
    dp, dcol = kfx.Buffer(16 * cap), kfx.Buffer(4 * cap)
    kfx.call("kfx_extract_slice", s, C.byref(kfx.point3i(mn)), C.byref(kfx.point3i(mx)), C.byref(rb), dv.ptr,
             dc.ptr if dc is not None else None, dp.ptr, dcol.ptr, cap, C.byref(n))
    kfx.synchronize()
    pts = dp.download(np.float32, (n.value, 4))
    cols = dcol.download(np.uint8, (n.value, 4))
    kfx.call("kfx_slicer_destroy", s)
    return pts, cols


@pytest.mark.parametrize("n,origin,mn,mx", CASES)
def test_extract_matches_oracle(gpu, orc, n, origin, mn, mx):
    kfx = gpu
    rb, vol, col = si.random_volume(n, origin, seed=n[0] + origin[1])
    dv = kfx.Buffer.from_array(vol)
    dc = kfx.Buffer.from_array(col)
    for c, d in ((col, dc), (None, None)):
        pts, cols = _gpu_extract(kfx, dv, d, rb, mn, mx)
        wp, wc = orc.extract_slice(vol, c, mn, mx, rb)
        assert len(wp) > 0 or mn[0] > mx[0]
        assert np.array_equal(pts, wp)
        assert np.array_equal(cols, wc)


@pytest.mark.parametrize("n,origin,mn,mx", CASES)
def test_extract_records_match_points(gpu, orc, n, origin, mn, mx):
    # (x, y, z, tsdf, weight) records: the oracle's points in the same order,
    # shifted by the cube's global voxel offset, with the stored value and
    # weight of each point's voxel.
    kfx = gpu
    rb, vol, col = si.random_volume(n, origin, seed=11 + n[1])
    rb.origin_GRID_global = kfx.Point3f(1000.0, -7.0, 33.0)
    dv, dc = kfx.Buffer.from_array(vol), kfx.Buffer.from_array(col)
    wp, wc = orc.extract_slice(vol, col, mn, mx, rb)
    s = C.c_void_p()
    kfx.call("kfx_slicer_create", C.byref(rb), C.byref(s))
    cap = max(1, len(wp))
    drec, dcol = kfx.Buffer(16 * cap), kfx.Buffer(4 * cap)
    cnt = C.c_uint64()
    kfx.call("kfx_extract_slice_records", s, C.byref(kfx.point3i(mn)), C.byref(kfx.point3i(mx)), C.byref(rb), dv.ptr,
             dc.ptr, drec.ptr, dcol.ptr, cap, C.byref(cnt))
    kfx.call("kfx_slicer_destroy", s)
    assert cnt.value == len(wp)
    if not len(wp):
        return
    rec = drec.download(kfx.SLICE_RECORD_DTYPE, len(wp))
    assert np.array_equal(dcol.download(np.uint8, (len(wp), 4)), wc)
    local = wp[:, :3].astype(np.int64)
    assert np.array_equal(rec["x"] - 1000, local[:, 0])
    assert np.array_equal(rec["y"] + 7, local[:, 1])
    assert np.array_equal(rec["z"] - 33, local[:, 2])
    addr = [orc.lib.orc_cyclic_address(int(x), int(y), int(z), C.byref(rb)) for x, y, z in local]
    assert np.array_equal(rec["tsdf"], vol["value"][addr])
    assert np.array_equal(rec["weight"], vol["weight"][addr])


@pytest.mark.parametrize("n,origin,mn,mx", CASES)
def test_clean_matches_oracle(gpu, orc, n, origin, mn, mx):
    kfx = gpu
    rb, vol, col = si.random_volume(n, origin, seed=7)
    dv = kfx.Buffer.from_array(vol)
    dc = kfx.Buffer.from_array(col)
    kfx.call("kfx_clean_tsdf_slice", C.byref(kfx.point3i(mn)), C.byref(kfx.point3i(mx)), C.byref(rb), dv.ptr)
    kfx.call("kfx_clean_color_slice", C.byref(kfx.point3i(mn)), C.byref(kfx.point3i(mx)), C.byref(rb), dc.ptr)
    orc.clean_slice(vol, mn, mx, rb)
    orc.clean_slice(col, mn, mx, rb)
    assert np.array_equal(dv.download(kfx.TSDF_DTYPE, rb.buffer_length), vol)
    assert np.array_equal(dc.download(orc.COLOR_DTYPE, rb.buffer_length), col)


def test_extract_after_integration_512(gpu, orc):
    # Full size: a 512^3 volume integrated from C1-style frames; shift by 100
    # voxels in x.  Count and first/last points agree with the oracle.
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    rb = kfx.round_buffer(512, 3.0, origin=(300, 17, 411))
    scene, poses = scenes.c1_scene(), scenes.c1_poses(30)
    vol = scenes.empty_volume(rb)
    for f in (0, 10, 20):
        R, t = poses[f].matrices()
        orc.integrate_f32(vol, kfx.render(scene, poses[f], intr, 1, f, noise=True), intr, R.T, t, 0.03, rb)
    mn, mx = (300, 10 ** 6, 10 ** 6), (399, -1, -1)
    dv = kfx.Buffer.from_array(vol)
    pts, cols = _gpu_extract(kfx, dv, None, rb, mn, mx)
    wp, wc = orc.extract_slice(vol, None, mn, mx, rb)
    assert len(wp) > 1000
    assert np.array_equal(pts, wp)
