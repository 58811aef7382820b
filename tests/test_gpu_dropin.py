"""A C++ program written only against the reference host API
(opencl_wrappers.h, gpu_array<>, kernel_manager, round_buffer, matrix<>)
runs integrate, raycast, Integrate_Colors, resize, bilateral, pyramid,
vertex/normal maps, Estimate_Combined, Transform_Maps, Extract_Slice (with
the colour volume), Clean_Tsdf_Slice and Clean_Color_Slice through
libkfx_host.so; every dumped result is compared with the oracle."""
import os
import subprocess

import numpy as np
import pytest

import icp_inputs as ii
import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "3deye_amd", "dropin_check")


def _load(d, name, dtype, shape=None):
    a = np.fromfile(os.path.join(d, name), dtype=dtype)
    return a.reshape(shape) if shape is not None else a


def test_reference_api_program(gpu, orc, tmp_path):
    kfx = gpu
    assert os.path.exists(BIN), "dropin_check not built (make -C 3deye_amd)"
    W, H, N = 320, 240, 128
    intr = kfx.kinect_intrinsics(W, H, 262.5)
    scene, poses = scenes.c1_scene(), scenes.c1_poses(30)
    R0, t0 = poses[10].matrices()
    d0 = kfx.render(scene, poses[10], intr, 1, 10, noise=True)
    d1 = kfx.render(scene, poses[11], intr, 1, 11, noise=True)
    rgb0 = kfx.render_rgb(scene, poses[10], intr)
    with open(tmp_path / "in.bin", "wb") as f:
        np.array([W, H, N], np.int32).tofile(f)
        np.array([intr.fx, intr.fy, intr.cx, intr.cy, 3.0], np.float32).tofile(f)
        np.concatenate([R0.reshape(9), t0]).astype(np.float32).tofile(f)
        d0.tofile(f)
        d1.tofile(f)
        rgb0.tofile(f)
    r = subprocess.run([BIN, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    d = str(tmp_path)
    R0f = R0.astype(np.float32).astype(np.float64)
    t0f = t0.astype(np.float32).astype(np.float64)
    rb = kfx.round_buffer(N, 3.0)
    # frame 0: integrate + raycast
    vol = scenes.empty_volume(rb)
    orc.integrate_f32(vol, d0, intr, R0f.T, t0f, 0.03, rb)
    assert np.array_equal(_load(d, "volume0.bin", kfx.TSDF_DTYPE), vol)
    rv, rn = orc.raycast_f32(vol, rb, t0f, R0f, 0.03, intr)
    gv = _load(d, "vmap0.bin", np.float32, (H, W, 3))
    assert np.array_equal(np.isnan(gv), np.isnan(rv))
    ok = ~np.isnan(rv)
    assert np.array_equal(gv[ok], rv[ok])
    # frame 0 colours at that raycast (Integrate_Colors, weight cap 1)
    col = np.zeros(rb.buffer_length, orc.COLOR_DTYPE)
    assert orc.integrate_colors(col, rgb0, rv, intr, R0f.T, t0f, 0.03, 1, rb) > 1000
    assert np.array_equal(_load(d, "colors0.bin", orc.COLOR_DTYPE), col)
    # frame 1: bilateral (within 1 mm), pyramid and vertex maps exact on the GPU's level 0
    g0 = _load(d, "depth1_l0.bin", np.uint16, (H, W))
    assert np.max(np.abs(g0.astype(int) - orc.bilateral(d1))) <= 1
    lv = [g0]
    for l in (1, 2):
        lv.append(orc.pyr_down(lv[-1]))
        assert np.array_equal(_load(d, f"depth1_l{l}.bin", np.uint16, lv[l].shape), lv[l])
    vcs = []
    for l in range(3):
        want = orc.vertex_map(lv[l], intr.level(l))
        got = _load(d, f"vmap1_l{l}.bin", np.float32, want.shape)
        ok = ~np.isnan(want)
        assert np.array_equal(np.isnan(got), ~ok) and np.array_equal(got[ok], want[ok])
        vcs.append(got)
    nc0 = _load(d, "nmap1_l0.bin", np.float32, (H, W, 3))
    # Estimate_Combined: the reference reduction, bit for bit, on the dumped maps
    rn0 = _load(d, "nmap0.bin", np.float32, (H, W, 3))
    m = orc.estimate_combined_f32(vcs[0], nc0, gv, rn0, R0f, t0f, R0f.T, t0f, intr, ii.DIST_THR, ii.ANGLE_THR)
    A, b = orc.unpack_sums(m.astype(np.float64))
    assert np.array_equal(_load(d, "A.bin", np.float64, (6, 6)), A)
    assert np.array_equal(_load(d, "b.bin", np.float64), b)
    # Transform_Maps back into the camera
    Ri = R0f.T.astype(np.float32)
    ti = -(Ri.astype(np.float64) @ t0f.astype(np.float32).astype(np.float64))
    wv, _ = orc.transform_maps(gv, rn0, Ri, ti.astype(np.float32))
    gt = _load(d, "vmap0_cam.bin", np.float32, (H, W, 3))
    ok = ~np.isnan(wv)
    assert np.array_equal(np.isnan(gt), ~ok)
    np.testing.assert_allclose(gt[ok], wv[ok], atol=1e-6)
    # Extract_Slice + Clean_Tsdf_Slice of a 16-voxel x slab
    mn, mx = (0, 1 << 30, 1 << 30), (15, -1, -1)
    pts, pcol = orc.extract_slice(vol, col, mn, mx, rb)
    gp = np.fromfile(os.path.join(d, "slice_points.bin"), np.float32).reshape(-1, 4)
    assert len(pts) > 0 and np.array_equal(gp, pts)
    assert np.array_equal(_load(d, "slice_colors.bin", np.uint8, (-1, 4)), pcol)
    # matrix::Save_To_File of both slices: the text holds the same values.
    head, body = open(os.path.join(d, "slice_points.tst")).read().split("\n")[:2]
    assert head == f"{len(pts)} 1 1 Point_XYZI" or head.startswith(f"{len(pts)} 1 1 ")
    txt = np.array([[float(v) for v in e.split("x")] for e in body.split()], np.float32)
    assert np.array_equal(txt, pts)  # grid coordinates print exactly under %g
    chead, cbody = open(os.path.join(d, "slice_colors.tst")).read().split("\n")[:2]
    assert chead.startswith(f"{len(pcol)} 1 1 ")
    assert np.array_equal(np.array([[int(v) for v in e.split(".")] for e in cbody.split()], np.uint8), pcol)
    # Extract_Tsdf_Volume: the whole cube's crossings, metric voxel centres;
    # Extract_Color_Volume: the colour of the voxel under each point.
    allp, _ = orc.extract_slice(vol, None, (0, 1, 1), (N - 1, 0, 0), rb)
    cell = np.float32(3.0 / N)
    want = allp.copy()
    want[:, :3] = (allp[:, :3] + np.float32(0.5)) * cell
    cloud = _load(d, "volume_cloud.bin", np.float32, (-1, 4))
    assert len(cloud) > len(pts) and np.array_equal(cloud, want)
    g = np.floor(cloud[:, :3] / cell).astype(np.int64)
    addr = g[:, 0] + N * g[:, 1] + N * N * g[:, 2]  # origin 0
    assert np.array_equal(_load(d, "volume_cloud_colors.bin", orc.COLOR_DTYPE), col[addr])
    orc.clean_slice(vol, mn, mx, rb)
    orc.clean_slice(col, mn, mx, rb)
    assert np.array_equal(_load(d, "volume_cleaned.bin", kfx.TSDF_DTYPE), vol)
    assert np.array_equal(_load(d, "colors_cleaned.bin", orc.COLOR_DTYPE), col)
