"""The larger configurations on the GPU, each checked against the oracle:

- the 512^3 walk (configuration 2) through the headline pipeline (learned
  denoiser, graph replay) over a window that crosses its first cyclic
  shift: the volume after the shift frame equals the oracle's integrate,
  extract and clean of the frame before it, bit for bit, the extracted
  points equal the oracle's extraction, and the raycast of the frame before
  equals the oracle's;
- the 1024^3 / 1280x720 case (configuration 5): integrate (64-voxel z runs)
  and raycast against the oracle;
- the precision study (configuration 3): the three ICP sum modes over the
  512^3 orbit, the compensated FP32 sums agreeing with the FP64 sums and
  the plain FP32 sums drifting away from them.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "3deye_amd"))
import workload as wl  # noqa: E402
import scenes  # noqa: E402

pytestmark = pytest.mark.gpu


def _slab(rb_pre, rb_post):
    """The black zone maybe_shift extracted and cleared, from the origin
    change (s = the whole-voxel shift per axis)."""
    mn, mx = [], []
    n = (rb_pre.voxels_size.x, rb_pre.voxels_size.y, rb_pre.voxels_size.z)
    o = (rb_pre.origin_GRID.x, rb_pre.origin_GRID.y, rb_pre.origin_GRID.z)
    g0 = (rb_pre.origin_GRID_global.x, rb_pre.origin_GRID_global.y, rb_pre.origin_GRID_global.z)
    g1 = (rb_post.origin_GRID_global.x, rb_post.origin_GRID_global.y, rb_post.origin_GRID_global.z)
    for k in range(3):
        s = int(round(g1[k] - g0[k]))
        if s > 0:
            mn.append(o[k]); mx.append(o[k] + s - 1)
        elif s < 0:
            mn.append(o[k] + n[k] + s); mx.append(o[k] + n[k] - 1)
        else:
            mn.append(1); mx.append(0)
    return tuple(mn), tuple(mx)


def _copy_rb(kfx, rb):
    out = kfx.RoundBuffer()
    C.memmove(C.byref(out), C.byref(rb), C.sizeof(rb))
    return out


def test_c2_walk_window_across_a_shift(gpu, orc):
    # The walk's first shift comes at frame 77 (the target point has moved
    # 1.5 m from the cube centre); its leaving slab holds what the camera
    # saw over the first frames.
    kfx = gpu
    frames, watch_from = 90, 68
    intr, scene, poses, origin, depths = wl.sequence(wl.SCENE_SEED, frames)
    dd = kfx.Buffer.from_array(depths)
    fb = depths[0].nbytes
    dn = kfx.Denoiser(wl.W, wl.H, kfx.Denoiser.load_params())
    cfg = wl.pipeline_config(mode=kfx.FUSED, use_graph=1, denoiser=1)
    pipe = kfx.Pipeline(cfg, poses[0], origin)
    pipe.set_denoiser(dn)
    n = wl.N_VOX ** 3
    prev_vol, prev_rb, prev_info, prev_vmap = None, None, None, None
    infos = []
    k = None
    for i in range(frames):
        info = pipe.process(C.c_void_p(dd.ptr.value + i * fb))
        infos.append(info)
        v = pipe.view()
        if info.shifted:
            k = i
            break
        if i + 1 < watch_from:
            continue
        prev_vol = kfx.download(v.volume, kfx.TSDF_DTYPE, n)
        prev_vmap = kfx.download(v.vmap_model, np.float32, (wl.H, wl.W, 3))
        prev_rb = _copy_rb(kfx, v.round_buffer)
        prev_info = info
        # The integrations keep every brick holding a negative value flagged
        # (the skipping raycast's condition for the full march's result).
        flags = kfx.download(v.brick_flags, np.uint8, (wl.N_VOX // 16) ** 3)
        exact = scenes.brick_flags_cpu(prev_vol, prev_rb)
        assert not np.any((exact != 0) & (flags == 0))
    print(f"first shift at frame {k}")
    assert k is not None and k >= watch_from, "the walk's first shift must fall inside the watched window"
    assert all(i.fallback == 0 for i in infos[1:])

    # The raycast of frame k-1 (the model the shift frame was tracked against).
    R, t = prev_info.pose_volume.matrices()
    rv, _ = orc.raycast_f32(prev_vol, prev_rb, t, R, wl.TRUNC, intr)
    assert np.array_equal(np.isnan(prev_vmap), np.isnan(rv))
    hit = ~np.isnan(rv[..., 0])
    assert hit.mean() > 0.3
    assert np.max(np.abs(prev_vmap[hit] - rv[hit])) <= 1e-4

    # Frame k: integrate into the frame k-1 volume at the pose ICP found,
    # then the slab that left the cube is extracted and cleared.
    info = infos[k]
    v = pipe.view()
    R, t = info.pose_volume.matrices()
    ref = prev_vol
    if info.integrated:
        orc.integrate_f32(ref, depths[k], intr, R.T, t, wl.TRUNC, prev_rb)
    mn, mx = _slab(prev_rb, v.round_buffer)
    want_pts, _ = orc.extract_slice(ref, None, mn, mx, prev_rb)
    assert info.extracted_points == len(want_pts) > 1000
    got_pts = kfx.download(v.last_points, np.float32, (info.extracted_points, 4))
    assert np.array_equal(got_pts, want_pts)
    orc.clean_slice(ref, mn, mx, prev_rb)
    got = kfx.download(v.volume, kfx.TSDF_DTYPE, n)
    same = got.view(np.uint32) == ref.view(np.uint32)
    assert same.all(), f"{np.count_nonzero(~same)} voxels differ after the shift"
    # The shift rebuilt the brick flags of the moved volume exactly.
    flags = kfx.download(v.brick_flags, np.uint8, (wl.N_VOX // 16) ** 3)
    assert np.array_equal(flags != 0, scenes.brick_flags_cpu(got, _copy_rb(kfx, v.round_buffer)) != 0)
    assert v.shifts == 1
    ate = kfx.ate_rmse([i.pose_world for i in infos], poses[:k + 1])
    print(f"{len(want_pts)} points extracted at frame {k}; ATE {ate * 1000:.2f} mm over {k + 1} frames")
    assert ate < 0.005
    pipe.close()
    dn.close()


def test_c5_1024_integrate_and_raycast_match_oracle(gpu, orc):
    # 1024^3 (4 mm voxels, 4.096 m cube, 4 GiB), 1280x720 depth (fx = fy =
    # 900): three frames of the walk integrated at their true poses, then a
    # raycast, all against the oracle.
    kfx = gpu
    nvox, vol_m = 1024, 4.096
    intr = kfx.Intrinsics(900.0, 900.0, 639.5, 359.5, 1280, 720)
    scene = wl.scene(wl.SCENE_SEED)
    poses = wl.trajectory(wl.SCENE_SEED, 11)
    origin = wl.cube_origin(poses[0], vol_m=vol_m)
    rb = kfx.round_buffer(nvox, vol_m, origin=(37, 900, 511))
    cs = kfx.cell_size(rb)
    ref = np.zeros(rb.buffer_length, dtype=kfx.TSDF_DTYPE)
    dv = kfx.Buffer(rb.buffer_length * 4)
    kfx.call("kfx_init_tsdf", dv.ptr, C.byref(rb))
    ds = kfx.Buffer(intr.cols * intr.rows * 4)
    total = 0
    for f in (0, 5, 10):
        R, t = poses[f].matrices()
        tv = t - origin
        depth = kfx.render(scene, poses[f], intr, wl.SCENE_SEED, f, noise=True)
        total += orc.integrate_f32(ref, depth, intr, R.T, tv, wl.TRUNC, rb)
        dd = kfx.Buffer.from_array(depth)
        kfx.call("kfx_integrate_tsdf", dd.ptr, ds.ptr, dv.ptr, C.byref(intr), C.byref(kfx.matrix3f(R.T)),
                 C.byref(kfx.point3f(tv)), wl.TRUNC, C.byref(cs), C.byref(rb))
    kfx.synchronize()
    got = dv.download(kfx.TSDF_DTYPE, rb.buffer_length)
    same = got.view(np.uint32) == ref.view(np.uint32)
    print(f"1024^3: N_upd over 3 frames {total}, {np.count_nonzero(~same)} voxels differ")
    assert total > 10_000_000
    assert same.all()
    del got, same
    R, t = poses[10].matrices()
    tv = t - origin
    dvm = kfx.Buffer(intr.cols * intr.rows * 12)
    dnm = kfx.Buffer(intr.cols * intr.rows * 12)
    kfx.call("kfx_raycast", dv.ptr, dvm.ptr, dnm.ptr, C.byref(rb), C.byref(kfx.point3f(tv)),
             C.byref(kfx.matrix3f(R)), wl.TRUNC, C.byref(intr), C.byref(cs))
    vm = dvm.download(np.float32, (intr.rows, intr.cols, 3))
    nm = dnm.download(np.float32, (intr.rows, intr.cols, 3))
    rv, rn = orc.raycast_f32(ref, rb, tv, R, wl.TRUNC, intr)
    assert np.array_equal(np.isnan(vm), np.isnan(rv))
    assert np.array_equal(np.isnan(nm), np.isnan(rn))
    hit = ~np.isnan(rv[..., 0])
    assert hit.mean() > 0.3
    assert np.max(np.abs(vm[hit] - rv[hit])) <= 1e-4


def _rot_angle(Ra, Rb):
    return float(2.0 * np.arcsin(min(1.0, np.linalg.norm(Ra - Rb) / (2.0 * np.sqrt(2.0)))))


def test_c3_sum_modes(gpu):
    # 120 frames of the precision-study orbit (512^3, no shifts) in each ICP
    # sum mode.  Compensated FP32 workgroup sums track the FP64 sums to a few
    # ulps of a double per frame and give the identical volume; plain FP32
    # sums drift from them by more than 1e-9 while staying within 1e-4.
    kfx = gpu
    import precision_study as ps
    frames = 120
    intr, poses, origin, depths = ps.orbit_sequence(wl.SCENE_SEED, frames)
    dd = kfx.Buffer.from_array(depths)
    fb = depths[0].nbytes
    res = {name: ps.run(m, poses, origin, dd, fb, frames) for name, m in ps.MODES.items()}
    ref_p, ref_v = res["fp64"]

    def diff(name):
        p = res[name][0]
        dt = max(float(np.linalg.norm(a[1] - b[1])) for a, b in zip(p, ref_p))
        dr = max(_rot_angle(a[0], b[0]) for a, b in zip(p, ref_p))
        return dt, dr

    kt, kr = diff("fp32_kahan")
    pt, pr = diff("fp32_plain")
    print(f"kahan vs fp64: {kt:.2e} m {kr:.2e} rad; plain vs fp64: {pt:.2e} m {pr:.2e} rad")
    assert kt <= 1e-12 and kr <= 1e-12
    assert np.array_equal(res["fp32_kahan"][1], ref_v)
    assert max(pt, pr) > 1e-9
    assert pt < 1e-4 and pr < 1e-4
    ate = kfx.ate_rmse([kfx.pose_from(*p) for p in ref_p], poses)
    assert ate < 0.01
