"""End-to-end pipeline on the GPU: plain (reference control flow) and fused
modes track the C1 orbit, agree with each other, the graph replay equals the
eager fused path bit for bit, and the cyclic shift keeps tracking while
extracting the leaving slab."""
import ctypes as C

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


def _run(kfx, pipe, depths):
    dd = kfx.Buffer.from_array(depths)
    return [pipe.process(C.c_void_p(dd.ptr.value + i * depths[0].nbytes)) for i in range(len(depths))]


def _c1_run(kfx, mode, use_graph=0, frames=30, n=256, **kw):
    # The C1 orbit looks at the cube centre from 1.5 m: no shifts.
    intr = kfx.kinect_intrinsics()
    scene, poses = scenes.c1_scene(), scenes.c1_poses(frames)
    depths = np.stack([kfx.render(scene, poses[i], intr, scenes.C1_SEED, i, noise=True) for i in range(frames)])
    kw.setdefault("shift_threshold", 100.0)
    cfg = kfx.default_config(intr=intr, voxels=kfx.Point3i(n, n, n), volume_size=kfx.Point3f(3.0, 3.0, 3.0),
                             mode=mode, use_graph=use_graph, **kw)
    pipe = kfx.Pipeline(cfg, poses[0], (0.0, 0.0, 0.0))
    return pipe, _run(kfx, pipe, depths), poses


@pytest.mark.parametrize("mode", [0, 1])
def test_pipeline_tracks_c1(gpu, mode):
    kfx = gpu
    pipe, infos, gt = _c1_run(kfx, mode)
    est = [i.pose_world for i in infos]
    ate = kfx.ate_rmse(est, gt)
    print(f"mode {mode}: ATE {ate * 1000:.2f} mm over {len(gt)} frames")
    assert ate < 0.01
    assert all(i.icp_singular == 0 for i in infos[1:])
    assert all(i.icp_iterations == 19 for i in infos[1:])
    pipe.close()


def test_fused_matches_plain_and_graph_is_exact(gpu):
    kfx = gpu
    p0, i0, gt = _c1_run(kfx, 0, frames=12)
    p1, i1, _ = _c1_run(kfx, 1, frames=12)
    p2, i2, _ = _c1_run(kfx, 1, use_graph=1, frames=12)
    for a, b, c in zip(i0, i1, i2):
        ta, tb = np.array(a.pose_world.t[:]), np.array(b.pose_world.t[:])
        assert np.linalg.norm(ta - tb) < 5e-4  # FP32 host-unpacked sums vs compensated/FP64
        assert np.array_equal(np.array(b.pose_world.t[:]), np.array(c.pose_world.t[:]))
        assert np.array_equal(np.array(b.pose_world.R[:]), np.array(c.pose_world.R[:]))
    for p in (p0, p1, p2):
        p.close()


@pytest.mark.parametrize("early", [0.0, 1e-5])
def test_one_launch_frame_icp_is_exact(gpu, early):
    # The fused pipeline (graph replay) with each frame's ICP as one launch
    # and as two launches per iteration: identical poses and iteration
    # counts, with and without the early exit.
    kfx = gpu
    kw = dict(converge_rotation=early, converge_translation=early)
    p1, i1, _ = _c1_run(kfx, 1, use_graph=1, frames=12, icp_launch=kfx.ICP_LAUNCH_FRAME, **kw)
    p2, i2, _ = _c1_run(kfx, 1, use_graph=1, frames=12, icp_launch=kfx.ICP_LAUNCH_TWO, **kw)
    for a, b in zip(i1, i2):
        assert np.array_equal(np.array(a.pose_world.t[:]), np.array(b.pose_world.t[:]))
        assert np.array_equal(np.array(a.pose_world.R[:]), np.array(b.pose_world.R[:]))
        assert a.icp_iterations == b.icp_iterations
    for p in (p1, p2):
        p.close()


def test_early_exit_reduces_iterations(gpu):
    kfx = gpu
    pipe, infos, gt = _c1_run(kfx, 1, frames=15, converge_rotation=1e-5, converge_translation=1e-5)
    its = [i.icp_iterations for i in infos[1:]]
    assert max(its) <= 19 and np.mean(its) < 19
    assert kfx.ate_rmse([i.pose_world for i in infos], gt) < 0.01
    pipe.close()


@pytest.mark.parametrize("mode,use_graph", [(0, 0), (1, 1)])
def test_cyclic_shift_keeps_tracking(gpu, mode, use_graph):
    # Random walk of 2 cm per frame (the 512^3 workload's trajectory) in a
    # 256^3, 3.072 m cube with a 0.4 m shift threshold on the target point:
    # several shifts in 80 frames, tracking continues across them.
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    scene = kfx.make_scene(2, 10, 10, (0.5, 9.0, -2.0, 2.0), 0.0)
    gt = kfx.walk(2, (0.0, -1.2, 0.0), 0.02, 0.35, 80)
    R0, t0 = gt[0].matrices()
    origin = t0 + 1.5 * R0[:, 2] - 1.536  # cube centred on the first target point
    depths = np.stack([kfx.render(scene, gt[i], intr, 2, i, noise=True) for i in range(80)])
    cfg = kfx.default_config(intr=intr, voxels=kfx.Point3i(256, 256, 256),
                             volume_size=kfx.Point3f(3.072, 3.072, 3.072), mode=mode, use_graph=use_graph,
                             shift_threshold=0.4)
    pipe = kfx.Pipeline(cfg, gt[0], origin)
    infos = _run(kfx, pipe, depths)
    v = pipe.view()
    shifted = [i for i in infos if i.shifted]
    assert len(shifted) >= 1 and v.shifts == len(shifted)
    assert v.extracted_points > 0
    o = v.round_buffer.origin_GRID
    assert (o.x, o.y, o.z) != (0, 0, 0)
    ate = kfx.ate_rmse([i.pose_world for i in infos], gt)
    print(f"mode {mode}: {len(shifted)} shifts, ATE {ate * 1000:.2f} mm")
    assert ate < 0.02
    pipe.close()


def test_shift_extracts_records(gpu):
    # The same walk with (x, y, z, tsdf, weight) records: the last shift's
    # records lie in the slab that left the cube (outside the new cube's
    # global grid range on some axis, or on its first layer: a crossing
    # emits the +1 neighbour of a slab voxel), carry observed voxels, and
    # number as many as the points the XYZI run extracted at every shift.
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    scene = kfx.make_scene(2, 10, 10, (0.5, 9.0, -2.0, 2.0), 0.0)
    gt = kfx.walk(2, (0.0, -1.2, 0.0), 0.02, 0.35, 60)
    R0, t0 = gt[0].matrices()
    origin = t0 + 1.5 * R0[:, 2] - 1.536
    depths = np.stack([kfx.render(scene, gt[i], intr, 2, i, noise=True) for i in range(60)])
    counts = {}
    for rec in (0, 1):
        cfg = kfx.default_config(intr=intr, voxels=kfx.Point3i(256, 256, 256),
                                 volume_size=kfx.Point3f(3.072, 3.072, 3.072), mode=kfx.PLAIN,
                                 shift_threshold=0.4, extract_records=rec)
        pipe = kfx.Pipeline(cfg, gt[0], origin)
        infos = _run(kfx, pipe, depths)
        last = max(i for i, f in enumerate(infos) if f.shifted)
        counts[rec] = [f.extracted_points for f in infos]
        v = pipe.view()
        if rec:
            assert v.last_points is None and v.last_records
            n = infos[last].extracted_points
            r = kfx.download(v.last_records, kfx.SLICE_RECORD_DTYPE, n)
            g = v.round_buffer.origin_GRID_global
            assert n > 1000
            assert np.all(r["weight"] > 0)
            outside = np.zeros(n, bool)
            for ax, g0 in (("x", g.x), ("y", g.y), ("z", g.z)):
                g0 = int(round(g0))
                outside |= (r[ax] <= g0) | (r[ax] >= g0 + 256)
            assert outside.all()
        pipe.close()
    assert counts[0] == counts[1]


def _c1_learned(kfx, params, use_graph, frames=20, **kw):
    intr = kfx.kinect_intrinsics()
    scene, poses = scenes.c1_scene(), scenes.c1_poses(frames)
    depths = np.stack([kfx.render(scene, poses[i], intr, scenes.C1_SEED, i, noise=True) for i in range(frames)])
    cfg = kfx.default_config(intr=intr, voxels=kfx.Point3i(256, 256, 256), volume_size=kfx.Point3f(3.0, 3.0, 3.0),
                             mode=1, use_graph=use_graph, denoiser=1, shift_threshold=100.0, **kw)
    dn = kfx.Denoiser(intr.cols, intr.rows, params)
    pipe = kfx.Pipeline(cfg, poses[0], (0.0, 0.0, 0.0))
    pipe.set_denoiser(dn)
    infos = _run(kfx, pipe, depths)
    pipe.close()
    dn.close()
    return infos, poses


@pytest.mark.parametrize("use_graph", [0, 1])
def test_learned_pipeline_identity_denoiser_tracks(gpu, use_graph):
    # All-zero parameters: the denoiser passes the raw depth through, which
    # agrees with the model far inside the 15 mm gate, so no frame falls back.
    kfx = gpu
    infos, gt = _c1_learned(kfx, np.zeros(kfx.DENOISER_PARAMS, np.float32), use_graph)
    assert all(i.fallback == 0 for i in infos[1:])
    rms = [i.denoise_rms_mm for i in infos[1:]]
    assert 0.5 < np.median(rms) < 10.0, rms
    assert kfx.ate_rmse([i.pose_world for i in infos], gt) < 0.01


@pytest.mark.parametrize("case", ["trained", "rms_gate"])
def test_learned_gate_in_one_launch_frame_is_exact(gpu, case):
    # The learned pipeline evaluates its fallback gate in the one-launch
    # frame ICP; with separate launches (ICP iterations, then the two gate
    # kernels) the poses, gate statistics and decisions are the same bits.
    kfx = gpu
    if case == "trained":
        params, kw = kfx.Denoiser.load_params(), {}
    else:
        params = np.zeros(kfx.DENOISER_PARAMS, np.float32)
        params[-1] = 15.0
        kw = dict(icp_iterations=(0, 0, 0))
    runs = [_c1_learned(kfx, params, 1, frames=10, icp_launch=launch, **kw)[0]
            for launch in (kfx.ICP_LAUNCH_FRAME, kfx.ICP_LAUNCH_TWO)]
    for a, b in zip(*runs):
        assert np.array_equal(np.array(a.pose_world.t[:]), np.array(b.pose_world.t[:]))
        assert np.array_equal(np.array(a.pose_world.R[:]), np.array(b.pose_world.R[:]))
        assert a.fallback == b.fallback and a.denoise_rms_mm == b.denoise_rms_mm
    if case == "rms_gate":
        assert all(i.fallback == 1 for i in runs[0][1:])


def test_learned_pipeline_rms_gate(gpu):
    # Output bias of 15 noise sigmas (about 5 cm at the C1 depths of 1.2-2 m)
    # with ICP switched off (the estimate stays at the previous pose, so the
    # bias cannot be absorbed into the pose): the denoised frames disagree
    # with the model by more than the 15 mm RMS gate.
    kfx = gpu
    params = np.zeros(kfx.DENOISER_PARAMS, np.float32)
    params[-1] = 15.0
    infos, _ = _c1_learned(kfx, params, 1, frames=8, icp_iterations=(0, 0, 0))
    assert all(i.fallback == 1 and i.denoise_rms_mm > 15.0 for i in infos[1:]), [i.denoise_rms_mm for i in infos]


def test_learned_pipeline_coverage_gate(gpu):
    # Output bias of 120 noise sigmas (over 30 cm at the C1 depths): no
    # denoised pixel lies within 10 cm of the model, ICP finds no
    # correspondences, and the coverage gate (fewer than half of the model
    # pixels within 10 cm) falls back on every frame; the trajectory is the
    # bilateral pipeline's.
    kfx = gpu
    params = np.zeros(kfx.DENOISER_PARAMS, np.float32)
    params[-1] = 120.0
    infos, gt = _c1_learned(kfx, params, 1)
    assert all(i.fallback == 1 for i in infos[1:]), [i.denoise_rms_mm for i in infos]
    assert kfx.ate_rmse([i.pose_world for i in infos], gt) < 0.01


def test_graph_replay_survives_a_larger_pipeline(gpu):
    # Regression for per-pipeline workspaces: the captured graph of a
    # 640x480 learned pipeline holds its depth-bound pyramid and gate
    # partials.  A 1280x720 learned pipeline created and run in between
    # needs larger ones; the first pipeline's later replays must still
    # equal an uninterrupted run bit for bit.
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    scene, poses = scenes.c1_scene(), scenes.c1_poses(12)
    depths = np.stack([kfx.render(scene, poses[i], intr, scenes.C1_SEED, i, noise=True) for i in range(12)])
    zero = np.zeros(kfx.DENOISER_PARAMS, np.float32)

    def learned(intr_, n):
        cfg = kfx.default_config(intr=intr_, voxels=kfx.Point3i(n, n, n), volume_size=kfx.Point3f(3.0, 3.0, 3.0),
                                 mode=kfx.FUSED, use_graph=1, denoiser=1, shift_threshold=100.0)
        dn = kfx.Denoiser(intr_.cols, intr_.rows, zero)
        pipe = kfx.Pipeline(cfg, poses[0], (0.0, 0.0, 0.0))
        pipe.set_denoiser(dn)
        return pipe, dn

    dd = kfx.Buffer.from_array(depths)
    fb = depths[0].nbytes
    ref_pipe, ref_dn = learned(intr, 256)
    ref = [ref_pipe.process(C.c_void_p(dd.ptr.value + i * fb)) for i in range(12)]
    ref_pipe.close()
    ref_dn.close()

    pipe, dn = learned(intr, 256)
    got = [pipe.process(C.c_void_p(dd.ptr.value + i * fb)) for i in range(6)]
    big_intr = kfx.Intrinsics(900.0, 900.0, 639.5, 359.5, 1280, 720)
    big_depths = np.stack([kfx.render(scene, poses[i], big_intr, scenes.C1_SEED, i, noise=True) for i in range(3)])
    bd = kfx.Buffer.from_array(big_depths)
    big, big_dn = learned(big_intr, 128)
    big_infos = [big.process(C.c_void_p(bd.ptr.value + i * big_depths[0].nbytes)) for i in range(3)]
    big.close()
    big_dn.close()
    assert all(i.integrated for i in big_infos)
    got += [pipe.process(C.c_void_p(dd.ptr.value + i * fb)) for i in range(6, 12)]
    pipe.close()
    dn.close()
    for a, b in zip(ref, got):
        assert np.array_equal(np.array(a.pose_world.t[:]), np.array(b.pose_world.t[:]))
        assert np.array_equal(np.array(a.pose_world.R[:]), np.array(b.pose_world.R[:]))
        assert a.fallback == b.fallback and a.denoise_rms_mm == b.denoise_rms_mm


@pytest.mark.parametrize("use_graph", [0, 1])
def test_fused_raycast_builds_the_map_pyramid(gpu, use_graph):
    # The fused pipeline's skipping raycast writes levels 1 and 2 of the
    # model maps from its own tiles; they equal a separate resize of its
    # level 0 (kfx_resize_pyramid) bit for bit, NaNs included.
    kfx = gpu
    pipe, infos, _gt = _c1_run(kfx, 1, use_graph=use_graph, frames=12)
    v = pipe.view()
    assert v.brick_flags  # the skipping raycast ran
    intr = kfx.kinect_intrinsics()
    W, H = intr.cols, intr.rows
    got = [kfx.download(v.vmap_model_levels[l], np.float32, (H >> l, W >> l, 3)) for l in range(3)]
    got_n = [kfx.download(v.nmap_model_levels[l], np.float32, (H >> l, W >> l, 3)) for l in range(3)]
    assert v.vmap_model_levels[0] == v.vmap_model and v.nmap_model_levels[0] == v.nmap_model
    out = [kfx.Buffer((W >> l) * (H >> l) * 12) for l in (1, 1, 2, 2)]
    kfx.call("kfx_resize_pyramid", C.c_void_p(v.vmap_model), C.c_void_p(v.nmap_model), W, H,
             out[0].ptr, out[1].ptr, out[2].ptr, out[3].ptr)
    want = [o.download(np.float32, (H >> l, W >> l, 3)) for o, l in zip(out, (1, 1, 2, 2))]
    assert np.isfinite(got[1][..., 0]).mean() > 0.3
    for a, b in zip((got[1], got_n[1], got[2], got_n[2]), want):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    pipe.close()


@pytest.mark.parametrize("use_graph", [0, 1])
def test_fused_maps_launch_builds_scaled_depth_and_bound_pyramid(gpu, orc, use_graph):
    # The fused pipeline builds the integration's scaled depth and depth-bound
    # pyramid of the raw frame inside its maps launch (the pre-pass after ICP
    # then computes flags only): the scaled depth equals the oracle's
    # scale_depth of the last frame bit for bit, and every pyramid texel of
    # levels 1..5 is the maximum of its 2^l x 2^l pixels (0 outside the image).
    kfx = gpu
    frames = 3
    pipe, _infos, _gt = _c1_run(kfx, 1, use_graph=use_graph, frames=frames)
    v = pipe.view()
    assert v.depth_bound
    intr = kfx.kinect_intrinsics()
    W, H = intr.cols, intr.rows
    d = kfx.render(scenes.c1_scene(), scenes.c1_poses(frames)[frames - 1], intr, scenes.C1_SEED, frames - 1,
                   noise=True)
    got = kfx.download(v.depth_scaled, np.float32, (H, W))
    want = orc.scale_depth_f32(d, intr).astype(np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    sizes = [((W + (1 << l) - 1) >> l, (H + (1 << l) - 1) >> l) for l in range(1, 6)]
    pyr = kfx.download(v.depth_bound, np.float32, (sum(w * h for w, h in sizes),))
    pad = np.zeros(((H + 31) // 32 * 32, (W + 31) // 32 * 32), np.float32)
    pad[:H, :W] = want
    off = 0
    for l, (w, h) in enumerate(sizes, start=1):
        s = 1 << l
        m = pad.reshape(pad.shape[0] // s, s, pad.shape[1] // s, s).max(axis=(1, 3))[:h, :w]
        assert np.array_equal(pyr[off:off + w * h].reshape(h, w), m), f"level {l}"
        off += w * h
    pipe.close()
