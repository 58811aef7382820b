"""Benchmark: frames/s of the per-frame KinectFusion pipeline on one MI355X.

Workload (the 512^3 configuration): 512^3 cyclic TSDF volume (6 mm voxels,
3.072 m cube, int16 tsdf + int16 weight), 640x480 Kinect depth (fx=fy=525,
cx=319.5, cy=239.5) rendered from a seeded analytic scene (ground plane, 10
boxes, 10 spheres, sizes 0.1-0.5 m) along a 300-frame seeded random walk of
2 cm per frame (6 m), Kinect noise sigma(z) = 1.2 + 1.9 (z - 0.4)^2 mm, 1 mm
quantisation, 2 % dropout; truncation 0.03 m; ICP 3 levels x 10/5/4;
cyclic shift when the camera's target point (1.5 m along its optical axis)
is more than 1.5 m from the cube centre, with the leaving slab extracted; the
cube starts centred on the first frame's target point.  Depth frames are rendered on the host before
timing and are resident in HBM when the timed region starts.

One step = one frame through the pipeline (preprocess, ICP, integrate,
raycast, shift).  The headline `value` is the fused pipeline (device-resident
pose, one graph replay per frame); the plain pipeline (the reference's control
flow: a synchronisation after every kernel, the ICP normal equations
downloaded and solved on the host every iteration) is timed on the same
frames and reported beside it, with both trajectories' ATE.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3deye_amd"))
import kfx  # noqa: E402

HBM_PEAK_GBS = 8000.0
W, H = 640, 480
N_VOX = 512
VOL_M = 3.072
TRUNC = 0.03
SCENE_SEED = 2
N_FRAMES = 300


def cube_origin(pose, target_distance=1.5):
    """World position of the cube's corner voxel such that the cube's centre
    is the camera's target point, target_distance along its optical axis
    (where the pipeline re-centres the cube on every cyclic shift)."""
    R, t = pose.matrices()
    return t + target_distance * R[:, 2] - VOL_M / 2


def workload_inputs(n_frames):
    intr = kfx.kinect_intrinsics(W, H, 525.0)
    scene = kfx.make_scene(SCENE_SEED, 10, 10, (0.5, 9.0, -2.0, 2.0), 0.0)
    poses = kfx.walk(SCENE_SEED, (0.0, -1.2, 0.0), 0.02, 0.35, n_frames)
    origin = cube_origin(poses[0])
    depths = np.stack([kfx.render(scene, poses[i], intr, SCENE_SEED, i, noise=True) for i in range(n_frames)])
    return intr, scene, poses, origin, depths


def run_pipeline(cfg, poses, origin, ddepth, frame_bytes, warmup, steps, profile=False):
    """Frames [0, warmup) untimed, [warmup, warmup+steps) timed.  Returns
    (elapsed seconds, frame infos, pipeline)."""
    pipe = kfx.Pipeline(cfg, poses[0], origin)
    infos = []
    for i in range(warmup):
        infos.append(pipe.process(C.c_void_p(ddepth.ptr.value + i * frame_bytes)))
    kfx.synchronize()
    if profile:
        kfx.profile_reset()
        kfx.profile_enable(True)
    t0 = time.perf_counter()
    for i in range(warmup, warmup + steps):
        infos.append(pipe.process(C.c_void_p(ddepth.ptr.value + i * frame_bytes)))
    kfx.synchronize()
    elapsed = time.perf_counter() - t0
    if profile:
        kfx.profile_enable(False)
    return elapsed, infos, pipe


def count_updates(intr, ddepth, frame_bytes, infos, first, rb):
    """N_upd of each integrated frame, from the pose it was integrated at."""
    scaled = kfx.Buffer(W * H * 4)
    cs = kfx.cell_size(rb)
    total = 0
    for k, info in enumerate(infos[first:], start=first):
        if not info.integrated:
            continue
        R, t = info.pose_volume.matrices()
        n = C.c_uint64()
        kfx.call("kfx_integrate_count_updates", C.c_void_p(ddepth.ptr.value + k * frame_bytes), scaled.ptr,
                 C.byref(intr), C.byref(kfx.matrix3f(R.T)), C.byref(kfx.point3f(t)), TRUNC, C.byref(cs),
                 C.byref(rb), C.byref(n))
        total += n.value
    return total


def cpu_baseline(intr, depths, poses, origin, frames=12):
    """The oracle's single-precision restatement of the per-frame pipeline
    (bilateral, pyramid, vertex/normal maps, 19 ICP iterations, integrate,
    raycast) on the host cores, over the first frames of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes as orc

    cores = min(16, os.cpu_count() or 1)
    orc.set_threads(cores)
    rb = kfx.round_buffer(N_VOX, VOL_M)
    vol = np.zeros(rb.buffer_length, dtype=kfx.TSDF_DTYPE)
    ang = float(np.sin(np.deg2rad(20.0)))
    R0, t0 = poses[0].matrices()
    R, t = R0, t0 - origin
    vp = np_ = None
    t_start = time.perf_counter()
    for i in range(frames):
        lv = [orc.bilateral(depths[i])]
        for _ in range(2):
            lv.append(orc.pyr_down(lv[-1]))
        vc = [orc.vertex_map(lv[l], intr.level(l)) for l in range(3)]
        nc = [orc.normal_map_f64(v) for v in vc]
        if i > 0:
            Rp, tp = R, t
            for l in (2, 1, 0):
                for _ in range((10, 5, 4)[l]):
                    s, _n = orc.icp_sums(vc[l], nc[l], vp[l], np_[l], R, t, Rp.T, tp, intr.level(l), 0.1, ang, 0)
                    _rc, R, t, _x = orc.icp_solve_update(s, R, t)
        orc.integrate_f32(vol, depths[i], intr, R.T, t, TRUNC, rb)
        v0, n0 = orc.raycast_f32(vol, rb, t, R, TRUNC, intr)
        vp, np_ = [v0], [n0]
        for _ in range(2):
            vp.append(orc.resize_map(vp[-1], False))
            np_.append(orc.resize_map(np_[-1], True))
    dt = time.perf_counter() - t_start
    return {"value": frames / dt, "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"oracle restatement of the full per-frame pipeline (bilateral, pyramid, vertex/normal maps, "
                      f"19 ICP iterations, integrate, raycast), first {frames} frames of the 512^3 workload"}


def kernel_probe(frames=20):
    """Short, fixed workload for hardware-counter runs (rocprofv3 --pmc):
    three 256 MiB streaming copies (known traffic, for calibrating the
    FETCH_SIZE / WRITE_SIZE counters), then `frames` integrations of the
    512^3 workload at the ground-truth poses, then their raycasts."""
    intr, scene, poses, origin, depths = workload_inputs(frames)
    kfx.init(0)
    nbytes = 256 << 20
    a, b = kfx.Buffer(nbytes), kfx.Buffer(nbytes)
    a.fill(1)
    for _ in range(3):
        kfx.call("kfx_copy_kernel", b.ptr, a.ptr, nbytes)
    kfx.synchronize()
    rb = kfx.round_buffer(N_VOX, VOL_M)
    cs = kfx.cell_size(rb)
    vol = kfx.Buffer(rb.buffer_length * 4)
    kfx.call("kfx_init_tsdf", vol.ptr, C.byref(rb))
    dd = kfx.Buffer.from_array(depths)
    scaled = kfx.Buffer(W * H * 4)
    vm, nm = kfx.Buffer(W * H * 12), kfx.Buffer(W * H * 12)
    n_upd = 0
    for i in range(frames):
        R, t = poses[i].matrices()
        tv = t - origin
        dp = C.c_void_p(dd.ptr.value + i * depths[0].nbytes)
        kfx.call("kfx_integrate_tsdf", dp, scaled.ptr, vol.ptr, C.byref(intr), C.byref(kfx.matrix3f(R.T)),
                 C.byref(kfx.point3f(tv)), TRUNC, C.byref(cs), C.byref(rb))
        kfx.call("kfx_raycast", vol.ptr, vm.ptr, nm.ptr, C.byref(rb), C.byref(kfx.point3f(tv)),
                 C.byref(kfx.matrix3f(R)), TRUNC, C.byref(intr), C.byref(cs))
    kfx.synchronize()
    for i in range(frames):
        R, t = poses[i].matrices()
        n = C.c_uint64()
        dp = C.c_void_p(dd.ptr.value + i * depths[0].nbytes)
        kfx.call("kfx_integrate_count_updates", dp, scaled.ptr, C.byref(intr), C.byref(kfx.matrix3f(R.T)),
                 C.byref(kfx.point3f(t - origin)), TRUNC, C.byref(cs), C.byref(rb), C.byref(n))
        n_upd += n.value
    print(json.dumps({"probe": "copy 3 x 256 MiB; integrate + raycast", "frames": frames,
                      "copy_bytes_per_launch": nbytes, "n_upd_per_frame": n_upd / frames,
                      "integrate_alg_bytes_per_frame": (8.0 * n_upd + 2.0 * W * H * frames) / frames}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=290)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--kernel-probe", action="store_true",
                    help="short fixed workload for hardware-counter runs (no JSON benchmark line)")
    args = ap.parse_args()
    if args.kernel_probe:
        kernel_probe()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the pipeline runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built).", file=sys.stderr)
        sys.exit(2)
    n = args.warmup + args.steps
    if n > N_FRAMES or args.warmup < 1:
        sys.exit(f"need warmup >= 1 and warmup + steps <= {N_FRAMES} (the trajectory length)")

    intr, scene, poses, origin, depths = workload_inputs(N_FRAMES)
    kfx.init(0)
    ddepth = kfx.Buffer.from_array(depths[:n])
    fb = depths[0].nbytes
    common = dict(intr=intr, voxels=kfx.Point3i(N_VOX, N_VOX, N_VOX), volume_size=kfx.Point3f(VOL_M, VOL_M, VOL_M))

    # Fused pipeline: headline.
    cfg_f = kfx.default_config(mode=kfx.FUSED, use_graph=1, **common)
    el_f, inf_f, pipe_f = run_pipeline(cfg_f, poses, origin, ddepth, fb, args.warmup, args.steps)
    v_f = pipe_f.view()
    pipe_f.close()
    ate_f = kfx.ate_rmse([i.pose_world for i in inf_f], poses[:n])

    # Plain pipeline, with per-kernel HIP events for the integrate roofline.
    plain = None
    el_p, inf_p = None, None
    if not args.skip_plain:
        cfg_p = kfx.default_config(mode=kfx.PLAIN, **common)
        el_p, inf_p, pipe_p = run_pipeline(cfg_p, poses, origin, ddepth, fb, args.warmup, args.steps, profile=True)
        pipe_p.close()
        ate_p = kfx.ate_rmse([i.pose_world for i in inf_p], poses[:n])
        plain = {"frames_per_s": args.steps / el_p, "ms_per_frame": 1e3 * el_p / args.steps, "ate_m": ate_p}
    int_ms, int_n = kfx.profile_query("integrate")
    kern = {}
    for name in ("bilateral", "pyr_down", "vertex_map", "normal_map", "estimate_combined", "transform_estimation",
                 "integrate", "raycast", "resize_map", "slice_count", "slice_emit", "clean_tsdf_slice"):
        ms, cnt = kfx.profile_query(name)
        if cnt:
            kern[name] = round(1e3 * ms / cnt, 2)

    roofline = None
    if int_n:
        rb = kfx.round_buffer(N_VOX, VOL_M)
        n_upd = count_updates(intr, ddepth, fb, inf_p, args.warmup, rb)
        launches = int_n
        alg_bytes = (8.0 * n_upd + 2.0 * W * H * launches) / launches
        avg_s = int_ms / 1e3 / launches
        achieved = alg_bytes / avg_s / 1e9
        roofline = {"bound": "hbm", "kernel": "integrate", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": achieved / HBM_PEAK_GBS, "traffic": None, "avg_launch_us": avg_s * 1e6,
                    "alg_bytes_per_launch": alg_bytes, "n_upd_per_launch": n_upd / launches,
                    "full_sweep_equiv_GBs": 8.0 * N_VOX ** 3 / avg_s / 1e9}

    base = None if args.no_cpu_baseline else cpu_baseline(intr, depths, poses, origin)
    fps = args.steps / el_f
    out = {
        "metric": "frames/s (512^3, 640x480) learned vs plain at <= plain ATE; integrate HBM GB/s",
        "value": fps,
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el_f / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic scene, Kinect noise model)",
        "config": {"workload": "512^3 TSDF (6 mm), 640x480 depth, 300-frame 6 m random walk with cyclic shifts",
                   "pipeline": "fused: device-resident pose, FP64 ICP sums and solve, "
                               "one HIP graph per frame; bilateral filter",
                   "frames": N_FRAMES, "icp_iterations": [10, 5, 4]},
        "fused": {"frames_per_s": fps, "ate_m": ate_f, "shifts": v_f.shifts,
                  "extracted_points": int(v_f.extracted_points)},
        "plain": plain,
        "speedup_vs_plain": (fps / plain["frames_per_s"]) if plain else None,
        "roofline": roofline,
        "kernels_avg_us_plain": kern,
        "cpu_baseline": base,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
