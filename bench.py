"""Benchmark: frames/s of the per-frame KinectFusion pipeline on one MI355X.

Workload (the 512^3 configuration, 3deye_amd/workload.py): 512^3 cyclic TSDF
volume (6 mm voxels, 3.072 m cube, int16 tsdf + int16 weight), 640x480
Kinect depth (fx=fy=525, cx=319.5, cy=239.5) rendered from a seeded analytic
scene (ground plane, 10 boxes, 10 spheres, sizes 0.1-0.5 m) along a
300-frame seeded random walk of 2 cm per frame (6 m) at 1.2 m height looking
0.6 rad down, Kinect noise sigma(z) = 1.2 + 1.9 (z - 0.4)^2 mm, 1 mm
quantisation, 2 % dropout; truncation 0.03 m; ICP 3 levels x 10/5/4; cyclic
shift when the camera's target point (1.5 m along its optical axis) is more
than 1.5 m from the cube centre, with the leaving slab extracted; the cube
starts centred on the first frame's target point, so the shifts fall at
frames 77, 150, 225 and 297.  Depth frames are rendered on the host before
timing and are resident in HBM when the timed region starts.

A run processes the walk's first PREROLL (60) frames untimed (they build the
map), then W warmup frames, then times K frames: with the defaults W = 10,
K = 230 the timed frames are 70-299 (four shifts); with W = 5, K = 20 they
are 65-84 (the shift at frame 77).  One step = one frame through the
pipeline (preprocess, ICP, integrate, raycast, shift with extraction).

The headline `value` is the learned pipeline (denoiser in place of the
bilateral filter with the bilateral fallback, device-resident pose,
one-launch frame ICP with FP32-compensated sums, raycast skipping bricks
without negative values), its kernels launched in order on one stream with
one host synchronisation per frame.  Every timed configuration first runs
the same window once untimed.  Timed on the same frames and reported beside
it: the same as one HIP graph per frame; with the ICP early exit; the fused
pipeline with the bilateral filter (launched, as a graph, with early exit);
and the plain pipeline (the reference's control flow: a synchronisation
after every kernel, the ICP normal equations downloaded and solved on the
host every iteration), each with its trajectory's ATE against the ground
truth over the whole walk (the frames after the timed ones run untimed).
Per-kernel tables: the plain pipeline's kernels and the learned pipeline's
kernels (a replay of the same frames with HIP events around the launches,
identical results), each with algorithmic bytes or FLOPs per launch.

Prints one JSON line.  Other modes: --kernel-probe (the learned pipeline,
eager, for hardware-counter runs), --summarise-pmc (turns two counter runs
of the probe into profiles/integrate_traffic.json), --summarise-counters
(per-kernel counter means of a probe run), --profile-learned N and
--profile-fused N (one pipeline, launched as in the headline, for a kernel
trace), --summarise-trace (kernel table of such a trace).
"""
import argparse
import ctypes as C
import glob
import json
import os
import sqlite3
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3deye_amd"))
import kfx  # noqa: E402
import workload as wl  # noqa: E402

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
W, H = wl.W, wl.H
N_VOX = wl.N_VOX
VOL_M = wl.VOL_M
TRUNC = wl.TRUNC
SCENE_SEED = wl.SCENE_SEED
N_FRAMES = 300
PREROLL = wl.PREROLL_FRAMES
DENOISER_FLOP_PER_PX = 10080
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "integrate_traffic.json")


def workload_inputs(n_frames):
    return wl.sequence(SCENE_SEED, n_frames)


def run_pipeline(cfg, poses, origin, ddepth, frame_bytes, warmup, steps, profile=False, denoiser=None,
                 rest=0):
    """Frames [0, PREROLL + warmup) untimed, then `steps` timed frames, then
    `rest` more frames untimed (the rest of the walk, for its ATE).
    Returns (elapsed seconds, frame infos, pipeline, index of the first timed
    frame)."""
    pipe = kfx.Pipeline(cfg, poses[0], origin)
    if denoiser is not None:
        pipe.set_denoiser(denoiser)
    first = PREROLL + warmup
    infos = [pipe.process(C.c_void_p(ddepth.ptr.value + i * frame_bytes)) for i in range(first)]
    kfx.synchronize()
    if profile:
        kfx.profile_reset()
        kfx.profile_enable(True)
    t0 = time.perf_counter()
    for i in range(first, first + steps):
        infos.append(pipe.process(C.c_void_p(ddepth.ptr.value + i * frame_bytes)))
    kfx.synchronize()
    elapsed = time.perf_counter() - t0
    if profile:
        kfx.profile_enable(False)
    for i in range(first + steps, first + steps + rest):
        infos.append(pipe.process(C.c_void_p(ddepth.ptr.value + i * frame_bytes)))
    kfx.synchronize()
    return elapsed, infos, pipe, first


def summary(elapsed, infos, pipe, first, steps, poses):
    """frames/s of the timed frames; ATE over every frame processed (the
    whole walk when the run went on past the timed frames) and over the
    frames up to the end of the timed ones; what the timed frames did
    (shifts, extracted points, fallbacks, ICP iterations)."""
    timed = infos[first:first + steps]
    n = len(infos)
    est = [i.pose_world for i in infos]
    out = {"frames_per_s": steps / elapsed, "ms_per_frame": 1e3 * elapsed / steps,
           "ate_m": kfx.ate_rmse(est, poses[:n]),
           "ate_frames": [0, n - 1],
           "ate_m_to_timed_end": kfx.ate_rmse(est[:first + steps], poses[:first + steps]),
           "frames_timed": [first, first + steps - 1],
           "shifts_timed": int(sum(i.shifted for i in timed)),
           "shift_frames_timed": [first + k for k, i in enumerate(timed) if i.shifted],
           "extracted_points_timed": int(sum(i.extracted_points for i in timed)),
           "fallbacks_timed": int(sum(i.fallback for i in timed)),
           "icp_iterations_per_frame": float(np.mean([i.icp_iterations for i in timed]))}
    pipe.close()
    return out


def count_updates(intr, ddepth, frame_bytes, infos, first, rb):
    """N_upd of each frame from `first` on, from the pose it was integrated at
    (0 for a frame the pipeline did not integrate)."""
    scaled = kfx.Buffer(W * H * 4)
    cs = kfx.cell_size(rb)
    out = []
    for k, info in enumerate(infos[first:], start=first):
        if not info.integrated:
            out.append(0)
            continue
        R, t = info.pose_volume.matrices()
        n = C.c_uint64()
        kfx.call("kfx_integrate_count_updates", C.c_void_p(ddepth.ptr.value + k * frame_bytes), scaled.ptr,
                 C.byref(intr), C.byref(kfx.matrix3f(R.T)), C.byref(kfx.point3f(t)), TRUNC, C.byref(cs),
                 C.byref(rb), C.byref(n))
        out.append(n.value)
    return out


def count_raycast_samples(cfg, poses, origin, ddepth, frame_bytes, first, n, denoiser=None):
    """S (the raycast algorithm's voxel reads) averaged over frames [first, n)
    of a pipeline: an untimed replay of the same deterministic pipeline,
    recounting each frame's raycast on its volume at its pose.  Frames
    followed by a shift are left out (their volume has moved on)."""
    intr = cfg.intr
    pipe = kfx.Pipeline(cfg, poses[0], origin)
    if denoiser is not None:
        pipe.set_denoiser(denoiser)
    vm, nm = kfx.Buffer(W * H * 12), kfx.Buffer(W * H * 12)
    total, frames = 0, 0
    for i in range(n):
        info = pipe.process(C.c_void_p(ddepth.ptr.value + i * frame_bytes))
        if i < first or info.shifted:
            continue
        v = pipe.view()
        R, t = info.pose_volume.matrices()
        ns = C.c_uint64()
        kfx.call("kfx_raycast_count_samples", C.c_void_p(v.volume), vm.ptr, nm.ptr, C.byref(v.round_buffer),
                 C.byref(kfx.point3f(t)), C.byref(kfx.matrix3f(R)), TRUNC, C.byref(intr), C.byref(v.cell_size),
                 C.byref(ns))
        total += ns.value
        frames += 1
    pipe.close()
    return total / max(frames, 1), frames


def level_sizes():
    px = [(W >> l) * (H >> l) for l in range(3)]
    return px


def kernel_row(name, nbytes_per_launch=None, flop_per_launch=None):
    """Launches, average duration from the HIP events of the profiled run,
    and the achieved rate against the roofline of the kernel's bound."""
    ms, cnt = kfx.profile_query(name)
    if not cnt:
        return None
    avg_s = ms / 1e3 / cnt
    row = {"launches": cnt, "avg_us": round(avg_s * 1e6, 2)}
    if nbytes_per_launch is not None:
        row["alg_bytes_per_launch"] = round(nbytes_per_launch)
        row["GBs"] = round(nbytes_per_launch / avg_s / 1e9, 1)
        row["frac_hbm_peak"] = round(nbytes_per_launch / avg_s / 1e9 / HBM_PEAK_GBS, 4)
    if flop_per_launch is not None:
        row["alg_flop_per_launch"] = round(flop_per_launch)
        row["TFLOPs"] = round(flop_per_launch / avg_s / 1e12, 2)
        row["frac_mfma_bf16_peak"] = round(flop_per_launch / avg_s / 1e12 / MFMA_BF16_PEAK_TFLOPS, 4)
    return row


def kernel_table_plain(n_upd_per_launch, samples_per_launch):
    """The plain pipeline's kernels.  Algorithmic bytes per launch: bilateral
    4 B/px; pyramid level i 2 (WH/4^(i-1)) + 2 (WH/4^i); vertex map 14 B/px;
    normal map 24 B/px; resize 12 B per input and output pixel; ICP sums per
    iteration 48 B/px + 108 B per workgroup partial, their reduction 108 B per
    partial; scale_depth 6 B/px; integrate 8 N_upd + 2 WH; raycast 4 S + 24 WH."""
    lv = [(W >> l, H >> l) for l in range(3)]
    px = level_sizes()
    nblk = [((w + 31) // 32) * ((h + 7) // 8) for w, h in lv]
    iters = [10, 5, 4]
    per_frame = {
        "bilateral": (1, 4 * px[0]),
        "pyr_down": (2, sum(2 * px[i - 1] + 2 * px[i] for i in (1, 2))),
        "vertex_map": (3, sum(14 * p for p in px)),
        "normal_map": (3, sum(24 * p for p in px)),
        "resize_map": (4, sum(2 * (12 * px[l - 1] + 12 * px[l]) for l in (1, 2))),
        "estimate_combined": (19, sum(iters[l] * (48 * px[l] + 108 * nblk[l]) for l in range(3))),
        "transform_estimation": (19, sum(iters[l] * (108 * nblk[l] + 108) for l in range(3))),
        "depth_bound": (1, 6 * px[0]),
        "integrate": (1, 8 * n_upd_per_launch + 2 * px[0]),
        "raycast": (1, 4 * samples_per_launch + 24 * px[0]),
    }
    out = {}
    for name, (per, nbytes) in per_frame.items():
        row = kernel_row(name, nbytes / per)
        if row:
            out[name] = row
    return out


def kernel_table_learned(n_upd_per_launch, samples_per_launch):
    """The learned pipeline's kernels (eager replay of the timed frames).
    Algorithmic work per launch: denoiser 10,080 FLOP/px (bf16 MFMA hidden
    layers, against 2.5 PFLOP/s dense) and 4 B/px; the depth pyramid with the
    vertex + normal maps of every level 26 B/px (level-0 depth in, the other
    levels' depth out, 24 B of maps); ICP sums at level l 48 B/px + 216 B per workgroup (27 FP64
    partials), the solve 216 B per partial + the 600 B state; the gate 14 B/px
    + 32 B per workgroup, its reduce 32 B per workgroup; the integration
    pre-pass 2 + 4 + 4/3 B/px (depth in, scaled depth and pyramid out);
    integrate 8 N_upd + 2 WH; raycast 4 S + 24 WH; model-map resize 24 B per
    level-0 pixel read, 24 B per level-1 and level-2 pixel written; the
    one-launch frame ICP 48 B/px per iteration and its exchange."""
    px = level_sizes()
    nblk_icp = [(px[0] + 511) // 512, (px[1] + 255) // 256, (px[2] + 255) // 256]
    nblk_gate = (px[0] + 255) // 256
    rows = {
        "denoise_forward": kernel_row("denoise_forward", 4 * px[0], DENOISER_FLOP_PER_PX * px[0]),
        "maps_pyramid": kernel_row("maps_pyramid", sum(26 * p for p in px)),
        "icp_begin_frame": kernel_row("icp_begin_frame"),
        "depth_bound": kernel_row("depth_bound", (2 + 4 + 4 / 3) * px[0]),
        "integrate": kernel_row("integrate", 8 * n_upd_per_launch + 2 * px[0]),
        "raycast": kernel_row("raycast", 4 * samples_per_launch + 24 * px[0]),
        "resize_pyramid": kernel_row("resize_pyramid", 24 * (px[0] + px[1] + px[2])),
        "consistency": kernel_row("consistency", 14 * px[0] + 32 * nblk_gate),
        "consistency_reduce": kernel_row("consistency_reduce", 32 * nblk_gate),
    }
    for l in range(3):
        rows[f"icp_sums@{W >> l}"] = kernel_row(f"icp_sums@{W >> l}", 48 * px[l] + 216 * nblk_icp[l])
    solve_bytes = sum(n_it * (216 * nb + 600) for n_it, nb in zip((10, 5, 4), nblk_icp)) / 19
    rows["icp_solve"] = kernel_row("icp_solve", solve_bytes)
    # A frame's ICP in one launch: per iteration the pixels' 48 B and the
    # exchange (27 values of each of 256 workgroups, written and read once by
    # the 4 leaders; 4 x 27 wave totals read by every workgroup).
    exch = 2 * 216 * 256 + 4 * 216 * 256
    rows["icp_frame"] = kernel_row("icp_frame", sum(n_it * (48 * p + exch) for n_it, p in zip((10, 5, 4), px)))
    return {k: v for k, v in rows.items() if v}


def measured_traffic(frames):
    """HBM bytes per integrate launch over the given frames, from the
    committed per-frame table of two rocprofv3 --pmc runs of --kernel-probe
    (FETCH_SIZE and WRITE_SIZE of each integrate dispatch, scaled by the
    factors that map a 256 MiB copy kernel's readings to 256 MiB), or None
    when the table does not cover the frames."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        t = json.load(f)
    per = t.get("frames", {})
    vals = [per[str(k)] for k in frames if str(k) in per]
    if len(vals) != len(frames) or not vals:
        return None
    return {"traffic_bytes_per_launch": float(np.mean([v["fetch"] + v["write"] for v in vals])),
            "fetch_bytes_per_launch": float(np.mean([v["fetch"] for v in vals])),
            "write_bytes_per_launch": float(np.mean([v["write"] for v in vals])),
            "source": os.path.relpath(TRAFFIC_FILE, ROOT) + " (" + t.get("command", "") + ")"}


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


def rgbd_cpu(intr, scene, poses, depths, pairs=4):
    """The oracle's RGBD odometry (one host thread) on the first pairs of the
    same frames: the CPU time per pair beside the GPU's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_ctypes as orc

    rgb = [kfx.render_rgb(scene, poses[k], intr).view(np.uint8) for k in range(pairs + 1)]
    t0 = time.perf_counter()
    for k in range(1, pairs + 1):
        orc.rgbd_estimate(rgb[k - 1], depths[k - 1], rgb[k], depths[k], intr, 4, (7, 7, 7, 10), 10.0, 0.07,
                          np.eye(3), np.zeros(3))
    return {"ms_per_pair": 1e3 * (time.perf_counter() - t0) / pairs, "cores": 1, "kind": "port",
            "sample": "oracle RGBD odometry, walk frames 0-%d" % pairs}


def rgbd_odometry(intr, scene, poses, depths, pairs=60, warmup=5):
    """RGBD photometric odometry (kfx_rgbd) on consecutive frames of the walk,
    with rendered colour: pairs per second (frame preparation, 4 levels x
    7/7/7/10 iterations from the identity, blocking result) and the error of
    each pair's motion against the rendered motion."""
    n = pairs + warmup + 1
    rgb = kfx.Buffer.from_array(np.stack([kfx.render_rgb(scene, poses[k], intr) for k in range(n)]))
    dep = kfx.Buffer.from_array(depths[:n])
    cb, db = W * H * 4, W * H * 2
    h = C.c_void_p()
    kfx.check(kfx.lib.kfx_rgbd_create(W, H, 4, C.byref(h)))
    prm = kfx.rgbd_params()
    ident = kfx.pose_from(np.eye(3), np.zeros(3))
    res = kfx.RgbdResult()

    def frame(slot, k):
        kfx.check(kfx.lib.kfx_rgbd_set_frame(h, slot, C.c_void_p(rgb.ptr.value + k * cb),
                                             C.c_void_p(dep.ptr.value + k * db)))

    frame(0, 0)
    et, er, solves = [], [], []
    t0 = None
    for k in range(1, n):
        if k == warmup + 1:
            kfx.synchronize()
            t0 = time.perf_counter()
        frame(1, k)
        kfx.check(kfx.lib.kfx_rgbd_estimate(h, C.byref(intr), C.byref(prm), C.byref(ident), C.byref(res)))
        if k > warmup:
            R0, tv0 = poses[k - 1].matrices()
            R1, tv1 = poses[k].matrices()
            Rt, tt = R1.T @ R0, R1.T @ (np.asarray(tv0) - np.asarray(tv1))
            R = np.array(res.pose.R[:]).reshape(3, 3)
            et.append(float(np.linalg.norm(np.array(res.pose.t[:]) - tt)))
            er.append(float(np.arccos(np.clip((np.trace(R @ Rt.T) - 1.0) / 2.0, -1.0, 1.0))))
            solves.append(res.solves)
        kfx.check(kfx.lib.kfx_rgbd_swap(h))
    el = time.perf_counter() - t0
    kfx.check(kfx.lib.kfx_rgbd_destroy(h))
    return {"pairs_per_s": pairs / el, "ms_per_pair": 1e3 * el / pairs, "pairs": pairs,
            "trans_err_median_m": float(np.median(et)), "rot_err_median_rad": float(np.median(er)),
            "solves_per_pair": float(np.mean(solves)),
            "config": "640x480 walk frames 1-%d, 4 levels x 7/7/7/10, identity start" % (n - 1)}


def kernel_probe(frames=N_FRAMES):
    """Short, fixed workload for hardware-counter runs (rocprofv3 --pmc):
    three 256 MiB streaming copies (known traffic, for calibrating the
    FETCH_SIZE / WRITE_SIZE counters), then the learned pipeline of the
    benchmark (the same kernels and results as the headline run), over
    the walk's frames.  Prints, in order, the frame of every integration
    dispatch with a device pose (frame 0 integrates at the host pose; a frame
    that falls back to the bilateral filter dispatches twice, the first one
    skipped on the device)."""
    intr, scene, poses, origin, depths = workload_inputs(frames)
    kfx.init(0)
    nbytes = 256 << 20
    a, b = kfx.Buffer(nbytes), kfx.Buffer(nbytes)
    a.fill(1)
    for _ in range(3):
        kfx.call("kfx_copy_kernel", b.ptr, a.ptr, nbytes)
    kfx.synchronize()
    a.free()
    b.free()
    dd = kfx.Buffer.from_array(depths)
    dn = kfx.Denoiser(W, H, kfx.Denoiser.load_params())
    cfg = wl.pipeline_config(mode=kfx.FUSED, use_graph=0, denoiser=1)
    pipe = kfx.Pipeline(cfg, poses[0], origin)
    pipe.set_denoiser(dn)
    dispatch_frames = []
    for i in range(frames):
        info = pipe.process(C.c_void_p(dd.ptr.value + i * depths[0].nbytes))
        if i > 0:
            dispatch_frames += [i] * (2 if info.fallback else 1)
    kfx.synchronize()
    pipe.close()
    dn.close()
    print(json.dumps({"probe": "copy 3 x 256 MiB; learned pipeline (eager) over the walk", "frames": frames,
                      "copy_bytes_per_launch": nbytes, "integrate_dispatch_frames": dispatch_frames}))


def _pmc_values(db_dir, counter):
    """(kernel name, counter value) per dispatch, in dispatch order, from a
    rocprofv3 --pmc output directory."""
    dbs = glob.glob(os.path.join(db_dir, "**", "*.db"), recursive=True)
    if not dbs:
        raise SystemExit(f"no rocprofv3 database under {db_dir}")
    con = sqlite3.connect(dbs[0])
    rows = con.execute("select dispatch_id, kernel_name, value from counters_collection where counter_name = ? "
                       "order by dispatch_id", (counter,)).fetchall()
    con.close()
    agg = {}
    for disp, name, value in rows:
        k = agg.setdefault(disp, [name, 0.0])
        k[1] += value
    return [tuple(agg[d]) for d in sorted(agg)]


def summarise_pmc(fetch_dir, write_dir, probe_log, out_json, command):
    """Per-frame HBM traffic of the learned pipeline's integrate dispatches
    from a FETCH_SIZE run and a WRITE_SIZE run of --kernel-probe, each counter
    scaled by the factor that maps the 256 MiB copy kernel's reading to
    256 MiB."""
    with open(probe_log) as f:
        probe = json.loads([ln for ln in f if ln.startswith("{")][-1])
    frames = probe["integrate_dispatch_frames"]
    copy_kib = probe["copy_bytes_per_launch"] / 1024.0
    out = {"command": command, "frames": {}}
    for label, d, counter in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        vals = _pmc_values(d, counter)
        copies = [v for n, v in vals if "copy16_kernel" in n]
        scale = copy_kib / float(np.mean(copies)) if copies else 1.0
        integ = [v for n, v in vals if "integrate_kernel<true>" in n]
        if len(integ) != len(frames):
            raise SystemExit(f"{counter}: {len(integ)} integrate dispatches, the probe made {len(frames)}")
        out[f"{label}_calibration"] = scale
        for fr, v in zip(frames, integ):
            e = out["frames"].setdefault(str(fr), {"fetch": 0.0, "write": 0.0})
            e[label] += v * 1024.0 * scale
    with open(out_json, "w") as f:
        json.dump(out, f, indent=0)
    print(f"wrote {out_json}: {len(out['frames'])} frames; calibration fetch {out['fetch_calibration']:.4f} "
          f"write {out['write_calibration']:.4f}")


def summarise_trace(db_dir, frames):
    """Kernel table of a rocprofv3 --kernel-trace run of --profile-learned or
    --profile-fused: per kernel, dispatches, mean and median span (us) and
    microseconds per frame over the last `frames` frames (the timed ones: the
    dispatches after the frame boundary that leaves `frames` pipeline
    frames, one denoiser or bilateral launch opening each)."""
    dbs = glob.glob(os.path.join(db_dir, "**", "*.db"), recursive=True)
    if not dbs:
        raise SystemExit(f"no rocprofv3 database under {db_dir}")
    con = sqlite3.connect(dbs[0])
    rows = con.execute("select S.display_name, K.start, K.end from rocpd_kernel_dispatch K join "
                       "rocpd_info_kernel_symbol S on S.id = K.kernel_id and S.guid = K.guid "
                       "order by K.start").fetchall()
    con.close()
    opens = [i for i, r in enumerate(rows) if "dn_fused_fwd" in r[0] or "bilateral_kernel" in r[0]]
    if len(opens) < frames:
        raise SystemExit(f"{len(opens)} frames in the trace, asked for {frames}")
    rows = rows[opens[-frames]:]
    per = {}
    for name, a, b in rows:
        per.setdefault(name, []).append((b - a) / 1e3)
    wall = (rows[-1][2] - rows[0][1]) / 1e3 / frames
    print(f"{frames} frames; first-dispatch to last-end wall time per frame {wall:.1f} us")
    print(f"{'kernel':60s} {'calls':>6s} {'mean_us':>8s} {'med_us':>8s} {'us/frame':>9s}")
    total = 0.0
    for name, d in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        total += sum(d) / frames
        print(f"{name[:60]:60s} {len(d):6d} {np.mean(d):8.2f} {np.median(d):8.2f} {sum(d) / frames:9.1f}")
    print(f"{'sum of kernel spans per frame':60s} {'':6s} {'':8s} {'':8s} {total:9.1f}")


def summarise_counters(db_dir, kernels):
    """Mean per-dispatch value of every counter of a rocprofv3 --pmc run of
    --kernel-probe for the named kernels (substring match), over their
    dispatches after the first 20 (a populated volume)."""
    dbs = glob.glob(os.path.join(db_dir, "**", "*.db"), recursive=True)
    if not dbs:
        raise SystemExit(f"no rocprofv3 database under {db_dir}")
    con = sqlite3.connect(dbs[0])
    rows = con.execute("select dispatch_id, kernel_name, counter_name, value from counters_collection").fetchall()
    con.close()
    per = {}
    for disp, name, counter, value in rows:
        per.setdefault((disp, name), {}).setdefault(counter, 0.0)
        per[(disp, name)][counter] += value
    for want in kernels:
        got = [m for (d, n), m in sorted(per.items()) if want in n]
        if not got:
            print(f"{want}: no dispatches")
            continue
        use = got[20:] if len(got) > 40 else got
        print(f"{want}: {len(use)} dispatches")
        for counter in sorted(use[0]):
            print(f"  {counter:28s} {np.mean([m[counter] for m in use]):16.1f}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=230)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-rgbd", action="store_true", help="no RGBD odometry measurement")
    ap.add_argument("--profile-learned", type=int, default=0, metavar="N",
                    help="only the learned pipeline (as in the headline): the pre-roll, 10 warmup and N timed frames, "
                         "for a kernel trace (prints frames/s, no JSON benchmark line)")
    ap.add_argument("--profile-fused", type=int, default=0, metavar="N",
                    help="the same for the fused pipeline with the bilateral filter")
    ap.add_argument("--kernel-probe", action="store_true",
                    help="the learned pipeline, eager, for hardware-counter runs (no JSON benchmark line)")
    ap.add_argument("--summarise-pmc", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "PROBE_LOG"),
                    help="write profiles/integrate_traffic.json from two --pmc runs of --kernel-probe")
    ap.add_argument("--summarise-trace", nargs=2, metavar=("TRACE_DIR", "FRAMES"),
                    help="kernel table of the last FRAMES frames of a --kernel-trace run of --profile-learned/fused")
    ap.add_argument("--summarise-counters", nargs="+", metavar=("PMC_DIR", "KERNEL"),
                    help="mean counter values per dispatch of the named kernels from a --pmc run of --kernel-probe")
    ap.add_argument("--pmc-command", default="", help="the counter command, recorded with --summarise-pmc")
    ap.add_argument("--traffic-out", default=TRAFFIC_FILE, help="where --summarise-pmc writes its table")
    args = ap.parse_args()
    if args.summarise_counters:
        summarise_counters(args.summarise_counters[0], args.summarise_counters[1:])
        return
    if args.summarise_trace:
        summarise_trace(args.summarise_trace[0], int(args.summarise_trace[1]))
        return
    if args.summarise_pmc:
        summarise_pmc(*args.summarise_pmc, args.traffic_out, args.pmc_command)
        return
    if args.kernel_probe:
        kernel_probe()
        return
    if args.profile_learned or args.profile_fused:
        learned_run = args.profile_learned > 0
        n_t = args.profile_learned if learned_run else args.profile_fused
        intr, scene, poses, origin, depths = workload_inputs(PREROLL + 10 + n_t)
        kfx.init(0)
        dd = kfx.Buffer.from_array(depths)
        dn = kfx.Denoiser(W, H, kfx.Denoiser.load_params()) if learned_run else None
        cfg = wl.pipeline_config(mode=kfx.FUSED, use_graph=0, denoiser=1 if learned_run else 0)
        el, _, pipe, _ = run_pipeline(cfg, poses, origin, dd, depths[0].nbytes, 10, n_t, denoiser=dn)
        pipe.close()
        print("%s pipeline: %d frames, %.1f frames/s" % ("learned" if learned_run else "fused", n_t, n_t / el))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the pipeline runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built).", file=sys.stderr)
        sys.exit(2)
    n = PREROLL + args.warmup + args.steps
    if n > N_FRAMES or args.warmup < 1 or args.steps < 1:
        sys.exit(f"need warmup >= 1, steps >= 1 and {PREROLL} + warmup + steps <= {N_FRAMES} (the walk)")

    intr, scene, poses, origin, depths = workload_inputs(N_FRAMES)
    kfx.init(0)
    ddepth = kfx.Buffer.from_array(depths)
    fb = depths[0].nbytes
    first = PREROLL + args.warmup
    rest = N_FRAMES - n  # frames after the timed ones: untimed, for the whole walk's ATE

    dn = kfx.Denoiser(W, H, kfx.Denoiser.load_params())
    early = dict(converge_rotation=1e-5, converge_translation=1e-5)
    def timed(cfg, den=None):
        """A run of the timed window, after an identical untimed pass (the
        first run of a configuration in a process starts slower)."""
        run_pipeline(cfg, poses, origin, ddepth, fb, args.warmup, args.steps, denoiser=den)[2].close()
        return run_pipeline(cfg, poses, origin, ddepth, fb, args.warmup, args.steps, denoiser=den, rest=rest)

    # Learned pipeline, kernels launched in order on one stream: headline.
    cfg_l = wl.pipeline_config(mode=kfx.FUSED, use_graph=0, denoiser=1)
    run_l = timed(cfg_l, dn)
    inf_l = run_l[1]
    learned = summary(*run_l, args.steps, poses)
    # Beside the headline, not as it: the same frames as one HIP graph per
    # frame (hipGraphLaunch costs more host time here than the frame's ten
    # launches); the KinectFusion early exit (a level stops once an increment
    # is below 1e-5 rad and 1e-5 m, the pose tolerance), at most 10/5/4
    # iterations, for both filters.
    learned["graph_replay"] = summary(*timed(wl.pipeline_config(mode=kfx.FUSED, use_graph=1, denoiser=1), dn),
                                      args.steps, poses)
    cfg_le = wl.pipeline_config(mode=kfx.FUSED, use_graph=0, denoiser=1, **early)
    learned["early_exit_1e-5"] = summary(*timed(cfg_le, dn), args.steps, poses)
    cfg_f = wl.pipeline_config(mode=kfx.FUSED, use_graph=0)
    fused = summary(*timed(cfg_f), args.steps, poses)
    fused["graph_replay"] = summary(*timed(wl.pipeline_config(mode=kfx.FUSED, use_graph=1)), args.steps, poses)
    cfg_fe = wl.pipeline_config(mode=kfx.FUSED, use_graph=0, **early)
    fused["early_exit_1e-5"] = summary(*timed(cfg_fe), args.steps, poses)

    # The learned pipeline again, eager, with HIP events around every launch:
    # the per-kernel table and the integrate roofline of the headline.
    cfg_lp = wl.pipeline_config(mode=kfx.FUSED, use_graph=0, denoiser=1)
    el_lp, inf_lp, pipe_lp, _ = run_pipeline(cfg_lp, poses, origin, ddepth, fb, args.warmup, args.steps,
                                             profile=True, denoiser=dn)
    pipe_lp.close()
    learned["profiled_replay_identical"] = all(
        a.pose_world.t[:] == b.pose_world.t[:] and a.pose_world.R[:] == b.pose_world.R[:]
        for a, b in zip(inf_l, inf_lp))
    int_ms, int_n = kfx.profile_query("integrate")
    rb = kfx.round_buffer(N_VOX, VOL_M)
    n_upd = count_updates(intr, ddepth, fb, inf_lp, first, rb)
    s_avg, s_frames = count_raycast_samples(cfg_lp, poses, origin, ddepth, fb, first, n, denoiser=dn)
    kern_l = kernel_table_learned(sum(n_upd) / int_n, s_avg)
    kern_l["raycast"]["samples_per_launch"] = round(s_avg)
    kern_l["raycast"]["samples_counted_frames"] = s_frames
    alg_bytes = (8.0 * sum(n_upd) + 2.0 * W * H * int_n) / int_n
    avg_s = int_ms / 1e3 / int_n
    achieved = alg_bytes / avg_s / 1e9
    traffic = measured_traffic(list(range(first, n)))
    roofline = {"bound": "hbm", "kernel": "integrate (device pose, learned pipeline)", "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic["traffic_bytes_per_launch"] if traffic else None,
                "traffic_source": traffic["source"] if traffic else None,
                "frames": [first, n - 1], "launches": int_n, "avg_launch_us": avg_s * 1e6,
                "alg_bytes_per_launch": alg_bytes, "n_upd_per_launch": sum(n_upd) / int_n,
                "full_sweep_equiv_GBs": 8.0 * N_VOX ** 3 / avg_s / 1e9}

    # Plain pipeline, with per-kernel HIP events.
    plain, kern_p = None, None
    if not args.skip_plain:
        cfg_p = wl.pipeline_config(mode=kfx.PLAIN)
        el_p, inf_p, pipe_p, _ = run_pipeline(cfg_p, poses, origin, ddepth, fb, args.warmup, args.steps,
                                              profile=True, rest=rest)
        p_int_ms, p_int_n = kfx.profile_query("integrate")
        p_upd = count_updates(intr, ddepth, fb, inf_p[:n], first, rb)
        ps_avg, ps_frames = count_raycast_samples(cfg_p, poses, origin, ddepth, fb, first, n)
        kfx.profile_enable(False)
        plain = summary(el_p, inf_p, pipe_p, first, args.steps, poses)
        # kernel_table_plain reads the plain run's events, still recorded.
        kern_p = kernel_table_plain(sum(p_upd) / max(p_int_n, 1), ps_avg)
        kern_p["raycast"]["samples_per_launch"] = round(ps_avg)
        kern_p["raycast"]["samples_counted_frames"] = ps_frames
    dn.close()

    rgbd = None if args.skip_rgbd else rgbd_odometry(intr, scene, poses, depths)
    if rgbd is not None and not args.no_cpu_baseline:
        rgbd["cpu_baseline"] = rgbd_cpu(intr, scene, poses, depths)
    base = None if args.no_cpu_baseline else cpu_baseline(intr, depths, poses, origin)
    out = {
        "metric": "frames/s (512^3, 640x480) learned vs plain at <= plain ATE; integrate HBM GB/s",
        "value": learned["frames_per_s"],
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": learned["ms_per_frame"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic scene, Kinect noise model)",
        "config": {"workload": f"512^3 TSDF (6 mm), 640x480 depth, 300-frame 6 m random walk with cyclic shifts "
                               f"at frames 77/150/225/297; timed frames {first}-{n - 1} after {PREROLL} pre-roll "
                               f"and {args.warmup} warmup frames",
                   "pipeline": "learned: denoiser (bf16 MFMA) with bilateral fallback at 15 mm RMS, device-resident "
                               "pose, one-launch frame ICP with FP64 workgroup and cross-workgroup sums, "
                               "one-wave Gauss-Jordan solve and pose update, raycast skipping bricks without negative values; "
                               "kernels launched in order on one stream, one host synchronisation per frame",
                   "frames_timed": [first, n - 1], "shifts_timed": learned["shifts_timed"],
                   "extracted_points_timed": learned["extracted_points_timed"],
                   "scene_seed": SCENE_SEED, "icp_iterations": [10, 5, 4]},
        "learned": learned,
        "fused": fused,
        "plain": plain,
        "speedup_vs_plain": (learned["frames_per_s"] / plain["frames_per_s"]) if plain else None,
        "ate_le_plain": (learned["ate_m"] <= plain["ate_m"]) if plain else None,
        "ate_note": "ate_m: RMSE of camera positions over the whole 300-frame walk (the frames after the timed "
                    "ones run untimed); ate_m_to_timed_end: over frames 0 to the last timed frame",
        "roofline": roofline,
        "kernels_learned": kern_l,
        "kernels_plain": kern_p,
        "rgbd_odometry": rgbd,
        "cpu_baseline": base,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
