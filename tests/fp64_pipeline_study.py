"""Pose and TSDF error of the GPU pipeline against the oracle's
double-precision pipeline (tests/fp64_pipeline.py), per configuration:

- c2: the 512^3 walk (640x480, 6 mm voxels, 300 frames, cyclic shifts at
  frames 77/150/225/297), 3deye_amd/workload.py;
- c3: the 512^3 precision-study orbit (1000 frames, no shifts),
  3deye_amd/precision_study.py.

Two comparisons, each on the GPU's fused pipeline with the bilateral filter:

- open loop (--part open): at every frame the oracle's FP64 ICP runs on the
  pipeline's own inputs (its current maps, its model maps and pose of the
  frame before) and its pose is compared with the pipeline's; the oracle's
  FP64 integration of every frame at the pipeline's poses (with its shifts)
  is compared with the pipeline's volume at the end: the error of the GPU's
  arithmetic per frame;
- closed loop (--part closed): the pipeline in each ICP sum mode
  (FP32-Kahan, plain FP32, FP64 workgroup sums; graph replay) against the
  oracle's FP64 pipeline run on its own (tests/fp64_pipeline.py): per-frame
  pose differences of the two trajectories (which also grow as the two
  trackers drift apart), the ATE of each against the ground truth, their
  shift frames and the end-of-run TSDF difference histogram.

Pose errors are reported as median, 99th percentile, maximum and decade
histograms; TSDF errors in quantisation steps over the observed voxels.

    python tests/fp64_pipeline_study.py --config c3 --frames 1000 --threads 16 --out gpurun_out/fp64_c3.json
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("3deye_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))
import kfx  # noqa: E402
import oracle_ctypes as orc  # noqa: E402
import workload as wl  # noqa: E402
import fp64_pipeline as fp  # noqa: E402

MODES = {"fp32_kahan": kfx.ICP_SUM_FP32_KAHAN, "fp32_plain": kfx.ICP_SUM_FP32, "fp64": kfx.ICP_SUM_FP64}


def log(msg):
    print(msg, flush=True)


def inputs(config, frames, threads):
    if config == "c2":
        intr, _scene, poses, origin, depths = wl.sequence(wl.SCENE_SEED, frames, threads=threads)
        return intr, poses, origin, depths, 1.5
    import precision_study as ps
    intr, poses, origin, depths = ps.orbit_sequence(wl.SCENE_SEED, frames, threads=threads)
    return intr, poses, origin, depths, 100.0


def gpu_run(mode, intr, poses, origin, dd, fb, frames, shift_threshold):
    cfg = wl.pipeline_config(mode=kfx.FUSED, use_graph=1, icp_sum_mode=mode, shift_threshold=shift_threshold)
    pipe = kfx.Pipeline(cfg, poses[0], origin)
    infos = [pipe.process(C.c_void_p(dd.ptr.value + i * fb)) for i in range(frames)]
    v = pipe.view()
    vol = kfx.download(v.volume, kfx.TSDF_DTYPE, v.round_buffer.buffer_length)
    pipe.close()
    return ([i.pose_world.matrices() for i in infos], [k for k, i in enumerate(infos) if i.shifted],
            [int(i.integrated) for i in infos], vol)


def _rb_copy(rb):
    out = kfx.RoundBuffer()
    C.memmove(C.byref(out), C.byref(rb), C.sizeof(rb))
    return out


def _slab(rb_pre, rb_post):
    """The slab a shift cleaned, from the round buffer before and after."""
    n = (rb_pre.voxels_size.x, rb_pre.voxels_size.y, rb_pre.voxels_size.z)
    o = (rb_pre.origin_GRID.x, rb_pre.origin_GRID.y, rb_pre.origin_GRID.z)
    g0 = (rb_pre.origin_GRID_global.x, rb_pre.origin_GRID_global.y, rb_pre.origin_GRID_global.z)
    g1 = (rb_post.origin_GRID_global.x, rb_post.origin_GRID_global.y, rb_post.origin_GRID_global.z)
    mn, mx = [], []
    for k in range(3):
        sk = int(round(g1[k] - g0[k]))
        if sk > 0:
            mn.append(o[k]); mx.append(o[k] + sk - 1)
        elif sk < 0:
            mn.append(o[k] + n[k] + sk); mx.append(o[k] + n[k] - 1)
        else:
            mn.append(1); mx.append(0)
    return tuple(mn), tuple(mx)


class GpuMaps:
    """The pipeline's current-frame vertex and normal maps of every level,
    rebuilt from its filtered depth: the oracle's pyramid (bit-exact with
    the GPU's) uploaded, then kfx_maps_pyramid (the pipeline's kernel)."""

    def __init__(self, intr):
        self.intr = intr
        self.lv = [intr.level(l) for l in range(3)]
        px = [i.cols * i.rows for i in self.lv]
        self.dep = [kfx.Buffer(2 * p) for p in px]
        self.vm = [kfx.Buffer(12 * p) for p in px]
        self.nm = [kfx.Buffer(12 * p) for p in px]

    def __call__(self, depth0):
        lv = [depth0]
        for _ in range(2):
            lv.append(orc.pyr_down(lv[-1]))
        for l in range(3):
            kfx.call("kfx_memcpy_htod", self.dep[l].ptr, np.ascontiguousarray(lv[l]).ctypes.data_as(C.c_void_p),
                     lv[l].nbytes)
        dep = (C.c_void_p * 3)(*[b.ptr.value for b in self.dep])
        vm = (C.c_void_p * 3)(*[b.ptr.value for b in self.vm])
        nm = (C.c_void_p * 3)(*[b.ptr.value for b in self.nm])
        kfx.call("kfx_maps_pyramid", dep, vm, nm, (kfx.Intrinsics * 3)(*self.lv), 3)
        kfx.synchronize()
        shp = [(i.rows, i.cols, 3) for i in self.lv]
        return ([self.vm[l].download(np.float32, shp[l]) for l in range(3)],
                [self.nm[l].download(np.float32, shp[l]) for l in range(3)])


def open_loop(intr, poses, origin, depths, frames, shift_thr):
    """The fused pipeline (FP32-Kahan sums, eager: the same poses as graph
    replay) with, at every frame, the oracle's FP64 ICP run on the
    pipeline's own inputs (its current maps, its model maps of the frame
    before and its previous pose) and compared with the pipeline's pose; and
    the oracle's FP64 integration of every frame at the pipeline's pose
    (with its shifts), compared with the pipeline's volume at the end."""
    dd = kfx.Buffer.from_array(depths)
    fb = depths[0].nbytes
    cfg = wl.pipeline_config(mode=kfx.FUSED, use_graph=0, shift_threshold=shift_thr)
    pipe = kfx.Pipeline(cfg, poses[0], origin)
    maps = GpuMaps(intr)
    v = pipe.view()
    orb = _rb_copy(v.round_buffer)
    ovol = np.zeros(orb.buffer_length, kfx.TSDF_DTYPE)
    ang = float(np.sin(np.deg2rad(20.0)))
    prev = None  # (Rp, tp volume coordinates, model maps per level)
    dt, dr, steps = [], [], []
    t1 = time.perf_counter()
    for i in range(frames):
        info = pipe.process(C.c_void_p(dd.ptr.value + i * fb))
        v = pipe.view()
        R, t = info.pose_volume.matrices()
        if prev is not None:
            Rp, tp, mv, mn_ = prev
            vc, nc = maps(kfx.download(v.depth_filtered, np.uint16, (intr.rows, intr.cols)))
            Ro, to = Rp.copy(), tp.copy()
            for l in (2, 1, 0):
                for _ in range((10, 5, 4)[l]):
                    sums, _n = orc.icp_sums(vc[l], nc[l], mv[l], mn_[l], Ro, to, Rp.T, tp, intr.level(l), 0.1, ang,
                                            1)
                    _rc, Ro, to, _x = orc.icp_solve_update(sums, Ro, to)
            dt.append(float(np.linalg.norm(t - to)))
            dr.append(fp.rot_angle(R, Ro))
            steps.append(float(np.linalg.norm(t - tp)))
        if info.integrated:
            orc.integrate_f64(ovol, depths[i], intr, R.T, t, wl.TRUNC, orb)
        if info.shifted:
            mn, mx = _slab(orb, v.round_buffer)
            orc.clean_slice(ovol, mn, mx, orb)
            orb = _rb_copy(v.round_buffer)
        mv = [kfx.download(v.vmap_model, np.float32, (intr.rows, intr.cols, 3))]
        mn_ = [kfx.download(v.nmap_model, np.float32, (intr.rows, intr.cols, 3))]
        for _ in range(2):
            mv.append(orc.resize_map(mv[-1], False))
            mn_.append(orc.resize_map(mn_[-1], True))
        Rw, tw = info.pose_world.matrices()
        prev = (Rw, tw - np.array(v.origin_world[:]), mv, mn_)
        if (i + 1) % 25 == 0:
            log(f"open loop frame {i + 1}/{frames} ({(time.perf_counter() - t1) / (i + 1):.2f} s/frame)")
    gvol = kfx.download(v.volume, kfx.TSDF_DTYPE, v.round_buffer.buffer_length)
    pipe.close()
    dt, dr = np.array(dt), np.array(dr)
    return {"icp_vs_fp64_on_same_inputs": fp.summarise(dt, dr),
            "frames_within_1e-5": int(np.count_nonzero((dt <= 1e-5) & (dr <= 1e-5))), "frames_compared": len(dt),
            "per_frame_trans_err_m": [float(x) for x in dt], "per_frame_rot_err_rad": [float(x) for x in dr],
            "median_frame_motion_m": float(np.median(steps)),
            "tsdf_vs_fp64_at_same_poses": fp.tsdf_histogram(gvol, ovol)}


def summary_text(paths, command):
    """A short text summary of study outputs (JSON files)."""
    lines = ["# GPU pipeline against the oracle's double-precision pipeline", "", "Command:", ""]
    lines += ["    " + c for c in command] + [""]
    for path in paths:
        with open(path) as f:
            d = json.load(f)
        lines.append(f"## {d['config']}: {d['workload']}, {d['frames']} frames")
        lines.append("")
        o = d.get("open_loop")
        if o:
            s = o["icp_vs_fp64_on_same_inputs"]
            lines.append("Open loop (the FP64 oracle ICP on the pipeline's own inputs each frame; FP32-Kahan sums):")
            lines.append(f"  pose error per frame: translation median {s['trans_median_m']:.2e} m, p99 "
                         f"{s['trans_p99_m']:.2e}, max {s['trans_max_m']:.2e}; rotation median "
                         f"{s['rot_median_rad']:.2e} rad, p99 {s['rot_p99_rad']:.2e}, max {s['rot_max_rad']:.2e}")
            lines.append(f"  frames within 1e-5 m and 1e-5 rad: {o['frames_within_1e-5']} of {o['frames_compared']} "
                         f"(median frame motion {o['median_frame_motion_m'] * 1e3:.1f} mm)")
            lines.append(f"  translation error decades: {s['trans_decades']}")
            lines.append(f"  rotation error decades:    {s['rot_decades']}")
            t = o["tsdf_vs_fp64_at_same_poses"]
            lines.append(f"  TSDF at the same poses, end of run: {t['voxels_observed']} observed voxels, "
                         f"{100 * t['within_1_lsb']:.4f} % within 1 LSB, weights differ on {t['weight_differs']}")
            lines.append(f"  |tsdf difference| in LSB: {t['lsb_histogram']}")
            lines.append("")
        c = d.get("closed_loop")
        if c:
            lines.append(f"Closed loop (the oracle's FP64 pipeline run on its own; shift frames "
                         f"{c['oracle']['shift_frames']}, ATE {c['oracle']['ate_m'] * 1e3:.3f} mm):")
            for name, m in c["modes"].items():
                s = m["vs_oracle"]
                t = m.get("tsdf_vs_oracle")
                lines.append(f"  {name}: trajectory difference median {s['trans_median_m']:.2e} m / "
                             f"{s['rot_median_rad']:.2e} rad, max {s['trans_max_m']:.2e} m / {s['rot_max_rad']:.2e} "
                             f"rad; ATE {m['ate_m'] * 1e3:.3f} mm; shifts match {m['shifts_match']}, integration "
                             f"decisions match {m['integration_decisions_match']}")
                if t:
                    lines.append(f"    TSDF: {100 * t['within_1_lsb']:.2f} % within 1 LSB, weights differ on "
                                 f"{t['weight_differs']} of {t['voxels_observed']}; LSB {t['lsb_histogram']}")
            lines.append("")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", choices=("c2", "c3"))
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--part", choices=("closed", "open", "both"), default="both")
    ap.add_argument("--out", default=None)
    ap.add_argument("--summarise", nargs="+", metavar="JSON", help="print a text summary of study outputs")
    ap.add_argument("--command", action="append", default=[], help="the study commands, for --summarise")
    args = ap.parse_args()
    if args.summarise:
        print(summary_text(args.summarise, args.command))
        return
    frames = args.frames or (300 if args.config == "c2" else 1000)
    orc.set_threads(args.threads)
    t0 = time.perf_counter()
    intr, poses, origin, depths, shift_thr = inputs(args.config, frames, args.threads)
    log(f"{args.config}: {frames} frames rendered in {time.perf_counter() - t0:.0f} s")
    kfx.init(0)
    out = {"config": args.config, "frames": frames,
           "workload": ("512^3 walk, 640x480, 300-frame 6 m random walk with cyclic shifts" if args.config == "c2"
                        else "512^3 orbit, 640x480, radius 1.5 m, 0.01 rad per frame, no shifts")}
    if args.part in ("open", "both"):
        out["open_loop"] = open_loop(intr, poses, origin, depths, frames, shift_thr)
        o = out["open_loop"]
        s = o["icp_vs_fp64_on_same_inputs"]
        log(f"open loop: ICP vs FP64 on the same inputs: trans median {s['trans_median_m']:.2e} max "
            f"{s['trans_max_m']:.2e} m, rot median {s['rot_median_rad']:.2e} max {s['rot_max_rad']:.2e} rad; "
            f"{o['frames_within_1e-5']}/{o['frames_compared']} frames within 1e-5; TSDF "
            f"{o['tsdf_vs_fp64_at_same_poses']}")
    if args.part == "open":
        if args.out:
            with open(args.out, "w") as f:
                f.write(json.dumps(out))
        log("done")
        return
    dd = kfx.Buffer.from_array(depths)
    fb = depths[0].nbytes
    gpu = {}
    for name, mode in MODES.items():
        gpu[name] = gpu_run(mode, intr, poses, origin, dd, fb, frames, shift_thr)
        log(f"gpu {name}: shifts at {gpu[name][1]}")
    del dd
    op = fp.OraclePipeline(orc, kfx, intr, wl.N_VOX, wl.VOL_M, poses[0], origin, trunc=wl.TRUNC,
                           shift_threshold=shift_thr)
    orc_poses, orc_shifts, orc_integrated = [], [], []
    t1 = time.perf_counter()
    for i in range(frames):
        R, t, integ, shift = op.process(depths[i])
        orc_poses.append((R, t))
        orc_integrated.append(int(integ))
        if shift is not None:
            orc_shifts.append(i)
        if (i + 1) % 25 == 0:
            log(f"oracle frame {i + 1}/{frames} ({(time.perf_counter() - t1) / (i + 1):.2f} s/frame)")
    gt = [p.matrices() for p in poses[:frames]]
    out["closed_loop"] = {"oracle": {"shift_frames": orc_shifts, "integrated_frames": int(sum(orc_integrated)),
                                     "ate_m": kfx.ate_rmse([kfx.pose_from(*p) for p in orc_poses], poses),
                                     "seconds_per_frame": (time.perf_counter() - t1) / frames,
                                     "threads": args.threads},
                          "modes": {}}
    closed = out["closed_loop"]
    for name, (gposes, gshifts, ginteg, gvol) in gpu.items():
        dt, dr = fp.pose_errors(gposes, orc_poses)
        gdt, gdr = fp.pose_errors(gposes, gt)
        rec = {"vs_oracle": fp.summarise(dt, dr), "vs_ground_truth_trans_max_m": float(gdt.max()),
               "ate_m": kfx.ate_rmse([kfx.pose_from(*p) for p in gposes], poses), "shift_frames": gshifts,
               "shifts_match": gshifts == orc_shifts,
               "integration_decisions_match": ginteg == orc_integrated,
               "per_frame_trans_err_m": [float(v) for v in dt], "per_frame_rot_err_rad": [float(v) for v in dr]}
        if gshifts == orc_shifts:
            rec["tsdf_vs_oracle"] = fp.tsdf_histogram(gvol, op.vol)
        closed["modes"][name] = rec
        s = rec["vs_oracle"]
        log(f"{name}: vs oracle trans median {s['trans_median_m']:.2e} max {s['trans_max_m']:.2e} m, rot median "
            f"{s['rot_median_rad']:.2e} max {s['rot_max_rad']:.2e} rad; ATE {rec['ate_m'] * 1e3:.3f} mm "
            f"(oracle {closed['oracle']['ate_m'] * 1e3:.3f}); TSDF {rec.get('tsdf_vs_oracle')}")
    txt = json.dumps(out)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt)
    log("done")


if __name__ == "__main__":
    main()
