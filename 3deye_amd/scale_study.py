"""The pipeline at scale: 1024^3 volume (4 mm voxels, 4.096 m cube, 4 GiB of
tsdf + weight), 1280x720 depth (fx = fy = 900, cx = 639.5, cy = 359.5), a
seeded random walk of the workload scene family (workload.py), cyclic shifts
at 1.5 m.  With --long, the long walk: the scene family over 13.5 m
(workload.long_scene) walked at 6 mm per frame, 2000 frames covering 12 m;
otherwise the 512^3 workload's scene at 2 cm per frame.  Times the learned
pipeline (denoiser, graph replay), the fused pipeline with the bilateral
filter and the plain pipeline on the same frames and reports frames/s,
ATE, shifts and extracted points for each, plus the integrate and raycast
kernels' average times (HIP events, plain run).

    python 3deye_amd/scale_study.py --long --frames 2000 --out gpurun_out/scale_long.json
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import kfx  # noqa: E402
import workload as wl  # noqa: E402

W, H = 1280, 720
N_VOX, VOL_M = 1024, 4.096


def summary_text(path, command):
    with open(path) as f:
        d = json.load(f)
    lines = [f"# The pipeline at 1024^3 / 1280x720: {d['workload']}", "", "Command:", "", "    " + command, ""]
    for k in ("learned", "fused", "plain"):
        x = d[k]
        lines.append(f"{k:8s} {x['frames_per_s']:7.1f} frames/s ({x['ms_per_frame']:.3f} ms/frame), ATE "
                     f"{x['ate_m'] * 1e3:.2f} mm, RPE {x['rpe_trans_m'] * 1e3:.3f} mm / {x['rpe_rot_rad']:.2e} rad per "
                     f"frame, final position error {x['final_position_error_m'] * 1e3:.1f} mm, {x['shifts']} shifts "
                     f"at {x['shift_frames']}, {x['extracted_points']} points extracted, {x['fallbacks']} fallbacks")
    lines.append("")
    lines.append(f"learned / plain: {d['speedup_learned_vs_plain']:.2f}x frames/s; plain-run kernel averages (HIP "
                 f"events): integrate {d['integrate_avg_us_plain']:.1f} us, raycast {d['raycast_avg_us_plain']:.1f} us")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=wl.SCENE_SEED)
    ap.add_argument("--long", action="store_true", help="the 12 m walk over the long scene at 6 mm per frame")
    ap.add_argument("--out", default=None)
    ap.add_argument("--summarise", metavar="JSON", help="print a text summary of a study output")
    ap.add_argument("--command", default="", help="the study command, for --summarise")
    args = ap.parse_args()
    if args.summarise:
        print(summary_text(args.summarise, args.command))
        return
    intr = kfx.Intrinsics(900.0, 900.0, 639.5, 359.5, W, H)
    scene = wl.long_scene(args.seed) if args.long else wl.scene(args.seed)
    poses = wl.trajectory(args.seed, args.frames, step=wl.LONG_STEP if args.long else wl.STEP)
    origin = wl.cube_origin(poses[0], vol_m=VOL_M)
    depths = np.zeros((args.frames, H, W), np.uint16)
    t_r = time.perf_counter()
    for i in range(args.frames):
        depths[i] = kfx.render(scene, poses[i], intr, args.seed, i, noise=True)
        if (i + 1) % 200 == 0:
            print(f"rendered {i + 1}/{args.frames} frames ({time.perf_counter() - t_r:.0f} s)", flush=True)
    kfx.init(0)
    dd = kfx.Buffer.from_array(depths)
    fb = depths[0].nbytes
    common = dict(intr=intr, voxels=kfx.Point3i(N_VOX, N_VOX, N_VOX), volume_size=kfx.Point3f(VOL_M, VOL_M, VOL_M))

    def run(name, cfg, dn=None, profile=False):
        pipe = kfx.Pipeline(cfg, poses[0], origin)
        if dn is not None:
            pipe.set_denoiser(dn)
        infos = [pipe.process(C.c_void_p(dd.ptr.value + i * fb)) for i in range(args.warmup)]
        kfx.synchronize()
        if profile:
            kfx.profile_reset()
            kfx.profile_enable(True)
        t = time.perf_counter()
        infos += [pipe.process(C.c_void_p(dd.ptr.value + i * fb)) for i in range(args.warmup, args.frames)]
        kfx.synchronize()
        el = time.perf_counter() - t
        if profile:
            kfx.profile_enable(False)
        v = pipe.view()
        pipe.close()
        steps = args.frames - args.warmup
        ate, rpe_t, rpe_r = kfx.trajectory_errors([i.pose_world for i in infos], poses)
        out = {"frames_per_s": steps / el, "ms_per_frame": 1e3 * el / steps, "ate_m": ate,
               "rpe_trans_m": rpe_t, "rpe_rot_rad": rpe_r, "shifts": v.shifts,
               "shift_frames": [k for k, i in enumerate(infos) if i.shifted],
               "extracted_points": int(sum(i.extracted_points for i in infos)),
               "fallbacks": int(sum(i.fallback for i in infos)),
               "final_position_error_m": float(np.linalg.norm(np.array(infos[-1].pose_world.t[:]) -
                                                              np.array(poses[args.frames - 1].t[:])))}
        print(f"{name}: {out['frames_per_s']:.1f} frames/s, ATE {ate * 1e3:.2f} mm, {v.shifts} shifts", flush=True)
        return out

    dn = kfx.Denoiser(W, H, kfx.Denoiser.load_params())
    path_m = float(sum(np.linalg.norm(np.array(poses[i + 1].t[:]) - np.array(poses[i].t[:]))
                       for i in range(args.frames - 1)))
    out = {"workload": f"1024^3 (4 mm), 1280x720, {args.frames}-frame walk over {path_m:.1f} m "
                       f"({'long scene, 6 mm' if args.long else '2 cm'} per frame), scene seed {args.seed}",
           "path_m": path_m,
           "learned": run("learned", kfx.default_config(mode=kfx.FUSED, use_graph=1, denoiser=1, **common), dn),
           "fused": run("fused", kfx.default_config(mode=kfx.FUSED, use_graph=1, **common)),
           "plain": run("plain", kfx.default_config(mode=kfx.PLAIN, **common), profile=True)}
    ms, n = kfx.profile_query("integrate")
    out["integrate_avg_us_plain"] = 1e3 * ms / n if n else None
    ms, n = kfx.profile_query("raycast")
    out["raycast_avg_us_plain"] = 1e3 * ms / n if n else None
    out["speedup_learned_vs_plain"] = out["learned"]["frames_per_s"] / out["plain"]["frames_per_s"]
    dn.close()
    txt = json.dumps(out, indent=1)
    print(txt)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt)


if __name__ == "__main__":
    main()
