"""The pipeline at scale: 1024^3 volume (4 mm voxels, 4.096 m cube, 4 GiB of
tsdf + weight), 1280x720 depth (fx = fy = 900, cx = 639.5, cy = 359.5), a
seeded random walk of the workload scene family (workload.py) at 2 cm per
frame, cyclic shifts at 1.5 m.  Times the learned pipeline (denoiser, graph
replay), the fused pipeline with the bilateral filter and the plain pipeline
on the same frames and reports frames/s and ATE for each, plus the integrate
kernel's average time (HIP events, plain run).

    python 3deye_amd/scale_study.py --frames 300 --out gpurun_out/scale.json
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--seed", type=int, default=wl.SCENE_SEED)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    intr = kfx.Intrinsics(900.0, 900.0, 639.5, 359.5, W, H)
    scene = wl.scene(args.seed)
    poses = wl.trajectory(args.seed, args.frames)
    R0, t0 = poses[0].matrices()
    origin = t0 + wl.TARGET_DISTANCE * R0[:, 2] - VOL_M / 2
    depths = np.stack([kfx.render(scene, poses[i], intr, args.seed, i, noise=True) for i in range(args.frames)])
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
        return {"frames_per_s": steps / el, "ms_per_frame": 1e3 * el / steps,
                "ate_m": kfx.ate_rmse([i.pose_world for i in infos], poses), "shifts": v.shifts,
                "fallbacks": int(sum(i.fallback for i in infos))}

    dn = kfx.Denoiser(W, H, kfx.Denoiser.load_params())
    out = {"workload": f"1024^3 (4 mm), 1280x720, {args.frames}-frame walk, scene seed {args.seed}",
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
