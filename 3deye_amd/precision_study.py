"""ICP precision study on the GPU: the fused pipeline with FP32-Kahan, plain
FP32 and FP64 workgroup sums (cross-workgroup sum, 6x6 solve and pose
composition in FP64 in every mode) over a long seeded trajectory.

Workload: 512^3 volume (3.072 m cube), 640x480 depth of a workload-family
scene (workload.py: ground plane, 10 boxes, 10 spheres), camera on an orbit of
radius 1.5 m at 1.2 m height around the scene point (3, 0, 0), 0.01 rad per
frame, looking at that point; the cube is centred on it, so the run has no
cyclic shifts.  Reports, per mode, the per-frame position and rotation
errors against the ground truth and against the FP64-sum run, and at the end
the TSDF voxel differences against the FP64-sum run as a histogram in
quantisation steps (LSB).

    python 3deye_amd/precision_study.py --frames 1000 --out gpurun_out/precision.json
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

CENTER = (3.0, 0.0, 0.0)
MODES = {"fp32_kahan": kfx.ICP_SUM_FP32_KAHAN, "fp32_plain": kfx.ICP_SUM_FP32, "fp64": kfx.ICP_SUM_FP64}


def orbit_sequence(seed, frames, threads=None):
    intr = wl.intrinsics()
    scene = wl.scene(seed)
    poses = kfx.orbit(CENTER, 1.5, -wl.HEIGHT, 0.0, 0.01, frames)
    origin = np.array(CENTER) + np.array([0.0, -0.6, 0.0]) - wl.VOL_M / 2
    kw = {} if threads is None else {"threads": threads}
    depths = np.stack([kfx.render(scene, poses[i], intr, seed, i, noise=True, **kw) for i in range(frames)])
    return intr, poses, origin, depths


def rot_angle(Ra, Rb):
    """Angle of Ra^T Rb from the Frobenius distance, ||Ra - Rb|| = 2 sqrt(2)
    sin(angle / 2), accurate for small angles (unlike the arccos of the
    trace)."""
    return float(2.0 * np.arcsin(min(1.0, np.linalg.norm(Ra - Rb) / (2.0 * np.sqrt(2.0)))))


def run(mode, poses, origin, dd, fb, frames):
    cfg = wl.pipeline_config(mode=kfx.FUSED, use_graph=1, icp_sum_mode=mode, shift_threshold=100.0)
    pipe = kfx.Pipeline(cfg, poses[0], origin)
    infos = [pipe.process(C.c_void_p(dd.ptr.value + i * fb)) for i in range(frames)]
    v = pipe.view()
    n = v.round_buffer.buffer_length
    vol = np.zeros(n, kfx.TSDF_DTYPE)
    kfx.call("kfx_memcpy_dtoh", vol.ctypes.data_as(C.c_void_p), C.c_void_p(v.volume), n * 4)
    pipe.close()
    return [i.pose_world.matrices() for i in infos], vol


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--frames", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=wl.SCENE_SEED)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    t0 = time.time()
    intr, poses, origin, depths = orbit_sequence(args.seed, args.frames)
    kfx.init(0)
    dd = kfx.Buffer.from_array(depths)
    fb = depths[0].nbytes
    gt = [p.matrices() for p in poses]
    res = {name: run(m, poses, origin, dd, fb, args.frames) for name, m in MODES.items()}
    ref_poses, ref_vol = res["fp64"]
    report = {"workload": f"512^3, 640x480, {args.frames}-frame orbit r=1.5 m, scene seed {args.seed}",
              "modes": {}}
    for name, (ps, vol) in res.items():
        et = np.array([np.linalg.norm(p[1] - g[1]) for p, g in zip(ps, gt)])
        er = np.array([rot_angle(p[0], g[0]) for p, g in zip(ps, gt)])
        dt = np.array([np.linalg.norm(p[1] - r[1]) for p, r in zip(ps, ref_poses)])
        dr = np.array([rot_angle(p[0], r[0]) for p, r in zip(ps, ref_poses)])
        dv = np.abs(vol["value"].astype(np.int64) - ref_vol["value"].astype(np.int64))
        observed = (vol["weight"] > 0) | (ref_vol["weight"] > 0)
        d = dv[observed]
        report["modes"][name] = {
            "ate_m": float(kfx.ate_rmse([kfx.pose_from(*p) for p in ps], poses)),
            "pos_err_vs_gt_max_m": float(et.max()), "rot_err_vs_gt_max_rad": float(er.max()),
            "pos_diff_vs_fp64_max_m": float(dt.max()), "rot_diff_vs_fp64_max_rad": float(dr.max()),
            "pos_diff_vs_fp64_median_m": float(np.median(dt)),
            "frames_within_1e-5_of_fp64": int(np.sum((dt <= 1e-5) & (dr <= 1e-5))),
            "tsdf_vs_fp64_lsb_hist": {"0": int(np.sum(d == 0)), "1": int(np.sum(d == 1)),
                                      "2-10": int(np.sum((d >= 2) & (d <= 10))), ">10": int(np.sum(d > 10))},
            "tsdf_vs_fp64_within_1lsb_frac": float(np.mean(d <= 1)) if d.size else 1.0,
            "observed_voxels": int(d.size),
        }
    report["seconds"] = time.time() - t0
    txt = json.dumps(report, indent=1)
    print(txt)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt)


if __name__ == "__main__":
    main()
