"""Train the learned depth denoiser on the GPU with the project's own kernels.

Training pairs are (raw depth frame, model depth) where the model depth is the
pipeline's own raycast of its TSDF at the frame's estimated pose, produced by
the plain pipeline (the reference's control flow, bilateral filter) running
on seeded random-walk sequences of the 512^3 / 640x480 workload family
(workload.py: ground plane, 10 boxes, 10 spheres per scene, 2 cm per frame).
Every pair contributes one forward + backward pass (kfx_denoiser_accumulate);
Adam steps every `--batch` pairs.  The first `--skip` frames of a sequence
are not used (the model is still thin there), nor are frames whose estimated
position is more than 1 cm from the ground truth.  When a run drifts 3 cm
from the ground truth, its model no longer matches the frames: the pipeline
is restarted at that frame's ground-truth pose (a relocalisation) and the
next `--skip` frames are again not used.

Held-out evaluation: on scenes with seeds not used for training, the RMSE of
raw, bilateral-filtered and denoised depth against the renderer's exact depth,
on pixels valid in all three.

    python 3deye_amd/train_denoiser.py --seeds 50 --frames 400 \
        --out 3deye_amd/denoiser_weights.f32 --log gpurun_out/denoiser_train.json
"""
import argparse
import concurrent.futures as cf
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

W, H = wl.W, wl.H
NP = kfx.DENOISER_PARAMS
TRAIN_SEED_BASE = 100
HELDOUT_SEED_BASE = 5000
KEEP_ERR_M = 0.01
STOP_ERR_M = 0.03


def sequence(seed, frames):
    intr = wl.intrinsics()
    scene = wl.scene(seed)
    poses = wl.trajectory(seed, frames)
    return intr, scene, poses, wl.cube_origin(poses[0])


def render_seq(seed, frames, threads):
    intr, scene, poses, origin = sequence(seed, frames)
    depths = np.stack([kfx.render(scene, poses[i], intr, seed, i, noise=True, threads=threads) for i in range(frames)])
    return depths


def init_params(rng):
    """He-normal weights for the ReLU layers (fan-in 2 * 9 and 16 * 9), a
    small output layer (the net starts close to the identity: output = raw
    depth + residual), zero biases.  Layout: W1 [16][2][9], b1, W2 [16][16][9],
    b2, W3, b3, W4 [16][9], b4."""
    p = np.zeros(NP, np.float32)
    p[0:288] = rng.normal(0.0, np.sqrt(2.0 / 18.0), 288)
    p[304:2608] = rng.normal(0.0, np.sqrt(2.0 / 144.0), 2304)
    p[2624:4928] = rng.normal(0.0, np.sqrt(2.0 / 144.0), 2304)
    p[4944:5088] = rng.normal(0.0, 1e-3, 144)
    return p


def load_params(path):
    p = np.fromfile(path, dtype="<f4")
    if p.size != NP:
        raise ValueError(f"{path}: {p.size} parameters, expected {NP}")
    return p


def heldout_eval(dn, seeds, frames, threads):
    """RMSE (mm) of raw / bilateral / denoised depth against exact depth."""
    sums = {"raw": 0.0, "bilateral": 0.0, "denoised": 0.0}
    count = 0
    intr = kfx.kinect_intrinsics(W, H, 525.0)
    src, bil, den = kfx.Buffer(W * H * 2), kfx.Buffer(W * H * 2), kfx.Buffer(W * H * 2)
    for s in seeds:
        _, scene, poses, _ = sequence(s, frames)
        for i in range(0, frames, max(1, frames // 10)):
            raw, clean = kfx.render(scene, poses[i], intr, s, i, noise=True, threads=threads, want_clean=True)
            src.upload(raw)
            kfx.call("kfx_bilateral_filter", src.ptr, bil.ptr, W, H, 4.5, 30.0, 0xFFFFFFFF)
            kfx.call("kfx_denoiser_forward", dn, src.ptr, den.ptr, None)
            b = bil.download(np.uint16, (H, W)).astype(np.float64)
            d = den.download(np.uint16, (H, W)).astype(np.float64)
            r = raw.astype(np.float64)
            ok = (raw > 0) & (clean > 0) & (np.abs(r - clean * 1000) < 50)
            gt = clean[ok] * 1000.0
            sums["raw"] += np.sum((r[ok] - gt) ** 2)
            sums["bilateral"] += np.sum((b[ok] - gt) ** 2)
            sums["denoised"] += np.sum((d[ok] - gt) ** 2)
            count += int(ok.sum())
    return {k: float(np.sqrt(v / count)) for k, v in sums.items()}, count


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=50)
    ap.add_argument("--frames", type=int, default=400)
    ap.add_argument("--skip", type=int, default=10)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--heldout", type=int, default=10)
    ap.add_argument("--init", default=None, help="start from these parameters")
    ap.add_argument("--init-seed", type=int, default=1, help="seed of the random initialisation")
    ap.add_argument("--output-scale", type=float, default=1.0,
                    help="scale of the trained residual (output layer weights and bias) in the saved network")
    ap.add_argument("--out", default=os.path.join(HERE, "denoiser_weights.f32"))
    ap.add_argument("--log", default=None)
    ap.add_argument("--threads", type=int, default=min(16, os.cpu_count() or 1))
    args = ap.parse_args()

    kfx.init(0)
    dn = C.c_void_p()
    kfx.call("kfx_denoiser_create", W, H, C.byref(dn))
    p0 = load_params(args.init) if args.init else init_params(np.random.default_rng(args.init_seed))
    kfx.call("kfx_denoiser_set_params", dn, p0.ctypes.data_as(C.c_void_p))

    cfg = wl.pipeline_config(mode=kfx.PLAIN)
    target = kfx.Buffer(W * H * 4)
    pool = cf.ThreadPoolExecutor(max_workers=1)
    seeds = [TRAIN_SEED_BASE + i for i in range(args.seeds)]
    fut = pool.submit(render_seq, seeds[0], args.frames, args.threads)
    history = []
    pairs = 0
    in_batch = 0
    batch_loss = 0.0
    t0 = time.time()
    total_pairs = args.seeds * (args.frames - args.skip)
    skipped = stopped = 0
    for si, seed in enumerate(seeds):
        depths = fut.result()
        if si + 1 < len(seeds):
            fut = pool.submit(render_seq, seeds[si + 1], args.frames, args.threads)
        intr, scene, poses, origin = sequence(seed, args.frames)
        dd = kfx.Buffer.from_array(depths)
        pipe = kfx.Pipeline(cfg, poses[0], origin)
        fb = depths[0].nbytes
        start = 0
        for k in range(args.frames):
            dptr = C.c_void_p(dd.ptr.value + k * fb)
            info = pipe.process(dptr)
            err = float(np.linalg.norm(info.pose_world.matrices()[1] - poses[k].matrices()[1]))
            if err > STOP_ERR_M:
                stopped += 1
                pipe.close()
                pipe = kfx.Pipeline(cfg, poses[k], wl.cube_origin(poses[k]))
                pipe.process(dptr)
                start = k
                continue
            if k - start < args.skip or err > KEEP_ERR_M:
                skipped += (k - start) >= args.skip
                continue
            v = pipe.view()
            kfx.call("kfx_model_depth", C.c_void_p(v.vmap_model), C.byref(v.pose_volume), W, H, target.ptr)
            loss = C.c_double()
            # cosine learning-rate decay over the run
            lr = args.lr * 0.5 * (1.0 + np.cos(np.pi * min(1.0, pairs / max(1, total_pairs))))
            kfx.call("kfx_denoiser_accumulate", dn, dptr, target.ptr, 0.05, C.byref(loss), None)
            batch_loss += loss.value
            in_batch += 1
            pairs += 1
            if in_batch == args.batch:
                kfx.call("kfx_denoiser_adam_step", dn, max(lr, 1e-5), in_batch)
                history.append(batch_loss / in_batch)
                in_batch = 0
                batch_loss = 0.0
        pipe.close()
        dd.free()
        recent = float(np.mean(history[-20:])) if history else float("nan")
        print(f"seed {si + 1}/{len(seeds)}: {pairs} pairs, loss {recent * 1e6:.2f} mm^2, "
              f"{time.time() - t0:.0f} s", flush=True)
    if in_batch:
        kfx.call("kfx_denoiser_adam_step", dn, 1e-5, in_batch)
    params = np.zeros(NP, np.float32)
    kfx.call("kfx_denoiser_get_params", dn, params.ctypes.data_as(C.c_void_p))
    # The residual is linear in the output layer: scaling its weights and bias
    # scales the correction (in inverse-depth noise units) applied to the raw
    # frame.
    params[4944:5089] *= np.float32(args.output_scale)
    kfx.call("kfx_denoiser_set_params", dn, params.ctypes.data_as(C.c_void_p))
    params.astype("<f4").tofile(args.out)

    heldout = [HELDOUT_SEED_BASE + i for i in range(args.heldout)]
    rmse, n = heldout_eval(dn, heldout, 200, args.threads)
    summary = {"pairs": pairs, "seeds": args.seeds, "frames_per_seed": args.frames, "batch": args.batch,
               "output_scale": args.output_scale,
               "frames_skipped_pose_error": skipped, "relocalisations": stopped,
               "lr": args.lr, "seconds": time.time() - t0, "loss_first": history[:5], "loss_last": history[-5:],
               "loss_history_every_10": history[::10], "heldout_seeds": heldout, "heldout_pixels": n,
               "heldout_rmse_mm": rmse}
    print(json.dumps({k: summary[k] for k in ("pairs", "seconds", "heldout_rmse_mm")}), flush=True)
    if args.log:
        with open(args.log, "w") as f:
            json.dump(summary, f, indent=1)
    kfx.call("kfx_denoiser_destroy", dn)


if __name__ == "__main__":
    main()
