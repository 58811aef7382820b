"""Rendered RGBD frame pairs of the C1 orbit and their true relative motion."""
import numpy as np

import scenes


def rgbd_frames(kfx, intr, frames, noise=False):
    scene = scenes.c1_scene()
    poses = scenes.c1_poses(30)
    out = []
    for f in frames:
        d = kfx.render(scene, poses[f], intr, scenes.C1_SEED, f, noise=noise)
        c = kfx.render_rgb(scene, poses[f], intr)
        out.append((np.ascontiguousarray(c), np.ascontiguousarray(d)))
    return out, poses


def true_motion(poses, f_prev, f_curr):
    """(R, t) taking previous-camera coordinates to current-camera ones."""
    R0, t0 = poses[f_prev].matrices()
    R1, t1 = poses[f_curr].matrices()
    return R1.T @ R0, R1.T @ (np.asarray(t0) - np.asarray(t1))


def rot_angle(R):
    return float(np.arccos(np.clip((np.trace(R) - 1.0) / 2.0, -1.0, 1.0)))
