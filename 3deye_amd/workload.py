"""The 512^3 walk workload and its scene family.

One place for the configuration the benchmark, the denoiser training and the
tests share: 512^3 cyclic TSDF volume (6 mm voxels, 3.072 m cube), 640x480
Kinect depth (fx = fy = 525, cx = 319.5, cy = 239.5), truncation 0.03 m,
scenes of a ground plane with N_BOXES boxes and N_SPHERES spheres (sizes and
radii 0.1-0.5 m) in REGION, and a seeded random walk of STEP metres per frame
at HEIGHT metres, looking PITCH radians down.  The pipeline tracks the
camera's target point, 1.5 m along the optical axis, and re-centres the cube
on it at every cyclic shift (when it is more than 1.5 m from the cube
centre).  The cube starts centred on the first target point; the first
shift falls at frame 77 of the benchmark walk, then every 73-75 frames.
A timed run starts after PREROLL_FRAMES untimed frames that build the map,
so that even a short timed window (from frame 65 on) holds a shift whose
leaving slab was observed, with its points extracted.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import kfx  # noqa: E402

W, H = 640, 480
N_VOX = 512
VOL_M = 3.072
TRUNC = 0.03
N_BOXES, N_SPHERES = 10, 10
REGION = (0.5, 7.5, -1.0, 1.0)      # x0, x1, z0, z1 of the object centres (metres)
HEIGHT = 1.2                        # camera height above the ground (metres)
PITCH = 0.6                         # look-down angle (radians)
STEP = 0.02                         # metres per frame
SCENE_SEED = 4                      # the benchmark's scene and trajectory
TARGET_DISTANCE = 1.5               # the cube's centre is this far along the optical axis
PREROLL_FRAMES = 60                 # untimed frames before a timed run's warmup


# The long walk (1024^3 / 1280x720 runs): the same scene family over 13.5 m
# at the same object density, walked at LONG_STEP metres per frame (2000
# frames cover 12 m).
LONG_REGION = (0.5, 14.0, -1.5, 1.5)
LONG_BOXES, LONG_SPHERES = 29, 29
LONG_STEP = 0.006


def intrinsics():
    return kfx.kinect_intrinsics(W, H, 525.0)


def scene(seed):
    return kfx.make_scene(seed, N_BOXES, N_SPHERES, REGION, 0.0)


def trajectory(seed, n_frames, step=STEP):
    return kfx.walk(seed, (0.0, -HEIGHT, 0.0), step, PITCH, n_frames)


def long_scene(seed):
    return kfx.make_scene(seed, LONG_BOXES, LONG_SPHERES, LONG_REGION, 0.0)


def target_point(pose, target_distance=TARGET_DISTANCE):
    R, t = pose.matrices()
    return t + target_distance * R[:, 2]


def cube_origin(pose, target_distance=TARGET_DISTANCE, vol_m=VOL_M):
    """World position of the cube's corner voxel such that the cube's centre
    is the camera's target point."""
    return target_point(pose, target_distance) - vol_m / 2


def pipeline_config(**kw):
    return kfx.default_config(intr=intrinsics(), voxels=kfx.Point3i(N_VOX, N_VOX, N_VOX),
                              volume_size=kfx.Point3f(VOL_M, VOL_M, VOL_M), **kw)


def sequence(seed, n_frames, threads=None, want_clean=False):
    """(intrinsics, scene, poses, cube origin, depth frames[, exact depths])."""
    intr = intrinsics()
    sc = scene(seed)
    poses = trajectory(seed, n_frames)
    kw = {} if threads is None else {"threads": threads}
    if want_clean:
        frames = [kfx.render(sc, poses[i], intr, seed, i, noise=True, want_clean=True, **kw) for i in range(n_frames)]
        depths = np.stack([f[0] for f in frames])
        clean = np.stack([f[1] for f in frames])
        return intr, sc, poses, cube_origin(poses[0]), depths, clean
    depths = np.stack([kfx.render(sc, poses[i], intr, seed, i, noise=True, **kw) for i in range(n_frames)])
    return intr, sc, poses, cube_origin(poses[0]), depths
