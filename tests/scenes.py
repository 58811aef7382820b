"""Synthetic configurations used by the tests and the benchmark.

C1: 256^3 TSDF over a 3 m cube, 640x480 Kinect intrinsics (fx=fy=525,
cx=319.5, cy=239.5), ground plane + 8 boxes + 6 spheres, camera on a smooth
orbit 1.5 m out, 30 frames, truncation 0.03 m.
"""
import numpy as np

import kfx

C1_SEED = 1
TRUNC = 0.03


def c1_scene(seed=C1_SEED):
    return kfx.make_scene(seed, 8, 6, (0.8, 2.2, 0.8, 2.2), 2.4)


def c1_poses(n=30):
    return kfx.orbit((1.5, 2.0, 1.5), 1.5, -0.6, -0.3, 0.02, n)


def c1_volume(n=256):
    return kfx.round_buffer(n, 3.0)


def pose_arrays(pose):
    R, t = pose.matrices()
    return R, t


def empty_volume(rb):
    return np.zeros(rb.buffer_length, dtype=kfx.TSDF_DTYPE)


def brick_flags_cpu(vol, rb, brick=16):
    """Exact brick flags of a volume on the CPU: 1 where a brick of local
    voxel coordinates (x fastest) holds a negative value."""
    n = rb.voxels_size.x
    o = np.array([rb.origin_GRID.x, rb.origin_GRID.y, rb.origin_GRID.z])
    v = vol["value"].reshape(n, n, n)  # [z][y][x] memory order
    local = np.roll(v, shift=(-o[2], -o[1], -o[0]), axis=(0, 1, 2))
    nb = n // brick
    neg = (local < 0).reshape(nb, brick, nb, brick, nb, brick).any(axis=(1, 3, 5))
    return neg.astype(np.uint8).reshape(-1)
