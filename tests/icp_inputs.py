"""Synthetic ICP inputs: current-frame maps in camera coordinates and
previous-frame maps in volume coordinates, built from the analytic scene with
the oracle's vertex/normal map functions."""
import numpy as np

import kfx
import scenes

ANGLE_THR = float(np.sin(np.deg2rad(20.0)))  # constants.h:50
DIST_THR = 0.1                                # common_defines.h:28


def camera_maps(orc, depth, intr):
    vm = orc.vertex_map(depth, intr)
    nm = orc.normal_map_f64(vm)
    return vm, nm


def icp_pair(orc, intr, frame_prev=10, frame_cur=11, noise=False, scale=1):
    """(vc, nc) of frame_cur in camera coordinates, (vp, np) of frame_prev in
    volume coordinates, and both ground-truth poses (C1 orbit)."""
    scene = scenes.c1_scene()
    poses = scenes.c1_poses(30)
    Rp, tp = poses[frame_prev].matrices()
    Rc, tc = poses[frame_cur].matrices()
    dp = kfx.render(scene, poses[frame_prev], intr, scenes.C1_SEED, frame_prev, noise=False)
    dc = kfx.render(scene, poses[frame_cur], intr, scenes.C1_SEED, frame_cur, noise=noise)
    vpc, npc = camera_maps(orc, dp, intr)
    vp, np_ = orc.transform_maps(vpc, npc, Rp, tp)
    vc, nc = camera_maps(orc, dc, intr)
    return (vc, nc, vp, np_), (Rp, tp), (Rc, tc)


def pyramid_maps(orc, maps, levels=3):
    """Resize (vc, nc, vp, np) for the coarser levels."""
    out = [maps]
    for _ in range(1, levels):
        vc, nc, vp, np_ = out[-1]
        out.append((orc.resize_map(vc, False), orc.resize_map(nc, True), orc.resize_map(vp, False),
                    orc.resize_map(np_, True)))
    return out


def oracle_icp(orc, pyr, intr, R0, t0, Rprev, tprev, iters=(10, 5, 4), precision=0):
    """Coarse-to-fine ICP in the oracle: returns (R, t)."""
    R, t = np.array(R0, np.float64), np.array(t0, np.float64)
    Rpi = np.asarray(Rprev).T
    for level in (2, 1, 0):
        li = intr.level(level)
        vc, nc, vp, np_ = pyr[level]
        for _ in range(iters[level]):
            s, n = orc.icp_sums(vc, nc, vp, np_, R, t, Rpi, tprev, li, DIST_THR, ANGLE_THR, precision)
            rc, R, t, x = orc.icp_solve_update(s, R, t)
    return R, t
