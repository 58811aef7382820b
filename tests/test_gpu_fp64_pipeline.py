"""The GPU pipeline against the oracle's double-precision pipeline
(tests/fp64_pipeline.py) on the first frames of the 512^3 walk: on the
pipeline's own inputs the FP64 ICP lands within 1e-5 m / 1e-5 rad of the
GPU's pose per frame, and the FP64 integration at the GPU's poses leaves
the volume within 1 quantisation step on at least 99.99 % of the observed
voxels; run on its own, the oracle pipeline tracks the same trajectory
(shifts and integration decisions) within a millimetre.  The full runs
(300 and 1000 frames) are tests/fp64_pipeline_study.py."""
import ctypes as C

import numpy as np
import pytest

import fp64_pipeline as fp
import fp64_pipeline_study as study
import workload as wl

pytestmark = pytest.mark.gpu

FRAMES = 40


@pytest.fixture(scope="module")
def walk(gpu):
    intr, _scene, poses, origin, depths = wl.sequence(wl.SCENE_SEED, FRAMES)
    return intr, poses, origin, depths


def test_open_loop_icp_and_tsdf_against_fp64(gpu, orc, walk):
    intr, poses, origin, depths = walk
    o = study.open_loop(intr, poses, origin, depths, FRAMES, 1.5)
    dt, dr = np.array(o["per_frame_trans_err_m"]), np.array(o["per_frame_rot_err_rad"])
    t = o["tsdf_vs_fp64_at_same_poses"]
    print(f"ICP vs FP64: median {np.median(dt):.2e} m {np.median(dr):.2e} rad, max {dt.max():.2e} m "
          f"{dr.max():.2e} rad; TSDF {100 * t['within_1_lsb']:.4f} % within 1 LSB")
    assert len(dt) == FRAMES - 1
    assert np.mean((dt <= 1e-5) & (dr <= 1e-5)) >= 0.95
    assert np.median(dt) < 1e-6 and np.median(dr) < 1e-6
    assert dt.max() < 1e-4 and dr.max() < 1e-4
    assert t["within_1_lsb"] >= 0.9999 and t["voxels_observed"] > 10_000_000


def test_closed_loop_follows_fp64_pipeline(gpu, orc, walk):
    kfx = gpu
    intr, poses, origin, depths = walk
    dd = kfx.Buffer.from_array(depths)
    gposes, gshifts, ginteg, _vol = study.gpu_run(kfx.ICP_SUM_FP32_KAHAN, intr, poses, origin, dd,
                                                  depths[0].nbytes, FRAMES, 1.5)
    op = fp.OraclePipeline(orc, kfx, intr, wl.N_VOX, wl.VOL_M, poses[0], origin, trunc=wl.TRUNC)
    res = [op.process(depths[i]) for i in range(FRAMES)]
    dt, dr = fp.pose_errors(gposes, [(r[0], r[1]) for r in res])
    print(f"trajectory difference max {dt.max():.2e} m {dr.max():.2e} rad")
    assert [int(r[2]) for r in res] == ginteg
    assert [i for i, r in enumerate(res) if r[3] is not None] == gshifts
    assert dt.max() < 1e-3 and dr.max() < 1e-3
