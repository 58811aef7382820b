"""Oracle of the colour volume (Integrate_Colors, common/src/opencl_wrappers.cpp:181-214;
its kernel integrate_color_volume is absent from the reference, so the rule
is the published KinectFusion colour volume, pinned here by closed forms and
brute force rather than by reference vectors):

- the FP32 running average equals exact round-half-even of the rational
  (c * w + c_new) / (w + 1), exhaustively over every 8-bit input;
- which voxels update, and from which pixel, agrees with an independent
  float64 evaluation of the geometry on a tilted camera and a non-square
  image (a transposed rotation, swapped u/v or a wrong distance test fails);
- channels, weights and the weight cap follow the closed form;
- a cyclic origin only rolls the volume."""
import numpy as np
import pytest

import scenes


def _rne_div(n, d):
    q, r = np.divmod(n, d)
    return q + ((2 * r > d) | ((2 * r == d) & (q % 2 == 1)))


def test_color_average_fp32_is_exact_round_half_even():
    c = np.arange(256, dtype=np.int64)
    for w in range(255):
        n = c[:, None] * w + c[None, :]
        f32 = (np.float32(w) * c[:, None].astype(np.float32) + c[None, :].astype(np.float32)) / np.float32(w + 1)
        assert f32.dtype == np.float32
        assert np.array_equal(np.rint(f32).astype(np.int64), _rne_div(n, w + 1)), w


def _rot(ax, ay):
    cx, sx, cy, sy = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    return ry @ rx


def _wall(kfx, n=40, size=1.2, wall_z=0.75):
    """Tilted camera in front of the plane z = wall_z (volume coordinates);
    the vertex map is the exact ray/plane intersection, rounded to float32."""
    intr = kfx.Intrinsics(60.0, 55.0, 23.5, 16.5, 48, 34)
    rb = kfx.round_buffer(n, size)
    R = _rot(0.12, -0.2)  # camera -> volume
    t = np.array([0.52 * size, 0.47 * size, 0.08])
    u, v = np.meshgrid(np.arange(intr.cols), np.arange(intr.rows))
    d = np.stack([(u - intr.cx) / intr.fx, (v - intr.cy) / intr.fy, np.ones_like(u, dtype=np.float64)], -1) @ R.T
    s = (wall_z - t[2]) / d[..., 2]
    vmap = (t + s[..., None] * d).astype(np.float32)
    return intr, rb, R, t, vmap


def _expected(intr, rb, R, t, vmap, trunc):
    """float64 geometry per voxel (z, y, x order): pixel index (-1 when the
    voxel does not project into the image), whether it updates, and whether
    the decision is too close to a rounding boundary to compare."""
    n = rb.voxels_size.x
    cs = rb.volume_size.x / n
    i = (np.arange(n) + 0.5) * cs
    gz, gy, gx = np.meshgrid(i, i, i, indexing="ij")
    g = np.stack([gx, gy, gz], -1)
    c = (g - t) @ R  # R^-1 (g - t) = R^T (g - t)
    with np.errstate(divide="ignore", invalid="ignore"):
        fu = intr.fx * c[..., 0] / c[..., 2] + intr.cx
        fv = intr.fy * c[..., 1] / c[..., 2] + intr.cy
    front = c[..., 2] > 1e-6
    u, v = np.rint(np.where(front, fu, -1e9)), np.rint(np.where(front, fv, -1e9))
    inside = front & (u >= 0) & (u < intr.cols) & (v >= 0) & (v < intr.rows)
    pix = np.where(inside, v * intr.cols + u, -1).astype(np.int64)
    p = vmap.reshape(-1, 3).astype(np.float64)[np.maximum(pix, 0)]
    dist = np.linalg.norm(p - g, axis=-1)
    upd = inside & (dist < trunc)
    frac_u, frac_v = np.abs(fu - np.floor(fu) - 0.5), np.abs(fv - np.floor(fv) - 0.5)
    ambiguous = front & ((frac_u < 1e-3) | (frac_v < 1e-3) | (np.abs(dist - trunc) < 1e-5))
    return pix, upd, ambiguous


def _cvol(orc, rb):
    return np.zeros(rb.buffer_length, orc.COLOR_DTYPE)


def _frame(orc, intr, r, g, b):
    f = np.zeros((intr.rows, intr.cols), orc.COLOR_DTYPE)
    f["r"], f["g"], f["b"], f["w"] = r, g, b, 1
    return f


def test_colors_update_set_and_pixels_match_float64_geometry(kfx, orc):
    intr, rb, R, t, vmap = _wall(kfx, n=64)
    u, v = np.meshgrid(np.arange(intr.cols), np.arange(intr.rows))
    rgb = _frame(orc, intr, u, v, 7)  # each pixel's colour names the pixel
    vol = _cvol(orc, rb)
    touched = orc.integrate_colors(vol, rgb, vmap, intr, R.T, t, scenes.TRUNC, 1, rb)
    pix, upd, amb = _expected(intr, rb, R, t, vmap, scenes.TRUNC)
    got = vol.reshape(pix.shape)
    sure = ~amb
    assert amb.mean() < 0.01
    assert 0 < touched == np.count_nonzero(got["w"])
    assert upd[sure].sum() > 1000
    assert np.array_equal(got["w"][sure] == 1, upd[sure])
    sel = sure & upd
    assert np.array_equal(got["r"][sel], (pix[sel] % intr.cols).astype(np.uint8))
    assert np.array_equal(got["g"][sel], (pix[sel] // intr.cols).astype(np.uint8))
    assert np.all(got["b"][sel] == 7)
    assert not np.any(got["r"][~upd & sure]) and not np.any(got["w"][~upd & sure])


def test_colors_running_average_channels_and_weight(kfx, orc):
    intr, rb, R, t, vmap = _wall(kfx)
    probe = _cvol(orc, rb)
    orc.integrate_colors(probe, _frame(orc, intr, 1, 1, 1), vmap, intr, R.T, t, scenes.TRUNC, 1, rb)
    upd = probe["w"] == 1
    rng = np.random.default_rng(3)
    vol = _cvol(orc, rb)
    for ch in "rgb":
        vol[ch] = rng.integers(0, 256, vol.size)
    vol["w"] = rng.integers(0, 10, vol.size)
    before = vol.copy()
    orc.integrate_colors(vol, _frame(orc, intr, 201, 99, 30), vmap, intr, R.T, t, scenes.TRUNC, 6, rb)
    assert np.array_equal(vol[~upd], before[~upd])
    w = before["w"][upd].astype(np.int64)
    for ch, new in (("r", 201), ("g", 99), ("b", 30)):
        want = _rne_div(before[ch][upd].astype(np.int64) * w + new, w + 1)
        assert np.array_equal(vol[ch][upd], want), ch
    assert np.array_equal(vol["w"][upd], np.minimum(w + 1, 6))


def test_colors_weight_cap_and_constant_colour(kfx, orc):
    intr, rb, R, t, vmap = _wall(kfx, n=24)
    vol = _cvol(orc, rb)
    rgb = _frame(orc, intr, 12, 240, 77)
    for _ in range(5):
        orc.integrate_colors(vol, rgb, vmap, intr, R.T, t, scenes.TRUNC, 3, rb)
    upd = vol["w"] > 0
    assert upd.sum() > 100
    assert np.all(vol["w"][upd] == 3)
    assert np.all(vol["r"][upd] == 12) and np.all(vol["g"][upd] == 240) and np.all(vol["b"][upd] == 77)


def test_colors_nan_vertices_and_far_planes_update_nothing(kfx, orc):
    intr, rb, R, t, vmap = _wall(kfx, n=24)
    vol = _cvol(orc, rb)
    rgb = _frame(orc, intr, 1, 2, 3)
    assert orc.integrate_colors(vol, rgb, np.full_like(vmap, np.nan), intr, R.T, t, scenes.TRUNC, 1, rb) == 0
    far = vmap.copy()
    far[..., 2] += 5.0  # every vertex far behind the volume
    assert orc.integrate_colors(vol, rgb, far, intr, R.T, t, scenes.TRUNC, 1, rb) == 0
    assert not np.any(vol.view(np.uint32))


def test_colors_cyclic_origin_rolls_the_volume(kfx, orc):
    intr, rb0, R, t, vmap = _wall(kfx, n=32)
    o = (5, 11, 29)
    rb1 = kfx.round_buffer(32, 1.2, origin=o)
    u, v = np.meshgrid(np.arange(intr.cols), np.arange(intr.rows))
    rgb = _frame(orc, intr, u, v, 5)
    a, b = _cvol(orc, rb0), _cvol(orc, rb1)
    orc.integrate_colors(a, rgb, vmap, intr, R.T, t, scenes.TRUNC, 1, rb0)
    orc.integrate_colors(b, rgb, vmap, intr, R.T, t, scenes.TRUNC, 1, rb1)
    assert np.count_nonzero(a["w"]) > 0
    rolled = np.roll(a.reshape(32, 32, 32), (o[2], o[1], o[0]), axis=(0, 1, 2))
    assert np.array_equal(b.reshape(32, 32, 32), rolled)


@pytest.mark.parametrize("frame", [0, 12])
def test_synth_rgb_albedo(kfx, frame):
    intr = kfx.kinect_intrinsics(160, 120, 131.25)
    scene, poses = scenes.c1_scene(), scenes.c1_poses(30)
    rgb = kfx.render_rgb(scene, poses[frame], intr)
    _, clean = kfx.render(scene, poses[frame], intr, 1, frame, noise=False, want_clean=True)
    assert np.array_equal(rgb, kfx.render_rgb(scene, poses[frame], intr, threads=1))
    assert np.all(rgb["w"] == 1)
    hit = clean > 0
    assert not np.any(rgb["r"][~hit] | rgb["g"][~hit] | rgb["b"][~hit])
    grey = (rgb["r"] == rgb["g"]) & (rgb["g"] == rgb["b"])
    assert np.all(np.isin(rgb["r"][hit & grey], [60, 200]))
    assert np.count_nonzero(hit & ~grey) > 100  # objects are coloured
