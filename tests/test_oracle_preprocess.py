"""Pins the oracle's preprocessing functions against closed forms and brute
force in double precision.  No GPU needed."""
import numpy as np

import scenes


def _bilateral_brute(d, ss=4.5, sr=30.0):
    """Direct double-precision evaluation of the reference window and weights
    (compute_bilateral_filtering.cl:24-53)."""
    H, W = d.shape
    out = np.zeros_like(d)
    for r in range(H):
        for c in range(W):
            v = float(d[r, c])
            ys = np.arange(max(r - 6, 0), min(r + 7, H - 1))
            xs = np.arange(max(c - 6, 0), min(c + 7, W - 1))
            t = d[np.ix_(ys, xs)].astype(np.float64)
            s2 = (c - xs[None, :]) ** 2 + (r - ys[:, None]) ** 2
            w = np.exp(-(s2 * 0.5 / ss ** 2 + (v - t) ** 2 * 0.5 / sr ** 2))
            out[r, c] = min(max(int(np.rint((w * t).sum() / w.sum())), 0), 32767)
    return out


def test_bilateral_constant_image_unchanged(orc):
    d = np.full((30, 40), 1234, np.uint16)
    assert np.array_equal(orc.bilateral(d), d)


def test_bilateral_matches_double_brute_force(orc):
    rng = np.random.default_rng(3)
    d = (1500 + rng.normal(0, 20, size=(24, 32))).astype(np.uint16)
    d[5:9, 10:20] = 2500  # a step that the range kernel must preserve
    d[15, 3] = 0
    got = orc.bilateral(d)
    want = _bilateral_brute(d)
    assert np.max(np.abs(got.astype(int) - want)) <= 1
    assert np.mean(got == want) > 0.99


def test_bilateral_large_sigma_r_is_spatial_gaussian(orc):
    # sigma_r -> infinity: the weights reduce to the spatial Gaussian.
    rng = np.random.default_rng(4)
    d = rng.integers(500, 3000, size=(20, 26)).astype(np.uint16)
    got = orc.bilateral(d, sigma_color=1e9)
    want = _bilateral_brute(d, sr=1e9)
    assert np.max(np.abs(got.astype(int) - want)) <= 1


def test_bilateral_preserves_valid_mask_and_edges(kfx, orc):
    intr = kfx.kinect_intrinsics(160, 120, 131.25)
    scene = scenes.c1_scene()
    pose = kfx.orbit((1.5, 2.0, 1.5), 1.5, -0.6, -0.3, 0.02, 1)[0]
    d = kfx.render(scene, pose, intr, 9, 0, noise=True)
    f = orc.bilateral(d)
    assert np.array_equal(f > 0, d > 0)
    # Far-edge row/column never enter a window (compute_bilateral_filtering.cl:24-25):
    # changing them leaves every other output untouched except through the
    # centre value of that row/column itself.
    d2 = d.copy()
    d2[-1, :] = 4000
    d2[:, -1] = 4000
    f2 = orc.bilateral(d2)
    assert np.array_equal(f2[:-1, :-1], f[:-1, :-1])


def test_bilateral_max_distance(orc):
    d = np.full((16, 16), 800, np.uint16)
    assert np.all(orc.bilateral(d, max_distance=700) == 0)
    assert np.all(orc.bilateral(d, max_distance=800) == 800)


def _pyr_brute(d, sr=30.0):
    H, W = d.shape
    w1 = {0: 0.375, 1: 0.25, 2: 0.0625}
    out = np.zeros((H // 2, W // 2), np.uint16)
    for y in range(H // 2):
        for x in range(W // 2):
            c = int(d[2 * y, 2 * x])
            s = wall = 0.0
            for yy in range(max(0, 2 * y - 2), min(H, 2 * y + 3)):
                for xx in range(max(0, 2 * x - 2), min(W, 2 * x + 3)):
                    v = int(d[yy, xx])
                    if abs(v - c) < 3 * sr:
                        w = w1[abs(xx - 2 * x)] * w1[abs(yy - 2 * y)]
                        s += v * w
                        wall += w
            out[y, x] = int(s / wall)
    return out


def test_pyr_down_closed_forms(orc):
    d = np.full((20, 30), 777, np.uint16)
    assert np.array_equal(orc.pyr_down(d), np.full((10, 15), 777, np.uint16))
    # Step larger than 3 sigma: the two sides never mix.
    s = np.zeros((20, 30), np.uint16)
    s[:, :15] = 1000
    s[:, 15:] = 2000
    p = orc.pyr_down(s)
    assert set(np.unique(p)) == {1000, 2000}
    rng = np.random.default_rng(5)
    r = (1500 + rng.normal(0, 40, size=(22, 34))).astype(np.uint16)
    r[rng.random(r.shape) < 0.05] = 0
    assert np.array_equal(orc.pyr_down(r), _pyr_brute(r))


def test_vertex_map_closed_form(kfx, orc):
    intr = kfx.Intrinsics(100.0, 90.0, 10.5, 7.5, 20, 16)
    rng = np.random.default_rng(6)
    d = rng.integers(0, 5000, size=(16, 20)).astype(np.uint16)
    d[3, 4] = 0
    v = orc.vertex_map(d, intr)
    u, vv = np.meshgrid(np.arange(20), np.arange(16))
    z = d / 1000.0
    want = np.stack([z * (u - 10.5) / 100.0, z * (vv - 7.5) / 90.0, z], -1)
    ok = d > 0
    np.testing.assert_allclose(v[ok], want[ok], rtol=1e-6, atol=1e-9)
    assert np.all(np.isnan(v[~ok]))


def test_normal_map_plane_and_sphere(kfx, orc):
    # A tilted plane n.x = d seen by the camera: every valid normal equals n
    # (oriented toward the camera).
    intr = kfx.Intrinsics(120.0, 120.0, 31.5, 23.5, 64, 48)
    n = np.array([0.2, -0.3, -1.0])
    n /= np.linalg.norm(n)
    u, v = np.meshgrid(np.arange(64), np.arange(48))
    ray = np.stack([(u - 31.5) / 120.0, (v - 23.5) / 120.0, np.ones_like(u, dtype=float)], -1)
    s = -1.5 / (ray @ n)  # plane n.x = -1.5 (camera side has n.x > -1.5)
    vm = (ray * s[..., None]).astype(np.float32)
    nm = orc.normal_map_f64(vm)
    ok = ~np.isnan(nm[..., 0])
    assert ok[3:-4, 3:-4].all()
    np.testing.assert_allclose(nm[ok], np.broadcast_to(n, nm[ok].shape), atol=1e-5)
    # Too few valid neighbours -> NaN.
    vm2 = vm.copy()
    vm2[10:20, 10:20] = np.nan
    vm2[15, 15] = vm[15, 15]
    assert np.isnan(orc.normal_map_f64(vm2)[15, 15, 0])


def test_transform_maps_closed_form(orc):
    rng = np.random.default_rng(7)
    vm = rng.normal(size=(8, 10, 3)).astype(np.float32)
    nm = rng.normal(size=(8, 10, 3)).astype(np.float32)
    vm[2, 3] = np.nan
    nm[4, 4, 0] = np.nan
    a = 0.3
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    t = np.array([0.1, -0.2, 0.3])
    vd, nd = orc.transform_maps(vm, nm, R, t)
    np.testing.assert_allclose(vd[0, 0], R @ vm[0, 0] + t, rtol=1e-5)
    np.testing.assert_allclose(nd[1, 1], R @ nm[1, 1], rtol=1e-5)
    assert np.all(np.isnan(vd[2, 3])) and np.all(np.isnan(nd[4, 4]))
    vi, ni = orc.transform_maps(vm, nm, np.eye(3), np.zeros(3))
    ok = ~np.isnan(vm[..., 0])
    assert np.array_equal(vi[ok], vm[ok])


def test_resize_map_closed_form(orc):
    rng = np.random.default_rng(8)
    m = rng.normal(size=(10, 12, 3)).astype(np.float32)
    m[3, 5] = np.nan
    r = orc.resize_map(m, False)
    want = m.reshape(5, 2, 6, 2, 3).mean(axis=(1, 3))
    ok = ~np.isnan(want[..., 0])
    np.testing.assert_allclose(r[ok], want[ok], rtol=1e-6)
    assert np.isnan(r[1, 2, 0]) and ok.sum() == 29
    rn = orc.resize_map(m, True)
    np.testing.assert_allclose(np.linalg.norm(rn[ok], axis=-1), 1.0, rtol=1e-6)
