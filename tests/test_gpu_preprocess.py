"""GPU parity of the preprocessing kernels against the oracle on C1 frames:
bilateral (within 1 mm, valid mask bit-exact), pyramid, vertex maps,
transform and resize (bit-exact), normal maps (mask bit-exact, direction
against the double-precision oracle)."""
import ctypes as C

import numpy as np
import pytest

import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c1_frame(gpu):
    kfx = gpu
    intr = kfx.kinect_intrinsics()
    pose = scenes.c1_poses(30)[12]
    return intr, kfx.render(scenes.c1_scene(), pose, intr, scenes.C1_SEED, 12, noise=True)


def test_bilateral_matches_oracle(gpu, orc, c1_frame):
    kfx = gpu
    intr, d = c1_frame
    src = kfx.Buffer.from_array(d)
    dst = kfx.Buffer(d.nbytes)
    kfx.call("kfx_bilateral_filter", src.ptr, dst.ptr, intr.cols, intr.rows, 4.5, 30.0, 0xFFFFFFFF)
    got = dst.download(np.uint16, d.shape)
    want = orc.bilateral(d)
    assert np.array_equal(got > 0, want > 0)
    assert np.array_equal(got > 0, d > 0)
    diff = np.abs(got.astype(int) - want)
    assert diff.max() <= 1
    assert np.mean(diff == 0) > 0.999


def test_pyramid_vertex_maps_bit_exact(gpu, orc, c1_frame):
    kfx = gpu
    intr, d = c1_frame
    lvl = [d]
    for i in range(1, 3):
        lvl.append(orc.pyr_down(lvl[-1]))
    src = kfx.Buffer.from_array(d)
    cur = src
    for i in range(1, 3):
        h, w = lvl[i - 1].shape
        nxt = kfx.Buffer(lvl[i].nbytes)
        kfx.call("kfx_pyr_down", cur.ptr, w, h, nxt.ptr, 30.0)
        got = nxt.download(np.uint16, lvl[i].shape)
        assert np.array_equal(got, lvl[i]), f"pyramid level {i}"
        cur = nxt
    for i in range(3):
        li = intr.level(i)
        db = kfx.Buffer.from_array(lvl[i])
        vb = kfx.Buffer(lvl[i].size * 12)
        kfx.call("kfx_vertex_map", db.ptr, vb.ptr, C.byref(li))
        got = vb.download(np.float32, lvl[i].shape + (3,))
        want = orc.vertex_map(lvl[i], li)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        assert np.array_equal(got[ok], want[ok]), f"vertex map level {i}"


def _eigen_gap(vm):
    """Per pixel (lambda_1 - lambda_0) / lambda_2 of the 7x7 neighbourhood
    covariance the normal is the smallest eigenvector of (window rows
    [v-3, min(v+4, rows-1)), columns likewise), in double precision."""
    H, W, _ = vm.shape
    pad = np.full((H + 6, W + 6, 3), np.nan)
    pad[3:H + 2, 3:W + 2] = vm[:H - 1, :W - 1]  # the last row/column is outside every window
    win = [pad[3 + dy:3 + dy + H, 3 + dx:3 + dx + W] for dy in range(-3, 4) for dx in range(-3, 4)]
    cnt = sum((~np.isnan(p[..., 0])).astype(np.float64) for p in win)
    mean = sum(np.nan_to_num(p) for p in win) / np.maximum(cnt, 1)[..., None]
    cov = np.zeros((H, W, 3, 3))
    for p in win:
        d = np.nan_to_num(p - mean)
        cov += d[..., :, None] * d[..., None, :]
    lam = np.linalg.eigvalsh(cov)
    return (lam[..., 1] - lam[..., 0]) / np.maximum(lam[..., 2], 1e-300)


def test_normal_map_matches_double_oracle(gpu, orc, c1_frame):
    kfx = gpu
    intr, d = c1_frame
    vm = orc.vertex_map(orc.bilateral(d), intr)
    vb = kfx.Buffer.from_array(vm)
    nb = kfx.Buffer(vm.nbytes)
    kfx.call("kfx_normal_map", vb.ptr, nb.ptr, intr.cols, intr.rows)
    got = nb.download(np.float32, vm.shape)
    want = orc.normal_map_f64(vm)
    assert np.array_equal(np.isnan(got[..., 0]), np.isnan(want[..., 0]))
    ok = ~np.isnan(want[..., 0])
    assert ok.mean() > 0.5
    cosang = np.sum(got[ok] * want[ok], axis=-1)
    # The direction is ill-conditioned where the two smallest eigenvalues of
    # the neighbourhood covariance nearly coincide (points close to a line),
    # and the orientation toward the camera where the normal is nearly
    # perpendicular to the viewing ray: those pixels are compared in bulk only.
    gap = _eigen_gap(vm)[ok]
    view = vm[ok] / np.linalg.norm(vm[ok], axis=-1, keepdims=True)
    grazing = np.abs(np.sum(want[ok] * view, axis=-1)) < 0.05
    wellcond = gap > 1e-2
    assert wellcond.mean() > 0.95
    assert np.all(np.abs(cosang[wellcond]) > 0.99), np.min(np.abs(cosang[wellcond]))
    assert np.all(cosang[wellcond & ~grazing] > 0.99), np.min(cosang[wellcond & ~grazing])
    # within 4.5 mrad on 99.9 % of the well-conditioned pixels and 95 % of all
    assert np.mean(cosang[wellcond] > 1 - 1e-5) > 0.999, np.mean(cosang[wellcond] > 1 - 1e-5)
    assert np.mean(cosang > 1 - 1e-5) > 0.95, np.mean(cosang > 1 - 1e-5)


def test_transform_and_resize_bit_exact(gpu, orc, c1_frame):
    kfx = gpu
    intr, d = c1_frame
    vm = orc.vertex_map(d, intr)
    nm = orc.normal_map_f64(vm)
    R = scenes.c1_poses(30)[5].matrices()[0]
    t = np.array([0.3, -0.1, 1.2])
    vb, nb = kfx.Buffer.from_array(vm), kfx.Buffer.from_array(nm)
    vo, no = kfx.Buffer(vm.nbytes), kfx.Buffer(vm.nbytes)
    kfx.call("kfx_transform_maps", vb.ptr, nb.ptr, C.byref(kfx.matrix3f(R)), C.byref(kfx.point3f(t)),
             intr.cols, intr.rows, vo.ptr, no.ptr)
    wv, wn = orc.transform_maps(vm, nm, R, t)
    for buf, want in ((vo, wv), (no, wn)):
        got = buf.download(np.float32, vm.shape)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        assert np.array_equal(got[ok], want[ok])
    for normalize, src, m in ((0, vb, vm), (1, nb, nm)):
        out = kfx.Buffer(vm.nbytes // 4)
        kfx.call("kfx_resize_map", src.ptr, intr.cols, intr.rows, out.ptr, normalize)
        got = out.download(np.float32, (intr.rows // 2, intr.cols // 2, 3))
        want = orc.resize_map(m, normalize)
        assert np.array_equal(np.isnan(got), np.isnan(want))
        ok = ~np.isnan(want)
        assert np.array_equal(got[ok], want[ok])


def test_maps_pyramid_equals_per_level_kernels(gpu, orc, c1_frame):
    # One launch for every level's vertex and normal maps: the same bits as
    # kfx_vertex_map + kfx_normal_map level by level (vertices also against
    # the oracle).
    kfx = gpu
    intr, d = c1_frame
    lvl = [d]
    for i in range(1, 3):
        lvl.append(orc.pyr_down(lvl[-1]))
    intrs = (kfx.Intrinsics * 3)(*[intr.level(i) for i in range(3)])
    db = [kfx.Buffer.from_array(x) for x in lvl]
    vb = [kfx.Buffer(x.size * 12) for x in lvl]
    nb = [kfx.Buffer(x.size * 12) for x in lvl]
    ptrs = lambda bufs: (C.c_void_p * 3)(*[b.ptr.value for b in bufs])
    kfx.call("kfx_maps_pyramid", ptrs(db), ptrs(vb), ptrs(nb), intrs, 3)
    for i in range(3):
        shape = lvl[i].shape + (3,)
        rv, rn = kfx.Buffer(lvl[i].size * 12), kfx.Buffer(lvl[i].size * 12)
        kfx.call("kfx_vertex_map", db[i].ptr, rv.ptr, C.byref(intrs[i]))
        kfx.call("kfx_normal_map", rv.ptr, rn.ptr, intrs[i].cols, intrs[i].rows)
        for got_b, want_b in ((vb[i], rv), (nb[i], rn)):
            got, want = got_b.download(np.float32, shape), want_b.download(np.float32, shape)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"level {i}"
        want_v = orc.vertex_map(lvl[i], intrs[i])
        got_v = vb[i].download(np.float32, shape)
        assert np.array_equal(np.isnan(got_v), np.isnan(want_v))
        ok = ~np.isnan(want_v)
        assert np.array_equal(got_v[ok], want_v[ok])


@pytest.mark.parametrize("sigma", [30.0, 29.7, 1.0 / 3.0, 5.5, 1e6])
def test_pyr_down_range_threshold_and_full_range_depth(gpu, orc, c1_frame, sigma):
    # The kernel's integer range test and sums against the oracle's float
    # ones: non-integer, exactly integer (3 sigma = 1) and unbounded
    # thresholds, on the frame with a block of full-range uint16 values
    # (sums up to 65535 x 256, the largest the exact-sum argument allows).
    kfx = gpu
    intr, d = c1_frame
    d = d.copy()
    rng = np.random.default_rng(7)
    d[100:220, 200:360] = rng.integers(0, 65536, size=(120, 160), dtype=np.uint16)
    d[300:340, 40:120] = 65535
    d[400:410, 500:620] = rng.integers(0, 3, size=(10, 120), dtype=np.uint16)
    h, w = d.shape
    src = kfx.Buffer.from_array(d)
    dst = kfx.Buffer(2 * (h // 2) * (w // 2))
    kfx.call("kfx_pyr_down", src.ptr, w, h, dst.ptr, C.c_float(sigma))
    got = dst.download(np.uint16, (h // 2, w // 2))
    want = orc.pyr_down(d, float(np.float32(sigma)))
    assert np.array_equal(got, want)


@pytest.mark.parametrize("crop,levels", [((480, 640), 3), ((478, 634), 3), ((101, 203), 3), ((480, 640), 2),
                                         ((480, 640), 1)])
def test_maps_pyramid_from_depth_equals_pyr_down_then_maps(gpu, orc, c1_frame, crop, levels):
    # Level 1 and 2 depth derived inside the maps launch: the same bits as
    # kfx_pyr_down level by level then kfx_maps_pyramid, on the full frame,
    # on crops whose levels are ragged against the 32 x 8 tiles, and with
    # fewer levels; the derived depth also against the oracle's pyramid.
    kfx = gpu
    intr0, d = c1_frame
    d = np.ascontiguousarray(d[:crop[0], :crop[1]])
    base = kfx.Intrinsics(intr0.fx, intr0.fy, intr0.cx, intr0.cy, crop[1], crop[0])
    intrs = (kfx.Intrinsics * 3)(*[base.level(i) for i in range(3)])
    sc = 30.0
    shapes = [(intrs[i].rows, intrs[i].cols) for i in range(3)]
    d0 = kfx.Buffer.from_array(d)
    db = [d0] + [kfx.Buffer(2 * r * c) for r, c in shapes[1:]]
    for i in range(1, levels):
        kfx.call("kfx_pyr_down", db[i - 1].ptr, intrs[i - 1].cols, intrs[i - 1].rows, db[i].ptr, C.c_float(sc))
    ptrs = lambda bufs: (C.c_void_p * 3)(*[b.ptr.value if b is not None else None for b in bufs])
    mk = lambda: [kfx.Buffer(12 * r * c) for r, c in shapes]
    vw, nw, vg, ng = mk(), mk(), mk(), mk()
    kfx.call("kfx_maps_pyramid", ptrs(db), ptrs(vw), ptrs(nw), intrs, levels)
    dout = [None] + [kfx.Buffer(2 * r * c) for r, c in shapes[1:]]
    kfx.call("kfx_maps_pyramid_from_depth", d0.ptr, ptrs(dout), ptrs(vg), ptrs(ng), intrs, levels, C.c_float(sc))
    want_d = [d]
    for i in range(1, levels):
        want_d.append(orc.pyr_down(want_d[-1], sc))
        got = dout[i].download(np.uint16, shapes[i])
        assert np.array_equal(got, db[i].download(np.uint16, shapes[i])), f"depth level {i}"
        assert np.array_equal(got, want_d[i]), f"depth level {i} vs oracle"
    for i in range(levels):
        for got_b, want_b in ((vg[i], vw[i]), (ng[i], nw[i])):
            got = got_b.download(np.float32, shapes[i] + (3,))
            want = want_b.download(np.float32, shapes[i] + (3,))
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"level {i} {crop}"
    # Without depth outputs the maps are the same.
    vg2, ng2 = mk(), mk()
    kfx.call("kfx_maps_pyramid_from_depth", d0.ptr, None, ptrs(vg2), ptrs(ng2), intrs, levels, C.c_float(sc))
    for i in range(levels):
        for a, b in ((vg[i], vg2[i]), (ng[i], ng2[i])):
            x = a.download(np.float32, shapes[i] + (3,))
            y = b.download(np.float32, shapes[i] + (3,))
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32))


def test_maps_pyramid_from_depth_rejects_unhalved_levels(gpu, c1_frame):
    kfx = gpu
    intr, d = c1_frame
    lv = [intr.level(i) for i in range(3)]
    lv[2] = kfx.Intrinsics(lv[2].fx, lv[2].fy, lv[2].cx, lv[2].cy, lv[2].cols + 1, lv[2].rows)
    intrs = (kfx.Intrinsics * 3)(*lv)
    d0 = kfx.Buffer.from_array(d)
    bufs = [kfx.Buffer(12 * x.cols * x.rows) for x in lv for _ in range(2)]
    vm = (C.c_void_p * 3)(*[b.ptr.value for b in bufs[0::2]])
    nm = (C.c_void_p * 3)(*[b.ptr.value for b in bufs[1::2]])
    with pytest.raises(kfx.KfxError, match="not half"):
        kfx.call("kfx_maps_pyramid_from_depth", d0.ptr, None, vm, nm, intrs, 3, C.c_float(30.0))


def test_resize_pyramid_equals_two_resizes(gpu, orc, c1_frame):
    kfx = gpu
    intr, d = c1_frame
    vm = orc.vertex_map(d, intr)
    nm = orc.normal_map_f64(vm)
    W, H = intr.cols, intr.rows
    vb, nb = kfx.Buffer.from_array(vm), kfx.Buffer.from_array(nm)
    v1, n1 = kfx.Buffer(vm.nbytes // 4), kfx.Buffer(vm.nbytes // 4)
    v2, n2 = kfx.Buffer(vm.nbytes // 16), kfx.Buffer(vm.nbytes // 16)
    kfx.call("kfx_resize_pyramid", vb.ptr, nb.ptr, W, H, v1.ptr, n1.ptr, v2.ptr, n2.ptr)
    for src, normalize, o1, o2, m in ((vb, 0, v1, v2, vm), (nb, 1, n1, n2, nm)):
        a, b = kfx.Buffer(vm.nbytes // 4), kfx.Buffer(vm.nbytes // 16)
        kfx.call("kfx_resize_map", src.ptr, W, H, a.ptr, normalize)
        kfx.call("kfx_resize_map", a.ptr, W // 2, H // 2, b.ptr, normalize)
        for got_b, want_b, shape in ((o1, a, (H // 2, W // 2, 3)), (o2, b, (H // 4, W // 4, 3))):
            got, want = got_b.download(np.float32, shape), want_b.download(np.float32, shape)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        w2 = orc.resize_map(orc.resize_map(m, normalize), normalize)
        g2 = o2.download(np.float32, (H // 4, W // 4, 3))
        ok = ~np.isnan(w2)
        assert np.array_equal(np.isnan(g2), np.isnan(w2)) and np.array_equal(g2[ok], w2[ok])

