"""GPU parity of the denoiser kernels against the double-precision oracle
(with bfloat16 storage emulated): forward output, the loss and every layer's
gradient; Adam lowers the loss; the consistency gate and model-depth helpers
against numpy."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NP = 5089
LAYERS = {"w1": (0, 288), "b1": (288, 304), "w2": (304, 2608), "b2": (2608, 2624), "w3": (2624, 4928),
          "b3": (4928, 4944), "w4": (4944, 5088), "b4": (5088, 5089)}


def _frame(rng, W, H):
    yy, xx = np.mgrid[0:H, 0:W]
    clean = 1.2 + 0.3 * np.sin(xx / 9.0) + 0.002 * yy
    d = np.rint((clean + rng.normal(0, 0.006, size=(H, W))) * 1000).astype(np.uint16)
    d[rng.random((H, W)) < 0.03] = 0
    t = clean.astype(np.float32)
    t[rng.random((H, W)) < 0.02] = 0.0
    return d, t


def _params(rng, scale=0.08):
    p = rng.normal(0, scale, size=NP).astype(np.float32)
    p[4944:5089] *= 0.1  # small output layer, as trained nets have
    return p


@pytest.fixture
def dn(gpu):
    h = C.c_void_p()
    gpu.call("kfx_denoiser_create", 72, 40, C.byref(h))
    yield h
    gpu.call("kfx_denoiser_destroy", h)


def test_param_count(gpu):
    assert gpu.lib.kfx_denoiser_param_count() == NP


def test_forward_matches_oracle(gpu, orc, dn):
    # W = 72 is not a multiple of the 16-pixel MFMA tile.
    kfx = gpu
    rng = np.random.default_rng(0)
    d, _ = _frame(rng, 72, 40)
    p = _params(rng)
    kfx.call("kfx_denoiser_set_params", dn, p.ctypes.data_as(C.c_void_p))
    dd = kfx.Buffer.from_array(d)
    ym = kfx.Buffer(d.size * 4)
    out = kfx.Buffer(d.size * 2)
    kfx.call("kfx_denoiser_forward", dn, dd.ptr, out.ptr, ym.ptr)
    y = ym.download(np.float32, d.shape)
    yo = orc.dn_forward(d, p, bf16=True)
    err = np.abs(y - yo)
    assert np.all(y[d == 0] == 0)
    assert np.median(err) < 1e-6 and np.max(err) < 5e-4, (np.median(err), np.max(err))
    mm = out.download(np.uint16, d.shape)
    assert np.array_equal(mm > 0, d > 0)  # the valid mask is preserved exactly
    ok = d > 0
    assert np.max(np.abs(mm[ok].astype(np.float64) - np.rint(yo[ok] * 1000))) <= 1


@pytest.mark.parametrize("size", [(72, 40), (640, 480), (130, 33)])
def test_fused_forward_equals_layer_by_layer(gpu, size):
    # The one-launch inference kernel (tiles of 32 x 16 with halos in LDS)
    # against the four-launch forward: identical bits, including the image
    # borders and sizes that are not multiples of the tile.
    kfx = gpu
    W, H = size
    rng = np.random.default_rng(W + H)
    d, _ = _frame(rng, W, H)
    h = C.c_void_p()
    kfx.call("kfx_denoiser_create", W, H, C.byref(h))
    for params in (_params(rng), kfx.Denoiser.load_params()):
        kfx.call("kfx_denoiser_set_params", h, np.ascontiguousarray(params, np.float32).ctypes.data_as(C.c_void_p))
        dd = kfx.Buffer.from_array(d)
        outs = []
        for name in ("kfx_denoiser_forward", "kfx_denoiser_forward_layers"):
            ym, mm = kfx.Buffer(d.size * 4), kfx.Buffer(d.size * 2)
            kfx.call(name, h, dd.ptr, mm.ptr, ym.ptr)
            outs.append((ym.download(np.float32, d.shape), mm.download(np.uint16, d.shape)))
        assert np.array_equal(outs[0][0].view(np.uint32), outs[1][0].view(np.uint32))
        assert np.array_equal(outs[0][1], outs[1][1])
        # The network ran: the output differs from the input depth in metres.
        assert np.count_nonzero(np.abs(outs[0][0] - d.astype(np.float32) * 1e-3) > 1e-6) > 0
    kfx.call("kfx_denoiser_destroy", h)


def test_zero_params_identity(gpu, dn):
    kfx = gpu
    rng = np.random.default_rng(1)
    d, _ = _frame(rng, 72, 40)
    kfx.call("kfx_denoiser_set_params", dn, np.zeros(NP, np.float32).ctypes.data_as(C.c_void_p))
    dd = kfx.Buffer.from_array(d)
    out = kfx.Buffer(d.size * 2)
    kfx.call("kfx_denoiser_forward", dn, dd.ptr, out.ptr, None)
    assert np.array_equal(out.download(np.uint16, d.shape), d)


def test_gradients_match_oracle(gpu, orc, dn):
    kfx = gpu
    rng = np.random.default_rng(2)
    d, t = _frame(rng, 72, 40)
    p = _params(rng)
    kfx.call("kfx_denoiser_set_params", dn, p.ctypes.data_as(C.c_void_p))
    dd, dt = kfx.Buffer.from_array(d), kfx.Buffer.from_array(t)
    loss = C.c_double()
    cnt = C.c_uint64()
    kfx.call("kfx_denoiser_accumulate", dn, dd.ptr, dt.ptr, 0.05, C.byref(loss), C.byref(cnt))
    g = np.zeros(NP, np.float32)
    kfx.call("kfx_denoiser_get_grad", dn, g.ctypes.data_as(C.c_void_p))
    lo, go, no = orc.dn_loss_grad(d, t, p, bf16=True, outlier=0.05)
    assert cnt.value == no
    assert abs(loss.value - lo) <= 1e-3 * lo
    for name, (a, b) in LAYERS.items():
        rel = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
        assert rel < 2e-3, (name, rel)


def test_adam_lowers_loss(gpu, dn):
    kfx = gpu
    rng = np.random.default_rng(3)
    frames = [_frame(rng, 72, 40) for _ in range(4)]
    bufs = [(kfx.Buffer.from_array(d), kfx.Buffer.from_array(t)) for d, t in frames]
    kfx.call("kfx_denoiser_set_params", dn, _params(rng, 0.02).ctypes.data_as(C.c_void_p))
    losses = []
    for step in range(60):
        tot = 0.0
        for dd, dt in bufs:
            loss = C.c_double()
            kfx.call("kfx_denoiser_accumulate", dn, dd.ptr, dt.ptr, 0.05, C.byref(loss), None)
            tot += loss.value
        kfx.call("kfx_denoiser_adam_step", dn, 2e-3, len(bufs))
        losses.append(tot / len(bufs))
    # The input noise (6 mm) sets a floor of ~3.6e-5 m^2 for the identity map;
    # 60 steps from a small random start take the loss well below the start.
    assert losses[-1] < 0.9 * losses[0], losses[::10]
    assert all(b <= a * 1.05 for a, b in zip(losses[::10], losses[10::10])), losses[::10]


def test_model_depth_and_consistency(gpu):
    kfx = gpu
    rng = np.random.default_rng(4)
    intr = kfx.kinect_intrinsics(64, 48, 52.5)
    # vertices of a wall at z = 2 m in front of a camera at the identity pose
    u, v = np.meshgrid(np.arange(64), np.arange(48))
    z = np.full((48, 64), 2.0)
    vm = np.stack([(u - intr.cx) / intr.fx * z, (v - intr.cy) / intr.fy * z, z], -1).astype(np.float32)
    vm[5, 5] = np.nan
    pose = kfx.Pose()
    pose.R[:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    pose.t[:] = [0, 0, 0]
    dvm = kfx.Buffer.from_array(vm)
    out = kfx.Buffer(64 * 48 * 4)
    kfx.call("kfx_model_depth", dvm.ptr, C.byref(pose), 64, 48, out.ptr)
    md = out.download(np.float32, (48, 64))
    assert md[5, 5] == 0 and np.allclose(np.delete(md.ravel(), 5 * 64 + 5), 2.0)
    depth = np.full((48, 64), 2010, np.uint16)
    depth[0, :] = 0
    dd = kfx.Buffer.from_array(depth)
    dpose = kfx.Buffer.from_array(np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32))
    st = kfx.Buffer(32)
    kfx.call("kfx_depth_consistency", dvm.ptr, dpose.ptr, dd.ptr, C.byref(intr), 0.1, st.ptr)
    s = st.download(np.float64, 4)
    assert s[1] == 64 * 48 - 1 - 64 and s[2] == s[1] and s[3] == 64 * 48 - 1
    np.testing.assert_allclose(np.sqrt(s[0] / s[1]), 0.010, rtol=1e-3)
    # 20 cm away: outside the 10 cm window, still on pixels with depth.
    far = kfx.Buffer.from_array(np.full((48, 64), 2200, np.uint16))
    kfx.call("kfx_depth_consistency", dvm.ptr, dpose.ptr, far.ptr, C.byref(intr), 0.1, st.ptr)
    s = st.download(np.float64, 4)
    assert s[0] == 0 and s[1] == 0 and s[2] == 64 * 48 - 1 and s[3] == 64 * 48 - 1


def test_trained_denoiser_gains_over_bilateral(gpu):
    # The committed network on frames of held-out scenes (seeds never used in
    # training): depth RMSE against the renderer's exact depth, on pixels
    # valid in the raw frame within 5 cm of the truth, for the raw frame, the
    # bilateral filter (sigma_s 4.5 px, sigma_r 30 mm) and the network.
    kfx = gpu
    import workload as wl
    W, H = wl.W, wl.H
    dn = kfx.Denoiser(W, H, kfx.Denoiser.load_params())
    src, bil, den = kfx.Buffer(W * H * 2), kfx.Buffer(W * H * 2), kfx.Buffer(W * H * 2)
    sums = {"raw": 0.0, "bilateral": 0.0, "denoised": 0.0}
    count = 0
    for seed in (5000, 5003, 5007):
        poses = wl.trajectory(seed, 60)
        sc = wl.scene(seed)
        for i in (5, 30, 55):
            raw, clean = kfx.render(sc, poses[i], wl.intrinsics(), seed, i, noise=True, want_clean=True)
            src.upload(raw)
            kfx.call("kfx_bilateral_filter", src.ptr, bil.ptr, W, H, 4.5, 30.0, 0xFFFFFFFF)
            kfx.call("kfx_denoiser_forward", dn.h, src.ptr, den.ptr, None)
            r = raw.astype(np.float64)
            ok = (raw > 0) & (clean > 0) & (np.abs(r - clean * 1000) < 50)
            gt = clean[ok] * 1000.0
            sums["raw"] += np.sum((r[ok] - gt) ** 2)
            sums["bilateral"] += np.sum((bil.download(np.uint16, (H, W)).astype(np.float64)[ok] - gt) ** 2)
            sums["denoised"] += np.sum((den.download(np.uint16, (H, W)).astype(np.float64)[ok] - gt) ** 2)
            count += int(ok.sum())
    rmse = {k: float(np.sqrt(v / count)) for k, v in sums.items()}
    print("held-out depth RMSE (mm):", rmse)
    dn.close()
    assert rmse["denoised"] < 0.6 * rmse["bilateral"] < 0.6 * rmse["raw"], rmse
