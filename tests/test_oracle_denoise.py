"""Pins the denoiser oracle: zero parameters give the identity, the
analytic gradient matches central finite differences of the double-precision
loss, and the bf16 emulation stays close to double.  No GPU needed."""
import numpy as np
import pytest


def _frame(rng, W=12, H=10):
    d = rng.integers(800, 1400, size=(H, W)).astype(np.uint16)
    d[rng.random((H, W)) < 0.1] = 0
    t = (d.astype(np.float64) * 1e-3 + rng.normal(0, 0.004, size=(H, W))).astype(np.float32)
    t[rng.random((H, W)) < 0.05] = 0.0
    return d, t


def _params(rng, n, scale=0.2):
    return (rng.normal(0, scale, size=n)).astype(np.float32)


def test_zero_params_is_identity(orc):
    rng = np.random.default_rng(0)
    d, _ = _frame(rng)
    y = orc.dn_forward(d, np.zeros(orc.lib.orc_dn_param_count(), np.float32), bf16=False)
    # y = 1 / (1000 / d): the depth up to the rounding of the inverse, and
    # exactly the input in millimetres.
    assert np.array_equal(y > 0, d > 0)
    np.testing.assert_allclose(y[d > 0], d[d > 0] * 1e-3, rtol=2e-7)
    assert np.array_equal(np.rint(y * 1000.0).astype(np.int64), d.astype(np.int64))


def test_param_count(orc):
    assert orc.lib.orc_dn_param_count() == 2 * 16 * 9 + 16 + 2 * (16 * 16 * 9 + 16) + 16 * 9 + 1 == 5089


@pytest.mark.parametrize("seed", [1, 2])
def test_gradient_matches_finite_differences(orc, seed):
    rng = np.random.default_rng(seed)
    d, t = _frame(rng)
    n = orc.lib.orc_dn_param_count()
    p = _params(rng, n)
    loss, g, cnt = orc.dn_loss_grad(d, t, p, bf16=False, outlier=1.0)
    assert cnt > 50
    # Parameters are fp32, so perturb by representable steps and divide by the
    # actual difference; check a sample from every layer and every bias block.
    idx = list(rng.choice(288, 8, replace=False)) + [288, 295] + list(304 + rng.choice(2304, 8, replace=False)) + \
        [2608, 2620] + list(2624 + rng.choice(2304, 8, replace=False)) + [4930] + \
        list(4944 + rng.choice(144, 8, replace=False)) + [5088]
    for i in idx:
        h = np.float32(2e-5) * max(1.0, abs(float(p[i])))
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = orc.dn_loss_grad(d, t, pp, bf16=False, outlier=1.0)[0]
        lm = orc.dn_loss_grad(d, t, pm, bf16=False, outlier=1.0)[0]
        fd = (lp - lm) / (float(pp[i]) - float(pm[i]))
        assert abs(fd - g[i]) <= 1e-4 * max(1e-3, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_emulation_is_close_to_double(orc):
    rng = np.random.default_rng(3)
    d, t = _frame(rng, 20, 16)
    p = _params(rng, orc.lib.orc_dn_param_count(), 0.1)
    y64 = orc.dn_forward(d, p, bf16=False)
    y16 = orc.dn_forward(d, p, bf16=True)
    assert np.max(np.abs(y64 - y16)) < 0.05 * np.max(np.abs(y64 - d * 1e-3)) + 1e-6
