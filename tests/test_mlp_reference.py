"""The CPU network reference used to check the GPU gradients is itself pinned
by central finite differences in float64 (no bf16 rounding)."""
import numpy as np

from mlp_reference import loss_grad, shapes, bf16


def test_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    shp = shapes(6, 16, 3, 5)
    npar = sum(n * k + n for n, k in shp)
    p = rng.normal(0, 0.4, npar)
    x = rng.normal(size=(32, 6))
    y = rng.normal(size=(32, 5))
    norm = (np.zeros(6), np.ones(6), np.zeros(5), np.ones(5))
    L, g = loss_grad(p, shp, x, y, *norm, emulate_bf16=False)
    for i in rng.choice(npar, 40, replace=False):
        h = 1e-6
        e = np.zeros(npar)
        e[i] = h
        fd = (loss_grad(p + e, shp, x, y, *norm, emulate_bf16=False)[0] -
              loss_grad(p - e, shp, x, y, *norm, emulate_bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.00390625, 1.01171875, 3.14159265, -2.5e-3, 65504.0], np.float32)
    r = bf16(v)
    # bf16 keeps 8 significant bits: ties go to the even mantissa
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1.015625)
    assert abs(r[3] - 3.140625) < 1e-7
    assert np.all(np.abs(r - v) <= np.abs(v) * 2.0 ** -8)
