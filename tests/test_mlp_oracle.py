"""Pins the MLP oracle's analytic gradient with central finite differences (float64)."""
import numpy as np

from oracle import mlp_oracle as M


def _blob(rng):
    b = np.zeros(M.NPARAM + 14)
    b[:M.NPARAM] = rng.normal(0, 0.3, M.NPARAM)
    b[M.NPARAM:M.NPARAM + 6] = rng.normal(0, 1, 6)
    b[M.NPARAM + 6:M.NPARAM + 12] = rng.uniform(0.5, 2, 6)
    b[-2], b[-1] = 0.1, 0.7
    return b


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    b = _blob(rng)
    x = rng.normal(0, 1, (50, 6))
    y = rng.normal(0, 1, 50)
    _, g = M.loss_grad(b, x, y)
    idx = rng.choice(M.NPARAM, 60, replace=False)
    for i in idx:
        h = 1e-6
        bp, bm = b.copy(), b.copy()
        bp[i] += h
        bm[i] -= h
        fd = (M.loss_grad(bp, x, y)[0] - M.loss_grad(bm, x, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.01171875, -3.14159, 0.0], np.float32)
    r = M.bf16(v)
    # 1 + 2^-8 and 1 + 3*2^-8 are halfway cases: ties go to the even mantissa
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.015625
    assert abs(r[3] - (-3.140625)) < 1e-12 and r[4] == 0.0
