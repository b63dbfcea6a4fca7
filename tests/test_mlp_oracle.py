"""CPU checks of the classifier oracle (oracle/mlp.py): its analytic
gradients against central finite differences of its own loss in float64, and
its BF16 rounding against the definition."""

import numpy as np
import numpy.testing as npt

from oracle import mlp


def _params(rng):
    p = np.zeros(mlp.N_PARAMS)
    p[mlp.OFF_W1:mlp.OFF_B1] = rng.normal(0, 0.5, mlp.N_H * mlp.N_IN)
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.normal(0, 0.1, mlp.N_H)
    p[mlp.OFF_W2:mlp.OFF_B2] = rng.normal(0, 0.2, mlp.N_H * mlp.N_H)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.normal(0, 0.1, mlp.N_H)
    p[mlp.OFF_W3:mlp.OFF_B3] = rng.normal(0, 0.3, mlp.N_H)
    p[mlp.OFF_B3] = 0.1
    return p


def test_gradient_formulas_match_finite_differences():
    rng = np.random.default_rng(0)
    p = _params(rng)
    feat = rng.normal(0, 1, (40, 12))
    t = rng.integers(0, 2, 40)
    mu, sigma = np.zeros(11), np.ones(11)
    g, _ = mlp.gradients(p, feat, t, mu, sigma, exact=True)
    # every parameter block, including the unused W1 columns (gradient 0 there
    # because the padded inputs are 0)
    picks = list(range(0, mlp.N_H * mlp.N_IN, 37)) + list(range(mlp.OFF_B1, mlp.N_PARAMS, 29)) + [mlp.OFF_B3]
    h = 1e-6
    for k in picks:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        yp, _ = mlp.forward(pp, feat, mu, sigma, exact=True)
        ym, _ = mlp.forward(pm, feat, mu, sigma, exact=True)
        fd = (mlp.loss(yp, t) - mlp.loss(ym, t)) / (2 * h)
        npt.assert_allclose(g[k], fd, rtol=1e-4, atol=1e-9, err_msg=f"parameter {k}")


def test_bf16_rounding():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-3, 65504.0], np.float32)
    r = mlp.bf16(x)
    # representable values stay; a value halfway between two BF16 numbers goes to the even one
    assert r[0] == 1.0 and r[3] == -2.5
    assert r[1] == 1.0  # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties to even (1)
    assert r[2] == np.float32(1.0078125)  # 1 + 1.5 * 2^-8 rounds up to 1 + 2^-7
    npt.assert_allclose(r, x, rtol=2 ** -8)
