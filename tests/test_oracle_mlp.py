"""Pins of the MLP oracle: its analytic gradient against central finite differences (float64, no BF16
rounding so the loss is smooth), and the BF16 rounding against known bit patterns."""
import numpy as np

from oracle import mlp_oracle as MO


def test_bf16_rounding_known_values():
    # bfloat16 keeps 7 mantissa bits: spacing 2^-7 at 1.0
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5, 3.14159265],
                    np.float32)
    got = MO.bf16_round(vals)
    assert got[0] == 1.0
    assert got[1] == 1.0                      # tie between 1 and 1 + 2^-7 -> even mantissa (1.0)
    assert got[2] == 1.0                      # below half a spacing
    assert got[3] == 1.0 + 2 ** -7            # above half a spacing
    assert got[4] == 1.0 + 2 ** -6            # tie between odd 1 + 2^-7 and even 1 + 2^-6
    assert got[5] == -2.5 and abs(got[6] - 3.140625) < 1e-12


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = (7, 6, 5, 4)
    din, h1, h2, dout = dims
    npar = din * h1 + h1 + h1 * h2 + h2 + h2 * dout + dout
    p = rng.standard_normal(npar) * 0.4
    x = rng.standard_normal((9, din))
    t = rng.standard_normal((9, dout))
    loss, g = MO.loss_and_grad(p, x, t, dims, bf16=False)
    eps = 1e-6
    num = np.zeros(npar)
    for i in range(npar):
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        num[i] = (MO.loss_and_grad(pp, x, t, dims, bf16=False)[0] - MO.loss_and_grad(pm, x, t, dims, bf16=False)[0]) / (2 * eps)
    assert np.max(np.abs(num - g)) <= 1e-8 * max(1.0, np.max(np.abs(g)))
    assert loss > 0
