"""The CPU oracle of the surrogate network, pinned by central finite
differences of its own FP64 forward pass (no BF16 emulation)."""
import numpy as np

import oracle.mlp_oracle as M


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    dims = [4, 7, 6, 5]
    npar = sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))
    flat = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((9, dims[0]))
    T = rng.standard_normal((9, dims[-1]))
    loss, g = M.loss_grad(flat, dims, X, T, emulate_bf16=False)
    h = 1e-6
    num = np.zeros(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        lp, _ = M.loss_grad(flat + e, dims, X, T, emulate_bf16=False)
        lm, _ = M.loss_grad(flat - e, dims, X, T, emulate_bf16=False)
        num[i] = (lp - lm) / (2 * h)
    assert np.abs(num - g).max() < 1e-7 * max(1.0, np.abs(g).max())


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.140625, 1e-3], dtype=np.float64)
    r = M.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie -> even (mantissa of 1.0 is even)
    assert r[2] == 1.0 + 2 ** -7                # 1.5 ulp rounds up
    assert r[3] == -2.5 and r[4] == 3.140625
    assert abs(r[5] - 1e-3) / 1e-3 < 2 ** -8
