"""Pins the NumPy MLP oracle (oracle/mlp_oracle.py) before it checks the
engine's surrogate kernels: its backward pass against central finite
differences in float64, and its BF16 rounding against exact cases."""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -7, 3.0e38],
                 dtype=np.float32)
    got = M.bf16(x)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0);
    # 1 + 3*2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6: even is 1 + 2^-6
    want = [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 1.0 + 2 ** -7, 1.765625 * 2.0 ** 127]
    np.testing.assert_allclose(got, want, rtol=1e-4)
    assert got[1] == 1.0 and got[2] == 1.0 + 2 ** -6


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, H, n_out, B = 5, 7, 3, 11
    nw = n_in * H + H + H * H + H + H * n_out + n_out
    w = rng.standard_normal(nw) * 0.5
    X = rng.standard_normal((B, n_in))
    T = rng.standard_normal((B, n_out))
    loss, g = M.loss_and_grad(w, X, T, n_in, H, n_out, bf=False)
    eps = 1e-6
    idx = rng.choice(nw, 40, replace=False)
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += eps
        wm[i] -= eps
        lp, _ = M.loss_and_grad(wp, X, T, n_in, H, n_out, bf=False)
        lm, _ = M.loss_and_grad(wm, X, T, n_in, H, n_out, bf=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_silu_derivative_closed_form():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    fd = (M.silu(z + h) - M.silu(z - h)) / (2 * h)
    np.testing.assert_allclose(M.dsilu(z), fd, atol=1e-8)
    assert M.silu(0.0) == 0.0
