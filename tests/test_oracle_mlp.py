"""Pin the CPU MLP oracle: its analytic gradient equals central finite
differences of its own loss (FP64), parameter by parameter on a sample."""
import numpy as np

import mlp_oracle as M


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, k_in, H, B = 5, 16, 8, 7
    sizes = [k_in * H, H, H * H, H, H * H, H, H, 1]
    p = rng.normal(scale=0.5, size=sum(sizes))
    # padded input rows of W1 carry no signal (inputs there are zero)
    x = rng.normal(size=(B, n_in))
    t = rng.normal(size=B)
    loss, g = M.loss_and_grad(p, x, t, k_in, H)
    idx = rng.choice(p.size, size=120, replace=False)
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = M.loss_and_grad(q, x, t, k_in, H)
        q[i] -= 2 * h
        lm, _ = M.loss_and_grad(q, x, t, k_in, H)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_gelu_matches_closed_values():
    assert M.gelu(np.array([0.0]))[0] == 0.0
    # GELU(1) = 0.5 (1 + erf(1/sqrt 2)) = Phi(1) = 0.8413447460685429
    assert abs(M.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    z = np.linspace(-3, 3, 13)
    fd = (M.gelu(z + 1e-6) - M.gelu(z - 1e-6)) / 2e-6
    assert np.max(np.abs(fd - M.gelu_grad(z))) < 1e-8


def test_bf16_rounding():
    v = np.array([1.0, 1.00390625, 1.005859375, -3.14159], dtype=np.float32)
    r = M.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0078125
    assert abs(r[3] + 3.140625) < 1e-12
