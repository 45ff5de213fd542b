"""The MLP oracle is pinned by finite differences and BF16 rounding rules."""
import numpy as np

from oracle import mlp as om


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38],
                 dtype=np.float32)
    r = om.bf16_round(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                      # tie -> even mantissa
    assert r[2] == 1.0 + 4 * 2 ** -8        # tie -> even (up)
    assert r[3] == 1.0                      # below half an ulp
    assert r[4] == -2.5
    assert np.isfinite(r[5])


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    widths = (3, 7, 5, 4)
    n = sum(a * b + b for a, b in zip(widths[:-1], widths[1:]))
    p = rng.normal(scale=0.5, size=n)
    x = rng.normal(size=(6, 3))
    t = rng.normal(size=(6, 4))
    loss, g, _ = om.forward_backward(p, widths, x, t, bf16=False)
    idx = list(range(0, n, 3))
    fd = om.finite_difference_grad(p, widths, x, t, idx)
    assert np.max(np.abs(fd - g[idx])) < 1e-7 * max(1.0, np.max(np.abs(g)))


def test_gelu_closed_form():
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    assert np.allclose(om.gelu(z), z * 0.5 * (1 + np.array([__import__("math").erf(v / 2 ** 0.5)
                                                             for v in z])))
    h = 1e-6
    assert np.allclose(om.gelu_d(z), (om.gelu(z + h) - om.gelu(z - h)) / (2 * h), atol=1e-8)
