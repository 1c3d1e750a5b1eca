"""The CPU reference of the viscosity surrogate network (tests/mlp_ref.py), pinned by
central finite differences of its loss and by exact bfloat16 rounding cases."""
import numpy as np

from tests import mlp_ref


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    w = mlp_ref.init(seed=1) + rng.normal(0.0, 0.05, mlp_ref.NPARAM)
    x = rng.normal(size=(40, mlp_ref.NIN))
    t = rng.normal(size=40)
    loss, g = mlp_ref.loss_grad(w, x, t)
    # every bias and a spread of weights of each layer
    idx = list(range(448, 512)) + list(range(4608, 4737)) + list(rng.choice(448, 30, replace=False)) + \
        list(512 + rng.choice(4096, 60, replace=False))
    h = 1e-6
    for p in idx:
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd = (mlp_ref.loss_grad(wp, x, t)[0] - mlp_ref.loss_grad(wm, x, t)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(1.0, abs(g[p])), (p, fd, g[p])


def test_forward_closed_form_for_a_single_unit():
    # all weights zero except one path: y = w3 * silu(w2 * silu(w1 * x0)) + b3
    d = {k: np.zeros(s) for k, s in mlp_ref.SIZES}
    d["W1"][5, 0], d["W2"][7, 5], d["W3"][7], d["b3"][0] = 0.7, -1.3, 2.0, 0.25
    w = mlp_ref.pack(d)
    x = np.array([[0.9, 1, 2, 3, 4, 5, 6]])
    s = lambda z: z / (1 + np.exp(-z))
    y, _ = mlp_ref.forward(w, x)
    assert np.isclose(y[0], 2.0 * s(-1.3 * s(0.7 * 0.9)) + 0.25, rtol=0, atol=1e-15)


def test_bfloat16_rounding():
    # bfloat16 keeps 8 significant bits: near 1 the spacing is 2^-7
    vals = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-8, 1.0 + 3 * 2**-9, -2.5, 3.0e38], dtype=np.float32)
    out = mlp_ref.bf16(vals)
    # 1+2^-8 is a tie and rounds to even (1.0); 1+3*2^-8 is a tie and rounds to even
    # (1+2^-6); 1+3*2^-9 rounds up to 1+2^-7
    assert list(out[:5]) == [1.0, 1.0, 1.0 + 2**-6, 1.0 + 2**-7, -2.5]
    assert np.isfinite(out[5])
