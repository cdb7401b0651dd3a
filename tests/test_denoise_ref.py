"""CPU pins of the denoiser's restatement (oracle/denoise_ref.py): the
backward pass is the derivative of the loss (central differences of the exact
float64 network), the adjoint used for input gradients is the transpose of
the periodic convolution, and the BF16 rounding is round-to-nearest-even."""
import numpy as np

import denoise_ref as D

SHAPE = (5, 4, 3)
V = 5 * 4 * 3


def _random(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(D.n_params()) * scale / np.sqrt(27 * 8)
    x = rng.standard_normal((V, 4)) * 0.5
    t = x + rng.standard_normal((V, 4)) * 0.1
    return p, x, t


def test_param_count():
    assert D.n_params() == (27 * 4 * 32 + 32) + 2 * (27 * 32 * 32 + 32) + (27 * 32 * 4 + 4)


def test_backward_is_the_derivative():
    p, x, t = _random(1)
    loss, _, g = D.forward_backward(p, x, t, SHAPE, bf=False)
    rng = np.random.default_rng(2)
    idx = rng.choice(p.size, 40, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (D.forward_backward(pp, x, t, SHAPE, bf=False)[0] -
              D.forward_backward(pm, x, t, SHAPE, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g).max()) + 1e-4 * abs(fd), (i, fd, g[i])
    assert loss > 0


def test_adjoint_is_the_transpose():
    rng = np.random.default_rng(3)
    nbr = D.neighbours(SHAPE)
    W = rng.standard_normal((27, 6, 5))
    X = rng.standard_normal((V, 6))
    G = rng.standard_normal((V, 5))
    lhs = np.sum(D.conv(X, W, nbr) * G)
    rhs = np.sum(X * D.conv_adj(G, W, nbr))
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)
    # the weight gradient is the derivative of <conv(X, W), G> in W
    assert np.allclose(D.wgrad(X, G, nbr), np.einsum("vtc,vd->tcd", X[nbr], G))


def test_conv_taps_are_periodic_shifts():
    nbr = D.neighbours(SHAPE)
    X = np.zeros((V, 1))
    X[0, 0] = 1.0  # voxel (0,0,0)
    W = np.zeros((27, 1, 1))
    W[2 + 3 * 1 + 9 * 1, 0, 0] = 1.0  # tap (dx, dy, dz) = (+1, 0, 0)
    Y = D.conv(X, W, nbr)[:, 0]
    # Y[v] = X[v + x^]: the one at voxel 0 shows up at x = -1 = nx - 1
    assert Y[SHAPE[0] - 1] == 1.0 and Y.sum() == 1.0


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-39], np.float32)
    r = D.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0                 # tie -> even mantissa
    assert r[2] == 1.0 + 4 * 2 ** -8   # tie -> even (rounds up)
    assert r[3] == -2.5
    assert np.all(np.abs(D.bf16(np.linspace(-3, 3, 101)) - np.linspace(-3, 3, 101)) <=
                  2 ** -8 * 3)
